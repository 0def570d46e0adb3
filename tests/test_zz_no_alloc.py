"""dcns_sw_batch_noalloc (pairs in device memory, the stripe boundaries in
the library's static device array, no allocation) against dcns_sw_batch on
the same pairs: two calls in a row, the second smaller, then the size checks.
The test's own device buffers come from the HIP runtime the library links
(libamdhip64.so.7, through ctypes), the process's only HIP runtime."""
import ctypes

import numpy as np
import pytest

import dcns_amd
from dcns_amd import synth

pytestmark = pytest.mark.gpu


class _Dev:
    """A device buffer holding a numpy array (hipMalloc / hipMemcpy)."""
    hip = None

    def __init__(self, arr):
        if _Dev.hip is None:
            _Dev.hip = ctypes.CDLL("libamdhip64.so.7")
        self.arr = np.ascontiguousarray(arr)
        self.p = ctypes.c_void_p()
        assert _Dev.hip.hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(max(self.arr.nbytes, 16))) == 0
        assert _Dev.hip.hipMemcpy(self.p, self.arr.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(self.arr.nbytes),
                                  1) == 0  # hipMemcpyHostToDevice

    def get(self):
        out = np.empty_like(self.arr)
        assert _Dev.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p, ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def __del__(self):
        if self.p:
            _Dev.hip.hipFree(self.p)


def _device_pairs(pairs):
    """Codes laid out as the entry point needs them: offsets multiples of
    16, round_up(len, 16) reference bytes and round_up(len, 32) query bytes
    readable (padded with N)."""
    off1, off2, l1, l2 = [], [], [], []
    o = 0
    for a, b in pairs:
        off1.append(o); l1.append(len(a)); o += (len(a) + 15) // 16 * 16 + 16
        off2.append(o); l2.append(len(b)); o += (len(b) + 31) // 32 * 32 + 32
    buf = np.full(o + 64, 2, np.int8)
    for (a, b), x, y in zip(pairs, off1, off2):
        buf[x:x + len(a)] = a
        buf[y:y + len(b)] = b
    dev = lambda v, t: _Dev(np.asarray(v, t))  # noqa: E731
    return (dev(buf, np.int8), dev(off1, np.int64), dev(l1, np.int32), dev(off2, np.int64), dev(l2, np.int32),
            max(l1) if l1 else 0, len(l1))


def _noalloc(gpu, d, sc, max_len1=None):
    seqs, o1, l1, o2, l2, m, n = d
    out = _Dev(np.full(max(n, 1) * 3, -7, np.int32))
    rc = gpu.lib.dcns_sw_batch_noalloc(gpu.h, seqs.p, o1.p, l1.p, o2.p, l2.p, n, m if max_len1 is None else max_len1,
                                       *sc, out.p)
    return rc, out.get()[:3 * n].reshape(n, 3)


def _pairs(seed, n, lo, hi):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        a = synth.random_seq(rng, int(rng.integers(lo, hi)))
        b = synth.random_seq(rng, int(rng.integers(lo, hi)))
        if len(a) > 60 and len(b) > 60 and rng.random() < 0.6:  # a shared stretch
            k = int(rng.integers(0, len(a) - 50))
            m = synth.mutate(rng, a[k:k + 50], 0.1, 0.02)[:len(b)]
            b[:len(m)] = m
        out.append((a, b))
    return out


@pytest.mark.parametrize("sc", [(2, -3, -4, -2), (1, -1, -2, -1)])
def test_noalloc_matches_sw_batch(gpu, sc):
    first = _pairs(11, 5000, 1, 2400)  # more boundary rows than one launch of the array holds
    second = _pairs(12, 200, 30, 700)   # then a smaller call
    for name, pairs in (("first", first), ("second", second)):
        print("sw_batch_noalloc", name, len(pairs), sc, flush=True)
        rc, got = _noalloc(gpu, _device_pairs(pairs), sc)
        assert rc == 0, gpu.lib.dcns_last_error(gpu.h).decode()
        want = gpu.sw_batch(pairs, sc)
        assert np.array_equal(got, want), name
        assert got[:, 0].max() > 0


def test_noalloc_size_checks(gpu):
    pairs = _pairs(13, 40, 100, 900)
    d = _device_pairs(pairs)
    # max_len1 above what the array holds: rejected before any launch
    rc, _ = _noalloc(gpu, d, (2, -3, -4, -2), max_len1=10 ** 8)
    assert rc == -1
    assert b"max_len1" in gpu.lib.dcns_last_error(gpu.h)
    # a pair longer than the max_len1 the caller promised: reported, the others computed
    short = min(len(a) for a, _ in pairs)
    rc, got = _noalloc(gpu, d, (2, -3, -4, -2), max_len1=short)
    assert rc == -1 and b"longer than max_len1" in gpu.lib.dcns_last_error(gpu.h)
    want = gpu.sw_batch(pairs)
    for i, (a, _) in enumerate(pairs):
        if len(a) <= short:
            assert tuple(got[i]) == tuple(want[i])
        else:
            assert tuple(got[i]) == (-7, -7, -7)
    # nothing to do
    rc = gpu.lib.dcns_sw_batch_noalloc(gpu.h, None, None, None, None, None, 0, 0, 2, -3, -4, -2, None)
    assert rc == 0
