"""ctypes bindings for the checker libraries.

TEST INFRASTRUCTURE ONLY.  ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this module; the product package
(``dcns_amd``) never does.

Two libraries share one calling convention:
  * ``oracle/liboracle.so``      -- this repo's CPU restatement (``orc_*``)
  * ``oracle/_ref/libdcnsref.so`` -- the reference's own sources compiled by
    ``oracle/Makefile`` (``ref_*``); present only where it has been built.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libdcnsref.so")
REF_CLI = os.path.join(HERE, "_ref", "dCNS_ref")

_P = ctypes.c_void_p
_I = ctypes.c_int32
_D = ctypes.c_double
_L = ctypes.c_int64


def _ptr(a):
    return a.ctypes.data if a is not None and a.size else None


class Checker:
    """One of the two CPU libraries, with numpy-friendly wrappers."""

    def __init__(self, path, prefix):
        self.path = path
        self.lib = ctypes.CDLL(path)
        self.prefix = prefix
        f = self._f
        f("sw").argtypes = [_P, _P, _I, _I, _I, _I, _I, _I, _P]
        f("sw_textbook").argtypes = [_P, _P, _I, _I, _I, _I, _I, _I, _P]
        f("xextend").argtypes = [_P, _P, _I, _I, _I, _I, _I, _I, _I, _I, _P]
        f("xextend_tb").argtypes = [_P, _P, _I, _I, _I, _I, _I, _I, _I, _I, _P, _P, _I]
        f("semiglobal").argtypes = [_P, _P, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _P, _P, _I]
        f("seed_set").argtypes = [_P, _I, _P, _I, _I, _I, _I, _I, _I, _I, _I, _I, _P, _I]
        if prefix == "orc":
            f("find_similar").argtypes = [_P, _I, _P, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _D, _D, _D,
                                          _P, _P, _I, _P, _P, _L, _P]
            f("sw_max_batch").argtypes = [_P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _P, _I]
            f("sw_uses_simd_variant").argtypes = [_I, _I, _I, _I, _I]
            f("find_similar_2gaps").argtypes = [_P, _I, _P, _I] + [_I] * 13 + [_D, _D, _D, _P, _P, _I, _P, _P, _L, _P]
        else:
            f("find_similar").argtypes = [_P, _I, _P, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _D, _D, _D,
                                          _P, _P, _I, _P, _P, _L]
            f("find_similar_2gaps").argtypes = [_P, _I, _P, _I] + [_I] * 13 + [_D, _D, _D, _P, _P, _I, _P, _P, _L]

    def _f(self, name):
        return getattr(self.lib, f"{self.prefix}_{name}")

    @staticmethod
    def _codes(a):
        return np.ascontiguousarray(a, dtype=np.int8)

    def sw(self, ref, read, sc=(2, -3, -4, -2), textbook=False):
        ref, read = self._codes(ref), self._codes(read)
        out = np.zeros(3, np.int32)
        fn = self._f("sw_textbook" if textbook else "sw")
        fn(_ptr(ref), _ptr(read), len(ref), len(read), *sc, out.ctypes.data)
        return tuple(int(x) for x in out)

    def xextend(self, s1, s2, sc=(2, -3, -4, -2), xdrop=20, w=10):
        s1, s2 = self._codes(s1), self._codes(s2)
        out = np.zeros(3, np.int32)
        self._f("xextend")(_ptr(s1), _ptr(s2), len(s1), len(s2), *sc, xdrop, w, out.ctypes.data)
        return tuple(int(x) for x in out)

    def xextend_tb(self, s1, s2, sc=(2, -3, -4, -2), xdrop=20, w=10):
        s1, s2 = self._codes(s1), self._codes(s2)
        out = np.zeros(5, np.int32)
        cap = 2 * (len(s1) + len(s2)) + 8
        cig = np.zeros(cap, np.uint32)
        n = self._f("xextend_tb")(_ptr(s1), _ptr(s2), len(s1), len(s2), *sc, xdrop, w, out.ctypes.data,
                                  cig.ctypes.data, cap)
        assert n >= 0
        return tuple(int(x) for x in out), [int(c) for c in cig[:n]]

    def semiglobal(self, s1, s2, sc=(2, -3, -4, -2, -45, 0), zdrop=50, w=20000, cigar=True):
        s1, s2 = self._codes(s1), self._codes(s2)
        out = np.zeros(3, np.int32)
        cap = 2 * (len(s1) + len(s2)) + 8
        cig = np.zeros(cap, np.uint32)
        n = self._f("semiglobal")(_ptr(s1), _ptr(s2), len(s1), len(s2), *sc, zdrop, w, int(cigar),
                                  out.ctypes.data, cig.ctypes.data, cap)
        assert n >= 0
        return tuple(int(x) for x in out), [int(c) for c in cig[:n]]

    def seed_set(self, s1, s2, sc=(2, -3, -4, -2), window=38, step=8, min_score=40, linked=False):
        s1, s2 = self._codes(s1), self._codes(s2)
        cap = 1 << 16
        out = np.zeros(4 * cap, np.int32)
        n = self._f("seed_set")(_ptr(s1), len(s1), _ptr(s2), len(s2), *sc, window, step, min_score,
                                int(linked), out.ctypes.data, cap)
        assert n >= 0
        return [tuple(int(v) for v in out[4 * i:4 * i + 4]) for i in range(n)]

    def link_seeds(self, seeds, insertion_order=False):
        """link_seeds of a seed list (inserted into the set in the order given):
        the linked set in set order, or (oracle only) with insertion_order the
        merged list in the order link_seeds inserts it back."""
        a = np.ascontiguousarray(np.asarray(seeds, np.int32).reshape(-1, 4))
        out = np.zeros((max(len(a), 1), 4), np.int32)
        fn = self._f("link_seeds")
        if self.prefix == "orc":
            fn.argtypes = [_P, _I, _I, _P, _I]
            n = fn(a.ctypes.data, len(a), int(insertion_order), out.ctypes.data, len(out))
        else:
            assert not insertion_order
            fn.argtypes = [_P, _I, _P, _I]
            n = fn(a.ctypes.data, len(a), out.ctypes.data, len(out))
        assert n >= 0
        return [tuple(int(v) for v in r) for r in out[:n]]

    def find_similar(self, s1, s2, sc=(2, -3, -4, -2), window=38, step=8, min_score=40, w=10, xdrop=20,
                     pvalues=0.1, lam=0.382291, k=0.006662, stats=None):
        """Returns a list of (start1, end1, start2, end2, score, pvalue, evalue, cigar)."""
        s1, s2 = self._codes(s1), self._codes(s2)
        cap = 4096
        cig_cap = 1 << 20
        ints = np.zeros(5 * cap, np.int32)
        dbls = np.zeros(2 * cap, np.float64)
        cig = np.zeros(cig_cap, np.uint32)
        off = np.zeros(cap + 1, np.int64)
        args = [_ptr(s1), len(s1), _ptr(s2), len(s2), *sc, window, step, min_score, w, xdrop, pvalues, lam, k,
                ints.ctypes.data, dbls.ctypes.data, cap, cig.ctypes.data, off.ctypes.data, cig_cap]
        if self.prefix == "orc":
            st = np.zeros(4, np.int64)
            args.append(st.ctypes.data)
        n = self._f("find_similar")(*args)
        assert n >= 0, "capacity"
        if stats is not None and self.prefix == "orc":
            stats.update(scan_cells=int(st[0]), xdrop_cells=int(st[1]), seeds=int(st[2]), linked=int(st[3]))
        res = []
        for i in range(n):
            a = [int(v) for v in ints[5 * i:5 * i + 5]]
            res.append((*a, float(dbls[2 * i]), float(dbls[2 * i + 1]),
                        [int(c) for c in cig[off[i]:off[i + 1]]]))
        return res


def _find_similar_2gaps(self, s1, s2, sc=(2, -3, -4, -2, -45, 0), window=38, step=8, min_score=40, w=10, xdrop=20,
                        zdrop=50, bandwidth=20000, pvalues=0.1, lam=0.382291, k=0.006662, stats=None):
    """findSimilarFragmentsForPairedSequence overload 4: a list of
    (start1, end1, start2, end2, score, pvalue, evalue, cigar)."""
    s1, s2 = self._codes(s1), self._codes(s2)
    cap, cig_cap = 4096, 1 << 22
    ints = np.zeros(5 * cap, np.int32)
    dbls = np.zeros(2 * cap, np.float64)
    cig = np.zeros(cig_cap, np.uint32)
    off = np.zeros(cap + 1, np.int64)
    args = [_ptr(s1), len(s1), _ptr(s2), len(s2), *sc, window, step, min_score, w, xdrop, zdrop, bandwidth, pvalues,
            lam, k, ints.ctypes.data, dbls.ctypes.data, cap, cig.ctypes.data, off.ctypes.data, cig_cap]
    if self.prefix == "orc":
        st = np.zeros(5, np.int64)
        args.append(st.ctypes.data)
    n = self._f("find_similar_2gaps")(*args)
    assert n >= 0, "capacity"
    if stats is not None and self.prefix == "orc":
        stats.update(scan_cells=int(st[0]), xdrop_cells=int(st[1]), seeds=int(st[2]), linked=int(st[3]),
                     semiglobal_cells=int(st[4]))
    return [(*[int(v) for v in ints[5 * i:5 * i + 5]], float(dbls[2 * i]), float(dbls[2 * i + 1]),
             [int(c) for c in cig[off[i]:off[i + 1]]]) for i in range(n)]


Checker.find_similar_2gaps = _find_similar_2gaps

# -- weighted (ZDP) paths ----------------------------------------------------
# profiles: {category: (match, mismatch, open1, ext1, open2, ext2, zdrop)},
# the seven lines of one score-folder file (model/Score.cpp:9-71).


def write_score_folder(path, profiles):
    """Writes a score folder (matrixList + one file per category)."""
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "matrixList"), "w") as f:
        f.write("\n".join(str(c) for c in profiles))
    for c, v in profiles.items():
        with open(os.path.join(path, str(c)), "w") as f:
            f.write("\n".join(str(int(x)) for x in v))
    return path


def _profile_arrays(profiles):
    cats = np.array(list(profiles), np.int16)
    vals = np.array([list(profiles[c]) for c in profiles], np.int32).reshape(-1)
    return cats, vals


def _semiglobal_weighted(self, s1, s2, wcat, profiles, w=20000, two_piece=True, cigar=True, folder=None):
    """Weighted SemiGlobal on one task.  wcat: categories of rows -1 .. L1
    (L1 + 2 values; the weights pointer is wcat + 1).  Returns (max, end1,
    end2, cigar, cells) -- cells from the oracle only (None for the reference)."""
    s1, s2 = self._codes(s1), self._codes(s2)
    wc = np.ascontiguousarray(wcat, np.int16)
    assert wc.size == len(s1) + 2
    out = np.zeros(3, np.int32)
    cap = len(s1) + len(s2) + 16
    cig = np.zeros(cap, np.uint32)
    wptr = wc.ctypes.data + 2
    if self.prefix == "orc":
        cats, vals = _profile_arrays(profiles)
        f = self._f("semiglobal_weighted")
        f.argtypes = [_P, _P, _I, _I, _P, _I, _P, _P, _I, _I, _I, _P, _P, _I, _P]
        cells = np.zeros(1, np.int64)
        n = f(_ptr(s1), _ptr(s2), len(s1), len(s2), wptr, len(cats), cats.ctypes.data, vals.ctypes.data, w,
              int(two_piece), int(cigar), out.ctypes.data, cig.ctypes.data, cap, cells.ctypes.data)
        cl = int(cells[0])
    else:
        assert folder is not None, "the reference reads its profiles from a score folder"
        f = self._f("semiglobal_weighted")
        f.argtypes = [_P, _P, _I, _I, _P, ctypes.c_char_p, _I, _I, _I, _P, _P, _I]
        n = f(_ptr(s1), _ptr(s2), len(s1), len(s2), wptr, folder.encode(), w, int(two_piece), int(cigar),
              out.ctypes.data, cig.ctypes.data, cap)
        cl = None
    assert n >= 0
    return int(out[0]), int(out[1]), int(out[2]), [int(c) for c in cig[:n]], cl


def _find_similar_weighted(self, s1, s2, wcat, profiles, two_piece=True, sc=(2, -3, -4, -2), window=38, step=8,
                           min_score=40, w=10, xdrop=20, bandwidth=20000, pvalues=0.1, lam=0.382291, k=0.006662,
                           folder=None, stats=None):
    """findSimilarFragmentsForPairedSequence_wighted (two_piece) / _wighted_1gap:
    a list of (start1, end1, start2, end2, score, pvalue, evalue, cigar)."""
    s1, s2 = self._codes(s1), self._codes(s2)
    wc = np.ascontiguousarray(wcat, np.int16)
    assert wc.size == len(s1) + 2
    cap, cig_cap = 4096, 1 << 22
    ints = np.zeros(5 * cap, np.int32)
    dbls = np.zeros(2 * cap, np.float64)
    cig = np.zeros(cig_cap, np.uint32)
    off = np.zeros(cap + 1, np.int64)
    head = [_ptr(s1), len(s1), _ptr(s2), len(s2), *sc, window, step, min_score, w, xdrop, bandwidth, pvalues, lam, k,
            wc.ctypes.data]
    tail = [ints.ctypes.data, dbls.ctypes.data, cap, cig.ctypes.data, off.ctypes.data, cig_cap]
    f = self._f("find_similar_weighted")
    if self.prefix == "orc":
        cats, vals = _profile_arrays(profiles)
        f.argtypes = [_P, _I, _P, _I] + [_I] * 10 + [_D, _D, _D, _P, _I, _P, _P, _I, _P, _P, _I, _P, _P, _L, _P]
        st = np.zeros(5, np.int64)
        n = f(*head, len(cats), cats.ctypes.data, vals.ctypes.data, int(two_piece), *tail, st.ctypes.data)
        if stats is not None:
            stats.update(scan_cells=int(st[0]), xdrop_cells=int(st[1]), seeds=int(st[2]), linked=int(st[3]),
                         semiglobal_cells=int(st[4]))
    else:
        assert folder is not None, "the reference reads its profiles from a score folder"
        f.argtypes = [_P, _I, _P, _I] + [_I] * 10 + [_D, _D, _D, _P, ctypes.c_char_p, _I, _P, _P, _I, _P, _P, _L]
        n = f(*head, folder.encode(), int(two_piece), *tail)
    assert n >= 0, "capacity"
    return [(*[int(v) for v in ints[5 * i:5 * i + 5]], float(dbls[2 * i]), float(dbls[2 * i + 1]),
             [int(c) for c in cig[off[i]:off[i + 1]]]) for i in range(n)]


def _score_profile(self, folder, category):
    """The reference's Score(folder) getters for one category (reference only)."""
    f = self._f("score_profile")
    f.argtypes = [ctypes.c_char_p, ctypes.c_int16, _P]
    v = np.zeros(7, np.int32)
    assert f(folder.encode(), category, v.ctypes.data) == 0
    return tuple(int(x) for x in v)


def _gff_categories(self, gff, chroms, which, start, count):
    """readGffFileWithEveryThing's categories of chromosome `which` at
    [start, start + count) (reference only); chroms = [(name, size), ...]."""
    f = self._f("gff_categories")
    f.argtypes = [ctypes.c_char_p, _P, _P, _I, _I, _L, _L, _P]
    names = (ctypes.c_char_p * len(chroms))(*[n.encode() for n, _ in chroms])
    sizes = np.array([s for _, s in chroms], np.int32)
    out = np.zeros(count, np.int16)
    assert f(gff.encode(), ctypes.cast(names, _P), sizes.ctypes.data, len(chroms), which, start, count,
             out.ctypes.data) == 0
    return out


def _find_similar_seeded(self, s1, s2, seeds, sc=(2, -3, -4, -2, -45, 0), zdrop=50, bandwidth=20000, max_len=140000,
                         lam=0.382291, k=0.006662, stats=None):
    """findSimilarFragmentsForPairedSequence overload 3 (cut2Gaps2): seeds =
    [(start1, end1, start2, end2), ...] (0-based).  A list of (start1, end1,
    start2, end2, score, pvalue, evalue, cigar)."""
    s1, s2 = self._codes(s1), self._codes(s2)
    sd = np.ascontiguousarray(np.asarray(seeds, np.int32).reshape(-1, 4))
    cap, cig_cap = 4096, 1 << 22
    ints = np.zeros(5 * cap, np.int32)
    dbls = np.zeros(2 * cap, np.float64)
    cig = np.zeros(cig_cap, np.uint32)
    off = np.zeros(cap + 1, np.int64)
    args = [_ptr(s1), len(s1), _ptr(s2), len(s2), sd.ctypes.data if sd.size else None, len(sd), *sc, zdrop, bandwidth,
            max_len, lam, k, ints.ctypes.data, dbls.ctypes.data, cap, cig.ctypes.data, off.ctypes.data, cig_cap]
    f = self._f("find_similar_seeded")
    base = [_P, _I, _P, _I, _P, _I] + [_I] * 9 + [_D, _D, _P, _P, _I, _P, _P, _L]
    if self.prefix == "orc":
        f.argtypes = base + [_P]
        cl = np.zeros(1, np.int64)
        n = f(*args, cl.ctypes.data)
        if stats is not None:
            stats["semiglobal_cells"] = int(cl[0])
    else:
        f.argtypes = base
        n = f(*args)
    assert n >= 0, "capacity"
    return [(*[int(v) for v in ints[5 * i:5 * i + 5]], float(dbls[2 * i]), float(dbls[2 * i + 1]),
             [int(c) for c in cig[off[i]:off[i + 1]]]) for i in range(n)]


Checker.find_similar_seeded = _find_similar_seeded
Checker.semiglobal_weighted = _semiglobal_weighted
Checker.find_similar_weighted = _find_similar_weighted
Checker.score_profile = _score_profile
Checker.gff_categories = _gff_categories


def oracle():
    """The CPU restatement (always available once ``make -C oracle oracle`` ran)."""
    return Checker(ORACLE_SO, "orc")


def reference():
    """The compiled reference, or None where oracle/_ref has not been built."""
    if not os.path.exists(REF_SO):
        return None
    return Checker(REF_SO, "ref")


def sw_max_batch(seqs, ref_off, ref_len, qry_off, qry_len, sc=(2, -3, -4, -2), nthreads=1):
    """Oracle max local score for a batch of pairs (the ranSco inner call)."""
    o = oracle()
    f = o.lib.orc_sw_max_batch
    f.argtypes = [_P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _P, _I]
    n = len(ref_len)
    out = np.zeros(n, np.int32)
    arrs = [np.ascontiguousarray(seqs, np.int8), np.ascontiguousarray(ref_off, np.int64),
            np.ascontiguousarray(ref_len, np.int32), np.ascontiguousarray(qry_off, np.int64),
            np.ascontiguousarray(qry_len, np.int32)]
    f(*[a.ctypes.data for a in arrs], n, *sc, out.ctypes.data, nthreads)
    return out


# -- maskGenome --------------------------------------------------------------

def mask_genome(d, opts, out="out.fa"):
    """The oracle's maskGenome on the files mask_args wrote into folder d
    (``opts`` as in dcns_amd.synth.mask_case).  Returns (status, stdout text)."""
    f = oracle().lib.orc_mask_genome
    f.argtypes = [ctypes.c_char_p] * 3 + [_I, ctypes.c_char_p, _I, ctypes.c_char_p, ctypes.c_char_p, _D,
                                          ctypes.c_char, _I, _I, ctypes.c_char_p, _I]
    p = lambda name: os.path.join(d, name).encode()  # noqa: E731
    log = ctypes.create_string_buffer(1 << 16)
    rc = f(p("genome.fa"), p(out), p("kmers.fa") if opts["use_dump"] else None, opts["freq"],
           p("ann.gff3") if opts["use_gff"] else None, int(opts["gene"]),
           p("cds.sam") if opts["use_sam"] else None, p("cds.fa") if opts["use_sam"] else None,
           opts["similarity"], opts["alt"].encode()[:1], int(opts["soft"]), opts["width"], log, len(log))
    return rc, log.value.decode()


def reference_mask_genome(d, args, timeout=120):
    """oracle/_ref/dCNS_ref maskGenome in folder d, or None where it is not built."""
    import subprocess
    if not os.path.exists(REF_CLI):
        return None
    return subprocess.run([REF_CLI, "maskGenome"] + list(args), cwd=d, capture_output=True, text=True,
                          timeout=timeout)


def kmer_mask(dump, freq, seqs):
    """The oracle's k-mer scan over byte strings seqs: (k, [0/1 uint8 arrays])."""
    f = oracle().lib.orc_kmer_mask
    f.argtypes = [ctypes.c_char_p, _I, _P, _P, _I, _P]
    buf = np.frombuffer(b"".join(seqs) or b"\0", np.uint8)
    off = np.zeros(len(seqs) + 1, np.int64)
    off[1:] = np.cumsum([len(s) for s in seqs])
    mask = np.zeros(max(1, int(off[-1])), np.uint8)
    k = f(dump.encode() if dump else None, freq, buf.ctypes.data, off.ctypes.data, len(seqs), mask.ctypes.data)
    return k, [mask[off[i]:off[i + 1]].copy() for i in range(len(seqs))]


def kmer_members(dump, freq):
    """The oracle's k-mer profile: (k, sorted list of members of length k)."""
    f = oracle().lib.orc_kmer_members
    f.argtypes = [ctypes.c_char_p, _I, _P, _L, _P]
    k = np.zeros(1, np.int32)
    cap = max(1, os.path.getsize(dump)) if dump and os.path.exists(dump) else 1
    out = np.zeros(cap, np.uint8)
    n = f(dump.encode() if dump else None, freq, out.ctypes.data, cap, k.ctypes.data)
    assert n >= 0, n
    kk = int(k[0])
    return kk, [out[i * kk:(i + 1) * kk].tobytes() for i in range(n)]
