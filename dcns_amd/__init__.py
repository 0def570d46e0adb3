"""dcns_amd -- MI355X (gfx950) implementation of dCNS's alignment hot path.

This module is a thin ctypes binding of the C-ABI declared in include/dcns.h
(library: dcns_amd/lib/libdcns.so, built by ``python -m dcns_amd.build``).
All computation runs in the library's HIP kernels; there is no CPU fallback:
``Context()`` raises when the library is missing or no gfx950 device is
visible.
"""
import ctypes
import weakref
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libdcns.so")
CLI_PATH = os.path.join(PKG, "bin", "dCNS")

_P = ctypes.c_void_p
_I = ctypes.c_int32
_L = ctypes.c_int64
_D = ctypes.c_double


class DcnsError(RuntimeError):
    pass


class Params(ctypes.Structure):
    """dcns_params (defaults: cut1Gap defaults, controlLayer4.cpp:28-52)."""
    _fields_ = [("match", _I), ("mismatch", _I), ("open", _I), ("ext", _I), ("open2", _I), ("ext2", _I),
                ("zdrop", _I), ("bandwidth", _I), ("window", _I), ("step", _I), ("min_score", _I), ("w", _I),
                ("xdrop", _I), ("pad", _I), ("pvalue", _D), ("lambda_", _D), ("k", _D)]


class Batch(ctypes.Structure):
    _fields_ = [("seqs", _P), ("off1", _P), ("len1", _P), ("off2", _P), ("len2", _P), ("n", _I), ("pad", _I)]


class Stats(ctypes.Structure):
    _fields_ = [("scan_cells", _L), ("xdrop_cells", _L), ("seeds", _L), ("linked_seeds", _L), ("fragments", _L),
                ("gpu_scan_cells", _L), ("levels", _I), ("pad", _I), ("ms_seed", _D), ("ms_replay", _D),
                ("ms_xdrop", _D), ("ms_total", _D), ("semiglobal_cells", _L), ("ms_semiglobal", _D)]


class Tuning(ctypes.Structure):
    """dcns_tuning: launch shapes of the seed phase and the x-drop lane kernels (0 = default)."""
    _fields_ = [("level0_mode", _I), ("level0_waves", _I), ("wave_tasks", _I), ("wave_rows", _I),
                ("graph_levels", _I), ("pool_min", _I), ("xdrop_lane_rows", _I), ("k1_lanes", _I),
                ("level0_tail", _I), ("level0_split", _I)]


_lib = None


def load_library():
    """Loads libdcns.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise DcnsError(f"{LIB_PATH} is missing: run `python -m dcns_amd.build` (hipcc, gfx950)")
    lib = ctypes.CDLL(LIB_PATH)
    lib.dcns_version.restype = ctypes.c_char_p
    lib.dcns_last_error.restype = ctypes.c_char_p
    lib.dcns_last_error.argtypes = [_P]
    lib.dcns_default_params.argtypes = [ctypes.POINTER(Params)]
    lib.dcns_ctx_create.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.POINTER(_P)]
    lib.dcns_ctx_destroy.argtypes = [_P]
    lib.dcns_device_info.argtypes = [_P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.c_char_p, _I]
    lib.dcns_last_stats.argtypes = [_P, ctypes.POINTER(Stats)]
    lib.dcns_kernel_time.argtypes = [_P, ctypes.c_char_p, ctypes.POINTER(_D), ctypes.POINTER(_L)]
    lib.dcns_kernel_time_reset.argtypes = [_P]
    lib.dcns_sw_bas_batch.argtypes = [_P, ctypes.POINTER(Batch), _I, _I, _I, _P, _P, _P, _L, _P]
    lib.dcns_sw_batch_noalloc.argtypes = [_P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _P]
    lib.dcns_get_tuning.argtypes = [_P, ctypes.POINTER(Tuning)]
    lib.dcns_set_tuning.argtypes = [_P, ctypes.POINTER(Tuning)]
    lib.dcns_sw_batch.argtypes = [_P, ctypes.POINTER(Batch), _I, _I, _I, _I, _P]
    lib.dcns_xdrop_extend.argtypes = [_P, ctypes.POINTER(Batch), ctypes.POINTER(Params), _P, _P, _I, _P, _P, _P, _L]
    lib.dcns_seed_sets.argtypes = [_P, ctypes.POINTER(Batch), ctypes.POINTER(Params), _I, _P, _P, _L]
    lib.dcns_find_similar_batch.argtypes = [_P, ctypes.POINTER(Batch), ctypes.POINTER(Params), ctypes.POINTER(_P)]
    lib.dcns_stage.argtypes = [_P, ctypes.POINTER(Batch), ctypes.POINTER(Params)]
    lib.dcns_find_similar_staged.argtypes = [_P, ctypes.POINTER(Params), ctypes.POINTER(_P)]
    lib.dcns_frag_set_total.restype = _L
    lib.dcns_frag_set_total.argtypes = [_P]
    lib.dcns_frag_set_pair.argtypes = [_P, _I, ctypes.POINTER(_L), ctypes.POINTER(_L), ctypes.POINTER(_L)]
    lib.dcns_frag_set_get.argtypes = [_P, _L, _P, _P, _P, _I, ctypes.POINTER(_I)]
    lib.dcns_frag_set_free.argtypes = [_P]
    lib.dcns_fit_lambda_k.argtypes = [_P, _L, _P, _D, _D, _D, _I, _P]
    lib.dcns_find_similar_2gaps_batch.argtypes = [_P, ctypes.POINTER(Batch), ctypes.POINTER(Params), ctypes.POINTER(_P)]
    lib.dcns_semiglobal_batch.argtypes = [_P, ctypes.POINTER(Batch), ctypes.POINTER(Params), _P, _P, _P, _I, _I, _P,
                                          _P, _P, _L]
    lib.dcns_read_score_folder.argtypes = [ctypes.c_char_p, _P, _I, ctypes.POINTER(_I)]
    lib.dcns_link_seeds.argtypes = [_P, ctypes.c_int64, _I, _P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
    lib.dcns_gff_categories.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _L, _L, _L, _P]
    lib.dcns_find_similar_weighted_batch.argtypes = [_P, ctypes.POINTER(Batch), ctypes.POINTER(Params), _P, _I, _P,
                                                     _P, _I, ctypes.POINTER(_P)]
    lib.dcns_semiglobal_weighted_batch.argtypes = [_P, ctypes.POINTER(Batch), ctypes.POINTER(Params), _P, _I, _P, _P,
                                                   _I, _P, _P, _P, _I, _I, _P, _P, _P, _L]
    lib.dcns_find_similar_seeded_batch.argtypes = [_P, ctypes.POINTER(Batch), ctypes.POINTER(Params), _P, _P, _I,
                                                   ctypes.POINTER(_P)]
    lib.dcns_calculate_lambda.restype = _D
    lib.dcns_calculate_lambda.argtypes = [_P, _P, ctypes.c_uint32, ctypes.c_uint32, _I, _I, _D, _D]
    lib.dcns_fit_lambda_k_batch.argtypes = [_P, _L, _I, _P, _D, _D, _I, _P]
    lib.dcns_sw_random_max.argtypes = [_P, _I, _I, _I, _I, _P, _P, _P, _P, _I, _I, _P]
    lib.dcns_kmer_set_read.argtypes = [ctypes.c_char_p, _I, ctypes.POINTER(_P)]
    lib.dcns_kmer_set_info.argtypes = [_P, ctypes.POINTER(_I), ctypes.POINTER(_L), ctypes.POINTER(_L)]
    lib.dcns_kmer_set_get.argtypes = [_P, _P, _L]
    lib.dcns_kmer_set_free.argtypes = [_P]
    lib.dcns_cds_mask_create.argtypes = [_P, _P, _I, ctypes.POINTER(_P)]
    lib.dcns_cds_mask_gff.argtypes = [_P, ctypes.c_char_p, _I]
    lib.dcns_cds_mask_sam.argtypes = [_P, ctypes.c_char_p, ctypes.c_char_p, _D, ctypes.POINTER(_I)]
    lib.dcns_cds_mask_get.argtypes = [_P, _I, _P, _L]
    lib.dcns_cds_mask_free.argtypes = [_P]
    lib.dcns_mask_stage.argtypes = [_P, _P, _P, _P, _I]
    lib.dcns_mask_body_size.restype = _L
    lib.dcns_mask_body_size.argtypes = [_L, _I]
    lib.dcns_mask_run.argtypes = [_P, _P, ctypes.c_char, _I, _I, _P, _P]
    lib.dcns_mlp_create.argtypes = [_P, ctypes.c_uint64, ctypes.POINTER(_P)]
    lib.dcns_mlp_destroy.argtypes = [_P]
    lib.dcns_mlp_num_params.restype = _L
    lib.dcns_mlp_num_params.argtypes = [_P]
    lib.dcns_mlp_set_scaling.argtypes = [_P, _P, _P, _P, _P]
    lib.dcns_mlp_get_params.argtypes = [_P, _P, _L]
    lib.dcns_mlp_set_params.argtypes = [_P, _P, _L]
    lib.dcns_mlp_predict.argtypes = [_P, _P, _L, _P]
    lib.dcns_mlp_loss_grad.argtypes = [_P, _P, _P, _L, _P, _P, _L]
    lib.dcns_mlp_train.argtypes = [_P, _P, _P, _L, _I, _I, ctypes.c_float, ctypes.c_float, ctypes.c_uint64, _P]
    _lib = lib
    return lib


# Entry points declared in include/dcns.h (the library must export all of them).
EXPORTS = ["dcns_version", "dcns_default_params", "dcns_ctx_create", "dcns_ctx_destroy", "dcns_last_error",
           "dcns_device_info", "dcns_last_stats", "dcns_kernel_time", "dcns_kernel_time_reset", "dcns_sw_batch",
           "dcns_xdrop_extend", "dcns_seed_sets", "dcns_link_seeds", "dcns_find_similar_batch", "dcns_stage",
           "dcns_find_similar_staged", "dcns_frag_set_total",
           "dcns_frag_set_pair", "dcns_frag_set_get", "dcns_frag_set_free", "dcns_cut1gap_main",
           "dcns_ransco_main", "dcns_fit_lambda_k", "dcns_mlp_create", "dcns_mlp_destroy", "dcns_mlp_num_params",
           "dcns_mlp_set_scaling", "dcns_mlp_get_params", "dcns_mlp_set_params", "dcns_mlp_predict",
           "dcns_mlp_loss_grad", "dcns_mlp_train", "dcns_find_similar_2gaps_batch", "dcns_semiglobal_batch",
           "dcns_cut2gaps_main", "dcns_paircns2gaps_main", "dcns_paircnsxextend_main", "dcns_sw_random_max",
           "dcns_fit_lambda_k_batch", "dcns_calculate_lambda", "dcns_read_score_folder", "dcns_gff_categories",
           "dcns_find_similar_weighted_batch", "dcns_semiglobal_weighted_batch", "dcns_weighted1gap_main",
           "dcns_weighted2gaps_main", "dcns_find_similar_seeded_batch", "dcns_cut2gaps2_main", "dcns_kmer_set_read",
           "dcns_kmer_set_info", "dcns_kmer_set_get", "dcns_kmer_set_free", "dcns_cds_mask_create",
           "dcns_cds_mask_gff", "dcns_cds_mask_sam", "dcns_cds_mask_get", "dcns_cds_mask_free", "dcns_mask_stage",
           "dcns_mask_body_size", "dcns_mask_run", "dcns_mask_genome_main", "dcns_batch_main", "dcns_get_tuning",
           "dcns_set_tuning", "dcns_sw_batch_noalloc", "dcns_sw_bas_batch"]


def default_params(**kw):
    p = Params()
    load_library().dcns_default_params(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, "lambda_" if k == "lambda" else k, v)
    return p


def fit_lambda_k(scores, log_space=None, fixed_space=1e6, k0=0.3, lambda0=0.2, maxiter=550):
    """FP64 Levenberg-Marquardt fit of (K, lambda) to maximum local scores
    (host code of the library; see dcns_fit_lambda_k).  Returns
    (K, lambda, rss, iterations)."""
    lib = load_library()
    sc = np.ascontiguousarray(scores, np.int32)
    ls = None if log_space is None else np.ascontiguousarray(log_space, np.float64)
    out = np.zeros(4, np.float64)
    rc = lib.dcns_fit_lambda_k(sc.ctypes.data, len(sc), None if ls is None else ls.ctypes.data, float(fixed_space),
                               float(k0), float(lambda0), int(maxiter), out.ctypes.data)
    if rc != 0:
        raise DcnsError(f"dcns_fit_lambda_k failed ({rc})")
    return float(out[0]), float(out[1]), float(out[2]), int(out[3])


def calculate_lambda(seq1, match=2, mismatch=-3, epsilon=1e-6, precision=1e-12):
    """Ungapped Karlin-Altschul lambda of seq1's composition
    (dcns_calculate_lambda)."""
    a = np.ascontiguousarray(seq1, np.int8)
    return float(load_library().dcns_calculate_lambda(a.ctypes.data, None, len(a), 0, int(match), int(mismatch),
                                                      float(epsilon), float(precision)))


def fit_lambda_k_batch(scores, space, k0=0.3, lambda0=0.2, maxiter=550):
    """One FP64 fit per row of scores [groups x n] with search space space[g]
    (dcns_fit_lambda_k_batch).  Returns [groups x 4] (K, lambda, rss, iterations)."""
    lib = load_library()
    sc = np.ascontiguousarray(scores, np.int32)
    sp = np.ascontiguousarray(space, np.float64).ravel()
    if sc.ndim != 2 or len(sp) != sc.shape[0]:
        raise ValueError("scores must be [groups x n] with one space per group")
    out = np.zeros((sc.shape[0], 4), np.float64)
    rc = lib.dcns_fit_lambda_k_batch(sc.ctypes.data, sc.shape[0], sc.shape[1], sp.ctypes.data, float(k0),
                                     float(lambda0), int(maxiter), out.ctypes.data)
    if rc != 0:
        raise DcnsError(f"dcns_fit_lambda_k_batch failed ({rc})")
    return out


def link_seeds(seeds, insertion_order=False):
    """link_seeds of a seed list (start1, end1, start2, end2 each), inserted into
    the set in the order given: the linked set in set order, or with
    insertion_order the merged list in the order link_seeds inserts it back
    (dcns_link_seeds; no device needed)."""
    lib = load_library()
    a = np.ascontiguousarray(np.asarray(seeds, np.int32).reshape(-1, 4))
    out = np.zeros((max(len(a), 1), 4), np.int32)
    n = ctypes.c_int64()
    rc = lib.dcns_link_seeds(a.ctypes.data, len(a), int(insertion_order), out.ctypes.data, len(out), ctypes.byref(n))
    if rc != 0:
        raise DcnsError(f"dcns_link_seeds failed ({rc}): {lib.dcns_last_error(None).decode()}")
    return [tuple(int(v) for v in r) for r in out[:n.value]]


# -- weighted (ZDP) inputs ----------------------------------------------------
# profiles: {category: (match, mismatch, open1, ext1, open2, ext2, zdrop)}


def read_score_folder(folder):
    """Score::Score(folder): {category: 7 values} in matrixList order
    (dcns_read_score_folder; no device needed)."""
    lib = load_library()
    out = np.zeros(8 * 256, np.int32)
    n = _I()
    rc = lib.dcns_read_score_folder(folder.encode(), out.ctypes.data, 256, ctypes.byref(n))
    if rc != 0:
        raise DcnsError(f"dcns_read_score_folder failed ({rc}): {lib.dcns_last_error(None).decode()}")
    rows = out[:8 * n.value].reshape(-1, 8)
    return {int(r[0]): tuple(int(x) for x in r[1:]) for r in rows}


def gff_categories(gff_path, chrom, chrom_size, start, count):
    """readGffFileWithEveryThing's categories of chrom at [start, start + count)
    (dcns_gff_categories; no device needed)."""
    lib = load_library()
    out = np.zeros(max(int(count), 1), np.int16)
    rc = lib.dcns_gff_categories(gff_path.encode(), chrom.encode(), int(chrom_size), int(start), int(count),
                                 out.ctypes.data)
    if rc != 0:
        raise DcnsError(f"dcns_gff_categories failed ({rc}): {lib.dcns_last_error(None).decode()}")
    return out[:int(count)]


# -- maskGenome (controlLayer4.cpp:1999-2084) ---------------------------------

def _last_error(lib):
    return lib.dcns_last_error(None).decode()


class KmerSet:
    """The k-mer profile of a jellyfish dump (dcns_kmer_set_read,
    impl/maskGenome.cpp:15-48): records counted more than ``freq`` times,
    upper-cased; ``k`` is the length of the dump's last record.  ``freq <= 0``
    reads nothing (k = 10, no members)."""

    def __init__(self, path, freq):
        self.lib = load_library()
        h = _P()
        rc = self.lib.dcns_kmer_set_read(path.encode() if path else None, int(freq), ctypes.byref(h))
        if rc != 0:
            raise DcnsError(f"dcns_kmer_set_read failed ({rc}): {_last_error(self.lib)}")
        self.h = h
        k, m, r = _I(), _L(), _L()
        self.lib.dcns_kmer_set_info(h, ctypes.byref(k), ctypes.byref(m), ctypes.byref(r))
        self.k, self.members, self.records = k.value, m.value, r.value

    def kmers(self):
        """The members of length k as bytes (a member listed twice appears twice)."""
        buf = np.zeros(max(1, self.members * self.k), np.uint8)
        rc = self.lib.dcns_kmer_set_get(self.h, buf.ctypes.data, buf.size)
        if rc != 0:
            raise DcnsError(f"dcns_kmer_set_get failed ({rc})")
        return [buf[i * self.k:(i + 1) * self.k].tobytes() for i in range(self.members)]

    def close(self):
        if getattr(self, "h", None):
            self.lib.dcns_kmer_set_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CdsMask:
    """The reference's ifCds over a genome of uniquely named sequences
    (dcns_cds_mask_*): gff() is gffToMask / gffToMaskGene, sam() samToMask
    (impl/getAllTheCdsPositions.cpp:9-173).  Host code, no device needed."""

    def __init__(self, names, sizes):
        self.lib = load_library()
        self._names = [ctypes.c_char_p(n.encode()) for n in names]
        arr = (ctypes.c_char_p * max(1, len(names)))(*self._names)
        sz = np.ascontiguousarray(sizes, np.int64)
        h = _P()
        rc = self.lib.dcns_cds_mask_create(ctypes.cast(arr, _P), sz.ctypes.data if len(sz) else None, len(names),
                                           ctypes.byref(h))
        if rc != 0:
            raise DcnsError(f"dcns_cds_mask_create failed ({rc}): {_last_error(self.lib)}")
        self.h = h
        self.sizes = [int(x) for x in sz]

    def gff(self, path, gene=False):
        rc = self.lib.dcns_cds_mask_gff(self.h, path.encode(), int(bool(gene)))
        if rc != 0:
            raise DcnsError(f"dcns_cds_mask_gff failed ({rc}): {_last_error(self.lib)}")

    def sam(self, sam_path, cds_fasta, similarity=0.8):
        """Returns True when the masks were first rebuilt empty."""
        r = _I()
        rc = self.lib.dcns_cds_mask_sam(self.h, sam_path.encode(), cds_fasta.encode(), float(similarity),
                                        ctypes.byref(r))
        if rc != 0:
            raise DcnsError(f"dcns_cds_mask_sam failed ({rc}): {_last_error(self.lib)}")
        return bool(r.value)

    def mask(self, s):
        out = np.zeros(max(1, self.sizes[s]), np.uint8)
        rc = self.lib.dcns_cds_mask_get(self.h, int(s), out.ctypes.data, out.size)
        if rc != 0:
            raise DcnsError(f"dcns_cds_mask_get failed ({rc}): {_last_error(self.lib)}")
        return out[:self.sizes[s]]

    def close(self):
        if getattr(self, "h", None):
            self.lib.dcns_cds_mask_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mask_body_size(length, width):
    """Bytes of one masked FASTA body (bases, line breaks, closing newline)."""
    return int(load_library().dcns_mask_body_size(int(length), int(width)))


def _profile_array(profiles):
    return np.array([[int(c), *[int(x) for x in v]] for c, v in profiles.items()], np.int32).reshape(-1)


def _category_arrays(pairs, wcats):
    """Concatenates per-pair category arrays (len(seq1) + 2 values each)."""
    off = np.zeros(len(pairs), np.int64)
    parts, o = [], 0
    for i, ((a, _), w) in enumerate(zip(pairs, wcats)):
        w = np.ascontiguousarray(w, np.int16)
        if w.size != len(a) + 2:
            raise ValueError("categories must cover reference positions -1 .. len1 (len1 + 2 values)")
        off[i] = o
        parts.append(w)
        o += w.size
    cat = np.concatenate(parts) if parts else np.zeros(1, np.int16)
    return cat, off


class _BatchArrays:
    """Keeps the numpy arrays of a Batch alive."""

    def __init__(self, pairs):
        s1s = [np.ascontiguousarray(a, np.int8) for a, _ in pairs]
        s2s = [np.ascontiguousarray(b, np.int8) for _, b in pairs]
        l1 = np.array([len(a) for a in s1s], np.int32)
        l2 = np.array([len(b) for b in s2s], np.int32)
        chunks, off1, off2, o = [], [], [], 0
        for a, b in zip(s1s, s2s):
            off1.append(o)
            chunks.append(a)
            o += len(a)
            off2.append(o)
            chunks.append(b)
            o += len(b)
        self.seqs = np.concatenate(chunks) if chunks else np.zeros(1, np.int8)
        if self.seqs.size == 0:
            self.seqs = np.zeros(1, np.int8)
        self.off1 = np.array(off1, np.int64)
        self.off2 = np.array(off2, np.int64)
        self.len1, self.len2 = l1, l2
        self.b = Batch(self.seqs.ctypes.data, self.off1.ctypes.data if len(pairs) else None, l1.ctypes.data
                       if len(pairs) else None, self.off2.ctypes.data if len(pairs) else None,
                       l2.ctypes.data if len(pairs) else None, len(pairs), 0)

    @classmethod
    def from_flat(cls, seqs, off1, len1, off2, len2):
        self = cls.__new__(cls)
        self.seqs = np.ascontiguousarray(seqs, np.int8)
        self.off1 = np.ascontiguousarray(off1, np.int64)
        self.len1 = np.ascontiguousarray(len1, np.int32)
        self.off2 = np.ascontiguousarray(off2, np.int64)
        self.len2 = np.ascontiguousarray(len2, np.int32)
        self.b = Batch(self.seqs.ctypes.data, self.off1.ctypes.data, self.len1.ctypes.data, self.off2.ctypes.data,
                       self.len2.ctypes.data, len(self.len1), 0)
        return self


class Context:
    """One HIP device + stream (dcns_ctx)."""

    def __init__(self, device=0, timing=False):
        self.lib = load_library()
        h = _P()
        rc = self.lib.dcns_ctx_create(device, 1 if timing else 0, ctypes.byref(h))
        if rc != 0:
            raise DcnsError(f"dcns_ctx_create failed ({rc}): {self.lib.dcns_last_error(None).decode()}")
        self.h = h
        self._children = weakref.WeakSet()  # networks bound to this context

    def close(self):
        if getattr(self, "h", None):
            for child in list(self._children):
                child.close()
            self.lib.dcns_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise DcnsError(f"{what} failed ({rc}): {self.lib.dcns_last_error(self.h).decode()}")

    def device_info(self):
        cu, clk = _I(), _I()
        name = ctypes.create_string_buffer(64)
        self._check(self.lib.dcns_device_info(self.h, ctypes.byref(cu), ctypes.byref(clk), name, 64), "device_info")
        return {"cu": cu.value, "clock_mhz": clk.value, "arch": name.value.decode()}

    def stats(self):
        s = Stats()
        self._check(self.lib.dcns_last_stats(self.h, ctypes.byref(s)), "last_stats")
        return {f: getattr(s, f) for f, _ in Stats._fields_ if f != "pad"}

    def kernel_time(self, name):
        ms, n = _D(), _L()
        self._check(self.lib.dcns_kernel_time(self.h, name.encode(), ctypes.byref(ms), ctypes.byref(n)),
                    "kernel_time")
        return ms.value, n.value

    def kernel_time_reset(self):
        self._check(self.lib.dcns_kernel_time_reset(self.h), "kernel_time_reset")

    def get_tuning(self):
        t = Tuning()
        self._check(self.lib.dcns_get_tuning(self.h, ctypes.byref(t)), "dcns_get_tuning")
        return {f: getattr(t, f) for f, _ in Tuning._fields_ if not f.startswith("pad")}

    def set_tuning(self, **kw):
        """Sets the given dcns_tuning fields (others keep their values);
        returns the previous settings."""
        old = self.get_tuning()
        t = Tuning(**{**old, **kw})
        self._check(self.lib.dcns_set_tuning(self.h, ctypes.byref(t)), "dcns_set_tuning")
        return old

    # -- K1 -----------------------------------------------------------------
    def sw_batch(self, pairs=None, sc=(2, -3, -4, -2), flat=None):
        """SmithWaterman (max, end1, end2) per pair; pairs = [(ref, read), ...]
        or flat = (seqs, off1, len1, off2, len2)."""
        ba = _BatchArrays.from_flat(*flat) if flat is not None else _BatchArrays(pairs)
        n = ba.b.n
        out = np.zeros(3 * max(n, 1), np.int32)
        self._check(self.lib.dcns_sw_batch(self.h, ctypes.byref(ba.b), *sc, out.ctypes.data), "dcns_sw_batch")
        return out[:3 * n].reshape(n, 3)

    def sw_bas(self, pairs, open=-2, ext=-1, reverse=False):
        """SmithWaterman_bas per pair: (maxScore, end1, end2) as uint32 and,
        with reverse, (cigar list, status) per pair."""
        ba = _BatchArrays(pairs)
        n = ba.b.n
        out = np.zeros(3 * max(n, 1), np.uint32)
        cap = sum(2 * (len(a) + len(b)) + 4 for a, b in pairs) + 16
        cig = np.zeros(cap, np.uint32)
        off = np.zeros(n + 1, np.int64)
        st = np.zeros(max(n, 1), np.int32)
        self._check(self.lib.dcns_sw_bas_batch(self.h, ctypes.byref(ba.b), int(open), int(ext), int(bool(reverse)),
                                               out.ctypes.data, cig.ctypes.data, off.ctypes.data, cap,
                                               st.ctypes.data), "dcns_sw_bas_batch")
        res = [tuple(int(v) for v in out[3 * i:3 * i + 3]) for i in range(n)]
        if not reverse:
            return res
        return [(r, [int(c) for c in cig[off[i]:off[i + 1]]], int(st[i])) for i, r in enumerate(res)]

    def sw_random_max(self, sc, gc, len1, len2, seed, pairs_per_point):
        """Max local scores of random pairs generated on the device
        (dcns_sw_random_max): returns [npoints x pairs_per_point] int32."""
        gc = np.ascontiguousarray(gc, np.float64).ravel()
        l1 = np.ascontiguousarray(len1, np.int32).ravel()
        l2 = np.ascontiguousarray(len2, np.int32).ravel()
        sd = np.ascontiguousarray(seed, np.uint64).ravel()
        n = len(gc)
        if not (len(l1) == len(l2) == len(sd) == n):
            raise ValueError("gc, len1, len2 and seed must have one entry per point")
        out = np.zeros((n, pairs_per_point), np.int32)
        self._check(self.lib.dcns_sw_random_max(self.h, *sc, gc.ctypes.data, l1.ctypes.data, l2.ctypes.data,
                                                sd.ctypes.data, n, int(pairs_per_point), out.ctypes.data),
                    "dcns_sw_random_max")
        return out

    # -- x-drop -------------------------------------------------------------
    def xdrop_extend(self, pairs, pair_of_seed, starts, params=None):
        """x_extend_seed with traceback.  starts[i] = (start1, start2) (seed start + 1).
        Returns a list of (score, start1, end1, start2, end2, ok, cigar)."""
        ba = _BatchArrays(pairs)
        p = params or default_params()
        pos = np.ascontiguousarray(pair_of_seed, np.int32)
        st = np.ascontiguousarray(np.asarray(starts, np.int32).reshape(-1, 2))
        n = len(pos)
        ints = np.zeros(6 * max(n, 1), np.int32)
        cap = int(sum(int(ba.len1[q]) + int(ba.len2[q]) + 4 for q in pos)) + 16
        cig = np.zeros(cap, np.uint32)
        off = np.zeros(n + 1, np.int64)
        self._check(self.lib.dcns_xdrop_extend(self.h, ctypes.byref(ba.b), ctypes.byref(p), pos.ctypes.data,
                                               st.ctypes.data, n, ints.ctypes.data, cig.ctypes.data,
                                               off.ctypes.data, cap), "dcns_xdrop_extend")
        return [(*[int(v) for v in ints[6 * i:6 * i + 6]], [int(c) for c in cig[off[i]:off[i + 1]]])
                for i in range(n)]

    # -- seeds --------------------------------------------------------------
    def seed_sets(self, pairs, params=None, linked=False):
        ba = _BatchArrays(pairs)
        p = params or default_params()
        n = ba.b.n
        counts = np.zeros(max(n, 1), np.int32)
        cap = 1 << 16
        while True:
            seeds = np.zeros(4 * cap, np.int32)
            rc = self.lib.dcns_seed_sets(self.h, ctypes.byref(ba.b), ctypes.byref(p), int(linked),
                                         counts.ctypes.data, seeds.ctypes.data, cap)
            if rc == -3:
                cap = int(counts[:n].sum()) + 16
                continue
            self._check(rc, "dcns_seed_sets")
            break
        out, o = [], 0
        for i in range(n):
            c = int(counts[i])
            out.append([tuple(int(v) for v in seeds[4 * k:4 * k + 4]) for k in range(o, o + c)])
            o += c
        return out

    # -- whole path ---------------------------------------------------------
    def find_similar(self, pairs=None, params=None, flat=None, with_cigar=True):
        """findSimilarFragmentsForPairedSequence for every pair.  Returns, per
        pair, (seeds_size, [(start1, end1, start2, end2, score, p, e, cigar), ...])."""
        ba = _BatchArrays.from_flat(*flat) if flat is not None else _BatchArrays(pairs)
        p = params or default_params()
        fs = _P()
        self._check(self.lib.dcns_find_similar_batch(self.h, ctypes.byref(ba.b), ctypes.byref(p), ctypes.byref(fs)),
                    "dcns_find_similar_batch")
        return self._collect(fs, ba.b.n, with_cigar)

    def find_similar_2gaps(self, pairs=None, params=None, flat=None):
        """findSimilarFragmentsForPairedSequence overload 4 (the two-piece path)
        per pair: [(seeds_size, [(start1, end1, start2, end2, score, p, e, cigar), ...]), ...]."""
        ba = _BatchArrays.from_flat(*flat) if flat is not None else _BatchArrays(pairs)
        p = params or default_params()
        fs = _P()
        self._check(self.lib.dcns_find_similar_2gaps_batch(self.h, ctypes.byref(ba.b), ctypes.byref(p),
                                                           ctypes.byref(fs)), "dcns_find_similar_2gaps_batch")
        return self._collect(fs, ba.b.n, True)

    def semiglobal(self, pairs, pair_of_task, anchors, reverse=None, params=None, with_cigar=True):
        """SemiGlobal for many tasks: (out3 [n x 3], cigars or None)."""
        ba = _BatchArrays(pairs)
        p = params or default_params()
        pt = np.ascontiguousarray(pair_of_task, np.int32)
        an = np.ascontiguousarray(anchors, np.int32).reshape(-1)
        n = len(pt)
        rv = None if reverse is None else np.ascontiguousarray(reverse, np.int32)
        out = np.zeros(3 * max(n, 1), np.int32)
        lens = [len(a) + len(b) + 2 for a, b in pairs]
        cap = max(int(sum(lens[int(i)] for i in pt)), 16) if with_cigar else 16
        cig = np.zeros(cap, np.uint32)
        off = np.zeros(n + 1, np.int64)
        self._check(self.lib.dcns_semiglobal_batch(self.h, ctypes.byref(ba.b), ctypes.byref(p), pt.ctypes.data,
                                                   an.ctypes.data, None if rv is None else rv.ctypes.data, n,
                                                   1 if with_cigar else 0, out.ctypes.data, cig.ctypes.data,
                                                   off.ctypes.data, cap), "dcns_semiglobal_batch")
        res = out[:3 * n].reshape(n, 3)
        if not with_cigar:
            return res, None
        return res, [[int(c) for c in cig[off[i]:off[i + 1]]] for i in range(n)]

    def find_similar_weighted(self, pairs, wcats, profiles, two_piece=True, params=None):
        """findSimilarFragmentsForPairedSequence_wighted (two_piece) / _wighted_1gap
        per pair: [(seeds_size, [(start1, end1, start2, end2, score, p, e, cigar), ...]), ...].
        wcats[i]: categories of pair i's reference positions -1 .. len1."""
        ba = _BatchArrays(pairs)
        p = params or default_params()
        pr = _profile_array(profiles)
        cat, off = _category_arrays(pairs, wcats)
        fs = _P()
        self._check(self.lib.dcns_find_similar_weighted_batch(self.h, ctypes.byref(ba.b), ctypes.byref(p),
                                                              pr.ctypes.data, len(profiles), cat.ctypes.data,
                                                              off.ctypes.data, int(bool(two_piece)),
                                                              ctypes.byref(fs)), "dcns_find_similar_weighted_batch")
        return self._collect(fs, ba.b.n, True)

    def find_similar_seeded(self, pairs, seeds, max_align_length=140000, params=None):
        """findSimilarFragmentsForPairedSequence overload 3 (cut2Gaps2) per pair:
        seeds[i] = [(start1, end1, start2, end2), ...] (0-based).  Returns
        [(n_seeds, [(start1, end1, start2, end2, score, p, e, cigar), ...]), ...]."""
        ba = _BatchArrays(pairs)
        p = params or default_params()
        first = np.zeros(len(pairs) + 1, np.int64)
        flat = []
        for i, s in enumerate(seeds):
            flat.extend(int(v) for t in s for v in t)
            first[i + 1] = first[i] + len(s)
        sd = np.array(flat if flat else [0, 0, 0, 0], np.int32)
        fs = _P()
        self._check(self.lib.dcns_find_similar_seeded_batch(self.h, ctypes.byref(ba.b), ctypes.byref(p),
                                                            sd.ctypes.data, first.ctypes.data, int(max_align_length),
                                                            ctypes.byref(fs)), "dcns_find_similar_seeded_batch")
        return self._collect(fs, ba.b.n, True)

    def semiglobal_weighted(self, pairs, wcats, profiles, pair_of_task, anchors, reverse=None, two_piece=True,
                            params=None, with_cigar=True):
        """The weighted SemiGlobal for many tasks: (out3 [n x 3], cigars or None)."""
        ba = _BatchArrays(pairs)
        p = params or default_params()
        pr = _profile_array(profiles)
        cat, coff = _category_arrays(pairs, wcats)
        pt = np.ascontiguousarray(pair_of_task, np.int32)
        an = np.ascontiguousarray(anchors, np.int32).reshape(-1)
        n = len(pt)
        rv = None if reverse is None else np.ascontiguousarray(reverse, np.int32)
        out = np.zeros(3 * max(n, 1), np.int32)
        lens = [len(a) + len(b) + 2 for a, b in pairs]
        cap = max(int(sum(lens[int(i)] for i in pt)), 16) if with_cigar else 16
        cig = np.zeros(cap, np.uint32)
        off = np.zeros(n + 1, np.int64)
        self._check(self.lib.dcns_semiglobal_weighted_batch(
            self.h, ctypes.byref(ba.b), ctypes.byref(p), pr.ctypes.data, len(profiles), cat.ctypes.data,
            coff.ctypes.data, int(bool(two_piece)), pt.ctypes.data, an.ctypes.data,
            None if rv is None else rv.ctypes.data, n, 1 if with_cigar else 0, out.ctypes.data, cig.ctypes.data,
            off.ctypes.data, cap), "dcns_semiglobal_weighted_batch")
        res = out[:3 * n].reshape(n, 3)
        if not with_cigar:
            return res, None
        return res, [[int(c) for c in cig[off[i]:off[i + 1]]] for i in range(n)]

    # -- maskGenome ----------------------------------------------------------
    def mask_stage(self, seqs, kmers=None):
        """Uploads sequences (bytes) and the k-mer set (KmerSet or None)."""
        buf = np.frombuffer(b"".join(seqs) or b"\0", np.uint8)
        off = np.zeros(len(seqs) + 1, np.int64)
        off[1:] = np.cumsum([len(x) for x in seqs])
        self._mask_off = off
        self._check(self.lib.dcns_mask_stage(self.h, kmers.h if kmers is not None else None, buf.ctypes.data,
                                             off.ctypes.data, len(seqs)), "mask_stage")

    def mask_run(self, cds=None, alt="n", soft=False, width=60, want_mask=True, want_bodies=True):
        """Masks the staged genome.  Returns (masks, bodies): per sequence a
        0/1 uint8 array and the FASTA body bytes (None when not wanted)."""
        off = self._mask_off
        n = len(off) - 1
        mask = np.zeros(max(1, int(off[-1])), np.uint8) if want_mask else None
        sizes = [mask_body_size(int(off[i + 1] - off[i]), width) for i in range(n)] if width != 0 else [0] * n
        boff = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        body = np.zeros(max(1, int(boff[-1])), np.uint8) if want_bodies else None
        a = alt.encode()[:1] if isinstance(alt, str) else bytes([alt])
        self._check(self.lib.dcns_mask_run(self.h, cds.h if cds is not None else None, a or b"\0", int(bool(soft)),
                                           int(width), mask.ctypes.data if mask is not None else None,
                                           body.ctypes.data if body is not None else None), "mask_run")
        masks = [mask[off[i]:off[i + 1]] for i in range(n)] if want_mask else None
        bodies = [body[boff[i]:boff[i + 1]].tobytes() for i in range(n)] if want_bodies else None
        return masks, bodies

    def mask_genome(self, seqs, kmers=None, cds=None, alt="n", soft=False, width=60):
        self.mask_stage(seqs, kmers)
        return self.mask_run(cds, alt, soft, width)

    def stage(self, pairs=None, params=None, flat=None):
        """Keeps a batch resident in device memory for find_similar_staged()."""
        ba = _BatchArrays.from_flat(*flat) if flat is not None else _BatchArrays(pairs)
        p = params or default_params()
        self._check(self.lib.dcns_stage(self.h, ctypes.byref(ba.b), ctypes.byref(p)), "dcns_stage")
        self._staged_n = ba.b.n

    def find_similar_staged(self, params=None, collect=True, with_cigar=True):
        p = params or default_params()
        fs = _P()
        self._check(self.lib.dcns_find_similar_staged(self.h, ctypes.byref(p), ctypes.byref(fs)),
                    "dcns_find_similar_staged")
        if not collect:
            n = self.lib.dcns_frag_set_total(fs)
            self.lib.dcns_frag_set_free(fs)
            return n
        return self._collect(fs, self._staged_n, with_cigar)

    def _collect(self, fs, npairs, with_cigar):
        try:
            res = []
            ints = np.zeros(5, np.int32)
            dbls = np.zeros(2, np.float64)
            cig = np.zeros(1 << 16, np.uint32)
            nc = _I()
            for i in range(npairs):
                first, count, ss = _L(), _L(), _L()
                self.lib.dcns_frag_set_pair(fs, i, ctypes.byref(first), ctypes.byref(count), ctypes.byref(ss))
                frs = []
                for k in range(first.value, first.value + count.value):
                    rc = self.lib.dcns_frag_set_get(fs, k, ints.ctypes.data, dbls.ctypes.data,
                                                    cig.ctypes.data if with_cigar else None, len(cig),
                                                    ctypes.byref(nc))
                    self._check(rc, "dcns_frag_set_get")
                    c = [int(x) for x in cig[:nc.value]] if with_cigar else None
                    frs.append((*[int(v) for v in ints], float(dbls[0]), float(dbls[1]), c))
                res.append((ss.value, frs))
            return res
        finally:
            self.lib.dcns_frag_set_free(fs)


class Mlp:
    """The lambda/K surrogate network on the device (dcns_mlp_*): 8 inputs ->
    4 x 256 GELU -> (ln lambda, ln K), BF16 matrix-core GEMMs with FP32
    accumulation and FP32 AdamW master weights."""

    N_IN, N_OUT = 8, 2

    def __init__(self, ctx, seed=0):
        self.ctx, self.lib = ctx, ctx.lib
        h = _P()
        ctx._check(self.lib.dcns_mlp_create(ctx.h, int(seed), ctypes.byref(h)), "dcns_mlp_create")
        self.h = h
        self.n_params = int(self.lib.dcns_mlp_num_params(h))
        ctx._children.add(self)

    def close(self):
        if getattr(self, "h", None):
            self.lib.dcns_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _f32(a, cols=None):
        a = np.ascontiguousarray(a, np.float32)
        if cols is not None and (a.ndim != 2 or a.shape[1] != cols):
            raise ValueError(f"expected an [n x {cols}] array, got shape {a.shape}")
        return a

    def set_scaling(self, in_mean, in_std, out_mean, out_std):
        arrs = [self._f32(v).ravel() for v in (in_mean, in_std, out_mean, out_std)]
        if [len(a) for a in arrs] != [8, 8, 2, 2]:
            raise ValueError("scaling takes 8 input and 2 output means / deviations")
        self.ctx._check(self.lib.dcns_mlp_set_scaling(self.h, *[a.ctypes.data for a in arrs]), "dcns_mlp_set_scaling")

    def get_params(self):
        out = np.zeros(self.n_params, np.float32)
        self.ctx._check(self.lib.dcns_mlp_get_params(self.h, out.ctypes.data, len(out)), "dcns_mlp_get_params")
        return out

    def set_params(self, flat):
        flat = self._f32(flat).ravel()
        self.ctx._check(self.lib.dcns_mlp_set_params(self.h, flat.ctypes.data, len(flat)), "dcns_mlp_set_params")

    def predict(self, X):
        X = self._f32(X, self.N_IN)
        Y = np.zeros((len(X), self.N_OUT), np.float32)
        self.ctx._check(self.lib.dcns_mlp_predict(self.h, X.ctypes.data, len(X), Y.ctypes.data), "dcns_mlp_predict")
        return Y

    def loss_grad(self, X, Y):
        X, Y = self._f32(X, self.N_IN), self._f32(Y, self.N_OUT)
        loss = ctypes.c_float()
        g = np.zeros(self.n_params, np.float32)
        self.ctx._check(self.lib.dcns_mlp_loss_grad(self.h, X.ctypes.data, Y.ctypes.data, len(X), ctypes.byref(loss),
                                                    g.ctypes.data, len(g)), "dcns_mlp_loss_grad")
        return float(loss.value), g

    def train(self, X, Y, steps, batch=4096, lr=1e-3, weight_decay=1e-4, seed=0):
        X, Y = self._f32(X, self.N_IN), self._f32(Y, self.N_OUT)
        losses = np.zeros(max(steps, 1), np.float32)
        self.ctx._check(self.lib.dcns_mlp_train(self.h, X.ctypes.data, Y.ctypes.data, len(X), int(steps), int(batch),
                                                float(lr), float(weight_decay), int(seed), losses.ctypes.data),
                        "dcns_mlp_train")
        return losses[:steps]
