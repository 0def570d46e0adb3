"""C-ABI library and host logic, CPU only: the library loads, exports every
symbol include/dcns.h declares, refuses to run without a gfx950 device, and
the host-side formulas match their closed forms."""
import math
import os
import re

import numpy as np
import pytest

import dcns_amd
from dcns_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "dcns.h")).read()
    return sorted(set(re.findall(r"\b(dcns_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = dcns_amd.load_library()
    names = _declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(dcns_amd.EXPORTS) == set(names)


def test_default_params_are_cut1gap_defaults():
    p = dcns_amd.default_params()
    assert (p.match, p.mismatch, p.open, p.ext, p.window, p.step, p.min_score, p.w, p.xdrop) == \
        (2, -3, -4, -2, 38, 8, 40, 10, 20)
    assert (p.open2, p.ext2, p.zdrop, p.bandwidth) == (-45, 0, 50, 20000)
    assert (p.pvalue, p.lambda_, p.k) == (0.1, 0.382291, 0.006662)


def test_no_silent_cpu_fallback():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(dcns_amd.DcnsError, match="no HIP device"):
        dcns_amd.Context(0)


def test_pvalue_closed_form():
    """E = K*L1*L2*exp(-lambda*S), p = 1-exp(-E) (findSimilarFragmentsForPairedSequence4.cpp:282-283):
    at the default lambda/K a 100 kb x 100 kb pair reaches p ~ 0.1 between scores 53 and 54."""
    lam, k = 0.382291, 0.006662
    p = lambda s, n: 1 - math.exp(-k * n * n * math.exp(-lam * s))
    assert abs(p(53, 1e5) - 0.1003) < 5e-4 and abs(p(54, 1e5) - 0.0696) < 5e-4


def test_window_schedule_matches_oracle_cells(oracle_lib):
    import bench
    rng = np.random.default_rng(3)
    for _ in range(5):
        a = synth.random_seq(rng, int(rng.integers(50, 400)))
        b = synth.random_seq(rng, int(rng.integers(1, 300)))
        st = {}
        oracle_lib.find_similar(a, b, stats=st)
        # random pairs have no seeds: the scan is exactly the window loop
        if st["seeds"] == 0:
            assert st["scan_cells"] == bench.window_cells([len(a)], [len(b)])


def test_encode_matches_reference_alphabet():
    assert list(synth.encode("ACGTUacgtuNnXR-")) == [0, 1, 3, 4, 4, 0, 1, 3, 4, 4, 2, 2, 2, 2, 2]
