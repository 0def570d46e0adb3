"""The drop-in dCNS command against the reference's own command output
(golden files from the compiled reference; live where oracle/_ref exists)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import dcns_amd
from dcns_amd import synth

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _run(args, cwd):
    print("dCNS", " ".join(args), flush=True)
    r = subprocess.run([dcns_amd.CLI_PATH] + args, cwd=cwd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r.stdout


@pytest.mark.parametrize("inp,out", [("in.fa", "out"), ("in_refminus.fa", "outm")])
def test_cut1gap_matches_reference_files(tmp_path, inp, out):
    src = os.path.join(GOLD, "cut1gap")
    for f in os.listdir(src):
        if not f.startswith("expected_"):
            shutil.copy(os.path.join(src, f), tmp_path)
    stdout = _run(["cut1Gap", "-i", inp, "-r", "ref", "-ra", "refgenome.fa", "-qa", "querygenome.fa", "-o", out],
                  tmp_path)
    assert stdout == open(os.path.join(src, f"expected_stdout_{out}.txt")).read()
    # species names all start with "sp", so out + "sp" separates the two prefixes
    expected = sorted(f[len("expected_"):] for f in os.listdir(src) if f.startswith("expected_" + out + "sp"))
    produced = sorted(f for f in os.listdir(tmp_path) if f.startswith(out + "sp"))
    assert produced == expected and expected
    for f in expected:
        assert open(os.path.join(tmp_path, f)).read() == open(os.path.join(src, "expected_" + f)).read(), f


@pytest.mark.parametrize("tag,args", [
    ("l_300_p_200_s_5", ["-l", "300", "-p", "200", "-s", "5"]),
    ("l_500_p_100_s_11_A_1_B_2_O_3_E_1", ["-l", "500", "-p", "100", "-s", "11", "-A", "1", "-B", "-2", "-O", "-3",
                                          "-E", "-1"]),
])
def test_ransco_matches_reference(tmp_path, tag, args):
    src = os.path.join(GOLD, "ransco")
    for f in ("refgenome.fa", "q1.fa", "q2.fa"):
        shutil.copy(os.path.join(src, f), tmp_path)
    stdout = _run(["ranSco", "-r", "refgenome.fa", "-i", "q1.fa", "q2.fa"] + args, tmp_path)
    assert stdout == open(os.path.join(src, f"expected_{tag}.txt")).read()


def test_cut1gap_live_against_reference_cli(tmp_path):
    """Larger regions with masked runs, both strands, against the compiled
    reference command (skipped where oracle/_ref is absent)."""
    from oracle import bind
    if not os.path.exists(bind.REF_CLI):
        pytest.skip("oracle/_ref not built")
    import importlib.util
    spec = importlib.util.spec_from_file_location("mg", os.path.join(GOLD, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    rng = np.random.default_rng(77)
    for d in ("ours", "ref"):
        os.makedirs(tmp_path / d)
    mg.write_cut1gap_inputs(str(tmp_path / "ours"), rng)
    for f in os.listdir(tmp_path / "ours"):
        shutil.copy(tmp_path / "ours" / f, tmp_path / "ref")
    args = ["cut1Gap", "-i", "in.fa", "-r", "ref", "-ra", "refgenome.fa", "-qa", "querygenome.fa", "-o", "o_"]
    a = _run(args, tmp_path / "ours")
    b = subprocess.run([bind.REF_CLI] + args, cwd=tmp_path / "ref", capture_output=True, text=True, timeout=300).stdout
    assert a == b
    outs = sorted(f for f in os.listdir(tmp_path / "ref") if f.startswith("o_"))
    assert outs == sorted(f for f in os.listdir(tmp_path / "ours") if f.startswith("o_"))
    for f in outs:
        assert open(tmp_path / "ours" / f).read() == open(tmp_path / "ref" / f).read(), f
