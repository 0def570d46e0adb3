"""The batch command (one process for a GNU-parallel command file of cut1Gap
lines) against the compiled reference's cut1Gap run line by line: every SAM
and o file, stdout and the failure of a missing input; and, on a larger
synthetic set split into small chunks, against this build's own cut1Gap run
line by line."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import dcns_amd
from dcns_amd import synth

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cut1gap_batch")


def _outputs(d):
    return sorted(f for f in os.listdir(d) if ".fa.5" in f and (f.endswith(".sam") or f.endswith("o")))


def test_batch_matches_reference_files(tmp_path):
    for f in os.listdir(GOLD):
        if not f.startswith("expected_"):
            shutil.copy(os.path.join(GOLD, f), tmp_path)
    print("dCNS batch command1", flush=True)
    r = subprocess.run([dcns_amd.CLI_PATH, "batch", "command1"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    status = json.load(open(os.path.join(GOLD, "expected_status.json")))
    assert r.returncode == (1 if any(status) else 0), r.stderr
    assert "error in opening fasta file missing.fa" in r.stderr
    assert r.stdout == open(os.path.join(GOLD, "expected_stdout.txt")).read()
    expected = sorted(f[len("expected_"):] for f in os.listdir(GOLD)
                      if f.startswith("expected_") and ".fa.5" in f)
    assert _outputs(tmp_path) == expected
    for f in expected:
        assert open(tmp_path / f).read() == open(os.path.join(GOLD, "expected_" + f)).read(), f


def _region_set(d, rng, n):
    """n region files over two masked 200 kb genomes, and their command lines."""
    def masked(codes):
        s = np.frombuffer(synth.to_text(codes).encode(), np.uint8).copy()
        for _ in range(60):
            p = int(rng.integers(0, len(s) - 200))
            s[p:p + int(rng.integers(5, 150))] = ord("n")
        return s.tobytes().decode()
    r = synth.random_seq(rng, 200000)
    q = synth.random_seq(rng, 200000)
    lines = []
    for t in range(n):
        rs, rl = int(rng.integers(1, 190000)), int(rng.integers(2000, 8000))
        rows = [("ref", "spA", "c0", int(rng.integers(0, 2)), rs, min(rs + rl, 200000))]
        for k in range(int(rng.integers(1, 5))):
            qs, ql = int(rng.integers(1, 190000)), int(rng.integers(2000, 8000))
            for _ in range(4):  # conserved fragments
                ln = int(rng.integers(50, 200))
                a = rs - 1 + int(rng.integers(0, max(1, rl - ln)))
                b = qs - 1 + int(rng.integers(0, max(1, ql - ln)))
                m = synth.mutate(rng, r[a:a + ln], 0.08, 0.02)[:ln]
                q[b:b + len(m)] = m[:len(q) - b]
            rows.append((f"q{k}", f"sp{chr(66 + k)}", "c0", int(rng.integers(0, 2)), qs, min(qs + ql, 200000)))
        with open(os.path.join(d, f"r{t:03d}.fa"), "w") as fh:
            for name, sp, c, strand, st, en in rows:
                fh.write(f">{name}\tspecies:{sp}\tchr:{c}\tstrand:{strand}\tstart:{st}\tend:{en}\nACGT\n")
        lines.append(["cut1Gap", "-ra", "rg.fa", "-qa", "qg.fa", "-i", f"r{t:03d}.fa", "-r", "ref", "-o",
                      f"r{t:03d}.fa.5"])
    for name, s in (("rg.fa", masked(r)), ("qg.fa", masked(q))):
        with open(os.path.join(d, name), "w") as fh:
            fh.write(">c0\n")
            fh.write("\n".join(s[i:i + 60] for i in range(0, len(s), 60)) + "\n")
        with open(os.path.join(d, name + ".fai"), "w") as fh:
            fh.write(f"c0\t{len(s)}\t4\t60\t61\n")
    return lines


def test_batch_chunks_equal_line_by_line(tmp_path):
    rng = np.random.default_rng(11)
    (tmp_path / "a").mkdir()
    lines = _region_set(str(tmp_path / "a"), rng, 30)
    shutil.copytree(tmp_path / "a", tmp_path / "b")
    with open(tmp_path / "a" / "cmds", "w") as fh:
        fh.write("".join("dCNS " + " ".join(l) + "\n" for l in lines))
    print("dCNS batch cmds -n 8 (30 regions)", flush=True)
    ra = subprocess.run([dcns_amd.CLI_PATH, "batch", "cmds", "-n", "8"], cwd=tmp_path / "a", capture_output=True,
                        text=True, timeout=300)
    assert ra.returncode == 0, ra.stderr
    out_b = ""
    for l in lines:
        rb = subprocess.run([dcns_amd.CLI_PATH] + l, cwd=tmp_path / "b", capture_output=True, text=True, timeout=120)
        assert rb.returncode == 0, rb.stderr
        out_b += rb.stdout
    assert ra.stdout == out_b
    fa, fb = _outputs(tmp_path / "a"), _outputs(tmp_path / "b")
    assert fa == fb and len(fa) > 30
    nonempty = 0
    for f in fa:
        ta = open(tmp_path / "a" / f).read()
        assert ta == open(tmp_path / "b" / f).read(), f
        nonempty += f.endswith(".sam") and len(ta) > 0
    assert nonempty > 5


def test_batch_bad_region_and_truncated_genome_fail_alone(tmp_path):
    """One line's region file has a malformed header (its own process would
    die on it) and one line's region runs past the end of a truncated genome
    file (the index still gives the full length): the malformed line fails
    alone, every other line's files and stdout equal this build's cut1Gap run
    line by line, one process each (fresh genome streams per line)."""
    rng = np.random.default_rng(23)
    (tmp_path / "a").mkdir()
    lines = _region_set(str(tmp_path / "a"), rng, 10)
    a = tmp_path / "a"
    with open(a / "r003.fa", "w") as fh:  # malformed start coordinate
        fh.write(">ref\tspecies:spA\tchr:c0\tstrand:1\tstart:x12\tend:3000\nACGT\n"
                 ">q0\tspecies:spB\tchr:c0\tstrand:1\tstart:1000\tend:4000\nACGT\n")
    # the query genome loses its last 9 kb of text, its index does not
    qg = (a / "qg.fa").read_text()
    (a / "qg.fa").write_text(qg[:len(qg) - 9000])
    with open(a / "r006.fa", "w") as fh:  # a query region inside the lost text
        fh.write(">ref\tspecies:spA\tchr:c0\tstrand:1\tstart:5000\tend:9000\nACGT\n"
                 ">q0\tspecies:spB\tchr:c0\tstrand:0\tstart:195000\tend:199000\nACGT\n"
                 ">q1\tspecies:spC\tchr:c0\tstrand:1\tstart:2000\tend:6000\nACGT\n")
    shutil.copytree(a, tmp_path / "b")
    with open(a / "cmds", "w") as fh:
        fh.write("".join("dCNS " + " ".join(l) + "\n" for l in lines))
    print("dCNS batch with a malformed region and a truncated genome", flush=True)
    ra = subprocess.run([dcns_amd.CLI_PATH, "batch", "cmds", "-n", "6"], cwd=a, capture_output=True, text=True,
                        timeout=300)
    assert ra.returncode == 1, ra.stderr
    assert "r003.fa" in ra.stderr and "stoi" in ra.stderr
    out_b, failed = "", []
    for l in lines:
        rb = subprocess.run([dcns_amd.CLI_PATH] + l, cwd=tmp_path / "b", capture_output=True, text=True, timeout=120)
        if rb.returncode != 0:
            failed.append(l[6])
        out_b += rb.stdout
    assert failed == ["r003.fa"]
    assert ra.stdout == out_b
    fa, fb = _outputs(a), _outputs(tmp_path / "b")
    assert fa == fb and not [f for f in fa if f.startswith("r003")]
    assert [f for f in fa if f.startswith("r006")] and [f for f in fa if f.startswith("r009")]
    for f in fa:
        assert open(a / f).read() == open(tmp_path / "b" / f).read(), f
