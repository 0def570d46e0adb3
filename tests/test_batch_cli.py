"""The batch command's handling of bad lines, on the CPU (no line reaches
the device): a missing genome index fails its group and the run goes on, an
unknown command is reported, comments and the program name are skipped."""
import os
import subprocess

import dcns_amd


def _batch(tmp_path, text):
    (tmp_path / "cmds").write_text(text)
    return subprocess.run([dcns_amd.CLI_PATH, "batch", "cmds"], cwd=tmp_path, capture_output=True, text=True,
                          timeout=60)


def test_batch_usage(tmp_path):
    r = subprocess.run([dcns_amd.CLI_PATH, "batch"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Usage: dCNS batch" in r.stderr
    r = subprocess.run([dcns_amd.CLI_PATH, "batch", "missing_file"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "error in opening command file" in r.stderr


def test_batch_bad_lines_fail_alone(tmp_path):
    r = _batch(tmp_path, "# comment line\n"
                         "dCNS cut1Gap -ra nope.fa -qa nope2.fa -i x.fa -r ref -o out\n"
                         "/opt/bin/dCNS notACommand -x 1\n"
                         "\n"
                         "cut1Gap -ra nope.fa -qa nope2.fa -i y.fa -r ref -o out2 -c 30\n")
    assert r.returncode == 1
    assert r.stderr.count("error in opening fasta index file nope.fa.fai.") == 2  # two groups (-c 30 differs)
    assert "command notACommand is not part of this build" in r.stderr
    assert not [f for f in os.listdir(tmp_path) if f.startswith("out")]


def test_cut1gap_malformed_region_header_fails_cleanly(tmp_path):
    """A region header whose start is not a number: the line reports the
    error and fails with status 1 (no exception escapes the C boundary, no
    abort), before any device work or output file."""
    (tmp_path / "g.fa").write_text(">c0\n" + "ACGT" * 50 + "\n")
    (tmp_path / "g.fa.fai").write_text("c0\t200\t4\t200\t201\n")
    (tmp_path / "bad.fa").write_text(">ref\tspecies:spA\tchr:c0\tstrand:1\tstart:1x\tend:100\nACGT\n"
                                     ">q\tspecies:spB\tchr:c0\tstrand:1\tstart:abc\tend:90\nACGT\n")
    r = subprocess.run([dcns_amd.CLI_PATH, "cut1Gap", "-ra", "g.fa", "-qa", "g.fa", "-i", "bad.fa", "-r", "ref",
                        "-o", "out"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1, (r.returncode, r.stderr)
    assert "dCNS: cut1Gap:" in r.stderr and "stoi" in r.stderr
    assert not [f for f in os.listdir(tmp_path) if f.startswith("out")]
