"""The reference-side binding INTEGRATION.md gives for overload 1 (the body a
maintainer puts in place of impl/findSimilarFragmentsForPairedSequence4.cpp:
433-524) compiles against the reference's own headers and links against
libdcns.so and the reference's PairedSimilarFragment and Scorei; the linked program
then reaches the library (here, with no GPU, its context creation fails with
the library's message and the stub's exit status).  Runs where the
reference tree and its build (oracle/_ref) are present."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
PSF = os.path.join(ROOT, "oracle", "_ref", "obj", "model", "PairedSimilarFragment.o")
SCORE = os.path.join(ROOT, "oracle", "_ref", "obj", "model", "Score.o")

MAIN = r'''
#include <iostream>
#include "findSimilarFragmentsForPairedSequence.h"
int main() {
    int8_t a[64], b[64];
    for (int i = 0; i < 64; ++i) a[i] = b[i] = (int8_t)((i * 7) % 5 == 2 ? 0 : (i * 7) % 5);
    int32_t l1 = 64, l2 = 64, win = 38, step = 8, w = 10, xd = 20;
    int32_t mini = 40, bound = 20, o = -4, e = -2, ma = 2, mi = -3;
    Scorei m(ma, mi);
    std::string s1, s2;
    std::vector<PairedSimilarFragment> f = findSimilarFragmentsForPairedSequence(
        a, a, b, b, l1, l2, win, mini, bound, o, e, ma, mi, m, step, s1, s2, 0.1, 0.382291, 0.006662, w, xd);
    std::cout << "fragments " << f.size() << std::endl;
    return 0;
}
'''


def _stub():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    sec = text[text.index("### Overload 1"):]
    return re.search(r"```cpp\n(.*?)```", sec, re.S).group(1)


@pytest.mark.skipif(not (os.path.isdir(REF) and os.path.exists(PSF)), reason="reference tree or oracle/_ref absent")
def test_overload1_stub_compiles_links_and_reaches_the_library(tmp_path):
    (tmp_path / "stub.cpp").write_text("#include <iostream>\n" + _stub())
    (tmp_path / "main.cpp").write_text(MAIN)
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + REF, "-I" + os.path.join(REF, "impl"),
           "-I" + os.path.join(REF, "model")]
    lib = os.path.join(ROOT, "dcns_amd", "lib")
    cmd = ["g++", "-std=gnu++14", "-O1", "-w"] + inc + [str(tmp_path / "stub.cpp"), str(tmp_path / "main.cpp"), PSF, SCORE,
                                                         "-L" + lib, "-ldcns", "-Wl,-rpath," + lib,
                                                         "-o", str(tmp_path / "prog")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([str(tmp_path / "prog")], capture_output=True, text=True, timeout=120)
    # no GPU in this container: the stub's context creation reports the library's error
    if r.returncode == 0:
        assert "seeds size" in r.stdout and "fragments" in r.stdout
    else:
        assert r.returncode == 1 and ("HIP" in r.stderr or "gfx950" in r.stderr or "device" in r.stderr), r.stderr
