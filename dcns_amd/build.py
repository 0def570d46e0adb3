"""Builds the native library (HIP kernels + C-ABI + C++ host layer) for gfx950.

    python -m dcns_amd.build          # -> dcns_amd/lib/libdcns.so, dcns_amd/bin/dCNS

Every translation unit is compiled with hipcc for gfx950 only; objects go to
dcns_amd/_build/.  Packed FP32 VALU forms are disabled for all device code.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
BINDIR = os.path.join(PKG, "bin")
LIB = os.path.join(LIBDIR, "libdcns.so")
CLI = os.path.join(BINDIR, "dCNS")

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-Wno-unused-command-line-argument", f"-I{os.path.join(ROOT, 'include')}",
          f"-I{CSRC}"]

LIB_SOURCES = ["seedscan.hip", "xdrop.hip", "swbatch.hip", "pipeline.cpp", "capi.cpp", "host/cli_io.cpp",
               "host/commands.cpp", "host/lambdak.cpp", "mlp.hip",
               "semiglobal.hip", "twogaps.cpp", "host/weights.cpp", "mask.hip", "maskgenome.cpp",
               "host/mask_io.cpp", "host/mask_cmd.cpp"]
CLI_SOURCES = ["cli/main.cpp"]


def _obj(src):
    return os.path.join(BUILD, src.replace("/", "_") + ".o")


def _compile(src):
    path = os.path.join(CSRC, src)
    out = _obj(src)
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC] + COMMON + lang + ["-c", path, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(BINDIR, exist_ok=True)
    sources = [s for s in LIB_SOURCES + CLI_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = dict(zip(sources, ex.map(_compile, sources)))
    lib_objs = [objs[s] for s in LIB_SOURCES if s in objs]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + lib_objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    cli_objs = [objs[s] for s in CLI_SOURCES if s in objs]
    if cli_objs:
        tmp = CLI + ".tmp"
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-o", tmp] + cli_objs + [f"-L{LIBDIR}", "-ldcns",
                                                                         f"-Wl,-rpath,$ORIGIN/../lib", "-lpthread"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(tmp, CLI)
    if verbose:
        print(f"built {LIB}" + (f" and {CLI}" if cli_objs else ""))
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
