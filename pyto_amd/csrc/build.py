"""Build libpytohip.so for gfx950 in-tree (pyto_amd/libpytohip.so).

Usage: python -m pyto_amd.csrc.build  (or call ``build()``).  Each .hip source
is compiled to an object in pyto_amd/csrc/_build/ in parallel, then linked.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
REPO = os.path.dirname(PKG)
SOURCES = ["operator.hip", "coarse.hip", "solver.hip", "postproc.hip", "optimizer.hip", "batch.hip",
           "thermo.hip", "mma.hip", "builders.hip", "topology.hip"]
LIB = os.path.join(PKG, "libpytohip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PYTO_ARCH", "gfx950")
ROCM_LIB = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")

FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    "-munsafe-fp-atomics",
    # no packed FP32 math: keep every FP32 op a plain VALU op
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(REPO, "include"),
    "-Wno-unused-command-line-argument",
]


def _compile(src):
    out = os.path.join(HERE, "_build", src.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(HERE, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, r.returncode, r.stdout + r.stderr, " ".join(cmd)


def build(verbose=False):
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(4, os.cpu_count() or 1)) as ex:
        results = list(ex.map(_compile, SOURCES))
    for src, rc, log, cmd in results:
        if rc != 0:
            raise RuntimeError("hipcc failed on %s:\n%s\n%s" % (src, cmd, log))
        if verbose and log.strip():
            print(log)
    objs = [os.path.join(HERE, "_build", s.replace(".hip", ".o")) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-Wl,--no-undefined", *objs,
           "-L" + ROCM_LIB, "-Wl,-rpath," + ROCM_LIB, "-lrocsolver", "-lrocblas", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stdout + r.stderr))
    shutil.move(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
