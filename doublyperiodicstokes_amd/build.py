"""Build the gfx950 shared library ``lib/libdpstokes_hip.so`` with hipcc.

Usage: ``python -m doublyperiodicstokes_amd.build`` (or ``build()`` from
Python).  Every source is compiled for gfx950 only; packed-FP32 instructions
are disabled for all device code.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIBNAME = "libdpstokes_hip.so"
SOURCES = ["dps_api.hip", "dps_kernels.hip", "dps_spectral.hip", "dps_fft.hip", "dps_zgemm.hip"]
HEADERS = ["dps_core.h", "dps_fft.h", "dps_internal.h", "dps_mma.h", "dps_seg.h", "dps_tables.h"]
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def lib_path():
    return os.path.join(LIBDIR, LIBNAME)


def _hipcc():
    p = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    if not os.path.exists(p):
        raise RuntimeError("hipcc not found; the HIP toolchain is required to build")
    return p


def build(verbose=False):
    """Compile every source to an object, then link the shared library."""
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "_obj")
    os.makedirs(objdir, exist_ok=True)
    flags = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics",
             "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [_hipcc()] + flags + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    errors = []
    for cmd, pr in procs:
        out, _ = pr.communicate()
        text = out.decode(errors="replace")
        if pr.returncode != 0:
            errors.append(text)
        elif verbose and text.strip():
            print(text)
    if errors:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errors))
    tmp = lib_path() + ".tmp"
    link = [_hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + \
           ["-L" + os.path.join(ROCM, "lib"), "-lrocfft", "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    if verbose:
        print(" ".join(link), flush=True)
    r = subprocess.run(link, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout.decode(errors="replace"))
    os.replace(tmp, lib_path())
    return lib_path()


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
