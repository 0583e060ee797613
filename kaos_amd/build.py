"""Build libkaos_hip.so in-tree for gfx950 with hipcc.

Flags: -ffp-contract=off so the FP64 path evaluates the same IEEE operations
as the host build of the shared arithmetic; packed FP32 ops disabled in the
device code generator.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build_lib(verbose=False, out=None, defines=()):
    """libkaos_hip.so (or `out`), with optional -D compile-time variants."""
    src = os.path.join(HERE, "csrc", "kaos_hip.hip")
    out = out or os.path.join(HERE, "libkaos_hip.so")
    tmp = out + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-O3", "-fPIC", "-shared", "-std=c++17",
           "-ffp-contract=off", *DEVICE_FLAGS, "-Wall", "-Wno-unused-function",
           *(f"-D{d}" for d in defines),
           "-I", os.path.join(HERE, "..", "include"), src, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stderr)
    if verbose and r.stderr:
        print(r.stderr)
    os.replace(tmp, out)
    return out


FLANG = "/opt/rocm/lib/llvm/bin/flang"


def build_fortran(verbose=False):
    """The iso_c_binding module and the Fortran driver kaos_fdriver, linked
    against libkaos_hip.so in the package directory (rpath $ORIGIN/..)."""
    fdir = os.path.join(HERE, "fortran")
    obj = os.path.join(fdir, "_build")
    os.makedirs(obj, exist_ok=True)
    out = os.path.join(fdir, "kaos_fdriver")
    tmp = out + ".tmp"
    cmd = [FLANG, "-O2", "-J", obj, os.path.join(fdir, "kaos_hip_iface.f90"),
           os.path.join(fdir, "kaos_fdriver.f90"), "-L", HERE, "-lkaos_hip",
           "-Wl,-rpath,$ORIGIN/..", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=obj)
    if r.returncode != 0:
        raise RuntimeError("flang failed:\n" + r.stderr)
    if verbose and r.stderr:
        print(r.stderr)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build_lib(verbose=True))
