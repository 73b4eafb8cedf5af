"""Builds libbigdft_hip.so (HIP, gfx950) in-tree with hipcc."""
import os
import subprocess

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libbigdft_hip.so")
SOURCES = ["bdft_api.hip", "bdft_pass.hip", "bdft_plane.hip", "bdft_mfma.hip",
           "bdft_xfer.hip", "bdft_bounds.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    "-Wno-unused-value",
    # keep hipcc from emitting packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose=False):
    """Compile every HIP source into lib/libbigdft_hip.so (always rebuilds)."""
    os.makedirs(LIBDIR, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB + ".tmp"
    cmd = [hipcc] + HIPCC_FLAGS + ["-o", tmp] + \
        [os.path.join(CSRC, s) for s in SOURCES]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stderr[-6000:])
    if verbose and r.stderr:
        print(r.stderr)
    os.replace(tmp, LIB)
    return LIB


FORTRAN = os.path.join(PKG, "fortran")
FDRIVER = os.path.join(FORTRAN, "locham_hip_driver")


def build_fortran():
    """Compile the iso_c_binding module and its driver with amdflang (when
    the compiler is present); returns the driver path or None."""
    fc = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")
    if not os.path.exists(fc):
        return None
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        mod = os.path.join(tmp, "module_locham_hip.o")
        r = subprocess.run([fc, "-O2", "-module-dir", tmp, "-c",
                            os.path.join(FORTRAN, "module_locham_hip.f90"),
                            "-o", mod], capture_output=True, text=True)
        if r.returncode == 0:
            r = subprocess.run([fc, "-O2", "-I", tmp,
                                os.path.join(FORTRAN, "locham_hip_driver.f90"),
                                mod, "-L" + LIBDIR, "-lbigdft_hip",
                                "-Wl,-rpath,$ORIGIN/../lib", "-o", FDRIVER],
                               capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("amdflang failed:\n" + r.stderr[-4000:])
    return FDRIVER


if __name__ == "__main__":
    print(build(verbose=True))
