"""Build of libhp3d_gpu.so (HIP, gfx950 only) in the package directory, and of
the Fortran binding (fortran/hp3d_gpu_mod.F90) with its driver program
fortran/elem_driver.bin when amdflang is available."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "hp3d_gpu.hip")
OUT = os.path.join(HERE, "libhp3d_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("FC_HP3D", "/opt/rocm/bin/amdflang")
FDIR = os.path.join(HERE, "fortran")
FDRIVER = os.path.join(FDIR, "elem_driver.bin")
ROCM_LIB = "/opt/rocm/lib"

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 VALU ops are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-value",
]


def build(verbose=False):
    tmp = OUT + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp, SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    if verbose:
        print(r.stderr)
    os.replace(tmp, OUT)
    build_fortran()
    return OUT


def build_fortran():
    """Module + driver, linked against libhp3d_gpu.so (rpath to this directory)."""
    if not os.path.exists(FC):
        return None
    mod = os.path.join(FDIR, "_mod")
    os.makedirs(mod, exist_ok=True)
    cmd = [FC, "-O2", "-J", mod, "-o", FDRIVER + ".tmp",
           os.path.join(FDIR, "hp3d_gpu_mod.F90"), os.path.join(FDIR, "elem_driver.F90"),
           "-L" + HERE, "-lhp3d_gpu", "-Wl,-rpath,$ORIGIN/..",
           # the batched modes stage device memory with the HIP runtime
           "-L" + ROCM_LIB, "-lamdhip64", "-Wl,-rpath," + ROCM_LIB]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("amdflang failed:\n" + r.stdout + r.stderr)
    os.replace(FDRIVER + ".tmp", FDRIVER)
    return FDRIVER


if __name__ == "__main__":
    print(build(verbose=True))
