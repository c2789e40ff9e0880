"""Build the native library in-tree.

hipcc compiles the HIP sources for gfx950 into
quantumvortices_amd/lib/libvikshobha.so (device code + C ABI).  Packed FP32
instructions are disabled for all device code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib", "libvikshobha.so")
SRC = [os.path.join(HERE, "csrc", f) for f in ("vk_gpe.hip", "vk_fno.hip", "vk_io.hip", "vk_sts.hip", "vk_patterns.hip")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-value", "-Wno-unused-result",
    "-I" + os.path.join(ROOT, "include"),
]


FORTRAN = os.environ.get("FC", "/opt/rocm/bin/amdflang")
FMOD = os.path.join(HERE, "fortran", "vikshobha_mod.f90")
FDRIVERS = {name: os.path.join(HERE, "fortran", name + ".f90") for name in ("driver_check", "driver_part")}
FBIN = os.path.join(HERE, "lib", "driver_check")


def build_fortran(verbose=False):
    """ISO_C_BINDING layer + the reference-style check drivers (amdflang)."""
    import tempfile
    with tempfile.TemporaryDirectory() as mod:
        for name, src in FDRIVERS.items():
            cmd = [FORTRAN, "-O2", "-J", mod, "-o", os.path.join(HERE, "lib", name), FMOD, src,
                   "-L" + os.path.dirname(LIB), "-lvikshobha", "-Wl,-rpath,$ORIGIN"]
            if verbose:
                print(" ".join(cmd), flush=True)
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError("amdflang failed:\n" + r.stderr)
    return FBIN


def build(verbose=False):
    """Compile every source to an object in parallel (each translation unit
    holds its own kernels), then link the shared library."""
    import tempfile
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + ".tmp"
    cflags = [f for f in FLAGS if f != "-shared"]
    with tempfile.TemporaryDirectory() as td:
        jobs = []
        for src in SRC:
            obj = os.path.join(td, os.path.basename(src) + ".o")
            cmd = [HIPCC] + cflags + ["-c", "-o", obj, src]
            if verbose:
                print(" ".join(cmd), flush=True)
            jobs.append((obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        errs = []
        for obj, pr in jobs:
            _, e = pr.communicate()
            if pr.returncode != 0:
                errs.append("\n".join(l for l in e.splitlines() if "packed-fp32-ops" not in l))
        if errs:
            raise RuntimeError("hipcc failed:\n" + "\n".join(errs))
        cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + [o for o, _ in jobs]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, LIB)
    if os.path.exists(FORTRAN):
        build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
