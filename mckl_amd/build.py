"""Build step for libmckl_hip.so (gfx950) and the C++ host-API test program.

Run as ``python -m mckl_amd.build`` or through ``__graft_entry__.build()``.
Every HIP source is compiled with -ffp-contract=off (each floating-point
operation rounds once, as in the reference build) and with packed-FP32
instruction selection disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmckl_hip.so")

HIP_SOURCES = ["rng.hip", "smc.hip", "primitives.hip", "smc_grad.hip",
               "resample_index.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def hipcc():
    cand = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")
    return cand if os.path.exists(cand) else "hipcc"


def build_lib(jobs=4):
    objs = []
    procs = []
    for src in HIP_SOURCES:
        s = os.path.join(CSRC, src)
        if not os.path.exists(s):
            continue
        o = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc()] + HIPCC_FLAGS + ["-c", s, "-o", o]
        print("+", " ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd), s))
        objs.append(o)
    for p, s in procs:
        if p.wait() != 0:
            raise RuntimeError("hipcc failed on %s" % s)
    tmp = LIB + ".tmp"
    _run([hipcc()] + HIPCC_FLAGS + ["-shared", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


def build_host_test():
    """C++ program exercising the host-side API mirror (tests/cpp)."""
    src = os.path.join(ROOT, "tests", "cpp", "host_api_test.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(ROOT, "tests", "cpp", "host_api_test")
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    _run([
        "g++", "-std=c++14", "-O2", "-Wall",
        "-D__HIP_PLATFORM_AMD__",
        "-I", os.path.join(ROOT, "include"),
        "-I", os.path.join(HERE, "include"),
        "-I", os.path.join(rocm, "include"),
        src, "-o", out,
        "-L", HERE, "-lmckl_hip", "-Wl,-rpath,$ORIGIN/../../mckl_amd",
        "-L", os.path.join(rocm, "lib"), "-lamdhip64",
        "-Wl,-rpath," + os.path.join(rocm, "lib"),
    ])
    return out


def build_pf_program():
    """The reference's particle-filter example on the device API
    (tests/cpp/algorithm_pf.hip), a hipcc program linked to the library."""
    src = os.path.join(ROOT, "tests", "cpp", "algorithm_pf.hip")
    if not os.path.exists(src):
        return None
    out = os.path.join(ROOT, "tests", "cpp", "algorithm_pf")
    _run([hipcc()] + [f for f in HIPCC_FLAGS if f != "-fPIC"] + [
        "-I", os.path.join(ROOT, "include"),
        "-I", os.path.join(HERE, "include"),
        src, "-o", out,
        "-L", HERE, "-lmckl_hip", "-Wl,-rpath,$ORIGIN/../../mckl_amd",
    ])
    return out


def build_fused_test():
    """Caller-written models on the fused samplers through
    mckl_amd/include/mckl_hip_fused.hpp (tests/cpp/fused_model_test.hip), a
    hipcc shared library linked to libmckl_hip.so."""
    src = os.path.join(ROOT, "tests", "cpp", "fused_model_test.hip")
    if not os.path.exists(src):
        return None
    out = os.path.join(ROOT, "tests", "cpp", "libfused_model_test.so")
    _run([hipcc()] + HIPCC_FLAGS + [
        "-shared",
        "-I", os.path.join(ROOT, "include"),
        "-I", os.path.join(HERE, "include"),
        src, "-o", out,
        "-L", HERE, "-lmckl_hip", "-Wl,-rpath,$ORIGIN/../../mckl_amd",
    ])
    return out


def build_crmath_host():
    """Host builds of device headers for the CPU tests (tests/cpp):
    csrc/crmath.hpp (correctly rounded log/sin/cos) and csrc/smc_index.hpp
    (the resampling kernels' index arithmetic)."""
    outs = []
    for name in ("crmath_host", "rs_guard_host"):
        src = os.path.join(ROOT, "tests", "cpp", name + ".cpp")
        if not os.path.exists(src):
            continue
        out = os.path.join(ROOT, "tests", "cpp", "lib%s.so" % name)
        _run(["g++", "-std=c++17", "-O2", "-ffp-contract=off",
              "-fno-fast-math", "-fPIC", "-shared", "-Wall", "-o", out, src])
        outs.append(out)
    return outs


def main():
    build_lib()
    build_host_test()
    build_pf_program()
    build_fused_test()
    build_crmath_host()


if __name__ == "__main__":
    sys.exit(main())
