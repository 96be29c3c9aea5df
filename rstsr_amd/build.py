"""Build the gfx950 device library, the C++ host front end and the CPU oracle, in tree.

    python rstsr_amd/build.py            # everything
    python rstsr_amd/build.py --lib      # librstsr_hip.so only

(run by path: the package itself refuses to import until the native modules exist)

Outputs:
    rstsr_amd/librstsr_hip.so            C ABI of include/rstsr_hip.h (hipcc, --offload-arch=gfx950)
    rstsr_amd/_rstsr<ext>.so             pybind11 module over the C++ front end (links the library)
    oracle/liboracle.so                  CPU restatement used only by tests and benchmarks
    tests/gpu_check/libdot2check.so      GPU Dot2 GEMM checker used only by tests and benchmarks
    tests/gpu_check/libstreamcheck.so    HBM stream-ceiling probe used only by the benchmark
"""
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "rstsr_amd")
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
LIB = os.path.join(PKG, "librstsr_hip.so")

# device sources and extra flags; elementwise/reduce keep IEEE operation order
DEVICE_SOURCES = [
    ("api.cpp", []),
    ("elementwise.hip", ["-ffp-contract=off"]),
    ("reduce.hip", ["-ffp-contract=off"]),
    ("unary.hip", ["-ffp-contract=off"]),
    ("gemm.hip", []),
    ("ozaki.hip", []),
    ("blas.hip", []),
]

HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed:\n  " + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _newer(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def _headers():
    hs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))]
    hs.append(os.path.join(ROOT, "include", "rstsr_hip.h"))
    return hs


def build_lib(force=True, jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    hdrs = _headers()
    for src, extra in DEVICE_SOURCES:
        sp = os.path.join(CSRC, src)
        if not os.path.exists(sp):
            continue
        op = os.path.join(BUILD, src + ".o")
        objs.append(op)
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        if force or _newer(op, [sp] + hdrs):
            cmds.append([HIPCC] + HIP_FLAGS + extra + lang + ["-c", sp, "-o", op])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lamdhip64"])
    os.replace(tmp, LIB)
    return LIB


def build_front(force=True):
    import pybind11

    ext = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(PKG, "_rstsr" + ext)
    src = os.path.join(CSRC, "front", "pymodule.cpp")
    deps = [src] + _headers() + [os.path.join(CSRC, "front", f) for f in os.listdir(os.path.join(CSRC, "front"))]
    if not force and not _newer(out, deps + [LIB]):
        return out
    py_inc = sysconfig.get_paths()["include"]
    tmp = out + ".tmp"
    _run([
        "g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden",
        "-I", pybind11.get_include(), "-I", py_inc, "-I", os.path.join(ROOT, "include"), "-I", CSRC,
        src, "-o", tmp, "-L", PKG, "-lrstsr_hip", "-Wl,-rpath,$ORIGIN",
    ])
    os.replace(tmp, out)
    return out


CHECKER_SRC = os.path.join(ROOT, "tests", "gpu_check", "dot2_check.hip")
CHECKER_LIB = os.path.join(ROOT, "tests", "gpu_check", "libdot2check.so")
STREAM_SRC = os.path.join(ROOT, "tests", "gpu_check", "stream_ceiling.hip")
STREAM_LIB = os.path.join(ROOT, "tests", "gpu_check", "libstreamcheck.so")


def build_checker(force=True):
    """GPU Dot2 GEMM checker and HBM stream-ceiling probe (test infrastructure for the
    GEMM error and roofline measurements)."""
    for src, out, extra in ((CHECKER_SRC, CHECKER_LIB, ["-ffp-contract=off"]), (STREAM_SRC, STREAM_LIB, [])):
        if not os.path.exists(src) or (not force and not _newer(out, [src])):
            continue
        tmp = out + ".tmp"
        _run([HIPCC] + HIP_FLAGS + extra + ["-shared", "-o", tmp, src])
        os.replace(tmp, out)
    return CHECKER_LIB


def build_oracle(force=True):
    return _run(["make", "-C", os.path.join(ROOT, "oracle")] + (["-B"] if force else []))


def build_all(force=True):
    build_lib(force=force)
    build_front(force=force)
    build_oracle(force=force)
    build_checker(force=force)


if __name__ == "__main__":
    args = sys.argv[1:]
    if "--lib" in args:
        print(build_lib())
    elif "--front" in args:
        print(build_front())
    elif "--oracle" in args:
        print(build_oracle())
    elif "--checker" in args:
        print(build_checker())
    else:
        build_all(force="--incremental" not in args)
        print("built", LIB)
