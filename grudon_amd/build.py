"""Build recipe for the gfx950 library, the C++ drop-in host and the oracle.

    python -m grudon_amd.build            # everything
    python -m grudon_amd.build --no-oracle
    python -m grudon_amd.build --sanitize # host code under ASan + UBSan, then the CPU tests

Outputs (in-tree, git-ignored, shipped with the working tree):
    grudon_amd/libgrudon_hip.so      HIP kernels + C-ABI (include/grudon_hip.h)
    grudon_amd/bin/grudon            CLI (-g -c -m -t -p -a), reference-compatible
    grudon_amd/libgrudon_probe.so    measurement probes used by bench.py only
    oracle/liboracle.so              CPU oracle (test infrastructure only)
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
BUILD = os.path.join(ROOT, "_build")
ARCH = os.environ.get("GRUDON_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# -packed-fp32-ops: no packed FP32 instructions are generated (they have
# given wrong results on these machines); -munsafe-fp-atomics: hardware
# float/double atomic add instead of CAS loops.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH, "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]
HOST_CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result"]


def _run(cmd, cwd=ROOT):
    p = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("build step failed (%d): %s\n%s" % (p.returncode, " ".join(cmd), p.stdout))
    return p.stdout


def _compile_hip(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    _run([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj])
    return obj


def _compile_host_cpp(src):
    # host-only translation units of the library still include the HIP runtime API
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    _run([HIPCC] + HOST_CXX_FLAGS + ["-c", src, "-o", obj])
    return obj


def build_library(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    hip_srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    cpp_srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".cpp"))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile_hip, s) for s in hip_srcs] + [ex.submit(_compile_host_cpp, s) for s in cpp_srcs]
        objs = [f.result() for f in futs]
    out = os.path.join(PKG, "libgrudon_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, out)
    return out


def build_probe():
    """Measurement-only library for bench.py (random-gather ceiling, streaming
    read of known size); not part of libgrudon_hip.so."""
    os.makedirs(BUILD, exist_ok=True)
    src = os.path.join(PKG, "measure", "probe.hip")
    out = os.path.join(PKG, "libgrudon_probe.so")
    _run([HIPCC] + HIP_FLAGS + ["-shared", src, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_host():
    """C++ drop-in operator API + CLI, linked against libgrudon_hip.so."""
    if not os.path.isdir(HOST):
        return None
    srcs = sorted(os.path.join(d, f) for d, _, fs in os.walk(HOST) for f in fs if f.endswith(".cpp"))
    if not srcs:
        return None
    bindir = os.path.join(PKG, "bin")
    os.makedirs(bindir, exist_ok=True)
    exe = os.path.join(bindir, "grudon")
    cxx = shutil.which("g++") or "g++"
    _run([cxx] + HOST_CXX_FLAGS + ["-I", os.path.join(ROOT, "include"), "-I", HOST] + srcs +
         ["-o", exe + ".tmp", "-L", PKG, "-lgrudon_hip", "-Wl,-rpath,$ORIGIN/..", "-pthread"])
    os.replace(exe + ".tmp", exe)
    return exe


def build_oracle():
    odir = os.path.join(ROOT, "oracle")
    out = os.path.join(odir, "liboracle.so")
    cc = shutil.which("gcc") or "gcc"
    _run([cc, "-O2", "-std=c11", "-Wall", "-Wextra", "-fPIC", "-shared", "-pthread",
          os.path.join(odir, "grudon_oracle.c"), "-o", out + ".tmp", "-lm"])
    os.replace(out + ".tmp", out)
    return out


CLANG = "/opt/rocm/llvm/bin/clang"
CLANGXX = "/opt/rocm/llvm/bin/clang++"
SAN = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined", "-g", "-O1"]


def asan_runtime():
    out = subprocess.run([CLANG, "-print-file-name=libclang_rt.asan-x86_64.so"], stdout=subprocess.PIPE,
                         text=True).stdout.strip()
    if not os.path.exists(out):
        import glob
        hits = glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")
        out = hits[0] if hits else ""
    return out


def build_sanitized(jobs=8):
    """Host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU
    only): the library's host translation units (C-ABI, graph readers, host
    layout build) compiled by clang++ with the sanitizers and linked with the
    unchanged device objects, and the oracle.  Outputs in _build/san/."""
    sdir = os.path.join(BUILD, "san")
    os.makedirs(sdir, exist_ok=True)
    hip_objs = [os.path.join(BUILD, f + ".o") for f in sorted(os.listdir(CSRC)) if f.endswith(".hip")]
    if not all(os.path.exists(o) for o in hip_objs):
        build_library(jobs)
    cpp_srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".cpp"))
    inc = ["-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", "-std=c++17", "-fPIC"]

    def cc(src):
        obj = os.path.join(sdir, os.path.basename(src) + ".o")
        _run([CLANGXX] + SAN + inc + ["-c", src, "-o", obj])
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(cc, cpp_srcs))
    lib = os.path.join(sdir, "libgrudon_hip.so")
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-shared-libasan", "-Xarch_host",
          "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-o", lib] + objs + hip_objs + ["-lpthread"])
    orc = os.path.join(sdir, "liboracle.so")
    _run([CLANG] + SAN + ["-std=c11", "-fPIC", "-shared", "-shared-libasan", "-pthread",
                          os.path.join(ROOT, "oracle", "grudon_oracle.c"), "-o", orc, "-lm"])
    return lib, orc


def run_sanitized_cpu_tests():
    """pytest -m 'not gpu' against the sanitizer builds, the ASan runtime
    preloaded into the interpreter; returns pytest's exit status."""
    lib, orc = build_sanitized()
    env = dict(os.environ)
    env.update({"LD_PRELOAD": asan_runtime(), "GRUDON_HIP_LIB": lib, "GRUDON_ORACLE_LIB": orc,
                "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1",
                "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1"})
    p = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-m", "not gpu", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests")], env=env, cwd=ROOT)
    return p.returncode


def build_all(oracle=True):
    outs = [build_library(), build_probe()]
    h = build_host()
    if h:
        outs.append(h)
    if oracle:
        outs.append(build_oracle())
    return outs


if __name__ == "__main__":
    if "--sanitize" in sys.argv:
        sys.exit(run_sanitized_cpu_tests())
    for o in build_all(oracle="--no-oracle" not in sys.argv):
        print("built", o)
