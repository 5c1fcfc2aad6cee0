"""Build recipes: libphc.so (HIP, gfx950), the C++ drop-in drivers, and the
CPU oracle used by the tests.

Everything is built in-tree so that the shared objects travel with the source
tree.  Object files go to _build/ (git-ignored).
"""
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
INCLUDE = os.path.join(ROOT, "include")
ORACLE = os.path.join(ROOT, "oracle")
BUILD = os.path.join(ROOT, "_build")

LIBPHC = os.path.join(PKG, "libphc.so")
LIBORACLE = os.path.join(ORACLE, "liboracle.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 in generated code.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + INCLUDE, "-I" + CSRC]

# Per-source additions.  mlp.hip: MFMA results in VGPRs (the correction
# pass's network reads each result with VALU, which AGPR results would first
# have to copy out, and takes its biases as the accumulator operand).
FILE_FLAGS = {"mlp.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]}

# roctx ranges around the host API calls (rocprofv3 --marker-trace)
ROCTX_LIBS = ["-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx"]

HIP_SOURCES = ["phc_common.hip", "setup.hip", "fd.hip", "sem.hip", "batch.hip", "mlp.hip", "fd64.hip", "sem64.hip", "sem2d.hip"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (r.returncode, " ".join(cmd), out))
    return out


def build_libphc(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(BUILD, s.replace(".hip", ".o")) for s in srcs]

    def compile_one(pair):
        src, obj = pair
        _run([HIPCC] + HIP_FLAGS + FILE_FLAGS.get(src, []) + ["-c", os.path.join(CSRC, src), "-o", obj])

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(compile_one, zip(srcs, objs)))
    tmp = LIBPHC + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ROCTX_LIBS)
    os.replace(tmp, LIBPHC)
    return LIBPHC


def build_oracle():
    srcs = [os.path.join(ORACLE, s) for s in ("fd_oracle.c", "sem_oracle.c", "sem2d_oracle.c")
            if os.path.exists(os.path.join(ORACLE, s))]
    tmp = LIBORACLE + ".tmp"
    # -ffp-contract=off: no FMA contraction, so float/double operations round
    # exactly where the reference's expressions round on x86-64.
    _run(["gcc", "-O2", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
          "-o", tmp] + srcs + ["-lm"])
    os.replace(tmp, LIBORACLE)
    return LIBORACLE


def build_host():
    """C++ drop-in drivers (fd_hip, sem_hip) linked against libphc.so."""
    outs = []
    for name, srcs in (("fd_hip", ["fd_main.cpp"]), ("sem_hip", ["sem_main.cpp", "SEMproxy.cpp"])):
        paths = [os.path.join(HOST, s) for s in srcs]
        exe = os.path.join(PKG, name)
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + INCLUDE, "-I" + HOST] + paths + ["-o", exe,
              "-L" + PKG, "-lphc", "-Wl,-rpath,$ORIGIN"])
        outs.append(exe)
    return outs


def build_checks():
    """Host check programs under tests/cpp/ (the drop-in classes' semantics),
    linked against libphc.so; and, where the reference tree is present, the
    reference's own drivers built unchanged on the drop-in headers
    (oracle/ref_drivers.mk, outputs in oracle/_ref/)."""
    outs = []
    src = os.path.join(ROOT, "tests", "cpp", "dropin_semantics.cpp")
    if os.path.exists(src):
        exe = src[:-4]
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + INCLUDE, "-I" + HOST, src, "-o", exe,
              "-L" + PKG, "-lphc", "-Wl,-rpath,$ORIGIN/../../proxy-geos-hc_amd"])
        outs.append(exe)
    ref = os.environ.get("PHC_REFERENCE", "/root/reference")
    if os.path.isdir(os.path.join(ref, "src", "fd")):
        _run(["make", "-s", "-B", "-f", os.path.join(ORACLE, "ref_drivers.mk"), "REF=" + ref])
        outs.append(os.path.join(ORACLE, "_ref"))
    return outs


def sanitize_host(jobs=8):
    """libphc.so with AddressSanitizer on its host code (the device code is
    built as usual; -Xarch_host keeps the sanitizer off the GPU side) and
    tests/cpp/host_asan_check linked against it, in _build/asan/.  Returns the
    check program's path; it runs without a GPU."""
    out = os.path.join(BUILD, "asan")
    os.makedirs(out, exist_ok=True)
    san = ["-O1", "-g", "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer"]
    flags = [f for f in HIP_FLAGS if f != "-O3"] + san
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(out, s.replace(".hip", ".o")) for s in srcs]

    def compile_one(pair):
        src, obj = pair
        _run([HIPCC] + flags + FILE_FLAGS.get(src, []) + ["-c", os.path.join(CSRC, src), "-o", obj])

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(compile_one, zip(srcs, objs)))
    lib = os.path.join(out, "libphc.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-Xarch_host", "-fsanitize=address", "-o", lib] + objs +
         ROCTX_LIBS)
    exe = os.path.join(out, "host_asan_check")
    _run([HIPCC, "-std=c++17"] + san + ["-I" + INCLUDE, "-I" + HOST,
                                       os.path.join(ROOT, "tests", "cpp", "host_asan_check.cpp"), "-o", exe,
                                       "-L" + out, "-lphc", "-Wl,-rpath," + out])
    return exe


def build_all(jobs=8):
    with cf.ThreadPoolExecutor(max_workers=2) as ex:
        f1 = ex.submit(build_libphc, jobs)
        f2 = ex.submit(build_oracle)
        f1.result()
        f2.result()
    build_host()
    build_checks()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"all": build_all, "phc": build_libphc, "oracle": build_oracle, "host": build_host,
     "checks": build_checks, "sanitize_host": sanitize_host}[what]()
