"""Build the native parts of os_d_amd in-tree.

libqjx.so  : HIP kernels + C++ engine + C ABI (include/qjx.h), gfx950 only.
qjx        : the drop-in command-line runner (config.txt / params.txt in, reference
             output files out), linked against libqjx.so.
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
LIBDIR = os.path.join(ROOT, "_lib")
LIB = os.path.join(LIBDIR, "libqjx.so")
CLI = os.path.join(LIBDIR, "qjx")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 instructions are disabled on purpose (see the design notes).
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
LIB_SOURCES = ["qjx_kernels.hip", "qjx_step.hip", "qjx_lindblad.hip", "qjx_rk.hip", "qjx_engine.cpp",
               "qjx_config.cpp", "qjx_pool.cpp", "qjx_lpn.cpp"]


def _run(cmd):
    res = subprocess.run(cmd, cwd=CSRC, capture_output=True, text=True)
    out = "\n".join(l for l in (res.stdout + res.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if res.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build_lib():
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [s for s in LIB_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-shared", "-I", os.path.join(REPO, "include")] + srcs + ["-o", tmp])
    os.replace(tmp, LIB)
    if os.path.exists(os.path.join(CSRC, "qjx_main.cpp")):
        tmp = CLI + ".tmp"
        _run([HIPCC, "-O2", "-std=c++17", "-I", os.path.join(REPO, "include"), "qjx_main.cpp",
              "-L", LIBDIR, "-lqjx", "-Wl,-rpath,$ORIGIN", "-o", tmp])
        os.replace(tmp, CLI)
    return LIB


if __name__ == "__main__":
    print(build_lib())
