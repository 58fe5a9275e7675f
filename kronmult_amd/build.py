"""Build libkronmult_amd.so in-tree: hipcc for gfx950, no JIT, no torch.

    python -m kronmult_amd.build          # or kronmult_amd.build.build()

Every .hip translation unit is compiled in parallel to an object file under
kronmult_amd/_build/, then linked with the host-only units (capi.cpp, the C
ABI; mixed_plan.cpp, the mixed-shape plan) into kronmult_amd/libkronmult_amd.so.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libkronmult_amd.so")
ARCH = os.environ.get("KRON_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd):
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))


def host_sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith(".cpp"))


def build(jobs=None, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    objs = []
    cmds = []
    for src in sources():
        obj = os.path.join(BUILD, src[:-4] + ".o")
        objs.append(obj)
        cmds.append(["hipcc"] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    # host-only translation units (C ABI, plans)
    for src in host_sources():
        obj = os.path.join(BUILD, src[:-4] + ".o")
        cmds.append(["hipcc", "-O2", "-std=c++17", "-fPIC", "-c",
                     os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run(["hipcc", "--offload-arch=" + ARCH, "-shared", "-fPIC"] + objs + ["-o", tmp])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
