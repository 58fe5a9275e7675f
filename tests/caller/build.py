"""Build the compiled drop-in caller (tests/caller/dropin_caller), a C++
program that includes include/kronmult.hpp and links libkronmult_amd.so
in-tree (rpath relative to the binary, so the tree can move)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(HERE, "dropin_caller")
ARCH = os.environ.get("KRON_OFFLOAD_ARCH", "gfx950")


def build_caller():
    cmd = ["hipcc", "--offload-arch=" + ARCH, "-O2", "-std=c++17",
           "-I" + os.path.join(ROOT, "include"), os.path.join(HERE, "dropin_caller.cpp"),
           "-L" + os.path.join(ROOT, "kronmult_amd"), "-lkronmult_amd",
           "-Wl,-rpath,$ORIGIN/../../kronmult_amd", "-o", BIN]
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        raise RuntimeError("drop-in caller build failed:\n" + res.stdout)
    return BIN


if __name__ == "__main__":
    print(build_caller())
