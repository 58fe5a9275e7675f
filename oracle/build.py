"""Build recipe for the oracle (test infrastructure only).

* ``oracle/liboracle.so``: the C restatement ``kron_oracle.c`` (gcc, OpenMP).
* ``oracle/_ref/libkron_ref.so``: the reference kronmult CPU path compiled
  from its own sources under /root/reference (host-only clang via hipcc),
  plus ``ref_shim.cpp``.  Built only where /root/reference exists; the
  resulting .so is git-ignored and travels with the working tree.

The reference headers name GPU builtins inside never-instantiated templates
(kgemm_nn.hpp:70,132; kgemm_nt.hpp:62,101), which a conforming compiler
rejects at definition time; ``-fdelayed-template-parsing`` defers parsing of
template bodies until instantiation, so the CPU path compiles unmodified.
``kron_flops`` (kroncommon.hpp:204) is a non-inline function defined in a
header, so linking several reference translation units needs
``--allow-multiple-definition``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.environ.get("KRON_REFERENCE_DIR", "/root/reference")
LLVM_LIB = "/opt/rocm/llvm/lib"  # clang's OpenMP runtime (libomp)

REF_SOURCES = (
    [f"kronmult{d}_batched.cpp" for d in range(1, 7)]
    + [f"kronmult{d}_xbatched.cpp" for d in range(1, 7)]
    # kronmult4_pbatched.cpp:3 is not valid C++ as shipped (a malformed
    # preprocessing directive); d = 4 pbatched is left out of this library
    + [f"kronmult{d}_pbatched.cpp" for d in (1, 2, 3, 5, 6)]
    + ["kgemm_nn_batched.cpp", "kgemm_nt_batched.cpp"]
)


def _run(cmd):
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        raise RuntimeError("oracle build failed: " + " ".join(cmd))


def build_oracle():
    out = os.path.join(HERE, "liboracle.so")
    _run(["gcc", "-O3", "-std=c11", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared",
          os.path.join(HERE, "kron_oracle.c"), "-o", out])
    return out


def build_ref():
    """Compile the reference CPU path; returns the .so path or None."""
    if not os.path.isdir(REF_DIR):
        return None
    outdir = os.path.join(HERE, "_ref")
    os.makedirs(outdir, exist_ok=True)
    out = os.path.join(outdir, "libkron_ref.so")
    srcs = [os.path.join(REF_DIR, s) for s in REF_SOURCES]
    cmd = ["hipcc", "-x", "hip", "--offload-host-only", "-fdelayed-template-parsing",
           "-O3", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared",
           "-w", "-I" + REF_DIR] + srcs + [os.path.join(HERE, "ref_shim.cpp"),
                                           "-Wl,--allow-multiple-definition",
                                           "-Wl,-rpath," + LLVM_LIB, "-o", out]
    _run(cmd)
    return out


if __name__ == "__main__":
    print(build_oracle())
    print(build_ref())
