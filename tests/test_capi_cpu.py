"""The C-ABI library without a GPU: it loads, exports every symbol that
include/*.h declares (the C ABI and the reference-signature C++ overloads),
and rejects invalid arguments on the host before any HIP call."""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "kronmult_amd", "libkronmult_amd.so")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        from kronmult_amd import build
        build.build()
    import kronmult_amd
    return kronmult_amd.lib()


def _declared_c_symbols():
    h = open(os.path.join(ROOT, "include", "kronmult_amd.h")).read()
    h = re.sub(r"/\*.*?\*/", "", h, flags=re.S)
    return sorted(set(re.findall(r"\b(kron_[a-z0-9_]+)\s*\(", h)))


def test_every_c_symbol_exported(lib):
    names = _declared_c_symbols()
    assert len(names) >= 17
    for name in names:
        assert hasattr(lib, name), name


def _exported_cpp():
    if not shutil.which("nm"):
        pytest.skip("nm not available")
    out = subprocess.run(["nm", "-DC", "--defined-only", LIB], stdout=subprocess.PIPE,
                         text=True, check=True).stdout
    return set(l.split(" ", 2)[2] for l in out.splitlines() if " T " in l)


def test_reference_signatures_exported(lib):
    """Demangled signatures as the reference build exports them (checked
    against oracle/_ref's nm output where that exists)."""
    syms = _exported_cpp()
    want = []
    for T in ("double", "float"):
        for d in range(1, 7):
            want.append("kronmult%d_batched(int, %s const*, %s*, %s*, %s*, int)" % (d, T, T, T, T))
            want.append("kronmult%d_xbatched(int, %s const* const*, int, %s**, %s**, %s**, int, int)"
                        % (d, T, T, T, T))
            want.append("kronmult%d_pbatched(int, %s const*, %s**, %s**, %s**, int)"
                        % (d, T, T, T, T))
            mn = ", ".join(["int"] * (2 * d))
            want.append("kronmult%d_vbatched(%s, %s const* const*, %s**, %s**, %s*, unsigned long, int)"
                        % (d, mn, T, T, T, T))
        for tr in ("nn", "nt"):
            want.append("kgemm_%s_batched(int, int, int, %s, %s* const*, int const*, %s* const*, "
                        "int const*, %s, %s* const*, int const*, int)" % (tr, T, T, T, T, T))
    missing = [w for w in want if w not in syms]
    assert not missing, missing
    ref = os.path.join(ROOT, "oracle", "_ref", "libkron_ref.so")
    if os.path.exists(ref):
        out = subprocess.run(["nm", "-DC", "--defined-only", ref], stdout=subprocess.PIPE,
                             text=True, check=True).stdout
        ref_syms = set(l.split(" ", 2)[2] for l in out.splitlines() if " T " in l)
        for w in want:
            # kronmult4_pbatched.cpp does not compile as shipped (not in oracle/_ref)
            if "double" in w and "vbatched" not in w and "kronmult4_pbatched" not in w:
                assert w in ref_syms, "reference exports a different signature: " + w


def test_host_side_validation(lib):
    vp = ctypes.c_void_p
    fake = vp(0x1000)  # never dereferenced: validation fails first
    assert lib.kron_batched_f64(7, 2, fake, fake, fake, fake, 1, None) == 1
    assert lib.kron_batched_f64(0, 2, fake, fake, fake, fake, 1, None) == 1
    assert lib.kron_batched_f64(3, 0, fake, fake, fake, fake, 1, None) == 1
    assert lib.kron_batched_f64(3, 2, fake, fake, fake, fake, -1, None) == 1
    assert lib.kron_batched_f64(3, 2, None, fake, fake, fake, 1, None) == 1
    assert lib.kron_batched_f64(3, 2, fake, fake, fake, fake, 0, None) == 0  # empty: no-op
    assert lib.kron_batched_f64(6, 1000, fake, fake, fake, fake, 1, None) == 1  # n^d overflow
    # xbatched: lda < n, negative subbatch
    assert lib.kron_xbatched_f64(3, 4, fake, 3, fake, fake, fake, 1, 0, None) == 1
    assert lib.kron_xbatched_f64(3, 4, fake, 4, fake, fake, fake, 1, -1, None) == 1
    # vbatched: zero extent, Wcapacity below one slot (kronmult_vbatched.hpp:157)
    m = (ctypes.c_int * 2)(2, 3)
    n = (ctypes.c_int * 2)(3, 0)
    assert lib.kron_vbatched_f64(2, m, n, fake, fake, fake, fake, 1 << 20, 1, None) == 1
    n = (ctypes.c_int * 2)(3, 4)
    # sizeX = 12, sizeW = max(2*4) = 8 -> slot = 2*12*8 = 192 bytes
    assert lib.kron_vbatched_f64(2, m, n, fake, fake, fake, fake, 191, 1, None) == 2
    assert lib.kron_vbatched_f64(2, m, n, fake, fake, fake, None, 192, 1, None) == 1
    # kgemm
    assert lib.kron_kgemm_nn_batched_f64(-1, 2, 2, 1.0, fake, fake, fake, fake, 0.0, fake,
                                         fake, 1, None) == 1
    assert lib.kron_kgemm_nt_batched_f64(2, 2, 2, 1.0, fake, fake, fake, fake, 0.0, fake,
                                         fake, 0, None) == 0
    assert lib.kron_status_string(2).decode().startswith("KRON_EWORKSPACE")
    assert lib.kron_version() >= 100


def test_adjoint_workspace_query(lib):
    """Scratch of the adjoint per entry: none when the entry fits on chip,
    two vectors when it streams (d = 6, n = 6); argument errors on the host."""
    n_out = ctypes.c_size_t()
    assert lib.kron_batched_adjoint_workspace(6, 6, 8, ctypes.byref(n_out)) == 0
    assert n_out.value == 2 * 6 ** 6
    assert lib.kron_batched_adjoint_workspace(2, 3, 8, ctypes.byref(n_out)) == 0
    assert n_out.value == 0
    # d = 6, n = 4 (4,096 elements) runs on chip in both precisions
    for elem in (8, 4):
        assert lib.kron_batched_adjoint_workspace(6, 4, elem, ctypes.byref(n_out)) == 0
        assert n_out.value == 0
    assert lib.kron_batched_adjoint_workspace(7, 2, 8, ctypes.byref(n_out)) == 1
    assert lib.kron_batched_adjoint_workspace(3, 2, 2, ctypes.byref(n_out)) == 1
