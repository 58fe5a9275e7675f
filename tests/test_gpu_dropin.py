"""A compiled C++ caller of the reference's entry points (tests/caller/
dropin_caller.cpp: includes include/kronmult.hpp, links libkronmult_amd.so,
calls kronmult6_batched, kronmult6_xbatched, kgemm_nn_batched and
kronmult3_vbatched on hipMalloc'd device memory on the null stream).  Its
outputs are checked bit-exact against the oracle on the same integer inputs."""
import os
import subprocess

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "caller", "dropin_caller")
O = oracle.oracle_lib()


def _read(path):
    out = []
    with open(path, "rb") as f:
        while True:
            h = f.read(8)
            if not h:
                break
            n = int(np.frombuffer(h, dtype=np.int64)[0])
            out.append(np.frombuffer(f.read(8 * n), dtype=np.float64).copy())
    return out


@pytest.fixture(scope="module")
def dumps(tmp_path_factory):
    if not os.path.exists(BIN):
        pytest.fail("tests/caller/dropin_caller is not built (python __graft_entry__.py build)")
    d = tmp_path_factory.mktemp("dropin")
    res = subprocess.run([BIN, str(d)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                         text=True, timeout=120)
    assert res.returncode == 0, res.stdout
    assert "dropin_caller ok" in res.stdout
    return d


def test_dropin_batched(dumps):
    A, X, Y0, Y = _read(os.path.join(dumps, "batched.bin"))
    d, n, batch = 6, 4, 37
    Yo = Y0.reshape(batch, -1).copy()
    oracle.batched(O, d, n, A.reshape(batch, d, n * n), X.reshape(batch, -1).copy(), Yo,
                   np.zeros((batch, n ** d)), batch)
    assert np.array_equal(Y.reshape(batch, -1), Yo)


@pytest.mark.parametrize("name,batch", [("xbatched_300", 300), ("xbatched_4096", 4096)])
def test_dropin_xbatched(dumps, name, batch):
    Apool, Xpool, xm, Y0, ym, Y = _read(os.path.join(dumps, name + ".bin"))
    d, n, lda, E = 6, 4, 5, 4096
    Ap = Apool.reshape(batch * d, lda * n)
    Xp = Xpool.reshape(-1, E)
    Yo = Y0.reshape(-1, E).copy()
    nslot = 2 * batch
    oracle.xbatched(O, d, n, Ap, np.arange(batch * d).reshape(batch, d), lda, Xp,
                    xm.astype(np.int64), Yo, ym.astype(np.int64), np.zeros((nslot, E)),
                    np.arange(nslot), batch)
    assert np.array_equal(Y.reshape(-1, E), Yo)


def test_dropin_kgemm(dumps):
    Ap, Bp, C0, cm, C = _read(os.path.join(dumps, "kgemm.bin"))
    mm, nn, kk, batch, ldA, ldB, ldC = 7, 5, 6, 9, 8, 6, 9
    want = C0.reshape(-1, ldC * nn).copy()
    for b in range(batch):
        A = Ap[b * ldA * kk:(b + 1) * ldA * kk].reshape(kk, ldA).T[:mm, :]
        B = Bp[b * ldB * nn:(b + 1) * ldB * nn].reshape(nn, ldB).T[:kk, :]
        Cm = want[int(cm[b])].reshape(nn, ldC).T
        Cm[:mm, :] += A @ B
        want[int(cm[b])] = Cm.T.reshape(-1)
    assert np.array_equal(C.reshape(-1, ldC * nn), want)


def test_dropin_vbatched(dumps):
    Apool, Xp, Y0, Y = _read(os.path.join(dumps, "vbatched.bin"))
    m, n, batch, ny = [3, 2, 4], [4, 5, 2], 10, 5
    Ap = Apool.reshape(batch * 3, 12)
    Yo = Y0.reshape(ny, -1).copy()
    sx = int(np.prod(n))
    Wcap = 4 * 80 * 8
    oracle.vbatched(O, m, n, Ap, np.arange(batch * 3).reshape(batch, 3), Xp.reshape(batch, sx),
                    np.arange(batch), Yo, np.arange(batch) % ny, np.zeros(Wcap // 8), Wcap, batch)
    assert np.array_equal(Y.reshape(ny, -1), Yo)
