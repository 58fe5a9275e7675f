"""The oracle against golden fixtures written by the reference's own CPU path
(tests/golden/make_golden.py runs oracle/_ref, compiled from the reference's
sources).  Single-threaded, without overlapping Y, the restatement follows
the reference's operation order, so every value must be bit-identical; this
pins the oracle on checkouts where the reference itself is absent."""
import glob
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))
O = oracle.oracle_lib()


def _load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_fixtures_present():
    kinds = {os.path.basename(p).split("_")[0] for p in FIXTURES}
    assert {"batched", "xbatched", "vbatched", "kgemm"} <= kinds
    assert sum(os.path.getsize(p) for p in FIXTURES) < (1 << 20)


def oracle_output(z, key, L=O):
    """The oracle's (or another oracle-interface library's) output for the
    inputs of fixture `key`."""
    if key.startswith("batched"):
        d, n = int(z["d"]), int(z["n"])
        Y = z["Y0"].copy()
        X = z["X"].copy()
        oracle.batched(L, d, n, z["A"].copy(), X, Y, np.zeros_like(X), X.shape[0],
                       use_omp=False)
        return Y
    if key.startswith("xbatched"):
        d, n, lda = int(z["d"]), int(z["n"]), int(z["lda"])
        batch = z["Y0"].shape[0]
        Y = z["Y0"].copy()
        W = np.zeros((2 * batch, z["Xp"].shape[1]))
        oracle.xbatched(L, d, n, z["A"], np.arange(batch * d).reshape(batch, d), lda, z["Xp"],
                        z["xm"], Y, np.arange(batch), W, np.arange(2 * batch), batch, 0,
                        use_omp=False)
        return Y
    if key.startswith("vbatched"):
        m, n = [int(v) for v in z["m"]], [int(v) for v in z["n"]]
        d, batch = len(m), z["X"].shape[0]
        sx = int(np.prod(n))
        sw = max(int(np.prod(m[:k + 1])) * int(np.prod(n[k + 1:])) for k in range(d))
        Wcap = 2 * max(sx, sw) * 8 * batch
        Y = z["Y0"].copy()
        oracle.vbatched(L, m, n, z["A"], np.arange(batch * d).reshape(batch, d), z["X"],
                        np.arange(batch), Y, np.arange(batch), np.zeros(Wcap // 8), Wcap, batch,
                        use_omp=False)
        return Y
    tr = "T" if int(z["trans"]) else "N"
    batch = z["C0"].shape[0]
    C = z["C0"].copy()
    oracle.kgemm(L, tr, int(z["mm"]), int(z["nn"]), int(z["kk"]), float(z["alpha"]), z["A"],
                 [int(z["ldA"])] * batch, z["B"], [int(z["ldB"])] * batch, float(z["beta"]), C,
                 [int(z["ldC"])] * batch, batch)
    return C


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_reproduces_reference(path):
    key = os.path.basename(path)[:-4]
    z = _load(path)
    want = z["C"] if key.startswith("kgemm") else z["Y"]
    got = oracle_output(z, key)
    assert np.array_equal(got, want), key


def test_planted_mistake_is_caught():
    """A transposed factor in one mode (a plausible indexing mistake) changes
    the batched d3n3 fixture's output."""
    z = _load(os.path.join(HERE, "golden", "batched_u_d3n3.npz"))
    A = z["A"].copy()
    A[:, 1] = A[:, 1].reshape(-1, 3, 3).transpose(0, 2, 1).reshape(A.shape[0], 9)
    z["A"] = A
    assert not np.array_equal(oracle_output(z, "batched_u_d3n3"), z["Y"])
