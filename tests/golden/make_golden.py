"""Write the golden fixtures of tests/golden/ from the reference's own CPU path.

The reference CPU/OpenMP kernels (hipdac-lab/kronmult, compiled from their
sources by oracle/build.py into oracle/_ref/libkron_ref.so) compute every
expected output here; inputs come from the portable splitmix64 generator of
tests/helpers.py or from the reference tests' closed-form formulas
(test_kronmult6_batched.cpp:148-175).  Each fixture is an .npz of inputs and
expected outputs (no pickles).  Run in a checkout where /root/reference has
been built:

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import helpers as H  # noqa: E402
import oracle  # noqa: E402


def ref():
    L = oracle.ref_lib()
    if L is None:
        raise SystemExit("make_golden: oracle/_ref/libkron_ref.so is not built "
                         "(python -c 'from oracle import build; build.build_ref()')")
    return L


def batched_cases(L, out):
    # uniform values (single-threaded reference: the order of every sum is
    # the reference's own) and the reference tests' closed-form inputs
    for d, n, batch in [(1, 5, 4), (2, 4, 3), (3, 3, 3), (4, 4, 2), (5, 3, 2), (6, 2, 3),
                        (6, 4, 2), (3, 7, 2)]:
        E = n ** d
        A = H.uniform(batch * d * n * n, 1000 + 10 * d + n).reshape(batch, d, n * n)
        X = H.uniform(batch * E, 2000 + 10 * d + n).reshape(batch, E)
        Y0 = H.uniform(batch * E, 3000 + 10 * d + n).reshape(batch, E)
        Y = Y0.copy()
        oracle.batched(L, d, n, A, X.copy(), Y, np.zeros_like(X), batch)
        out["batched_u_d%dn%d" % (d, n)] = dict(d=d, n=n, A=A, X=X, Y0=Y0, Y=Y)
    for d, n, batch in [(6, 3, 4), (4, 4, 4), (2, 3, 16)]:
        E = n ** d
        A, X = H.ref_a(n, d, batch), H.ref_x(E, batch)
        Y = np.zeros((batch, E))
        oracle.batched(L, d, n, A, X.copy(), Y, np.zeros_like(X), batch)
        out["batched_ref_d%dn%d" % (d, n)] = dict(d=d, n=n, A=A, X=X, Y0=np.zeros_like(Y), Y=Y)


def xbatched_cases(L, out):
    # distinct Y per entry (the value of Y does not depend on the order of the
    # atomic adds), padded leading dimension lda = n + 3 with NaN rows
    for d, n, batch, nx in [(3, 4, 12, 5), (4, 3, 9, 4), (6, 3, 5, 3), (2, 6, 7, 7)]:
        E, lda = n ** d, n + 3
        A = H.ref_a(n, d, batch, lda=lda).reshape(batch * d, lda * n)
        Xp = H.uniform(nx * E, 4000 + d).reshape(nx, E)
        xm = (H.splitmix64(batch, 4100 + d) % np.uint64(nx)).astype(np.int64)
        Y0 = H.uniform(batch * E, 4200 + d).reshape(batch, E)
        Y = Y0.copy()
        W = np.zeros((2 * batch, E))
        oracle.xbatched(L, d, n, A, np.arange(batch * d).reshape(batch, d), lda, Xp, xm, Y,
                        np.arange(batch), W, np.arange(2 * batch), batch, 0, use_omp=False)
        out["xbatched_d%dn%d" % (d, n)] = dict(d=d, n=n, lda=lda, A=A, Xp=Xp, xm=xm, Y0=Y0, Y=Y)


def vbatched_cases(L, out):
    for m, n in [([3, 2, 4], [4, 5, 2]), ([2, 3], [5, 1]), ([2, 2, 2, 3], [3, 2, 2, 2])]:
        d, batch = len(m), 4
        A = np.zeros((batch * d, max(a * b for a, b in zip(m, n))))
        for b in range(batch):
            for k in range(d):
                A[b * d + k, : m[k] * n[k]] = H.uniform(m[k] * n[k], 5000 + 31 * b + k)
        sx, sy = int(np.prod(n)), int(np.prod(m))
        X = H.uniform(batch * sx, 5100 + d).reshape(batch, sx)
        Y0 = H.uniform(batch * sy, 5200 + d).reshape(batch, sy)
        Y = Y0.copy()
        sw = max(int(np.prod(m[:k + 1])) * int(np.prod(n[k + 1:])) for k in range(d))
        Wcap = 2 * max(sx, sw) * 8 * batch
        oracle.vbatched(L, m, n, A, np.arange(batch * d).reshape(batch, d), X, np.arange(batch),
                        Y, np.arange(batch), np.zeros(Wcap // 8), Wcap, batch, use_omp=False)
        key = "vbatched_" + "x".join("%d%d" % (a, b) for a, b in zip(m, n))
        out[key] = dict(m=np.array(m), n=np.array(n), A=A, X=X, Y0=Y0, Y=Y)


def kgemm_cases(L, out):
    # the reference test's alpha = 1.3, beta = 1.2 and beta = 0 / 1
    for tr, mm, nn, kk, beta in [("N", 7, 5, 6, 1.2), ("T", 8, 4, 9, 1.2), ("N", 3, 10, 2, 0.0),
                                 ("T", 5, 5, 5, 1.0)]:
        batch, ld = 3, 2
        ldA, ldB, ldC = mm + ld, (kk if tr == "N" else nn) + ld, mm + ld
        Ap = H.uniform(batch * ldA * kk, 6000 + mm).reshape(batch, -1)
        bcols = nn if tr == "N" else kk
        Bp = H.uniform(batch * ldB * bcols, 6100 + nn).reshape(batch, -1)
        C0 = H.uniform(batch * ldC * nn, 6200 + kk).reshape(batch, -1)
        C = C0.copy()
        oracle.kgemm(L, tr, mm, nn, kk, 1.3, Ap, [ldA] * batch, Bp, [ldB] * batch, beta, C,
                     [ldC] * batch, batch)
        out["kgemm_%s_%dx%dx%d_b%g" % (tr, mm, nn, kk, beta)] = dict(
            trans=np.array(0 if tr == "N" else 1), mm=mm, nn=nn, kk=kk, alpha=1.3, beta=beta,
            ldA=ldA, ldB=ldB, ldC=ldC, A=Ap, B=Bp, C0=C0, C=C)


def main():
    L = ref()
    oracle.set_threads(L, 1)
    cases = {}
    batched_cases(L, cases)
    xbatched_cases(L, cases)
    vbatched_cases(L, cases)
    kgemm_cases(L, cases)
    total = 0
    for key, arrays in cases.items():
        path = os.path.join(HERE, key + ".npz")
        np.savez_compressed(path, **{k: np.asarray(v) for k, v in arrays.items()})
        total += os.path.getsize(path)
    print("wrote %d fixtures, %.0f KB" % (len(cases), total / 1024))


if __name__ == "__main__":
    main()
