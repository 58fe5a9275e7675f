"""The HIP library through its C ABI against the golden fixtures of the
reference's CPU path (tests/golden/): FP64 within 1e-12 in the reference
tests' relative measure (the GPU kernels sum in another order)."""
import glob
import os

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import kronmult_amd as K  # noqa: E402
from gpu_util import DEV, dev, host, ptrs  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def _load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_library_matches_reference(path):
    key = os.path.basename(path)[:-4]
    z = _load(path)
    if key.startswith("batched"):
        d, n = int(z["d"]), int(z["n"])
        Xd, Yd = dev(z["X"]), dev(z["Y0"])
        K.kronmult_batched(d, n, dev(z["A"]), Xd, Yd, torch.empty_like(Xd), z["X"].shape[0])
        got, want = host(Yd), z["Y"]
    elif key.startswith("xbatched"):
        d, n, lda = int(z["d"]), int(z["n"]), int(z["lda"])
        batch = z["Y0"].shape[0]
        Ad, Xd, Yd = dev(z["A"]), dev(z["Xp"]), dev(z["Y0"])
        Wd = torch.zeros((2 * batch, z["Xp"].shape[1]), dtype=torch.float64, device=DEV)
        K.kronmult_xbatched(d, n, ptrs(Ad, np.arange(batch * d)), lda, ptrs(Xd, z["xm"]),
                            ptrs(Yd, np.arange(batch)), ptrs(Wd, np.arange(2 * batch)), batch)
        got, want = host(Yd), z["Y"]
    elif key.startswith("vbatched"):
        m, n = [int(v) for v in z["m"]], [int(v) for v in z["n"]]
        d, batch = len(m), z["X"].shape[0]
        sx = int(np.prod(n))
        sw = max(int(np.prod(m[:k + 1])) * int(np.prod(n[k + 1:])) for k in range(d))
        Wcap = 2 * max(sx, sw) * 8 * batch
        Ad, Xd, Yd = dev(z["A"]), dev(z["X"]), dev(z["Y0"])
        Wd = torch.empty(Wcap // 8, dtype=torch.float64, device=DEV)
        K.kronmult_vbatched(m, n, ptrs(Ad, np.arange(batch * d)), ptrs(Xd, np.arange(batch)),
                            ptrs(Yd, np.arange(batch)), Wd, Wcap, batch)
        got, want = host(Yd), z["Y"]
    else:
        batch = z["C0"].shape[0]
        Ad, Bd, Cd = dev(z["A"]), dev(z["B"]), dev(z["C0"])
        f = K.kgemm_nt_batched if int(z["trans"]) else K.kgemm_nn_batched
        ld = lambda v: torch.full((batch,), int(v), dtype=torch.int32, device=DEV)
        f(int(z["mm"]), int(z["nn"]), int(z["kk"]), float(z["alpha"]),
          ptrs(Ad, np.arange(batch)), ld(z["ldA"]), ptrs(Bd, np.arange(batch)), ld(z["ldB"]),
          float(z["beta"]), ptrs(Cd, np.arange(batch)), ld(z["ldC"]), batch)
        got, want = host(Cd), z["C"]
    assert H.rel_err(got, want) < 1e-12, key
