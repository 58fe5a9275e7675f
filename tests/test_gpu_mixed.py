"""The mixed-(d, n) persistent launch (kron_mixed_plan_*) against the oracle:
every class of a plan is checked as the reference would compute it, one
kronmult_xbatched call per class into the same Y arrays
(kronmult_xbatched.hpp:19-157), with integer-valued inputs (exact in any
summation order, so the results must match bit for bit) and with uniform
inputs at 1e-12 relative."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import kronmult_amd as K  # noqa: E402
import oracle  # noqa: E402
from gpu_util import DEV, dev, host, ptrs  # noqa: E402

O = oracle.oracle_lib()

# (d, n, batch, nx, ny, lda): every path of the kernel plus a separate launch
CLASSES = [
    (1, 5, 70, 20, 9, 5),      # packed, d = 1
    (2, 3, 200, 40, 30, 3),    # packed
    (2, 8, 90, 30, 40, 9),     # packed
    (3, 4, 120, 30, 50, 4),    # packed
    (3, 8, 60, 20, 30, 8),     # packed
    (4, 4, 100, 25, 40, 6),    # packed
    (4, 6, 50, 20, 20, 6),     # one entry per workgroup
    (5, 3, 80, 20, 30, 3),     # packed
    (5, 5, 40, 15, 14, 5),     # one entry per workgroup
    (5, 6, 20, 8, 7, 7),       # one entry per workgroup (7,776 elements)
    (6, 2, 90, 20, 25, 2),     # packed
    (6, 4, 60, 20, 19, 4),     # one entry per workgroup
    (6, 5, 10, 5, 4, 5),       # slabs (mappings 0, 1) or two passes (mapping 2)
    (5, 7, 10, 5, 4, 8),       # slabs / two passes
    (5, 8, 6, 4, 3, 8),        # slabs / two passes
    (6, 6, 6, 3, 3, 6),        # slabs / two passes
    (3, 9, 12, 6, 5, 9),       # outside the compiled set: separate launch
]


def _class_data(d, n, batch, nx, ny, lda, seed, kind, dtype):
    E = n ** d
    if kind == "int":
        A = H.integers(batch * d * n * n, H.sub_seed(seed, 1), -2, 2)
        X = H.integers(nx * E, H.sub_seed(seed, 2), -4, 4)
        Y0 = H.integers(ny * E, H.sub_seed(seed, 5), -8, 8)
    else:
        A = H.uniform(batch * d * n * n, H.sub_seed(seed, 1))
        X = H.uniform(nx * E, H.sub_seed(seed, 2))
        Y0 = H.uniform(ny * E, H.sub_seed(seed, 5))
    A_pool = np.full((batch * d, n, lda), np.nan)
    A_pool[:, :, :n] = A.reshape(batch * d, n, n)
    A_pool = A_pool.reshape(batch * d, lda * n).astype(dtype)
    X_pool = X.reshape(nx, E).astype(dtype)
    Y0_pool = Y0.reshape(ny, E).astype(dtype)
    x_map = (H.splitmix64(batch, H.sub_seed(seed, 3)) % np.uint64(nx)).astype(np.int64)
    y_map = (H.splitmix64(batch, H.sub_seed(seed, 4)) % np.uint64(ny)).astype(np.int64)
    return A_pool, X_pool, Y0_pool, x_map, y_map


def _oracle_class(d, n, lda, A_pool, X_pool, x_map, Y_pool, y_map):
    """One reference xbatched call into Y_pool (float64, in place)."""
    batch = len(x_map)
    W = np.zeros((max(2, batch), X_pool.shape[1]))
    oracle.xbatched(O, d, n, A_pool.astype(np.float64), np.arange(batch * d).reshape(batch, d),
                    lda, X_pool.astype(np.float64), x_map, Y_pool, y_map, W,
                    np.arange(W.shape[0]), batch)


class _Case:
    """Device and host data of a list of classes; each class has its own Y
    pool unless share_y maps it onto another class's pool."""

    def __init__(self, classes, seed=1, kind="int", dtype=np.float64, share_y=None):
        self.classes = classes
        self.host, self.devs, self.mixed = [], [], []
        share_y = share_y or {}
        for i, (d, n, batch, nx, ny, lda) in enumerate(classes):
            A_pool, X_pool, Y0_pool, x_map, y_map = _class_data(d, n, batch, nx, ny, lda,
                                                                seed * 100 + i, kind, dtype)
            if i in share_y:
                j = share_y[i]
                Y0_pool = self.host[j][2]
                Yd = self.devs[j][2]
                y_map = y_map % Y0_pool.shape[0]
            else:
                Yd = dev(Y0_pool)
            Ad, Xd = dev(A_pool), dev(X_pool)
            pA, pX, pY = ptrs(Ad, np.arange(batch * d)), ptrs(Xd, x_map), ptrs(Yd, y_map)
            self.host.append((A_pool, X_pool, Y0_pool, x_map, y_map))
            self.devs.append((Ad, Xd, Yd, pA, pX, pY))
            self.mixed.append(K.MixedClass(d, n, pA, pX, pY, batch, lda=lda))
        self.dtype = torch.float64 if dtype == np.float64 else torch.float32

    def plan(self, **kw):
        return K.MixedPlan(self.mixed, dtype=self.dtype, **kw)

    def expected(self, applies=1):
        """Y pools after `applies` reference applies (float64), keyed by pool id."""
        pools = {}
        for (d, n, batch, nx, ny, lda), (A_pool, X_pool, Y0_pool, x_map, y_map) in zip(
                self.classes, self.host):
            pools.setdefault(id(Y0_pool), Y0_pool.astype(np.float64))
        for _ in range(applies):
            for (d, n, batch, nx, ny, lda), (A_pool, X_pool, Y0_pool, x_map, y_map) in zip(
                    self.classes, self.host):
                _oracle_class(d, n, lda, A_pool, X_pool, x_map, pools[id(Y0_pool)], y_map)
        return pools

    def got(self):
        out = {}
        for (_, _, Y0_pool, _, _), (_, _, Yd, _, _, _) in zip(self.host, self.devs):
            out[id(Y0_pool)] = host(Yd).astype(np.float64)
        return out


def _assert_pools(got, want, exact, tol=1e-12):
    assert got.keys() == want.keys()
    for k in want:
        if exact:
            assert np.array_equal(got[k], want[k])
        else:
            assert H.rel_err(got[k], want[k]) < tol


@pytest.mark.parametrize("flags", [0, K.KRON_MIXED_ATOMIC, K.KRON_MIXED_FORCE_V0,
                                   K.KRON_MIXED_FORCE_V1, K.KRON_MIXED_FORCE_V2])
def test_mixed_integer_bit_exact(flags):
    case = _Case(CLASSES, seed=3)
    plan = case.plan(flags=flags)
    plan.execute()
    _assert_pools(case.got(), case.expected(), exact=True)
    kinds = {(c[0], c[1]): plan.class_info(i)["kind"] for i, c in enumerate(CLASSES)}
    assert kinds[(3, 9)] == "launch"
    big = "twopass" if flags == K.KRON_MIXED_FORCE_V2 else "slab"
    assert kinds[(6, 6)] == big and kinds[(5, 8)] == big
    # one persistent launch, plus the separate class (two-pass: two launches)
    assert plan.launch_info()["launches"] == (3 if big == "twopass" else 2)
    assert kinds[(2, 3)] == "packed" and kinds[(1, 5)] == "packed"
    if flags not in (K.KRON_MIXED_FORCE_V1, K.KRON_MIXED_FORCE_V2):
        assert kinds[(6, 4)] == "onchip" and kinds[(5, 6)] == "onchip"
    for i in range(len(CLASSES)):
        info = plan.class_info(i)
        if flags == K.KRON_MIXED_ATOMIC or info["kind"] == "packed":
            assert info["atomic"] == 1
    plan.close()


@pytest.mark.parametrize("flags", [K.KRON_MIXED_FORCE_V0, K.KRON_MIXED_FORCE_V1,
                                   K.KRON_MIXED_FORCE_V0 | K.KRON_MIXED_ATOMIC,
                                   K.KRON_MIXED_FORCE_V1 | K.KRON_MIXED_ATOMIC])
def test_mixed_slab_classes_walk(flags):
    """The four classes too large for LDS as slab items only: as many slabs
    per item as LDS holds (3, 2, 2 and 1 of 7, 8, 5 and 6: a last item with
    fewer) or one slab per item (more items than resident workgroups), Y
    groups of several entries and padded factors; integer-valued, bit-exact."""
    classes = [(6, 6, 150, 60, 40, 7), (5, 8, 200, 80, 50, 8), (5, 7, 260, 90, 70, 9),
               (6, 5, 300, 100, 60, 5)]
    case = _Case(classes, seed=31)
    plan = case.plan(flags=flags)
    for i in range(len(classes)):
        assert plan.class_info(i)["kind"] == "slab"
    if flags & K.KRON_MIXED_FORCE_V1:  # one slab per item: every workgroup walks items
        assert plan.launch_info()["items"] > plan.launch_info()["grid"]
    plan.execute()
    _assert_pools(case.got(), case.expected(), exact=True)


def test_mixed_uniform_f64():
    case = _Case(CLASSES, seed=5, kind="uniform")
    plan = case.plan()
    plan.execute()
    _assert_pools(case.got(), case.expected(), exact=False, tol=1e-12)


def test_mixed_f32():
    classes = [c for c in CLASSES if c[1] ** c[0] <= 20000]
    case = _Case(classes, seed=7, kind="uniform", dtype=np.float32)
    plan = case.plan()
    plan.execute()
    _assert_pools(case.got(), case.expected(), exact=False, tol=2e-5)


def test_mixed_cross_class_shared_y_goes_atomic():
    """(2,8), (3,4) and (6,2) all have 64-element vectors; the second and third
    write into the first one's Y pool, so every Y vector has writers from
    three classes (three different kernel paths)."""
    classes = [(2, 8, 80, 20, 10, 8), (3, 4, 90, 20, 10, 4), (6, 2, 70, 20, 10, 2),
               (6, 4, 40, 10, 8, 4)]
    case = _Case(classes, seed=11, share_y={1: 0, 2: 0})
    plan = case.plan()
    for i in range(3):
        assert plan.class_info(i)["atomic"] == 1
    assert plan.class_info(3)["atomic"] == 0  # its own pool, targets disjoint
    plan.execute()
    _assert_pools(case.got(), case.expected(), exact=True)


def test_mixed_partially_overlapping_targets():
    """Y vectors that overlap without being equal (rows 40 elements apart,
    each 64 long): the plan must fall back to atomics."""
    d, n, batch, ny = 6, 2, 50, 12
    E = n ** d
    A_pool, X_pool, _, x_map, y_map = _class_data(d, n, batch, 10, ny, n, 13, "int", np.float64)
    flat = H.integers(40 * (ny - 1) + E, 99, -8, 8)
    Yflat_d = dev(flat)
    Yd = torch.as_strided(Yflat_d, (ny, E), (40, 1))
    Ad, Xd = dev(A_pool), dev(X_pool)
    pA, pX, pY = ptrs(Ad, np.arange(batch * d)), ptrs(Xd, x_map), ptrs(Yd, y_map)
    plan = K.MixedPlan([K.MixedClass(d, n, pA, pX, pY, batch)])
    assert plan.class_info(0)["atomic"] == 1
    plan.execute()
    want = flat.copy()
    Yw = np.lib.stride_tricks.as_strided(want, shape=(ny, E), strides=(40 * 8, 8))
    _oracle_class(d, n, n, A_pool, X_pool, x_map, Yw, y_map)
    assert np.array_equal(host(Yflat_d), want)


def test_mixed_long_group_is_split():
    """200 entries into one Y vector: the group exceeds the per-item limit,
    the class goes atomic and the group is spread over several items."""
    classes = [(6, 4, 200, 30, 1, 4), (4, 6, 150, 20, 1, 6)]
    case = _Case(classes, seed=17)
    plan = case.plan()
    assert plan.class_info(0)["atomic"] == 1 and plan.class_info(1)["atomic"] == 1
    plan.execute()
    _assert_pools(case.got(), case.expected(), exact=True)


def test_mixed_graph_replays():
    """Two applies captured in a hipGraph, replayed three times with eager
    work and synchronisations between; every replay checked."""
    classes = [CLASSES[i] for i in (1, 4, 6, 9, 11, 12, 16)]
    case = _Case(classes, seed=19)
    plan = case.plan()
    plan.capture(2)
    for r in range(1, 4):
        plan.replay()
        torch.cuda.synchronize()
        scratch = torch.ones(1 << 20, dtype=torch.float64, device=DEV)
        scratch.mul_(2.0)  # unrelated eager work between replays
        torch.cuda.synchronize()
        _assert_pools(case.got(), case.expected(applies=2 * r), exact=True)
    plan.execute()  # eager execute after replays: the queue was reset
    _assert_pools(case.got(), case.expected(applies=7), exact=True)


def test_mixed_autotune_and_cache(tmp_path):
    path = str(tmp_path / "tune.txt")
    classes = [(2, 5, 100, 20, 30, 5), (4, 6, 60, 20, 20, 6), (6, 5, 8, 4, 4, 5),
               (3, 9, 6, 3, 3, 9)]
    case = _Case(classes, seed=23)
    plan = case.plan(tune_cache=path, tune=True)
    lines = open(path).read().split("\n")
    lines = [l for l in lines if l.strip()]
    assert len(lines) == 3  # the separately launched class is not tuned
    for i in range(3):
        info = plan.class_info(i)
        assert info["variant"] in (0, 1, 2) and info["cost"] > 0
    plan.execute()
    _assert_pools(case.got(), case.expected(), exact=True)
    # a second plan takes the mappings from the cache without re-timing
    plan2 = case.plan(tune_cache=path)
    for i in range(3):
        assert plan2.class_info(i)["variant"] == plan.class_info(i)["variant"]
    assert len([l for l in open(path).read().split("\n") if l.strip()]) == 3


def test_mixed_empty_and_errors():
    plan = K.MixedPlan([])
    plan.execute()
    assert plan.launch_info()["items"] == 0
    z = torch.zeros(1, dtype=torch.int64, device=DEV)
    plan = K.MixedPlan([K.MixedClass(4, 3, z, z, z, 0)])
    plan.execute()
    with pytest.raises(K.KronError) as e:
        K.MixedPlan([K.MixedClass(7, 2, z, z, z, 1)])
    assert e.value.status == K.KRON_EINVAL
    with pytest.raises(K.KronError) as e:
        K.MixedPlan([K.MixedClass(3, 4, z, z, z, 1, lda=3)])
    assert e.value.status == K.KRON_EINVAL
    with pytest.raises(K.KronError) as e:
        K.MixedPlan([K.MixedClass(3, 4, None, z, z, 1)])
    assert e.value.status == K.KRON_EINVAL


def test_mixed_sparse_grid_workload_shape():
    """The 8,192-entry workload over 28 (d, n) classes (d in 2..6, n in 3..8,
    n^d <= 65,536), one X per entry, Y targets from a pool of 1,024 per class,
    uniform inputs, against the oracle at 1e-12."""
    counts = H.mixed_class_counts(8192, 2024)
    assert len(counts) == 28
    classes = [(d, n, c, c, 1024, n) for (d, n), c in counts.items()]
    case = _Case(classes, seed=29, kind="uniform")
    plan = case.plan()
    info = plan.launch_info()
    assert info["items"] > 0 and info["grid"] > 0
    assert info["launches"] == 1  # every class in the one persistent launch
    plan.execute()
    _assert_pools(case.got(), case.expected(), exact=False, tol=1e-12)


@pytest.mark.parametrize("m,n", [([3, 4], [4, 2]), ([2, 3, 4], [3, 2, 2]), ([4] * 6, [4] * 6),
                                 ([1, 2, 1, 2, 1], [2, 1, 2, 1, 2])])
def test_plain_per_entry_kernel(m, n):
    """The reference-structured baseline (kron_vbatched_plain_*) against the
    oracle's vbatched, rectangular shapes, shared Y, W room for 3 slots (so
    several sub-batches run)."""
    d, batch, ny = len(m), 40, 7
    sx, sy = int(np.prod(n)), int(np.prod(m))
    sizes = [int(np.prod(m[:k + 1]) * np.prod(n[k + 1:])) for k in range(d - 1)]
    sxw = max([sx] + sizes)
    A_pool = H.integers(batch * d * 64, 41, -2, 2).reshape(batch * d, 64)
    for k in range(d):
        A_pool[k::d, m[k] * n[k]:] = np.nan
    X_pool = H.integers(batch * sx, 42, -4, 4).reshape(batch, sx)
    Y0 = H.integers(ny * sy, 43, -8, 8).reshape(ny, sy)
    y_map = (H.splitmix64(batch, 44) % np.uint64(ny)).astype(np.int64)
    want = Y0.copy()
    Wh = np.zeros(2 * sxw * batch)
    oracle.vbatched(O, m, n, A_pool, np.arange(batch * d).reshape(batch, d), X_pool,
                    np.arange(batch), want, y_map, Wh, Wh.nbytes, batch)
    Ad, Xd, Yd = dev(A_pool), dev(X_pool), dev(Y0)
    Wd = torch.full((3 * 2 * sxw,), float("nan"), dtype=torch.float64, device=DEV)
    K.kronmult_vbatched_plain(m, n, ptrs(Ad, np.arange(batch * d)), ptrs(Xd, np.arange(batch)),
                              ptrs(Yd, y_map), Wd, Wd.numel() * 8, batch)
    assert np.array_equal(host(Yd), want)
