// kron_fast.hpp -- fused on-chip kernel for square factors, compile-time (D, N).
//
// Replaces the reference's tail recursion (kronmultv.hpp:19-148), which
// issues one small GEMM per vector per level with every intermediate going
// through global memory, by one pass per mode over an LDS-resident copy of
// the entry:
//
//   mode 2      : X (HBM, coalesced)      -> registers -> LDS
//   modes D..3  : LDS -> registers -> LDS (in place, one fiber per thread)
//   mode 1      : LDS -> registers -> Y   (coalesced; atomic when Y is shared)
//
// A "fiber" of mode k is the N elements that differ only in index k
// (stride N^(D-k)).  Each thread owns whole fibers, so a mode pass is N
// loads, N*N FMAs with the factor held in registers, and N stores, with no
// barrier inside the pass.  X is read once and Y updated once per entry; all
// other traffic stays in LDS.  The contraction order differs from the
// reference's (1, 2, ..., D), which changes only the rounding.
//
// Small entries share a 256-thread workgroup: G = pow2ceil(N^(D-1)) or
// N^(D-1) threads (at most 256) per entry, 256/G entries per workgroup; large
// ones (see FastGeom) take a 1,024-thread workgroup each.
#pragma once

#include "kron_device.hpp"

namespace kron {

// An entry whose LDS alone exceeds 80 KB leaves room for one workgroup per
// CU (two above 40 KB); with 256 threads that is 4 (8) waves, too few to hide
// the passes' LDS and HBM latency.  Such entries get 1,024 threads when the factor fits the
// scalar registers (a thread then needs well under the 128 VGPRs of four
// waves per SIMD), else 512 (two waves per SIMD, 256 VGPRs).
template <typename T, int D, int N>
struct FastGeom {
    static constexpr int E = cpow(N, D);        // elements per entry
    static constexpr int F = E / N;              // fibers per mode
    // FP64 from 40 KB (d5n6: 0.198 -> 0.165 ms); in FP32 the extra threads
    // did not pay below 80 KB (d5n7, d6n5 3-5% slower)
    static constexpr bool BIG =
        (size_t)E * sizeof(T) > (sizeof(T) == 8 ? 40960 : 81920) && F >= 1024;
    static constexpr int THREADS = !BIG ? 256 : (size_t)N * N * sizeof(T) <= 200 ? 1024 : 512;
    // threads per entry: F rounded up to a power of two, or exactly F when
    // that packs more entries into the workgroup (F < 128 only, so P stays
    // above 1); threads from P * G on then sit idle (d3n6: 7 entries of 36
    // threads instead of 4 of 64)
    static constexpr int G2 = F >= THREADS ? THREADS : pow2ceil(F);
    static constexpr bool EXACT = !BIG && F < 128 && THREADS / F > THREADS / G2;
    // F in (64, 128] with no exact packing (d4n5: 125 fibers) takes a whole
    // workgroup: one entry per workgroup reads its factors into SGPRs
    static constexpr bool WHOLE = !BIG && !EXACT && F > 64 && F <= 128;
    static constexpr int G = EXACT ? F : WHOLE ? THREADS : G2;
    static constexpr int P = THREADS / G;        // entries per workgroup
    static constexpr size_t lds_bytes(size_t elem) { return (size_t)P * E * elem; }
};

// a[i][j] = A(i,j) (or A(j,i) when TRANS), A column-major with leading dim lda.
// With one entry per workgroup Ap is a constant-address-space pointer: a
// factor of at most SBYTES bytes is read with scalar loads into SGPRs
// (v_fma_f64 takes one scalar operand), saving N*N*2 VGPRs.  A larger one
// would overflow the SGPR file (spills through v_writelane / v_readlane), so
// it is read with vector loads of the same uniform addresses instead.
template <typename T, typename Acc, int N, bool TRANS, bool UNI, int SBYTES = 288,
          typename PtrT>
__device__ __forceinline__ void load_factor(Acc (&a)[N][N], PtrT Ap, int lda)
{
    if constexpr (UNI && (size_t)N * N * sizeof(Acc) > (size_t)SBYTES) {
        const gptr<const T> Av = (gptr<const T>)(uintptr_t)Ap;
#pragma unroll
        for (int j = 0; j < N; ++j)
#pragma unroll
            for (int i = 0; i < N; ++i)
                a[i][j] = Acc(TRANS ? Av[j + i * lda] : Av[i + j * lda]);
    } else {
#pragma unroll
        for (int j = 0; j < N; ++j)
#pragma unroll
            for (int i = 0; i < N; ++i)
                a[i][j] = Acc(TRANS ? Ap[j + i * lda] : Ap[i + j * lda]);
    }
}

// One pass over F fibers of stride S (fiber f holds the N elements
// base + j*S, base = f % S + (f / S) * S * N) owned by thread g of G.
template <typename Acc, int N, int S, int F, int G, typename LoadF, typename StoreF>
__device__ __forceinline__ void tile_pass(int g, const Acc (&a)[N][N], LoadF &&ld, StoreF &&st)
{
    constexpr int IT = (F + G - 1) / G;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int f = g + it * G;
        if ((F % G) == 0 || f < F) {
            const int base = (f % S) + (f / S) * (S * N);
            Acc in[N];
#pragma unroll
            for (int j = 0; j < N; ++j) in[j] = Acc(ld(base + j * S));
#pragma unroll
            for (int i = 0; i < N; ++i) {
                Acc s = a[i][0] * in[0];
#pragma unroll
                for (int j = 1; j < N; ++j) s = kfma(a[i][j], in[j], s);
                st(base + i * S, s);
            }
        }
    }
}

// Fiber iterations whose loads are batched together: about 16 values (32
// VGPRs in FP64) per batch.
template <int N>
constexpr int batch_iters()
{
    return 16 / N > 0 ? 16 / N : 1;
}

// tile_pass for passes that read HBM: the thread's fibers are taken in
// batches of B, and the loads of batch b+1 are issued before batch b is
// computed, so one batch's latency hides behind the previous batch's FMAs.
template <typename T, typename Acc, int N, int S, int F, int G, typename LoadF, typename StoreF>
__device__ __forceinline__ void tile_pass_in(int g, const Acc (&a)[N][N], LoadF &&ld, StoreF &&st)
{
    constexpr int IT = (F + G - 1) / G;
    constexpr int B = batch_iters<N>();
    constexpr int NB = (IT + B - 1) / B;
    T in[2][B][N];
    auto fetch = [&](int bi, T(&dst)[B][N]) {
#pragma unroll
        for (int q = 0; q < B; ++q) {
            const int f = g + (bi * B + q) * G;
            if (bi * B + q < IT && ((F % G) == 0 || f < F)) {
                const int base = (f % S) + (f / S) * (S * N);
#pragma unroll
                for (int j = 0; j < N; ++j) dst[q][j] = ld(base + j * S);
            }
        }
    };
    fetch(0, in[0]);
#pragma unroll
    for (int bi = 0; bi < NB; ++bi) {
        if (bi + 1 < NB) fetch(bi + 1, in[(bi + 1) & 1]);
#pragma unroll
        for (int q = 0; q < B; ++q) {
            const int f = g + (bi * B + q) * G;
            if (bi * B + q < IT && ((F % G) == 0 || f < F)) {
                const int base = (f % S) + (f / S) * (S * N);
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    Acc s = a[i][0] * Acc(in[bi & 1][q][0]);
#pragma unroll
                    for (int j = 1; j < N; ++j)
                        s = kfma(a[i][j], Acc(in[bi & 1][q][j]), s);
                    st(base + i * S, s);
                }
            }
        }
    }
}

// tile_pass whose outputs are added into Y (y[yi(e)] += value for tile
// element e, skipped where !ok(e)).  Without atomics the Y values are read in
// batches, batch b+1 issued before batch b is stored: the read-modify-write
// costs about one round trip instead of one per fiber (the compiler cannot
// move a Y load above an earlier Y store by itself; batches touch disjoint
// fibers, so this order is safe).
template <typename T, typename Acc, int N, int S, int F, int G, typename LoadF, typename YP,
          typename IdxF, typename OkF>
__device__ __forceinline__ void tile_pass_y(int g, const Acc (&a)[N][N], LoadF &&ld, YP y,
                                            IdxF &&yi, OkF &&ok, bool atomic)
{
    constexpr int IT = (F + G - 1) / G;
    constexpr int B = batch_iters<N>();
    constexpr int NB = (IT + B - 1) / B;
    if (atomic) {
        tile_pass<Acc, N, S, F, G>(g, a, static_cast<LoadF &&>(ld), [&](int e, Acc v) {
            if (ok(e)) atomic_add(y + yi(e), T(v));
        });
        return;
    }
    T yv[2][B][N];
    auto fetch = [&](int bi, T(&dst)[B][N]) {
#pragma unroll
        for (int q = 0; q < B; ++q) {
            const int f = g + (bi * B + q) * G;
            if (bi * B + q < IT && ((F % G) == 0 || f < F)) {
                const int base = (f % S) + (f / S) * (S * N);
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    const int e = base + i * S;
                    dst[q][i] = ok(e) ? y[yi(e)] : T(0);
                }
            }
        }
    };
    fetch(0, yv[0]);
#pragma unroll
    for (int bi = 0; bi < NB; ++bi) {
        if (bi + 1 < NB) fetch(bi + 1, yv[(bi + 1) & 1]);
#pragma unroll
        for (int q = 0; q < B; ++q) {
            const int f = g + (bi * B + q) * G;
            if (bi * B + q < IT && ((F % G) == 0 || f < F)) {
                const int base = (f % S) + (f / S) * (S * N);
                Acc in[N];
#pragma unroll
                for (int j = 0; j < N; ++j) in[j] = Acc(ld(base + j * S));
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    Acc s = a[i][0] * in[0];
#pragma unroll
                    for (int j = 1; j < N; ++j) s = kfma(a[i][j], in[j], s);
                    const int e = base + i * S;
                    if (ok(e)) y[yi(e)] = T(Acc(yv[bi & 1][q][i]) + s);
                }
            }
        }
    }
}

// One pass over the fibers of mode K of an N^D entry owned by thread g of G.
template <typename Acc, int D, int N, int K, int G, typename LoadF, typename StoreF>
__device__ __forceinline__ void fiber_pass(int g, const Acc (&a)[N][N], LoadF &&ld, StoreF &&st)
{
    tile_pass<Acc, N, cpow(N, D - K), cpow(N, D - 1), G>(g, a, static_cast<LoadF &&>(ld),
                                                           static_cast<StoreF &&>(st));
}

// LDS address of element i.  For N = 4 (and at least 64 elements) the two
// low 2-bit digits are XORed with the third: the 16 lanes that a 64-bit LDS
// access serves together then hit 16 distinct 8-byte bank slots for every
// mode's fiber pattern (strides 1, 4, 16, 64, ...).  Other N: identity.
template <int D, int N>
__device__ __forceinline__ int lds_index(int i)
{
    if constexpr (N == 4 && D >= 3) {
        const int h = (i >> 4) & 3;
        return i ^ (h | (h << 2));
    } else {
        return i;
    }
}

// Registers of one thread's mode-2 fibers (the first pass reads X here).
template <typename T, int D, int N, int G>
struct XFibers {
    static constexpr int F = cpow(N, D - 1);
    static constexpr int S = cpow(N, D - 2);
    static constexpr int IT = (F + G - 1) / G;
    T v[IT][N];
    template <typename PtrT>
    __device__ __forceinline__ void load(PtrT x, int g)
    {
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int f = g + it * G;
            if ((F % G) == 0 || f < F) {
                const int base = (f % S) + (f / S) * (S * N);
#pragma unroll
                for (int j = 0; j < N; ++j) v[it][j] = x[base + j * S];
            }
        }
    }
};

// Prefetching the next entry's X costs IT*N registers; worth it while that
// stays within 32 VGPRs.
template <typename T, int D, int N>
constexpr bool fast_pipelined()
{
    using X = XFibers<T, D, N, FastGeom<T, D, N>::G>;
    return D >= 2 && (size_t)X::IT * N * sizeof(T) <= 128;
}

// Grouped calls (GroupOrder, kron_internal.hpp): workgroup w takes the
// positions [w*chunk, (w+1)*chunk) of the Y-target order.

// Grouping needs one entry per workgroup and the mode-1 outputs of a thread
// (the running Y sum) in registers: 32 VGPRs, or 80 / 160 when the entry's
// LDS already limits a CU to two workgroups / one (a wave may then hold 256
// / 512 VGPRs).  Without it every element of every entry is an atomic add:
// at d=6 n=5, 65,536 entries, a billion of them.
template <typename T, int D, int N>
constexpr bool fast_groupable()
{
    using Geo = FastGeom<T, D, N>;
    constexpr size_t lds = Geo::lds_bytes(sizeof(T));
    // 1,024 threads: four waves per SIMD, 128 VGPRs each; 512: two, 256
    constexpr size_t budget = Geo::THREADS == 1024 ? 160
                              : Geo::THREADS == 512 ? 320
                              : lds > 81920        ? 640
                              : lds > 40960        ? 320
                                                   : 128;
    return D >= 2 && Geo::P == 1 &&
           (size_t)((Geo::F + Geo::G - 1) / Geo::G) * N * sizeof(T) <= budget;
}

// Persistent workgroups walk the batch with stride gridDim.x * P.  With
// PIPE the loads of the next entry's X are issued before the current entry's
// LDS passes, so HBM latency overlaps the on-chip work.
// GROUPED: entries are visited in Y-target order (GroupOrder); the mode-1
// results of consecutive entries with the same target are summed in registers
// and added to Y once per run ("batched atomicAdd merges").
template <typename T, typename Acc, int D, int N, Layout LAY, bool SHARED, bool TRANS,
          bool GROUPED>
__global__ void __launch_bounds__((FastGeom<T, D, N>::THREADS))
kron_fast_kernel(Entries<T, LAY> ent, int lda, int batch, GroupOrder ord)
{
    using Geo = FastGeom<T, D, N>;
    constexpr int E = Geo::E;
    constexpr int G = Geo::G;
    constexpr int P = Geo::P;
    constexpr bool PIPE = fast_pipelined<T, D, N>();
    constexpr bool UNI = (P == 1);  // one entry per workgroup: pointers are uniform
    extern __shared__ __align__(16) unsigned char kron_smem[];
    T *smem = reinterpret_cast<T *>(kron_smem);

    const int tid = threadIdx.x;
    const int slot = (P == 1) ? 0 : tid / G;
    const int g0 = (P == 1) ? tid : tid % G;
    // threads past P * G (G not a power of two) own no entry: entry() gives
    // them batch, which every work guard (b < batch) skips
    const bool idle = P * G < Geo::THREADS && slot >= P;
    const int stride = gridDim.x * P;
    Acc a[N][N];
    static_assert(!GROUPED || fast_groupable<T, D, N>(), "grouping needs P == 1, D >= 2");

    // positions r in [rbeg, rend) with step rstep; entry(r) is the batch index
    int rbeg, rend, rstep;
    if constexpr (GROUPED) {
        rbeg = blockIdx.x * ord.chunk;
        rend = min(batch, rbeg + ord.chunk);
        rstep = 1;
    } else {
        rbeg = blockIdx.x * P;
        rend = batch;
        rstep = stride;
    }
    auto entry = [&](int r) -> int {
        if constexpr (GROUPED)
            return to_const(ord.perm)[r];
        else
            return idle ? batch : r + slot;
    };

    if constexpr (D == 1) {
        const int g = g0;
        for (int b = idle ? batch : blockIdx.x * P + slot; b < batch; b += stride) {
            load_factor<T, Acc, N, TRANS, UNI>(a, ent.template factor_g<UNI>(b, 0), lda);
            const auto x = ent.template x_g<UNI>(b);
            const auto y = ent.template y_g<UNI>(b);
            fiber_pass<Acc, D, N, 1, G>(
                g, a, [&](int i) { return x[i]; },
                [&](int i, Acc v) { y_accumulate<SHARED, T, Acc>(y, i, v); });
        }
    } else {
        T *buf = smem + slot * E;
        auto L = [&](int i) -> T & { return buf[lds_index<D, N>(i)]; };
        XFibers<T, D, N, (PIPE ? G : 1 << 30)> nxt;  // empty unless PIPE
        if constexpr (PIPE) {
            if (rbeg < rend) {
                const int b = entry(rbeg);
                if (b < batch) nxt.load(ent.template x_g<UNI>(b), g0);
            }
        }
        constexpr int F1 = cpow(N, D - 1);   // mode-1 fibers (= stride of mode 1)
        constexpr int IT1 = (F1 + G - 1) / G;
        Acc acc[GROUPED ? IT1 : 1][N];       // running Y sum of the current run
        if constexpr (GROUPED) {
#pragma unroll
            for (int q = 0; q < IT1; ++q)
#pragma unroll
                for (int i = 0; i < N; ++i) acc[q][i] = 0;
        }
        for (int r = rbeg; r < rend; r += rstep) {
            const int b = entry(r);
            const bool active = GROUPED || b < batch;
            // Keep the lane index opaque per iteration: otherwise every LDS and
            // global address of every mode pass is hoisted out of the loop and
            // held in ~100 VGPRs, halving occupancy.
            int g = g0;
            asm volatile("" : "+v"(g));
            if constexpr (PIPE) {
                // mode 2 from the registers loaded one iteration ahead; the
                // same registers then receive the next entry's X, whose loads
                // stay in flight during the remaining passes.
                if (active) {
                    load_factor<T, Acc, N, TRANS, UNI>(a, ent.template factor_g<UNI>(b, 1), lda);
                    constexpr int S2 = cpow(N, D - 2);
                    constexpr int F = cpow(N, D - 1);
#pragma unroll
                    for (int q = 0; q < XFibers<T, D, N, G>::IT; ++q) {
                        const int f = g + q * G;
                        if ((F % G) == 0 || f < F) {
                            const int base = (f % S2) + (f / S2) * (S2 * N);
#pragma unroll
                            for (int i = 0; i < N; ++i) {
                                Acc s = a[i][0] * Acc(nxt.v[q][0]);
#pragma unroll
                                for (int j = 1; j < N; ++j)
                                    s = kfma(a[i][j], Acc(nxt.v[q][j]), s);
                                L(base + i * S2) = T(s);
                            }
                        }
                    }
                }
                if (r + rstep < rend) {
                    const int bn = entry(r + rstep);
                    if (GROUPED || bn < batch) nxt.load(ent.template x_g<UNI>(bn), g);
                }
            } else {
                if (active) {
                    load_factor<T, Acc, N, TRANS, UNI>(a, ent.template factor_g<UNI>(b, 1), lda);
                    const auto x = ent.template x_g<UNI>(b);
                    fiber_pass<Acc, D, N, 2, G>(
                        g, a, [&](int i) { return x[i]; }, [&](int i, Acc v) { L(i) = T(v); });
                }
            }
            __syncthreads();
            for_modes_desc<D, 3>([&](auto kc) {
                constexpr int K = decltype(kc)::value;
                if (active) {
                    load_factor<T, Acc, N, TRANS, UNI>(a, ent.template factor_g<UNI>(b, K - 1),
                                                       lda);
                    fiber_pass<Acc, D, N, K, G>(
                        g, a, [&](int i) { return L(i); }, [&](int i, Acc v) { L(i) = T(v); });
                }
                __syncthreads();
            });
            if constexpr (GROUPED) {
                load_factor<T, Acc, N, TRANS, UNI>(a, ent.template factor_g<UNI>(b, 0), lda);
#pragma unroll
                for (int q = 0; q < IT1; ++q) {
                    const int f = g + q * G;
                    if ((F1 % G) == 0 || f < F1) {
                        Acc in[N];
#pragma unroll
                        for (int j = 0; j < N; ++j) in[j] = Acc(L(f + j * F1));
#pragma unroll
                        for (int i = 0; i < N; ++i) {
                            Acc s = a[i][0] * in[0];
#pragma unroll
                            for (int j = 1; j < N; ++j) s = kfma(a[i][j], in[j], s);
                            acc[q][i] += s;
                        }
                    }
                }
                const auto keys = to_const(ord.keys);
                const unsigned long long key = keys[r];
                if (r + 1 >= rend || keys[r + 1] != key) {
                    const auto y = to_global(reinterpret_cast<T *>(key));
#pragma unroll
                    for (int q = 0; q < IT1; ++q) {
                        const int f = g + q * G;
                        if ((F1 % G) == 0 || f < F1) {
#pragma unroll
                            for (int i = 0; i < N; ++i) {
                                atomic_add(y + (f + i * F1), T(acc[q][i]));
                                acc[q][i] = 0;
                            }
                        }
                    }
                }
            } else {
                if (active) {
                    load_factor<T, Acc, N, TRANS, UNI>(a, ent.template factor_g<UNI>(b, 0), lda);
                    const auto y = ent.template y_g<UNI>(b);
                    fiber_pass<Acc, D, N, 1, G>(
                        g, a, [&](int i) { return L(i); },
                        [&](int i, Acc v) { y_accumulate<SHARED, T, Acc>(y, i, v); });
                }
            }
            __syncthreads();  // LDS is rewritten by the next entry
        }
    }
}

// Largest N with a compiled fast kernel for dimension D and element size.
// The bound keeps N^D elements of T within 160 KiB of LDS.
template <typename T>
constexpr int fast_max_n(int D)
{
    return (sizeof(T) == 8) ? (D == 6 ? 5 : D == 5 ? 7 : 8)
                            : (D == 6 ? 5 : 8);
}

}  // namespace kron

// (after load_factor, which the plane kernel uses)
#include "kron_plane.hpp"

namespace kron {

// The d = 6, n = 4 entry runs the two-modes-per-pass kernel (kron_plane.hpp).
template <int D, int N>
constexpr bool fast_uses_plane()
{
    return D == 6 && N == 4;
}

template <typename T, typename Acc, int D, int N, Layout LAY, bool SHARED, bool TRANS,
          bool GROUPED>
constexpr auto fast_kernel_for()
{
    if constexpr (fast_uses_plane<D, N>())
        return kron_plane_kernel<T, Acc, LAY, SHARED, TRANS, GROUPED>;
    else
        return kron_fast_kernel<T, Acc, D, N, LAY, SHARED, TRANS, GROUPED>;
}

template <typename T, typename Acc, int D, int N, Layout LAY, bool SHARED, bool TRANS>
kron_status launch_fast_one(const Entries<T, LAY> &ent, int lda, int batch, hipStream_t stream,
                            const GroupOrder *ord)
{
    using Geo = FastGeom<T, D, N>;
    constexpr int THREADS = fast_uses_plane<D, N>() ? 256 : Geo::THREADS;
    const size_t lds = (D == 1)                    ? 0
                       : fast_uses_plane<D, N>() ? (size_t)kPlaneLds * sizeof(T)
                                                 : Geo::lds_bytes(sizeof(T));
    if ((long long)lds > (long long)max_lds_bytes()) return KRON_ENOTSUP;
    constexpr bool CAN_GROUP = SHARED && fast_groupable<T, D, N>();
    const bool grouped = CAN_GROUP && ord != nullptr;
    const void *kern;
    if constexpr (CAN_GROUP) {
        kern = grouped ? reinterpret_cast<const void *>(
                             fast_kernel_for<T, Acc, D, N, LAY, SHARED, TRANS, true>())
                       : reinterpret_cast<const void *>(
                             fast_kernel_for<T, Acc, D, N, LAY, SHARED, TRANS, false>());
    } else {
        kern = reinterpret_cast<const void *>(
            fast_kernel_for<T, Acc, D, N, LAY, SHARED, TRANS, false>());
    }
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds);
        if (e != hipSuccess) {
            set_last_hip_error((int)e);
            return KRON_EHIP;
        }
    }
    const long long resident = persistent_blocks(kern, lds, THREADS);
    GroupOrder o{nullptr, nullptr, 0};
    long long nblocks = ((long long)batch + Geo::P - 1) / Geo::P;
    if (grouped) {
        // equal contiguous chunks of the target order, one per resident block
        const long long blocks = resident > 0 ? resident : 1024;
        const long long chunk = ((long long)batch + blocks - 1) / blocks;
        o = *ord;
        o.chunk = (int)chunk;
        nblocks = ((long long)batch + chunk - 1) / chunk;
    } else if (resident > 0 && nblocks > resident) {
        nblocks = resident;
    }
    if (nblocks > 0x7fffffffLL) return KRON_EINVAL;
    if constexpr (CAN_GROUP) {
        if (grouped)
            hipLaunchKernelGGL((fast_kernel_for<T, Acc, D, N, LAY, SHARED, TRANS, true>()),
                               dim3((unsigned)nblocks), dim3(THREADS), lds, stream, ent, lda, batch,
                               o);
        else
            hipLaunchKernelGGL((fast_kernel_for<T, Acc, D, N, LAY, SHARED, TRANS, false>()),
                               dim3((unsigned)nblocks), dim3(THREADS), lds, stream, ent, lda, batch,
                               o);
    } else {
        hipLaunchKernelGGL((fast_kernel_for<T, Acc, D, N, LAY, SHARED, TRANS, false>()),
                           dim3((unsigned)nblocks), dim3(THREADS), lds, stream, ent, lda, batch, o);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_last_hip_error((int)e);
        return KRON_EHIP;
    }
    return KRON_OK;
}

// Whether (D, N) has a grouped (Y-merging) fast kernel for element type T.
template <typename T>
bool fast_groupable_rt(int D, int N)
{
#define KRON_G_CASE(DD, NN) \
    if (D == DD && N == NN) return fast_groupable<T, DD, NN>();
#define KRON_G_D(DD) \
    KRON_G_CASE(DD, 1) KRON_G_CASE(DD, 2) KRON_G_CASE(DD, 3) KRON_G_CASE(DD, 4) \
    KRON_G_CASE(DD, 5) KRON_G_CASE(DD, 6) KRON_G_CASE(DD, 7) KRON_G_CASE(DD, 8)
    KRON_G_D(2) KRON_G_D(3) KRON_G_D(4) KRON_G_D(5) KRON_G_D(6)
#undef KRON_G_D
#undef KRON_G_CASE
    return false;
}

// Runtime (D, N) -> compiled kernel.  KRON_ENOTSUP when N > fast_max_n.
template <typename T, typename Acc, Layout LAY, bool SHARED, bool TRANS>
kron_status launch_fast(int D, int N, const Entries<T, LAY> &ent, int lda, int batch,
                        hipStream_t stream, const GroupOrder *ord = nullptr)
{
    if (N < 1 || N > fast_max_n<T>(D)) return KRON_ENOTSUP;
#define KRON_FAST_CASE(DD, NN)                                                           \
    if (D == DD && N == NN) {                                                            \
        if constexpr (NN <= fast_max_n<T>(DD))                                           \
            return launch_fast_one<T, Acc, DD, NN, LAY, SHARED, TRANS>(ent, lda, batch,  \
                                                                       stream, ord);     \
        else                                                                             \
            return KRON_ENOTSUP;                                                         \
    }
#define KRON_FAST_D(DD)                                                                  \
    KRON_FAST_CASE(DD, 1) KRON_FAST_CASE(DD, 2) KRON_FAST_CASE(DD, 3)                    \
    KRON_FAST_CASE(DD, 4) KRON_FAST_CASE(DD, 5) KRON_FAST_CASE(DD, 6)                    \
    KRON_FAST_CASE(DD, 7) KRON_FAST_CASE(DD, 8)
    KRON_FAST_D(1)
    KRON_FAST_D(2)
    KRON_FAST_D(3)
    KRON_FAST_D(4)
    KRON_FAST_D(5)
    KRON_FAST_D(6)
#undef KRON_FAST_D
#undef KRON_FAST_CASE
    return KRON_ENOTSUP;
}

}  // namespace kron
