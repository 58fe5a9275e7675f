// kron_mixed.hip -- mixed-(d, n) batched Kronecker matvec in one persistent
// launch.
//
// The reference runs one kronmult{d}_vbatched launch per shape class
// (kronmult_vbatched.hpp:19-248, one (m_k, n_k) set per launch).  A caller
// with many classes (a sparse-grid operator with 28 (d, n) classes, say) then
// pays one launch, one ramp-up and one drain per class, and small classes
// leave most of the GPU idle.  Here every entry of every class is a work item
// of one persistent kernel:
//
//   * items: a Y-target group of an on-chip class (its entries run one after
//     another in one workgroup, so Y is updated with plain read-modify-write),
//     P entries of a small class (packed into one workgroup, atomic Y), or
//     one tile of a large class's slow-mode pass (X -> scratch) or fast-mode
//     pass (scratch -> Y, for a Y-target group);
//   * with large classes present an apply is two launches of the kernel:
//     the first holds every slow-mode tile, the second every fast-mode tile,
//     and the whole-entry items are split between them to balance the two;
//   * the items are sorted by estimated cost, largest first, and taken from
//     an atomic counter (longest-processing-time-first list scheduling);
//   * each class runs a compile-time (D, N) path with a tile and thread
//     mapping variant chosen by the autotuner (threads per entry for on-chip
//     classes, the slow/fast mode split for large ones);
//   * the counter resets itself (the last workgroup out clears it), so the
//     launch needs no memset and replays inside a hipGraph.
//
// Y update semantics are the reference's (Y += K X, element-wise atomic where
// entries may share Y): the plan proves Y targets pairwise identical or
// disjoint before it drops atomics for a class.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "kron_fast.hpp"
#include "kron_mixed.hpp"

namespace kron {

// ---------------------------------------------------------------------------
// Compile-time paths
// ---------------------------------------------------------------------------

// Slow/fast split of a large entry: pass 1 contracts modes 1..h over chunks of
// c1 consecutive fast indices (an S x c1 tile, S = N^h), pass 2 contracts
// modes h+1..D over chunks of c2 slow rows (a c2 x F tile, F = N^(D-h)).
struct MixSplit {
    int h, c1, c2, cost;
};

constexpr MixSplit mix_split(int D, int N, int L, int v)
{
    MixSplit cand[kMaxDim] = {};
    int nc = 0;
    for (int h = 1; h < D; ++h) {
        const int S = cpow(N, h), F = cpow(N, D - h);
        if (S > L || F > L) continue;
        const int k1 = (F + L / S - 1) / (L / S);
        const int c1 = (F + k1 - 1) / k1;
        const int k2 = (S + L / F - 1) / (L / F);
        const int c2 = (S + k2 - 1) / k2;
        // short column runs of pass 1 coalesce poorly; ties go to the most
        // balanced split
        const int bal = 2 * h > D ? 2 * h - D : D - 2 * h;
        const int cost = 16 * (k1 + k2 + (c1 < 16 ? 4 * k1 : 0)) + bal;
        cand[nc++] = MixSplit{h, c1, c2, cost};
    }
    for (int i = 0; i < nc; ++i)
        for (int j = i + 1; j < nc; ++j)
            if (cand[j].cost < cand[i].cost) {
                const MixSplit t = cand[i];
                cand[i] = cand[j];
                cand[j] = t;
            }
    return v < nc ? cand[v] : MixSplit{0, 0, 0, -1};
}

template <typename T, int D, int N, int V>
struct MixPath {
    static constexpr int L = kMixLdsBytes / (int)sizeof(T);  // LDS elements
    static constexpr int E = cpow(N, D);
    static constexpr int F = E / N;
    static constexpr int G0 = F >= kMixThreads ? kMixThreads : pow2ceil(F);
    static constexpr bool BIG = E > L;
    // on-chip classes: variant v halves the threads per entry v times
    static constexpr int G = BIG ? kMixThreads : (G0 >> V) > 0 ? (G0 >> V) : 1;
    static constexpr int P = kMixThreads / G;
    static constexpr MixSplit SP = mix_split(D, N, L, BIG ? V : 0);
    static constexpr bool VALID =
        BIG ? SP.cost >= 0 : ((G0 >> V) >= 1 && (size_t)P * E <= (size_t)L);
    static constexpr size_t lds_bytes()
    {
        if (BIG) {
            const int S = cpow(N, SP.h), Fs = cpow(N, D - SP.h);
            return (size_t)std::max(S * SP.c1, SP.c2 * Fs) * sizeof(T);
        }
        return D == 1 ? 0 : (size_t)P * E * sizeof(T);
    }
    static constexpr long long scratch_elems() { return BIG ? E : 0; }
};

// Largest factor (bytes) held in SGPRs: the kernel keeps more scalar state
// live than a single-shape kernel does.
constexpr int kMixSBytes = 128;

// One entry per workgroup, the entry resident in LDS (the fast kernel's mode
// order: 2 from HBM, D..3 in place, 1 into Y).  The entries of the item share
// one Y target when the class updates Y without atomics; they run in order,
// and each thread updates the same Y elements every time.
template <typename T, typename Acc, int D, int N>
__device__ __forceinline__ void mix_onchip(const MixedClass<T> &c, const int *order, int count,
                                           T *smem, int tid)
{
    const cptr<int> ord = to_const(order);
    const Entries<T, Layout::Ptr> ent{c.pA, c.pX, c.pY, D};
    const bool atomic = c.atomic_y != 0;
    Acc a[N][N];
    auto L = [&](int i) -> T & { return smem[lds_index<D, N>(i)]; };
    for (int i = 0; i < count; ++i) {
        const int b = ord[i];
        int g = tid;
        asm volatile("" : "+v"(g));
        load_factor<T, Acc, N, false, true, kMixSBytes>(a, ent.template factor_g<true>(b, 1), c.lda);
        constexpr int F = cpow(N, D - 1);
        const auto x = ent.template x_g<true>(b);
        tile_pass_in<T, Acc, N, cpow(N, D - 2), F, kMixThreads>(
            g, a, [&](int e) { return x[e]; }, [&](int e, Acc v) { L(e) = T(v); });
        __syncthreads();
        for_modes_desc<D, 3>([&](auto kc) {
            constexpr int K = decltype(kc)::value;
            load_factor<T, Acc, N, false, true, kMixSBytes>(a, ent.template factor_g<true>(b, K - 1), c.lda);
            fiber_pass<Acc, D, N, K, kMixThreads>(
                g, a, [&](int e) { return L(e); }, [&](int e, Acc v) { L(e) = T(v); });
            __syncthreads();
        });
        load_factor<T, Acc, N, false, true, kMixSBytes>(a, ent.template factor_g<true>(b, 0), c.lda);
        const auto y = ent.template y_g<true>(b);
        tile_pass_y<T, Acc, N, F, F, kMixThreads>(
            g, a, [&](int e) { return L(e); }, y, [](int e) { return e; },
            [](int) { return true; }, atomic);
        __syncthreads();
    }
}

// P = 256 / G entries per workgroup, G threads each (small entries); Y is
// updated atomically (entries of one item may share Y).
template <typename T, typename Acc, int D, int N, int G>
__device__ __forceinline__ void mix_packed(const MixedClass<T> &c, const int *order, int count,
                                           T *smem, int tid)
{
    constexpr int E = cpow(N, D);
    const Entries<T, Layout::Ptr> ent{c.pA, c.pX, c.pY, D};
    const int slot = tid / G;
    const bool active = slot < count;
    const int b = active ? to_global(order)[slot] : 0;
    const int g = tid % G;
    Acc a[N][N];
    if constexpr (D == 1) {
        if (active) {
            load_factor<T, Acc, N, false, false>(a, ent.template factor_g<false>(b, 0), c.lda);
            const auto x = ent.template x_g<false>(b);
            const auto y = ent.template y_g<false>(b);
            fiber_pass<Acc, 1, N, 1, G>(
                g, a, [&](int e) { return x[e]; }, [&](int e, Acc v) { atomic_add(y + e, T(v)); });
        }
    } else {
        T *buf = smem + slot * E;
        auto L = [&](int i) -> T & { return buf[lds_index<D, N>(i)]; };
        if (active) {
            load_factor<T, Acc, N, false, false>(a, ent.template factor_g<false>(b, 1), c.lda);
            const auto x = ent.template x_g<false>(b);
            fiber_pass<Acc, D, N, 2, G>(
                g, a, [&](int e) { return x[e]; }, [&](int e, Acc v) { L(e) = T(v); });
        }
        __syncthreads();
        for_modes_desc<D, 3>([&](auto kc) {
            constexpr int K = decltype(kc)::value;
            if (active) {
                load_factor<T, Acc, N, false, false>(a, ent.template factor_g<false>(b, K - 1),
                                                     c.lda);
                fiber_pass<Acc, D, N, K, G>(
                    g, a, [&](int e) { return L(e); }, [&](int e, Acc v) { L(e) = T(v); });
            }
            __syncthreads();
        });
        if (active) {
            load_factor<T, Acc, N, false, false>(a, ent.template factor_g<false>(b, 0), c.lda);
            const auto y = ent.template y_g<false>(b);
            fiber_pass<Acc, D, N, 1, G>(
                g, a, [&](int e) { return L(e); }, [&](int e, Acc v) { atomic_add(y + e, T(v)); });
        }
        __syncthreads();
    }
}

// A large entry in two passes through its own scratch vector W (n^d
// elements; entry b of the class at scratch + wofs + b * n^d).  Pass 1
// contracts the slow modes 1..H of one S x C1 column tile (X -> W), pass 2
// the fast modes H+1..D of one C2 x F slab tile (W -> Y); each tile is a work
// item, and the two passes are separate launches (the launch boundary makes
// W visible across the chip).  Tile elements past the end of the entry
// compute on stale LDS and are not stored.
template <typename T, typename Acc, int D, int N, int H, int C1>
__device__ __forceinline__ void mix_pass1(const MixedClass<T> &c, int b, int chunk, T *smem,
                                          T *scratch, int tid)
{
    constexpr int S = cpow(N, H), F = cpow(N, D - H);
    constexpr int T1 = S * C1;
    const Entries<T, Layout::Ptr> ent{c.pA, c.pX, c.pY, D};
    const auto W = to_global(scratch + c.wofs + (long long)b * (S * F));
    const auto x = ent.template x_g<true>(b);
    const int c0 = chunk * C1;
    Acc a[N][N];
    auto L = [&](int i) -> T & { return smem[i]; };
    const int g = tid;
    // tile element t = s * C1 + c is entry element s * F + c0 + c
    auto gidx = [&](int t) { return (t / C1) * F + c0 + t % C1; };
    auto valid = [&](int t) { return c0 + t % C1 < F; };
    if constexpr (H == 1) {
        load_factor<T, Acc, N, false, true, kMixSBytes>(a, ent.template factor_g<true>(b, 0),
                                                        c.lda);
        tile_pass_in<T, Acc, N, C1, T1 / N, kMixThreads>(
            g, a, [&](int t) { return valid(t) ? x[gidx(t)] : T(0); },
            [&](int t, Acc v) {
                if (valid(t)) W[gidx(t)] = T(v);
            });
    } else {
        load_factor<T, Acc, N, false, true, kMixSBytes>(a, ent.template factor_g<true>(b, H - 1),
                                                        c.lda);
        tile_pass_in<T, Acc, N, C1, T1 / N, kMixThreads>(
            g, a, [&](int t) { return valid(t) ? x[gidx(t)] : T(0); },
            [&](int t, Acc v) { L(t) = T(v); });
        __syncthreads();
        for_modes_desc<H - 1, 2>([&](auto kc) {
            constexpr int K = decltype(kc)::value;
            load_factor<T, Acc, N, false, true, kMixSBytes>(
                a, ent.template factor_g<true>(b, K - 1), c.lda);
            tile_pass<Acc, N, cpow(N, H - K) * C1, T1 / N, kMixThreads>(
                g, a, [&](int t) { return L(t); }, [&](int t, Acc v) { L(t) = T(v); });
            __syncthreads();
        });
        load_factor<T, Acc, N, false, true, kMixSBytes>(a, ent.template factor_g<true>(b, 0),
                                                        c.lda);
        tile_pass<Acc, N, cpow(N, H - 1) * C1, T1 / N, kMixThreads>(
            g, a, [&](int t) { return L(t); },
            [&](int t, Acc v) {
                if (valid(t)) W[gidx(t)] = T(v);
            });
    }
    __syncthreads();  // LDS is rewritten by the next item
}

// Pass 2 of slab tile `chunk` for the entries of one item, in order (they
// share one Y target when the class updates Y without atomics, and this
// workgroup owns that slab of it).
template <typename T, typename Acc, int D, int N, int H, int C2>
__device__ __forceinline__ void mix_pass2(const MixedClass<T> &c, const int *order, int count,
                                          int chunk, T *smem, T *scratch, int tid)
{
    const cptr<int> ord = to_const(order);
    constexpr int S = cpow(N, H), F = cpow(N, D - H);
    constexpr int T2 = C2 * F;
    const Entries<T, Layout::Ptr> ent{c.pA, c.pX, c.pY, D};
    const bool atomic = c.atomic_y != 0;
    const int r0 = chunk * C2;
    const int nv = (S - r0 < C2 ? S - r0 : C2) * F;  // valid tile elements
    Acc a[N][N];
    auto L = [&](int i) -> T & { return smem[i]; };
    for (int i = 0; i < count; ++i) {
        const int b = ord[i];
        int g = tid;
        asm volatile("" : "+v"(g));
        // tile element t = r * F + f is entry element r0 * F + t
        const auto w = to_global(scratch + c.wofs + (long long)b * (S * F) + (long long)r0 * F);
        const auto yt = ent.template y_g<true>(b) + (long long)r0 * F;
        auto ok = [&](int t) { return t < nv; };
        auto ident = [](int t) { return t; };
        if constexpr (D - H == 1) {
            load_factor<T, Acc, N, false, true, kMixSBytes>(
                a, ent.template factor_g<true>(b, D - 1), c.lda);
            tile_pass_y<T, Acc, N, 1, T2 / N, kMixThreads>(
                g, a, [&](int t) { return t < nv ? w[t] : T(0); }, yt, ident, ok, atomic);
        } else {
            load_factor<T, Acc, N, false, true, kMixSBytes>(
                a, ent.template factor_g<true>(b, D - 1), c.lda);
            tile_pass_in<T, Acc, N, 1, T2 / N, kMixThreads>(
                g, a, [&](int t) { return t < nv ? w[t] : T(0); },
                [&](int t, Acc v) { L(t) = T(v); });
            __syncthreads();
            for_modes_desc<D - 1, H + 2>([&](auto kc) {
                constexpr int K = decltype(kc)::value;
                load_factor<T, Acc, N, false, true, kMixSBytes>(
                    a, ent.template factor_g<true>(b, K - 1), c.lda);
                tile_pass<Acc, N, cpow(N, D - K), T2 / N, kMixThreads>(
                    g, a, [&](int t) { return L(t); }, [&](int t, Acc v) { L(t) = T(v); });
                __syncthreads();
            });
            load_factor<T, Acc, N, false, true, kMixSBytes>(a, ent.template factor_g<true>(b, H),
                                                            c.lda);
            tile_pass_y<T, Acc, N, cpow(N, D - H - 1), T2 / N, kMixThreads>(
                g, a, [&](int t) { return L(t); }, yt, ident, ok, atomic);
        }
        __syncthreads();  // LDS is rewritten next
    }
}

// chunk: -1 for a whole-entry item, c >= 0 for pass-1 tile c, -2 - c for
// pass-2 tile c of a large entry.
template <typename T, typename Acc, int D, int N, int V>
__device__ __forceinline__ void mix_run(const MixedClass<T> &c, const int *ord, int count,
                                        int chunk, T *smem, T *scratch, int tid)
{
    using Pth = MixPath<T, D, N, V>;
    if constexpr (!Pth::VALID) {
        return;
    } else if constexpr (Pth::BIG) {
        if (chunk >= 0)
            mix_pass1<T, Acc, D, N, Pth::SP.h, Pth::SP.c1>(c, to_const(ord)[0], chunk, smem,
                                                           scratch, tid);
        else
            mix_pass2<T, Acc, D, N, Pth::SP.h, Pth::SP.c2>(c, ord, count, -2 - chunk, smem,
                                                           scratch, tid);
    } else if constexpr (Pth::P == 1 && D >= 2) {
        mix_onchip<T, Acc, D, N>(c, ord, count, smem, tid);
    } else {
        mix_packed<T, Acc, D, N, Pth::G>(c, ord, count, smem, tid);
    }
}

// The compiled (D, N, variant) set: D = 1..6, N = 1..8 with N^D <= 65536.
#define KRON_MIX_FOR_EACH(X)                                                                  \
    X(1, 1) X(1, 2) X(1, 3) X(1, 4) X(1, 5) X(1, 6) X(1, 7) X(1, 8)                           \
    X(2, 1) X(2, 2) X(2, 3) X(2, 4) X(2, 5) X(2, 6) X(2, 7) X(2, 8)                           \
    X(3, 1) X(3, 2) X(3, 3) X(3, 4) X(3, 5) X(3, 6) X(3, 7) X(3, 8)                           \
    X(4, 1) X(4, 2) X(4, 3) X(4, 4) X(4, 5) X(4, 6) X(4, 7) X(4, 8)                           \
    X(5, 1) X(5, 2) X(5, 3) X(5, 4) X(5, 5) X(5, 6) X(5, 7) X(5, 8)                           \
    X(6, 1) X(6, 2) X(6, 3) X(6, 4) X(6, 5) X(6, 6)

constexpr int mix_code(int D, int N, int V) { return (D * 16 + N) * kMixVariants + V; }

template <typename T>
__global__ void __launch_bounds__(kMixThreads, 2)
kron_mixed_kernel(const MixedClass<T> *cls, const MixedItem *items, const int *order,
                  int nitems, unsigned *queue, T *scratch)
{
    extern __shared__ __align__(16) unsigned char kron_smem[];
    __shared__ int s_next;
    T *smem = reinterpret_cast<T *>(kron_smem);
    // the next item's index is requested one item ahead, so the atomic's
    // round trip overlaps the current item
    unsigned ahead = 0;
    if (threadIdx.x == 0) ahead = atomicAdd(queue, 1u);
    for (;;) {
        if (threadIdx.x == 0) {
            s_next = (int)ahead;
            ahead = atomicAdd(queue, 1u);
        }
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(s_next);
        __syncthreads();  // s_next is rewritten by the next dequeue
        if (it >= nitems) break;
        // the item and its class, field by field through scalar loads
        const auto ip = to_const(items) + it;
        const int ci = ip->cls, count = ip->count, chunk = ip->chunk;
        const int *ord = order + ip->first;
        const auto cp = to_const(cls) + ci;
        MixedClass<T> c;
        c.pA = cp->pA;
        c.pX = cp->pX;
        c.pY = cp->pY;
        c.d = cp->d;
        c.n = cp->n;
        c.lda = cp->lda;
        c.path = cp->path;
        c.atomic_y = cp->atomic_y;
        c.wofs = cp->wofs;
        // lane-dependent index math of every path derives from this per-item
        // opaque copy, so none of it is hoisted out of the loop (where it
        // would stay live across all paths)
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        switch (c.path) {
#define KRON_MIX_CASE_V(D, N, V)                                                              \
    case mix_code(D, N, V):                                                                   \
        mix_run<T, T, D, N, V>(c, ord, count, chunk, smem, scratch, tid);                     \
        break;
#define KRON_MIX_CASE(D, N) KRON_MIX_CASE_V(D, N, 0) KRON_MIX_CASE_V(D, N, 1)
            KRON_MIX_FOR_EACH(KRON_MIX_CASE)
#undef KRON_MIX_CASE
#undef KRON_MIX_CASE_V
        default: break;
        }
    }
    // every workgroup has taken its last (empty) items; the last one out
    // clears the counters for the next launch on the stream
    if (threadIdx.x == 0) {
        __threadfence();
        const unsigned done = atomicAdd(queue + 1, 1u);
        if (done == gridDim.x - 1) {
            atomicExch(queue, 0u);
            atomicExch(queue + 1, 0u);
        }
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

template <typename T>
static bool path_info(int D, int N, int V, MixPathInfo &out)
{
#define KRON_MIX_INFO_V(DD, NN, VV)                                                           \
    if (D == DD && N == NN && V == VV) {                                                      \
        using Pth = MixPath<T, DD, NN, VV>;                                                   \
        out.valid = Pth::VALID;                                                               \
        out.big = Pth::BIG;                                                                   \
        out.packed = !Pth::BIG && (Pth::P > 1 || DD == 1);                                    \
        out.per_item = Pth::BIG ? 1 : Pth::P;                                                 \
        out.lds = Pth::VALID ? Pth::lds_bytes() : 0;                                          \
        out.scratch = Pth::VALID ? Pth::scratch_elems() : 0;                                  \
        out.split = Pth::BIG ? Pth::SP.h : 0;                                                 \
        out.tiles1 = Pth::BIG ? (cpow(NN, DD - Pth::SP.h) + Pth::SP.c1 - 1) / Pth::SP.c1 : 0;  \
        out.tiles2 = Pth::BIG ? (cpow(NN, Pth::SP.h) + Pth::SP.c2 - 1) / Pth::SP.c2 : 0;       \
        out.threads = Pth::G;                                                                 \
        return true;                                                                          \
    }
#define KRON_MIX_INFO(DD, NN) KRON_MIX_INFO_V(DD, NN, 0) KRON_MIX_INFO_V(DD, NN, 1)
    KRON_MIX_FOR_EACH(KRON_MIX_INFO)
#undef KRON_MIX_INFO
#undef KRON_MIX_INFO_V
    out = MixPathInfo{};
    return false;
}

template <typename T>
bool mixed_path_info(int d, int n, int variant, MixPathInfo &out)
{
    if (variant < 0 || variant >= kMixVariants) return false;
    return path_info<T>(d, n, variant, out) && out.valid;
}

int mixed_path_code(int d, int n, int variant) { return mix_code(d, n, variant); }

template <typename T>
const void *mixed_kernel() { return reinterpret_cast<const void *>(kron_mixed_kernel<T>); }

template <typename T>
kron_status mixed_launch(const MixedLaunch &L, hipStream_t stream)
{
    if (L.nitems == 0) return KRON_OK;
    const void *kern = mixed_kernel<T>();
    if (L.lds > 65536) {
        hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)L.lds);
        if (e != hipSuccess) {
            set_last_hip_error((int)e);
            return KRON_EHIP;
        }
    }
    hipLaunchKernelGGL(kron_mixed_kernel<T>, dim3((unsigned)L.grid), dim3(kMixThreads), L.lds,
                       stream, static_cast<const MixedClass<T> *>(L.classes), L.items, L.order,
                       L.nitems, L.queue, static_cast<T *>(L.scratch));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_last_hip_error((int)e);
        return KRON_EHIP;
    }
    return KRON_OK;
}

template bool mixed_path_info<double>(int, int, int, MixPathInfo &);
template bool mixed_path_info<float>(int, int, int, MixPathInfo &);
template const void *mixed_kernel<double>();
template const void *mixed_kernel<float>();
template kron_status mixed_launch<double>(const MixedLaunch &, hipStream_t);
template kron_status mixed_launch<float>(const MixedLaunch &, hipStream_t);

}  // namespace kron
