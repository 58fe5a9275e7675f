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
//     P entries of a small class (packed into one workgroup, atomic Y), or,
//     for a class too large for LDS, S slabs of its first mode for a
//     Y-target group (every slab reads X and adds its own rows of Y, so
//     items never wait for one another and an apply is one launch);
//   * the items are spread over one queue per XCD (the slab items of an
//     entry on one queue, so their reads of its X share one L2), sorted by
//     estimated cost, largest first, and taken from atomic counters
//     (longest-processing-time-first list scheduling); a workgroup drains
//     its own XCD's queue, then takes from the others;
//   * each class runs a compile-time (D, N) path with a tile and thread
//     mapping variant chosen by the autotuner (threads per entry for on-chip
//     classes, slabs per item for large ones);
//   * a third mapping of the large classes, two passes through scratch
//     (slow modes, then fast modes), needs the apply split into two
//     launches; it is kept for comparison and never chosen by the tuner;
//   * the counters reset themselves (the last workgroup out clears them), so the
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
    // large classes: variants 0 and 1 run an entry as N slabs of its first
    // mode (no scratch) -- 0 with as many slabs per work item as LDS holds
    // (fewer reads of X), 1 with one (more, smaller items) -- and variant 2
    // in two passes through scratch
    static constexpr int SMAX = BIG ? (L / F < N ? L / F : N) : 0;
    static constexpr bool SLAB = BIG && V <= 1;
    static constexpr int SLABS = V == 0 ? SMAX : 1;  // slabs per item
    static constexpr bool TWOPASS = BIG && V >= 2;
    // on-chip classes: variant v halves the threads per entry v times
    static constexpr int G = BIG ? kMixThreads : (G0 >> V) > 0 ? (G0 >> V) : 1;
    static constexpr int P = kMixThreads / G;
    // elements of a two-pass tile
    static constexpr int LT = L;
    static constexpr MixSplit SP = mix_split(D, N, LT, TWOPASS ? V - 2 : 0);
    static constexpr bool VALID = SLAB      ? SMAX >= 1 && (V == 0 || SMAX > 1)
                                  : TWOPASS ? SP.cost >= 0
                                            : ((G0 >> V) >= 1 && (size_t)P * E <= (size_t)L);
    static constexpr size_t lds_bytes()
    {
        if (SLAB) return (size_t)SLABS * F * sizeof(T);
        if (TWOPASS) {
            const int S = cpow(N, SP.h), Fs = cpow(N, D - SP.h);
            return (size_t)std::max(S * SP.c1, SP.c2 * Fs) * sizeof(T);
        }
        return D == 1 ? 0 : (size_t)P * E * sizeof(T);
    }
    static constexpr long long scratch_elems() { return TWOPASS ? E : 0; }
};

// Largest factor (bytes) held in SGPRs: up to 6 x 6 in FP64.  The kernel keeps
// more scalar state live than a single-shape kernel, so some of it spills to
// VGPR lanes; that still beats holding the factor in VGPRs through vector
// loads, whose waits share the in-order memory counter with the data loads.
constexpr int kMixSBytes = 288;

// Paths with N <= 6 and at most 24 first-pass values per thread keep the
// next unit's input in flight through their remaining passes (factor plus
// input fit the 256 VGPRs of two waves per SIMD); the others stream their
// first pass in batches.
template <int N>
constexpr bool mix_pipelined(int regs)
{
    return N <= 6 && regs <= 24;
}

// Launch state shared by the paths.
template <typename T>
struct MixCtx {
    const MixedClass<T> *cls;
    const MixedItem *items;
    const MixedEntry *recs;  // per-entry pointers, in item order
    int nitems;
    unsigned *queue;   // the queues' take counters, then the exit counter
    MixQueues qs;
    T *scratch;
    T *smem;
    int *s_next;       // LDS slot broadcasting a dequeued index
    // thread 0: the queue it takes from, how many queues it has found empty,
    // and the position requested in that queue for the item after next
    int q, empty;
    unsigned ahead;
};

// The XCD this wave runs on (hardware register XCC_ID; 0..7 on MI355X).
__device__ __forceinline__ int mix_xcd()
{
    unsigned v;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
    return (int)(v & (kMixXcds - 1));
}

// Thread 0: request the next position of queue x.q.
template <typename T>
__device__ __forceinline__ void mix_request(MixCtx<T> &x)
{
    x.ahead = atomicAdd(x.queue + kMixQueueStride * x.q, 1u);
}

// The next item index (workgroup-uniform; nitems once every queue is empty).
// A workgroup takes from its own XCD's queue, then from the following ones
// in turn.  Thread 0 keeps one request in flight, so the atomic's round trip
// overlaps the current item; a request past the end of its queue moves it
// to the next queue.  Every position a request returns below its queue's
// length is run, so each item runs exactly once.
template <typename T>
__device__ __forceinline__ int mix_dequeue(MixCtx<T> &x)
{
    if (threadIdx.x == 0) {
        int it = x.nitems;
        while (x.empty < kMixXcds) {
            const int beg = x.qs.off[x.q], len = x.qs.off[x.q + 1] - beg;
            if ((int)x.ahead < len) {
                it = beg + (int)x.ahead;
                mix_request(x);
                break;
            }
            ++x.empty;
            x.q = x.q + 1 == kMixXcds ? 0 : x.q + 1;
            if (x.empty < kMixXcds) mix_request(x);
        }
        *x.s_next = it;
    }
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(*x.s_next);
    __syncthreads();  // s_next is rewritten by the next dequeue
    return it;
}

// Entry r's factor k, X and Y from its record: with r uniform, scalar loads
// of one 64-byte record; the _v forms take a per-lane r.
template <typename T>
__device__ __forceinline__ cptr<T> mix_factor(const MixCtx<T> &x, int r, int k)
{
    return to_const(static_cast<const T *>((to_const(x.recs) + r)->A[k]));
}
template <typename T>
__device__ __forceinline__ gptr<const T> mix_x(const MixCtx<T> &x, int r)
{
    return to_global(static_cast<const T *>((to_const(x.recs) + r)->X));
}
template <typename T>
__device__ __forceinline__ gptr<T> mix_y(const MixCtx<T> &x, int r)
{
    return to_global(static_cast<T *>((to_const(x.recs) + r)->Y));
}
template <typename T>
__device__ __forceinline__ gptr<const T> mix_factor_v(const MixCtx<T> &x, int r, int k)
{
    return to_global(static_cast<const T *>((to_global(x.recs) + r)->A[k]));
}
template <typename T>
__device__ __forceinline__ gptr<const T> mix_x_v(const MixCtx<T> &x, int r)
{
    return to_global(static_cast<const T *>((to_global(x.recs) + r)->X));
}
template <typename T>
__device__ __forceinline__ gptr<T> mix_y_v(const MixCtx<T> &x, int r)
{
    return to_global(static_cast<T *>((to_global(x.recs) + r)->Y));
}

// A work item with its class, read through scalar loads.
template <typename T>
struct MixUnit {
    MixedClass<T> c;
    int first, count, chunk;
};

template <typename T>
__device__ __forceinline__ void mix_read(const MixCtx<T> &x, int it, MixUnit<T> &u)
{
    const auto ip = to_const(x.items) + it;
    u.first = ip->first;
    u.count = ip->count;
    u.chunk = ip->chunk;
    const auto cp = to_const(x.cls) + ip->cls;
    u.c.pA = cp->pA;
    u.c.pX = cp->pX;
    u.c.pY = cp->pY;
    u.c.d = cp->d;
    u.c.n = cp->n;
    u.c.lda = cp->lda;
    u.c.path = cp->path;
    u.c.atomic_y = cp->atomic_y;
    u.c.wofs = cp->wofs;
}

template <typename T>
__device__ __forceinline__ int mix_path_of(const MixCtx<T> &x, int it)
{
    if (it >= x.nitems) return -1;
    const auto ip = to_const(x.items) + it;
    return (to_const(x.cls) + ip->cls)->path;
}

// Runs units of one path back to back.  A unit is one entry of an item
// (Path::PER_ENTRY) or the whole item.  The first-pass input of the next unit
// -- the item's next entry, or the next queue item when it uses the same
// path -- is loaded into the registers the current unit's first pass has
// just consumed, so its HBM latency overlaps the current unit's remaining
// passes (the fast kernel's pipeline, across work items).  Returns the next
// item index of another path (or >= nitems).
template <typename T, typename Path>
__device__ __forceinline__ int mix_pipeline(MixCtx<T> &x, int it, int code, int tid)
{
    MixUnit<T> u;
    mix_read(x, it, u);
    int k = 0;
    typename Path::Regs r;
    // tid is made opaque again in every iteration: otherwise the address math
    // of all passes is hoisted out of the unit loop and held in VGPRs
    if constexpr (!Path::PIPE) {
        // registers too tight to keep a prefetch in flight: one unit at a time
        for (;;) {
            int t = tid;
            asm volatile("" : "+v"(t));
            Path::direct(x, u, k, t);
            Path::rest(x, u, k, t);
            if (Path::PER_ENTRY && k + 1 < u.count) {
                ++k;
                continue;
            }
            it = mix_dequeue(x);
            if (mix_path_of(x, it) != code) return it;
            mix_read(x, it, u);
            k = 0;
        }
    }
    Path::load(r, x, u, k, tid);
    for (;;) {
        int t = tid;
        asm volatile("" : "+v"(t));
        MixUnit<T> nu = u;
        int nk = k + 1, nit = it;
        bool same = true;
        if (!Path::PER_ENTRY || nk >= u.count) {
            nit = mix_dequeue(x);
            same = mix_path_of(x, nit) == code;
            if (same) mix_read(x, nit, nu);
            nk = 0;
        }
        Path::first(r, x, u, k, t);
        if (same) Path::load(r, x, nu, nk, t);
        Path::rest(x, u, k, t);
        if (!same) return nit;
        u = nu;
        k = nk;
        it = nit;
    }
}

// One entry per workgroup, resident in LDS (the fast kernel's mode order:
// 2 from HBM, D..3 in place, 1 into Y).  An item's entries share one Y target
// when the class updates Y without atomics; they run in order and each
// thread updates the same Y elements every time.
template <typename T, typename Acc, int D, int N>
struct PathOnchip {
    static constexpr bool PER_ENTRY = true;
    static constexpr int F = cpow(N, D - 1);
    static constexpr int S2 = cpow(N, D - 2);
    static constexpr int IT = (F + kMixThreads - 1) / kMixThreads;
    static constexpr bool PIPE = mix_pipelined<N>(IT * N);
    struct Regs {
        T v[IT][N];
    };
    static __device__ __forceinline__ T &L(const MixCtx<T> &x, int i)
    {
        return x.smem[lds_index<D, N>(i)];
    }
    static __device__ __forceinline__ void load(Regs &rg, const MixCtx<T> &x,
                                                const MixUnit<T> &u, int k, int tid)
    {
        const int r = u.first + k;
        const auto xp = mix_x<T>(x, r);
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int f = tid + it * kMixThreads;
            if ((F % kMixThreads) == 0 || f < F) {
                const int base = (f % S2) + (f / S2) * (S2 * N);
#pragma unroll
                for (int j = 0; j < N; ++j) rg.v[it][j] = xp[base + j * S2];
            }
        }
    }
    static __device__ __forceinline__ void first(const Regs &rg, const MixCtx<T> &x,
                                                 const MixUnit<T> &u, int k, int tid)
    {
        const int r = u.first + k;
        Acc a[N][N];
        load_factor<T, Acc, N, false, true, kMixSBytes>(a, mix_factor<T>(x, r, 1),
                                                        u.c.lda);
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int f = tid + it * kMixThreads;
            if ((F % kMixThreads) == 0 || f < F) {
                const int base = (f % S2) + (f / S2) * (S2 * N);
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    Acc s = a[i][0] * Acc(rg.v[it][0]);
#pragma unroll
                    for (int j = 1; j < N; ++j) s = kfma(a[i][j], Acc(rg.v[it][j]), s);
                    L(x, base + i * S2) = T(s);
                }
            }
        }
    }
    // load + first without holding the whole input: batched loads
    static __device__ __forceinline__ void direct(const MixCtx<T> &x, const MixUnit<T> &u, int k,
                                                  int tid)
    {
        const int r = u.first + k;
        Acc a[N][N];
        load_factor<T, Acc, N, false, true, kMixSBytes>(a, mix_factor<T>(x, r, 1),
                                                        u.c.lda);
        const auto xp = mix_x<T>(x, r);
        tile_pass_in<T, Acc, N, S2, F, kMixThreads>(
            tid, a, [&](int e) { return xp[e]; }, [&](int e, Acc v) { L(x, e) = T(v); });
    }
    static __device__ __forceinline__ void rest(const MixCtx<T> &x, const MixUnit<T> &u, int k,
                                                int tid)
    {
        const int r = u.first + k;
        Acc a[N][N];
        __syncthreads();
        for_modes_desc<D, 3>([&](auto kc) {
            constexpr int K = decltype(kc)::value;
            load_factor<T, Acc, N, false, true, kMixSBytes>(
                a, mix_factor<T>(x, r, K - 1), u.c.lda);
            fiber_pass<Acc, D, N, K, kMixThreads>(
                tid, a, [&](int e) { return L(x, e); }, [&](int e, Acc v) { L(x, e) = T(v); });
            __syncthreads();
        });
        load_factor<T, Acc, N, false, true, kMixSBytes>(a, mix_factor<T>(x, r, 0),
                                                        u.c.lda);
        tile_pass_y<T, Acc, N, F, F, kMixThreads>(
            tid, a, [&](int e) { return L(x, e); }, mix_y<T>(x, r),
            [](int e) { return e; }, [](int) { return true; }, u.c.atomic_y != 0);
        __syncthreads();  // LDS is rewritten by the next unit
    }
};

// P = 256 / G entries per workgroup, G threads each (small entries, or
// d = 1); the unit is the whole item.  Y is updated atomically (entries of
// one item may share Y).
template <typename T, typename Acc, int D, int N, int G>
struct PathPacked {
    static constexpr bool PER_ENTRY = false;
    static constexpr int E = cpow(N, D);
    static constexpr int F = E / N;
    static constexpr int S2 = D >= 2 ? cpow(N, D - 2) : 1;
    static constexpr int IT = (F + G - 1) / G;
    static constexpr bool PIPE = false;
    struct Regs {
        T v[IT][N];
    };
    // this thread's slot entry (b < 0: no entry)
    static __device__ __forceinline__ int slot_entry(const MixCtx<T> &x, const MixUnit<T> &u,
                                                     int tid)
    {
        const int slot = tid / G;
        return slot < u.count ? u.first + slot : -1;
    }
    static __device__ __forceinline__ void load(Regs &rg, const MixCtx<T> &x,
                                                const MixUnit<T> &u, int, int tid)
    {
        const int b = slot_entry(x, u, tid);
        if (b < 0) return;
        const int g = tid % G;
        const auto xp = mix_x_v<T>(x, b);
        constexpr int FS = D >= 2 ? S2 : 1;  // stride of the first pass's fibers
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int f = g + it * G;
            if ((F % G) == 0 || f < F) {
                const int base = (f % FS) + (f / FS) * (FS * N);
#pragma unroll
                for (int j = 0; j < N; ++j) rg.v[it][j] = xp[base + j * FS];
            }
        }
    }
    static __device__ __forceinline__ void first(const Regs &rg, const MixCtx<T> &x,
                                                 const MixUnit<T> &u, int, int tid)
    {
        const int b = slot_entry(x, u, tid);
        const int g = tid % G;
        T *buf = x.smem + (tid / G) * E;
        Acc a[N][N];
        if (b >= 0) {
            load_factor<T, Acc, N, false, false>(a, mix_factor_v<T>(x, b, D >= 2 ? 1 : 0),
                                                 u.c.lda);
            constexpr int FS = D >= 2 ? S2 : 1;
            const auto y = mix_y_v<T>(x, b);
#pragma unroll
            for (int it = 0; it < IT; ++it) {
                const int f = g + it * G;
                if ((F % G) == 0 || f < F) {
                    const int base = (f % FS) + (f / FS) * (FS * N);
#pragma unroll
                    for (int i = 0; i < N; ++i) {
                        Acc s = a[i][0] * Acc(rg.v[it][0]);
#pragma unroll
                        for (int j = 1; j < N; ++j)
                            s = kfma(a[i][j], Acc(rg.v[it][j]), s);
                        if constexpr (D == 1)
                            atomic_add(y + (base + i * FS), T(s));
                        else
                            buf[lds_index<D, N>(base + i * FS)] = T(s);
                    }
                }
            }
        }
    }
    static __device__ __forceinline__ void direct(const MixCtx<T> &x, const MixUnit<T> &u, int,
                                                  int tid)
    {
        const int b = slot_entry(x, u, tid);
        if (b < 0) return;
        const int g = tid % G;
        T *buf = x.smem + (tid / G) * E;
        Acc a[N][N];
        load_factor<T, Acc, N, false, false>(a, mix_factor_v<T>(x, b, D >= 2 ? 1 : 0),
                                             u.c.lda);
        constexpr int FS = D >= 2 ? S2 : 1;
        const auto xp = mix_x_v<T>(x, b);
        const auto y = mix_y_v<T>(x, b);
        // one fiber at a time: batching loads through per-slot (non-uniform)
        // pointers would hold a 64-bit address per load
        tile_pass<Acc, N, FS, F, G>(g, a, [&](int e) { return xp[e]; }, [&](int e, Acc v) {
            if constexpr (D == 1)
                atomic_add(y + e, T(v));
            else
                buf[lds_index<D, N>(e)] = T(v);
        });
    }
    static __device__ __forceinline__ void rest(const MixCtx<T> &x, const MixUnit<T> &u, int,
                                                int tid)
    {
        if constexpr (D >= 2) {
            const int b = slot_entry(x, u, tid);
                const int g = tid % G;
            T *buf = x.smem + (tid / G) * E;
            auto L = [&](int i) -> T & { return buf[lds_index<D, N>(i)]; };
            Acc a[N][N];
            __syncthreads();
            for_modes_desc<D, 3>([&](auto kc) {
                constexpr int K = decltype(kc)::value;
                if (b >= 0) {
                    load_factor<T, Acc, N, false, false>(
                        a, mix_factor_v<T>(x, b, K - 1), u.c.lda);
                    fiber_pass<Acc, D, N, K, G>(
                        g, a, [&](int e) { return L(e); }, [&](int e, Acc v) { L(e) = T(v); });
                }
                __syncthreads();
            });
            if (b >= 0) {
                load_factor<T, Acc, N, false, false>(a, mix_factor_v<T>(x, b, 0),
                                                     u.c.lda);
                const auto y = mix_y_v<T>(x, b);
                fiber_pass<Acc, D, N, 1, G>(
                    g, a, [&](int e) { return L(e); },
                    [&](int e, Acc v) { atomic_add(y + e, T(v)); });
            }
            __syncthreads();
        }
    }
};

// A large entry in two passes through its own scratch vector W (n^d
// elements; entry b of the class at scratch + wofs + b * n^d).  Pass 1
// contracts the slow modes 1..H of one S x C1 column tile (X -> W), pass 2
// the fast modes H+1..D of one C2 x F slab tile (W -> Y); each tile is a work
// item, and the two passes are separate launches (the launch boundary makes
// W visible across the chip).  Tile elements past the end of the entry
// compute on stale LDS and are not stored.
template <typename T, typename Acc, int D, int N, int H, int C1>
struct PathPass1 {
    static constexpr bool PER_ENTRY = false;
    static constexpr int S = cpow(N, H), F = cpow(N, D - H);
    static constexpr int T1 = S * C1;
    static constexpr int NF = T1 / N;  // fibers of each tile mode
    static constexpr int IT = (NF + kMixThreads - 1) / kMixThreads;
    static constexpr bool PIPE = mix_pipelined<N>(IT * N);
    struct Regs {
        T v[IT][N];
    };
    // tile element t = s * C1 + c is entry element s * F + c0 + c
    static __device__ __forceinline__ int gidx(int t, int c0) { return (t / C1) * F + c0 + t % C1; }
    static __device__ __forceinline__ bool valid(int t, int c0) { return c0 + t % C1 < F; }
    static __device__ __forceinline__ void load(Regs &rg, const MixCtx<T> &x,
                                                const MixUnit<T> &u, int, int tid)
    {
        const int r = u.first;
        const int c0 = u.chunk * C1;
        const auto xp = mix_x<T>(x, r);
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int f = tid + it * kMixThreads;
            if ((NF % kMixThreads) == 0 || f < NF) {
                const int base = (f % C1) + (f / C1) * (C1 * N);
#pragma unroll
                for (int j = 0; j < N; ++j) {
                    const int t = base + j * C1;
                    rg.v[it][j] = valid(t, c0) ? xp[gidx(t, c0)] : T(0);
                }
            }
        }
    }
    static __device__ __forceinline__ void first(const Regs &rg, const MixCtx<T> &x,
                                                 const MixUnit<T> &u, int, int tid)
    {
        const int r = u.first;
        const int c0 = u.chunk * C1;
        const auto W = to_global(x.scratch + u.c.wofs + (long long)r * (S * F));
        Acc a[N][N];
        load_factor<T, Acc, N, false, true, kMixSBytes>(a, mix_factor<T>(x, r, H - 1),
                                                        u.c.lda);
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int f = tid + it * kMixThreads;
            if ((NF % kMixThreads) == 0 || f < NF) {
                const int base = (f % C1) + (f / C1) * (C1 * N);
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    Acc s = a[i][0] * Acc(rg.v[it][0]);
#pragma unroll
                    for (int j = 1; j < N; ++j) s = kfma(a[i][j], Acc(rg.v[it][j]), s);
                    const int t = base + i * C1;
                    if constexpr (H == 1) {
                        if (valid(t, c0)) W[gidx(t, c0)] = T(s);
                    } else {
                        x.smem[t] = T(s);
                    }
                }
            }
        }
    }
    static __device__ __forceinline__ void direct(const MixCtx<T> &x, const MixUnit<T> &u, int,
                                                  int tid)
    {
        const int r = u.first;
        const int c0 = u.chunk * C1;
        const auto W = to_global(x.scratch + u.c.wofs + (long long)r * (S * F));
        const auto xp = mix_x<T>(x, r);
        Acc a[N][N];
        load_factor<T, Acc, N, false, true, kMixSBytes>(a, mix_factor<T>(x, r, H - 1),
                                                        u.c.lda);
        tile_pass_in<T, Acc, N, C1, NF, kMixThreads>(
            tid, a, [&](int t) { return valid(t, c0) ? xp[gidx(t, c0)] : T(0); },
            [&](int t, Acc v) {
                if constexpr (H == 1) {
                    if (valid(t, c0)) W[gidx(t, c0)] = T(v);
                } else {
                    x.smem[t] = T(v);
                }
            });
    }
    static __device__ __forceinline__ void rest(const MixCtx<T> &x, const MixUnit<T> &u, int,
                                                int tid)
    {
        if constexpr (H > 1) {
            const int r = u.first;
            const int c0 = u.chunk * C1;
                const auto W = to_global(x.scratch + u.c.wofs + (long long)r * (S * F));
            auto L = [&](int i) -> T & { return x.smem[i]; };
            Acc a[N][N];
            __syncthreads();
            for_modes_desc<H - 1, 2>([&](auto kc) {
                constexpr int K = decltype(kc)::value;
                load_factor<T, Acc, N, false, true, kMixSBytes>(
                    a, mix_factor<T>(x, r, K - 1), u.c.lda);
                tile_pass<Acc, N, cpow(N, H - K) * C1, NF, kMixThreads>(
                    tid, a, [&](int t) { return L(t); }, [&](int t, Acc v) { L(t) = T(v); });
                __syncthreads();
            });
            load_factor<T, Acc, N, false, true, kMixSBytes>(a, mix_factor<T>(x, r, 0),
                                                            u.c.lda);
            tile_pass<Acc, N, cpow(N, H - 1) * C1, NF, kMixThreads>(
                tid, a, [&](int t) { return L(t); },
                [&](int t, Acc v) {
                    if (valid(t, c0)) W[gidx(t, c0)] = T(v);
                });
            __syncthreads();  // LDS is rewritten by the next unit
        }
    }
};

// Pass 2 of slab tile -2 - chunk for the entries of one item, in order (they
// share one Y target when the class updates Y without atomics, and this
// workgroup owns that slab of it).
template <typename T, typename Acc, int D, int N, int H, int C2>
struct PathPass2 {
    static constexpr bool PER_ENTRY = true;
    static constexpr int S = cpow(N, H), F = cpow(N, D - H);
    static constexpr int T2 = C2 * F;
    static constexpr int NF = T2 / N;
    static constexpr int IT = (NF + kMixThreads - 1) / kMixThreads;
    static constexpr bool PIPE = mix_pipelined<N>(IT * N) && D - H > 1;
    struct Regs {
        T v[IT][N];
    };
    static __device__ __forceinline__ int rows(const MixUnit<T> &u) { return (-2 - u.chunk) * C2; }
    // valid tile elements (tile element t = r * F + f is entry element r0 * F + t)
    static __device__ __forceinline__ int nvalid(int r0) { return (S - r0 < C2 ? S - r0 : C2) * F; }
    static __device__ __forceinline__ void load(Regs &rg, const MixCtx<T> &x,
                                                const MixUnit<T> &u, int k, int tid)
    {
        const int r = u.first + k;
        const int r0 = rows(u), nv = nvalid(r0);
        const auto w = to_global(x.scratch + u.c.wofs + (long long)r * (S * F) + (long long)r0 * F);
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int f = tid + it * kMixThreads;
            if ((NF % kMixThreads) == 0 || f < NF) {
#pragma unroll
                for (int j = 0; j < N; ++j) {
                    const int t = f * N + j;
                    rg.v[it][j] = t < nv ? w[t] : T(0);
                }
            }
        }
    }
    static __device__ __forceinline__ void first(const Regs &rg, const MixCtx<T> &x,
                                                 const MixUnit<T> &u, int k, int tid)
    {
        const int r = u.first + k;
        const int r0 = rows(u), nv = nvalid(r0);
        Acc a[N][N];
        load_factor<T, Acc, N, false, true, kMixSBytes>(a, mix_factor<T>(x, r, D - 1),
                                                        u.c.lda);
        if constexpr (D - H == 1) {
            // the only fast mode: straight into Y
            const auto yt = mix_y<T>(x, r) + (long long)r0 * F;
            const bool atomic = u.c.atomic_y != 0;
            T yv[IT][N];
            if (!atomic) {
#pragma unroll
                for (int it = 0; it < IT; ++it) {
                    const int f = tid + it * kMixThreads;
                    if ((NF % kMixThreads) == 0 || f < NF) {
#pragma unroll
                        for (int i = 0; i < N; ++i) {
                            const int t = f * N + i;
                            yv[it][i] = t < nv ? yt[t] : T(0);
                        }
                    }
                }
            }
#pragma unroll
            for (int it = 0; it < IT; ++it) {
                const int f = tid + it * kMixThreads;
                if ((NF % kMixThreads) == 0 || f < NF) {
#pragma unroll
                    for (int i = 0; i < N; ++i) {
                        Acc s = a[i][0] * Acc(rg.v[it][0]);
#pragma unroll
                        for (int j = 1; j < N; ++j)
                            s = kfma(a[i][j], Acc(rg.v[it][j]), s);
                        const int t = f * N + i;
                        if (t < nv) {
                            if (atomic)
                                atomic_add(yt + t, T(s));
                            else
                                yt[t] = T(Acc(yv[it][i]) + s);
                        }
                    }
                }
            }
        } else {
#pragma unroll
            for (int it = 0; it < IT; ++it) {
                const int f = tid + it * kMixThreads;
                if ((NF % kMixThreads) == 0 || f < NF) {
#pragma unroll
                    for (int i = 0; i < N; ++i) {
                        Acc s = a[i][0] * Acc(rg.v[it][0]);
#pragma unroll
                        for (int j = 1; j < N; ++j)
                            s = kfma(a[i][j], Acc(rg.v[it][j]), s);
                        x.smem[f * N + i] = T(s);
                    }
                }
            }
        }
    }
    static __device__ __forceinline__ void direct(const MixCtx<T> &x, const MixUnit<T> &u, int k,
                                                  int tid)
    {
        const int r = u.first + k;
        const int r0 = rows(u), nv = nvalid(r0);
        const auto w = to_global(x.scratch + u.c.wofs + (long long)r * (S * F) + (long long)r0 * F);
        Acc a[N][N];
        load_factor<T, Acc, N, false, true, kMixSBytes>(a, mix_factor<T>(x, r, D - 1),
                                                        u.c.lda);
        auto ld = [&](int t) { return t < nv ? w[t] : T(0); };
        if constexpr (D - H == 1) {
            tile_pass_y<T, Acc, N, 1, NF, kMixThreads>(
                tid, a, ld, mix_y<T>(x, r) + (long long)r0 * F,
                [](int t) { return t; }, [&](int t) { return t < nv; }, u.c.atomic_y != 0);
        } else {
            tile_pass_in<T, Acc, N, 1, NF, kMixThreads>(tid, a, ld,
                                                        [&](int t, Acc v) { x.smem[t] = T(v); });
        }
    }
    static __device__ __forceinline__ void rest(const MixCtx<T> &x, const MixUnit<T> &u, int k,
                                                int tid)
    {
        if constexpr (D - H > 1) {
            const int r = u.first + k;
            const int r0 = rows(u), nv = nvalid(r0);
                auto L = [&](int i) -> T & { return x.smem[i]; };
            Acc a[N][N];
            __syncthreads();
            for_modes_desc<D - 1, H + 2>([&](auto kc) {
                constexpr int K = decltype(kc)::value;
                load_factor<T, Acc, N, false, true, kMixSBytes>(
                    a, mix_factor<T>(x, r, K - 1), u.c.lda);
                tile_pass<Acc, N, cpow(N, D - K), NF, kMixThreads>(
                    tid, a, [&](int t) { return L(t); }, [&](int t, Acc v) { L(t) = T(v); });
                __syncthreads();
            });
            load_factor<T, Acc, N, false, true, kMixSBytes>(a, mix_factor<T>(x, r, H),
                                                            u.c.lda);
            tile_pass_y<T, Acc, N, cpow(N, D - H - 1), NF, kMixThreads>(
                tid, a, [&](int t) { return L(t); },
                mix_y<T>(x, r) + (long long)r0 * F, [](int t) { return t; },
                [&](int t) { return t < nv; }, u.c.atomic_y != 0);
            __syncthreads();  // LDS is rewritten by the next unit
        }
    }
};

// A large entry as N slabs of its first mode: slab i of Y is
//     Y[i * F + s] += ((A_2 (x) .. (x) A_D) Z_i)[s],  Z_i[s] = sum_j A_1(i, j) X[j * F + s],
// F = N^(D-1).  A work item takes S consecutive slabs, i0 = its chunk:
// Z_i0 .. Z_(i0+S-1) are formed straight from X into LDS (S FMAs per X
// element read), modes D..3 are contracted there in place for all of them
// and mode 2 goes into Y.  An entry's X is read ceil(N / S) times; the
// entry's items are neighbours in the queue, so the re-reads find X in the
// caches more often than not.  No scratch and no order between work items:
// the classes that would need a second pass through scratch run in the same
// launch as the rest.  The units of an item are the entries of its group
// (they share the Y target when the class updates Y without atomics; each
// thread updates the same Y elements every time).
template <typename T, typename Acc, int D, int N, int S>
struct PathSlab {
    static constexpr bool PER_ENTRY = true;
    static constexpr bool PIPE = false;
    static constexpr int F = cpow(N, D - 1);
    static constexpr int IT = (F + kMixThreads - 1) / kMixThreads;
    // fiber iterations whose loads are in flight together (about 36 values)
    static constexpr int B = 36 / N > 0 ? 36 / N : 1;
    static constexpr int NB = (IT + B - 1) / B;
    static constexpr int S2 = cpow(N, D - 2);
    struct Regs {};
    static __device__ __forceinline__ T &L(const MixCtx<T> &x, int i)
    {
        return x.smem[lds_index<D - 1, N>(i)];
    }
    static __device__ __forceinline__ void load(Regs &, const MixCtx<T> &, const MixUnit<T> &,
                                                int, int)
    {
    }
    static __device__ __forceinline__ void first(const Regs &, const MixCtx<T> &,
                                                 const MixUnit<T> &, int, int)
    {
    }
    static __device__ __forceinline__ void direct(const MixCtx<T> &x, const MixUnit<T> &u, int k,
                                                  int tid)
    {
        const int r = u.first + k, i0 = u.chunk;
        const auto fp = mix_factor<T>(x, r, 0);
        // rows i0 .. i0+S-1 of A_1; rows past N contribute zero slabs, which
        // are not stored
        Acc a1[S][N];
#pragma unroll
        for (int q = 0; q < S; ++q)
#pragma unroll
            for (int j = 0; j < N; ++j)
                a1[q][j] = i0 + q < N ? Acc(fp[i0 + q + j * u.c.lda]) : Acc(0);
        const auto xp = mix_x<T>(x, r);
        T in[2][B][N];
        auto fetch = [&](int bi, T(&dst)[B][N]) {
#pragma unroll
            for (int q = 0; q < B; ++q) {
                const int e = tid + (bi * B + q) * kMixThreads;
                if (bi * B + q < IT && ((F % kMixThreads) == 0 || e < F)) {
#pragma unroll
                    for (int j = 0; j < N; ++j) dst[q][j] = xp[e + j * F];
                }
            }
        };
        fetch(0, in[0]);
#pragma unroll
        for (int bi = 0; bi < NB; ++bi) {
            if (bi + 1 < NB) fetch(bi + 1, in[(bi + 1) & 1]);
#pragma unroll
            for (int q = 0; q < B; ++q) {
                const int e = tid + (bi * B + q) * kMixThreads;
                if (bi * B + q < IT && ((F % kMixThreads) == 0 || e < F)) {
#pragma unroll
                    for (int sl = 0; sl < S; ++sl) {
                        Acc v = a1[sl][0] * Acc(in[bi & 1][q][0]);
#pragma unroll
                        for (int j = 1; j < N; ++j)
                            v = kfma(a1[sl][j], Acc(in[bi & 1][q][j]), v);
                        L(x, sl * F + e) = T(v);
                    }
                }
            }
        }
    }
    static __device__ __forceinline__ void rest(const MixCtx<T> &x, const MixUnit<T> &u, int k,
                                                int tid)
    {
        const int r = u.first + k, i0 = u.chunk;
        const int nvalid = (N - i0 < S ? N - i0 : S) * F;
        Acc a[N][N];
        __syncthreads();
        // original mode K is mode K - 1 of each slab's D - 1 modes; a pass
        // covers the fibers of all S slabs
        for_modes_desc<D, 3>([&](auto kc) {
            constexpr int K = decltype(kc)::value;
            load_factor<T, Acc, N, false, true, kMixSBytes>(a, mix_factor<T>(x, r, K - 1),
                                                            u.c.lda);
            tile_pass<Acc, N, cpow(N, D - K), S * S2, kMixThreads>(
                tid, a, [&](int e) { return L(x, e); }, [&](int e, Acc v) { L(x, e) = T(v); });
            __syncthreads();
        });
        load_factor<T, Acc, N, false, true, kMixSBytes>(a, mix_factor<T>(x, r, 1), u.c.lda);
        tile_pass_y<T, Acc, N, S2, S * S2, kMixThreads>(
            tid, a, [&](int e) { return L(x, e); }, mix_y<T>(x, r) + (long long)i0 * F,
            [](int e) { return e; }, [&](int e) { return e < nvalid; }, u.c.atomic_y != 0);
        __syncthreads();  // LDS is rewritten by the next unit
    }
};

// The paths of (D, N, variant); chunk selects the slab of slab classes, and
// pass 1 (>= 0) or pass 2 (<= -2) of two-pass classes.
template <typename T, int D, int N, int V>
__device__ __forceinline__ int mix_run(MixCtx<T> &x, int it, int code, int tid)
{
    using Pth = MixPath<T, D, N, V>;
    if constexpr (!Pth::VALID) {
        return mix_dequeue(x);
    } else if constexpr (Pth::SLAB) {
        return mix_pipeline<T, PathSlab<T, T, D, N, Pth::SLABS>>(x, it, code, tid);
    } else if constexpr (Pth::BIG) {
        if (to_const(x.items)[it].chunk >= 0)
            return mix_pipeline<T, PathPass1<T, T, D, N, Pth::SP.h, Pth::SP.c1>>(x, it, code, tid);
        return mix_pipeline<T, PathPass2<T, T, D, N, Pth::SP.h, Pth::SP.c2>>(x, it, code, tid);
    } else if constexpr (Pth::P == 1 && D >= 2) {
        return mix_pipeline<T, PathOnchip<T, T, D, N>>(x, it, code, tid);
    } else {
        return mix_pipeline<T, PathPacked<T, T, D, N, Pth::G>>(x, it, code, tid);
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
kron_mixed_kernel(const MixedClass<T> *cls, const MixedItem *items, const MixedEntry *recs,
                  int nitems, unsigned *queue, MixQueues qs, T *scratch)
{
    extern __shared__ __align__(16) unsigned char kron_smem[];
    __shared__ int s_next;
    MixCtx<T> x{cls, items, recs, nitems, queue, qs, scratch,
                reinterpret_cast<T *>(kron_smem), &s_next, 0, 0, 0u};
    if (threadIdx.x == 0) {
        x.q = mix_xcd();
        mix_request(x);
    }
    int it = mix_dequeue(x);
    while (it < nitems) {
        // lane-dependent index math of every path derives from this opaque
        // copy, so none of it is hoisted out of the loop (where it would stay
        // live across all paths)
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int code = mix_path_of(x, it);
        switch (code) {
#define KRON_MIX_CASE_V(D, N, V)                                                              \
    case mix_code(D, N, V): it = mix_run<T, D, N, V>(x, it, code, tid); break;
#define KRON_MIX_CASE(D, N) \
    KRON_MIX_CASE_V(D, N, 0) KRON_MIX_CASE_V(D, N, 1) KRON_MIX_CASE_V(D, N, 2)
            KRON_MIX_FOR_EACH(KRON_MIX_CASE)
#undef KRON_MIX_CASE
#undef KRON_MIX_CASE_V
        default: it = mix_dequeue(x); break;
        }
    }
    // every workgroup has taken its last (empty) items; the last one out
    // clears the counters for the next launch on the stream
    if (threadIdx.x == 0) {
        __threadfence();
        const unsigned done = atomicAdd(queue + kMixQueueExit, 1u);
        if (done == gridDim.x - 1) {
            for (int q = 0; q < kMixXcds; ++q) atomicExch(queue + kMixQueueStride * q, 0u);
            atomicExch(queue + kMixQueueExit, 0u);
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
        out.big = Pth::TWOPASS;                                                               \
        out.slab = Pth::SLAB;                                                                 \
        out.slabs = Pth::SLAB ? Pth::SLABS : 0;                                               \
        out.packed = !Pth::BIG && (Pth::P > 1 || DD == 1);                                    \
        out.per_item = Pth::BIG ? 1 : Pth::P;                                                 \
        out.lds = Pth::VALID ? Pth::lds_bytes() : 0;                                          \
        out.scratch = Pth::VALID ? Pth::scratch_elems() : 0;                                  \
        out.split = Pth::TWOPASS ? Pth::SP.h : 0;                                             \
        out.tiles1 = Pth::TWOPASS ? (cpow(NN, DD - Pth::SP.h) + Pth::SP.c1 - 1) / Pth::SP.c1   \
                                  : Pth::SLAB ? (NN + Pth::SLABS - 1) / Pth::SLABS : 0;       \
        out.tiles2 = Pth::TWOPASS ? (cpow(NN, Pth::SP.h) + Pth::SP.c2 - 1) / Pth::SP.c2 : 0;   \
        out.threads = Pth::G;                                                                 \
        return true;                                                                          \
    }
#define KRON_MIX_INFO(DD, NN) \
    KRON_MIX_INFO_V(DD, NN, 0) KRON_MIX_INFO_V(DD, NN, 1) KRON_MIX_INFO_V(DD, NN, 2)
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
                       stream, static_cast<const MixedClass<T> *>(L.classes), L.items, L.recs,
                       L.nitems, L.queue, L.q, static_cast<T *>(L.scratch));
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
