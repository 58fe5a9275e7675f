// kron_plane.hpp -- the d = 6, n = 4 entry (4,096 elements) with two modes
// per LDS pass.
//
// The fiber scheme of kron_fast.hpp makes one LDS pass per mode (six passes,
// ten LDS accesses per element).  Here each of the 256 threads owns a 4 x 4
// "plane" of the entry (two modes varying, the other four fixed) and
// contracts both modes in registers, so an entry takes three passes and four
// LDS accesses per element:
//
//   pass A: X plane (modes 2, 4), prefetched from HBM one entry ahead -> LDS
//   pass B: LDS plane (modes 6, 5) -> LDS, in place
//   pass C: LDS plane (modes 3, 1) -> Y (or the running sum of a Y run)
//
// LDS layout: element e at e + e / 16 (one element of padding after every
// 16).  The lane digits are m6, m5 (or m4, m3 in pass B): with the padding
// the 16 lanes a 64-bit LDS access serves together hit 16 distinct bank
// slots, the HBM accesses of passes A and C are runs of 16 (X) and 64 (Y)
// consecutive elements, and the 16 elements of a thread's plane sit at
// compile-time offsets from one base address (pass A: 272 i + 17 j, pass B:
// i + 4 j, pass C: 68 i + 1088 j).
// Element index e = 1024 m1 + 256 m2 + 64 m3 + 16 m4 + 4 m5 + m6 (A1 acts on
// m1, the slowest index, as in the reference's layout).
#pragma once

#include "kron_fast.hpp"  // load_factor, Entries

namespace kron {

constexpr int kPlaneElems = 4096;
constexpr int kPlaneLds = kPlaneElems + kPlaneElems / 16;

__device__ __forceinline__ int plane_pad(int e) { return e + (e >> 4); }

// v[i][j] <- sum_k a[i][k] v[k][j]: the mode along the first plane index.
template <typename Acc>
__device__ __forceinline__ void plane_first(Acc (&v)[4][4], const Acc (&a)[4][4])
{
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        Acc t[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            Acc s = a[i][0] * v[0][j];
#pragma unroll
            for (int k = 1; k < 4; ++k) s = kfma(a[i][k], v[k][j], s);
            t[i] = s;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i][j] = t[i];
    }
}

// v[i][j] <- sum_k a[j][k] v[i][k]: the mode along the second plane index.
template <typename Acc>
__device__ __forceinline__ void plane_second(Acc (&v)[4][4], const Acc (&a)[4][4])
{
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        Acc t[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            Acc s = a[j][0] * v[i][0];
#pragma unroll
            for (int k = 1; k < 4; ++k) s = kfma(a[j][k], v[i][k], s);
            t[j] = s;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) v[i][j] = t[j];
    }
}

// Addresses of one entry, wave-uniform: X, the six factors, Y (in target
// order: the Y key).
struct PlaneMeta {
    unsigned long long x, f[6], y;
};

__device__ __forceinline__ unsigned long long readlane_u64(unsigned long long v, int lane)
{
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), lane);
    return ((unsigned long long)hi << 32) | lo;
}

// The addresses of 32 consecutive iterations of a workgroup, spread over the
// lanes of each wave: lane i < 32 holds iteration wb + i's X and factor 0..2
// addresses, lane 32 + i its factor 3..5 and Y addresses.  Refilled once per
// 32 iterations with vector loads (the target order, then gathers of the
// pointer arrays); an iteration reads its addresses with readlane, so no
// load sits between an entry and its data loads.  The batched layout
// computes its addresses and keeps no window.
template <typename T, Layout LAY, bool GROUPED>
struct PlaneWindow {
    static constexpr bool USED = LAY == Layout::Ptr;
    unsigned long long v[4];

    __device__ __forceinline__ void refill(const Entries<T, LAY> &ent, const GroupOrder &ord,
                                           int wb, int cnt, int rbeg, int rstep, int lane)
    {
        if constexpr (USED) {
            const int jj = wb + (lane & 31);
            v[0] = v[1] = v[2] = v[3] = 0;
            if (jj < cnt) {
                const int rr = rbeg + jj * rstep;
                const int bb = GROUPED ? to_global(ord.perm)[rr] : rr;
                using U = unsigned long long;
                const auto pa = to_global(reinterpret_cast<const U *>(ent.pA)) + (long long)bb * 6;
                if (lane < 32) {
                    v[0] = to_global(reinterpret_cast<const U *>(ent.pX))[bb];
                    v[1] = pa[0];
                    v[2] = pa[1];
                    v[3] = pa[2];
                } else {
                    v[0] = pa[3];
                    v[1] = pa[4];
                    v[2] = pa[5];
                    if constexpr (GROUPED)
                        v[3] = to_global(ord.keys)[rr];
                    else
                        v[3] = to_global(reinterpret_cast<const U *>(ent.pY))[bb];
                }
            }
        }
    }

    // iteration j (wb <= j < wb + 32)
    __device__ __forceinline__ PlaneMeta at(const Entries<T, LAY> &ent, int j, int wb, int rbeg,
                                            int rstep) const
    {
        PlaneMeta m;
        if constexpr (USED) {
            const int i = j - wb;
            m.x = readlane_u64(v[0], i);
            m.f[0] = readlane_u64(v[1], i);
            m.f[1] = readlane_u64(v[2], i);
            m.f[2] = readlane_u64(v[3], i);
            m.f[3] = readlane_u64(v[0], 32 + i);
            m.f[4] = readlane_u64(v[1], 32 + i);
            m.f[5] = readlane_u64(v[2], 32 + i);
            m.y = readlane_u64(v[3], 32 + i);
        } else {
            const int b = rbeg + j * rstep;
            m.x = (unsigned long long)ent.x(b);
#pragma unroll
            for (int k = 0; k < 6; ++k) m.f[k] = (unsigned long long)ent.factor(b, k);
            m.y = (unsigned long long)ent.y(b);
        }
        return m;
    }
};

// Persistent workgroups, one entry at a time (as kron_fast_kernel with
// P == 1): GROUPED walks the Y-target order and adds each run's summed
// result to Y once; otherwise Y is updated per entry (atomically if SHARED).
// Iteration j of a workgroup is position rbeg + j * rstep.  While entry j is
// contracted, its successor's X plane is loaded into registers.  In the
// pointer layout wave 0 also pulls the successor's six factor lines into L2
// (one vector load per line), so the successor's scalar factor loads find
// them there (in the batched layout this measured 3% slower, so it is off).
template <typename T, typename Acc, Layout LAY, bool SHARED, bool TRANS, bool GROUPED>
__global__ void __launch_bounds__(256)
kron_plane_kernel(Entries<T, LAY> ent, int lda, int batch, GroupOrder ord)
{
    extern __shared__ __align__(16) unsigned char kron_smem[];
    T *const lds = reinterpret_cast<T *>(kron_smem);
    const int tid = threadIdx.x;
    const int l = tid & 63, w = tid >> 6;
    const int d0 = l & 3, d1 = (l >> 2) & 3, d2 = (l >> 4) & 3;
    // plane origins (lane digit d0 varies fastest): A (m1 = w, m3 = d2,
    // m5 = d1, m6 = d0), B (m1 = w, m2 = d2, m3 = d1, m4 = d0),
    // C (m2 = w, m4 = d2, m5 = d1, m6 = d0)
    const int eA = w * 1024 + d2 * 64 + d1 * 4 + d0;
    const int eB = w * 1024 + d2 * 256 + d1 * 64 + d0 * 16;
    const int eC = w * 256 + d2 * 16 + d1 * 4 + d0;
    T *const LA = lds + plane_pad(eA);
    T *const LB = lds + plane_pad(eB);
    T *const LC = lds + plane_pad(eC);

    int rbeg, rstep, cnt;
    if constexpr (GROUPED) {
        rbeg = blockIdx.x * ord.chunk;
        rstep = 1;
        cnt = max(0, min(batch, rbeg + ord.chunk) - rbeg);
    } else {
        rbeg = blockIdx.x;
        rstep = gridDim.x;
        cnt = rbeg < batch ? (batch - 1 - rbeg) / rstep + 1 : 0;
    }
    PlaneWindow<T, LAY, GROUPED> win;
    int wb = 0;
    Acc a[4][4];
    T nx[4][4];  // the next entry's X plane (modes 2, 4)
    auto load_x = [&](unsigned long long xa) {
        const auto x = to_global(reinterpret_cast<const T *>(xa)) + eA;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) nx[i][j] = x[i * 256 + j * 16];
    };
    auto factor = [&](const PlaneMeta &m, int k) {
        return to_const(reinterpret_cast<const T *>(m.f[k]));
    };
    Acc acc[GROUPED ? 4 : 1][4];  // running Y sum of the current run
    if constexpr (GROUPED) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = 0;
    }
    if (cnt > 0) {
        win.refill(ent, ord, 0, cnt, rbeg, rstep, l);
        load_x(win.at(ent, 0, 0, rbeg, rstep).x);
    }
    // The factor-line loads' values are folded into `touch` one iteration
    // later (no wait right after the load) and kept live to the end, so the
    // loads are issued.
    unsigned touch = 0, tval = 0;

    for (int j = 0; j < cnt; ++j) {
        const PlaneMeta cur = win.at(ent, j, wb, rbeg, rstep);
        const bool more = j + 1 < cnt;
        Acc v[4][4];
        // pass A: modes 2 and 4 of X, into LDS
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) v[i][q] = Acc(nx[i][q]);
        if (more) {
            if (PlaneWindow<T, LAY, GROUPED>::USED && j + 1 - wb == 32) {
                wb += 32;
                win.refill(ent, ord, wb, cnt, rbeg, rstep, l);
            }
            load_x(win.at(ent, j + 1, wb, rbeg, rstep).x);
        }
        load_factor<T, Acc, 4, TRANS, true>(a, factor(cur, 1), lda);
        plane_first(v, a);
        load_factor<T, Acc, 4, TRANS, true>(a, factor(cur, 3), lda);
        plane_second(v, a);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) LA[272 * i + 17 * q] = T(v[i][q]);
        __syncthreads();
        if (PlaneWindow<T, LAY, GROUPED>::USED && more && w == 0) {
            const PlaneMeta nm = win.at(ent, j + 1, wb, rbeg, rstep);
            unsigned long long p = nm.f[0];
#pragma unroll
            for (int k = 1; k < 6; ++k) p = l == k ? nm.f[k] : p;
            // every lane loads (lanes 6.. repeat factor 0's line): a masked
            // load would make the compiler wait for it to merge the lanes
            touch ^= tval;
            tval = *to_global(reinterpret_cast<const unsigned *>(p));
        }
        // pass B: modes 6 and 5, in place
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) v[i][q] = Acc(LB[i + 4 * q]);
        load_factor<T, Acc, 4, TRANS, true>(a, factor(cur, 5), lda);
        plane_first(v, a);
        load_factor<T, Acc, 4, TRANS, true>(a, factor(cur, 4), lda);
        plane_second(v, a);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) LB[i + 4 * q] = T(v[i][q]);
        __syncthreads();
        // pass C: modes 3 and 1, into Y.  A private Y (read-modify-write) is
        // read before the pass's math, so the loads overlap it.
        constexpr bool RMW = !GROUPED && !SHARED;
        T yv[RMW ? 4 : 1][4];
        if constexpr (RMW) {
            const auto y = to_global(reinterpret_cast<T *>(cur.y)) + eC;
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int q = 0; q < 4; ++q) yv[i][q] = y[64 * i + 1024 * q];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) v[i][q] = Acc(LC[68 * i + 1088 * q]);
        load_factor<T, Acc, 4, TRANS, true>(a, factor(cur, 2), lda);
        plane_first(v, a);
        load_factor<T, Acc, 4, TRANS, true>(a, factor(cur, 0), lda);
        if constexpr (GROUPED) {
            // the run's running sum is the mode-1 contraction's starting
            // value (no separate adds)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                Acc t[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    Acc s = kfma(a[q][0], v[i][0], acc[i][q]);
#pragma unroll
                    for (int k = 1; k < 4; ++k) s = kfma(a[q][k], v[i][k], s);
                    t[q] = s;
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[i][q] = t[q];
            }
            if (!more || win.at(ent, j + 1, wb, rbeg, rstep).y != cur.y) {
                const auto y = to_global(reinterpret_cast<T *>(cur.y)) + eC;
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        atomic_add(y + (64 * i + 1024 * q), T(acc[i][q]));
                        acc[i][q] = 0;
                    }
            }
        } else {
            plane_second(v, a);
            const auto y = to_global(reinterpret_cast<T *>(cur.y)) + eC;
            if constexpr (SHARED) {
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int q = 0; q < 4; ++q) atomic_add(y + (64 * i + 1024 * q), T(v[i][q]));
            } else {
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        y[64 * i + 1024 * q] = T(Acc(yv[i][q]) + v[i][q]);
            }
        }
        __syncthreads();  // LDS is rewritten by the next entry
    }
    asm volatile("" ::"v"(touch ^ tval));
}

}  // namespace kron
