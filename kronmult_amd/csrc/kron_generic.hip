// kron_generic.hip -- runtime-shape kernel and the forward dispatcher.
//
// The generic kernel covers every shape the compile-time fast kernel does
// not: rectangular factors (vbatched, kronmult_vbatched.hpp:63-68), n > 8,
// and entries whose n^d does not fit in LDS.  One 256-thread workgroup owns an
// entry at a time.
//
//   single pass (h == 0): stage X in LDS, contract modes 1..d with LDS
//       ping-pong buffers, add the result into Y.
//   two passes  (0 < h < d), when the entry is larger than LDS:
//       pass 1: for chunks of consecutive low indices, contract the slow modes
//               1..h in LDS and store the chunk to the entry's scratch vector;
//       pass 2: for chunks of slow-index slabs, contract the fast modes
//               h+1..d in LDS and add the contiguous slabs into Y.
//   The scratch holds m1..mh x n(h+1)..nd elements, one of the reference's
//   own W sizes (kronmult_vbatched.hpp:129-153), so the reference's workspace
//   rules suffice.  HBM traffic is X + 2 * scratch + Y RMW, against the
//   reference's 2*d passes over the whole vector.
//
// Inside a pass every output element is one work item: n_k LDS reads and n_k
// FMAs, consecutive threads on consecutive fast indices.
#include <algorithm>

#include "kron_dispatch.hpp"
#include "kron_fast.hpp"

namespace kron {

struct GenShape {
    int d;
    int h;                  // split point (0 = single pass)
    int m[kMaxDim];         // output extent of mode k (0-based, slowest first)
    int n[kMaxDim];         // input extent of mode k
    int ld[kMaxDim];        // leading dimension of factor k as stored
    int chunk1;             // low indices per pass-1 chunk
    int chunk2;             // slabs per pass-2 chunk
    int lds_elems;          // elements per LDS ping-pong buffer
    int fac_off[kMaxDim];   // LDS offset of each staged factor (elements)
    int fac_elems;          // LDS elements of all staged factors
    long long nlo_in;       // prod_{k>=h} n_k
    long long nhi_in;       // prod_{k<h}  n_k
    long long nhi_out;      // prod_{k<h}  m_k
    long long nlo_out;      // prod_{k>=h} m_k
};


// Elements of one work item's input held per thread while the previous item
// is contracted (the item's LDS chunk is at most kStagePer * 256 elements).
constexpr int kStagePer = 8;

// A work item is one chunk of pass 1 (slow modes over the consecutive low
// indices [c0, c0+C)) or of pass 2 (fast modes over the slabs [s0, s0+C)) of
// one entry.  Its input is first loaded into registers (RegStage), so the
// loads of item r+1 are in flight while item r is contracted in LDS.
template <typename T>
struct RegStage {
    T v[kStagePer];
    T y[kStagePer];  // pass 2 without atomics: the Y chunk to update
};

template <typename T, typename Acc, Layout LAY, bool SHARED, bool TRANS, int SQ>
struct GenericPasses {
    const Entries<T, LAY> &ent;
    const GenShape &sh;
    const ScratchDesc<T> &scr;
    int phase, b0;
    long long nck;
    T *buf0, *buf1, *facs;

    // LDS buffer b (0 or 1) of the ping-pong pair
    __device__ __forceinline__ T *buf(int b) const { return b ? buf1 : buf0; }
    // modes contracted by this pass
    __device__ __forceinline__ int modes() const { return phase == 1 ? sh.h : sh.d - sh.h; }

    __device__ __forceinline__ void decode(long long r, int &b, long long &c0, int &C) const
    {
        b = b0 + (int)(r / nck);
        const long long ck = r % nck;
        if (phase == 1) {
            c0 = ck * sh.chunk1;
            C = (int)min((long long)sh.chunk1, sh.nlo_in - c0);
        } else {
            c0 = ck * sh.chunk2;
            C = (int)min((long long)sh.chunk2, sh.nhi_out - c0);
        }
    }
    __device__ __forceinline__ const T *src_of(int b) const
    {
        if (phase == 1 || sh.h == 0) return ent.x(b);
        return scr.get(b - b0, b);
    }
    __device__ __forceinline__ int count(int C) const
    {
        return (phase == 1) ? (int)sh.nhi_in * C : C * (int)sh.nlo_in;
    }
    __device__ __forceinline__ void load(RegStage<T> &st, long long r) const
    {
        int b, C;
        long long c0;
        decode(r, b, c0, C);
        const auto src = to_global(src_of(b));
        const int cnt = count(C);
        if (phase == 1) {
            // rows of C consecutive low indices, nlo_in apart
            const auto s1 = src + c0;
            const FastDiv dc = fast_div(C);
        // The loads are unconditional (element indices clamped into the
        // chunk): every item issues the same number, so the wait for one item
        // can leave the next item's loads in flight.
#pragma unroll
            for (int k = 0; k < kStagePer; ++k) {
                const int w = min((int)threadIdx.x + k * 256, cnt - 1);
                int c;
                const int hi = split_fiber(w, dc, c);
                st.v[k] = s1[(long long)hi * sh.nlo_in + c];
            }
        } else {
            const auto s2 = src + c0 * sh.nlo_in;
#pragma unroll
            for (int k = 0; k < kStagePer; ++k) st.v[k] = s2[min((int)threadIdx.x + k * 256, cnt - 1)];
        }
        if constexpr (!SHARED) {
            // the item's Y chunk is read now, so its latency overlaps too
            if (phase == 2) {
                const auto y = to_global(ent.y(b)) + c0 * sh.nlo_out;
                const int nout = C * (int)sh.nlo_out;
#pragma unroll
                for (int k = 0; k < kStagePer; ++k)
                    st.y[k] = y[min((int)threadIdx.x + k * 256, nout - 1)];
            }
        }
    }
    __device__ __forceinline__ void to_lds(const RegStage<T> &st, int cnt, T *dst) const
    {
#pragma unroll
        for (int k = 0; k < kStagePer; ++k) {
            const int w = threadIdx.x + k * 256;
            if (w < cnt) dst[w] = st.v[k];
        }
    }
    __device__ __forceinline__ void stage_factors(int b) const
    {
        for (int k = 0; k < sh.d; ++k) {
            const int rows = TRANS ? sh.n[k] : sh.m[k], cols = TRANS ? sh.m[k] : sh.n[k];
            stage_factor(facs + sh.fac_off[k], ent.template factor_g<true>(b, k), rows, cols,
                         sh.ld[k]);
        }
    }
    // mode k of entry b: square factors of at most 288 bytes are read with
    // scalar loads into SGPRs (no VGPRs or LDS reads for them); otherwise
    // from the entry's factors staged in LDS
    __device__ __forceinline__ void contract(const T *src, T *dst, int outer, int inner, int b,
                                             int k) const
    {
        if constexpr (SQ > 0 && (size_t)SQ * SQ * sizeof(Acc) <= 288) {
            if ((int)threadIdx.x >= outer * inner) return;
            Acc a[SQ][SQ];
            load_factor<T, Acc, SQ, TRANS, true>(a, ent.template factor_g<true>(b, k), sh.ld[k]);
            contract_fibers_reg<T, Acc, SQ>(src, dst, outer, inner, a);
        } else {
            contract_mode_lds<T, Acc, TRANS, SQ>(src, dst, outer, sh.n[k], inner, sh.m[k],
                                                 facs + sh.fac_off[k],
                                                 TRANS ? sh.n[k] : sh.m[k]);
        }
    }
    // contract the item's modes in LDS, input in buf(in); the result ends in
    // buf(in ^ (modes() & 1))
    __device__ __forceinline__ void compute(int b, int C, int in) const
    {
        const int h = sh.h, d = sh.d;
        T *src = buf(in), *dst = buf(in ^ 1);
        if (phase == 1) {
            // the chunk is n_0 x .. x n_{h-1} x C; mode q sees outer = m_0..m_{q-1}
            // (already contracted) and inner = n_{q+1}..n_{h-1} x C
            int outer = 1;
            int inner = C;
            for (int k = 1; k < h; ++k) inner *= sh.n[k];
            for (int q = 0; q < h; ++q) {
                contract(src, dst, outer, inner, b, q);
                outer *= sh.m[q];
                if (q + 1 < h) inner /= sh.n[q + 1];
                __syncthreads();
                T *t = src; src = dst; dst = t;
            }
        } else {
            // the chunk is C x n_h x .. x n_{d-1}; mode k sees outer = C x m_h..m_{k-1}
            // and inner = n_{k+1}..n_{d-1}
            int outer = C;
            int inner = 1;
            for (int k = h + 1; k < d; ++k) inner *= sh.n[k];
            for (int k = h; k < d; ++k) {
                contract(src, dst, outer, inner, b, k);
                outer *= sh.m[k];
                if (k + 1 < d) inner /= sh.n[k + 1];
                __syncthreads();
                T *t = src; src = dst; dst = t;
            }
        }
    }
    // store the item's result (in res): the scratch chunk (pass 1) or Y
    __device__ __forceinline__ void store(int b, long long c0, int C, const T *res,
                                          const RegStage<T> &ycur) const
    {
        const int tid = threadIdx.x, nthr = blockDim.x;
        const T *src = res;
        if (phase == 1) {
            const auto S = to_global(scr.get(b - b0, b) + c0);
            const int nout = (int)sh.nhi_out * C;
            const FastDiv dc = fast_div(C);
#pragma unroll
            for (int k = 0; k < kStagePer; ++k) {
                const int w = tid + k * 256;
                if (w < nout) {
                    int c;
                    const int hi = split_fiber(w, dc, c);
                    S[(long long)hi * sh.nlo_in + c] = src[w];
                }
            }
        } else {
            const auto y = to_global(ent.y(b));
            const int nout = C * (int)sh.nlo_out;
            const long long ybase = c0 * sh.nlo_out;
            if constexpr (SHARED) {
                for (int w = tid; w < nout; w += nthr) atomic_add(y + (ybase + w), src[w]);
            } else {
#pragma unroll
                for (int k = 0; k < kStagePer; ++k) {
                    const int w = tid + k * 256;
                    if (w < nout) y[ybase + w] = T(Acc(ycur.y[k]) + Acc(src[w]));
                }
            }
        }
    }
};

// One pass (phase 1: slow modes into the scratch, phase 2: fast modes into Y)
// over the work items (entry, chunk) of the entries [b0, b1).  Each workgroup
// walks a contiguous range of items, so consecutive items usually belong to
// the same entry (factors are restaged only when the entry changes); entry b
// uses scratch slot b - b0.
template <typename T, typename Acc, Layout LAY, bool SHARED, bool TRANS, int SQ>
__global__ void __launch_bounds__(256)
kron_generic_kernel(Entries<T, LAY> ent, GenShape sh, ScratchDesc<T> scr, int phase, int b0,
                    int b1)
{
    extern __shared__ __align__(16) unsigned char kron_smem[];
    T *const buf0 = reinterpret_cast<T *>(kron_smem);
    T *const buf1 = buf0 + sh.lds_elems;
    T *const facs = buf1 + sh.lds_elems;  // the entry's factors, packed (ld = rows)
    const long long nck = (phase == 1) ? (sh.nlo_in + sh.chunk1 - 1) / sh.chunk1
                                       : (sh.nhi_out + sh.chunk2 - 1) / sh.chunk2;
    const GenericPasses<T, Acc, LAY, SHARED, TRANS, SQ> P{ent, sh, scr, phase, b0, nck,
                                                          buf0, buf1, facs};
    const long long items = (long long)(b1 - b0) * nck;
    const long long per = (items + gridDim.x - 1) / gridDim.x;
    const long long r0 = blockIdx.x * per;
    const long long r1 = min(items, r0 + per);
    // Two items in flight per workgroup: item r+2 is loaded into the
    // register stage item r came from while item r+1's loads are landing.
    RegStage<T> st[2];
    auto clampr = [&](long long q) { return q < r1 ? q : r1 - 1; };
    if (r0 < r1) {
        P.load(st[0], r0);
        P.load(st[1], clampr(r0 + 1));
    }
    int cur_b = -1;
    int in = 0;                      // LDS buffer holding the current item's input
    const int flip = P.modes() & 1;  // the result lands in buf(in ^ flip)
    auto item = [&](long long r, auto slot) __attribute__((always_inline)) {
        constexpr int S = decltype(slot)::value;  // register stage of item r
        int b, C;
        long long c0;
        P.decode(r, b, c0, C);
        if (b != cur_b) {
            P.stage_factors(b);
            cur_b = b;
        }
        if (r == r0) P.to_lds(st[S], P.count(C), P.buf(in));
        RegStage<T> ycur;
        if constexpr (!SHARED) {
#pragma unroll
            for (int k = 0; k < kStagePer; ++k) ycur.y[k] = st[S].y[k];
        }
        __syncthreads();
        P.load(st[S], clampr(r + 2));  // a repeat of the last item past the end
        P.compute(b, C, in);
        // The next item's input goes into the buffer the result is not in,
        // before this item's stores are issued: waiting for its loads then
        // does not also wait for those stores (the memory counter is in order).
        const int res = in ^ flip, next = res ^ 1;
        if (r + 1 < r1) {
            int nb, nC;
            long long nc0;
            P.decode(r + 1, nb, nc0, nC);
            P.to_lds(st[S ^ 1], P.count(nC), P.buf(next));
        }
        P.store(b, c0, C, P.buf(res), ycur);
        in = next;
        // (the barrier at the top of the next item orders these LDS reads
        // before that item's contractions overwrite buf(res))
    };
    for (long long r = r0; r < r1; r += 2) {
        item(r, std::integral_constant<int, 0>{});
        if (r + 1 < r1) item(r + 1, std::integral_constant<int, 1>{});
    }
}

// Choose the split and chunking for an LDS budget of `budget` elements per
// ping-pong buffer.  Returns false when no split fits.
static bool plan_generic(GenShape &sh, long long budget)
{
    const int d = sh.d;
    // largest intermediate of contracting modes [a, b) in ascending order on a
    // tensor whose other modes keep their input extents (mixed m-prefix).
    auto peak = [&](int a, int b) {
        long long best = 0;
        for (int k = a - 1; k < b; ++k) {
            long long s = 1;
            for (int l = a; l < b; ++l) s *= (l <= k) ? sh.m[l] : sh.n[l];
            best = std::max(best, s);
        }
        return best;
    };
    long long prod_n = 1;
    for (int k = 0; k < d; ++k) prod_n *= sh.n[k];
    const long long single = peak(0, d);
    if (single <= budget) {
        sh.h = 0;
        sh.chunk1 = 0;
        sh.chunk2 = 1;
        sh.lds_elems = (int)single;
        sh.nhi_in = sh.nhi_out = 1;
        sh.nlo_in = prod_n;
        long long mo = 1;
        for (int k = 0; k < d; ++k) mo *= sh.m[k];
        sh.nlo_out = mo;
        return true;
    }
    bool found = false;
    long long best_iters = 0;
    for (int h = 1; h < d; ++h) {
        long long nhi_in = 1, nhi_out = 1, nlo_in = 1, nlo_out = 1;
        for (int k = 0; k < h; ++k) { nhi_in *= sh.n[k]; nhi_out *= sh.m[k]; }
        for (int k = h; k < d; ++k) { nlo_in *= sh.n[k]; nlo_out *= sh.m[k]; }
        const long long per1 = peak(0, h), per2 = peak(h, d);
        long long c1 = std::min(nlo_in, budget / per1);
        long long c2 = std::min(nhi_out, budget / per2);
        if (c1 < 1 || c2 < 1) continue;
        // pass 1 reads rows of c1 consecutive elements: short rows waste HBM
        // bursts, so they count four times
        const long long iters = ((nlo_in + c1 - 1) / c1) * ((c1 < 16 && c1 < nlo_in) ? 4 : 1) +
                                (nhi_out + c2 - 1) / c2;
        if (!found || iters < best_iters) {
            found = true;
            best_iters = iters;
            sh.h = h;
            sh.chunk1 = (int)c1;
            sh.chunk2 = (int)c2;
            sh.lds_elems = (int)std::max(per1 * c1, per2 * c2);
            sh.nhi_in = nhi_in;
            sh.nhi_out = nhi_out;
            sh.nlo_in = nlo_in;
            sh.nlo_out = nlo_out;
        }
    }
    return found;
}

template <typename T>
static bool make_generic_shape(const KronArgs<T> &a, GenShape &sh)
{
    sh.d = a.d;
    for (int k = 0; k < kMaxDim; ++k) {
        // the operator applied is A_k (m_k x n_k) or A_k^T (n_k x m_k)
        sh.m[k] = a.trans ? a.n[k] : a.m[k];
        sh.n[k] = a.trans ? a.m[k] : a.n[k];
        sh.ld[k] = (a.lda > 0) ? a.lda : a.m[k];
    }
    int off = 0;
    for (int k = 0; k < kMaxDim; ++k) {
        sh.fac_off[k] = off;
        if (k < a.d) off += sh.m[k] * sh.n[k];
    }
    sh.fac_elems = (off + 1) & ~1;  // keep the buffers 16-byte aligned for doubles
    // Ping-pong buffers of at most kStagePer*256 elements (the register stage
    // of one work item): 2 x 16 KiB for doubles, four workgroups per CU.
    const long long budget = std::min<long long>(kStagePer * 256,
                                                 (max_lds_bytes() / 2) / (long long)sizeof(T));
    if (!plan_generic(sh, budget)) return false;
    return (2LL * sh.lds_elems + sh.fac_elems) * (long long)sizeof(T) <= max_lds_bytes();
}

template <typename T, Layout LAY, bool SHARED, bool TRANS, int SQ>
static kron_status launch_generic_sq(const Entries<T, LAY> &ent, const GenShape &sh,
                                     const ScratchDesc<T> &scr, int group, int batch,
                                     hipStream_t stream)
{
    auto kern = kron_generic_kernel<T, T, LAY, SHARED, TRANS, SQ>;
    const size_t lds = (2 * (size_t)sh.lds_elems + sh.fac_elems) * sizeof(T);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) { set_last_hip_error((int)e); return KRON_EHIP; }
    }
    const long long resident = persistent_blocks(reinterpret_cast<const void *>(kern), lds);
    const long long cap = resident > 0 ? 8 * resident : 65536;
    auto launch = [&](int phase, int b0, int b1, long long items) -> kron_status {
        const long long grid = std::max(1LL, std::min(items, cap / 8 > 0 ? cap / 8 : cap));
        hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(256), lds, stream, ent, sh, scr, phase,
                           b0, b1);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) { set_last_hip_error((int)e); return KRON_EHIP; }
        return KRON_OK;
    };
    const long long nck1 = sh.h > 0 ? (sh.nlo_in + sh.chunk1 - 1) / sh.chunk1 : 0;
    const long long nck2 = (sh.nhi_out + sh.chunk2 - 1) / sh.chunk2;
    // entries are processed in groups whose scratch vectors are all distinct
    for (int b0 = 0; b0 < batch; b0 += group) {
        const int b1 = std::min(batch, b0 + group);
        kron_status st;
        if (sh.h > 0) {
            st = launch(1, b0, b1, (long long)(b1 - b0) * nck1);
            if (st != KRON_OK) return st;
        }
        st = launch(2, b0, b1, (long long)(b1 - b0) * nck2);
        if (st != KRON_OK) return st;
    }
    return KRON_OK;
}

// Square shapes with n <= 8 get the compile-time contraction (factor in
// registers); n = 9..15 the row-block one (contract_rowblocks) in FP32, and
// in FP64 at d >= 4 and at d = 3 for n = 9, 12, where it measured 6-16%
// faster (batched; at FP64 d = 3, n = 10, 13 it was 9-11% slower, at n = 11,
// 15 even; in FP32 it was 3-30% faster at every d = 3 n); everything else the
// runtime-extent one.
template <typename T, Layout LAY, bool SHARED, bool TRANS>
static kron_status launch_generic_t(const Entries<T, LAY> &ent, const GenShape &sh,
                                    const ScratchDesc<T> &scr, int group, int batch,
                                    hipStream_t stream)
{
    bool square = true;
    for (int k = 0; k < sh.d; ++k) square = square && sh.m[k] == sh.n[0] && sh.n[k] == sh.n[0];
    const int n0 = sh.n[0];
    const bool rowblocks = sizeof(T) == 4 || sh.d >= 4 || n0 == 9 || n0 == 12;
    const int N = square && (n0 <= 8 || rowblocks) ? n0 : 0;
    switch (N) {
#define KRON_SQ(NN)                                                                          \
    case NN:                                                                                 \
        return launch_generic_sq<T, LAY, SHARED, TRANS, NN>(ent, sh, scr, group, batch, stream);
        KRON_SQ(2) KRON_SQ(3) KRON_SQ(4) KRON_SQ(5) KRON_SQ(6) KRON_SQ(7) KRON_SQ(8)
        KRON_SQ(9) KRON_SQ(10) KRON_SQ(11) KRON_SQ(12) KRON_SQ(13) KRON_SQ(14) KRON_SQ(15)
#undef KRON_SQ
    default:
        return launch_generic_sq<T, LAY, SHARED, TRANS, 0>(ent, sh, scr, group, batch, stream);
    }
}

template <typename T>
size_t kron_scratch_elems(const KronArgs<T> &a)
{
    bool square = true;
    for (int k = 0; k < a.d; ++k) square = square && (a.m[k] == a.n[k]) && a.m[k] == a.m[0];
    if (square && a.n[0] <= fast_max_n<T>(a.d)) {
        long long e = 1;
        for (int k = 0; k < a.d; ++k) e *= a.n[0];
        if (a.d == 1 || (long long)sizeof(T) * e <= max_lds_bytes()) return 0;
    }
    GenShape sh{};
    if (!make_generic_shape(a, sh)) return (size_t)-1;
    return sh.h == 0 ? 0 : (size_t)(sh.nhi_out * sh.nlo_in);
}

// Entries of a pointer-layout call that share Y targets are merged when the
// batch is large enough for the sort to pay off and the shape has a grouped
// kernel.  Failure to get the workspace falls back to per-entry atomics.
constexpr int kGroupMinBatch = 2048;

template <typename T>
static bool group_order(const KronArgs<T> &a, int N, GroupOrder &ord, hipStream_t s)
{
    if (!a.y_shared || a.trans || !fast_groupable_rt<T>(a.d, N)) return false;
    if (a.order) {
        ord = *a.order;
        return true;
    }
    if (a.batch < kGroupMinBatch) return false;
    const size_t bytes = group_workspace_bytes(a.batch);
    bool fresh = false;
    void *ws = nullptr;
    if (a.sort_ws) {
        // caller's scratch: its contents are whatever the last call left
        if (a.sort_ws_bytes < bytes) return false;
        ws = a.sort_ws;
        fresh = true;
    } else {
        ws = stream_workspace(s, bytes, &fresh);
    }
    if (!ws) return false;
    const unsigned long long *keys = nullptr;
    const int *perm = nullptr;
    if (group_by_target(reinterpret_cast<const void *const *>(a.pY), a.batch, ws, bytes, fresh,
                        &keys, &perm, s) != KRON_OK)
        return false;
    ord.perm = perm;
    ord.keys = keys;
    return true;
}

static kron_status fast_dispatch(const KronArgs<double> &a, int N, hipStream_t s)
{
    if (a.layout == Layout::Batched) {
        Entries<double, Layout::Batched> e{a.A, a.X, a.Y, a.d, (long long)N * N, 0, 0};
        long long E = 1;
        for (int k = 0; k < a.d; ++k) E *= N;
        e.sx = e.sy = E;
        return fast_batched_f64(a.d, N, e, a.trans, a.batch, s);
    }
    Entries<double, Layout::Ptr> e{a.pA, a.pX, a.pY, a.d};
    GroupOrder ord{nullptr, nullptr, 0};
    const bool grouped = group_order(a, N, ord, s);
    return fast_ptr_f64(a.d, N, e, a.lda > 0 ? a.lda : N, a.batch, s, grouped ? &ord : nullptr);
}

static kron_status fast_dispatch(const KronArgs<float> &a, int N, hipStream_t s)
{
    if (a.layout == Layout::Batched) {
        Entries<float, Layout::Batched> e{a.A, a.X, a.Y, a.d, (long long)N * N, 0, 0};
        long long E = 1;
        for (int k = 0; k < a.d; ++k) E *= N;
        e.sx = e.sy = E;
        return fast_batched_f32(a.d, N, e, a.trans, a.batch, s);
    }
    Entries<float, Layout::Ptr> e{a.pA, a.pX, a.pY, a.d};
    GroupOrder ord{nullptr, nullptr, 0};
    const bool grouped = group_order(a, N, ord, s);
    return fast_ptr_f32(a.d, N, e, a.lda > 0 ? a.lda : N, a.batch, s, grouped ? &ord : nullptr);
}

template <typename T>
bool kron_groupable(const KronArgs<T> &a)
{
    bool square = true;
    for (int k = 0; k < a.d; ++k) square = square && (a.m[k] == a.n[k]) && a.m[k] == a.m[0];
    if (!square || a.layout != Layout::Ptr || !a.y_shared || a.trans) return false;
    const int N = a.n[0];
    if (N > fast_max_n<T>(a.d)) return false;
    long long E = 1;
    for (int k = 0; k < a.d; ++k) E *= N;
    if ((long long)sizeof(T) * E > max_lds_bytes()) return false;
    return fast_groupable_rt<T>(a.d, N);
}

// The caller's scratch as the kernels address it (ScratchDesc), and how many
// consecutive entries may run at once with distinct scratch vectors (group):
// W(:,b) in the batched layout (wstride elements apart), pW[b] in pbatched,
// pW[2*slot] / W + slot*stride over subbatchCount slots in xbatched /
// vbatched.
template <typename T>
static kron_status scratch_of(const KronArgs<T> &a, long long wstride, ScratchDesc<T> &scr,
                              int &group)
{
    scr = ScratchDesc<T>{a.scratch, nullptr, 0, nullptr, nullptr, 0};
    group = a.batch;
    switch (a.scratch) {
    case Scratch::BatchedW:
        if (!a.W) return KRON_EWORKSPACE;
        scr.W = a.W;
        scr.wstride = wstride;
        break;
    case Scratch::PtrSlots:
    case Scratch::PtrSlot0:
        if (!a.pW) return KRON_EWORKSPACE;
        scr.pW = a.pW;
        group = std::min(a.batch, a.scratch == Scratch::PtrSlot0 ? 1 : a.nslots);
        break;
    case Scratch::PtrEntry:
        if (!a.pW) return KRON_EWORKSPACE;
        scr.pW = a.pW;
        break;
    case Scratch::Linear:
        if (!a.Wbase || a.nslots < 1) return KRON_EWORKSPACE;
        scr.Wbase = a.Wbase;
        scr.slot_stride = (long long)a.slot_stride;
        group = std::min(a.batch, a.nslots);
        break;
    default:
        return KRON_EWORKSPACE;
    }
    return group < 1 ? KRON_EWORKSPACE : KRON_OK;
}

template <typename T>
kron_status launch_kron(const KronArgs<T> &a)
{
    if (a.batch == 0) return KRON_OK;
    bool square = true;
    for (int k = 0; k < a.d; ++k) square = square && (a.m[k] == a.n[k]) && a.m[k] == a.m[0];
    // Fast path: square, compile-time (d, n), fits in LDS.  Transposed factors
    // are only compiled for the batched layout (the adjoint).
    // d = 1 from n = 8: the matrix-vector kernel beats the fused one (batched,
    // n = 8: 0.60 vs 0.80 ms for 4M entries; n = 4: 0.46 vs 0.42 ms)
    if (a.d == 1 && !a.trans && a.n[0] >= 8) return launch_gemv(a);
    if (square && (!a.trans || a.layout == Layout::Batched)) {
        kron_status st = fast_dispatch(a, a.n[0], a.stream);
        if (st != KRON_ENOTSUP) return st;
    }
    // square entries larger than LDS with factors in SGPRs (d = 6, n = 6):
    // two compile-time passes through the scratch
    if (square && !a.trans && split_supported(a.d, a.n[0])) {
        long long E = 1;
        for (int k = 0; k < a.d; ++k) E *= a.n[0];
        ScratchDesc<T> scr{Scratch::None, nullptr, 0, nullptr, nullptr, 0};
        int group = a.batch;
        kron_status st = scratch_of(a, E, scr, group);
        if (st != KRON_OK) return st;
        st = launch_split<T>(a, scr, group);
        if (st != KRON_ENOTSUP) return st;
    }
    // n >= 16: every mode a dense GEMM on the matrix cores
    if (square && !a.trans && a.layout == Layout::Batched && a.lda <= a.n[0] &&
        mfma_supported(a.d, a.n[0], nullptr)) {
        Entries<T, Layout::Batched> e{a.A, a.X, a.Y, a.d, (long long)a.n[0] * a.n[0], 0, 0};
        long long E = 1;
        for (int k = 0; k < a.d; ++k) E *= a.n[0];
        e.sx = e.sy = E;
        kron_status st = launch_mfma<T>(a.d, a.n[0], e, a.W, a.batch, a.stream);
        if (st != KRON_ENOTSUP) return st;
    }
    // pointer layouts, n = 16 / 32: also on the matrix cores (two-pass shapes
    // through the caller's scratch, group by group)
    if (square && !a.trans && a.layout == Layout::Ptr && mfma_supported(a.d, a.n[0], nullptr)) {
        Entries<T, Layout::Ptr> e{a.pA, a.pX, a.pY, a.d};
        bool needs_w = false;
        mfma_supported(a.d, a.n[0], &needs_w);
        ScratchDesc<T> scr{Scratch::None, nullptr, 0, nullptr, nullptr, 0};
        int group = a.batch;
        if (needs_w) {
            const kron_status st = scratch_of(a, 0, scr, group);
            if (st != KRON_OK) return st;
        }
        kron_status st = launch_mfma_ptr<T>(a.d, a.n[0], e, needs_w ? &scr : nullptr, group,
                                            a.lda > 0 ? a.lda : a.n[0], a.batch, a.stream);
        if (st != KRON_ENOTSUP) return st;
    }
    // d = 1: a batched matrix-vector product, one lane per output row
    if (a.d == 1 && !a.trans) return launch_gemv(a);
    // d = 2, extents up to 32: a wave per entry
    if (a.d == 2 && !a.trans) {
        kron_status st = launch_d2(a);
        if (st != KRON_ENOTSUP) return st;
    }
    // d >= 3 with mixed per-mode extents (vbatched): a wave per small entry
    if (!square && a.d >= 3 && !a.trans) {
        kron_status st = launch_small(a);
        if (st != KRON_ENOTSUP) return st;
    }
    GenShape sh{};
    if (!make_generic_shape(a, sh)) return KRON_ENOTSUP;
    ScratchDesc<T> scr{Scratch::None, nullptr, 0, nullptr, nullptr, 0};
    int grid = a.batch;  // entries whose scratch vectors are distinct (one launch group)
    if (sh.h > 0) {
        // batched is square here: W(:,b) holds sizeX elements
        const kron_status st = scratch_of(a, sh.nlo_in * sh.nhi_in, scr, grid);
        if (st != KRON_OK) return st;
    }
    if (a.layout == Layout::Batched) {
        Entries<T, Layout::Batched> e{a.A, a.X, a.Y, a.d, 0, 0, 0};
        e.fstride = (long long)a.m[0] * a.n[0];
        long long sx = 1, sy = 1;
        for (int k = 0; k < a.d; ++k) { sx *= a.n[k]; sy *= a.m[k]; }
        e.sx = sx;
        e.sy = sy;
        if (a.trans) {
            std::swap(e.sx, e.sy);
            return launch_generic_t<T, Layout::Batched, false, true>(e, sh, scr, grid, a.batch,
                                                                    a.stream);
        }
        return launch_generic_t<T, Layout::Batched, false, false>(e, sh, scr, grid, a.batch,
                                                                 a.stream);
    }
    Entries<T, Layout::Ptr> e{a.pA, a.pX, a.pY, a.d};
    if (a.trans) return KRON_ENOTSUP;
    return launch_generic_t<T, Layout::Ptr, true, false>(e, sh, scr, grid, a.batch, a.stream);
}

template <typename T>
__global__ void __launch_bounds__(256)
kron_factor_table(const T *A, long long fstride, long long count, const T **out)
{
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < count; i += (long long)gridDim.x * 256)
        out[i] = A + i * fstride;
}

template <typename T>
const T *const *factor_table(const T *A, long long fstride, long long count, hipStream_t s)
{
    const T **tab = static_cast<const T **>(
        stream_workspace_slot(s, (size_t)count * sizeof(T *), 1));
    if (!tab) return nullptr;
    const long long blocks = std::min<long long>((count + 255) / 256, 4096);
    hipLaunchKernelGGL(kron_factor_table<T>, dim3((unsigned)blocks), dim3(256), 0, s, A, fstride,
                       count, tab);
    if (hipGetLastError() != hipSuccess) return nullptr;
    return tab;
}

template const double *const *factor_table<double>(const double *, long long, long long,
                                                   hipStream_t);
template const float *const *factor_table<float>(const float *, long long, long long,
                                                 hipStream_t);
template kron_status launch_kron<double>(const KronArgs<double> &);
template kron_status launch_kron<float>(const KronArgs<float> &);
template size_t kron_scratch_elems<double>(const KronArgs<double> &);
template bool kron_groupable<double>(const KronArgs<double> &);
template bool kron_groupable<float>(const KronArgs<float> &);
template size_t kron_scratch_elems<float>(const KronArgs<float> &);

}  // namespace kron
