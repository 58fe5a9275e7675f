// kron_mfma.hip -- the large-n path: every mode contraction a dense GEMM on
// the matrix cores (MFMA), FP64 and FP32, batched layout, n in {16, 32}.
//
// At n >= 16 a mode contraction out(o, i, ii) = sum_j A(i, j) in(o, j, ii) is a
// real GEMM: M = K = n, and (o, ii) gives n^(d-1) columns.  The fused kernels'
// fiber-per-thread scheme would need the whole n x n factor per thread; here
// a wave computes 16 x 16 output blocks with v_mfma_{f64,f32}_16x16x4, the
// factor held as MFMA A-fragments (n/16 x n/4 values per lane) and the data
// read as B-fragments from an LDS tile.
//
// An entry (n^d elements, A1 on the slowest index) is split into slow modes
// 1..H and fast modes H+1..d, viewed as S x F (S = n^H, F = n^(d-H)):
//   pass 1 (row tiles): R consecutive slow rows, contiguous in HBM; contract
//          the fast modes in LDS; store the tile to the entry's W slice;
//   pass 2 (column tiles): all S rows x C consecutive fast columns of W;
//          contract the slow modes in LDS; add the tile into Y.
// An entry that fits one tile (n^d <= kMfmaTile) is done in one pass.  Every
// tile is an LDS array with PAD elements after each n (bank-conflict-free
// fragment reads and writes for every mode stride), filled from and drained
// to HBM with coalesced accesses; the next tile's HBM data is loaded into
// registers while the current one is contracted.  The two passes are
// separate launches (the launch boundary publishes W).  W is the reference's
// (n^d, batch) workspace (kronmult6_batched.cpp:5-11).
#include <algorithm>

#include "kron_device.hpp"
#include "kron_dispatch.hpp"

namespace kron {

constexpr int kMfmaTile = 8192;  // tile elements (64 KB in FP64 before padding)

template <typename T>
struct Mfma;

// v_mfma_f64_16x16x4f64: A[i][k] in lane i + 16k, B[k][j] in lane j + 16k,
// D[i][j] in lane j + 16*(i % 4), register i / 4.
template <>
struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(double a, double b, v4 c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
    static constexpr int PAD = 2;
};

// v_mfma_f32_16x16x4f32: A and B as above, D[i][j] in lane j + 16*(i / 4),
// register i % 4.
template <>
struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(float a, float b, v4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
    static constexpr int PAD = 3;
};

// Padded LDS position of tile element t.
template <typename T, int N>
__device__ __forceinline__ int mfma_pos(int t)
{
    return t + (t / N) * Mfma<T>::PAD;
}

// A-fragments of an n x n factor (column-major, leading dimension lda).
template <typename T, int N, typename P>
__device__ __forceinline__ void mfma_afrag(T (&af)[N / 16][N / 4], P A, int lda, int lane)
{
#pragma unroll
    for (int rb = 0; rb < N / 16; ++rb)
#pragma unroll
        for (int ks = 0; ks < N / 4; ++ks)
            af[rb][ks] = A[(rb * 16 + (lane & 15)) + (ks * 4 + (lane >> 4)) * lda];
}

// Contract the digit of stride INNER of an LDS tile of TOT elements in place:
// out(o, i, ii) = sum_j A(i, j) in(o, j, ii), column c = o * INNER + ii.  Each
// wave owns whole 16-column blocks: it reads all of a block's B-fragments
// before writing its outputs, and no other wave touches those columns.
template <typename T, int N, int TOT, int INNER>
__device__ __forceinline__ void mfma_mode(T *tile, const T (&af)[N / 16][N / 4], int wave,
                                          int lane)
{
    using M = Mfma<T>;
    constexpr int CB = TOT / N / 16;
    static_assert(TOT % (16 * N) == 0, "tile columns must fill 16-column blocks");
    for (int cb = wave; cb < CB; cb += 4) {
        const int c = cb * 16 + (lane & 15);
        const int base = (c / INNER) * N * INNER + c % INNER;
        T b[N / 4];
#pragma unroll
        for (int ks = 0; ks < N / 4; ++ks)
            b[ks] = tile[mfma_pos<T, N>(base + (ks * 4 + (lane >> 4)) * INNER)];
#pragma unroll
        for (int rb = 0; rb < N / 16; ++rb) {
            typename M::v4 acc = {0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < N / 4; ++ks) acc = M::mma(af[rb][ks], b[ks], acc);
#pragma unroll
            for (int r = 0; r < 4; ++r)
                tile[mfma_pos<T, N>(base + (rb * 16 + M::row(lane, r)) * INNER)] = acc[r];
        }
    }
}

// Geometry of (D, N): the split H and the tiles of both passes.
template <typename T, int D, int N>
struct MfmaGeom {
    static constexpr int E = cpow(N, D);
    static constexpr bool ONE = E <= kMfmaTile;  // single pass, one entry per tile
    static constexpr int pick_h()
    {
        for (int h = 1; h < D; ++h)
            if (cpow(N, D - h) <= kMfmaTile && cpow(N, h) * 16 <= kMfmaTile) return h;
        return -1;
    }
    static constexpr int H = ONE ? 0 : pick_h();
    static constexpr bool OK = ONE || H > 0;
    static constexpr int S = cpow(N, H > 0 ? H : 0), F = E / S;
    static constexpr int R = ONE ? 1 : kMfmaTile / F;   // rows per pass-1 tile
    static constexpr int C = ONE ? 0 : kMfmaTile / S;   // columns per pass-2 tile
    static constexpr int TOT1 = ONE ? E : R * F;
    static constexpr int TOT2 = ONE ? 0 : S * C;
    static constexpr int tiles(int pass) { return ONE ? 1 : pass == 1 ? S / R : F / C; }
    static constexpr size_t lds_bytes()
    {
        const int tot = TOT1 > TOT2 ? TOT1 : TOT2;
        return (size_t)(tot + (tot / N) * Mfma<T>::PAD) * sizeof(T);
    }
};

// pass 0: whole entries (X -> Y); 1: fast modes (X -> W); 2: slow modes (W -> Y)
// over the entries [b0, b1), entry b's scratch vector scr.get(b - b0, b).
// SHARED: entries may share Y, which is then updated with device atomics.
template <typename T, int D, int N, int PASS, Layout LAY = Layout::Batched, bool SHARED = false>
__global__ void __launch_bounds__(256, 2)
kron_mfma_kernel(Entries<T, LAY> ent, ScratchDesc<T> scr, int lda, int b0, int b1)
{
    using G = MfmaGeom<T, D, N>;
    constexpr int TOT = PASS == 2 ? G::TOT2 : G::TOT1;
    constexpr int PER = TOT / 256;  // elements per thread
    static_assert(TOT % 256 == 0, "tile must spread over 256 threads");
    constexpr int NT = G::tiles(PASS == 2 ? 2 : 1);
    extern __shared__ __align__(16) unsigned char kron_smem[];
    T *tile = reinterpret_cast<T *>(kron_smem);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const long long items = (long long)(b1 - b0) * NT;
    const long long E = G::E;

    // HBM offset (within the entry) of tile element t of tile k
    auto gofs = [&](int k, int t) -> long long {
        if constexpr (PASS == 2)
            return (long long)(t / G::C) * G::F + (long long)k * G::C + t % G::C;
        else
            return (long long)k * G::TOT1 + t;
    };
    // entry b's scratch vector (batched: W(:,b), a compile-time stride)
    auto w_of = [&](int b) -> T * {
        if constexpr (LAY == Layout::Batched)
            return scr.W + (long long)b * E;
        else
            return scr.get(b - b0, b);
    };
    auto src_of = [&](int b) -> gptr<const T> {
        if constexpr (PASS == 2)
            return to_global((const T *)w_of(b));
        else
            return to_global(ent.x(b));
    };
    // Pass 2 walks the entries last to first, so it starts on the W slices
    // pass 1 wrote last (the ones the Infinity Cache may still hold); about
    // 1% faster at d=3 n=32 and d=4 n=16.
    auto entry_of = [&](long long it) -> int {
        const int q = (int)(it / NT);
        return PASS == 2 ? b1 - 1 - q : b0 + q;
    };
    T v[PER];
    auto fetch = [&](long long it) {
        const int b = entry_of(it), k = (int)(it % NT);
        const auto src = src_of(b);
#pragma unroll
        for (int q = 0; q < PER; ++q) v[q] = src[gofs(k, tid + q * 256)];
    };
    long long it = blockIdx.x;
    if (it < items) fetch(it);
    for (; it < items; it += gridDim.x) {
        const int b = entry_of(it), k = (int)(it % NT);
#pragma unroll
        for (int q = 0; q < PER; ++q) tile[mfma_pos<T, N>(tid + q * 256)] = v[q];
        __syncthreads();
        if (it + gridDim.x < items) fetch(it + gridDim.x);
        // digits of the tile: pass 0/1 contract modes [first, D) with inner
        // strides N^(D-1-k) inside each row; pass 2 modes [0, H) with strides
        // N^(H-1-k) * C
        T af[N / 16][N / 4];
        constexpr int K0 = PASS == 2 ? 0 : G::H;
        constexpr int K1 = PASS == 2 ? G::H : D;
        for_modes_desc<K1 - 1, K0>([&](auto kc) {
            constexpr int KM = decltype(kc)::value;
            constexpr int INNER = PASS == 2 ? cpow(N, G::H - 1 - KM) * G::C : cpow(N, D - 1 - KM);
            // the batched layout's factors are n x n (ld = N at compile time:
            // immediate load offsets, ~30 fewer VGPRs than a runtime ld)
            mfma_afrag<T, N>(af, to_global(ent.factor(b, KM)),
                             LAY == Layout::Batched ? N : lda, lane);
            // FP32: keep the fragment loads together ahead of the MFMAs; left
            // to itself the scheduler may pair each load with its MFMA behind
            // a full wait (d2n32: 0.211 ms, 0.135 ms grouped).  In FP64 the
            // barrier cost d = 2 about 15% (d2n16 0.247 -> 0.316 ms).
            if constexpr (sizeof(T) == 4) __builtin_amdgcn_sched_barrier(0);
            mfma_mode<T, N, TOT, INNER>(tile, af, wave, lane);
            __syncthreads();
        });
        if constexpr (PASS == 1) {
            const auto w = to_global(w_of(b));
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                const int t = tid + q * 256;
                w[gofs(k, t)] = tile[mfma_pos<T, N>(t)];
            }
        } else if constexpr (SHARED) {
            const auto y = to_global(ent.y(b));
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                const int t = tid + q * 256;
                atomic_add(y + gofs(k, t), tile[mfma_pos<T, N>(t)]);
            }
        } else {
            // Y read-modify-write in chunks of 8 (the next tile's prefetch
            // already holds PER registers)
            const auto y = to_global(ent.y(b));
            constexpr int CH = PER < 8 ? PER : 8;
#pragma unroll
            for (int q0 = 0; q0 < PER; q0 += CH) {
                T yv[CH];
#pragma unroll
                for (int q = 0; q < CH; ++q) yv[q] = y[gofs(k, tid + (q0 + q) * 256)];
#pragma unroll
                for (int q = 0; q < CH; ++q) {
                    const int t = tid + (q0 + q) * 256;
                    y[gofs(k, t)] = yv[q] + tile[mfma_pos<T, N>(t)];
                }
            }
        }
        __syncthreads();  // the tile is rewritten next
    }
}

template <typename T, int D, int N, int PASS, Layout LAY = Layout::Batched, bool SHARED = false>
static kron_status mfma_launch(const Entries<T, LAY> &e, const ScratchDesc<T> &scr, int lda,
                               int b0, int b1, hipStream_t s)
{
    using G = MfmaGeom<T, D, N>;
    const size_t lds = G::lds_bytes();
    const void *kern =
        reinterpret_cast<const void *>(kron_mfma_kernel<T, D, N, PASS, LAY, SHARED>);
    if (lds > 65536) {
        hipError_t err = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)lds);
        if (err != hipSuccess) {
            set_last_hip_error((int)err);
            return KRON_EHIP;
        }
    }
    const long long items = (long long)(b1 - b0) * G::tiles(PASS == 2 ? 2 : 1);
    const long long resident = persistent_blocks(kern, lds);
    const long long grid = std::max(1LL, std::min(items, resident > 0 ? resident : 1024LL));
    hipLaunchKernelGGL((kron_mfma_kernel<T, D, N, PASS, LAY, SHARED>), dim3((unsigned)grid),
                       dim3(256), lds, s, e, scr, lda, b0, b1);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        set_last_hip_error((int)err);
        return KRON_EHIP;
    }
    return KRON_OK;
}

template <typename T, int D, int N>
static kron_status mfma_run(const Entries<T, Layout::Batched> &e, T *W, int batch,
                            hipStream_t s)
{
    using G = MfmaGeom<T, D, N>;
    if constexpr (!G::OK) {
        return KRON_ENOTSUP;
    } else {
        const ScratchDesc<T> scr{Scratch::BatchedW, W, G::E, nullptr, nullptr, 0};
        if constexpr (G::ONE) {
            return mfma_launch<T, D, N, 0>(e, scr, N, 0, batch, s);
        } else {
            if (!W) return KRON_EWORKSPACE;
            kron_status st = mfma_launch<T, D, N, 1>(e, scr, N, 0, batch, s);
            if (st != KRON_OK) return st;
            return mfma_launch<T, D, N, 2>(e, scr, N, 0, batch, s);
        }
    }
}

// Pointer layout, Y possibly shared (atomics).  Two-pass shapes run in groups
// of `group` consecutive entries whose scratch vectors are distinct: pass 1
// and pass 2 of one group, then the next group.
template <typename T, int D, int N>
static kron_status mfma_run_ptr(const Entries<T, Layout::Ptr> &e, const ScratchDesc<T> *scr,
                                int group, int lda, int batch, hipStream_t s)
{
    using G = MfmaGeom<T, D, N>;
    if constexpr (!G::OK) {
        return KRON_ENOTSUP;
    } else if constexpr (G::ONE) {
        const ScratchDesc<T> none{Scratch::None, nullptr, 0, nullptr, nullptr, 0};
        return mfma_launch<T, D, N, 0, Layout::Ptr, true>(e, none, lda, 0, batch, s);
    } else {
        if (!scr || group < 1) return KRON_EWORKSPACE;
        for (int b0 = 0; b0 < batch; b0 += group) {
            const int b1 = std::min(batch, b0 + group);
            kron_status st = mfma_launch<T, D, N, 1, Layout::Ptr, true>(e, *scr, lda, b0, b1, s);
            if (st != KRON_OK) return st;
            st = mfma_launch<T, D, N, 2, Layout::Ptr, true>(e, *scr, lda, b0, b1, s);
            if (st != KRON_OK) return st;
        }
        return KRON_OK;
    }
}

// Whether (d, n) has an MFMA kernel (and whether it needs W).
bool mfma_supported(int d, int n, bool *needs_w)
{
    bool ok = false, w = false;
#define KRON_MFMA_Q(DD, NN)                                                                  \
    if (d == DD && n == NN) {                                                                \
        ok = MfmaGeom<double, DD, NN>::OK;                                                   \
        w = !MfmaGeom<double, DD, NN>::ONE;                                                  \
    }
    KRON_MFMA_Q(2, 16) KRON_MFMA_Q(3, 16) KRON_MFMA_Q(4, 16) KRON_MFMA_Q(5, 16)
    KRON_MFMA_Q(2, 32) KRON_MFMA_Q(3, 32)
#undef KRON_MFMA_Q
    if (needs_w) *needs_w = w;
    return ok;
}

template <typename T>
kron_status launch_mfma(int d, int n, const Entries<T, Layout::Batched> &e, T *W, int batch,
                        hipStream_t s)
{
    if (batch == 0) return KRON_OK;
#define KRON_MFMA_CASE(DD, NN) \
    if (d == DD && n == NN) return mfma_run<T, DD, NN>(e, W, batch, s);
    KRON_MFMA_CASE(2, 16) KRON_MFMA_CASE(3, 16) KRON_MFMA_CASE(4, 16) KRON_MFMA_CASE(5, 16)
    KRON_MFMA_CASE(2, 32) KRON_MFMA_CASE(3, 32)
#undef KRON_MFMA_CASE
    return KRON_ENOTSUP;
}

template <typename T>
kron_status launch_mfma_ptr(int d, int n, const Entries<T, Layout::Ptr> &e, const ScratchDesc<T> *scr,
                            int group, int lda, int batch, hipStream_t s)
{
    if (batch == 0) return KRON_OK;
#define KRON_MFMA_PCASE(DD, NN) \
    if (d == DD && n == NN) return mfma_run_ptr<T, DD, NN>(e, scr, group, lda, batch, s);
    KRON_MFMA_PCASE(2, 16) KRON_MFMA_PCASE(3, 16) KRON_MFMA_PCASE(4, 16) KRON_MFMA_PCASE(5, 16)
    KRON_MFMA_PCASE(2, 32) KRON_MFMA_PCASE(3, 32)
#undef KRON_MFMA_PCASE
    return KRON_ENOTSUP;
}

template kron_status launch_mfma<double>(int, int, const Entries<double, Layout::Batched> &,
                                         double *, int, hipStream_t);
template kron_status launch_mfma<float>(int, int, const Entries<float, Layout::Batched> &,
                                        float *, int, hipStream_t);

template kron_status launch_mfma_ptr<double>(int, int, const Entries<double, Layout::Ptr> &,
                                             const ScratchDesc<double> *, int, int, int,
                                             hipStream_t);
template kron_status launch_mfma_ptr<float>(int, int, const Entries<float, Layout::Ptr> &,
                                            const ScratchDesc<float> *, int, int, int,
                                            hipStream_t);

}  // namespace kron
