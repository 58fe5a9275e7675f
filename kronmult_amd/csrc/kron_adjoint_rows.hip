// kron_adjoint_rows.hip -- the d = 6, n = 6 adjoint with every extent known
// at compile time.
//
// The adjoint of Y = K X (kron_adjoint.hip) splits the modes into slow
// S = {1, 2} and fast F = {3..6} and views an entry as 36 rows x 1,296
// contiguous columns.  Its scratch holds two vectors per entry, V then U,
// each column-chunk-major: column chunk k (all 36 rows of its C2 = 36
// columns) is one contiguous 36 x 36 block.  The passes are:
//   pass 1: V = (I_S (x) A_F^T) dY, one workgroup per row: the row in LDS,
//           the four fast modes contracted in place with transposed factors;
//   pass 3: the fast-mode gradients dA_3..dA_6 from U and dY.  For fast mode
//           k, P_k is U contracted by the fast factors after k and G_k is dY
//           contracted by the transposed fast factors before k; dA_k(p, q) is
//           the sum over every other index of G_k[..p..] P_k[..q..].  A
//           workgroup walks a contiguous range of (entry, row) items with the
//           next row's loads in flight; each of its four waves owns one
//           gradient and keeps a 6 x 6 block of FP64 sums per lane (outer
//           products of a G fiber and a P fiber), flushed when the entry
//           changes.
//   pass 2: U, dX and the slow-mode gradients dA_1, dA_2 over column chunks
//           of 36 columns (kron_adj6_cols).
#include <algorithm>

#include "kron_adjoint.hpp"
#include "kron_split.hpp"

namespace kron {

namespace {

constexpr int kA6N = 6, kA6R = 36, kA6C = 1296, kA6E = 46656, kA6F = kA6C / kA6N;
constexpr int kA6Per = (kA6C + 255) / 256;  // row elements per thread

// Scratch offset of element (r, c) of a column-chunk-major vector.
__device__ __forceinline__ int a6_chunk_off(int r, int c, int C2)
{
    const int k = c / C2;
    return (k * kA6R + r) * C2 + (c - k * C2);
}

// Factor m of entry b (or its transpose) in SGPRs; contractions run in the
// element type, the gradients' sums in FP64.
// The address goes through an empty asm, so a factor read again later (the
// passes use A4 and A5 twice) is loaded again instead of being kept in
// SGPRs from its first use.
// The batched layout's operands as the d6n6 kernels take them: base
// pointers only, every stride a compile-time constant (runtime ones, and
// unused fields, cost SGPRs that the factors need).
template <typename T>
struct A6Batched {
    const T *A, *X, *dY;
    T *dX, *dA;
    static constexpr bool kAccumulate = false;
    __device__ __forceinline__ static void put_dx(gptr<T> p, int i, T v) { p[i] = v; }
};
template <typename T, Layout LAY>
using A6IO = std::conditional_t<LAY == Layout::Batched, A6Batched<T>, AdjIO<T, Layout::Ptr>>;

// Per-entry addresses (IO: A6Batched or AdjIO<T, Ptr>).
template <typename IO>
__device__ __forceinline__ auto a6_fac(const IO &io, int b, int m)
{
    if constexpr (!IO::kAccumulate)
        return io.A + ((long long)b * 6 + m) * 36;
    else
        return io.fac(b, m);
}
template <typename IO>
__device__ __forceinline__ auto a6_x(const IO &io, int b)
{
    if constexpr (!IO::kAccumulate)
        return to_global(io.X + (long long)b * kA6E);
    else
        return io.x(b);
}
template <typename IO>
__device__ __forceinline__ auto a6_dy(const IO &io, int b)
{
    if constexpr (!IO::kAccumulate)
        return to_global(io.dY + (long long)b * kA6E);
    else
        return io.dy(b);
}
template <typename IO>
__device__ __forceinline__ auto a6_dx(const IO &io, int b)
{
    if constexpr (!IO::kAccumulate)
        return to_global(io.dX + (long long)b * kA6E);
    else
        return io.dx(b);
}
// dA_k(i, j) += v, w = 36 k + i + 6 j
template <typename IO, typename T>
__device__ __forceinline__ void a6_add_da(const IO &io, int b, int w, T v)
{
    if constexpr (!IO::kAccumulate)
        atomic_add(to_global(io.dA + (long long)b * 216 + w), v);
    else
        io.add_da(b, w, v);
}

template <typename T, bool TRANS, typename IO>
__device__ __forceinline__ void a6_factor(T (&a)[kA6N][kA6N], const IO &io, int b, int m)
{
    const T *p = a6_fac(io, b, m);
    asm volatile("" : "+s"(p));
    if constexpr (IO::kAccumulate) {
        if (io.lda() != kA6N) {
            load_factor<T, T, kA6N, TRANS, true>(a, to_const(p), io.lda());
            return;
        }
    }
    // 36 contiguous elements as 64-byte scalar loads (s_load_dwordx16; left
    // to itself the compiler splits them into many narrow loads, which
    // lengthens every workgroup's wait for its factor).  The vector types
    // keep the element's alignment.
    constexpr int V = 64 / sizeof(T);
    typedef T vec __attribute__((ext_vector_type(V), aligned(sizeof(T))));
    typedef T vec4 __attribute__((ext_vector_type(4), aligned(sizeof(T))));
    T f[kA6R];
    const cptr<vec> pv = (cptr<vec>)to_const(p);
#pragma unroll
    for (int c = 0; c < kA6R / V; ++c) {
        const vec x = pv[c];
#pragma unroll
        for (int e = 0; e < V; ++e) f[c * V + e] = x[e];
    }
#pragma unroll
    for (int c = kA6R / V * V; c < kA6R; c += 4) {
        const vec4 x = ((cptr<vec4>)to_const(p + c))[0];
#pragma unroll
        for (int e = 0; e < 4; ++e) f[c + e] = x[e];
    }
#pragma unroll
    for (int j = 0; j < kA6N; ++j)
#pragma unroll
        for (int i = 0; i < kA6N; ++i) a[i][j] = TRANS ? f[j + i * kA6N] : f[i + j * kA6N];
}

}  // namespace

// Pass 1: workgroup (entry blockIdx.x / 36, row blockIdx.x % 36).
template <typename T, Layout LAY>
__global__ void __launch_bounds__(256)
kron_adj6_rows_v(A6IO<T, LAY> io, T *__restrict__ W, int C2)
{
    __shared__ T row[kA6C];
    // Workgroups are dealt to the 8 XCDs round robin; a contiguous range of
    // rows per XCD lets the L2 of one XCD merge the 288-byte runs that
    // neighbouring rows write into the same lines of V.
    int bid = blockIdx.x;
    if ((gridDim.x & 7) == 0) bid = (bid & 7) * (gridDim.x >> 3) + (bid >> 3);
    const int b = bid / kA6R, r = bid % kA6R, tid = threadIdx.x;
    const auto g = a6_dy(io, b) + r * kA6C;
    T v[kA6Per];
#pragma unroll
    for (int k = 0; k < kA6Per; ++k) {
        const int i = tid + k * 256;
        if (i < kA6C) v[k] = g[i];
    }
#pragma unroll
    for (int k = 0; k < kA6Per; ++k) {
        const int i = tid + k * 256;
        if (i < kA6C) row[i] = v[k];
    }
    __syncthreads();
    T a[kA6N][kA6N];
    a6_factor<T, true>(a, io, b, 2);
    split_contract_lds<T, T, kA6N, 216, kA6C>(row, a);
    __syncthreads();
    a6_factor<T, true>(a, io, b, 3);
    split_contract_lds<T, T, kA6N, 36, kA6C>(row, a);
    __syncthreads();
    a6_factor<T, true>(a, io, b, 4);
    split_contract_lds<T, T, kA6N, 6, kA6C>(row, a);
    __syncthreads();
    a6_factor<T, true>(a, io, b, 5);
    split_contract_lds<T, T, kA6N, 1, kA6C>(row, a);
    __syncthreads();
    const auto V = to_global(W + 2LL * kA6E * b);
#pragma unroll
    for (int k = 0; k < kA6Per; ++k) {
        const int i = tid + k * 256;
        if (i < kA6C) V[a6_chunk_off(r, i, C2)] = row[i];
    }
}

// LDS layouts of a row (slot of element e).  A wave's 64-bit LDS accesses
// serve 16 lanes at a time over 16 bank pairs; the fibers of one stride S
// that 16 consecutive threads own must then sit in 16 distinct bank pairs.
//   Plain: e.  Stride 216 is conflict-free, stride 1 too when a fiber is
//          read as three 16-byte accesses (8 lanes at a time over 8 bank
//          quads); strides 6 and 36 put two lanes on one pair.
//   Pad:   e + 2 (e / 36) (1,366 slots).  Strides 6 and 36 are conflict-free
//          (fiber f's first slot is f plus a multiple of 16 modulo 16); 1
//          and 216 are not.
// Every fiber keeps compile-time offsets between its elements in both.
enum class A6Lay { Plain, Pad };
constexpr int kA6Row = kA6C + 2 * (kA6C / 36 - 1);  // slots of a Pad row

template <A6Lay L>
__device__ __forceinline__ int a6_slot(int e)
{
    return L == A6Lay::Pad ? e + 2 * (e / 36) : e;
}
// slot distance between consecutive elements of a stride-S fiber
template <A6Lay L, int S>
constexpr int a6_step()
{
    return L == A6Lay::Plain ? S : S == 1 ? 1 : S == 6 ? 6 : S == 36 ? 38 : 228;
}
// 16-byte accesses: stride-1 FP64 fibers of a Plain row (first slot 6 f)
template <typename T, A6Lay L, int S>
constexpr bool a6_wide()
{
    return L == A6Lay::Plain && S == 1 && sizeof(T) == 8;
}
typedef double a6_d2 __attribute__((ext_vector_type(2)));

template <typename T, A6Lay L, int S>
__device__ __forceinline__ void a6_read_fiber(const T *buf, int base, T (&v)[kA6N])
{
    const T *p = buf + a6_slot<L>(base);
    if constexpr (a6_wide<T, L, S>()) {
#pragma unroll
        for (int k = 0; k < kA6N / 2; ++k) {
            const a6_d2 t = reinterpret_cast<const a6_d2 *>(p)[k];
            v[2 * k] = t.x;
            v[2 * k + 1] = t.y;
        }
    } else {
#pragma unroll
        for (int j = 0; j < kA6N; ++j) v[j] = p[j * a6_step<L, S>()];
    }
}

template <typename T, A6Lay L, int S>
__device__ __forceinline__ void a6_write_fiber(T *buf, int base, const T (&v)[kA6N])
{
    T *p = buf + a6_slot<L>(base);
    if constexpr (a6_wide<T, L, S>()) {
#pragma unroll
        for (int k = 0; k < kA6N / 2; ++k) {
            a6_d2 t;
            t.x = v[2 * k];
            t.y = v[2 * k + 1];
            reinterpret_cast<a6_d2 *>(p)[k] = t;
        }
    } else {
#pragma unroll
        for (int j = 0; j < kA6N; ++j) p[j * a6_step<L, S>()] = v[j];
    }
}

// dst fiber f (stride S) = F src fiber f, f < 216 = thread id (in place when
// dst == src and the layouts agree); src in layout LI, dst in LO.
template <typename T, int S, A6Lay LI = A6Lay::Plain, A6Lay LO = A6Lay::Plain>
__device__ __forceinline__ void a6_contract(const T *src, T *dst, const T (&a)[kA6N][kA6N])
{
    const int f = threadIdx.x;
    if (f >= kA6F) return;
    const int base = (f / S) * (S * kA6N) + f % S;
    T in[kA6N], out[kA6N];
    a6_read_fiber<T, LI, S>(src, base, in);
#pragma unroll
    for (int i = 0; i < kA6N; ++i) {
        T s = a[i][0] * in[0];
#pragma unroll
        for (int j = 1; j < kA6N; ++j) s = kfma(a[i][j], in[j], s);
        out[i] = s;
    }
    a6_write_fiber<T, LO, S>(dst, base, out);
}

// Half of one gradient's positions (fibers of stride S) of the row,
// lane + 64 i of the half `half` of each pair of waves: acc[p][q] += G[p] P[q].
template <typename T, int S>
__device__ __forceinline__ void a6_gram(const T *G, const T *P, int half,
                                        double (&acc)[kA6N][kA6N])
{
    constexpr int H2 = kA6F / 2;  // 108 positions per wave
    const int lane = threadIdx.x & 63;
#pragma unroll 1
    for (int f0 = 0; f0 < H2; f0 += 64) {
        const int fl = f0 + lane;
        if (fl < H2) {
            const int f = half * H2 + fl;
            const int base = (f / S) * (S * kA6N) + f % S;
            double g[kA6N], p[kA6N];
#pragma unroll
            for (int j = 0; j < kA6N; ++j) {
                g[j] = double(G[base + j * S]);
                p[j] = double(P[base + j * S]);
            }
#pragma unroll
            for (int i = 0; i < kA6N; ++i)
#pragma unroll
                for (int j = 0; j < kA6N; ++j) acc[i][j] = kfma(g[i], p[j], acc[i][j]);
        }
    }
}

// Rows p0 .. p0+2 of one gradient over every position (fiber of stride S)
// of the row, lane + 64 i: acc[p][q] += G[p0 + p] P[q]; G in layout LG, P
// in LP.
template <typename T, int S, A6Lay LG = A6Lay::Plain, A6Lay LP = A6Lay::Plain>
__device__ __forceinline__ void a6_gram_rows(const T *G, const T *P, int p0,
                                             double (&acc)[3][kA6N])
{
    const int lane = threadIdx.x & 63;
#pragma unroll 1
    for (int f = lane; f < kA6F; f += 64) {
        const int base = (f / S) * (S * kA6N) + f % S;
        double g[3], p[kA6N];
        const T *gp = G + a6_slot<LG>(base);
        if constexpr (a6_wide<T, LG, S>()) {
            // elements p0 .. p0 + 2: one 16-byte pair and a single
            if (p0 == 0) {
                const a6_d2 t = reinterpret_cast<const a6_d2 *>(gp)[0];
                g[0] = t.x;
                g[1] = t.y;
                g[2] = gp[2];
            } else {
                const a6_d2 t = reinterpret_cast<const a6_d2 *>(gp)[2];
                g[0] = gp[3];
                g[1] = t.x;
                g[2] = t.y;
            }
        } else {
#pragma unroll
            for (int i = 0; i < 3; ++i) g[i] = double(gp[(p0 + i) * a6_step<LG, S>()]);
        }
        {
            T pv[kA6N];
            a6_read_fiber<T, LP, S>(P, base, pv);
#pragma unroll
            for (int j = 0; j < kA6N; ++j) p[j] = double(pv[j]);
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < kA6N; ++j) acc[i][j] = kfma(g[i], p[j], acc[i][j]);
    }
}

// Pass 3 over the items [r0, r1) of a workgroup, item = entry * 36 + row.
// Each gradient is split over a pair of waves by rows (p = 0..2 and 3..5):
// waves 0, 1 take dA_4 then dA_3, waves 2, 3 dA_5 then dA_6, so that every
// wave works in both gradient phases and keeps two 3 x 6 blocks of FP64 sums
// per lane.  Four LDS rows of 1,366 slots (44 KB: three workgroups per CU),
// each intermediate in the layout its accesses favour (Plain / Pad):
//   stage     B0 = P6 (U, Plain), B1 = G3 (dY, Plain)
//   step 1    B2 = P5 = A6 . P6 (mode 6, Pad),   B3 = G4 = A3^T . G3 (mode 3, Pad)
//   step 2    B0 = P4 = A5 . P5 (mode 5, Pad),   B1 = G5 = A4^T . G4 (mode 4, Pad)
//   grams     dA_4 (G4, P4), dA_5 (G5, P5)
//   step 3    B2 = P3 = A4 . P4 (mode 4, Plain), B3 = G6 = A5^T . G5 (mode 5, Plain)
//   restage   B0 = U, B1 = dY again, from the registers they were staged from
//   grams     dA_3 (G3, P3), dA_6 (G6, P6)
// The next item's rows are loaded into those registers after the restage.
template <typename T, Layout LAY>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
kron_adj6_rows_grams(A6IO<T, LAY> io, const T *__restrict__ W, int C2, int batch)
{
    constexpr A6Lay PL = A6Lay::Plain, PD = A6Lay::Pad;
    __shared__ __align__(16) T bufs[4][kA6Row];
    __shared__ double dAs[4 * 36];
    const int tid = threadIdx.x, wave = tid >> 6, pair = wave >> 1, p0 = (wave & 1) * 3;
    // items (entry * 36 + row) below 2^31: checked by the launcher
    const int items = batch * kA6R;
    const int per = (items + (int)gridDim.x - 1) / (int)gridDim.x;
    const int r0 = (int)blockIdx.x * per, r1 = min(items, r0 + per);
    if (tid < 4 * 36) dAs[tid] = 0;
    double acc[2][3][kA6N];  // [0]: dA_4 / dA_5, [1]: dA_3 / dA_6 (rows p0..p0+2)
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < kA6N; ++j) acc[g][i][j] = 0;
    T vu[kA6Per], vy[kA6Per];
    auto load = [&](int b, int r) {
        const auto u = to_global(W + 2LL * kA6E * b + kA6E);
        const auto y = a6_dy(io, b) + r * kA6C;
#pragma unroll
        for (int k = 0; k < kA6Per; ++k) {
            const int i = tid + k * 256;
            if (i < kA6C) {
                vu[k] = u[a6_chunk_off(r, i, C2)];
                vy[k] = y[i];
            }
        }
    };
    auto stage = [&]() {
#pragma unroll
        for (int k = 0; k < kA6Per; ++k) {
            const int i = tid + k * 256;
            if (i < kA6C) {
                bufs[0][i] = vu[k];
                bufs[1][i] = vy[k];
            }
        }
    };
    // this workgroup's sums for entry b into dA (cleared beforehand)
    // dAs[m * 36 + p + q * 6] holds dA_{3+m}
    const int gm0 = pair == 0 ? 1 : 2, gm1 = pair == 0 ? 0 : 3;
    auto flush = [&](int b) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < kA6N; ++j) {
                unsafeAtomicAdd(&dAs[gm0 * 36 + p0 + i + j * kA6N], acc[0][i][j]);
                unsafeAtomicAdd(&dAs[gm1 * 36 + p0 + i + j * kA6N], acc[1][i][j]);
                acc[0][i][j] = 0;
                acc[1][i][j] = 0;
            }
        __syncthreads();
        if (tid < 4 * 36) {
            a6_add_da(io, b, 2 * 36 + tid, T(dAs[tid]));
            dAs[tid] = 0;
        }
    };
    int b = r0 / kA6R, r = r0 - b * kA6R;
    if (r0 < r1) load(b, r);
    for (int it = r0; it < r1; ++it) {
        stage();
        __syncthreads();
        int nb = b, nr = r + 1;
        if (nr == kA6R) {
            nr = 0;
            ++nb;
        }
        T a[kA6N][kA6N];
        // One factor (72 SGPRs in FP64) at a time: the scheduling barriers
        // keep the next factor's scalar loads behind the current contraction.
        a6_factor<T, false>(a, io, b, 5);
        a6_contract<T, 1, PL, PD>(bufs[0], bufs[2], a);  // P5
        __builtin_amdgcn_sched_barrier(0);
        a6_factor<T, true>(a, io, b, 2);
        a6_contract<T, 216, PL, PD>(bufs[1], bufs[3], a);  // G4
        __syncthreads();
        a6_factor<T, false>(a, io, b, 4);
        a6_contract<T, 6, PD, PD>(bufs[2], bufs[0], a);  // P4
        __builtin_amdgcn_sched_barrier(0);
        a6_factor<T, true>(a, io, b, 3);
        a6_contract<T, 36, PD, PD>(bufs[3], bufs[1], a);  // G5
        __syncthreads();
        if (pair == 0) a6_gram_rows<T, 36, PD, PD>(bufs[3], bufs[0], p0, acc[0]);  // dA_4
        else a6_gram_rows<T, 6, PD, PD>(bufs[1], bufs[2], p0, acc[0]);            // dA_5
        __syncthreads();
        a6_factor<T, false>(a, io, b, 3);
        a6_contract<T, 36, PD, PL>(bufs[0], bufs[2], a);  // P3
        __builtin_amdgcn_sched_barrier(0);
        a6_factor<T, true>(a, io, b, 4);
        a6_contract<T, 6, PD, PL>(bufs[1], bufs[3], a);  // G6
        __syncthreads();
        stage();
        __syncthreads();
        if (it + 1 < r1) load(nb, nr);
        if (pair == 0) a6_gram_rows<T, 216, PL, PL>(bufs[1], bufs[2], p0, acc[1]);  // dA_3
        else a6_gram_rows<T, 1, PL, PL>(bufs[3], bufs[0], p0, acc[1]);              // dA_6
        __syncthreads();  // the next item's rows overwrite B0 and B1
        if (it + 1 == r1 || nb != b) flush(b);
        b = nb;
        r = nr;
    }
}

// Pass 2 over the items [r0, r1) of a workgroup, item = entry * 36 + chunk:
// column chunk k of X (rows r, columns 36 k .. 36 k + 35) and of V (one
// contiguous 36 x 36 block) give
//   P1 = (I (x) A2) X, U = (A1 (x) I) P1, G2 = (A1^T (x) I) V,
//   dX = (I (x) A2^T) G2, dA_1 = gram(V, P1) over mode 1,
//   dA_2 = gram(G2, X) over mode 2.
// A chunk in LDS is 36 x 36 row-major: mode 1 has stride 216, mode 2 stride
// 36, as modes 3 and 4 of a row.  Waves 0, 1 accumulate dA_1, waves 2, 3 dA_2.
// LDS: B0 = X, B1 = V, B2 = P1 (then U), B3 = G2 (then dX).
template <typename T, Layout LAY>
__global__ void __launch_bounds__(256)
kron_adj6_cols(A6IO<T, LAY> io, T *__restrict__ W, int batch)
{
    constexpr int CW = 36;  // chunk width (columns)
    __shared__ T bufs[4][kA6C];
    __shared__ double dAs[2 * 36];
    const int tid = threadIdx.x, wave = tid >> 6, pair = wave >> 1, half = wave & 1;
    const int items = batch * kA6R;
    const int per = (items + (int)gridDim.x - 1) / (int)gridDim.x;
    const int r0 = (int)blockIdx.x * per, r1 = min(items, r0 + per);
    if (tid < 2 * 36) dAs[tid] = 0;
    double acc[kA6N][kA6N];
#pragma unroll
    for (int i = 0; i < kA6N; ++i)
#pragma unroll
        for (int j = 0; j < kA6N; ++j) acc[i][j] = 0;
    T vx[kA6Per], vv[kA6Per];
    // element i of a chunk: row i / 36, column i % 36 of the chunk
    auto load = [&](int b, int k) {
        const auto x = a6_x(io, b) + k * CW;
        const auto v = to_global(W + 2LL * kA6E * b + k * kA6C);
#pragma unroll
        for (int q = 0; q < kA6Per; ++q) {
            const int i = tid + q * 256;
            if (i < kA6C) {
                const int r = i / CW;
                vx[q] = x[r * kA6C + (i - r * CW)];
                vv[q] = v[i];
            }
        }
    };
    auto flush = [&](int b) {
#pragma unroll
        for (int i = 0; i < kA6N; ++i)
#pragma unroll
            for (int j = 0; j < kA6N; ++j) {
                unsafeAtomicAdd(&dAs[pair * 36 + i + j * kA6N], acc[i][j]);
                acc[i][j] = 0;
            }
        __syncthreads();
        if (tid < 2 * 36) {
            a6_add_da(io, b, tid, T(dAs[tid]));
            dAs[tid] = 0;
        }
    };
    int b = r0 / kA6R, k = r0 - b * kA6R;
    if (r0 < r1) load(b, k);
    for (int it = r0; it < r1; ++it) {
#pragma unroll
        for (int q = 0; q < kA6Per; ++q) {
            const int i = tid + q * 256;
            if (i < kA6C) {
                bufs[0][i] = vx[q];
                bufs[1][i] = vv[q];
            }
        }
        __syncthreads();
        int nb = b, nk = k + 1;
        if (nk == kA6R) {
            nk = 0;
            ++nb;
        }
        if (it + 1 < r1) load(nb, nk);
        T a[kA6N][kA6N];
        a6_factor<T, false>(a, io, b, 1);
        a6_contract<T, 36>(bufs[0], bufs[2], a);  // P1: mode 2 of X
        __builtin_amdgcn_sched_barrier(0);
        a6_factor<T, true>(a, io, b, 0);
        a6_contract<T, 216>(bufs[1], bufs[3], a);  // G2: mode 1 of V
        __syncthreads();
        if (pair == 0) a6_gram<T, 216>(bufs[1], bufs[2], half, acc);  // dA_1
        else a6_gram<T, 36>(bufs[3], bufs[0], half, acc);             // dA_2
        __syncthreads();
        a6_factor<T, false>(a, io, b, 0);
        a6_contract<T, 216>(bufs[2], bufs[2], a);  // U: mode 1 of P1
        __builtin_amdgcn_sched_barrier(0);
        a6_factor<T, true>(a, io, b, 1);
        a6_contract<T, 36>(bufs[3], bufs[3], a);  // dX: mode 2 of G2
        __syncthreads();
        const auto u = to_global(W + 2LL * kA6E * b + kA6E + k * kA6C);
        const auto dx = a6_dx(io, b) + k * CW;
#pragma unroll
        for (int q = 0; q < kA6Per; ++q) {
            const int i = tid + q * 256;
            if (i < kA6C) {
                const int r = i / CW;
                u[i] = bufs[2][i];
                io.put_dx(dx, r * kA6C + (i - r * CW), bufs[3][i]);
            }
        }
        if (it + 1 == r1 || nb != b) flush(b);
        b = nb;
        k = nk;
    }
}

template <typename T, Layout LAY>
static kron_status launch_rows6_io(int pass, const AdjointArgs<T> &a, const A6IO<T, LAY> &io)
{
    constexpr int C2 = 36;  // column-chunk width of the scratch layout
    const int batch = a.batch;
    if ((long long)batch * kA6R > 0x7fffffffLL) return KRON_ENOTSUP;
    if (pass == 1) {
        hipLaunchKernelGGL((kron_adj6_rows_v<T, LAY>), dim3((unsigned)(batch * kA6R)), dim3(256),
                           0, a.stream, io, a.W, C2);
    } else {
        const void *kern = pass == 2
                               ? reinterpret_cast<const void *>(kron_adj6_cols<T, LAY>)
                               : reinterpret_cast<const void *>(kron_adj6_rows_grams<T, LAY>);
        const long long resident = persistent_blocks(kern, 0);
        const long long grid =
            std::max(1LL, std::min((long long)batch * kA6R, resident > 0 ? resident : 512));
        if (pass == 2)
            hipLaunchKernelGGL((kron_adj6_cols<T, LAY>), dim3((unsigned)grid), dim3(256), 0,
                               a.stream, io, a.W, batch);
        else
            hipLaunchKernelGGL((kron_adj6_rows_grams<T, LAY>), dim3((unsigned)grid), dim3(256), 0,
                               a.stream, io, a.W, C2, batch);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_last_hip_error((int)e);
        return KRON_EHIP;
    }
    return KRON_OK;
}

template <typename T>
kron_status launch_adjoint_rows6(int pass, const AdjointArgs<T> &a)
{
    if (a.layout == Layout::Batched)
        return launch_rows6_io<T, Layout::Batched>(pass, a,
                                                   A6Batched<T>{a.A, a.X, a.dY, a.dX, a.dA});
    return launch_rows6_io<T, Layout::Ptr>(pass, a, ptr_io(a));
}

template kron_status launch_adjoint_rows6<double>(int, const AdjointArgs<double> &);
template kron_status launch_adjoint_rows6<float>(int, const AdjointArgs<float> &);

}  // namespace kron
