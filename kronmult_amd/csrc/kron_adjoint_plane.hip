// kron_adjoint_plane.hip -- the d = 6, n = 4 adjoint with every entry on
// chip (4,096 elements), on the plane layout of kron_plane.hpp.
//
// With Y = K X per entry and upstream gradient dY, the forward's three plane
// passes are
//   X1 = A_a X0 (modes 2, 4), X2 = A_b X1 (modes 6, 5), Y = A_c X2 (modes 3, 1)
// and the adjoint runs the transposed chain on dY:
//   D1 = A_c^T dY, D2 = A_b^T D1, dX = A_a^T D2.
// The gradient of factor k is dA_k(p, q) = sum over the other five indices of
// G[.. i_k = p ..] P[.. j_k = q ..], where every other factor is applied on
// one side (transposed to dY, or to X).  Each pair of modes takes its
// gradients in the orientation whose planes already hold both sides:
//   modes 5, 6 (B planes): D1 and X1;
//   modes 3, 1 (C planes): dY and X2;
//   modes 2, 4 (A planes): D2 and X0;
// with the pair's other mode (the partner) contracted into one side.
// A thread's plane gives a 4 x 4 partial sum per gradient (64 FMAs); the
// workgroup's 256 partials are summed by a wave reduce-scatter (each lane
// keeps half of its values and adds its partner's half: 8 and 4 pairs by lane
// swaps across 32 and 16 lanes, 2 and 1 by DPP lane moves, then two full
// steps within quads) and one LDS add per output and wave.
//
// Per entry: X0 (A planes) and dY (C planes) come from HBM into registers,
// X1 and D1 go to two LDS buffers, are read back as B planes, X2 and D2 are
// written in place and read as C and A planes, beside dY and X0 read a second
// time (from cache); dX leaves from the A planes.
// Of the six partner contractions, two are a chain's own next step (A6^T D1
// and A4^T D2).
// In the pointer layout the entries are taken in the order of their dX
// targets (as the forward orders by Y), so a run of entries adding into one
// X gradient sums it in registers and flushes once; dA goes out per entry
// (96 atomic adds).  Nothing goes through the scratch W.
#include "kron_adjoint.hpp"
#include "kron_fast.hpp"  // load_factor, kron_plane.hpp

namespace kron {

namespace {

// v[i][j] <- sum_k f(i, k) v[k][j], f = a or its transpose
template <bool TR, typename Acc>
__device__ __forceinline__ void ap_first(Acc (&v)[4][4], const Acc (&a)[4][4])
{
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        Acc t[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            Acc s = (TR ? a[0][i] : a[i][0]) * v[0][j];
#pragma unroll
            for (int k = 1; k < 4; ++k) s = kfma(TR ? a[k][i] : a[i][k], v[k][j], s);
            t[i] = s;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i][j] = t[i];
    }
}

// v[i][j] <- sum_k f(j, k) v[i][k]
template <bool TR, typename Acc>
__device__ __forceinline__ void ap_second(Acc (&v)[4][4], const Acc (&a)[4][4])
{
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        Acc t[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            Acc s = (TR ? a[0][j] : a[j][0]) * v[i][0];
#pragma unroll
            for (int k = 1; k < 4; ++k) s = kfma(TR ? a[k][j] : a[j][k], v[i][k], s);
            t[j] = s;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) v[i][j] = t[j];
    }
}

template <typename Acc>
__device__ __forceinline__ void ap_copy(Acc (&d)[4][4], const Acc (&s)[4][4])
{
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) d[i][j] = s[i][j];
}

// Halves of a double, for the 32-bit cross-lane instructions.
__device__ __forceinline__ void ap_split(double x, unsigned &lo, unsigned &hi)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
    lo = (unsigned)u;
    hi = (unsigned)(u >> 32);
}

__device__ __forceinline__ double ap_join(unsigned lo, unsigned hi)
{
    return __builtin_bit_cast(double, (unsigned long long)hi << 32 | lo);
}

// a <- a + b after exchanging the upper 32 lanes of a with the lower 32 of b
// (v_permlane32_swap), or the odd 16-lane rows of a with the even rows of b
// (v_permlane16_swap): lanes with the bit clear end with the pair sum of a,
// lanes with it set with the pair sum of b.  No lane selects.
template <int M>
__device__ __forceinline__ void ap_swap_add(double &a, double b)
{
    static_assert(M == 32 || M == 16, "lane swaps exist across 32 and 16 lanes");
    unsigned al, ah, bl, bh;
    ap_split(a, al, ah);
    ap_split(b, bl, bh);
    if constexpr (M == 32) {
        const auto l = __builtin_amdgcn_permlane32_swap(al, bl, false, false);
        const auto h = __builtin_amdgcn_permlane32_swap(ah, bh, false, false);
        a = ap_join(l[0], h[0]) + ap_join(l[1], h[1]);
    } else {
        const auto l = __builtin_amdgcn_permlane16_swap(al, bl, false, false);
        const auto h = __builtin_amdgcn_permlane16_swap(ah, bh, false, false);
        a = ap_join(l[0], h[0]) + ap_join(l[1], h[1]);
    }
}

// x of the partner lane under DPP control CTRL (a permutation within each
// row of 16 lanes)
template <int CTRL>
__device__ __forceinline__ double ap_dpp(double x)
{
    unsigned lo, hi;
    ap_split(x, lo, hi);
    return ap_join((unsigned)__builtin_amdgcn_mov_dpp((int)lo, CTRL, 0xF, 0xF, true),
                   (unsigned)__builtin_amdgcn_mov_dpp((int)hi, CTRL, 0xF, 0xF, true));
}

// a <- a + b by lane pairs (DPP CTRL pairs every lane with one whose bit
// `up` differs, inside its group of 8 or 16): lanes with the bit clear keep
// the sum of a, the others the sum of b.
template <int CTRL>
__device__ __forceinline__ void ap_pair_add(double &a, double b, bool up)
{
    const double mine = up ? b : a, other = up ? a : b;
    a = mine + ap_dpp<CTRL>(other);
}

// Sum of v[0..15] over the wave's 64 lanes, reduce-scatter: afterwards every
// lane holds the wave's total of output o = 8 b5 + 4 b4 + 2 b3 + b2 (bk the
// lane's bit k).  Written out index by index so that every select is between
// two registers.
__device__ __forceinline__ double ap_wave_reduce16(double (&v)[16], int lane)
{
    ap_swap_add<32>(v[0], v[8]);
    ap_swap_add<32>(v[1], v[9]);
    ap_swap_add<32>(v[2], v[10]);
    ap_swap_add<32>(v[3], v[11]);
    ap_swap_add<32>(v[4], v[12]);
    ap_swap_add<32>(v[5], v[13]);
    ap_swap_add<32>(v[6], v[14]);
    ap_swap_add<32>(v[7], v[15]);
    ap_swap_add<16>(v[0], v[4]);
    ap_swap_add<16>(v[1], v[5]);
    ap_swap_add<16>(v[2], v[6]);
    ap_swap_add<16>(v[3], v[7]);
    // row_ror:8 pairs lanes L, L ^ 8; row_half_mirror pairs L, L ^ 7 inside
    // each 8 lanes (bit 2 differs), and the quad steps below then cover the
    // partner's quad
    ap_pair_add<0x128>(v[0], v[2], (lane & 8) != 0);
    ap_pair_add<0x128>(v[1], v[3], (lane & 8) != 0);
    ap_pair_add<0x141>(v[0], v[1], (lane & 4) != 0);
    const double r = v[0] + ap_dpp<0x4E>(v[0]);  // quad_perm [2,3,0,1]: L ^ 2
    return r + ap_dpp<0xB1>(r);                  // quad_perm [1,0,3,2]: L ^ 1
}

// The workgroup's gradient of factor kf (outputs p + 4 q) added into LDS:
// the lanes that hold output o after the wave reduction add it to dAs.
__device__ __forceinline__ void ap_gram_out(double (&v)[16], double *dAs, int kf, int lane)
{
    const double r = ap_wave_reduce16(v, lane);
    if ((lane & 3) == 0) {
        const int o = ((lane >> 5) & 1) * 8 + ((lane >> 4) & 1) * 4 + ((lane >> 3) & 1) * 2 +
                      ((lane >> 2) & 1);
        unsafeAtomicAdd(dAs + kf * 16 + o, r);
    }
}

// gradient over the first plane index: out(p, q) = sum_j g[p][j] x[q][j]
template <typename Acc>
__device__ __forceinline__ void ap_gram_first(const Acc (&g)[4][4], const Acc (&x)[4][4],
                                              double *dAs, int kf, int lane)
{
    double v[16];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            double s = double(g[p][0]) * double(x[q][0]);
#pragma unroll
            for (int j = 1; j < 4; ++j) s = kfma(double(g[p][j]), double(x[q][j]), s);
            v[p + 4 * q] = s;
        }
    ap_gram_out(v, dAs, kf, lane);
}

// gradient over the second plane index: out(p, q) = sum_i g[i][p] x[i][q]
template <typename Acc>
__device__ __forceinline__ void ap_gram_second(const Acc (&g)[4][4], const Acc (&x)[4][4],
                                               double *dAs, int kf, int lane)
{
    double v[16];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            double s = double(g[0][p]) * double(x[0][q]);
#pragma unroll
            for (int i = 1; i < 4; ++i) s = kfma(double(g[i][p]), double(x[i][q]), s);
            v[p + 4 * q] = s;
        }
    ap_gram_out(v, dAs, kf, lane);
}

}  // namespace

// Persistent workgroups, one entry at a time.  GROUPED (pointer layout):
// position r of the dX-target order holds entry perm[r]; a run of positions
// with one dX target sums its contribution in registers.  Otherwise entry
// rbeg + j * rstep, with dX stored (batched layout: each entry owns its dX)
// or added per element (pointer layout).
template <typename T, Layout LAY, bool GROUPED>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
kron_adjoint_plane_kernel(AdjIO<T, LAY> io, int batch, GroupOrder ord)
{
    using Acc = T;
    __shared__ __align__(16) T lx[kPlaneLds];
    __shared__ __align__(16) T ld[kPlaneLds];
    __shared__ double dAs[96];
    const int tid = threadIdx.x;
    const int l = tid & 63, w = tid >> 6;
    const int d0 = l & 3, d1 = (l >> 2) & 3, d2 = (l >> 4) & 3;
    // plane origins, as in kron_plane_kernel: A (m1 = w, m3 = d2, m5 = d1,
    // m6 = d0; first index m2, second m4), B (m1 = w, m2 = d2, m3 = d1,
    // m4 = d0; first m6, second m5), C (m2 = w, m4 = d2, m5 = d1, m6 = d0;
    // first m3, second m1)
    const int eA = w * 1024 + d2 * 64 + d1 * 4 + d0;
    const int eB = w * 1024 + d2 * 256 + d1 * 64 + d0 * 16;
    const int eC = w * 256 + d2 * 16 + d1 * 4 + d0;
    if (tid < 96) dAs[tid] = 0.0;

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
    const int lda = io.lda();
    auto fac = [&](int b, int k) { return to_const(io.fac(b, k)); };
    Acc dxs[GROUPED ? 4 : 1][4];  // the running dX sum of a run (GROUPED)
    if constexpr (GROUPED) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) dxs[i][j] = 0;
    }
    __syncthreads();

    for (int jj = 0; jj < cnt; ++jj) {
        const int rr = rbeg + jj * rstep;
        const int b = GROUPED ? to_global(ord.perm)[rr] : rr;
        Acc a[4][4], c[4][4], v[4][4];
        // phase 1: X0 (A planes) -> X1 = A2 A4 X0 -> LDS (A planes), and
        // dY (C planes) -> D1 = A3^T A1^T dY -> LDS (C planes)
        {
            Acc w[4][4];
            const auto x = io.x(b) + eA;
            const auto y = io.dy(b) + eC;
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    v[i][j] = Acc(x[256 * i + 16 * j]);
                    w[i][j] = Acc(y[64 * i + 1024 * j]);
                }
            load_factor<T, Acc, 4, false, true>(a, fac(b, 3), lda);
            ap_second<false>(v, a);
            load_factor<T, Acc, 4, false, true>(c, fac(b, 1), lda);
            ap_first<false>(v, c);
            T *const LA = lx + plane_pad(eA);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) LA[272 * i + 17 * j] = T(v[i][j]);
            load_factor<T, Acc, 4, false, true>(a, fac(b, 2), lda);
            ap_first<true>(w, a);
            load_factor<T, Acc, 4, false, true>(c, fac(b, 0), lda);
            ap_second<true>(w, c);
            T *const LC = ld + plane_pad(eC);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) LC[68 * i + 1088 * j] = T(w[i][j]);
        }
        __syncthreads();

        // phase 2 (B planes: first m6 = factor 5, second m5 = factor 4)
        {
            T *const LXB = lx + plane_pad(eB);
            T *const LDB = ld + plane_pad(eB);
            Acc g[4][4], p[4][4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    v[i][j] = Acc(LXB[i + 4 * j]);
                    g[i][j] = Acc(LDB[i + 4 * j]);
                }
            load_factor<T, Acc, 4, false, true>(a, fac(b, 5), lda);  // A6
            load_factor<T, Acc, 4, false, true>(c, fac(b, 4), lda);  // A5
            // Each gradient takes its partner mode from a chain's own first
            // step: dA6 (mode 6, the first index) D1 with A5 X1; dA5 (mode 5,
            // the second index) A6^T D1 with X1.
            ap_copy(p, v);
            ap_second<false>(p, c);  // A5 X1
            ap_gram_first(g, p, dAs, 5, l);
            ap_first<true>(g, a);    // A6^T D1
            ap_gram_second(g, v, dAs, 4, l);
            // X2 = A6 (A5 X1), D2 = A5^T (A6^T D1)
            ap_first<false>(p, a);
            ap_second<true>(g, c);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    LXB[i + 4 * j] = T(p[i][j]);
                    LDB[i + 4 * j] = T(g[i][j]);
                }
        }
        __syncthreads();

        // phase 3, C planes (first m3 = factor 2, second m1 = factor 0): X2
        // with dY contracted by the partner mode
        {
            // dY again from global memory (the cache still holds it): keeping
            // X0 and dY in registers from phase 1 would halve occupancy
            const T *const LXC = lx + plane_pad(eC);
            const auto y = io.dy(b) + eC;
            Acc y0[4][4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    y0[i][j] = Acc(y[64 * i + 1024 * j]);
                    v[i][j] = Acc(LXC[68 * i + 1088 * j]);
                }
            load_factor<T, Acc, 4, false, true>(a, fac(b, 2), lda);  // A3
            load_factor<T, Acc, 4, false, true>(c, fac(b, 0), lda);  // A1
            // dA3: A1^T dY with X2 over mode 3 (first); dA1: A3^T dY over mode 1
            Acc q[4][4];
            ap_copy(q, y0);
            ap_second<true>(q, c);
            ap_gram_first(q, v, dAs, 2, l);
            ap_first<true>(y0, a);
            ap_gram_second(y0, v, dAs, 0, l);
        }
        // A planes (first m2 = factor 1, second m4 = factor 3): D2 and X0,
        // and dX = A2^T A4^T D2
        {
            const T *const LDA = ld + plane_pad(eA);
            const auto x = io.x(b) + eA;
            Acc g[4][4], x0[4][4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    x0[i][j] = Acc(x[256 * i + 16 * j]);
                    g[i][j] = Acc(LDA[272 * i + 17 * j]);
                }
            load_factor<T, Acc, 4, false, true>(a, fac(b, 1), lda);  // A2
            load_factor<T, Acc, 4, false, true>(c, fac(b, 3), lda);  // A4
            // dA4 (mode 4, the second index): D2 with A2 X0; dA2 (mode 2, the
            // first index): A4^T D2, the first step of dX, with X0
            ap_copy(v, x0);
            ap_first<false>(v, a);  // A2 X0
            ap_gram_second(g, v, dAs, 3, l);
            ap_second<true>(g, c);  // A4^T D2
            ap_gram_first(g, x0, dAs, 1, l);
            ap_first<true>(g, a);   // dX = A2^T (A4^T D2)
            if constexpr (GROUPED) {
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) dxs[i][j] += g[i][j];
                if (jj + 1 == cnt || to_global(ord.keys)[rr + 1] != to_global(ord.keys)[rr]) {
                    const auto dx = io.dx(b) + eA;
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            atomic_add(dx + (256 * i + 16 * j), T(dxs[i][j]));
                            dxs[i][j] = 0;
                        }
                }
            } else {
                const auto dx = io.dx(b) + eA;
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) io.put_dx(dx, 256 * i + 16 * j, T(g[i][j]));
            }
        }
        __syncthreads();  // every wave's dA sums are in dAs; LDS is rewritten next
        if (tid < 96) {
            io.add_da(b, tid, T(dAs[tid]));
            dAs[tid] = 0.0;
        }
    }
}

template <typename T, Layout LAY>
static kron_status launch_plane_io(const AdjIO<T, LAY> &io, int batch, const GroupOrder *ord,
                                   hipStream_t s)
{
    const void *kern = ord ? reinterpret_cast<const void *>(
                                 kron_adjoint_plane_kernel<T, LAY, true>)
                           : reinterpret_cast<const void *>(
                                 kron_adjoint_plane_kernel<T, LAY, false>);
    const long long resident = persistent_blocks(kern, 0);
    long long grid = std::max(1LL, std::min((long long)batch, resident > 0 ? resident : 512));
    if (ord) {
        GroupOrder o = *ord;
        o.chunk = (int)((batch + grid - 1) / grid);
        grid = (batch + o.chunk - 1) / o.chunk;
        hipLaunchKernelGGL((kron_adjoint_plane_kernel<T, LAY, true>), dim3((unsigned)grid),
                           dim3(256), 0, s, io, batch, o);
    } else {
        hipLaunchKernelGGL((kron_adjoint_plane_kernel<T, LAY, false>), dim3((unsigned)grid),
                           dim3(256), 0, s, io, batch, GroupOrder{nullptr, nullptr, 0});
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_last_hip_error((int)e);
        return KRON_EHIP;
    }
    return KRON_OK;
}

template <typename T>
kron_status launch_adjoint_plane(const AdjointArgs<T> &a, const GroupOrder *ord)
{
    if (a.layout == Layout::Batched) return launch_plane_io<T>(batched_io(a), a.batch, nullptr,
                                                               a.stream);
    return launch_plane_io<T>(ptr_io(a), a.batch, ord, a.stream);
}

template kron_status launch_adjoint_plane<double>(const AdjointArgs<double> &,
                                                  const GroupOrder *);
template kron_status launch_adjoint_plane<float>(const AdjointArgs<float> &, const GroupOrder *);

}  // namespace kron
