// kron_adjoint.hip -- gradients of the batched Kronecker matvec.
//
// For Y = K X with K = A1 (x) ... (x) Ad per entry and upstream dY:
//     dX   = K^T dY = (A1^T (x) ... (x) Ad^T) dY
//     dA_i(p, q) = sum over all other indices of  G_i[.., p, ..] * P_i[.., q, ..]
// where P_i is X with the modes after i contracted by their factors (suffix)
// and G_i is dY with the modes before i contracted by the transposed factors
// (prefix); both then live in the same mixed index space.  The reference has
// no adjoint; this is a project addition built on the forward's layout
// (kronmult6_batched.hpp:51-71).
//
// Streaming in three passes per entry, with the modes split into slow S =
// [0, h) and fast F = [h, d) and the entry viewed as E_S x E_F:
//   pass 1 (contiguous slabs, fixed slow index):  V = (I_S (x) A_F^T) dY
//   pass 2 (column chunks, fixed fast indices):   the P chain of X gives
//           U = (A_S (x) I_F) X and P_i for i in S; the G chain of V gives
//           dX = (A_S^T (x) I_F) V and G_i for i in S; grams -> dA_S
//   pass 3 (slabs):  P chain of U and G chain of dY over the fast modes;
//           grams -> dA_F
// V and U go through the entry's scratch (2 * n^d elements); HBM traffic is
// 8 vectors per entry (forward two-pass: 5).  Intermediates are recomputed,
// never stored beyond V and U.  Each pass is one launch over (entry, chunk)
// work items; dA partial sums are added atomically into the cleared dA.
#include <algorithm>

#include "kron_device.hpp"

namespace kron {

struct AdjShape {
    int d, n, h;
    long long E, ES, EF;   // n^d, n^h, n^(d-h)
    int C1, C2, C3;        // chunk sizes of passes 1, 2, 3
    int buf;               // elements per LDS buffer
    int nbuf;              // LDS buffers
};

// n^k for the small uniform exponents of an entry's mode extents.
__device__ __forceinline__ int ipow(int n, int k)
{
    int p = 1;
    for (int i = 0; i < k; ++i) p *= n;
    return p;
}

// dA(p, r) += sum_{o, ii} G[o, p, ii] * P[o, r, ii] over tensors viewed as
// outer x n x inner (dA in LDS, column-major n x n, double accumulation).
// Thread t takes the pair t % npair and every (t / npair)-th (o, ii)
// position, so the lanes of a wave read a handful of broadcast addresses;
// indices advance without division.
template <typename T>
__device__ __forceinline__ void adj_gram(const T *__restrict__ G, const T *__restrict__ P,
                                         int outer, int n, int inner, double *dA)
{
    const int npair = n * n;
    const int rest = outer * inner;
    const int nthr = blockDim.x;
    for (int pair0 = 0; pair0 < npair; pair0 += nthr) {
        const int pairs = min(nthr, npair - pair0);
        const int per = nthr / pairs;  // threads sharing a pair
        const int t = threadIdx.x;
        if (t < pairs * per) {
            const int pair = pair0 + t % pairs;
            const int sub = t / pairs;
            const int p = pair % n, r = pair / n;
            const int d_o = per / inner, d_ii = per % inner;
            int o = sub / inner, ii = sub % inner;
            double s = 0;
            for (int k = sub; k < rest; k += per) {
                const int base = (o * n) * inner + ii;
                s = __builtin_fma(double(G[base + p * inner]), double(P[base + r * inner]), s);
                o += d_o;
                ii += d_ii;
                if (ii >= inner) {
                    ii -= inner;
                    ++o;
                }
            }
            unsafeAtomicAdd(&dA[pair], s);
        }
    }
}

// adj_gram for a compile-time n: thread t owns row p = t % N of dA and the
// columns part, part + PARTS, ... (part = t / N); per column it reads one G
// value and N P values into N independent accumulators (N-way ILP, about
// (N+1)/N LDS reads per FMA), the column index advancing without branches.
// The row is then added into the LDS accumulator (about 256/N lanes per
// address; sharing rows between fewer threads serialises these adds more).
template <typename T, int N>
__device__ __forceinline__ void adj_gram_rows(const T *__restrict__ G, const T *__restrict__ P,
                                              int outer, int inner, double *dA)
{
    constexpr int PARTS = 256 / N;
    const int t = threadIdx.x;
    if (t >= PARTS * N) return;
    const int p = t % N, part = t / N;
    const int cols = outer * inner;
    double acc[N];
#pragma unroll
    for (int r = 0; r < N; ++r) acc[r] = 0;
    int o = part / inner, ii = part % inner;
    const int d_o = PARTS / inner, d_ii = PARTS % inner;
#pragma unroll 2
    for (int c = part; c < cols; c += PARTS) {
        const int base = (o * N) * inner + ii;
        const double g = double(G[base + p * inner]);
#pragma unroll
        for (int r = 0; r < N; ++r) acc[r] = __builtin_fma(g, double(P[base + r * inner]), acc[r]);
        ii += d_ii;
        const int wrap = ii >= inner ? 1 : 0;
        ii -= wrap * inner;
        o += d_o + wrap;
    }
#pragma unroll
    for (int r = 0; r < N; ++r) unsafeAtomicAdd(&dA[p + r * N], acc[r]);
}

// dA(p, r) += G P^T over outer x n x inner tensors: the row form for a
// compile-time n, else the pair form.
template <typename T, int SQ>
__device__ __forceinline__ void adj_gram_any(const T *__restrict__ G, const T *__restrict__ P,
                                             int outer, int n, int inner, double *dA)
{
    if constexpr (SQ > 0)
        adj_gram_rows<T, SQ>(G, P, outer, inner, dA);
    else
        adj_gram<T>(G, P, outer, n, inner, dA);
}

// Elements of one work item's input held per thread per input stream while
// the previous item is worked on (an LDS buffer is at most kAdjPer*256).
constexpr int kAdjPer = 8;

// One pass of the adjoint over work items (entry, chunk):
//   phase 1: slab chunk   -> V = (I_S (x) A_F^T) dY
//   phase 2: column chunk -> U, dX, dA_S partial sums
//   phase 3: slab chunk   -> dA_F partial sums
// Each workgroup walks a contiguous range of items with the next item's
// inputs prefetched into registers; dA partial sums stay in LDS while the
// entry does not change and are then added to dA (cleared beforehand).
template <typename T, typename Acc, int SQ>
__global__ void __launch_bounds__(256)
kron_adjoint_kernel(const T *__restrict__ A, const T *__restrict__ X, const T *__restrict__ dY,
                    T *__restrict__ dX, T *__restrict__ dA, T *__restrict__ W, AdjShape sh,
                    int batch, int phase)
{
    extern __shared__ __align__(16) unsigned char kron_smem[];
    const int d = sh.d, n = sh.n, h = sh.h;
    double *dAacc = reinterpret_cast<double *>(kron_smem);  // d*n*n
    T *facL = reinterpret_cast<T *>(kron_smem + ((size_t)d * n * n * sizeof(double) + 15) / 16 * 16);
    T *bufs = facL + ((size_t)d * n * n + 1) / 2 * 2;  // after the staged factors
    auto B = [&](int i) { return bufs + (size_t)i * sh.buf; };
    auto fac = [&](int k) { return facL + k * n * n; };
    const int tid = threadIdx.x, nthr = blockDim.x;
    const long long E = sh.E, ES = sh.ES, EF = sh.EF;
    const int nf = d - h;  // number of fast modes
    const int ddn = d * n * n;
    const long long nck = (phase == 2) ? (EF + sh.C2 - 1) / sh.C2 : (ES + sh.C1 - 1) / sh.C1;
    const long long items = (long long)batch * nck;
    const long long per = (items + gridDim.x - 1) / gridDim.x;
    const long long r0 = blockIdx.x * per, r1 = min(items, r0 + per);

    // item r -> entry, chunk start, chunk size, element count
    auto decode = [&](long long r, int &b, long long &s0, int &C, int &cnt)
                      __attribute__((always_inline)) {
        b = (int)(r / nck);
        const long long ck = r % nck;
        if (phase == 2) {
            s0 = ck * sh.C2;
            C = (int)min((long long)sh.C2, EF - s0);
            cnt = (int)ES * C;
        } else {
            s0 = ck * sh.C1;
            C = (int)min((long long)sh.C1, ES - s0);
            cnt = C * (int)EF;
        }
    };
    T v0[kAdjPer], v1[kAdjPer];
    auto load = [&](long long r) __attribute__((always_inline)) {
        int b, C, cnt;
        long long s0;
        decode(r, b, s0, C, cnt);
        // phase 1: dY slab; phase 2: X and V (dY when every mode is slow)
        // columns; phase 3: U and dY slabs
        const T *dy = dY + b * E;
        const T *V = W + (long long)b * 2 * E;
        const T *p0 = (phase == 1) ? dy : (phase == 2) ? X + b * E : V + E;
        const T *p1 = (phase == 2 && nf > 0) ? V : dy;
        if (phase != 2) {
            p0 += s0 * EF;
            p1 += s0 * EF;
        }
        const auto g0 = to_global(p0), g1 = to_global(p1);
#pragma unroll
        for (int k = 0; k < kAdjPer; ++k) {
            const int w = tid + k * 256;
            if (w < cnt) {
                const long long g = (phase == 2) ? (long long)(w / C) * EF + s0 + w % C : w;
                v0[k] = g0[g];
                if (phase != 1) v1[k] = g1[g];
            }
        }
    };
    auto flush = [&](int b) __attribute__((always_inline)) {
        // this entry's share of dA: modes [0, h) in phase 2, [h, d) in phase 3
        const int k0 = (phase == 2) ? 0 : h, k1 = (phase == 2) ? h : d;
        const auto dAb = to_global(dA + (long long)b * ddn);
        for (int w = k0 * n * n + tid; w < k1 * n * n; w += nthr) atomic_add(dAb + w, T(dAacc[w]));
    };

    for (int w = tid; w < ddn; w += nthr) dAacc[w] = 0;
    if (r0 < r1) load(r0);
    int cur_b = -1;
    for (long long r = r0; r < r1; ++r) {
        int b, C, cnt;
        long long s0;
        decode(r, b, s0, C, cnt);
        if (b != cur_b) {
            if (cur_b >= 0 && phase != 1) flush(cur_b);
            for (int w = tid; w < ddn; w += nthr) {
                dAacc[w] = 0;
                facL[w] = A[(long long)b * ddn + w];
            }
            cur_b = b;
        }
        // the prefetched inputs into LDS
        const int i0 = (phase == 1) ? 0 : (phase == 2) ? h - 1 : nf - 1;
#pragma unroll
        for (int k = 0; k < kAdjPer; ++k) {
            const int w = tid + k * 256;
            if (w < cnt) {
                B(i0)[w] = v0[k];
                if (phase != 1) B(i0 + 1)[w] = v1[k];
            }
        }
        __syncthreads();
        if (r + 1 < r1) load(r + 1);

        if (phase == 1) {
            // the slab chunk is C x n^nf; fast mode t sees outer C n^t, inner n^(nf-1-t)
            int cur = 0;
            for (int t = 0; t < nf; ++t) {
                contract_mode_lds<T, Acc, true, SQ>(B(cur), B(1 - cur), C * ipow(n, t), n,
                                                    ipow(n, nf - 1 - t), n, fac(h + t), n);
                __syncthreads();
                cur = 1 - cur;
            }
            T *V = W + (long long)b * 2 * E;
            for (int w = tid; w < cnt; w += nthr) V[s0 * EF + w] = B(cur)[w];
        } else if (phase == 2) {
            // P_i in B(i) (B(h-1) holds X), G ping-pong in B(h), B(h+1); the
            // column chunk is n^h x C, slow mode i sees outer n^i, inner n^(h-1-i) C
            for (int i = h - 2; i >= 0; --i) {
                contract_mode_lds<T, Acc, false, SQ>(B(i + 1), B(i), ipow(n, i + 1), n,
                                                     ipow(n, h - 2 - i) * C, n, fac(i + 1), n);
                __syncthreads();
            }
            if (nf > 0) {
                // U chunk = A_0 x_0 P_0, through B(h+1) (the G chain starts in B(h))
                contract_mode_lds<T, Acc, false, SQ>(B(0), B(h + 1), 1, n, ipow(n, h - 1) * C, n,
                                                     fac(0), n);
                __syncthreads();
                T *U = W + (long long)b * 2 * E + E;
                for (int w = tid; w < cnt; w += nthr) {
                    const int is = w / C, c = w % C;
                    U[(long long)is * EF + s0 + c] = B(h + 1)[w];
                }
                __syncthreads();
            }
            int g = h;
            for (int i = 0; i < h; ++i) {
                const int outer = ipow(n, i), inner = ipow(n, h - 1 - i) * C;
                adj_gram_any<T, SQ>(B(g), B(i), outer, n, inner, dAacc + i * n * n);
                const int gn = (g == h) ? h + 1 : h;
                contract_mode_lds<T, Acc, true, SQ>(B(g), B(gn), outer, n, inner, n, fac(i), n);
                __syncthreads();
                g = gn;
            }
            T *dx = dX + b * E;
            for (int w = tid; w < cnt; w += nthr) {
                const int is = w / C, c = w % C;
                dx[(long long)is * EF + s0 + c] = B(g)[w];
            }
        } else {
            // P_t in B(t) (B(nf-1) holds U), G ping-pong in B(nf), B(nf+1); the
            // slab chunk is C x n^nf, fast mode t sees outer C n^t, inner n^(nf-1-t)
            for (int t = nf - 2; t >= 0; --t) {
                contract_mode_lds<T, Acc, false, SQ>(B(t + 1), B(t), C * ipow(n, t + 1), n,
                                                     ipow(n, nf - 2 - t), n, fac(h + t + 1), n);
                __syncthreads();
            }
            int g = nf;
            for (int t = 0; t < nf; ++t) {
                const int outer = C * ipow(n, t), inner = ipow(n, nf - 1 - t);
                adj_gram_any<T, SQ>(B(g), B(t), outer, n, inner, dAacc + (h + t) * n * n);
                if (t + 1 < nf) {
                    const int gn = (g == nf) ? nf + 1 : nf;
                    contract_mode_lds<T, Acc, true, SQ>(B(g), B(gn), outer, n, inner, n,
                                                        fac(h + t), n);
                    g = gn;
                }
                __syncthreads();
            }
        }
        __syncthreads();
    }
    if (cur_b >= 0 && phase != 1) flush(cur_b);
}

template <typename T>
__global__ void __launch_bounds__(256) kron_fill_zero(T *p, long long count)
{
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < count; i += (long long)gridDim.x * 256)
        p[i] = T(0);
}

// LDS per workgroup: half of what a workgroup may use (two workgroups per
// CU, as the kernel's registers allow), at most 80 KB.
static long long adjoint_lds_cap()
{
    return std::min<long long>(max_lds_bytes() / 2 >= 32768 ? max_lds_bytes() / 2 : 65536,
                               81920);
}

// Chunking for an LDS budget of `budget` elements (after the dA accumulator).
static bool plan_adjoint(AdjShape &sh, long long budget)
{
    const int d = sh.d, n = sh.n;
    bool found = false;
    long long best = 0;
    AdjShape pick = sh;
    for (int h = 1; h <= d; ++h) {
        AdjShape s = sh;
        s.h = h;
        s.ES = 1;
        for (int k = 0; k < h; ++k) s.ES *= n;
        s.EF = s.E / s.ES;
        const int nf = d - h;
        const int nb2 = h + 2, nb3 = nf + 2;
        const int nbuf = std::max(nb2, nf > 0 ? std::max(2, nb3) : 0);
        const long long per = std::min<long long>(budget / nbuf, kAdjPer * 256);
        if (per < std::max(s.ES, nf > 0 ? s.EF : 1)) continue;
        s.C1 = nf > 0 ? (int)std::min<long long>(s.ES, per / s.EF) : 1;
        s.C2 = (int)std::min<long long>(s.EF, per / s.ES);
        s.C3 = s.C1;
        s.nbuf = nbuf;
        s.buf = (int)std::max<long long>((long long)s.C2 * s.ES, nf > 0 ? (long long)s.C1 * s.EF : 0);
        const long long iters = (nf > 0 ? 2 * ((s.ES + s.C1 - 1) / s.C1) : 0) +
                                (s.EF + s.C2 - 1) / s.C2;
        // favour column runs of at least 8 elements (pass 2 reads and writes
        // rows of C2 consecutive elements: shorter rows waste HBM lines)
        const long long cost = iters * (s.C2 < 8 && s.C2 < s.EF ? 4 : 1);
        if (!found || cost < best) {
            found = true;
            best = cost;
            pick = s;
        }
    }
    if (found) sh = pick;
    return found;
}

template <typename T>
size_t adjoint_scratch_elems(int d, int n)
{
    long long E = 1;
    for (int k = 0; k < d; ++k) E *= n;
    AdjShape sh{d, n, 0, E, 0, 0, 0, 0, 0, 0, 0};
    const long long fixed = ((long long)d * n * n * 8 + 15) / 16 * 16 +
                            ((long long)d * n * n + 1) / 2 * 2 * (long long)sizeof(T);
    const long long budget = (adjoint_lds_cap() - fixed) / (long long)sizeof(T);
    if (budget <= 0 || !plan_adjoint(sh, budget)) return (size_t)-1;
    return (sh.h < d) ? (size_t)(2 * E) : 0;
}

template <typename T>
kron_status launch_adjoint(int d, int n, const T *A, const T *X, const T *dY, T *dX, T *dA, T *W,
                           int batch, hipStream_t stream)
{
    if (batch == 0) return KRON_OK;
    long long E = 1;
    for (int k = 0; k < d; ++k) E *= n;
    AdjShape sh{d, n, 0, E, 0, 0, 0, 0, 0, 0, 0};
    const long long fixed = ((long long)d * n * n * 8 + 15) / 16 * 16 +
                            ((long long)d * n * n + 1) / 2 * 2 * (long long)sizeof(T);
    const long long lds_cap = adjoint_lds_cap();
    const long long budget = (lds_cap - fixed) / (long long)sizeof(T);
    if (budget <= 0 || !plan_adjoint(sh, budget)) return KRON_ENOTSUP;
    if (sh.h < d && !W) return KRON_EWORKSPACE;
    const size_t lds = (size_t)fixed + (size_t)sh.nbuf * sh.buf * sizeof(T);
    const void *kern;
    switch (n) {
    case 2: kern = reinterpret_cast<const void *>(kron_adjoint_kernel<T, T, 2>); break;
    case 3: kern = reinterpret_cast<const void *>(kron_adjoint_kernel<T, T, 3>); break;
    case 4: kern = reinterpret_cast<const void *>(kron_adjoint_kernel<T, T, 4>); break;
    case 5: kern = reinterpret_cast<const void *>(kron_adjoint_kernel<T, T, 5>); break;
    case 6: kern = reinterpret_cast<const void *>(kron_adjoint_kernel<T, T, 6>); break;
    case 7: kern = reinterpret_cast<const void *>(kron_adjoint_kernel<T, T, 7>); break;
    case 8: kern = reinterpret_cast<const void *>(kron_adjoint_kernel<T, T, 8>); break;
    default: kern = reinterpret_cast<const void *>(kron_adjoint_kernel<T, T, 0>); break;
    }
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds);
        if (e != hipSuccess) { set_last_hip_error((int)e); return KRON_EHIP; }
    }
    const long long resident = persistent_blocks(kern, lds);
    const long long cap = resident > 0 ? 2 * resident : 4096;
    // dA is accumulated atomically by the work items: clear it first (a fill
    // kernel, so the sequence stays valid under stream capture)
    const long long nA = (long long)batch * d * n * n;
    hipLaunchKernelGGL(kron_fill_zero<T>, dim3((unsigned)std::min<long long>((nA + 255) / 256, 4096)),
                       dim3(256), 0, stream, dA, nA);
    const int nf = d - sh.h;
    for (int phase = 1; phase <= 3; ++phase) {
        if (phase != 2 && nf == 0) continue;
        const long long nck = (phase == 2) ? (sh.EF + sh.C2 - 1) / sh.C2
                                           : (sh.ES + sh.C1 - 1) / sh.C1;
        const long long grid = std::max(1LL, std::min((long long)batch * nck, cap));
        void *args[] = {(void *)&A, (void *)&X, (void *)&dY, (void *)&dX, (void *)&dA, (void *)&W,
                        (void *)&sh, (void *)&batch, (void *)&phase};
        hipError_t le = hipLaunchKernel(kern, dim3((unsigned)grid), dim3(256), args, lds, stream);
        if (le != hipSuccess) { set_last_hip_error((int)le); return KRON_EHIP; }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { set_last_hip_error((int)e); return KRON_EHIP; }
    return KRON_OK;
}

template kron_status launch_adjoint<double>(int, int, const double *, const double *,
                                            const double *, double *, double *, double *, int,
                                            hipStream_t);
template kron_status launch_adjoint<float>(int, int, const float *, const float *, const float *,
                                           float *, float *, float *, int, hipStream_t);
template size_t adjoint_scratch_elems<double>(int, int);
template size_t adjoint_scratch_elems<float>(int, int);

}  // namespace kron
