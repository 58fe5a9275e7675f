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

// dA(p, r) += sum_{o, ii} G[o, p, ii] * P[o, r, ii] for mode q; dA in LDS
// (column-major n x n, double accumulation, LDS atomics once per thread).
template <typename T>
__device__ __forceinline__ void adj_gram(const T *__restrict__ G, const T *__restrict__ P,
                                         const int *ext, int R, int q, double *dA)
{
    int inner = 1;
    for (int r = q + 1; r < R; ++r) inner *= ext[r];
    int outer = 1;
    for (int r = 0; r < q; ++r) outer *= ext[r];
    const int n = ext[q];
    const int npair = n * n;
    const int rest = outer * inner;
    const int nthr = blockDim.x;
    for (int pair0 = 0; pair0 < npair; pair0 += nthr) {
        const int pairs = min(nthr, npair - pair0);
        const int per = nthr / pairs;  // threads per (p, r) pair
        const int t = threadIdx.x;
        if (t < pairs * per) {
            const int pair = pair0 + t / per;
            const int sub = t % per;
            const int p = pair % n, r = pair / n;
            double s = 0;
            for (int k = sub; k < rest; k += per) {
                const int o = k / inner, ii = k % inner;
                const int base = (o * n) * inner + ii;
                s = __builtin_fma(double(G[base + p * inner]), double(P[base + r * inner]), s);
            }
            unsafeAtomicAdd(&dA[pair], s);
        }
    }
}

// One pass of the adjoint over work items (entry, chunk):
//   phase 1: slab chunk  -> V = (I_S (x) A_F^T) dY
//   phase 2: column chunk -> U, dX, dA_S partial sums
//   phase 3: slab chunk  -> dA_F partial sums
// Partial dA sums of a work item are accumulated in LDS and added to dA (zeroed
// beforehand) with one atomic per element.
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
    int ext[kMaxDim + 1];
    const long long nck = (phase == 2) ? (EF + sh.C2 - 1) / sh.C2 : (ES + sh.C1 - 1) / sh.C1;
    const long long items = (long long)batch * nck;

    for (long long it = blockIdx.x; it < items; it += gridDim.x) {
        const int b = (int)(it / nck);
        const long long ck = it % nck;
        const T *Ab = A + (long long)b * d * n * n;
        for (int w = tid; w < d * n * n; w += nthr) {
            facL[w] = Ab[w];
            dAacc[w] = 0;
        }
        const T *x = X + b * E;
        const T *dy = dY + b * E;
        T *V = (nf > 0) ? W + (long long)b * 2 * E : nullptr;
        T *U = (nf > 0) ? V + E : nullptr;

        if (phase == 1) {
            const long long s0 = ck * sh.C1;
            const int C = (int)min((long long)sh.C1, ES - s0);
            const int cnt = C * (int)EF;
            for (int w = tid; w < cnt; w += nthr) B(0)[w] = dy[s0 * EF + w];
            __syncthreads();
            ext[0] = C;
            for (int t = 0; t < nf; ++t) ext[1 + t] = n;
            int cur = 0;
            for (int t = 0; t < nf; ++t) {
                contract_mode_lds<T, Acc, true, SQ>(B(cur), B(1 - cur), ext, 1 + nf, 1 + t, n,
                                                 fac(h + t), n);
                __syncthreads();
                cur = 1 - cur;
            }
            for (int w = tid; w < cnt; w += nthr) V[s0 * EF + w] = B(cur)[w];
            __syncthreads();
            continue;
        }

        if (phase == 2) {
            // buffers: P_i in B(i) (B(h-1) holds X), G ping-pong in B(h), B(h+1)
            const T *Vr = (nf > 0) ? V : dy;
            const long long c0 = ck * sh.C2;
            const int C = (int)min((long long)sh.C2, EF - c0);
            const int cnt = (int)ES * C;
            for (int w = tid; w < cnt; w += nthr) {
                const int is = w / C, c = w % C;
                B(h - 1)[w] = x[is * EF + c0 + c];
                B(h)[w] = Vr[is * EF + c0 + c];
            }
            __syncthreads();
            for (int k = 0; k < h; ++k) ext[k] = n;
            ext[h] = C;
            for (int i = h - 2; i >= 0; --i) {
                contract_mode_lds<T, Acc, false, SQ>(B(i + 1), B(i), ext, h + 1, i + 1, n,
                                                  fac(i + 1), n);
                __syncthreads();
            }
            if (nf > 0) {
                // U chunk = A_0 x_0 P_0, through B(h+1) (the G chain starts in B(h))
                contract_mode_lds<T, Acc, false, SQ>(B(0), B(h + 1), ext, h + 1, 0, n, fac(0), n);
                __syncthreads();
                for (int w = tid; w < cnt; w += nthr) {
                    const int is = w / C, c = w % C;
                    U[is * EF + c0 + c] = B(h + 1)[w];
                }
                __syncthreads();
            }
            int g = h;
            for (int i = 0; i < h; ++i) {
                adj_gram<T>(B(g), B(i), ext, h + 1, i, dAacc + i * n * n);
                const int gn = (g == h) ? h + 1 : h;
                contract_mode_lds<T, Acc, true, SQ>(B(g), B(gn), ext, h + 1, i, n, fac(i), n);
                __syncthreads();
                g = gn;
            }
            T *dx = dX + b * E;
            for (int w = tid; w < cnt; w += nthr) {
                const int is = w / C, c = w % C;
                dx[is * EF + c0 + c] = B(g)[w];
            }
        } else {
            // phase 3 -- buffers: P_t in B(t) (B(nf-1) holds U), G ping-pong in
            // B(nf), B(nf+1)
            const long long s0 = ck * sh.C1;
            const int C = (int)min((long long)sh.C1, ES - s0);
            const int cnt = C * (int)EF;
            for (int w = tid; w < cnt; w += nthr) {
                B(nf - 1)[w] = U[s0 * EF + w];
                B(nf)[w] = dy[s0 * EF + w];
            }
            __syncthreads();
            ext[0] = C;
            for (int t = 0; t < nf; ++t) ext[1 + t] = n;
            for (int t = nf - 2; t >= 0; --t) {
                contract_mode_lds<T, Acc, false, SQ>(B(t + 1), B(t), ext, 1 + nf, 2 + t, n,
                                                  fac(h + t + 1), n);
                __syncthreads();
            }
            int g = nf;
            for (int t = 0; t < nf; ++t) {
                adj_gram<T>(B(g), B(t), ext, 1 + nf, 1 + t, dAacc + (h + t) * n * n);
                if (t + 1 < nf) {
                    const int gn = (g == nf) ? nf + 1 : nf;
                    contract_mode_lds<T, Acc, true, SQ>(B(g), B(gn), ext, 1 + nf, 1 + t, n,
                                                     fac(h + t), n);
                    g = gn;
                }
                __syncthreads();
            }
        }
        __syncthreads();
        // this item's share of dA: modes [0, h) in phase 2, [h, d) in phase 3
        const int k0 = (phase == 2) ? 0 : h, k1 = (phase == 2) ? h : d;
        const auto dAb = to_global(dA + (long long)b * d * n * n);
        for (int w = k0 * n * n + tid; w < k1 * n * n; w += nthr) atomic_add(dAb + w, T(dAacc[w]));
        __syncthreads();
    }
}

template <typename T>
__global__ void __launch_bounds__(256) kron_fill_zero(T *p, long long count)
{
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < count; i += (long long)gridDim.x * 256)
        p[i] = T(0);
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
        const long long per = budget / nbuf;  // elements per buffer
        if (per < std::max(s.ES, nf > 0 ? s.EF : 1)) continue;
        s.C1 = nf > 0 ? (int)std::min<long long>(s.ES, per / s.EF) : 1;
        s.C2 = (int)std::min<long long>(s.EF, per / s.ES);
        s.C3 = s.C1;
        s.nbuf = nbuf;
        s.buf = (int)std::max<long long>((long long)s.C2 * s.ES, nf > 0 ? (long long)s.C1 * s.EF : 0);
        const long long iters = (nf > 0 ? 2 * ((s.ES + s.C1 - 1) / s.C1) : 0) +
                                (s.EF + s.C2 - 1) / s.C2;
        // favour column runs of at least 8 elements (coalescing of pass 2)
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
    const long long budget = (std::min(max_lds_bytes(), 65536) - fixed) / (long long)sizeof(T);
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
    const long long lds_cap = std::min(max_lds_bytes(), 65536);
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
    const long long cap = resident > 0 ? 8 * resident : 65536;
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
