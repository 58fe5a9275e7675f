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
// [0, h) and fast F = [h, d) and the entry viewed as E_S x E_F (rows x
// columns, row-major):
//   pass 1 (slab chunks: rows [s0, s0+C1)):       V = (I_S (x) A_F^T) dY
//   pass 2 (column chunks: columns [s0, s0+C2)):  the P chain of X gives
//           U = (A_S (x) I_F) X and P_i for i in S; the G chain of V gives
//           dX = (A_S^T (x) I_F) V and G_i for i in S; grams -> dA_S
//   pass 3 (slab chunks):  P chain of U and G chain of dY over the fast
//           modes; grams -> dA_F
// V and U live in the entry's scratch (2 * n^d elements) in column-chunk-major
// order: column chunk c (ES rows x its width, row-major) is contiguous, so
// pass 2 reads V and writes U as contiguous runs and passes 1 and 3 move one
// run per chunk and slab.  HBM traffic is 8 vectors per entry (forward
// two-pass: 5).  Each pass is one launch over (entry, chunk) work items; a
// thread's gram rows stay in registers while its workgroup works on one
// entry and are then added into dA (cleared beforehand).
#include <algorithm>
#include <cmath>

#include "kron_adjoint.hpp"

namespace kron {

struct AdjShape {
    int d, n, h;
    long long E, ES, EF;   // n^d, n^h, n^(d-h)
    int C1, C2, C3;        // pass-1 slab rows, pass-2 column-chunk width, pass-3 slab rows
    int nck2;              // column chunks, ceil(EF / C2)
};

// n^k for the small uniform exponents of an entry's mode extents.
__device__ __forceinline__ int ipow(int n, int k)
{
    int p = 1;
    for (int i = 0; i < k; ++i) p *= n;
    return p;
}

// dA(p, r) += sum_{o, ii} G[o, p, ii] * P[o, r, ii] over tensors viewed as
// outer x n x inner (dA in LDS, column-major n x n, double accumulation), for
// a runtime n.  Thread t takes the pair t % npair and every (t / npair)-th
// (o, ii) position; indices advance without division.
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
                s = kfma(double(G[base + p * inner]), double(P[base + r * inner]), s);
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

// Row p = t % N of the gram G P^T over the columns (o, ii) of two
// outer x N x inner chunks, added into the thread's accumulators: thread t
// takes the columns part, part + PARTS, ... (part = t / N) and reads one G
// value and N P values per column (N independent FMAs).
template <typename T, int N>
__device__ __forceinline__ void adj_gram_row(const T *__restrict__ G, const T *__restrict__ P,
                                             int outer, int inner, double (&acc)[N])
{
    constexpr int PARTS = 256 / N;
    const int t = threadIdx.x;
    if (t >= PARTS * N) return;
    const int p = t % N, part = t / N;
    const int cols = outer * inner;
    const FastDiv di = fast_div(inner);
    const T *Gp = G + p * inner;
#pragma unroll 2
    for (int c = part; c < cols; c += PARTS) {
        int ii;
        const int o = split_fiber(c, di, ii);
        const int base = fiber_base(o, ii, N * inner);
        const T *Pb = P + base;
        const double g = double(Gp[base]);
#pragma unroll
        for (int r = 0; r < N; ++r) acc[r] = kfma(g, double(Pb[r * inner]), acc[r]);
    }
}

// One mode contraction of a chunk held in LDS as outer x N x inner:
// dst(o, i, ii) = sum_j F(i, j) src(o, j, ii), F = A_k or A_k^T (column-major,
// read through Al: the factor in global memory through the constant address
// space -- scalar loads into SGPRs -- or its LDS copy).  Thread t contracts
// the fibers t, t + 256, ...
template <typename T, int N, bool TRANS, typename FacP>
__device__ __forceinline__ void adj_contract(const T *__restrict__ src, T *__restrict__ dst,
                                             int outer, int inner, FacP Al, int lda)
{
    const int fibers = outer * inner;
    if ((int)threadIdx.x >= fibers) return;
    T a[N][N];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) a[i][j] = TRANS ? Al[j + i * lda] : Al[i + j * lda];
    const FastDiv di = fast_div(inner);
    for (int f = threadIdx.x; f < fibers; f += 256) {
        int ii;
        const int o = split_fiber(f, di, ii);
        const int base = fiber_base(o, ii, N * inner);
        const T *s = src + base;
        T *dd = dst + base;
        T in[N];
#pragma unroll
        for (int j = 0; j < N; ++j) in[j] = s[j * inner];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            T acc = a[i][0] * in[0];
#pragma unroll
            for (int j = 1; j < N; ++j) acc = kfma(a[i][j], in[j], acc);
            dd[i * inner] = acc;
        }
    }
}

// Elements of one work item's input held per thread per input stream while
// the previous item is worked on (a chunk is at most kAdjPer * 256 elements).
constexpr int kAdjPer = 8;
constexpr int kAdjChunk = kAdjPer * 256;

// Position of element q (0 <= q < rows * EF, in scratch order) of the slab
// rows [s0, s0 + rows) of an entry's column-chunk-major scratch: returns the
// scratch offset; w is the element's index in the row-major slab (rows x EF).
struct SlabMap {
    FastDiv blk, full, last;  // rows * C2, C2, width of the last chunk
    int EF, C2, nck2, s0, ES;
};
// Offsets within an entry are below n^d <= 2^22 (the planner's chunk limits),
// so 24-bit multiplies (mul24) suffice.
__device__ __forceinline__ int slab_map(const SlabMap &m, int q, int &w)
{
    const int c = quot(q, m.blk);
    const int rem = q - (int)mul24(c, m.blk.d);
    const bool last = c == m.nck2 - 1;
    const int Cc = last ? m.last.d : m.C2;
    const int r = quot(rem, last ? m.last : m.full);
    const int cc0 = (int)mul24(c, m.C2);
    w = (int)mul24(r, m.EF) + cc0 + (rem - (int)mul24(r, Cc));
    return (int)mul24(m.ES, cc0) + (int)mul24(m.s0, Cc) + rem;
}

// One pass of the adjoint (PHASE 1, 2 or 3) over work items (entry, chunk).
// Each workgroup walks a contiguous range of items: it issues the next item's
// loads, works on the current item in LDS, stages the next item's inputs into
// LDS and only then issues the current item's stores, so no wait for loaded
// data also waits for stores of the same item.  Input buffers are never the
// buffers results are stored from.  N > 0 is the compile-time factor size
// (thread gram rows in registers for N <= 6); N == 0 handles any n.
template <typename T, int N, int PHASE, int NFC, Layout LAY>
__global__ void __launch_bounds__(256)
kron_adjoint_kernel(AdjIO<T, LAY> io, T *__restrict__ W, AdjShape sh, int batch)
{
    extern __shared__ __align__(16) unsigned char kron_smem[];
    constexpr bool KEEP = N > 0 && N <= 6;  // gram rows kept across an entry's chunks
    constexpr int GA = KEEP ? (PHASE == 2 ? kMaxDim : kMaxDim - 1) : 1;
    constexpr int NR = N > 0 ? N : 1;
    // NFC > 0: the number of fast modes (passes 1 and 3) as a compile-time
    // constant, so every stride n^k of those passes is one
    const int d = sh.d, n = N > 0 ? N : sh.n, h = sh.h, nf = NFC > 0 ? NFC : d - h;
    const int nn = n * n, ddn = d * nn;
    double *dAacc = reinterpret_cast<double *>(kron_smem);  // d*n*n
    T *facL = reinterpret_cast<T *>(kron_smem + ((size_t)ddn * sizeof(double) + 15) / 16 * 16);
    T *bufs = facL + (ddn + 1) / 2 * 2;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const long long E = sh.E, ES = sh.ES, EF = sh.EF;
    const int C = PHASE == 1 ? sh.C1 : PHASE == 2 ? sh.C2 : sh.C3;
    const int bufsz = PHASE == 2 ? (int)ES * C : C * (int)EF;
    auto B = [&](int i) __attribute__((always_inline)) { return bufs + (size_t)i * bufsz; };
    auto fac = [&](int k) __attribute__((always_inline)) { return facL + k * nn; };
    const int nck = PHASE == 2 ? sh.nck2 : (int)((ES + C - 1) / C);
    const long long items = (long long)batch * nck;
    const long long per = (items + gridDim.x - 1) / gridDim.x;
    const long long r0 = blockIdx.x * per, r1 = min(items, r0 + per);
    // modes whose dA this pass accumulates
    const int k0 = PHASE == 3 ? h : 0, ng = PHASE == 3 ? nf : h;

    // chunk geometry: rows (passes 1, 3) or columns (pass 2) [s0, s0 + Cx)
    auto geom = [&](int ck, int &s0, int &Cx, int &cnt) __attribute__((always_inline)) {
        s0 = ck * C;
        if (PHASE == 2) {
            Cx = (int)min((long long)C, EF - s0);
            cnt = (int)ES * Cx;
        } else {
            Cx = (int)min((long long)C, ES - s0);
            cnt = Cx * (int)EF;
        }
    };
    auto slab = [&](int s0, int rows) __attribute__((always_inline)) {
        SlabMap m;
        m.blk = fast_div(rows * sh.C2);
        m.full = fast_div(sh.C2);
        m.last = fast_div((int)(EF - (long long)(sh.nck2 - 1) * sh.C2));
        m.EF = (int)EF;
        m.C2 = sh.C2;
        m.nck2 = sh.nck2;
        m.s0 = s0;
        m.ES = (int)ES;
        return m;
    };

    T v0[kAdjPer], v1[kAdjPer];
    // loads of item (b, ck) into v0 / v1:
    //   pass 1: dY slab;  pass 2: X columns and V chunk (dY columns when every
    //   mode is slow);  pass 3: U slab (scratch order) and dY slab
    auto load = [&](int b, int ck) __attribute__((always_inline)) {
        int s0, Cx, cnt;
        geom(ck, s0, Cx, cnt);
        const long long eb = (long long)b * E;
        const auto gdy = io.dy(b);
        if constexpr (PHASE == 1) {
            const int sb = (int)mul24(s0, (int)EF);
#pragma unroll
            for (int k = 0; k < kAdjPer; ++k) {
                const int q = tid + k * 256;
                if (q < cnt) v0[k] = gdy[sb + q];
            }
        } else if constexpr (PHASE == 2) {
            const auto gx = io.x(b);
            const auto gv = to_global(W + 2 * eb + ES * s0);
            const FastDiv dc = fast_div(Cx);
#pragma unroll
            for (int k = 0; k < kAdjPer; ++k) {
                const int q = tid + k * 256;
                if (q < cnt) {
                    const int is = quot(q, dc);
                    const int g = (int)mul24(is, (int)EF) + s0 + (q - (int)mul24(is, Cx));
                    v0[k] = gx[g];
                    v1[k] = nf > 0 ? gv[q] : gdy[g];
                }
            }
        } else {
            const auto gu = to_global(W + 2 * eb + E);
            const SlabMap m = slab(s0, Cx);
            const int sb = (int)mul24(s0, (int)EF);
#pragma unroll
            for (int k = 0; k < kAdjPer; ++k) {
                const int q = tid + k * 256;
                if (q < cnt) {
                    int w;
                    v0[k] = gu[slab_map(m, q, w)];
                    v1[k] = gdy[sb + q];
                }
            }
        }
    };
    // the loaded item into its input buffers: pass 1 B(0); pass 2 B(h-1) (X),
    // B(h) (V); pass 3 B(nf-1) (U), B(nf) (dY)
    auto stage = [&](int ck) __attribute__((always_inline)) {
        int s0, Cx, cnt;
        geom(ck, s0, Cx, cnt);
        if constexpr (PHASE == 1) {
#pragma unroll
            for (int k = 0; k < kAdjPer; ++k) {
                const int q = tid + k * 256;
                if (q < cnt) B(0)[q] = v0[k];
            }
        } else if constexpr (PHASE == 2) {
            T *bx = B(h - 1), *bv = B(h);
#pragma unroll
            for (int k = 0; k < kAdjPer; ++k) {
                const int q = tid + k * 256;
                if (q < cnt) {
                    bx[q] = v0[k];
                    bv[q] = v1[k];
                }
            }
        } else {
            T *bu = B(nf - 1), *by = B(nf);
            const SlabMap m = slab(s0, Cx);
#pragma unroll
            for (int k = 0; k < kAdjPer; ++k) {
                const int q = tid + k * 256;
                if (q < cnt) {
                    int w;
                    slab_map(m, q, w);
                    bu[w] = v0[k];
                    by[q] = v1[k];
                }
            }
        }
    };

    double acc[GA][NR];
#pragma unroll
    for (int g = 0; g < GA; ++g)
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[g][r] = 0;
    // gram g of this pass (mode k0 + g): G P^T of outer x n x inner chunks
    auto gram = [&](const T *G, const T *P, int outer, int inner, auto gc)
                    __attribute__((always_inline)) {
        constexpr int g = decltype(gc)::value;
        if constexpr (N == 0) {
            adj_gram<T>(G, P, outer, n, inner, dAacc + (k0 + g) * nn);
        } else if constexpr (KEEP) {
            adj_gram_row<T, N>(G, P, outer, inner, acc[g]);
        } else {
            adj_gram_row<T, N>(G, P, outer, inner, acc[0]);
            if (tid < (256 / N) * N) {
                const int p = tid % N;
#pragma unroll
                for (int r = 0; r < N; ++r) {
                    unsafeAtomicAdd(&dAacc[(k0 + g) * nn + p + r * N], acc[0][r]);
                    acc[0][r] = 0;
                }
            }
        }
    };
    int cur_b = -1;  // the entry whose factors are staged
    auto contract = [&](const T *src, T *dst, int outer, int inner, int k, auto tr)
                        __attribute__((always_inline)) {
        constexpr bool TR = decltype(tr)::value;
        if constexpr (N == 0)
            contract_fibers<T, T, TR, 8>(src, dst, outer, n, inner, n, fac(k), n);
        else if constexpr (N * N * sizeof(T) <= 288)
            // the factor through scalar loads (SGPRs): no LDS reads for it
            adj_contract<T, N, TR>(src, dst, outer, inner, to_const(io.fac(cur_b, k)),
                                   io.lda());
        else
            adj_contract<T, N, TR>(src, dst, outer, inner, fac(k), N);
    };
    // this entry's share of dA: registers -> LDS -> one atomic per element
    auto flush = [&](int b) __attribute__((always_inline)) {
        if constexpr (KEEP) {
            if (tid < (256 / NR) * NR) {
                const int p = tid % NR;
#pragma unroll
                for (int g = 0; g < GA; ++g)
                    if (g < ng) {
#pragma unroll
                        for (int r = 0; r < NR; ++r) {
                            unsafeAtomicAdd(&dAacc[(k0 + g) * nn + p + r * NR], acc[g][r]);
                            acc[g][r] = 0;
                        }
                    }
            }
        }
        __syncthreads();
        for (int w = k0 * nn + tid; w < (k0 + ng) * nn; w += nthr) {
            io.add_da(b, w, T(dAacc[w]));
            dAacc[w] = 0;
        }
    };
    using I0 = std::integral_constant<int, 0>;
    using TRUE_ = std::integral_constant<bool, true>;
    using FALSE_ = std::integral_constant<bool, false>;

    for (int w = tid; w < ddn; w += nthr) dAacc[w] = 0;
    int b = (int)(r0 / nck), ck = (int)(r0 % nck);
    if (r0 < r1) load(b, ck);
    bool staged = false;
    for (long long r = r0; r < r1; ++r) {
        if (b != cur_b) {
            if (cur_b >= 0 && PHASE != 1) flush(cur_b);
            for (int w = tid; w < ddn; w += nthr) facL[w] = io.fac_elem(b, w);
            cur_b = b;
        }
        if (!staged) stage(ck);
        __syncthreads();
        int nb = b, nck_ = ck + 1;
        if (nck_ == nck) {
            nck_ = 0;
            ++nb;
        }
        const bool more = r + 1 < r1;
        if (more) load(nb, nck_);

        int s0, Cx, cnt;
        geom(ck, s0, Cx, cnt);
        const T *res = nullptr;  // the buffer the item's result is stored from
        if constexpr (PHASE == 1) {
            // slab chunk Cx x n^nf; fast mode t sees outer Cx n^t, inner n^(nf-1-t)
            int src = 0;
            for (int t = 0; t < nf; ++t) {
                const int dst = 1 + (t & 1);
                contract(B(src), B(dst), Cx * ipow(n, t), ipow(n, nf - 1 - t), h + t, TRUE_{});
                __syncthreads();
                src = dst;
            }
            res = B(src);
        } else if constexpr (PHASE == 2) {
            // P_i in B(i) (B(h-1) holds X); column chunk n^h x Cx, slow mode i
            // sees outer n^i, inner n^(h-1-i) Cx
            for (int i = h - 2; i >= 0; --i) {
                contract(B(i + 1), B(i), ipow(n, i + 1), ipow(n, h - 2 - i) * Cx, i + 1, FALSE_{});
                __syncthreads();
            }
            // G_0 = V in B(h); G_1 -> B(h+2), G_(i+1) -> B(i-1) (P_(i-1) is
            // dead by then); U = A_0 P_0 -> B(h+1)
            int gb = h;
            auto step = [&](auto ic) __attribute__((always_inline)) {
                constexpr int i = decltype(ic)::value;
                if (i >= h) return;
                const int outer = ipow(n, i), inner = ipow(n, h - 1 - i) * Cx;
                gram(B(gb), B(i), outer, inner, ic);
                if (i == 0 && nf > 0) contract(B(0), B(h + 1), 1, inner, 0, FALSE_{});
                const int dst = (i == 0) ? h + 2 : i - 1;
                contract(B(gb), B(dst), outer, inner, i, TRUE_{});
                __syncthreads();
                gb = dst;
            };
            step(I0{});
            step(std::integral_constant<int, 1>{});
            step(std::integral_constant<int, 2>{});
            step(std::integral_constant<int, 3>{});
            step(std::integral_constant<int, 4>{});
            step(std::integral_constant<int, 5>{});
            res = B(gb);
        } else {
            // P_t in B(t) (B(nf-1) holds U); slab chunk Cx x n^nf, fast mode t
            // sees outer Cx n^t, inner n^(nf-1-t); G_0 = dY in B(nf),
            // G_1 -> B(nf+1), G_(t+1) -> B(t-1)
            for (int t = nf - 2; t >= 0; --t) {
                contract(B(t + 1), B(t), Cx * ipow(n, t + 1), ipow(n, nf - 2 - t), h + t + 1,
                         FALSE_{});
                __syncthreads();
            }
            int gb = nf;
            auto step = [&](auto tc) __attribute__((always_inline)) {
                constexpr int t = decltype(tc)::value;
                if (t >= nf) return;
                const int outer = Cx * ipow(n, t), inner = ipow(n, nf - 1 - t);
                gram(B(gb), B(t), outer, inner, tc);
                if (t + 1 < nf) {
                    const int dst = (t == 0) ? nf + 1 : t - 1;
                    contract(B(gb), B(dst), outer, inner, h + t, TRUE_{});
                    gb = dst;
                }
                __syncthreads();
            };
            step(I0{});
            step(std::integral_constant<int, 1>{});
            step(std::integral_constant<int, 2>{});
            step(std::integral_constant<int, 3>{});
            step(std::integral_constant<int, 4>{});
        }

        if (more) stage(nck_);
        staged = more;

        // stores of this item
        const long long eb = (long long)b * E;
        if constexpr (PHASE == 1) {
            const auto gv = to_global(W + 2 * eb);
            const SlabMap m = slab(s0, Cx);
#pragma unroll
            for (int k = 0; k < kAdjPer; ++k) {
                const int q = tid + k * 256;
                if (q < cnt) {
                    int w;
                    const int o = slab_map(m, q, w);
                    gv[o] = res[w];
                }
            }
        } else if constexpr (PHASE == 2) {
            const auto gu = to_global(W + 2 * eb + E + ES * s0);
            const auto gx = io.dx(b);
            const T *bu = B(h + 1);
            const FastDiv dc = fast_div(Cx);
#pragma unroll
            for (int k = 0; k < kAdjPer; ++k) {
                const int q = tid + k * 256;
                if (q < cnt) {
                    if (nf > 0) gu[q] = bu[q];
                    const int is = quot(q, dc);
                    io.put_dx(gx, (int)mul24(is, (int)EF) + s0 + (q - (int)mul24(is, Cx)), res[q]);
                }
            }
        }
        (void)res;
        b = nb;
        ck = nck_;
    }
    if (cur_b >= 0 && PHASE != 1) flush(cur_b);
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

static long long adj_fixed_bytes(int d, int n, int elem)
{
    const long long ddn = (long long)d * n * n;
    return (ddn * 8 + 15) / 16 * 16 + (ddn + 1) / 2 * 2 * elem;
}

// LDS buffers of each pass (chunk-sized): pass 1 a ping-pong pair plus the
// input (two when one fast mode), pass 2 h + 3, pass 3 nf + 2.
static int adj_nbuf(int phase, int h, int nf)
{
    if (phase == 1) return nf == 1 ? 2 : 3;
    if (phase == 2) return h + 3;
    return nf + 2;
}

// Estimated cost of one chunk of a pass, in CU cycles: the larger of the
// instruction time (a contraction round: n^2 FMAs, 2n LDS accesses and index
// arithmetic per fiber; a gram column: n FMAs and n+1 LDS reads; 4 cycles per
// wave instruction) and the HBM time of its streams at about 8 bytes per CU
// cycle.  A strided stream of runs of L bytes fetches about (L + 120) / L
// times its size (the runs start anywhere in 128-byte lines).
struct AdjCostIn {
    int n, elem;
    int contr;         // contractions per chunk
    int grams;         // grams per chunk
    long long fibers;  // fibers per contraction (= columns per gram)
    int barriers;
    double bytes;      // HBM bytes per chunk, run penalties included
};
static double adj_chunk_cost(const AdjCostIn &c)
{
    const int n = c.n;
    const double rounds = (double)((c.fibers + 255) / 256);
    const double parts = n > 0 ? (double)(256 / n) : 256.0;
    const double gcols = std::ceil((double)c.fibers / parts);
    const double instr = c.contr * (rounds * (n * n + 3.0 * n + 14) + n * n / 2.0 + 20) +
                         c.grams * (gcols * (2.0 * n + 10) + 16) + c.barriers * 40.0;
    const double compute = 4.0 * instr;
    const double hbm = c.bytes / 8.0;
    return std::max(compute, hbm) + 400.0;
}
static double run_penalty(double L) { return L >= 4096 ? 1.0 : (L + 120.0) / L; }

// Chunking of an entry for the LDS budget `cap` (bytes): the split h and the
// chunk sizes of minimal estimated cost.  False when no split fits.
static bool plan_adjoint(AdjShape &sh, int elem, long long cap)
{
    const int d = sh.d, n = sh.n;
    const long long fixed = adj_fixed_bytes(d, n, elem);
    bool found = false;
    double best = 0;
    AdjShape pick = sh;
    for (int h = 1; h <= d; ++h) {
        AdjShape s = sh;
        s.h = h;
        s.ES = 1;
        for (int k = 0; k < h; ++k) s.ES *= n;
        s.EF = s.E / s.ES;
        const int nf = d - h;
        if (s.ES > kAdjChunk) break;
        // pass 2: column chunks of width C2 (X and dX strided, V and U contiguous)
        double c2best = -1;
        for (long long C2 = 1; C2 <= std::min<long long>(s.EF, kAdjChunk / s.ES); ++C2) {
            if (fixed + (long long)adj_nbuf(2, h, nf) * s.ES * C2 * elem > cap) break;
            if (nf == 0 && C2 != 1) break;
            const double runb = (double)C2 * elem;
            const double vec = (double)s.ES * C2 * elem;
            AdjCostIn c{n, elem, (h - 1) + h + (nf > 0 ? 1 : 0), h, s.ES * C2 / n, 2 * h - 1,
                        vec * (run_penalty(runb) * 2 + (nf > 0 ? 2 : 0))};
            const long long nck = (s.EF + C2 - 1) / C2;
            const double cost = nck * adj_chunk_cost(c);
            if (c2best < 0 || cost < c2best) {
                c2best = cost;
                s.C2 = (int)C2;
                s.nck2 = (int)nck;
            }
        }
        if (c2best < 0) continue;
        double total = c2best;
        s.C1 = s.C3 = 0;
        if (nf > 0) {
            if (s.EF > kAdjChunk) continue;
            double c1best = -1, c3best = -1;
            for (long long C = 1; C <= std::min<long long>(s.ES, kAdjChunk / s.EF); ++C) {
                const double vec = (double)C * s.EF * elem;
                const double runb = (double)C * s.C2 * elem;  // scratch runs per column chunk
                const long long nck = (s.ES + C - 1) / C;
                if (fixed + (long long)adj_nbuf(1, h, nf) * C * s.EF * elem <= cap) {
                    AdjCostIn c{n, elem, nf, 0, C * s.EF / n, nf, vec * (1 + run_penalty(runb))};
                    const double cost = nck * adj_chunk_cost(c);
                    if (c1best < 0 || cost < c1best) {
                        c1best = cost;
                        s.C1 = (int)C;
                    }
                }
                if (fixed + (long long)adj_nbuf(3, h, nf) * C * s.EF * elem <= cap) {
                    AdjCostIn c{n, elem, 2 * nf - 2, nf, C * s.EF / n, 2 * nf - 1,
                                vec * (1 + run_penalty(runb))};
                    const double cost = nck * adj_chunk_cost(c);
                    if (c3best < 0 || cost < c3best) {
                        c3best = cost;
                        s.C3 = (int)C;
                    }
                }
            }
            if (c1best < 0 || c3best < 0) continue;
            total += c1best + c3best;
        }
        if (!found || total < best) {
            found = true;
            best = total;
            pick = s;
        }
    }
    if (found) sh = pick;
    return found;
}

static bool adjoint_shape(int d, int n, int elem, AdjShape &sh)
{
    long long E = 1;
    for (int k = 0; k < d; ++k) E *= n;
    sh = AdjShape{d, n, 0, E, 0, 0, 0, 0, 0, 0};
    return plan_adjoint(sh, elem, adjoint_lds_cap());
}

template <typename T>
size_t adjoint_scratch_elems(int d, int n)
{
    if (d == 6 && n == 4) return 0;  // on chip (kron_adjoint_plane.hip)
    AdjShape sh;
    if (!adjoint_shape(d, n, (int)sizeof(T), sh)) return (size_t)-1;
    return (sh.h < d) ? (size_t)(2 * sh.E) : 0;
}

// Entries of a pointer-layout d = 6, n = 4 adjoint ordered by dX target
// (pdX), so that runs adding into one X gradient sum it in registers; from
// 2,048 entries, as the forward groups by Y.  False: per-entry atomics.
template <typename T>
static bool adjoint_plane_order(const AdjointArgs<T> &a, GroupOrder &ord)
{
    if (a.layout != Layout::Ptr || a.batch < 2048) return false;
    const size_t bytes = group_workspace_bytes(a.batch);
    bool fresh = false;
    void *ws = stream_workspace(a.stream, bytes, &fresh);
    if (!ws) return false;
    const unsigned long long *keys = nullptr;
    const int *perm = nullptr;
    if (group_by_target(reinterpret_cast<const void *const *>(a.pdX), a.batch, ws, bytes, fresh,
                        &keys, &perm, a.stream) != KRON_OK)
        return false;
    ord.perm = perm;
    ord.keys = keys;
    ord.chunk = 0;
    return true;
}

template <typename T, int N, int P, Layout LAY>
static const void *adjoint_kernel_nf(int nf)
{
    // compile-time fast-mode counts for the common split (passes 1 and 3)
    if constexpr (P != 2 && N >= 4 && N <= 6) {
        switch (nf) {
        case 2: return reinterpret_cast<const void *>(kron_adjoint_kernel<T, N, P, 2, LAY>);
        case 3: return reinterpret_cast<const void *>(kron_adjoint_kernel<T, N, P, 3, LAY>);
        case 4: return reinterpret_cast<const void *>(kron_adjoint_kernel<T, N, P, 4, LAY>);
        default: break;
        }
    }
    return reinterpret_cast<const void *>(kron_adjoint_kernel<T, N, P, 0, LAY>);
}

template <typename T, int P, Layout LAY>
static const void *adjoint_kernel_for(int n, int nf)
{
    switch (n) {
    case 2: return adjoint_kernel_nf<T, 2, P, LAY>(nf);
    case 3: return adjoint_kernel_nf<T, 3, P, LAY>(nf);
    case 4: return adjoint_kernel_nf<T, 4, P, LAY>(nf);
    case 5: return adjoint_kernel_nf<T, 5, P, LAY>(nf);
    case 6: return adjoint_kernel_nf<T, 6, P, LAY>(nf);
    case 7: return adjoint_kernel_nf<T, 7, P, LAY>(nf);
    case 8: return adjoint_kernel_nf<T, 8, P, LAY>(nf);
    default: return reinterpret_cast<const void *>(kron_adjoint_kernel<T, 0, P, 0, LAY>);
    }
}

template <typename T, Layout LAY>
static kron_status launch_adjoint_io(const AdjointArgs<T> &a, const AdjIO<T, LAY> &io)
{
    const int d = a.d, n = a.n, batch = a.batch;
    T *W = a.W;
    AdjShape sh;
    if (!adjoint_shape(d, n, (int)sizeof(T), sh)) return KRON_ENOTSUP;
    const bool plane = d == 6 && n == 4;  // on chip, no scratch
    if (!plane && sh.h < d && !W) return KRON_EWORKSPACE;
    const int nf = d - sh.h;
    if constexpr (!AdjIO<T, LAY>::kAccumulate) {
        // dA is accumulated atomically by the work items: clear it first (a
        // fill kernel, so the sequence stays valid under stream capture)
        const long long nA = (long long)batch * d * n * n;
        hipLaunchKernelGGL(kron_fill_zero<T>,
                           dim3((unsigned)std::min<long long>((nA + 255) / 256, 4096)), dim3(256),
                           0, a.stream, a.dA, nA);
    }
    // d = 6, n = 4: every entry on chip, one launch (kron_adjoint_plane.hip)
    if (plane) {
        GroupOrder ord{nullptr, nullptr, 0};
        const bool grouped = adjoint_plane_order<T>(a, ord);
        return launch_adjoint_plane<T>(a, grouped ? &ord : nullptr);
    }
    const long long fixed = adj_fixed_bytes(d, n, (int)sizeof(T));
    // d = 6, n = 6: compile-time kernels for the three passes
    if (d == 6 && n == 6 && (long long)batch * 36 <= 0x7fffffffLL) {
        for (int phase = 1; phase <= 3; ++phase) {
            const kron_status st = launch_adjoint_rows6<T>(phase, a);
            if (st != KRON_OK) return st;
        }
        return KRON_OK;
    }
    for (int phase = 1; phase <= 3; ++phase) {
        if (phase != 2 && nf == 0) continue;
        const long long chunk = phase == 1   ? (long long)sh.C1 * sh.EF
                                : phase == 2 ? sh.ES * sh.C2
                                             : (long long)sh.C3 * sh.EF;
        const size_t lds = (size_t)(fixed + adj_nbuf(phase, sh.h, nf) * chunk * (long long)sizeof(T));
        const void *kern = phase == 1   ? adjoint_kernel_for<T, 1, LAY>(n, nf)
                           : phase == 2 ? adjoint_kernel_for<T, 2, LAY>(n, nf)
                                        : adjoint_kernel_for<T, 3, LAY>(n, nf);
        if (lds > 65536) {
            hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)lds);
            if (e != hipSuccess) { set_last_hip_error((int)e); return KRON_EHIP; }
        }
        const long long resident = persistent_blocks(kern, lds);
        const long long cap = resident > 0 ? 2 * resident : 4096;
        const long long nck = phase == 2 ? sh.nck2 : (sh.ES + (phase == 1 ? sh.C1 : sh.C3) - 1) /
                                                         (phase == 1 ? sh.C1 : sh.C3);
        const long long grid = std::max(1LL, std::min((long long)batch * nck, cap));
        AdjIO<T, LAY> iov = io;
        AdjShape shv = sh;
        int bv = batch;
        void *args[] = {(void *)&iov, (void *)&W, (void *)&shv, (void *)&bv};
        hipError_t le = hipLaunchKernel(kern, dim3((unsigned)grid), dim3(256), args, lds, a.stream);
        if (le != hipSuccess) { set_last_hip_error((int)le); return KRON_EHIP; }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { set_last_hip_error((int)e); return KRON_EHIP; }
    return KRON_OK;
}

template <typename T>
kron_status launch_adjoint(const AdjointArgs<T> &a)
{
    if (a.batch == 0) return KRON_OK;
    if (a.layout == Layout::Batched) return launch_adjoint_io<T>(a, batched_io(a));
    return launch_adjoint_io<T>(a, ptr_io(a));
}

template kron_status launch_adjoint<double>(const AdjointArgs<double> &);
template kron_status launch_adjoint<float>(const AdjointArgs<float> &);
template size_t adjoint_scratch_elems<double>(int, int);
template size_t adjoint_scratch_elems<float>(int, int);

}  // namespace kron
