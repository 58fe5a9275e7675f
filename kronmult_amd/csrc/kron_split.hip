// kron_split.hip -- square entries too large for LDS whose factors fit the
// SGPRs (n <= 6; in practice d = 6, n = 6: 46,656 elements), in two passes
// through the entry's scratch vector with every extent known at compile time.
//
// The entry is viewed as R x C, R = n^H rows over the slow modes 1..H and
// C = n^(D-H) contiguous columns over the fast modes (the element index is
// row * C + column, A1 on the slowest index as in kroncommon.hpp:97-199):
//
//   pass 1 (columns): one thread per column holds the column's R elements in
//       registers and contracts the slow modes there, from X into the
//       scratch T.  Consecutive threads take consecutive columns, so every
//       load and store of a wave is one contiguous 512-byte run; no LDS, no
//       barrier.
//   pass 2 (rows): one workgroup per row stages the row of T in LDS,
//       contracts the fast modes one after another in place (a thread owns
//       whole fibers, the factor in SGPRs) and adds the row into Y
//       (atomically where entries may share Y).
//
// HBM traffic: X, T written and read, Y read and written -- five vectors,
// each streamed contiguously.  The scratch is R x C in the same index order,
// one of the reference's own W sizes (kronmult_vbatched.hpp:129-153), so the
// reference's workspace arguments suffice.
#include "kron_dispatch.hpp"
#include "kron_split.hpp"

namespace kron {


// Slow modes Q..H-1 of a column in registers (mode Q has stride n^(H-1-Q)
// in the row index); fac[q] is factor q's address (wave-uniform).
template <typename T, typename Acc, int N, int H, int Q, int LEN, typename FP>
__device__ __forceinline__ void split_cols_modes(Acc (&v)[LEN], const FP (&fac)[H], int lda)
{
    if constexpr (Q < H) {
        Acc a[N][N];
        load_factor<T, Acc, N, false, true>(a, fac[Q], lda);
        split_contract_regs<Acc, N, cpow(N, H - 1 - Q), LEN>(v, a);
        // the next factor's scalar loads stay behind this mode's FMAs: both
        // factors at once would not fit the SGPRs
        __builtin_amdgcn_sched_barrier(0);
        split_cols_modes<T, Acc, N, H, Q + 1, LEN>(v, fac, lda);
    }
}

// Pass 1: workgroup (entry b0 + blockIdx.x / NB1, column block), one column
// per thread: T[r*C + c] = (A_1 (x) .. (x) A_H contracted over r) X[. * C + c].
// Every per-entry address is formed first, so the argument block's SGPRs
// are free for the factor (n*n doubles in SGPRs).
template <typename T, typename Acc, int D, int N, Layout LAY>
__global__ void __launch_bounds__(kSplitTB1)
kron_split_cols(Entries<T, LAY> ent, ScratchDesc<T> scr, int lda, int b0)
{
    using G = SplitGeom<D, N>;
    const int slot = blockIdx.x / G::NB1;
    const int b = b0 + slot;
    const int c = (blockIdx.x % G::NB1) * G::TB1 + threadIdx.x;
    const bool act = c < G::C;
    const int cc = act ? c : G::C - 1;  // idle lanes read a valid column
    constexpr unsigned kBytes = (unsigned)G::R * G::C * sizeof(T);
    const auto xr = vec_rsrc(uniform_ptr(ent.x(b)), kBytes);
    const auto wr = vec_rsrc(uniform_ptr(scr.get(slot, b)), kBytes);
    decltype(ent.template factor_g<true>(b, 0)) fac[G::H];
#pragma unroll
    for (int q = 0; q < G::H; ++q) fac[q] = ent.template factor_g<true>(b, q);
    Acc v[G::R];
#pragma unroll
    for (int r = 0; r < G::R; ++r)
        v[r] = Acc(buf_load((T *)nullptr, xr, cc * (int)sizeof(T), r * G::C * (int)sizeof(T)));
    split_cols_modes<T, Acc, N, G::H, 0, G::R>(v, fac, lda);
    if (act) {
#pragma unroll
        for (int r = 0; r < G::R; ++r)
            buf_store(T(v[r]), wr, c * (int)sizeof(T), r * G::C * (int)sizeof(T));
    }
}

template <typename T, typename Acc, int D, int N, int Q, Layout LAY>
__device__ __forceinline__ void split_rows_modes(T *row, const Entries<T, LAY> &ent, int b,
                                                 int lda)
{
    using G = SplitGeom<D, N>;
    if constexpr (Q < D) {
        Acc a[N][N];
        load_factor<T, Acc, N, false, true>(a, ent.template factor_g<true>(b, Q), lda);
        split_contract_lds<T, Acc, N, cpow(N, D - 1 - Q), G::C>(row, a);
        __syncthreads();
        split_rows_modes<T, Acc, D, N, Q + 1, LAY>(row, ent, b, lda);
    }
}

// Pass 2: workgroup (entry b0 + blockIdx.x / R, row blockIdx.x % R):
// Y[row] += (A_{H+1} (x) .. (x) A_D) T[row].
template <typename T, typename Acc, int D, int N, Layout LAY, bool SHARED>
__global__ void __launch_bounds__(256)
kron_split_rows(Entries<T, LAY> ent, ScratchDesc<T> scr, int lda, int b0)
{
    using G = SplitGeom<D, N>;
    __shared__ T row[G::C];
    const int slot = blockIdx.x / G::R;
    const int b = b0 + slot;
    const int r = blockIdx.x % G::R;
    const int tid = threadIdx.x;
    const auto t = to_global(scr.get(slot, b)) + (long long)r * G::C;
    const auto y = ent.template y_g<true>(b) + (long long)r * G::C;
    T tv[G::PER2];
    T yv[SHARED ? 1 : G::PER2];
#pragma unroll
    for (int k = 0; k < G::PER2; ++k) {
        const int i = tid + k * 256;
        if (G::C % 256 == 0 || i < G::C) {
            tv[k] = t[i];
            if constexpr (!SHARED) yv[k] = y[i];
        }
    }
#pragma unroll
    for (int k = 0; k < G::PER2; ++k) {
        const int i = tid + k * 256;
        if (G::C % 256 == 0 || i < G::C) row[i] = tv[k];
    }
    __syncthreads();
    split_rows_modes<T, Acc, D, N, G::H, LAY>(row, ent, b, lda);
#pragma unroll
    for (int k = 0; k < G::PER2; ++k) {
        const int i = tid + k * 256;
        if (G::C % 256 == 0 || i < G::C) {
            if constexpr (SHARED)
                atomic_add(y + i, row[i]);
            else
                y[i] = T(Acc(yv[k]) + Acc(row[i]));
        }
    }
}

template <typename T, int D, int N, Layout LAY, bool SHARED>
static kron_status launch_split_one(const Entries<T, LAY> &ent, const ScratchDesc<T> &scr,
                                    int group, int lda, int batch, hipStream_t s)
{
    using G = SplitGeom<D, N>;
    for (int b0 = 0; b0 < batch; b0 += group) {
        const int cnt = std::min(group, batch - b0);
        hipLaunchKernelGGL((kron_split_cols<T, T, D, N, LAY>), dim3((unsigned)(cnt * G::NB1)),
                           dim3(G::TB1), 0, s, ent, scr, lda, b0);
        hipLaunchKernelGGL((kron_split_rows<T, T, D, N, LAY, SHARED>),
                           dim3((unsigned)(cnt * G::R)), dim3(256), 0, s, ent, scr, lda, b0);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            set_last_hip_error((int)e);
            return KRON_EHIP;
        }
    }
    return KRON_OK;
}

bool split_supported(int d, int n) { return d == 6 && n == 6; }

template <typename T>
kron_status launch_split(const KronArgs<T> &a, const ScratchDesc<T> &scr, int group)
{
    if (!split_supported(a.d, a.n[0])) return KRON_ENOTSUP;
    const int lda = a.lda > 0 ? a.lda : a.n[0];
    // (batch * per-entry workgroups stays far below 2^31 for int batchCount)
    if (a.layout == Layout::Batched) {
        Entries<T, Layout::Batched> e{a.A, a.X, a.Y, a.d, 36, 46656, 46656};
        return launch_split_one<T, 6, 6, Layout::Batched, false>(e, scr, group, lda, a.batch,
                                                                 a.stream);
    }
    Entries<T, Layout::Ptr> e{a.pA, a.pX, a.pY, a.d};
    if (a.y_shared)
        return launch_split_one<T, 6, 6, Layout::Ptr, true>(e, scr, group, lda, a.batch, a.stream);
    return launch_split_one<T, 6, 6, Layout::Ptr, false>(e, scr, group, lda, a.batch, a.stream);
}

template kron_status launch_split<double>(const KronArgs<double> &, const ScratchDesc<double> &,
                                          int);
template kron_status launch_split<float>(const KronArgs<float> &, const ScratchDesc<float> &,
                                         int);

}  // namespace kron
