// kron_internal.hpp -- host-side interface between the C ABI (capi.cpp) and
// the HIP kernels (kron_*.hip).  Not installed; no torch types anywhere.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "../../include/kronmult_amd.h"

namespace kron {

constexpr int kMaxDim = 6;

// Where a call's entries live.  Batched: contiguous arrays (A (n,n,d,batch),
// X/Y/W (n^d,batch)).  Ptr: per-entry pointer arrays in device memory
// (Aarray[k + b*d], pX[b], pY[b]); Y targets may coincide.
enum class Layout : int { Batched = 0, Ptr = 1 };

// Where the per-entry scratch of a shape that does not fit in LDS comes from
// (mirrors the reference's workspace rules).
enum class Scratch : int {
    None = 0,      // shape fits on chip; no scratch is touched
    BatchedW = 1,  // W + b*sizeX (batched layout: W is (n^d, batch))
    PtrSlots = 2,  // pW[2*slot] (xbatched, batch > 1; kronmult_xbatched.hpp:125-126)
    PtrSlot0 = 3,  // pW[0]      (xbatched, batch == 1)
    Linear = 4,    // Wbase + slot*slot_stride (vbatched; kronmult_vbatched.hpp:184-185)
    PtrEntry = 5   // pW[b], one vector per entry (pbatched; kronmult6_pbatched.hpp:63-70)
};

// Entry order for merged (grouped) pointer-layout calls: perm[r] is the r-th
// entry in Y-target order, keys[r] = pY[perm[r]]; chunk is set by the launcher.
struct GroupOrder {
    const int *perm;
    const unsigned long long *keys;
    int chunk;
};

template <typename T>
struct KronArgs {
    int d = 0;
    int m[kMaxDim] = {1, 1, 1, 1, 1, 1};  // output extents (rows of A_k)
    int n[kMaxDim] = {1, 1, 1, 1, 1, 1};  // input extents (cols of A_k)
    int lda = 0;          // leading dim of every factor; 0 -> ld_k = m_k
    bool trans = false;   // apply A_k^T (adjoint); factors are then n_k x m_k
    Layout layout = Layout::Batched;
    // batched layout
    const T *A = nullptr;
    const T *X = nullptr;
    T *Y = nullptr;
    T *W = nullptr;
    // pointer layout
    const T *const *pA = nullptr;
    const T *const *pX = nullptr;
    T *const *pY = nullptr;
    T *const *pW = nullptr;
    T *Wbase = nullptr;
    size_t slot_stride = 0;  // elements between Linear scratch slots
    int nslots = 0;          // number of scratch slots available
    Scratch scratch = Scratch::None;
    bool y_shared = true;    // entries may share Y -> atomic adds
    int batch = 0;
    hipStream_t stream = nullptr;
    // pre-computed Y-target order (an xbatched plan); nullptr -> decide per call
    const GroupOrder *order = nullptr;
    // caller-provided device scratch for the Y-target sort (cleared by every
    // call that uses it); nullptr -> the library's per-stream workspace
    void *sort_ws = nullptr;
    size_t sort_ws_bytes = 0;
};

// Validates kron_xbatched_* arguments and fills `a` (the reference's
// subbatchCount and pW slot rules); KRON_OK or the status to return.
template <typename T>
kron_status xbatched_args(KronArgs<T> &a, int d, int n, const T *const *Aarray, int lda,
                          const T *const *pX, T *const *pY, T *const *pW, int batch,
                          int subbatch, kron_stream_t stream);

// d = 1 entries (a batched matrix-vector product, kron_gemv.hip): KRON_ENOTSUP
// for other d or transposed factors.
template <typename T>
kron_status launch_gemv(const KronArgs<T> &a);

// d = 2 entries with factor extents up to 32 (kron_d2.hip): KRON_ENOTSUP for
// other shapes or transposed factors.
template <typename T>
kron_status launch_d2(const KronArgs<T> &a);
// d >= 3 entries with mixed per-mode extents whose intermediates fit a
// wave's LDS slice (kron_small.hip): a wave per entry.  KRON_ENOTSUP when
// the shape does not fit.
template <typename T>
kron_status launch_small(const KronArgs<T> &a);

// Launch the forward kernel family.  Returns KRON_OK, KRON_ENOTSUP when no
// kernel covers the shape, KRON_EWORKSPACE when scratch is needed but absent,
// or KRON_EHIP on a launch failure.
template <typename T>
kron_status launch_kron(const KronArgs<T> &a);

// Whether launch_kron merges this call's entries by Y target (given an order).
template <typename T>
bool kron_groupable(const KronArgs<T> &a);

// Scratch elements per entry that launch_kron needs for this shape (0 when the
// shape runs entirely on chip).
template <typename T>
size_t kron_scratch_elems(const KronArgs<T> &a);

// Batched kgemm (kgemm_nn_batched.hpp:8-79 / kgemm_nt_batched.hpp:8-75).
template <typename T>
kron_status launch_kgemm(bool transB, int mm, int nn, int kk, T alpha,
                         const T *const *A, const int *ldA, const T *const *B,
                         const int *ldB, T beta, T *const *C, const int *ldC,
                         int batch, hipStream_t stream);

// Y-target ordering of pointer-layout entries (kron_group.hip).  The start
// of the workspace must be zero before the first call on it (pass fresh =
// true for a new workspace: the call clears it); every call leaves it zero.
// Workspace: the bucket counts (sized for the largest bucket count, so the
// zeroed region does not move with the batch), the finishing ticket and the
// used-bucket count, then bucket bases and the used-bucket list, per-entry
// ranks, the order and the sorted keys.
constexpr int kGroupMaxBits = 16;
constexpr int group_bits(int batch)
{
    int bits = 10;
    while (bits < kGroupMaxBits && (1 << bits) < batch) ++bits;
    return bits;
}
constexpr size_t group_align256(size_t v) { return (v + 255) & ~(size_t)255; }
constexpr size_t kGroupCountBytes = (size_t)4 << kGroupMaxBits;
constexpr size_t kGroupZeroBytes = kGroupCountBytes + 256;
constexpr size_t group_workspace_bytes(int batch)
{
    return kGroupZeroBytes + 2 * group_align256((size_t)(1 << group_bits(batch)) * 4) +
           2 * group_align256((size_t)batch * 4) + group_align256((size_t)batch * 8);
}
kron_status group_by_target(const void *const *pY, int batch, void *ws, size_t ws_bytes,
                            bool fresh, const unsigned long long **keys_sorted, const int **perm,
                            hipStream_t stream);
// Device scratch owned by the library, one buffer per (device, stream),
// grown on demand and kept for the life of the process.  Work on one stream
// is ordered, so calls on the same stream may reuse it.  nullptr on failure;
// *fresh (if given) tells whether the buffer was just allocated.
void *stream_workspace(hipStream_t stream, size_t bytes, bool *fresh = nullptr);
// The same with an independent buffer per slot (0: the Y-target sort,
// 1: pointer tables built on the device).
void *stream_workspace_slot(hipStream_t stream, size_t bytes, int slot, bool *fresh = nullptr);

// Device table out[i] = A + i * fstride, i < count (pbatched factor pointers),
// in the stream's slot-1 workspace.  nullptr on failure.
template <typename T>
const T *const *factor_table(const T *A, long long fstride, long long count, hipStream_t s);

// One adjoint call (kron_adjoint.hip): dX = K^T dY and every dA_k per entry.
//   Batched: A (n,n,d,batch), X/dY/dX (n^d,batch), dA (n,n,d,batch); dX and
//            dA are overwritten.
//   Ptr:     factor k of entry b at pA[b*d+k] (leading dimension lda), X and
//            dY vectors at pX[b] and pdY[b]; K_b^T dY is added into pdX[b]
//            and dA_k into pdA[b*d+k] (leading dimension ldda), atomically:
//            entries may share any of these.
// W: adjoint_scratch_elems(d, n) elements per entry, entry b's at
// W + b * that (nullptr when it is 0).
template <typename T>
struct AdjointArgs {
    Layout layout = Layout::Batched;
    int d = 0, n = 0, batch = 0;
    const T *A = nullptr, *X = nullptr, *dY = nullptr;
    T *dX = nullptr, *dA = nullptr;
    const T *const *pA = nullptr, *const *pX = nullptr, *const *pdY = nullptr;
    T *const *pdX = nullptr, *const *pdA = nullptr;
    int lda = 0, ldda = 0;
    T *W = nullptr;
    hipStream_t stream = nullptr;
};
template <typename T>
kron_status launch_adjoint(const AdjointArgs<T> &a);
template <typename T>
size_t adjoint_scratch_elems(int d, int n);
// The d = 6, n = 6 adjoint (kron_adjoint_rows.hip), pass 1 (V into W's
// first vector of each entry), 2 (U, dX, dA_1, dA_2) or 3 (dA_3..dA_6); the
// dA sums are added into dA (cleared beforehand in the batched layout).
// KRON_ENOTSUP when batchCount * 36 exceeds the int range.
template <typename T>
kron_status launch_adjoint_rows6(int pass, const AdjointArgs<T> &a);
// The d = 6, n = 4 adjoint on chip (kron_adjoint_plane.hip); ord: the
// pointer layout's entries in dX-target order (nullptr: per-entry order).
// dA must be cleared beforehand in the batched layout.
template <typename T>
kron_status launch_adjoint_plane(const AdjointArgs<T> &a, const GroupOrder *ord);

// The reference's per-entry vbatched structure (kron_plain.hip): one
// workgroup per entry, every mode a pass through the entry's scratch slot of
// 2*slot elements in W, nslots entries per launch.
template <typename T>
kron_status launch_plain(int d, const int *m, const int *n, const T *const *pA,
                         const T *const *pX, T *const *pY, T *W, long long slot, int nslots,
                         int batch, hipStream_t stream);

// Per-thread record of the last HIP error (kron_last_hip_error()).
void set_last_hip_error(int code);
int max_lds_bytes();
// Workgroups of `kernel` (256 threads, `lds` dynamic bytes) that fit on the
// current device at once: CUs x resident blocks per CU (0 if unknown).
long long persistent_blocks(const void *kernel, size_t lds, int threads = 256);

}  // namespace kron
