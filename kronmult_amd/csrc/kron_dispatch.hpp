// kron_dispatch.hpp -- entry points of the separately compiled fast-kernel
// translation units (kron_fast_*.hip), one per (element type, layout).
#pragma once

#include "kron_device.hpp"
#include "kron_fast.hpp"

namespace kron {

// Batched layout (Y private per entry).  trans selects the adjoint factors.
kron_status fast_batched_f64(int D, int N, const Entries<double, Layout::Batched> &e,
                             bool trans, int batch, hipStream_t s);
kron_status fast_batched_f32(int D, int N, const Entries<float, Layout::Batched> &e,
                             bool trans, int batch, hipStream_t s);
// Pointer layout (Y may be shared -> atomics), leading dimension lda.  With
// ord != nullptr the entries are visited in Y-target order and merged.
kron_status fast_ptr_f64(int D, int N, const Entries<double, Layout::Ptr> &e, int lda,
                         int batch, hipStream_t s, const GroupOrder *ord);
kron_status fast_ptr_f32(int D, int N, const Entries<float, Layout::Ptr> &e, int lda,
                         int batch, hipStream_t s, const GroupOrder *ord);

// The MFMA large-n path (kron_mfma.hip), batched layout, n in {16, 32}.
bool mfma_supported(int d, int n, bool *needs_w);
template <typename T>
kron_status launch_mfma(int d, int n, const Entries<T, Layout::Batched> &e, T *W, int batch,
                        hipStream_t s);
// The pointer layouts on the matrix cores, the same (d, n) set: factor
// leading dimension lda, Y updated with atomics.  Two-pass shapes need the
// caller's scratch (scr, entries per launch group); KRON_ENOTSUP for shapes
// outside the set.
template <typename T>
kron_status launch_mfma_ptr(int d, int n, const Entries<T, Layout::Ptr> &e,
                            const ScratchDesc<T> *scr, int group, int lda, int batch,
                            hipStream_t s);

// Square entries larger than LDS with n <= 6 (kron_split.hip): d = 6, n = 6.
// Two passes through the caller's scratch (scr), `group` entries at a time
// (entries of one group have distinct scratch vectors).
bool split_supported(int d, int n);
template <typename T>
kron_status launch_split(const KronArgs<T> &a, const ScratchDesc<T> &scr, int group);

}  // namespace kron
