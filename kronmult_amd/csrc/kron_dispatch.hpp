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
// The pointer layout on the matrix cores: single-pass shapes (d, n) in
// {(2, 16), (2, 32), (3, 16)}, factor leading dimension lda, Y updated with
// atomics when entries may share it; KRON_ENOTSUP for other shapes.
template <typename T>
kron_status launch_mfma_ptr(int d, int n, const Entries<T, Layout::Ptr> &e, int lda,
                            bool shared, int batch, hipStream_t s);

}  // namespace kron
