// Fast fused kernels, float, pointer layout (kronmult{1..6}_xbatched, square vbatched).
#include "kron_fast.hpp"
#include "kron_dispatch.hpp"

namespace kron {

kron_status fast_ptr_f32(int D, int N, const Entries<float, Layout::Ptr> &e, int lda,
                         int batch, hipStream_t s, const GroupOrder *ord)
{
    return launch_fast<float, float, Layout::Ptr, true, false>(D, N, e, lda, batch, s, ord);
}

}  // namespace kron
