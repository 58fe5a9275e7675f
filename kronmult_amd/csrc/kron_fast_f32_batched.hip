// Fast fused kernels, float, batched layout (kronmult{1..6}_batched).
#include "kron_fast.hpp"
#include "kron_dispatch.hpp"

namespace kron {

kron_status fast_batched_f32(int D, int N, const Entries<float, Layout::Batched> &e,
                             bool trans, int batch, hipStream_t s)
{
    if (trans)
        return launch_fast<float, float, Layout::Batched, false, true>(D, N, e, N, batch, s);
    return launch_fast<float, float, Layout::Batched, false, false>(D, N, e, N, batch, s);
}

}  // namespace kron
