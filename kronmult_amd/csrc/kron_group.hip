// kron_group.hip -- order the entries of a pointer-layout call by Y target.
//
// The reference merges every entry into its Y vector with one atomic add per
// element (kgemm_nn.hpp:392-394).  On MI355X FP64 atomics run at about
// 1.9e11 element-updates/s, so at d=6, n=4 the atomics alone bound a batch of
// 65,536 entries to ~1.4 ms.  Sorting the entries by Y pointer lets a
// workgroup accumulate consecutive entries with the same target in registers
// and flush once per run.  Correctness never depends on the order: a flush
// happens whenever the full 64-bit target changes, so an imperfect order only
// costs extra flushes.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "kron_internal.hpp"

namespace kron {

// Target pointers are at least 4-byte aligned and below 2^48 in the device
// address space: bits [2, 48) carry the order.
constexpr unsigned kKeyBeginBit = 2;
constexpr unsigned kKeyEndBit = 48;

static size_t sort_temp_bytes(int batch)
{
    size_t bytes = 0;
    rocprim::counting_iterator<int> iota(0);
    (void)rocprim::radix_sort_pairs(nullptr, bytes, (const unsigned long long *)nullptr,
                                    (unsigned long long *)nullptr, iota, (int *)nullptr,
                                    (unsigned)batch, kKeyBeginBit, kKeyEndBit, hipStream_t(0));
    return bytes;
}

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

size_t group_workspace_bytes(int batch)
{
    return align256((size_t)batch * 8) + align256((size_t)batch * 4) +
           align256(sort_temp_bytes(batch));
}

kron_status group_by_target(const void *const *pY, int batch, void *ws, size_t ws_bytes,
                            const unsigned long long **keys_sorted, const int **perm,
                            hipStream_t stream)
{
    unsigned char *p = static_cast<unsigned char *>(ws);
    unsigned long long *keys_out = reinterpret_cast<unsigned long long *>(p);
    p += align256((size_t)batch * 8);
    int *perm_out = reinterpret_cast<int *>(p);
    p += align256((size_t)batch * 4);
    size_t temp_bytes = sort_temp_bytes(batch);
    if (ws_bytes < group_workspace_bytes(batch)) return KRON_EWORKSPACE;
    rocprim::counting_iterator<int> iota(0);
    hipError_t e = rocprim::radix_sort_pairs(
        p, temp_bytes, reinterpret_cast<const unsigned long long *>(pY), keys_out, iota, perm_out,
        (unsigned)batch, kKeyBeginBit, kKeyEndBit, stream);
    if (e != hipSuccess) {
        set_last_hip_error((int)e);
        return KRON_EHIP;
    }
    *keys_sorted = keys_out;
    *perm = perm_out;
    return KRON_OK;
}

}  // namespace kron
