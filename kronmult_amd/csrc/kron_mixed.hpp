// kron_mixed.hpp -- types shared by the mixed-(d, n) persistent kernel
// (kron_mixed.hip) and its host-side plan (mixed_plan.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>

#include "kron_internal.hpp"

namespace kron {

constexpr int kMixThreads = 256;       // threads per workgroup
constexpr int kMixLdsBytes = 65536;    // LDS of the largest path (2 workgroups per CU)
constexpr int kMixVariants = 3;        // tile / thread mapping variants per (d, n)

// Work queues of one launch: one per XCD (MI355X has 8, each with its own
// L2), so that the items a plan puts on one queue (the slab items of one
// entry, which all read its X) run on one XCD and share its L2.  Counter q
// sits at kMixQueueStride * q (one 64-byte line each), the exit counter at
// kMixQueueExit; kMixQueueWords words per launch, zero between launches.
constexpr int kMixXcds = 8;
constexpr int kMixQueueStride = 16;
constexpr int kMixQueueExit = kMixXcds * kMixQueueStride;
constexpr int kMixQueueWords = kMixQueueExit + kMixQueueStride;

// Item ranges of the queues: queue q holds items [off[q], off[q + 1]).
struct MixQueues {
    int off[kMixXcds + 1];
};

// One shape class as the kernel reads it (device table).
template <typename T>
struct MixedClass {
    const T *const *pA;  // Aarray[k + b*d]
    const T *const *pX;  // pX[b]
    T *const *pY;        // pY[b]
    int d, n, lda;
    int path;            // compiled path code (mixed_path_code)
    int atomic_y;        // 1: Y updated with atomics
    int pad;
    long long wofs;      // scratch of record r at wofs + r * n^d (two-pass classes)
};

// The pointers of one entry, gathered by the plan from the caller's pointer
// arrays (one 64-byte record, read with scalar loads).
struct MixedEntry {
    const void *A[kMaxDim];
    const void *X;
    void *Y;
};

// A work item: entry records [first, first+count) of class cls (one entry
// for a pass-1 tile).
struct MixedItem {
    int cls;
    int first;
    int count;
    int chunk;  // -1: whole entries; c >= 0: slab c, or pass-1 tile c; -2 - c: pass-2 tile c
};

// What a compiled (d, n, variant) path needs.
struct MixPathInfo {
    bool valid = false;
    bool big = false;       // two passes through per-entry scratch
    bool slab = false;      // an entry as n slabs of its first mode (no scratch)
    int slabs = 0;          // slabs per work item (slab paths)
    bool packed = false;    // several entries per workgroup, atomic Y
    int per_item = 1;       // entries per item when packed
    size_t lds = 0;         // dynamic LDS bytes
    long long scratch = 0;  // scratch elements per workgroup
    int split = 0;          // slow modes of pass 1 (big paths)
    int tiles1 = 0;         // pass-1 tiles (big paths) or slab items (slab paths) per entry
    int tiles2 = 0;         // pass-2 tiles per entry (big paths)
    int threads = 0;        // threads per entry
};

struct MixedLaunch {
    const void *classes = nullptr;  // MixedClass<T>[]
    const MixedItem *items = nullptr;
    const MixedEntry *recs = nullptr;
    int nitems = 0;
    unsigned *queue = nullptr;      // kMixQueueWords counters, zero between launches
    MixQueues q{};                  // the queues' item ranges
    void *scratch = nullptr;        // per-entry scratch of two-pass classes
    int grid = 0;
    size_t lds = 0;
};

// The plan (mixed_plan.cpp), behind the C ABI's kron_mplan.
struct MixedPlanCore;
kron_status mixed_create_f64(MixedPlanCore *&out, int nclass, const kron_mixed_class_f64 *c,
                             const char *cache, int flags, hipStream_t s);
kron_status mixed_create_f32(MixedPlanCore *&out, int nclass, const kron_mixed_class_f32 *c,
                             const char *cache, int flags, hipStream_t s);
kron_status mixed_execute(MixedPlanCore *p, hipStream_t s);
kron_status mixed_capture(MixedPlanCore *p, int applies);
kron_status mixed_replay(MixedPlanCore *p, hipStream_t s);
kron_status mixed_class_info(MixedPlanCore *p, int cls, int *kind, int *variant, int *atomic,
                             double *cost);
kron_status mixed_launch_info(MixedPlanCore *p, int *items, int *grid, int *lds, int *launches);
void mixed_destroy(MixedPlanCore *p);

template <typename T>
bool mixed_path_info(int d, int n, int variant, MixPathInfo &out);
int mixed_path_code(int d, int n, int variant);
template <typename T>
const void *mixed_kernel();
template <typename T>
kron_status mixed_launch(const MixedLaunch &L, hipStream_t stream);

}  // namespace kron
