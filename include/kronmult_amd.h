/*
 * kronmult_amd.h -- C ABI of the MI355X (gfx950) batched Kronecker
 * matrix-vector library.
 *
 *   Y(:,k) += kron(A1(k), ..., Ad(k)) * X(:,k),   k = 1..batchCount, d <= 6
 *
 * Conventions shared by every entry point (they are the reference's):
 *   - column-major storage; in kron(A1,...,Ad) A1 acts on the slowest index
 *     of X and Ad on the fastest (kroncommon.hpp:97-199);
 *   - Y is accumulated (+=).  Wherever two batch entries may name the same Y
 *     vector (xbatched, vbatched, kgemm with beta == 1) the update is an
 *     element-wise device atomic add, as in the reference
 *     (kgemm_nn.hpp:392-394);
 *   - every data pointer and every pointer/leading-dimension array is device
 *     memory owned by the caller; calls never synchronise the stream;
 *   - targets of atomic adds (Y where entries may share it, C with beta == 1,
 *     the pointer-layout adjoint's dX and dA) must be coarse-grained device
 *     memory (hipMalloc, torch): the adds are the hardware's no-return FP64 /
 *     FP32 atomics, which fine-grained memory (hipHostMalloc, coherent
 *     managed memory) does not support.  The reference's atomicAdd compiles
 *     to a compare-and-swap loop that does; this library does not accept
 *     such memory there;
 *   - pointer-layout calls with batchCount >= 2048 order their entries by Y
 *     target on the device and merge runs before the atomic adds (the
 *     pointer-layout adjoint at d = 6, n = 4 orders them by dX target); the
 *     sort scratch (about 16 bytes per entry) is a library-owned device
 *     buffer per (device, stream), allocated on first use and kept for the
 *     process;
 *   - which calls allocate device memory (hipMalloc, never hipFree, never a
 *     synchronisation): kron_xbatched_* (the adjoint included) /
 *     kron_pbatched_* / kron_vbatched_* when the stream's sort workspace
 *     must grow (the first grouped call on a stream, or a larger batch than
 *     before), kron_pbatched_* when its
 *     factor-pointer table must grow, and the plan create calls.  A call
 *     that allocates must not be made under stream capture; warm the stream
 *     up with an uncaptured call of the largest batch first, or use
 *     kron_xbatched_noalloc_f64, or a plan.  Outgrown workspaces are kept
 *     until the process ends, so graphs captured earlier stay valid;
 *   - X is never written and W is used only as scratch when a shape does not
 *     fit on chip (the reference clobbers X in its batched layouts; callers
 *     must still treat X/W contents as unspecified after a call there).
 *   - FP32 (the _f32 entry points and the float overloads) sums each mode's
 *     n terms in FP32 (FP32 MFMA accumulation for n >= 16), where the
 *     reference's float templates sum each small GEMM in double and round
 *     per mode (kgemm_nn.hpp:688-701).  This is a deliberate contract
 *     change: on the same float inputs the results stay within 8 times the
 *     reference's own float error against FP64 (1.4x-4.5x measured, 1e-7 to
 *     2.4e-7 relative in norm; tests/test_gpu_fp32_contract.py), and below it
 *     where many entries merge into one Y.  The adjoint's gradient sums and
 *     the kgemm entry points keep FP64 accumulation;
 *   - stream is a hipStream_t (NULL = the null stream).
 *
 * All functions return a kron_status.  Argument errors are detected on the
 * host before anything is launched; KRON_EHIP reports a failed launch.
 */
#ifndef KRONMULT_AMD_H
#define KRONMULT_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum kron_status {
    KRON_OK = 0,
    KRON_EINVAL = 1,      /* invalid argument (d, n, sizes, null pointers) */
    KRON_EWORKSPACE = 2,  /* workspace (W / pW / Wcapacity) too small */
    KRON_EHIP = 3,        /* HIP runtime reported an error */
    KRON_ENOTSUP = 4,     /* shape outside every compiled kernel */
    KRON_ETOOBIG = 5      /* batchCount above KRON_NOALLOC_MAX_BATCH */
} kron_status;

typedef void *kron_stream_t; /* hipStream_t */

const char *kron_status_string(kron_status s);
/* Last HIP error code seen by the library on this thread (0 if none). */
int kron_last_hip_error(void);
/* Library version, major*10000 + minor*100 + patch. */
int kron_version(void);

/*
 * Contiguous batched layout.  Replaces kronmult{1..6}_batched
 * (kronmult6_batched.cpp:5-25, kernel kronmult6_batched.hpp:14-94).
 *   A: (n, n, d, batchCount), factor k of entry b at ((b*d)+k)*n*n
 *   X, Y, W: (n^d, batchCount)
 * W is scratch (needed only when n^d does not fit on chip).
 */
kron_status kron_batched_f64(int d, int n, const double *A, const double *X,
                             double *Y, double *W, int batchCount,
                             kron_stream_t stream);
kron_status kron_batched_f32(int d, int n, const float *A, const float *X,
                             float *Y, float *W, int batchCount,
                             kron_stream_t stream);

/*
 * Adjoint of the batched layout (no counterpart in the reference).  With
 * K = kron(A1, ..., Ad) per entry and an upstream gradient dY (n^d, batch):
 *   dX = K^T dY                      (n^d, batch), overwritten
 *   dA(:,:,k,b) = d<dY, K X>/dA_k    (n, n, d, batch), overwritten
 * Intermediates are recomputed on chip; only two vectors per in-flight entry
 * go through W, which must hold 2*n^d*batchCount elements (W may be NULL
 * when the whole entry fits on chip; KRON_EWORKSPACE otherwise).
 */
kron_status kron_batched_adjoint_f64(int d, int n, const double *A, const double *X,
                                     const double *dY, double *dX, double *dA, double *W,
                                     int batchCount, kron_stream_t stream);
kron_status kron_batched_adjoint_f32(int d, int n, const float *A, const float *X,
                                     const float *dY, float *dX, float *dA, float *W,
                                     int batchCount, kron_stream_t stream);

/*
 * Scratch the adjoint needs (no counterpart in the reference): elements of W
 * per entry for kron_batched_adjoint_f64 (elem_bytes 8) or _f32 (4) and
 * kron_xbatched_adjoint_* at (d, n) -- 0 when the entry fits on chip, else
 * 2*n^d (two vectors per entry stream through W).  KRON_ENOTSUP for shapes
 * without an adjoint kernel.
 */
kron_status kron_batched_adjoint_workspace(int d, int n, int elem_bytes,
                                           size_t *elems_per_entry);

/*
 * Contiguous factors, pointer vectors.  Replaces kronmult{1..6}_pbatched
 * (kronmult6_pbatched.cpp:5-26, kernel kronmult6_pbatched.hpp:14-81).
 *   A: (n, n, d, batchCount) as in kron_batched_*
 *   pX[b], pY[b], pW[b] -> vectors of n^d (device pointer arrays); pY entries
 *   may coincide (atomic add, as the reference's last kgemm step);
 *   pW[b] is entry b's scratch, used only when n^d does not fit on chip.
 * X is not written (the reference may clobber it).  The device table of
 * factor pointers is built in a library-owned buffer per (device, stream).
 */
kron_status kron_pbatched_f64(int d, int n, const double *A, const double *const *pX,
                              double *const *pY, double *const *pW, int batchCount,
                              kron_stream_t stream);
kron_status kron_pbatched_f32(int d, int n, const float *A, const float *const *pX,
                              float *const *pY, float *const *pW, int batchCount,
                              kron_stream_t stream);

/*
 * Pointer-array layout with shared Y.  Replaces kronmult{1..6}_xbatched
 * (kronmult6_xbatched.cpp:5-39, body kronmult_xbatched.hpp:19-157).
 *   Aarray[k + b*d] -> n x n factor k of entry b, leading dimension lda >= n
 *   pX[b], pY[b]     -> vectors of n^d; pY entries may coincide (atomic add)
 *   pW               -> scratch vectors of n^d: 2*subbatchCount of them when
 *                       batchCount > 1, else 1 (kronmult_xbatched.hpp:125-126);
 *                       used only when n^d does not fit on chip.
 *   subbatchCount == 0 means max(1, batchCount/2) (kronmult_xbatched.hpp:73-75).
 */
kron_status kron_xbatched_f64(int d, int n, const double *const *Aarray, int lda,
                              const double *const *pX, double *const *pY,
                              double *const *pW, int batchCount,
                              int subbatchCount, kron_stream_t stream);
kron_status kron_xbatched_f32(int d, int n, const float *const *Aarray, int lda,
                              const float *const *pX, float *const *pY,
                              float *const *pW, int batchCount,
                              int subbatchCount, kron_stream_t stream);

/*
 * Adjoint of the pointer-array layout (no counterpart in the reference; the
 * layout of kronmult_xbatched.hpp:79-83).  With K_b = kron(A_1, ..., A_d) of
 * entry b (factors Aarray[k + b*d], leading dimension lda) and the upstream
 * gradient of its Y target at pdY[b] (entries sharing a Y target share its
 * gradient), for every entry:
 *   pdX[b]          += K_b^T pdY[b]                 (vector of n^d)
 *   pdA[k + b*d]    += d<pdY[b], K_b pX[b]>/dA_k    (n x n, leading dim ldda)
 * Results are added (device atomics), never overwritten: entries may share
 * X vectors, their gradients and factors, and the caller clears the targets
 * where a sum should start from zero.  W: kron_batched_adjoint_workspace's
 * elements per entry times batchCount, contiguous (entry b at b times that);
 * NULL when that is 0.
 */
kron_status kron_xbatched_adjoint_f64(int d, int n, const double *const *Aarray, int lda,
                                      const double *const *pX, const double *const *pdY,
                                      double *const *pdX, double *const *pdA, int ldda,
                                      double *W, int batchCount, kron_stream_t stream);
kron_status kron_xbatched_adjoint_f32(int d, int n, const float *const *Aarray, int lda,
                                      const float *const *pX, const float *const *pdY,
                                      float *const *pdX, float *const *pdA, int ldda,
                                      float *W, int batchCount, kron_stream_t stream);

/*
 * kron_xbatched_f64 without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under hipGraph stream
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, semantics and kernels as kron_xbatched_f64; the scratch of the
 * Y-target sort (calls with batchCount >= 2048) lives in a fixed device
 * array of the library, about 16.8 MB per GPU, held until the process ends,
 * instead of the library's lazily allocated per-stream workspace.  The call
 * allocates and frees nothing and never synchronises.
 *   batchCount > KRON_NOALLOC_MAX_BATCH: KRON_ETOOBIG, nothing launched.
 * Every call on one GPU uses that GPU's one array, so calls that overlap on
 * one GPU (on different streams, or different threads) race; calls ordered
 * on one stream do not.
 */
#define KRON_NOALLOC_MAX_BATCH 1048576
kron_status kron_xbatched_noalloc_f64(int d, int n, const double *const *Aarray, int lda,
                                      const double *const *pX, double *const *pY,
                                      double *const *pW, int batchCount,
                                      int subbatchCount, kron_stream_t stream);

/*
 * Variable-shape layout: factor k is m[k] x n[k] with leading dimension m[k]
 * (one shape set per call).  Replaces the kernel templates
 * kronmult{1..6}_vbatched (kronmult6_vbatched.hpp:14-58, body
 * kronmult_vbatched.hpp:19-248), which have no host wrapper.
 *   Aarray[k + b*d], pX[b] (prod n), pY[b] (prod m, may coincide)
 *   W: scratch of Wcapacity_bytes.  As in the reference
 *   (kronmult_vbatched.hpp:136-157) the call needs room for one slot of
 *   2*sizeXW elements; a smaller Wcapacity returns KRON_EWORKSPACE where the
 *   reference would overrun.
 */
kron_status kron_vbatched_f64(int d, const int *m, const int *n,
                              const double *const *Aarray,
                              const double *const *pX, double *const *pY,
                              double *W, size_t Wcapacity_bytes, int batchCount,
                              kron_stream_t stream);
kron_status kron_vbatched_f32(int d, const int *m, const int *n,
                              const float *const *Aarray,
                              const float *const *pX, float *const *pY,
                              float *W, size_t Wcapacity_bytes, int batchCount,
                              kron_stream_t stream);

/*
 * The same update with the reference's own per-entry structure on the GPU
 * (kronmult_vbatched.hpp:164-247): one workgroup per entry, X copied into
 * the entry's W slot, one pass through W per mode, the last mode added into
 * Y atomically, sub-batches of Wcapacity / (2*sizeXW) entries.  Arguments
 * and workspace rule as kron_vbatched_*.  A comparison baseline (the plain
 * per-entry kernel the mixed-shape plan is measured against), not a fast
 * path.
 */
kron_status kron_vbatched_plain_f64(int d, const int *m, const int *n,
                                    const double *const *Aarray,
                                    const double *const *pX, double *const *pY,
                                    double *W, size_t Wcapacity_bytes, int batchCount,
                                    kron_stream_t stream);
kron_status kron_vbatched_plain_f32(int d, const int *m, const int *n,
                                    const float *const *Aarray,
                                    const float *const *pX, float *const *pY,
                                    float *W, size_t Wcapacity_bytes, int batchCount,
                                    kron_stream_t stream);

/*
 * xbatched plan (no counterpart in the reference; the reference re-derives
 * nothing between calls).  For repeated applies with an unchanged pointer
 * map, plan_create validates the arguments of kron_xbatched_* and computes,
 * on `stream`, the entries' order by Y target; plan_execute then performs
 * exactly the kron_xbatched_* update (Y += ..., atomic merges) using that
 * order.  The contents of A, X and Y may change between executes; the
 * pointer arrays Aarray, pX, pY must not.  plan_create allocates device
 * memory (about 12 bytes per entry plus sort scratch); plan_destroy
 * synchronises the device and frees it.
 */
typedef struct kron_xplan_s *kron_xplan;
kron_status kron_xbatched_plan_create_f64(kron_xplan *plan, int d, int n,
                                          const double *const *Aarray, int lda,
                                          const double *const *pX, double *const *pY,
                                          double *const *pW, int batchCount,
                                          int subbatchCount, kron_stream_t stream);
kron_status kron_xbatched_plan_create_f32(kron_xplan *plan, int d, int n,
                                          const float *const *Aarray, int lda,
                                          const float *const *pX, float *const *pY,
                                          float *const *pW, int batchCount,
                                          int subbatchCount, kron_stream_t stream);
kron_status kron_xbatched_plan_execute(kron_xplan plan, kron_stream_t stream);
kron_status kron_xbatched_plan_destroy(kron_xplan plan);

/*
 * Mixed-shape plan: many square (d, n) classes applied in one persistent
 * launch (no counterpart in the reference, where each class is a separate
 * kronmult{d}_vbatched launch, kronmult_vbatched.hpp:19-248).  Every class
 * has kronmult_xbatched / vbatched semantics for square factors:
 *   Aarray[k + b*d] -> n x n factor k of entry b, leading dimension lda
 *   pX[b] -> n^d input, pY[b] -> n^d output, Y += kron(A) X
 * with element-wise atomic adds wherever entries (of any classes) may share
 * a Y vector.  plan_create copies the pY arrays to the host once (it
 * synchronises `stream`), groups entries by Y target, checks that distinct
 * targets do not overlap, picks each class's tile / thread mapping and sorts
 * the work items by cost, largest first.  Classes outside the compiled set
 * (d = 6 with n >= 7, n > 8) run as extra launches on the same stream.
 *
 * tune_cache: NULL, or a text file of tuned mappings keyed by GPU
 * architecture, element type, d and n (created if absent).
 * flags: KRON_MIXED_TUNE times every mapping of each class not in the
 *   cache on synthetic data (the caller's arrays are not touched) and adds
 *   the result to the cache; without it uncached classes use the default
 *   mapping and a byte-count cost model.
 *   KRON_MIXED_ATOMIC forces atomic Y updates everywhere;
 *   KRON_MIXED_FORCE_V0 / _V1 / _V2 force mapping 0 / 1 / 2 where it exists
 *   (tests).
 * The contents of A, X and Y may change between executes; the pointer
 * arrays must not.  plan_capture records `applies` executes in a hipGraph
 * (on a plan-owned stream); plan_replay launches it on `stream`.
 * plan_destroy synchronises the device and frees the plan's memory.
 */
typedef struct kron_mixed_class_f64 {
    int d;                          /* factors per entry, 1..6 */
    int n;                          /* factor size (n x n) */
    int lda;                        /* leading dimension, >= n (0 means n) */
    int batchCount;                 /* entries in the class (>= 0) */
    const double *const *Aarray;    /* device: d*batchCount factor pointers */
    const double *const *pX;        /* device: batchCount input pointers */
    double *const *pY;              /* device: batchCount output pointers */
} kron_mixed_class_f64;
typedef struct kron_mixed_class_f32 {
    int d, n, lda, batchCount;
    const float *const *Aarray;
    const float *const *pX;
    float *const *pY;
} kron_mixed_class_f32;

#define KRON_MIXED_TUNE 0x1
#define KRON_MIXED_ATOMIC 0x2
#define KRON_MIXED_FORCE_V0 0x100
#define KRON_MIXED_FORCE_V1 0x200
#define KRON_MIXED_FORCE_V2 0x400

/* How a class of a mixed plan runs (kron_mixed_plan_class_info). */
#define KRON_MIXED_KIND_LAUNCH 0   /* separate launch (outside the compiled set) */
#define KRON_MIXED_KIND_ONCHIP 1   /* one entry per workgroup, entry in LDS */
#define KRON_MIXED_KIND_PACKED 2   /* several entries per workgroup */
#define KRON_MIXED_KIND_TWOPASS 3  /* slow / fast mode passes through scratch */
#define KRON_MIXED_KIND_SLAB 4     /* n slabs of the first mode per entry, one item each */

typedef struct kron_mplan_s *kron_mplan;
kron_status kron_mixed_plan_create_f64(kron_mplan *plan, int nclass,
                                       const kron_mixed_class_f64 *classes,
                                       const char *tune_cache, int flags,
                                       kron_stream_t stream);
kron_status kron_mixed_plan_create_f32(kron_mplan *plan, int nclass,
                                       const kron_mixed_class_f32 *classes,
                                       const char *tune_cache, int flags,
                                       kron_stream_t stream);
kron_status kron_mixed_plan_execute(kron_mplan plan, kron_stream_t stream);
kron_status kron_mixed_plan_capture(kron_mplan plan, int applies);
kron_status kron_mixed_plan_replay(kron_mplan plan, kron_stream_t stream);
/* kind: KRON_MIXED_KIND_*; variant: mapping index; atomic: 1 if Y updates
 * of the class are atomic; cost: estimated cost of one entry (microseconds
 * of the whole GPU when tuned, else bytes); any output pointer may be NULL. */
kron_status kron_mixed_plan_class_info(kron_mplan plan, int cls, int *kind, int *variant,
                                       int *atomic, double *cost);
/* Work items, persistent workgroups and dynamic LDS bytes of the persistent
 * kernel, and the kernel launches one execute makes (1 unless a class is
 * forced onto the two-pass mapping or runs outside the compiled set). */
kron_status kron_mixed_plan_launch_info(kron_mplan plan, int *items, int *grid,
                                        int *lds_bytes, int *launches);
kron_status kron_mixed_plan_destroy(kron_mplan plan);

/*
 * Batched small GEMM with per-entry pointers and leading dimensions (device
 * arrays), uniform mm, nn, kk.  Replaces kgemm_nn_batched
 * (kgemm_nn_batched.cpp:9-42) and kgemm_nt_batched (kgemm_nt_batched.cpp:8-41).
 *   nn: C_b = alpha*A_b*B_b   + beta*C_b,  A mm x kk, B kk x nn, C mm x nn
 *   nt: C_b = alpha*A_b*B_b^T + beta*C_b,  A mm x kk, B nn x kk, C mm x nn
 * beta == 1: element-wise atomic add into C (C may be shared between entries);
 * beta == 0: C is overwritten and never read;
 * otherwise: C = beta*C + alpha*AB (kgemm_nn.hpp:391-400).
 */
kron_status kron_kgemm_nn_batched_f64(int mm, int nn, int kk, double alpha,
                                      const double *const *Aarray, const int *ldAarray,
                                      const double *const *Barray, const int *ldBarray,
                                      double beta, double *const *Carray,
                                      const int *ldCarray, int batchCount,
                                      kron_stream_t stream);
kron_status kron_kgemm_nt_batched_f64(int mm, int nn, int kk, double alpha,
                                      const double *const *Aarray, const int *ldAarray,
                                      const double *const *Barray, const int *ldBarray,
                                      double beta, double *const *Carray,
                                      const int *ldCarray, int batchCount,
                                      kron_stream_t stream);
kron_status kron_kgemm_nn_batched_f32(int mm, int nn, int kk, float alpha,
                                      const float *const *Aarray, const int *ldAarray,
                                      const float *const *Barray, const int *ldBarray,
                                      float beta, float *const *Carray,
                                      const int *ldCarray, int batchCount,
                                      kron_stream_t stream);
kron_status kron_kgemm_nt_batched_f32(int mm, int nn, int kk, float alpha,
                                      const float *const *Aarray, const int *ldAarray,
                                      const float *const *Barray, const int *ldBarray,
                                      float beta, float *const *Carray,
                                      const int *ldCarray, int batchCount,
                                      kron_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* KRONMULT_AMD_H */
