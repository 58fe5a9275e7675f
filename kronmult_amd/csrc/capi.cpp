// capi.cpp -- C ABI (include/kronmult_amd.h) and the drop-in C++ overloads
// (include/kronmult.hpp).  Host code only: validates arguments, applies the
// reference's workspace rules, and hands a KronArgs to the kernel launchers.
#include <cstdio>
#include <map>
#include <mutex>
#include <tuple>
#include <utility>
#include <vector>

#include "kron_internal.hpp"
#include "kron_mixed.hpp"
#include "../../include/kronmult.hpp"

namespace kron {

static thread_local int g_last_hip_error = 0;
void set_last_hip_error(int code) { g_last_hip_error = code; }

int max_lds_bytes()
{
    static std::mutex mu;
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    std::lock_guard<std::mutex> lk(mu);
    if (cached[dev] == 0) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
                hipSuccess || v <= 0)
            v = 65536;
        cached[dev] = v;
    }
    return cached[dev];
}

long long persistent_blocks(const void *kernel, size_t lds, int threads)
{
    static std::mutex mu;
    static std::map<std::pair<const void *, int>, long long> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_pair(kernel, (dev * 1048576 + (int)lds) * 2 + (threads > 256));
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int cus = 0, per_cu = 0;
    long long r = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds) == hipSuccess &&
        cus > 0 && per_cu > 0)
        r = (long long)cus * per_cu;
    cache[key] = r;
    return r;
}

void *stream_workspace(hipStream_t stream, size_t bytes, bool *fresh)
{
    return stream_workspace_slot(stream, bytes, 0, fresh);
}

void *stream_workspace_slot(hipStream_t stream, size_t bytes, int slot, bool *fresh)
{
    struct Buf {
        void *ptr = nullptr;
        size_t bytes = 0;
    };
    static std::mutex mu;
    static std::map<std::tuple<int, hipStream_t, int>, Buf> bufs;
    // Outgrown buffers are never freed: queued work on the stream, or a graph
    // captured earlier, may still use them.  Sizes double, so the retired
    // buffers of a slot add up to less than its live one.
    static std::vector<void *> retired;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    Buf &b = bufs[std::make_tuple(dev, stream, slot)];
    if (fresh) *fresh = false;
    if (b.bytes < bytes) {
        size_t want = b.bytes ? 2 * b.bytes : bytes + bytes / 4;
        while (want < bytes) want *= 2;
        void *p = nullptr;
        if (hipMalloc(&p, want) != hipSuccess) return nullptr;
        if (b.ptr) retired.push_back(b.ptr);
        b.ptr = p;
        b.bytes = want;
        if (fresh) *fresh = true;
    }
    return b.ptr;
}

// n^d with overflow check against the reference's int sizes.
static bool ipow_ok(int n, int d, long long &out)
{
    long long r = 1;
    for (int k = 0; k < d; ++k) {
        r *= n;
        if (r > 0x7fffffffLL) return false;
    }
    out = r;
    return true;
}

template <typename T>
static kron_status batched_impl(int d, int n, const T *A, const T *X, T *Y, T *W, int batch,
                                kron_stream_t stream)
{
    long long E = 0;
    if (d < 1 || d > kMaxDim || n < 1 || batch < 0 || !ipow_ok(n, d, E)) return KRON_EINVAL;
    if (batch == 0) return KRON_OK;
    if (!A || !X || !Y) return KRON_EINVAL;
    KronArgs<T> a;
    a.d = d;
    for (int k = 0; k < d; ++k) a.m[k] = a.n[k] = n;
    a.lda = n;
    a.layout = Layout::Batched;
    a.A = A;
    a.X = X;
    a.Y = Y;
    a.W = W;
    a.scratch = Scratch::BatchedW;
    a.y_shared = false;
    a.batch = batch;
    a.stream = static_cast<hipStream_t>(stream);
    if (kron_scratch_elems(a) > 0 && !W) return KRON_EWORKSPACE;
    return launch_kron(a);
}

// kronmult6_pbatched.hpp:14-81: factor k of entry b at A + (b*d + k)*n*n,
// X/Y/W through pointer arrays, Y possibly shared (atomic last step).
template <typename T>
static kron_status pbatched_impl(int d, int n, const T *A, const T *const *pX, T *const *pY,
                                 T *const *pW, int batch, kron_stream_t stream)
{
    long long E = 0;
    if (d < 1 || d > kMaxDim || n < 1 || batch < 0 || !ipow_ok(n, d, E)) return KRON_EINVAL;
    if (batch == 0) return KRON_OK;
    if (!A || !pX || !pY) return KRON_EINVAL;
    KronArgs<T> a;
    a.d = d;
    for (int k = 0; k < d; ++k) a.m[k] = a.n[k] = n;
    a.lda = n;
    a.layout = Layout::Ptr;
    a.pX = pX;
    a.pY = pY;
    a.pW = pW;
    a.scratch = Scratch::PtrEntry;
    a.y_shared = true;
    a.batch = batch;
    a.stream = static_cast<hipStream_t>(stream);
    if (kron_scratch_elems(a) > 0 && !pW) return KRON_EWORKSPACE;
    a.pA = factor_table<T>(A, (long long)n * n, (long long)d * batch, a.stream);
    if (!a.pA) return KRON_EHIP;
    return launch_kron(a);
}

template <typename T>
kron_status xbatched_args(KronArgs<T> &a, int d, int n, const T *const *Aarray, int lda,
                                 const T *const *pX, T *const *pY, T *const *pW, int batch,
                                 int subbatch, kron_stream_t stream)
{
    long long E = 0;
    if (d < 1 || d > kMaxDim || n < 1 || lda < n || batch < 0 || subbatch < 0 ||
        !ipow_ok(n, d, E))
        return KRON_EINVAL;
    if (batch > 0 && (!Aarray || !pX || !pY)) return KRON_EINVAL;
    // kronmult_xbatched.hpp:73-75
    if (subbatch == 0) subbatch = (batch / 2 > 1) ? batch / 2 : 1;
    a.d = d;
    for (int k = 0; k < d; ++k) a.m[k] = a.n[k] = n;
    a.lda = lda;
    a.layout = Layout::Ptr;
    a.pA = Aarray;
    a.pX = pX;
    a.pY = pY;
    a.pW = pW;
    // kronmult_xbatched.hpp:125-126: slots pW[2*(iz-1)], pW[2*(iz-1)+1] for a
    // sub-batch of subbatch entries; a single entry uses pW[0].
    a.scratch = (batch > 1) ? Scratch::PtrSlots : Scratch::PtrSlot0;
    a.nslots = (batch > 1) ? (subbatch < batch ? subbatch : batch) : 1;
    a.y_shared = true;
    a.batch = batch;
    a.stream = static_cast<hipStream_t>(stream);
    if (batch > 0 && kron_scratch_elems(a) > 0 && !pW) return KRON_EWORKSPACE;
    return KRON_OK;
}

template kron_status xbatched_args<double>(KronArgs<double> &, int, int, const double *const *,
                                          int, const double *const *, double *const *,
                                          double *const *, int, int, kron_stream_t);
template kron_status xbatched_args<float>(KronArgs<float> &, int, int, const float *const *, int,
                                         const float *const *, float *const *, float *const *,
                                         int, int, kron_stream_t);

template <typename T>
static kron_status xbatched_impl(int d, int n, const T *const *Aarray, int lda,
                                 const T *const *pX, T *const *pY, T *const *pW, int batch,
                                 int subbatch, kron_stream_t stream)
{
    KronArgs<T> a;
    kron_status st = xbatched_args(a, d, n, Aarray, lda, pX, pY, pW, batch, subbatch, stream);
    if (st != KRON_OK || batch == 0) return st;
    return launch_kron(a);
}

template <typename T>
static kron_status vbatched_impl(int d, const int *m, const int *n, const T *const *Aarray,
                                 const T *const *pX, T *const *pY, T *W, size_t Wcap,
                                 int batch, kron_stream_t stream, bool plain = false)
{
    if (d < 1 || d > kMaxDim || batch < 0 || !m || !n) return KRON_EINVAL;
    int mm[kMaxDim], nn[kMaxDim];
    for (int k = 0; k < kMaxDim; ++k) {
        mm[k] = (k < d) ? m[k] : 1;
        nn[k] = (k < d) ? n[k] : 1;
        if (mm[k] < 1 || nn[k] < 1) return KRON_EINVAL;
    }
    // kronmult_vbatched.hpp:136-157
    long long sizeX = 1;
    for (int k = 0; k < kMaxDim; ++k) sizeX *= nn[k];
    long long sizeW = 0;
    for (int k = 0; k < kMaxDim - 1; ++k) {
        long long s = 1;
        for (int l = 0; l < kMaxDim; ++l) s *= (l <= k) ? mm[l] : nn[l];
        if (s > sizeW) sizeW = s;
    }
    long long sizeY = 1;
    for (int k = 0; k < kMaxDim; ++k) sizeY *= mm[k];
    if (sizeX > 0x7fffffffLL || sizeW > 0x7fffffffLL || sizeY > 0x7fffffffLL)
        return KRON_EINVAL;
    const long long sizeXW = sizeX > sizeW ? sizeX : sizeW;
    const size_t slot_bytes = 2 * (size_t)sizeXW * sizeof(T);
    if (batch == 0) return KRON_OK;
    if (!Aarray || !pX || !pY) return KRON_EINVAL;
    // The reference clamps subbatch to 1 and overruns W when Wcap is too small.
    if (Wcap < slot_bytes) return KRON_EWORKSPACE;
    if (!W) return KRON_EINVAL;
    long long subbatch = (long long)(Wcap / slot_bytes);
    if (subbatch > batch) subbatch = batch;
    if (subbatch < 1) subbatch = 1;
    if (plain)
        return launch_plain<T>(d, mm, nn, Aarray, pX, pY, W, sizeXW, (int)subbatch, batch,
                               static_cast<hipStream_t>(stream));
    KronArgs<T> a;
    a.d = d;
    for (int k = 0; k < kMaxDim; ++k) {
        a.m[k] = mm[k];
        a.n[k] = nn[k];
    }
    a.lda = 0;  // ld_k = m_k (kronmult_vbatched.hpp:206-211)
    a.layout = Layout::Ptr;
    a.pA = Aarray;
    a.pX = pX;
    a.pY = pY;
    a.Wbase = W;
    a.slot_stride = 2 * (size_t)sizeXW;
    a.nslots = (int)subbatch;
    a.scratch = Scratch::Linear;
    a.y_shared = true;
    a.batch = batch;
    a.stream = static_cast<hipStream_t>(stream);
    return launch_kron(a);
}

template <typename T>
static kron_status adjoint_impl(int d, int n, const T *A, const T *X, const T *dY, T *dX, T *dA,
                                T *W, int batch, kron_stream_t stream)
{
    long long E = 0;
    if (d < 1 || d > kMaxDim || n < 1 || batch < 0 || !ipow_ok(n, d, E)) return KRON_EINVAL;
    if (batch == 0) return KRON_OK;
    if (!A || !X || !dY || !dX || !dA) return KRON_EINVAL;
    const size_t need = adjoint_scratch_elems<T>(d, n);
    if (need == (size_t)-1) return KRON_ENOTSUP;
    if (need > 0 && !W) return KRON_EWORKSPACE;
    AdjointArgs<T> a;
    a.layout = Layout::Batched;
    a.d = d;
    a.n = n;
    a.batch = batch;
    a.A = A;
    a.X = X;
    a.dY = dY;
    a.dX = dX;
    a.dA = dA;
    a.W = W;
    a.stream = static_cast<hipStream_t>(stream);
    return launch_adjoint<T>(a);
}

template <typename T>
static kron_status xadjoint_impl(int d, int n, const T *const *pA, int lda, const T *const *pX,
                                 const T *const *pdY, T *const *pdX, T *const *pdA, int ldda,
                                 T *W, int batch, kron_stream_t stream)
{
    long long E = 0;
    if (d < 1 || d > kMaxDim || n < 1 || lda < n || ldda < n || batch < 0 || !ipow_ok(n, d, E))
        return KRON_EINVAL;
    if (batch == 0) return KRON_OK;
    if (!pA || !pX || !pdY || !pdX || !pdA) return KRON_EINVAL;
    const size_t need = adjoint_scratch_elems<T>(d, n);
    if (need == (size_t)-1) return KRON_ENOTSUP;
    if (need > 0 && !W) return KRON_EWORKSPACE;
    AdjointArgs<T> a;
    a.layout = Layout::Ptr;
    a.d = d;
    a.n = n;
    a.batch = batch;
    a.pA = pA;
    a.lda = lda;
    a.pX = pX;
    a.pdY = pdY;
    a.pdX = pdX;
    a.pdA = pdA;
    a.ldda = ldda;
    a.W = W;
    a.stream = static_cast<hipStream_t>(stream);
    return launch_adjoint<T>(a);
}

template <typename T>
static kron_status kgemm_impl(bool transB, int mm, int nn, int kk, T alpha,
                              const T *const *A, const int *ldA, const T *const *B,
                              const int *ldB, T beta, T *const *C, const int *ldC, int batch,
                              kron_stream_t stream)
{
    if (mm < 0 || nn < 0 || kk < 0 || batch < 0) return KRON_EINVAL;
    if (batch == 0 || mm == 0 || nn == 0) return KRON_OK;
    if (!A || !ldA || !B || !ldB || !C || !ldC) return KRON_EINVAL;
    return launch_kgemm<T>(transB, mm, nn, kk, alpha, A, ldA, B, ldB, beta, C, ldC, batch,
                           static_cast<hipStream_t>(stream));
}

}  // namespace kron

using namespace kron;

// An xbatched plan: the call's arguments plus its Y-target order, computed
// once on the device and reused by every execute.
struct kron_xplan_s {
    int elem = 0;  // 8 or 4
    KronArgs<double> a64;
    KronArgs<float> a32;
    void *ws = nullptr;
    GroupOrder ord{nullptr, nullptr, 0};
    bool grouped = false;
    int device = 0;
};

template <typename T>
static KronArgs<T> &plan_args(kron_xplan_s *p);
template <>
KronArgs<double> &plan_args<double>(kron_xplan_s *p) { return p->a64; }
template <>
KronArgs<float> &plan_args<float>(kron_xplan_s *p) { return p->a32; }

template <typename T>
static kron_status plan_create(kron_xplan *out, int d, int n, const T *const *Aarray, int lda,
                               const T *const *pX, T *const *pY, T *const *pW, int batch,
                               int subbatch, kron_stream_t stream)
{
    if (!out) return KRON_EINVAL;
    *out = nullptr;
    kron_xplan_s *p = new kron_xplan_s();
    p->elem = (int)sizeof(T);
    KronArgs<T> &a = plan_args<T>(p);
    kron_status st = xbatched_args(a, d, n, Aarray, lda, pX, pY, pW, batch, subbatch, stream);
    if (st != KRON_OK) {
        delete p;
        return st;
    }
    (void)hipGetDevice(&p->device);
    if (batch > 0 && kron_groupable(a)) {
        const size_t bytes = group_workspace_bytes(batch);
        hipError_t e = hipMalloc(&p->ws, bytes);
        if (e != hipSuccess) {
            set_last_hip_error((int)e);
            delete p;
            return KRON_EHIP;
        }
        st = group_by_target(reinterpret_cast<const void *const *>(pY), batch, p->ws, bytes,
                             true, &p->ord.keys, &p->ord.perm, static_cast<hipStream_t>(stream));
        if (st != KRON_OK) {
            (void)hipFree(p->ws);
            delete p;
            return st;
        }
        p->grouped = true;
    }
    *out = p;
    return KRON_OK;
}

template <typename T>
static kron_status plan_execute(kron_xplan p, kron_stream_t stream)
{
    KronArgs<T> a = plan_args<T>(p);
    if (a.batch == 0) return KRON_OK;
    a.stream = static_cast<hipStream_t>(stream);
    a.order = p->grouped ? &p->ord : nullptr;
    return launch_kron(a);
}

extern "C" {

const char *kron_status_string(kron_status s)
{
    switch (s) {
    case KRON_OK: return "KRON_OK";
    case KRON_EINVAL: return "KRON_EINVAL: invalid argument";
    case KRON_EWORKSPACE: return "KRON_EWORKSPACE: workspace too small or missing";
    case KRON_EHIP: return "KRON_EHIP: HIP runtime error";
    case KRON_ENOTSUP: return "KRON_ENOTSUP: shape not supported";
    case KRON_ETOOBIG:
        return "KRON_ETOOBIG: batchCount above the no-allocation entry point's limit "
               "(KRON_NOALLOC_MAX_BATCH)";
    }
    return "unknown kron_status";
}

int kron_last_hip_error(void) { return g_last_hip_error; }
int kron_version(void) { return 100; }

kron_status kron_batched_f64(int d, int n, const double *A, const double *X, double *Y,
                             double *W, int batch, kron_stream_t s)
{
    return batched_impl<double>(d, n, A, X, Y, W, batch, s);
}
kron_status kron_batched_f32(int d, int n, const float *A, const float *X, float *Y, float *W,
                             int batch, kron_stream_t s)
{
    return batched_impl<float>(d, n, A, X, Y, W, batch, s);
}
kron_status kron_batched_adjoint_f64(int d, int n, const double *A, const double *X,
                                     const double *dY, double *dX, double *dA, double *W,
                                     int batch, kron_stream_t s)
{
    return adjoint_impl<double>(d, n, A, X, dY, dX, dA, W, batch, s);
}
kron_status kron_batched_adjoint_workspace(int d, int n, int elem_bytes, size_t *elems)
{
    long long E = 0;
    if (!elems || d < 1 || d > kMaxDim || n < 1 || !ipow_ok(n, d, E) ||
        (elem_bytes != 8 && elem_bytes != 4))
        return KRON_EINVAL;
    const size_t need = elem_bytes == 8 ? adjoint_scratch_elems<double>(d, n)
                                        : adjoint_scratch_elems<float>(d, n);
    if (need == (size_t)-1) return KRON_ENOTSUP;
    *elems = need;
    return KRON_OK;
}
kron_status kron_batched_adjoint_f32(int d, int n, const float *A, const float *X,
                                     const float *dY, float *dX, float *dA, float *W, int batch,
                                     kron_stream_t s)
{
    return adjoint_impl<float>(d, n, A, X, dY, dX, dA, W, batch, s);
}
kron_status kron_xbatched_adjoint_f64(int d, int n, const double *const *Aarray, int lda,
                                      const double *const *pX, const double *const *pdY,
                                      double *const *pdX, double *const *pdA, int ldda,
                                      double *W, int batch, kron_stream_t s)
{
    return xadjoint_impl<double>(d, n, Aarray, lda, pX, pdY, pdX, pdA, ldda, W, batch, s);
}
kron_status kron_xbatched_adjoint_f32(int d, int n, const float *const *Aarray, int lda,
                                      const float *const *pX, const float *const *pdY,
                                      float *const *pdX, float *const *pdA, int ldda, float *W,
                                      int batch, kron_stream_t s)
{
    return xadjoint_impl<float>(d, n, Aarray, lda, pX, pdY, pdX, pdA, ldda, W, batch, s);
}
kron_status kron_xbatched_f64(int d, int n, const double *const *A, int lda,
                              const double *const *pX, double *const *pY, double *const *pW,
                              int batch, int subbatch, kron_stream_t s)
{
    return xbatched_impl<double>(d, n, A, lda, pX, pY, pW, batch, subbatch, s);
}
kron_status kron_xbatched_f32(int d, int n, const float *const *A, int lda,
                              const float *const *pX, float *const *pY, float *const *pW,
                              int batch, int subbatch, kron_stream_t s)
{
    return xbatched_impl<float>(d, n, A, lda, pX, pY, pW, batch, subbatch, s);
}
kron_status kron_vbatched_f64(int d, const int *m, const int *n, const double *const *A,
                              const double *const *pX, double *const *pY, double *W,
                              size_t Wcap, int batch, kron_stream_t s)
{
    return vbatched_impl<double>(d, m, n, A, pX, pY, W, Wcap, batch, s);
}
kron_status kron_vbatched_f32(int d, const int *m, const int *n, const float *const *A,
                              const float *const *pX, float *const *pY, float *W,
                              size_t Wcap, int batch, kron_stream_t s)
{
    return vbatched_impl<float>(d, m, n, A, pX, pY, W, Wcap, batch, s);
}

kron_status kron_pbatched_f64(int d, int n, const double *A, const double *const *pX,
                              double *const *pY, double *const *pW, int batch, kron_stream_t s)
{
    return pbatched_impl<double>(d, n, A, pX, pY, pW, batch, s);
}
kron_status kron_pbatched_f32(int d, int n, const float *A, const float *const *pX,
                              float *const *pY, float *const *pW, int batch, kron_stream_t s)
{
    return pbatched_impl<float>(d, n, A, pX, pY, pW, batch, s);
}

kron_status kron_vbatched_plain_f64(int d, const int *m, const int *n, const double *const *A,
                                   const double *const *pX, double *const *pY, double *W,
                                   size_t Wcap, int batch, kron_stream_t s)
{
    return vbatched_impl<double>(d, m, n, A, pX, pY, W, Wcap, batch, s, true);
}
kron_status kron_vbatched_plain_f32(int d, const int *m, const int *n, const float *const *A,
                                   const float *const *pX, float *const *pY, float *W,
                                   size_t Wcap, int batch, kron_stream_t s)
{
    return vbatched_impl<float>(d, m, n, A, pX, pY, W, Wcap, batch, s, true);
}

kron_status kron_xbatched_plan_create_f64(kron_xplan *plan, int d, int n,
                                          const double *const *A, int lda,
                                          const double *const *pX, double *const *pY,
                                          double *const *pW, int batch, int subbatch,
                                          kron_stream_t s)
{
    return plan_create<double>(plan, d, n, A, lda, pX, pY, pW, batch, subbatch, s);
}
kron_status kron_xbatched_plan_create_f32(kron_xplan *plan, int d, int n,
                                          const float *const *A, int lda,
                                          const float *const *pX, float *const *pY,
                                          float *const *pW, int batch, int subbatch,
                                          kron_stream_t s)
{
    return plan_create<float>(plan, d, n, A, lda, pX, pY, pW, batch, subbatch, s);
}
kron_status kron_xbatched_plan_execute(kron_xplan plan, kron_stream_t s)
{
    if (!plan) return KRON_EINVAL;
    return plan->elem == 8 ? plan_execute<double>(plan, s) : plan_execute<float>(plan, s);
}
kron_status kron_xbatched_plan_destroy(kron_xplan plan)
{
    if (!plan) return KRON_OK;
    if (plan->ws) {
        int cur = 0;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(plan->device);
        (void)hipDeviceSynchronize();
        (void)hipFree(plan->ws);
        (void)hipSetDevice(cur);
    }
    delete plan;
    return KRON_OK;
}

struct kron_mplan_s {
    MixedPlanCore *core = nullptr;
};

kron_status kron_mixed_plan_create_f64(kron_mplan *plan, int nclass,
                                       const kron_mixed_class_f64 *classes,
                                       const char *tune_cache, int flags, kron_stream_t stream)
{
    if (!plan) return KRON_EINVAL;
    *plan = nullptr;
    MixedPlanCore *core = nullptr;
    kron_status st = mixed_create_f64(core, nclass, classes, tune_cache, flags,
                                      static_cast<hipStream_t>(stream));
    if (st != KRON_OK) return st;
    *plan = new kron_mplan_s{core};
    return KRON_OK;
}
kron_status kron_mixed_plan_create_f32(kron_mplan *plan, int nclass,
                                       const kron_mixed_class_f32 *classes,
                                       const char *tune_cache, int flags, kron_stream_t stream)
{
    if (!plan) return KRON_EINVAL;
    *plan = nullptr;
    MixedPlanCore *core = nullptr;
    kron_status st = mixed_create_f32(core, nclass, classes, tune_cache, flags,
                                      static_cast<hipStream_t>(stream));
    if (st != KRON_OK) return st;
    *plan = new kron_mplan_s{core};
    return KRON_OK;
}
kron_status kron_mixed_plan_execute(kron_mplan plan, kron_stream_t stream)
{
    if (!plan) return KRON_EINVAL;
    return mixed_execute(plan->core, static_cast<hipStream_t>(stream));
}
kron_status kron_mixed_plan_capture(kron_mplan plan, int applies)
{
    if (!plan) return KRON_EINVAL;
    return mixed_capture(plan->core, applies);
}
kron_status kron_mixed_plan_replay(kron_mplan plan, kron_stream_t stream)
{
    if (!plan) return KRON_EINVAL;
    return mixed_replay(plan->core, static_cast<hipStream_t>(stream));
}
kron_status kron_mixed_plan_class_info(kron_mplan plan, int cls, int *kind, int *variant,
                                       int *atomic, double *cost)
{
    if (!plan) return KRON_EINVAL;
    return mixed_class_info(plan->core, cls, kind, variant, atomic, cost);
}
kron_status kron_mixed_plan_launch_info(kron_mplan plan, int *items, int *grid, int *lds_bytes,
                                        int *launches)
{
    if (!plan) return KRON_EINVAL;
    return mixed_launch_info(plan->core, items, grid, lds_bytes, launches);
}
kron_status kron_mixed_plan_destroy(kron_mplan plan)
{
    if (!plan) return KRON_OK;
    mixed_destroy(plan->core);
    delete plan;
    return KRON_OK;
}

#define KGEMM_C(NAME, T, TRANS)                                                            \
    kron_status NAME(int mm, int nn, int kk, T alpha, const T *const *A, const int *ldA,   \
                     const T *const *B, const int *ldB, T beta, T *const *C,               \
                     const int *ldC, int batch, kron_stream_t s)                           \
    {                                                                                      \
        return kgemm_impl<T>(TRANS, mm, nn, kk, alpha, A, ldA, B, ldB, beta, C, ldC,       \
                             batch, s);                                                    \
    }
KGEMM_C(kron_kgemm_nn_batched_f64, double, false)
KGEMM_C(kron_kgemm_nt_batched_f64, double, true)
KGEMM_C(kron_kgemm_nn_batched_f32, float, false)
KGEMM_C(kron_kgemm_nt_batched_f32, float, true)
#undef KGEMM_C

}  // extern "C"

// ---------------------------------------------------------------------------
// Drop-in C++ overloads (null stream), reference signatures.
// ---------------------------------------------------------------------------
static thread_local kron_status g_last_status = KRON_OK;
kron_status kronmult_last_status() { return g_last_status; }

static void report(const char *fn, kron_status st)
{
    g_last_status = st;
    if (st != KRON_OK) std::fprintf(stderr, "kronmult_amd: %s: %s\n", fn, kron_status_string(st));
}

#define DEF_BATCHED(D, T, SFX)                                                             \
    void kronmult##D##_batched(int const n, T const Aarray_[], T Xarray_[], T Yarray_[],  \
                               T Warray_[], int const batchCount)                          \
    {                                                                                      \
        report("kronmult" #D "_batched",                                                   \
               kron_batched_##SFX(D, n, Aarray_, Xarray_, Yarray_, Warray_, batchCount,    \
                                  nullptr));                                               \
    }
#define DEF_XBATCHED(D, T, SFX)                                                            \
    void kronmult##D##_xbatched(int const n, T const *const Aarray_[], int const lda,     \
                                T *Xarray_[], T *Yarray_[], T *Warray_[],                  \
                                int const batchCount, int const subbatchCount)             \
    {                                                                                      \
        report("kronmult" #D "_xbatched",                                                  \
               kron_xbatched_##SFX(D, n, Aarray_, lda, Xarray_, Yarray_, Warray_,          \
                                   batchCount, subbatchCount, nullptr));                   \
    }
#define DEF_PBATCHED(D, T, SFX)                                                            \
    void kronmult##D##_pbatched(int const n, T const Aarray_[], T *Xarray_[],             \
                                T *Yarray_[], T *Warray_[], int const batchCount)          \
    {                                                                                      \
        report("kronmult" #D "_pbatched",                                                  \
               kron_pbatched_##SFX(D, n, Aarray_, Xarray_, Yarray_, Warray_, batchCount,   \
                                   nullptr));                                              \
    }
#define DEF_ALL(T, SFX)                                                                    \
    DEF_PBATCHED(1, T, SFX) DEF_PBATCHED(2, T, SFX) DEF_PBATCHED(3, T, SFX)                \
    DEF_PBATCHED(4, T, SFX) DEF_PBATCHED(5, T, SFX) DEF_PBATCHED(6, T, SFX)                \
    DEF_BATCHED(1, T, SFX) DEF_BATCHED(2, T, SFX) DEF_BATCHED(3, T, SFX)                   \
    DEF_BATCHED(4, T, SFX) DEF_BATCHED(5, T, SFX) DEF_BATCHED(6, T, SFX)                   \
    DEF_XBATCHED(1, T, SFX) DEF_XBATCHED(2, T, SFX) DEF_XBATCHED(3, T, SFX)                \
    DEF_XBATCHED(4, T, SFX) DEF_XBATCHED(5, T, SFX) DEF_XBATCHED(6, T, SFX)
DEF_ALL(double, f64)
DEF_ALL(float, f32)

#define DEF_VBATCHED(T, SFX)                                                               \
    void kronmult1_vbatched(int const m1, int const n1, KRONMULT_VB_TAIL(T))               \
    {                                                                                      \
        const int m[1] = {m1}, n[1] = {n1};                                                \
        report("kronmult1_vbatched", kron_vbatched_##SFX(1, m, n, Aarray_, pX_, pY_, W_,   \
                                                         Wcapcity, batchCount, nullptr));  \
    }                                                                                      \
    void kronmult2_vbatched(int const m1, int const n1, int const m2, int const n2,        \
                            KRONMULT_VB_TAIL(T))                                           \
    {                                                                                      \
        const int m[2] = {m1, m2}, n[2] = {n1, n2};                                        \
        report("kronmult2_vbatched", kron_vbatched_##SFX(2, m, n, Aarray_, pX_, pY_, W_,   \
                                                         Wcapcity, batchCount, nullptr));  \
    }                                                                                      \
    void kronmult3_vbatched(int const m1, int const n1, int const m2, int const n2,        \
                            int const m3, int const n3, KRONMULT_VB_TAIL(T))               \
    {                                                                                      \
        const int m[3] = {m1, m2, m3}, n[3] = {n1, n2, n3};                                \
        report("kronmult3_vbatched", kron_vbatched_##SFX(3, m, n, Aarray_, pX_, pY_, W_,   \
                                                         Wcapcity, batchCount, nullptr));  \
    }                                                                                      \
    void kronmult4_vbatched(int const m1, int const n1, int const m2, int const n2,        \
                            int const m3, int const n3, int const m4, int const n4,        \
                            KRONMULT_VB_TAIL(T))                                           \
    {                                                                                      \
        const int m[4] = {m1, m2, m3, m4}, n[4] = {n1, n2, n3, n4};                        \
        report("kronmult4_vbatched", kron_vbatched_##SFX(4, m, n, Aarray_, pX_, pY_, W_,   \
                                                         Wcapcity, batchCount, nullptr));  \
    }                                                                                      \
    void kronmult5_vbatched(int const m1, int const n1, int const m2, int const n2,        \
                            int const m3, int const n3, int const m4, int const n4,        \
                            int const m5, int const n5, KRONMULT_VB_TAIL(T))               \
    {                                                                                      \
        const int m[5] = {m1, m2, m3, m4, m5}, n[5] = {n1, n2, n3, n4, n5};                \
        report("kronmult5_vbatched", kron_vbatched_##SFX(5, m, n, Aarray_, pX_, pY_, W_,   \
                                                         Wcapcity, batchCount, nullptr));  \
    }                                                                                      \
    void kronmult6_vbatched(int const m1, int const n1, int const m2, int const n2,        \
                            int const m3, int const n3, int const m4, int const n4,        \
                            int const m5, int const n5, int const m6, int const n6,        \
                            KRONMULT_VB_TAIL(T))                                           \
    {                                                                                      \
        const int m[6] = {m1, m2, m3, m4, m5, m6}, n[6] = {n1, n2, n3, n4, n5, n6};        \
        report("kronmult6_vbatched", kron_vbatched_##SFX(6, m, n, Aarray_, pX_, pY_, W_,   \
                                                         Wcapcity, batchCount, nullptr));  \
    }
DEF_VBATCHED(double, f64)
DEF_VBATCHED(float, f32)

#define DEF_KGEMM(NAME, T, CNAME)                                                          \
    void NAME(int const mm, int const nn, int const kk, T const alpha, T *const Aarray_[], \
              int const ldAarray_[], T *const Barray_[], int const ldBarray_[],            \
              T const beta, T *const Carray_[], int const ldCarray_[],                     \
              int const batchCount)                                                        \
    {                                                                                      \
        report(#NAME, CNAME(mm, nn, kk, alpha, Aarray_, ldAarray_, Barray_, ldBarray_,     \
                            beta, Carray_, ldCarray_, batchCount, nullptr));               \
    }
DEF_KGEMM(kgemm_nn_batched, double, kron_kgemm_nn_batched_f64)
DEF_KGEMM(kgemm_nt_batched, double, kron_kgemm_nt_batched_f64)
DEF_KGEMM(kgemm_nn_batched, float, kron_kgemm_nn_batched_f32)
DEF_KGEMM(kgemm_nt_batched, float, kron_kgemm_nt_batched_f32)
