// dropin_caller.cpp -- a C++ caller written against the reference's own entry
// points (include/kronmult.hpp declares them with the reference's signatures:
// kronmult6_batched.cpp:5-11, kronmult6_xbatched.cpp:5-13,
// kgemm_nn_batched.cpp:9-18, the kronmult3_vbatched kernel's parameter list
// kronmult3_vbatched.hpp), linked against libkronmult_amd.so.  It fills device
// memory allocated with hipMalloc with small integers (every sum is exact in
// FP64), calls each entry point on the null stream, synchronises as the
// reference's tests do, and writes every input and output to <dir>/<case>.bin
// for the test (tests/test_gpu_dropin.py) to check against the oracle.
//
// Usage: dropin_caller <dir>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "kronmult.hpp"

static uint64_t g_state = 0x243F6A8885A308D3ull;
static int next_int(int lo, int hi)
{
    g_state = g_state * 6364136223846793005ull + 1442695040888963407ull;
    return lo + (int)((g_state >> 33) % (uint64_t)(hi - lo + 1));
}
static std::vector<double> ints(size_t n, int lo, int hi)
{
    std::vector<double> v(n);
    for (auto &x : v) x = next_int(lo, hi);
    return v;
}

#define HIPCHECK(x)                                                                    \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            std::fprintf(stderr, "%s:%d %s\n", __FILE__, __LINE__, hipGetErrorString(e_)); \
            std::exit(2);                                                              \
        }                                                                              \
    } while (0)

template <typename T>
static T *to_dev(const std::vector<T> &h)
{
    T *d = nullptr;
    HIPCHECK(hipMalloc(&d, h.size() * sizeof(T) + 8));
    HIPCHECK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}
template <typename T>
static std::vector<T> to_host(const T *d, size_t n)
{
    std::vector<T> h(n);
    HIPCHECK(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost));
    return h;
}

// <dir>/<name>.bin: a sequence of arrays, each an int64 element count then
// that many doubles
struct Dump {
    FILE *f;
    Dump(const std::string &dir, const char *name)
    {
        f = std::fopen((dir + "/" + name + ".bin").c_str(), "wb");
        if (!f) std::exit(3);
    }
    ~Dump() { std::fclose(f); }
    void put(const std::vector<double> &v)
    {
        const int64_t n = (int64_t)v.size();
        std::fwrite(&n, 8, 1, f);
        std::fwrite(v.data(), 8, v.size(), f);
    }
    void put_idx(const std::vector<int> &v)
    {
        std::vector<double> d(v.begin(), v.end());
        put(d);
    }
};

static void check_status(const char *what)
{
    const kron_status st = kronmult_last_status();
    if (st != KRON_OK) {
        std::fprintf(stderr, "%s: %s\n", what, kron_status_string(st));
        std::exit(4);
    }
}

// kronmult6_batched: d = 6, n = 4, private Y per entry
static void case_batched(const std::string &dir)
{
    const int n = 4, d = 6, batch = 37, E = 4096;
    auto A = ints((size_t)n * n * d * batch, -2, 2);
    auto X = ints((size_t)E * batch, -8, 8);
    auto Y0 = ints((size_t)E * batch, -9, 9);
    double *dA = to_dev(A), *dX = to_dev(X), *dY = to_dev(Y0);
    double *dW = to_dev(std::vector<double>((size_t)E * batch, 0.0));
    kronmult6_batched(n, dA, dX, dY, dW, batch);
    HIPCHECK(hipDeviceSynchronize());
    check_status("kronmult6_batched");
    Dump o(dir, "batched");
    o.put(A);
    o.put(X);
    o.put(Y0);
    o.put(to_host(dY, Y0.size()));
    for (double *p : {dA, dX, dY, dW}) HIPCHECK(hipFree(p));
}

// kronmult6_xbatched: n = 4, lda = 5 (padded factors), X and Y pools shared
// by many entries; 300 entries (per-entry atomics) and 4,096 (entries merged
// by Y target on the device)
static void case_xbatched(const std::string &dir, int batch, const char *name)
{
    const int n = 4, d = 6, lda = 5, E = 4096, nx = 50, ny = 7;
    auto Apool = ints((size_t)lda * n * d * batch, -2, 2);
    auto Xpool = ints((size_t)E * nx, -8, 8);
    auto Y0 = ints((size_t)E * ny, -9, 9);
    std::vector<int> xm(batch), ym(batch);
    for (int b = 0; b < batch; ++b) {
        xm[b] = next_int(0, nx - 1);
        ym[b] = next_int(0, ny - 1);
    }
    double *dA = to_dev(Apool), *dX = to_dev(Xpool), *dY = to_dev(Y0);
    const int nslot = 2 * batch;
    double *dW = to_dev(std::vector<double>((size_t)E * nslot, 0.0));
    std::vector<const double *> pA((size_t)d * batch);
    std::vector<double *> pX(batch), pY(batch), pW(nslot);
    for (int b = 0; b < batch; ++b) {
        for (int k = 0; k < d; ++k) pA[(size_t)b * d + k] = dA + ((size_t)b * d + k) * lda * n;
        pX[b] = dX + (size_t)xm[b] * E;
        pY[b] = dY + (size_t)ym[b] * E;
    }
    for (int s = 0; s < nslot; ++s) pW[s] = dW + (size_t)s * E;
    const double **dpA = to_dev(pA);
    double **dpX = to_dev(pX), **dpY = to_dev(pY), **dpW = to_dev(pW);
    kronmult6_xbatched(n, dpA, lda, dpX, dpY, dpW, batch);  // subbatchCount defaults to 0
    HIPCHECK(hipDeviceSynchronize());
    check_status("kronmult6_xbatched");
    Dump o(dir, name);
    o.put(Apool);
    o.put(Xpool);
    o.put_idx(xm);
    o.put(Y0);
    o.put_idx(ym);
    o.put(to_host(dY, Y0.size()));
    HIPCHECK(hipFree((void *)dpA));
    for (void *p : {(void *)dpX, (void *)dpY, (void *)dpW, (void *)dA, (void *)dX, (void *)dY,
                    (void *)dW})
        HIPCHECK(hipFree(p));
}

// kgemm_nn_batched: C_b += A_b B_b (alpha = beta = 1) with C shared by entries
static void case_kgemm(const std::string &dir)
{
    const int mm = 7, nn = 5, kk = 6, batch = 9, nc = 3, ldA = 8, ldB = 6, ldC = 9;
    auto Ap = ints((size_t)ldA * kk * batch, -3, 3);
    auto Bp = ints((size_t)ldB * nn * batch, -3, 3);
    auto C0 = ints((size_t)ldC * nn * nc, -5, 5);
    std::vector<int> cm(batch);
    for (int b = 0; b < batch; ++b) cm[b] = b % nc;
    double *dA = to_dev(Ap), *dB = to_dev(Bp), *dC = to_dev(C0);
    std::vector<double *> pA(batch), pB(batch), pC(batch);
    for (int b = 0; b < batch; ++b) {
        pA[b] = dA + (size_t)b * ldA * kk;
        pB[b] = dB + (size_t)b * ldB * nn;
        pC[b] = dC + (size_t)cm[b] * ldC * nn;
    }
    std::vector<int> lA(batch, ldA), lB(batch, ldB), lC(batch, ldC);
    double **dpA = to_dev(pA), **dpB = to_dev(pB), **dpC = to_dev(pC);
    int *dlA = to_dev(lA), *dlB = to_dev(lB), *dlC = to_dev(lC);
    kgemm_nn_batched(mm, nn, kk, 1.0, dpA, dlA, dpB, dlB, 1.0, dpC, dlC, batch);
    HIPCHECK(hipDeviceSynchronize());
    check_status("kgemm_nn_batched");
    Dump o(dir, "kgemm");
    o.put(Ap);
    o.put(Bp);
    o.put(C0);
    o.put_idx(cm);
    o.put(to_host(dC, C0.size()));
    for (void *p : {(void *)dpA, (void *)dpB, (void *)dpC, (void *)dlA, (void *)dlB, (void *)dlC,
                    (void *)dA, (void *)dB, (void *)dC})
        HIPCHECK(hipFree(p));
}

// kronmult3_vbatched (host overload with the kernel's parameter list):
// rectangular factors 3x4, 2x5, 4x2, Y shared by pairs of entries
static void case_vbatched(const std::string &dir)
{
    const int m[3] = {3, 2, 4}, nn[3] = {4, 5, 2}, batch = 10, ny = 5;
    const int sx = 4 * 5 * 2, sy = 3 * 2 * 4;
    // sizeW = max(m1 n2 n3, m1 m2 n3, m1 m2 m3) = max(30, 12, 24); slot = 2 * max(sx, sizeW)
    const size_t slot = 2 * 40;
    auto Apool = ints((size_t)3 * 12 * batch, -2, 2);  // factor k of entry b at (b*3+k)*12
    auto Xp = ints((size_t)sx * batch, -4, 4);
    auto Y0 = ints((size_t)sy * ny, -5, 5);
    double *dA = to_dev(Apool), *dX = to_dev(Xp), *dY = to_dev(Y0);
    double *dW = to_dev(std::vector<double>(slot * 4, 0.0));
    std::vector<const double *> pA((size_t)3 * batch);
    std::vector<double *> pX(batch), pY(batch);
    for (int b = 0; b < batch; ++b) {
        for (int k = 0; k < 3; ++k) pA[(size_t)b * 3 + k] = dA + ((size_t)b * 3 + k) * 12;
        pX[b] = dX + (size_t)b * sx;
        pY[b] = dY + (size_t)(b % ny) * sy;
    }
    const double **dpA = to_dev(pA);
    double **dpX = to_dev(pX), **dpY = to_dev(pY);
    kronmult3_vbatched(m[0], nn[0], m[1], nn[1], m[2], nn[2], dpA, dpX, dpY, dW,
                       slot * 4 * sizeof(double), batch);
    HIPCHECK(hipDeviceSynchronize());
    check_status("kronmult3_vbatched");
    Dump o(dir, "vbatched");
    o.put(Apool);
    o.put(Xp);
    o.put(Y0);
    o.put(to_host(dY, Y0.size()));
    HIPCHECK(hipFree((void *)dpA));
    for (void *p : {(void *)dpX, (void *)dpY, (void *)dA, (void *)dX, (void *)dY, (void *)dW})
        HIPCHECK(hipFree(p));
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <dir>\n", argv[0]);
        return 1;
    }
    const std::string dir = argv[1];
    case_batched(dir);
    case_xbatched(dir, 300, "xbatched_300");
    case_xbatched(dir, 4096, "xbatched_4096");
    case_kgemm(dir);
    case_vbatched(dir);
    std::printf("dropin_caller ok\n");
    return 0;
}
