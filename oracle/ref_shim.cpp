// ref_shim.cpp -- TEST INFRASTRUCTURE ONLY.
//
// extern "C" entry points over the reference kronmult CPU/OpenMP path,
// compiled from the reference sources where they lie (oracle/build.py) into
// oracle/_ref/libkron_ref.so.  Used in this container to validate the C
// restatement (kron_oracle.c) and, where present, as the CPU baseline.
//
// The double-precision calls go through the reference's host wrappers
// (kronmultD_batched.cpp:5-25, kronmultD_xbatched.cpp:5-39,
// kgemm_nn_batched.cpp:9-42, kgemm_nt_batched.cpp:8-41).  Float calls and
// vbatched (which has no host wrapper) call the reference templates the way
// its own test drivers do on the CPU (test_kronmult6_batched.cpp:262-305,
// test_kronmult6_vbatched.cpp CPU branch).
#include "kronmult1_batched.hpp"
#include "kronmult2_batched.hpp"
#include "kronmult3_batched.hpp"
#include "kronmult4_batched.hpp"
#include "kronmult5_batched.hpp"
#include "kronmult6_batched.hpp"
#include "kronmult1_xbatched.hpp"
#include "kronmult2_xbatched.hpp"
#include "kronmult3_xbatched.hpp"
#include "kronmult4_xbatched.hpp"
#include "kronmult5_xbatched.hpp"
#include "kronmult6_xbatched.hpp"
#include "kronmult1_vbatched.hpp"
#include "kronmult2_vbatched.hpp"
#include "kronmult3_vbatched.hpp"
#include "kronmult4_vbatched.hpp"
#include "kronmult5_vbatched.hpp"
#include "kronmult6_vbatched.hpp"
#include "kgemm_nn_batched.hpp"
#include "kgemm_nt_batched.hpp"

#include <omp.h>

// Host wrappers defined in the reference .cpp files (C++ linkage).
#define DECL_BATCHED(D)                                                      \
    void kronmult##D##_batched(int const n, double const Aarray_[],         \
                               double Xarray_[], double Yarray_[],          \
                               double Warray_[], int const batchCount);
#define DECL_XBATCHED(D)                                                     \
    void kronmult##D##_xbatched(int const n, double const *const Aarray_[], \
                                int const lda, double *Xarray_[],           \
                                double *Yarray_[], double *Warray_[],        \
                                int const batchCount, int const subbatchCount);
#define DECL_PBATCHED(D)                                                     \
    void kronmult##D##_pbatched(int const n, double const Aarray_[],        \
                                double *Xarray_[], double *Yarray_[],        \
                                double *Warray_[], int const batchCount);
DECL_PBATCHED(1) DECL_PBATCHED(2) DECL_PBATCHED(3)
DECL_PBATCHED(5) DECL_PBATCHED(6)
DECL_BATCHED(1) DECL_BATCHED(2) DECL_BATCHED(3)
DECL_BATCHED(4) DECL_BATCHED(5) DECL_BATCHED(6)
DECL_XBATCHED(1) DECL_XBATCHED(2) DECL_XBATCHED(3)
DECL_XBATCHED(4) DECL_XBATCHED(5) DECL_XBATCHED(6)
void kgemm_nn_batched(int const mm, int const nn, int const kk, double const alpha,
                      double *const Aarray_[], int const ldAarray_[],
                      double *const Barray_[], int const ldBarray_[],
                      double const beta, double *const Carray_[],
                      int const ldCarray_[], int const batchCount);
void kgemm_nt_batched(int const mm, int const nn, int const kk, double const alpha,
                      double *const Aarray_[], int const ldAarray_[],
                      double *const Barray_[], int const ldBarray_[],
                      double const beta, double *const Carray_[],
                      int const ldCarray_[], int const batchCount);

#define VB_ARGS(T)                                                           \
    m[0], n[0], m[1], n[1], m[2], n[2], m[3], n[3], m[4], n[4], m[5], n[5]

template <typename T>
static int ref_vbatched(int d, const int *m_in, const int *n_in,
                        const T *const *A, T **pX, T **pY, T *W, size_t Wcap,
                        int batch)
{
    int m[6], n[6];
    for (int k = 0; k < 6; k++) {
        m[k] = (k < d) ? m_in[k] : 1;
        n[k] = (k < d) ? n_in[k] : 1;
    }
    switch (d) {
    case 1: kronmult1_vbatched<T>(m[0], n[0], A, pX, pY, W, Wcap, batch); return 0;
    case 2: kronmult2_vbatched<T>(m[0], n[0], m[1], n[1], A, pX, pY, W, Wcap, batch); return 0;
    case 3: kronmult3_vbatched<T>(m[0], n[0], m[1], n[1], m[2], n[2], A, pX, pY, W, Wcap, batch); return 0;
    case 4: kronmult4_vbatched<T>(m[0], n[0], m[1], n[1], m[2], n[2], m[3], n[3], A, pX, pY, W, Wcap, batch); return 0;
    case 5: kronmult5_vbatched<T>(m[0], n[0], m[1], n[1], m[2], n[2], m[3], n[3], m[4], n[4], A, pX, pY, W, Wcap, batch); return 0;
    case 6: kronmult6_vbatched<T>(VB_ARGS(T), A, pX, pY, W, Wcap, batch); return 0;
    }
    return 1;
}

extern "C" {

int ref_batched_f64(int d, int n, const double *A, double *X, double *Y,
                    double *W, int batch)
{
    switch (d) {
    case 1: kronmult1_batched(n, A, X, Y, W, batch); return 0;
    case 2: kronmult2_batched(n, A, X, Y, W, batch); return 0;
    case 3: kronmult3_batched(n, A, X, Y, W, batch); return 0;
    case 4: kronmult4_batched(n, A, X, Y, W, batch); return 0;
    case 5: kronmult5_batched(n, A, X, Y, W, batch); return 0;
    case 6: kronmult6_batched(n, A, X, Y, W, batch); return 0;
    }
    return 1;
}

int ref_batched_f32(int d, int n, const float *A, float *X, float *Y, float *W,
                    int batch)
{
    switch (d) {
    case 1: kronmult1_batched<float>(n, A, X, Y, W, batch); return 0;
    case 2: kronmult2_batched<float>(n, A, X, Y, W, batch); return 0;
    case 3: kronmult3_batched<float>(n, A, X, Y, W, batch); return 0;
    case 4: kronmult4_batched<float>(n, A, X, Y, W, batch); return 0;
    case 5: kronmult5_batched<float>(n, A, X, Y, W, batch); return 0;
    case 6: kronmult6_batched<float>(n, A, X, Y, W, batch); return 0;
    }
    return 1;
}

// kronmult{1..6}_pbatched (kronmult6_pbatched.cpp:5-26): double only in the
// reference's host wrappers.  kronmult4_pbatched.cpp does not compile as
// shipped (line 3), so d = 4 returns 1 (unavailable).
int ref_pbatched_f64(int d, int n, const double *A, double **pX, double **pY, double **pW,
                     int batch)
{
    switch (d) {
    case 1: kronmult1_pbatched(n, A, pX, pY, pW, batch); return 0;
    case 2: kronmult2_pbatched(n, A, pX, pY, pW, batch); return 0;
    case 3: kronmult3_pbatched(n, A, pX, pY, pW, batch); return 0;
    case 5: kronmult5_pbatched(n, A, pX, pY, pW, batch); return 0;
    case 6: kronmult6_pbatched(n, A, pX, pY, pW, batch); return 0;
    }
    return 1;
}

int ref_xbatched_f64(int d, int n, const double *const *A, int lda, double **pX,
                     double **pY, double **pW, int batch, int subbatch)
{
    switch (d) {
    case 1: kronmult1_xbatched(n, A, lda, pX, pY, pW, batch, subbatch); return 0;
    case 2: kronmult2_xbatched(n, A, lda, pX, pY, pW, batch, subbatch); return 0;
    case 3: kronmult3_xbatched(n, A, lda, pX, pY, pW, batch, subbatch); return 0;
    case 4: kronmult4_xbatched(n, A, lda, pX, pY, pW, batch, subbatch); return 0;
    case 5: kronmult5_xbatched(n, A, lda, pX, pY, pW, batch, subbatch); return 0;
    case 6: kronmult6_xbatched(n, A, lda, pX, pY, pW, batch, subbatch); return 0;
    }
    return 1;
}

int ref_xbatched_f32(int d, int n, const float *const *A, int lda, float **pX,
                     float **pY, float **pW, int batch, int subbatch)
{
    switch (d) {
    case 1: kronmult1_xbatched<float>(n, A, lda, pX, pY, pW, batch, subbatch); return 0;
    case 2: kronmult2_xbatched<float>(n, A, lda, pX, pY, pW, batch, subbatch); return 0;
    case 3: kronmult3_xbatched<float>(n, A, lda, pX, pY, pW, batch, subbatch); return 0;
    case 4: kronmult4_xbatched<float>(n, A, lda, pX, pY, pW, batch, subbatch); return 0;
    case 5: kronmult5_xbatched<float>(n, A, lda, pX, pY, pW, batch, subbatch); return 0;
    case 6: kronmult6_xbatched<float>(n, A, lda, pX, pY, pW, batch, subbatch); return 0;
    }
    return 1;
}

int ref_vbatched_f64(int d, const int *m, const int *n, const double *const *A,
                     double **pX, double **pY, double *W, size_t Wcap, int batch)
{
    return ref_vbatched<double>(d, m, n, A, pX, pY, W, Wcap, batch);
}

int ref_vbatched_f32(int d, const int *m, const int *n, const float *const *A,
                     float **pX, float **pY, float *W, size_t Wcap, int batch)
{
    return ref_vbatched<float>(d, m, n, A, pX, pY, W, Wcap, batch);
}

int ref_kgemm_batched_f64(char trans, int mm, int nn, int kk, double alpha,
                          double **A, const int *ldA, double **B, const int *ldB,
                          double beta, double **C, const int *ldC, int batch)
{
    if (trans == 'T' || trans == 't')
        kgemm_nt_batched(mm, nn, kk, alpha, A, ldA, B, ldB, beta, C, ldC, batch);
    else
        kgemm_nn_batched(mm, nn, kk, alpha, A, ldA, B, ldB, beta, C, ldC, batch);
    return 0;
}

int ref_kgemm_batched_f32(char trans, int mm, int nn, int kk, float alpha,
                          float **A, const int *ldA, float **B, const int *ldB,
                          float beta, float **C, const int *ldC, int batch)
{
    if (trans == 'T' || trans == 't')
        kgemm_nt_batched<float>(mm, nn, kk, alpha, A, ldA, B, ldB, beta, C, ldC, batch);
    else
        kgemm_nn_batched<float>(mm, nn, kk, alpha, A, ldA, B, ldB, beta, C, ldC, batch);
    return 0;
}

// Thread count of the reference's OpenMP loops (the CPU baseline is timed
// at all host cores and at one thread).
void ref_set_threads(int t) { omp_set_num_threads(t > 0 ? t : 1); }

}  // extern "C"
