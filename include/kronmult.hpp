// kronmult.hpp -- drop-in C++ entry points with the reference's exact
// signatures (C++ linkage), implemented over the C ABI in kronmult_amd.h on
// the null stream.  A caller of hipdac-lab/kronmult relinks against
// libkronmult_amd.so unchanged.
//
// Behaviour kept from the reference: asynchronous launch on the null stream,
// no return value, Y += with atomic adds wherever entries may share Y,
// subbatchCount == 0 -> max(1, batchCount/2).  Differences: X is never written
// (the reference clobbers it in the batched layouts), and argument errors are
// detected and reported on stderr (the reference does not check; device
// asserts are compiled out by -DNDEBUG, make.inc.hip_amd:3).  The status of
// the last call on this thread is kronmult_last_status().
#ifndef KRONMULT_DROPIN_HPP
#define KRONMULT_DROPIN_HPP

#include <cstddef>

#include "kronmult_amd.h"

// kronmult6_batched.cpp:5-11 (and d = 1..5 analogues)
#define KRONMULT_DECL_BATCHED(D, T)                                                   \
    void kronmult##D##_batched(int const n, T const Aarray_[], T Xarray_[],          \
                               T Yarray_[], T Warray_[], int const batchCount);
// kronmult6_xbatched.cpp:5-13
#define KRONMULT_DECL_XBATCHED(D, T)                                                  \
    void kronmult##D##_xbatched(int const n, T const *const Aarray_[], int const lda, \
                                T *Xarray_[], T *Yarray_[], T *Warray_[],             \
                                int const batchCount, int const subbatchCount = 0);

// kronmult6_pbatched.cpp:5-11: contiguous A (n,n,d,batch), pointer X/Y/W
#define KRONMULT_DECL_PBATCHED(D, T)                                                  \
    void kronmult##D##_pbatched(int const n, T const Aarray_[], T *Xarray_[],        \
                                T *Yarray_[], T *Warray_[], int const batchCount);

KRONMULT_DECL_PBATCHED(1, double) KRONMULT_DECL_PBATCHED(2, double)
KRONMULT_DECL_PBATCHED(3, double) KRONMULT_DECL_PBATCHED(4, double)
KRONMULT_DECL_PBATCHED(5, double) KRONMULT_DECL_PBATCHED(6, double)
KRONMULT_DECL_PBATCHED(1, float) KRONMULT_DECL_PBATCHED(2, float)
KRONMULT_DECL_PBATCHED(3, float) KRONMULT_DECL_PBATCHED(4, float)
KRONMULT_DECL_PBATCHED(5, float) KRONMULT_DECL_PBATCHED(6, float)
KRONMULT_DECL_BATCHED(1, double) KRONMULT_DECL_BATCHED(2, double)
KRONMULT_DECL_BATCHED(3, double) KRONMULT_DECL_BATCHED(4, double)
KRONMULT_DECL_BATCHED(5, double) KRONMULT_DECL_BATCHED(6, double)
KRONMULT_DECL_XBATCHED(1, double) KRONMULT_DECL_XBATCHED(2, double)
KRONMULT_DECL_XBATCHED(3, double) KRONMULT_DECL_XBATCHED(4, double)
KRONMULT_DECL_XBATCHED(5, double) KRONMULT_DECL_XBATCHED(6, double)
// float overloads: the reference templates accept float
// (kgemm_nn.hpp:688-701) but its host wrappers are double only.
KRONMULT_DECL_BATCHED(1, float) KRONMULT_DECL_BATCHED(2, float)
KRONMULT_DECL_BATCHED(3, float) KRONMULT_DECL_BATCHED(4, float)
KRONMULT_DECL_BATCHED(5, float) KRONMULT_DECL_BATCHED(6, float)
KRONMULT_DECL_XBATCHED(1, float) KRONMULT_DECL_XBATCHED(2, float)
KRONMULT_DECL_XBATCHED(3, float) KRONMULT_DECL_XBATCHED(4, float)
KRONMULT_DECL_XBATCHED(5, float) KRONMULT_DECL_XBATCHED(6, float)

// Host overloads for the vbatched kernels, with the kernels' parameter lists
// (kronmult6_vbatched.hpp:14-29; the reference has no host wrapper).
#define KRONMULT_VB_TAIL(T)                                                          \
    T const *const Aarray_[], T *pX_[], T *pY_[], T *W_, size_t const Wcapcity,     \
        int const batchCount
#define KRONMULT_DECL_VBATCHED(T)                                                     \
    void kronmult1_vbatched(int const m1, int const n1, KRONMULT_VB_TAIL(T));         \
    void kronmult2_vbatched(int const m1, int const n1, int const m2, int const n2,    \
                            KRONMULT_VB_TAIL(T));                                     \
    void kronmult3_vbatched(int const m1, int const n1, int const m2, int const n2,    \
                            int const m3, int const n3, KRONMULT_VB_TAIL(T));         \
    void kronmult4_vbatched(int const m1, int const n1, int const m2, int const n2,    \
                            int const m3, int const n3, int const m4, int const n4,   \
                            KRONMULT_VB_TAIL(T));                                     \
    void kronmult5_vbatched(int const m1, int const n1, int const m2, int const n2,    \
                            int const m3, int const n3, int const m4, int const n4,   \
                            int const m5, int const n5, KRONMULT_VB_TAIL(T));         \
    void kronmult6_vbatched(int const m1, int const n1, int const m2, int const n2,    \
                            int const m3, int const n3, int const m4, int const n4,   \
                            int const m5, int const n5, int const m6, int const n6,   \
                            KRONMULT_VB_TAIL(T));
KRONMULT_DECL_VBATCHED(double)
KRONMULT_DECL_VBATCHED(float)

// kgemm_nn_batched.cpp:9-18, kgemm_nt_batched.cpp:8-17
#define KRONMULT_DECL_KGEMM(NAME, T)                                                  \
    void NAME(int const mm, int const nn, int const kk, T const alpha,               \
              T *const Aarray_[], int const ldAarray_[], T *const Barray_[],          \
              int const ldBarray_[], T const beta, T *const Carray_[],                \
              int const ldCarray_[], int const batchCount);
KRONMULT_DECL_KGEMM(kgemm_nn_batched, double)
KRONMULT_DECL_KGEMM(kgemm_nt_batched, double)
KRONMULT_DECL_KGEMM(kgemm_nn_batched, float)
KRONMULT_DECL_KGEMM(kgemm_nt_batched, float)

// Status of the last drop-in call made on this thread.
kron_status kronmult_last_status();

#endif  // KRONMULT_DROPIN_HPP
