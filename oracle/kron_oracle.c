/*
 * kron_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference kronmult algorithm (hipdac-lab/kronmult),
 * used as the parity checker for the HIP library and as the "port" CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load this library.  It is never part of the product path.
 *
 * Every function follows the reference's operation order, so that on inputs
 * without overlapping Y the results are bit-identical to the reference CPU
 * build (checked against oracle/_ref by tests/test_oracle_vs_ref.py):
 *
 *   kgemm_nt_v1   kgemm_nt.hpp:215-385   C = alpha*A*B^T + beta*C
 *   kgemm_nn_v1   kgemm_nn.hpp:227-406   C = alpha*A*B   + beta*C
 *   kronmultv     kronmultv.hpp:19-148   tail recursion W = X*A1^T, recurse
 *                 kronmultv.hpp:152-238  base case Y += A1*X (beta=1, atomic)
 *   kronmult1     kronmult1.hpp:17-64    d=1 calls kgemm_nn directly
 *   batched       kronmult6_batched.hpp:14-94 (omp only for d=6, :75-79)
 *   xbatched      kronmult_xbatched.hpp:19-157
 *   vbatched      kronmult_vbatched.hpp:19-248
 *   kgemm_*_batched  kgemm_nn_batched.hpp:8-79, kgemm_nt_batched.hpp:8-75
 *
 * Accumulation type: double for both T=double and T=float
 * (kgemm_nn.hpp:688-701, kgemm_nt.hpp:557-571).  The beta rules are
 * beta==1 -> atomic add, beta==0 -> overwrite without reading C,
 * otherwise C = beta*C + alpha*cij (kgemm_nn.hpp:391-400,
 * kgemm_nt.hpp:371-379).
 *
 * Indexing is column-major; A1 acts on the slowest index of X and Ad on the
 * fastest (kroncommon.hpp:97-199).
 *
 * Build: gcc -O2 -ffp-contract=off -fopenmp -fPIC -shared (see oracle/build.py).
 */
#include <stddef.h>
#include <string.h>
#include <stdint.h>
#include <omp.h>

#define MINI(a, b) (((a) < (b)) ? (a) : (b))
#define MAXI(a, b) (((a) > (b)) ? (a) : (b))

/* OpenMP atomic add, the CPU stand-in for atomicAdd (kroncommon.hpp:36-64). */
static inline void atomic_add_d(double *p, double v)
{
#pragma omp atomic
    *p += v;
}
static inline void atomic_add_f(float *p, float v)
{
#pragma omp atomic
    *p += v;
}

/*
 * ---------------------------------------------------------------------------
 * Type-generic bodies.  ORACLE_T is the storage type, accumulation is double.
 * ---------------------------------------------------------------------------
 */
#define ORACLE_DEFINE(T, SFX, ATOMIC_ADD)                                          \
                                                                                   \
/* kgemm_nt.hpp:306-384 with ij_start=1, ij_size=1 (the CPU path).            */ \
static void kgemm_nt_##SFX(int mm, int nn, int kk, T alpha_in,                     \
                           const T *A_, int ldA, const T *B_, int ldB,             \
                           T beta_in, T *C_, int ldC)                              \
{                                                                                  \
    const double alpha = alpha_in, beta = beta_in;                                 \
    const int nb = 64;                                                             \
    for (int jstart = 1; jstart <= nn; jstart += nb) {                             \
        const int jend = MINI(nn, jstart + nb - 1);                                \
        const int jsize = jend - jstart + 1;                                       \
        for (int istart = 1; istart <= mm; istart += nb) {                         \
            const int iend = MINI(mm, istart + nb - 1);                            \
            const int isize = iend - istart + 1;                                   \
            for (int ij0 = 0; ij0 < isize * jsize; ij0++) {                        \
                const int i = (ij0 % isize) + 1;                                   \
                const int j = (ij0 - (i - 1)) / isize + 1;                         \
                const int ia = (istart - 1) + i;                                   \
                const int ib = (jstart - 1) + j;                                   \
                const T *Ap = &A_[(ia - 1)];   /* A(ia,1) */                       \
                const T *Bp = &B_[(ib - 1)];   /* B(ib,1) */                       \
                double cij = 0;                                                    \
                for (int k = 0; k < kk; k++) {                                     \
                    const double aik = *Ap;                                        \
                    const double bjk = *Bp;                                        \
                    cij += aik * bjk;                                              \
                    Ap += ldA;                                                     \
                    Bp += ldB;                                                     \
                }                                                                  \
                T *c = &C_[(ia - 1) + (size_t)(ib - 1) * ldC];                     \
                if (beta == 1) {                                                   \
                    ATOMIC_ADD(c, (T)(alpha * cij));                               \
                } else if (beta == 0) {                                            \
                    *c = (T)(alpha * cij);                                         \
                } else {                                                           \
                    *c = (T)(beta * (double)(*c) + alpha * cij);                   \
                }                                                                  \
            }                                                                      \
        }                                                                          \
    }                                                                              \
}                                                                                  \
                                                                                   \
/* kgemm_nn.hpp:321-405 with ij_start=1, ij_size=1 (the CPU path).            */ \
static void kgemm_nn_##SFX(int mm, int nn, int kk, T alpha_in,                     \
                           const T *A_, int ldA, const T *B_, int ldB,             \
                           T beta_in, T *C_, int ldC)                              \
{                                                                                  \
    const double alpha = alpha_in, beta = beta_in;                                 \
    const int nb = 64;                                                             \
    for (int istart = 1; istart <= mm; istart += nb) {                             \
        const int iend = MINI(mm, istart + nb - 1);                                \
        const int isize = iend - istart + 1;                                       \
        for (int jstart = 1; jstart <= nn; jstart += nb) {                         \
            const int jend = MINI(nn, jstart + nb - 1);                            \
            const int jsize = jend - jstart + 1;                                   \
            for (int ij0 = 0; ij0 < isize * jsize; ij0++) {                        \
                const int i = (ij0 % isize) + 1;                                   \
                const int j = ((ij0 - (i - 1)) / isize) + 1;                       \
                const int ia = (istart - 1) + i;                                   \
                const int jb = (jstart - 1) + j;                                   \
                const T *Ap = &A_[(ia - 1)];                     /* A(ia,1) */     \
                const T *Bp = &B_[(size_t)(jb - 1) * ldB];       /* B(1,jb) */     \
                double cij = 0;                                                    \
                for (int k = 0; k < kk; k++) {                                     \
                    const double aik = *Ap;                                        \
                    const double bkj = *Bp;                                        \
                    cij += aik * bkj;                                              \
                    Ap += ldA;                                                     \
                    Bp += 1;                                                       \
                }                                                                  \
                T *c = &C_[(ia - 1) + (size_t)(jb - 1) * ldC];                     \
                const T alpha_cij = (T)(alpha * cij);                              \
                if (beta == 1) {                                                   \
                    ATOMIC_ADD(c, alpha_cij);                                      \
                } else if (beta == 0) {                                            \
                    *c = alpha_cij;                                                \
                } else {                                                           \
                    *c = (T)(beta * (double)(*c) + (double)alpha_cij);             \
                }                                                                  \
            }                                                                      \
        }                                                                          \
    }                                                                              \
}                                                                                  \
                                                                                   \
/* kronmultv.hpp:19-238.  m[k], n[k], A[k], ld[k] for k = 0..ndim-1.          */ \
static void kronmultv_##SFX(int ndim, const int *m, const int *n,                  \
                            const T *const *A, const int *ld, int nvec,            \
                            T *X_, T *Y_, T *W_)                                   \
{                                                                                  \
    if (ndim == 1) {                                                               \
        /* kronmultv.hpp:152-195: Y(m1,nvec) += A1 * X(n1,nvec), beta=1 */         \
        kgemm_nn_##SFX(m[0], nvec, n[0], (T)1, A[0], ld[0], X_, n[0],              \
                       (T)1, Y_, m[0]);                                            \
        return;                                                                    \
    }                                                                              \
    int n26 = 1;                                                                   \
    for (int k = 1; k < ndim; k++) n26 *= n[k];                                    \
    const size_t ldX = (size_t)n26 * n[0];                                         \
    const size_t ldW = (size_t)n26 * m[0];                                         \
    for (int i = 0; i < nvec; i++) {                                               \
        /* kronmultv.hpp:105-127: Wi = Xi * A1^T, beta = 0 */                      \
        kgemm_nt_##SFX(n26, m[0], n[0], (T)1, X_ + i * ldX, n26,                   \
                       A[0], ld[0], (T)0, W_ + i * ldW, n26);                      \
    }                                                                              \
    /* kronmultv.hpp:130-145: recurse with nvec*m1, buffers (W, Y, X) */           \
    kronmultv_##SFX(ndim - 1, m + 1, n + 1, A + 1, ld + 1, nvec * m[0],            \
                    W_, Y_, X_);                                                   \
}                                                                                  \
                                                                                   \
/* Square front end kronmultx.hpp:20-58 / kronmult1.hpp:17-64.               */ \
static void kronmult_sq_##SFX(int ndim, int nsq, const T *const *A, int lda_in,    \
                              T *X_, T *Y_, T *W_)                                 \
{                                                                                  \
    const int lda = (lda_in == 0) ? nsq : lda_in;                                  \
    int m[6], n[6], ld[6];                                                         \
    for (int k = 0; k < 6; k++) { m[k] = nsq; n[k] = nsq; ld[k] = lda; }          \
    if (ndim == 1) {                                                               \
        /* kronmult1.hpp:43-62: kgemm_nn(n, nvec=1, n, 1, A1, lda, X, n, 1, Y, n) */\
        kgemm_nn_##SFX(nsq, 1, nsq, (T)1, A[0], lda, X_, nsq, (T)1, Y_, nsq);      \
        return;                                                                    \
    }                                                                              \
    kronmultv_##SFX(ndim, m, n, A, ld, 1, X_, Y_, W_);                             \
}                                                                                  \
                                                                                   \
/* kronmult{1..6}_batched.hpp, e.g. kronmult6_batched.hpp:14-94.             */ \
/* A is (n,n,ndim,batch); X, Y, W are (n^ndim, batch).  X is overwritten.    */ \
int oracle_batched_##SFX(int ndim, int nsq, const T *Aarray, T *X, T *Y, T *W,     \
                         int batchCount, int use_omp)                              \
{                                                                                  \
    if (ndim < 1 || ndim > 6 || nsq < 1 || batchCount < 0) return 1;              \
    size_t nd = 1;                                                                 \
    for (int k = 0; k < ndim; k++) nd *= (size_t)nsq;                              \
    const size_t n2 = (size_t)nsq * nsq;                                           \
    _Pragma("omp parallel for if(use_omp)")                                        \
    for (int ib = 0; ib < batchCount; ib++) {                                      \
        const T *A[6];                                                             \
        for (int k = 0; k < ndim; k++)                                             \
            A[k] = Aarray + ((size_t)ib * ndim + k) * n2;                          \
        kronmult_sq_##SFX(ndim, nsq, A, 0, X + ib * nd, Y + ib * nd,               \
                          W + ib * nd);                                            \
    }                                                                              \
    return 0;                                                                      \
}                                                                                  \
                                                                                   \
/* kronmult_xbatched.hpp:19-157 (CPU path: iz_start=1, iz_size=1).           */ \
int oracle_xbatched_##SFX(int ndim, int nsq, const T *const *Aarray, int lda,      \
                          T *const *pX, T *const *pY, T *const *pW,                \
                          int batchCount_in, int subbatchCount, int use_omp)       \
{                                                                                  \
    if (ndim < 1 || ndim > 6 || nsq < 1 || batchCount_in < 0) return 1;          \
    if (subbatchCount == 0) subbatchCount = MAXI(1, batchCount_in / 2);            \
    if (subbatchCount < 1) return 1;                                               \
    size_t nd = 1;                                                                 \
    for (int k = 0; k < ndim; k++) nd *= (size_t)nsq;                              \
    for (int ibs = 1; ibs <= batchCount_in; ibs += subbatchCount) {                \
        const int ibe = MINI(batchCount_in, ibs + subbatchCount - 1);              \
        const int bc = ibe - ibs + 1;                                              \
        _Pragma("omp parallel for if(use_omp)")                                    \
        for (int iz = 1; iz <= bc; iz++) {                                         \
            const int ibatch = (ibs - 1) + iz;                                     \
            T *Xin = pX[ibatch - 1];                                               \
            T *Xp = (batchCount_in > 1) ? pW[2 * (iz - 1)] : Xin;                  \
            T *Wp = (batchCount_in > 1) ? pW[2 * (iz - 1) + 1] : pW[ibatch - 1];   \
            if (Xp != Xin) memcpy(Xp, Xin, nd * sizeof(T));                        \
            const T *A[6];                                                         \
            for (int k = 0; k < ndim; k++)                                         \
                A[k] = Aarray[k + (size_t)(ibatch - 1) * ndim];                    \
            kronmult_sq_##SFX(ndim, nsq, A, lda, Xp, pY[ibatch - 1], Wp);          \
        }                                                                          \
    }                                                                              \
    return 0;                                                                      \
}                                                                                  \
                                                                                   \
/* kronmult_vbatched.hpp:19-248.  m_in/n_in hold ndim extents.               */ \
int oracle_vbatched_##SFX(int ndim, const int *m_in, const int *n_in,              \
                          const T *const *Aarray, T *const *pX, T *const *pY,      \
                          T *W_, size_t Wcapacity_bytes, int batchCount_in,        \
                          int use_omp)                                             \
{                                                                                  \
    if (ndim < 1 || ndim > 6 || batchCount_in < 0) return 1;                       \
    int m[6], n[6], ld[6];                                                         \
    for (int k = 0; k < 6; k++) {                                                  \
        m[k] = (k < ndim) ? m_in[k] : 1;                                           \
        n[k] = (k < ndim) ? n_in[k] : 1;                                           \
        if (m[k] < 1 || n[k] < 1) return 1;                                        \
        ld[k] = m[k];                                                              \
    }                                                                              \
    /* kronmult_vbatched.hpp:136-157 */                                            \
    const int n56 = n[4] * n[5], n456 = n[3] * n56, n3456 = n[2] * n456;           \
    const int n23456 = n[1] * n3456;                                               \
    const int sizeX = n[0] * n23456;                                               \
    const int m12 = m[0] * m[1], m123 = m12 * m[2], m1234 = m123 * m[3];           \
    const int m12345 = m1234 * m[4];                                               \
    const int sizeW = MAXI(MAXI(MAXI(m[0] * n23456, m12 * n3456),                  \
                                MAXI(m123 * n456, m1234 * n56)), m12345 * n[5]);   \
    const int sizeXW = MAXI(sizeX, sizeW);                                         \
    const size_t cap = Wcapacity_bytes / (2 * (size_t)sizeXW * sizeof(T));         \
    const int subbatch = MAXI(1, (int)MINI((size_t)batchCount_in, cap));           \
    for (int ibs = 1; ibs <= batchCount_in; ibs += subbatch) {                     \
        const int ibe = MINI(batchCount_in, ibs + subbatch - 1);                   \
        const int bc = ibe - ibs + 1;                                              \
        _Pragma("omp parallel for if(use_omp)")                                    \
        for (int iz = 1; iz <= bc; iz++) {                                         \
            const int ibatch = (ibs - 1) + iz;                                     \
            T *Xp = W_ + (size_t)(iz - 1) * 2 * sizeXW;                            \
            T *Wp = Xp + sizeXW;                                                   \
            memcpy(Xp, pX[ibatch - 1], (size_t)sizeX * sizeof(T));                 \
            const T *A[6];                                                         \
            for (int k = 0; k < ndim; k++)                                         \
                A[k] = Aarray[k + (size_t)(ibatch - 1) * ndim];                    \
            kronmultv_##SFX(ndim, m, n, A, ld, 1, Xp, pY[ibatch - 1], Wp);         \
        }                                                                          \
    }                                                                              \
    return 0;                                                                      \
}                                                                                  \
                                                                                   \
/* kgemm_nn_batched.hpp:8-79 / kgemm_nt_batched.hpp:8-75.  trans = 'N'|'T'.  */ \
int oracle_kgemm_batched_##SFX(char trans, int mm, int nn, int kk, T alpha,        \
                               T *const *Aarray, const int *ldA,                   \
                               T *const *Barray, const int *ldB, T beta,           \
                               T *const *Carray, const int *ldC, int batchCount)   \
{                                                                                  \
    if (batchCount < 0) return 1;                                                  \
    for (int ib = 0; ib < batchCount; ib++) {                                      \
        if (trans == 'T' || trans == 't')                                          \
            kgemm_nt_##SFX(mm, nn, kk, alpha, Aarray[ib], ldA[ib], Barray[ib],     \
                           ldB[ib], beta, Carray[ib], ldC[ib]);                    \
        else                                                                       \
            kgemm_nn_##SFX(mm, nn, kk, alpha, Aarray[ib], ldA[ib], Barray[ib],     \
                           ldB[ib], beta, Carray[ib], ldC[ib]);                    \
    }                                                                              \
    return 0;                                                                      \
}

ORACLE_DEFINE(double, f64, atomic_add_d)
ORACLE_DEFINE(float, f32, atomic_add_f)

/*
 * Textbook dense-Kronecker brute force (the reference tests' own checker,
 * test_kronmult6_batched.cpp:345-483, test_kronmult6_xbatched.cpp:554-585):
 * Y(ic) += sum_jc prod_k A_k(i_k, j_k) * X(jc), accumulated in double,
 * with rectangular factors A_k of m[k] x n[k] at leading dimension ld[k].
 * This is independent of the recursion above and pins it.
 */
void oracle_dense_kron_f64(int ndim, const int *m, const int *n,
                           const double *const *A, const int *ld,
                           const double *X, double *Y)
{
    size_t sizeX = 1, sizeY = 1;
    for (int k = 0; k < ndim; k++) { sizeX *= n[k]; sizeY *= m[k]; }
    for (size_t ic = 0; ic < sizeY; ic++) {
        int iidx[6];
        size_t r = ic;
        for (int k = ndim - 1; k >= 0; k--) { iidx[k] = (int)(r % m[k]); r /= m[k]; }
        double acc = 0;
        for (size_t jc = 0; jc < sizeX; jc++) {
            size_t s = jc;
            double c = 1;
            for (int k = ndim - 1; k >= 0; k--) {
                const int jk = (int)(s % n[k]);
                s /= n[k];
                c *= A[k][iidx[k] + (size_t)jk * ld[k]];
            }
            acc += c * X[jc];
        }
        Y[ic] += acc;
    }
}

/* The portable PRNG used for all synthetic inputs: splitmix64. */
static inline uint64_t splitmix64_next(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Uniform [-1,1) from the top 53 bits: u = 2*(k * 2^-53) - 1. */
void oracle_fill_uniform_f64(double *out, size_t count, uint64_t seed)
{
    uint64_t s = seed;
    for (size_t i = 0; i < count; i++) {
        const uint64_t k = splitmix64_next(&s) >> 11;
        out[i] = 2.0 * ((double)k * (1.0 / 9007199254740992.0)) - 1.0;
    }
}

/* Thread count of the OpenMP loops above (CPU-baseline timing). */
void oracle_set_threads(int t)
{
    omp_set_num_threads(t > 0 ? t : 1);
}
