/* hjoracle.h -- CPU FP64 restatement of the hierarchically blocked one-sided
 * (hyperbolic) Jacobi SVD of venovako/GPUJACHx (level-2 driver 1/HypJacL2.cu,
 * level-1 kernels share/device_code_{nosort,cdsort}_{solveV,accumV}.hpp).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (libhypjac.so) never links or calls it.
 *
 * Arithmetic order.  Every floating-point operation is an IEEE-754 double
 * operation (fma, +, *, /, sqrt) in the exact order the gfx950 kernels use, so
 * the GPU path is expected to match this oracle bit for bit:
 *   - Gram partial of a row chunk: split into hjo_gram_waves() equal
 *     sub-chunks, one per wave, whose results are added in wave order; in a
 *     sub-chunk, 16-row slabs ascending; inside a slab, 4 MFMA steps
 *     (j = 0..3), each a chain over k = 0..3 of fma(x,y,acc) on row
 *     16g+4j+k (the gfx950 FP64 MFMA 16x16x4 is a sequential fma chain in k,
 *     measured on the device; row 4j+k of a slab is where the block
 *     update's MFMA leaves it, so the update can feed the next Gram from
 *     registers); partials of the chunks are added in order.
 *   - Block update: out(r,j) = fma chain over k=0..31 of G(r,k)*W(k,j) from +0.
 *   - Column norms in initD: 64 lane-strided fma chains + xor butterfly over
 *     64 lanes (32,16,8,4,2,1).
 * Everything else (Cholesky, inner Jacobi, rotations, sort/swap rules, the
 * 32-lane xor butterfly 16,8,4,2,1 of the dot products, convergence counting)
 * follows the reference's own order.  The only departure from the reference
 * is the summation order inside the Gram and the update (the reference's is
 * a skewed per-thread chain, share/device_code_common_Cholesky.hpp:31-40 and
 * share/device_code_common.hpp:166-178); both are exact in real arithmetic.
 */
#ifndef HJORACLE_H
#define HJORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* constants of the method (share/HypJacL.hpp:11-51) */
#define HJO_SQRT_HEPS   0x1.6a09e667f3bcdp-27 /* sqrt(2^-53) */
#define HJO_SQRT_2_HEPS 0x1p27                /* sqrt(2/2^-53) */
#define HJO_MYTOL       0x1.6a09e667f3bcdp-51 /* sqrt(32)*2^-53 */
#define HJO_COSHI_FIX   0x1.279a74590331cp+0  /* 2/sqrt(3), correctly rounded */
#define HJO_TANH1_FIX   0.5
#define HJO_NSWEEP      30u

/* ALG bits (share/HypJacL.hpp:88-110) */
#define HJO_BLK_ORI  8u
#define HJO_FULL_SVD 2u
#define HJO_CDSORT_1 1u

/* correctly rounded 1/sqrt(x) for positive normal x (own method: IEEE
 * 1/sqrt(x), then a neighbour test with a double-double residual). */
double hjo_rsqrt(double x);
void hjo_rsqrt_n(const double *x, double *y, long n);

/* 2x2 rotations, reference share/device_code_common_rotate.hpp:15-64
 * (USE_RSQRT branch).  Return 1 for a "big" rotation. */
int hjo_rot_t(double Apq, double Dp, double Dq, double *c, double *t);
int hjo_rot_h(double Apq, double Dp, double Dq, double *c, double *t);

/* sum of 32 lane values by the xor butterfly 16,8,4,2,1
 * (share/device_code_common_Kepler.hpp:26-74) */
double hjo_sum32(const double v[32]);

/* number of 64-row-aligned row chunks the Gram of one block pair is split
 * into; identical to the device library's choice for (nrow, pairs). */
unsigned hjo_gram_nchunk(unsigned nrow, unsigned npairs);

/* Waves sharing one Gram chunk of nr rows: 4, 2 or 1, each taking a
 * multiple of 64 rows. */
unsigned hjo_gram_waves(unsigned nr);

/* Gram partial of rows [r0, r0+nr) (nr % 64 == 0) of the column pair
 * (G0: 16 columns of block i, G1: 16 columns of block j, leading dim ld).
 * out[768]: tiles (00,10,11) x lane(64) x reg(4) in the MFMA D layout. */
void hjo_gram_partial(const double *G0, const double *G1, unsigned ld,
                      unsigned r0, unsigned nr, double out[768]);

/* one block pair: sum nchunk partials, Cholesky, inner Jacobi (level 0),
 * producing the 32x32 column-major transformation W.
 * kind: 0 = definite (definite = +1/-1), 1 = hyperbolic (npos local).
 * Returns the block's rotation count; *big = big rotations;
 * *upd = 1 if the block update must be applied: unless W is the identity
 * (no pivot pair rotated or swapped). */
unsigned hjo_pair(const double *partials, unsigned nchunk, int sort,
                  int hyp, int definite, unsigned npos,
                  const uint8_t *strat0, unsigned s0steps, unsigned nswp,
                  double W[1024], unsigned *big, int *upd, double R[1024]);

/* level-0 sweeps on R (32x32 upper triangular, column-major) accumulating
 * W (identity on entry); returns the rotation counts and (swapped_out, may
 * be NULL) whether a pivot pair was swapped by the sort rule. */
void hjo_l0(double *R, double *W, int sort, int hyp, int definite,
            unsigned npos, const uint8_t *strat0, unsigned s0steps,
            unsigned nswp, unsigned *blk_s_out, unsigned *blk_b_out,
            int *swapped_out);

/* Factorization of the block pairs: Cholesky of the Gram (the reference's
 * default build), QR by 32-row panels (USE_QR), QR with Drmac's sorted
 * level-0 (USE_DRMAC).  Returns the previous setting. */
#define HJO_FACT_CHOLESKY 0
#define HJO_FACT_QR       1
#define HJO_FACT_QR_DRMAC 2
int hjo_set_factorization(int f);
/* the QR path's pieces (hjoracle_qr.c) */
double hjo_nrm2_32(const double x[32]);
double hjo_cosa(double Ap, double Aq, const double x[32], const double y[32]);
void hjo_givens(double f, double g, double *c, double *s, double *r);
void hjo_qr32(double A[1024]);
void hjo_peeloff(double R0[1024], double R1[1024]);
void hjo_factorize_qr(const double *G0, const double *G1, unsigned ld, unsigned nrow,
                      double R[1024]);
void hjo_l0_drmac(double *R, double *W, int definite, const uint8_t *strat0,
                  unsigned s0steps, unsigned nswp, unsigned *blk_s_out,
                  unsigned *blk_b_out);
unsigned hjo_pair_qr(const double *G0, const double *G1, unsigned ld, unsigned nrow,
                     int sort, int hyp, int definite, unsigned npos, int drmac,
                     const uint8_t *strat0, unsigned s0steps, unsigned nswp,
                     double W[1024], unsigned *big, int *upd);

/* [G0 G1](rows 0..nr-1) <- [G0 G1] * W  (device order) */
void hjo_update(double *G0, double *G1, unsigned ld, unsigned nr,
                const double W[1024]);

/* column normalisation: G <- U, D = +-||g||^2 (device order) */
void hjo_initD(double *G, unsigned ld, unsigned nrow, unsigned ncol,
               unsigned nplus, double *D);

/* hjoracle_wide.c: the method with blocks of w = 16 or 32 columns (pairs of
 * 2w columns); w = 16 equals hjo_hypJacL2 bit for bit */
void hjo_gram_partial_w(const double *G, unsigned ld, unsigned w, unsigned a, unsigned b,
                        unsigned r0, unsigned nr, double *out);
unsigned hjo_pair_w(const double *partials, unsigned nchunk, unsigned w, int sort, int hyp,
                    int definite, unsigned npos, const uint8_t *strat0, unsigned s0steps,
                    unsigned nswp, double *W, unsigned *big, int *upd);
void hjo_update_w(double *M, unsigned ld, unsigned w, unsigned a, unsigned b, unsigned nr,
                  const double *W);
int hjo_hypJacL2_w(unsigned routine, unsigned nrow, unsigned ncol, unsigned nplus, double *G,
                   unsigned ldG, double *V, unsigned ldV, double *D, unsigned *glbSwp,
                   unsigned long long *glb_s, unsigned long long *glb_b, unsigned w,
                   const uint8_t *strat0, unsigned s0steps, const uint16_t *strat1,
                   unsigned s1steps, unsigned max_sweeps, int nthreads, unsigned *rot_b_trace);

/* Whole level-2 driver, reference 1/HypJacL2.cu:13-202 (returns the same
 * argument codes).  strat0: [s0steps][16][2] uint8, strat1: [s1steps][n1/2][2]
 * uint16.  nthreads <= 0: all OpenMP threads.  max_steps > 0 stops after
 * that many level-1 steps in total (bounded CPU-timing sample; the result is
 * then not converged).  rot_s_trace/rot_b_trace (optional, as long as the
 * sweep cap) receive the per-sweep ROT_S / ROT_B. */
int hjo_hypJacL2(unsigned routine, unsigned nrow, unsigned ncol, unsigned nplus,
                 double *G, unsigned ldG, double *V, unsigned ldV, double *D,
                 unsigned *glbSwp, unsigned long long *glb_s,
                 unsigned long long *glb_b,
                 const uint8_t *strat0, unsigned s0steps,
                 const uint16_t *strat1, unsigned s1steps,
                 int nthreads, long max_steps,
                 unsigned *rot_s_trace, unsigned *rot_b_trace);

/* level-2 sweep cap (30 = the reference's HYPJAC_NSWEEP); returns the old */
unsigned hjo_set_max_sweeps(unsigned n);

/* Per-visit trace for the following hjo_hypJacL2 calls: entry
 * (sweep*s1steps + step)*npairs + pair holds (rotations, big rotations) at
 * buf[2*entry], buf[2*entry+1]; cap_entries bounds the entries written.
 * buf == NULL switches tracing off. */
void hjo_set_trace(unsigned *buf, size_t cap_entries);

/* Gram order of the batched launch (gpujachx_amd/csrc/hj_batch.hip): the
 * pair's Gram by one wave over all rows (no chunks).  Returns the old mode. */
int hjo_set_single_wave_gram(int on);

#ifdef __cplusplus
}
#endif

#endif /* HJORACLE_H */
