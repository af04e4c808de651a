/* hjoracle_wide.c -- the blocked one-sided (hyperbolic) Jacobi method with
 * the block width w as a parameter: w = 16 is the reference's
 * (share/HypJacL.hpp:112-116, HYPJACL1_NCOLB), w = 32 pairs two 32-column
 * blocks into 64-column block pairs (a 64x64 Gram, Cholesky and level-0).
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the library's wide-block form.
 * Only tests/ may load it.
 *
 * Every step is the w = 16 arithmetic of hjoracle.c written for any width,
 * in the operation order of the gfx950 kernels:
 *   - Gram: 16x16 tiles (ti >= tj) of the 2w x 2w pair Gram, each element a
 *     chain of fma over the rows of a wave's sub-chunk (16-row slabs, rows
 *     4j+k in order), the waves' results added in wave order, the chunks'
 *     in chunk order; partial layout per chunk: tile ti(ti+1)/2 + tj, then
 *     lane l, register r for element (4r + l/16, l%16) -- for w = 16 the
 *     [a00 a10 a11] layout of hjo_gram_partial;
 *   - Cholesky of the 2w x 2w Gram as dCholesky32
 *     (share/device_code_common_Cholesky.hpp:58-158, USE_RSQRT);
 *   - level-0 sweeps on R with a table of order 2w (w pivot pairs a step),
 *     the dot products summed by the xor butterfly w, w/2, ..., 1 over the
 *     2w rows; convergence test |Apq| >= sqrt(Dp Dq) sqrt(2w) 2^-53 (for
 *     w = 16 the reference's MYTOL, share/HypJacL.hpp:11-51);
 *   - block update out(r, j) = fma chain over k = 0..2w-1 from +0;
 *   - local J structure of a pair, share/device_code_nosort_accumV.hpp:53-76,
 *     with blocks of w columns.
 * With w = 16 the result equals hjo_hypJacL2's bit for bit (tested), which
 * pins every step of the generic code on the reference's own width.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "hjoracle.h"

#define MAXN2 64u  /* widest pair: 64 columns */
#define MAXT 10u   /* 16x16 tiles of its lower triangle */

/* sum of n (a power of two) values by the xor butterfly n/2, ..., 1 */
static double bsum(const double *v, unsigned n)
{
  double s[MAXN2], u[MAXN2];
  memcpy(s, v, sizeof(double) * n);
  for (unsigned off = n / 2u; off; off >>= 1u) {
    for (unsigned a = 0u; a < n; ++a)
      u[a] = s[a] + s[a ^ off];
    memcpy(s, u, sizeof(double) * n);
  }
  return s[0];
}

/* Gram tiles of the pair columns col[0..n2) over rows [r0, r0+nr), one wave */
static void gram_wave_w(const double *G, unsigned ld, const size_t *col, unsigned n2,
                        unsigned r0, unsigned nr, double *out)
{
  const unsigned nt = n2 / 16u, T = nt * (nt + 1u) / 2u;
  static __thread double acc[MAXT][16][16];
  memset(acc, 0, sizeof(acc));
  for (unsigned g = 0u; g < nr / 16u; ++g)
    for (unsigned j = 0u; j < 4u; ++j)
      for (unsigned k = 0u; k < 4u; ++k) {
        const size_t row = (size_t)r0 + 16u * g + 4u * j + k;
        double x[MAXN2];
        for (unsigned c = 0u; c < n2; ++c)
          x[c] = G[col[c] * ld + row];
        unsigned tt = 0u;
        for (unsigned ti = 0u; ti < nt; ++ti)
          for (unsigned tj = 0u; tj <= ti; ++tj, ++tt)
            for (unsigned i = 0u; i < 16u; ++i)
              for (unsigned jj = 0u; jj < 16u; ++jj)
                acc[tt][i][jj] = fma(x[16u * ti + i], x[16u * tj + jj], acc[tt][i][jj]);
      }
  for (unsigned tt = 0u; tt < T; ++tt)
    for (unsigned l = 0u; l < 64u; ++l)
      for (unsigned r = 0u; r < 4u; ++r)
        out[tt * 256u + l * 4u + r] = acc[tt][4u * r + l / 16u][l % 16u];
}

/* pair columns: block a's w columns, then block b's */
static void pair_cols(unsigned w, unsigned a, unsigned b, unsigned ld, size_t *col)
{
  (void)ld;
  for (unsigned c = 0u; c < w; ++c) {
    col[c] = (size_t)a * w + c;
    col[w + c] = (size_t)b * w + c;
  }
}

/* Gram partial of rows [r0, r0+nr) of the pair (a, b): hjo_gram_waves(nr)
 * sub-chunks, added in wave order; n2 * (n2/16 + 1) * 8 doubles */
void hjo_gram_partial_w(const double *G, unsigned ld, unsigned w, unsigned a, unsigned b,
                        unsigned r0, unsigned nr, double *out)
{
  const unsigned n2 = 2u * w, nt = n2 / 16u, T = nt * (nt + 1u) / 2u;
  size_t col[MAXN2];
  pair_cols(w, a, b, ld, col);
  const unsigned nw = hjo_gram_waves(nr), sub = nr / nw;
  gram_wave_w(G, ld, col, n2, r0, sub, out);
  for (unsigned v = 1u; v < nw; ++v) {
    double t[MAXT * 256u];
    gram_wave_w(G, ld, col, n2, r0 + v * sub, sub, t);
    for (unsigned e = 0u; e < T * 256u; ++e)
      out[e] = out[e] + t[e];
  }
}

#define AW(A, r, c) (A)[(size_t)(c) * n2 + (r)]

/* inner sweeps on R (n2 x n2) accumulating W: hjo_l0 for any width */
static void l0_w(double *R, double *W, unsigned n2, int sort, int hyp, int definite,
                 unsigned npos, const uint8_t *strat0, unsigned s0steps, unsigned nswp,
                 double tol, unsigned *blk_s_out, unsigned *blk_b_out, int *swapped_out)
{
  const unsigned npv = n2 / 2u;
  unsigned blk_s = 0u, blk_b = 0u;
  int swapped = 0;
  for (unsigned swp = 0u; swp < nswp; ++swp) {
    unsigned swp_s = 0u, swp_b = 0u;
    for (unsigned step = 0u; step < s0steps; ++step) {
      for (unsigned y = 0u; y < npv; ++y) {
        const unsigned p = strat0[(step * npv + y) * 2u + 0u];
        const unsigned q = strat0[(step * npv + y) * 2u + 1u];
        double Gp[MAXN2], Gq[MAXN2], Vp[MAXN2], Vq[MAXN2], v[MAXN2];
        for (unsigned x = 0u; x < n2; ++x) {
          Gp[x] = AW(R, x, p); Gq[x] = AW(R, x, q);
          Vp[x] = AW(W, x, p); Vq[x] = AW(W, x, q);
        }
        for (unsigned x = 0u; x < n2; ++x) v[x] = fma(Gp[x], Gp[x], 0.0);
        double Dp = bsum(v, n2);
        for (unsigned x = 0u; x < n2; ++x) v[x] = fma(Gq[x], Gq[x], 0.0);
        double Dq = bsum(v, n2);
        for (unsigned x = 0u; x < n2; ++x) v[x] = fma(Gp[x], Gq[x], 0.0);
        const double Apq = bsum(v, n2);

        const double Dp_ = sqrt(Dp), Dq_ = sqrt(Dq), Apq_ = fabs(Apq), Dpq_ = Dp_ * Dq_;
        const int transf_s = !(Apq_ < (Dpq_ * tol));
        swp_s += (unsigned)transf_s;

        int def = definite;
        if (hyp) {
          if ((p < npos) && (q < npos)) def = +1;
          else if ((p >= npos) && (q >= npos)) def = -1;
          else def = 0;
        }
        int transf_b = 0;
        if (transf_s) {
          double c, t, t_;
          /* unsorted: share/device_code_nosort.hpp:59-81, :165-196; sorted:
           * share/device_code_cdsort_KISS.hpp:226-274, :355-403 and
           * share/device_code_cdsort.hpp:88-141 */
          const int trig = sort ? (def != 0) : (!hyp || def);
          if (trig) {
            transf_b = hjo_rot_t(Apq, Dp, Dq, &c, &t);
            t_ = -t;
            if (sort) {
              Dp = fma(t, Apq, Dp);
              Dq = fma(t_, Apq, Dq);
            }
          }
          else {
            transf_b = hjo_rot_h(Apq, Dp, Dq, &c, &t);
            t_ = t;
          }
          int keep = 1;
          if (sort)
            keep = (!def) || ((def > 0) && (Dp >= Dq)) || ((def < 0) && (Dp <= Dq));
          const unsigned P = keep ? p : q, Q = keep ? q : p;
          for (unsigned x = 0u; x < n2; ++x) {
            const double rp = fma(t_, Gq[x], Gp[x]), rq = fma(t, Gp[x], Gq[x]);
            const double wp = fma(t_, Vq[x], Vp[x]), wq = fma(t, Vp[x], Vq[x]);
            AW(R, x, P) = transf_b ? c * rp : rp;
            AW(R, x, Q) = transf_b ? c * rq : rq;
            AW(W, x, P) = transf_b ? c * wp : wp;
            AW(W, x, Q) = transf_b ? c * wq : wq;
          }
        }
        else if (sort && (((def > 0) && (Dp < Dq)) || ((def < 0) && (Dq < Dp)))) {
          for (unsigned x = 0u; x < n2; ++x) {
            AW(R, x, p) = Gq[x]; AW(R, x, q) = Gp[x];
            AW(W, x, p) = Vq[x]; AW(W, x, q) = Vp[x];
          }
          swapped = 1;
        }
        swp_b += (unsigned)transf_b;
      }
    }
    if (swp_s) {
      blk_s += swp_s;
      blk_b += swp_b;
    }
    else
      break;
  }
  *blk_s_out = blk_s;
  *blk_b_out = blk_b;
  *swapped_out = swapped;
}

/* one block pair of width w from its Gram partials: Cholesky, level-0,
 * W (n2 x n2, column-major); returns the rotations */
unsigned hjo_pair_w(const double *partials, unsigned nchunk, unsigned w, int sort, int hyp,
                    int definite, unsigned npos, const uint8_t *strat0, unsigned s0steps,
                    unsigned nswp, double *W, unsigned *big, int *upd)
{
  const unsigned n2 = 2u * w, nt = n2 / 16u, T = nt * (nt + 1u) / 2u;
  double P[MAXT * 256u];
  memcpy(P, partials, sizeof(double) * T * 256u);
  for (unsigned c = 1u; c < nchunk; ++c)
    for (unsigned e = 0u; e < T * 256u; ++e)
      P[e] = P[e] + partials[(size_t)c * T * 256u + e];
  double A[MAXN2 * MAXN2];
  memset(A, 0, sizeof(A));
  unsigned tt = 0u;
  for (unsigned ti = 0u; ti < nt; ++ti)
    for (unsigned tj = 0u; tj <= ti; ++tj, ++tt)
      for (unsigned l = 0u; l < 64u; ++l)
        for (unsigned r = 0u; r < 4u; ++r)
          AW(A, 16u * ti + 4u * r + l / 16u, 16u * tj + l % 16u) = P[tt * 256u + l * 4u + r];
  /* dCholesky32 (USE_RSQRT) at n2 pivots */
  for (unsigned k = 0u; k < n2; ++k) {
    const double Akk = AW(A, k, k);
    const double rk = hjo_rsqrt(Akk);
    for (unsigned x = k + 1u; x < n2; ++x)
      AW(A, x, k) = AW(A, x, k) * rk;
    AW(A, k, k) = sqrt(Akk);
    for (unsigned j = k + 1u; j < n2; ++j)
      for (unsigned x = j; x < n2; ++x)
        AW(A, x, j) = fma(-AW(A, x, k), AW(A, j, k), AW(A, x, j));
  }
  double R[MAXN2 * MAXN2];
  for (unsigned y = 0u; y < n2; ++y)
    for (unsigned x = 0u; x < n2; ++x) {
      AW(R, y, x) = (x >= y) ? AW(A, x, y) : 0.0;
      AW(W, x, y) = (x == y) ? 1.0 : 0.0;
    }
  const double tol = sqrt((double)n2) * 0x1p-53;
  unsigned s = 0u, b = 0u;
  int swapped = 0;
  l0_w(R, W, n2, sort, hyp, definite, npos, strat0, s0steps, nswp, tol, &s, &b, &swapped);
  *big = b;
  *upd = (s != 0u) || swapped;
  return s;
}

/* rows [0, nr) of the pair (a, b) of M times W: fma chain over k from +0 */
void hjo_update_w(double *M, unsigned ld, unsigned w, unsigned a, unsigned b, unsigned nr,
                  const double *W)
{
  const unsigned n2 = 2u * w;
  size_t col[MAXN2];
  pair_cols(w, a, b, ld, col);
  for (unsigned r = 0u; r < nr; ++r) {
    double g[MAXN2], o[MAXN2];
    for (unsigned k = 0u; k < n2; ++k)
      g[k] = M[col[k] * ld + r];
    for (unsigned j = 0u; j < n2; ++j) {
      double acc = 0.0;
      for (unsigned k = 0u; k < n2; ++k)
        acc = fma(g[k], AW(W, k, j), acc);
      o[j] = acc;
    }
    for (unsigned k = 0u; k < n2; ++k)
      M[col[k] * ld + r] = o[k];
  }
}

/* local J structure of a pair of w-column blocks,
 * share/device_code_nosort_accumV.hpp:53-76 */
static void local_npos_w(unsigned nplus, unsigned w, unsigned iBlk, unsigned jBlk,
                         unsigned *npos, int *definite)
{
  const unsigned ifc0 = iBlk * w, ifc1 = jBlk * w;
  if (nplus <= ifc0) { *npos = 0u; *definite = -1; }
  else if (nplus <= (ifc0 + w)) { *npos = nplus - ifc0; *definite = 0; }
  else if (nplus <= ifc1) { *npos = w; *definite = 0; }
  else if (nplus <= (ifc1 + w)) { *npos = w + (nplus - ifc1); *definite = 0; }
  else { *npos = 2u * w; *definite = +1; }
}

/* 1/HypJacL2.cu:13-202 with blocks of w columns (w = 16 or 32): strat0 of
 * order 2w (s0steps steps of w pairs), strat1 of order ncol / w; sweeps
 * stop after max_sweeps.  Argument codes as hjo_hypJacL2, -3 also for ncol
 * not a multiple of 2w. */
int hjo_hypJacL2_w(unsigned routine, unsigned nrow, unsigned ncol, unsigned nplus, double *G,
                   unsigned ldG, double *V, unsigned ldV, double *D, unsigned *glbSwp,
                   unsigned long long *glb_s, unsigned long long *glb_b, unsigned w,
                   const uint8_t *strat0, unsigned s0steps, const uint16_t *strat1,
                   unsigned s1steps, unsigned max_sweeps, int nthreads, unsigned *rot_b_trace)
{
  if (routine >= 16u) return -1;
  if (w != 16u && w != 32u) return -3;
  const int cdsort = (routine & HJO_CDSORT_1) != 0;
  const int full_svd = (routine & HJO_FULL_SVD) != 0;
  const int blk_ori = (routine & HJO_BLK_ORI) != 0;
  if (!nrow || (nrow % 64u)) return -2;
  if (!ncol || (ncol > nrow) || (ncol % (2u * w))) return -3;
  if (nplus > ncol) return -4;
  if (!G) return -5;
  if (ldG < nrow) return -6;
  if (full_svd) {
    if (!V) return -7;
    if (ldV < ncol) return -8;
  }
  if (!D) return -9;
  if (!glbSwp) return -10;
  if (!glb_s) return -11;
  if (!glb_b) return -12;
  (void)nthreads;

  const unsigned n2 = 2u * w, nt = n2 / 16u, T = nt * (nt + 1u) / 2u;
  const unsigned nswp = blk_ori ? 1u : HJO_NSWEEP;
  const unsigned np = ncol / n2;
  const unsigned nchunk = hjo_gram_nchunk(nrow, np);
  const unsigned CH = nrow / nchunk;
  const int definite = nplus ? ((nplus < ncol) ? 0 : +1) : -1;

  if (full_svd)
    for (unsigned j = 0u; j < ncol; ++j)
      for (unsigned i = 0u; i < ncol; ++i)
        V[(size_t)j * ldV + i] = (i == j) ? 1.0 : 0.0;

  double *const part = (double *)malloc(sizeof(double) * T * 256u * nchunk * np);
  double *const Wb = (double *)malloc(sizeof(double) * n2 * n2 * np);
  unsigned *const ps = (unsigned *)malloc(sizeof(unsigned) * 3u * np);
  if (!part || !Wb || !ps) { free(part); free(Wb); free(ps); return -100; }

  *glb_s = 0ull;
  *glb_b = 0ull;
  unsigned blk_swp = 0u;
  while (blk_swp < max_sweeps) {
    unsigned long long cvg = 0ull;
    for (unsigned step = 0u; step < s1steps; ++step) {
#pragma omp parallel for schedule(dynamic, 1)
      for (unsigned pi = 0u; pi < np; ++pi) {
        const unsigned iBlk = strat1[((size_t)step * np + pi) * 2u + 0u];
        const unsigned jBlk = strat1[((size_t)step * np + pi) * 2u + 1u];
        double *const pp = part + (size_t)pi * T * 256u * nchunk;
        int hyp = 0, def = definite;
        unsigned npos = 0u;
        if (!definite) {
          local_npos_w(nplus, w, iBlk, jBlk, &npos, &def);
          hyp = (def == 0);
        }
        for (unsigned c = 0u; c < nchunk; ++c)
          hjo_gram_partial_w(G, ldG, w, iBlk, jBlk, c * CH, CH, pp + (size_t)c * T * 256u);
        double *const W = Wb + (size_t)pi * n2 * n2;
        unsigned b = 0u;
        int upd = 0;
        const unsigned s = hjo_pair_w(pp, nchunk, w, cdsort, hyp, def, npos, strat0, s0steps,
                                      nswp, W, &b, &upd);
        ps[3u * pi] = s;
        ps[3u * pi + 1u] = b;
        ps[3u * pi + 2u] = (unsigned)upd;
        if (upd) {
          hjo_update_w(G, ldG, w, iBlk, jBlk, nrow, W);
          if (full_svd)
            hjo_update_w(V, ldV, w, iBlk, jBlk, ncol, W);
        }
      }
      /* counter semantics of share/device_code_nosort.hpp:94-103 */
      for (unsigned pi = 0u; pi < np; ++pi) {
        const unsigned s = ps[3u * pi], b = ps[3u * pi + 1u];
        if (s) {
          if (b)
            cvg += (((unsigned long long)b) << 32u) | s;
          else
            cvg = (cvg & 0xffffffff00000000ull) | (unsigned)((unsigned)cvg + s);
        }
      }
    }
    ++blk_swp;
    const unsigned cvg_s = (unsigned)cvg, cvg_b = (unsigned)(cvg >> 32u);
    *glb_s += cvg_s;
    *glb_b += cvg_b;
    if (rot_b_trace) rot_b_trace[blk_swp - 1u] = cvg_b;
    if (!cvg_b)
      break;
  }
  *glbSwp = blk_swp;
  hjo_initD(G, ldG, nrow, ncol, nplus, D);
  free(part);
  free(Wb);
  free(ps);
  return 0;
}
