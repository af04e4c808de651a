/* hjoracle.c -- CPU FP64 restatement of GPUJACHx's blocked one-sided
 * (hyperbolic) Jacobi SVD.  TEST INFRASTRUCTURE ONLY (the checker); see
 * hjoracle.h for the arithmetic-order contract.  Each function cites the
 * reference file:line it follows. */
#include "hjoracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline uint64_t d2u(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static inline double u2d(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }

/* ---------------------------------------------------------------------- */
/* Correctly rounded reciprocal square root.  Replaces the reference's
 * my_drsqrt_rn (share/device_code_common.hpp:89-133), whose contract is the
 * correctly rounded 1/sqrt(x).  Own method: y = RN(1/RN(sqrt(m))) for the
 * mantissa m in [1,4) is within 2 ulps; each neighbouring midpoint mid is
 * tested by the sign of m*mid^2 - 1, evaluated with exact products (fma)
 * and a final sum whose error is below 2^-100, far under the gap any
 * non-midpoint case leaves. */
static double rs_resid(double m, double y, double d)
{
  /* m*(y+d)^2 - 1 = (m*y^2 - 1) + 2*m*y*d + m*d^2 */
  const double h = y * y, l = fma(y, y, -h);
  const double p = m * h, pe = fma(m, h, -p);
  const double q = m * l;
  const double my = m * y;
  double T = (p - 1.0) + pe;
  T = T + q;
  T = T + (2.0 * d) * my;
  T = T + (m * d) * d;
  return T;
}

double hjo_rsqrt(double x)
{
  uint64_t b = d2u(x);
  if (isnan(x))
    return x + x;
  if (x == 0.0)
    return u2d((b & 0x8000000000000000ull) | 0x7ff0000000000000ull);
  if (x < 0.0)
    return u2d(0xfff8000000000000ull);
  if (isinf(x))
    return 0.0;
  int kadj = 0;
  if (!(b >> 52)) { /* subnormal: scale by 2^108 */
    x = x * 0x1p108;
    b = d2u(x);
    kadj = 54;
  }
  const int e = (int)(b >> 52) - 1023;
  const int k = (e >= 0) ? (e / 2) : -((1 - e) / 2); /* floor(e/2) */
  const double m = u2d((b & 0x000fffffffffffffull) | ((uint64_t)(1023 + e - 2 * k) << 52));
  double y = 1.0 / sqrt(m);
  for (int it = 0; it < 4; ++it) {
    const double up = u2d(d2u(y) + 1u);
    if (rs_resid(m, y, (up - y) * 0.5) < 0.0)
      y = up;
    else
      break;
  }
  for (int it = 0; it < 4; ++it) {
    const double dn = u2d(d2u(y) - 1u);
    if (rs_resid(m, y, -((y - dn) * 0.5)) > 0.0)
      y = dn;
    else
      break;
  }
  /* y * 2^-(k - kadj): exact power-of-two scaling */
  const int s = kadj - k;
  return y * u2d((uint64_t)(1023 + s) << 52);
}

void hjo_rsqrt_n(const double *x, double *y, long n)
{
  for (long i = 0; i < n; ++i)
    y[i] = hjo_rsqrt(x[i]);
}

/* ---------------------------------------------------------------------- */
/* share/device_code_common_rotate.hpp:189-213 (dRotT, USE_RSQRT) */
int hjo_rot_t(double Apq, double Dp, double Dq, double *c, double *t)
{
  const double h = Dq - Dp;
  const double ct2 = h / (2.0 * Apq);
  const double CT2 = fabs(ct2);
  if (CT2 < HJO_SQRT_HEPS) {
    const double cot = CT2 + 1.0;
    *t = copysign(1.0 / cot, ct2);
    *c = hjo_rsqrt(fma(*t, *t, 1.0));
    return 1;
  }
  else if (CT2 >= HJO_SQRT_2_HEPS) {
    const double cot = 2.0 * CT2;
    *c = 1.0;
    *t = copysign(1.0 / cot, ct2);
    return 0;
  }
  else {
    const double cot = CT2 + sqrt(fma(CT2, CT2, 1.0));
    *t = copysign(1.0 / cot, ct2);
    *c = hjo_rsqrt(fma(*t, *t, 1.0));
    return (*c < 1.0);
  }
}

/* share/device_code_common_rotate.hpp:214-237 (dRotH, USE_RSQRT) */
int hjo_rot_h(double Apq, double Dp, double Dq, double *c, double *t)
{
  const double h = Dq + Dp;
  const double ct2 = -(h / (2.0 * Apq));
  const double CT2 = fabs(ct2);
  if (CT2 >= HJO_SQRT_2_HEPS) {
    const double cot = 2.0 * CT2;
    *c = 1.0;
    *t = copysign(1.0 / cot, ct2);
    return 0;
  }
  else if (CT2 > 1.0) {
    const double cot = CT2 + sqrt(fma(CT2, CT2, -1.0));
    *t = copysign(1.0 / cot, ct2);
    *c = hjo_rsqrt(fma(-*t, *t, 1.0));
    return (*c > 1.0);
  }
  else {
    *c = HJO_COSHI_FIX;
    *t = HJO_TANH1_FIX;
    return 1;
  }
}

/* share/device_code_common_Kepler.hpp:26-74 (dSum32) */
double hjo_sum32(const double v[32])
{
  double s[32], u[32];
  memcpy(s, v, sizeof(s));
  for (unsigned off = 16u; off; off >>= 1u) {
    for (unsigned a = 0u; a < 32u; ++a)
      u[a] = s[a] + s[a ^ off];
    memcpy(s, u, sizeof(s));
  }
  return s[0];
}

static double sum64(const double v[64])
{
  double s[64], u[64];
  memcpy(s, v, sizeof(s));
  for (unsigned off = 32u; off; off >>= 1u) {
    for (unsigned a = 0u; a < 64u; ++a)
      u[a] = s[a] + s[a ^ off];
    memcpy(s, u, sizeof(s));
  }
  return s[0];
}

/* ---------------------------------------------------------------------- */
unsigned hjo_gram_nchunk(unsigned nrow, unsigned npairs)
{
  unsigned target = npairs ? (2048u / npairs) : 1u;
  if (target > 16u)
    target = 16u;
  if (!target)
    target = 1u;
  const unsigned g = nrow / 64u;
  unsigned nc = 1u;
  while ((nc * 2u) <= target && g && !(g % (nc * 2u)))
    nc *= 2u;
  return nc;
}

/* Waves sharing one chunk of nr rows (a multiple of 64 rows each). */
unsigned hjo_gram_waves(unsigned nr)
{
  unsigned nw = 4u;
  while (nw > 1u && (nr % (64u * nw)))
    nw >>= 1u;
  return nw;
}

/* Gram of the column pair over rows [r0, r0+nr) by one wave; reference
 * dGtG, share/device_code_common_Cholesky.hpp:4-56 (only the summation
 * order differs, see hjoracle.h). */
static void gram_wave(const double *G0, const double *G1, unsigned ld,
                      unsigned r0, unsigned nr, double out[768])
{
  double a00[16][16], a10[16][16], a11[16][16];
  memset(a00, 0, sizeof(a00));
  memset(a10, 0, sizeof(a10));
  memset(a11, 0, sizeof(a11));
  for (unsigned g = 0u; g < nr / 16u; ++g)
    for (unsigned j = 0u; j < 4u; ++j)
        for (unsigned k = 0u; k < 4u; ++k) {
        const size_t row = (size_t)r0 + 16u * g + 4u * j + k;
        double x0[16], x1[16];
        for (unsigned i = 0u; i < 16u; ++i) {
          x0[i] = G0[(size_t)i * ld + row];
          x1[i] = G1[(size_t)i * ld + row];
        }
        for (unsigned i = 0u; i < 16u; ++i)
          for (unsigned j = 0u; j < 16u; ++j) {
            a00[i][j] = fma(x0[i], x0[j], a00[i][j]);
            a10[i][j] = fma(x1[i], x0[j], a10[i][j]);
            a11[i][j] = fma(x1[i], x1[j], a11[i][j]);
          }
        }
  /* MFMA D layout: lane l, reg r holds D[4r + l/16][l%16] */
  for (unsigned l = 0u; l < 64u; ++l)
    for (unsigned r = 0u; r < 4u; ++r) {
      const unsigned i = 4u * r + l / 16u, j = l % 16u;
      out[0 * 256 + l * 4 + r] = a00[i][j];
      out[1 * 256 + l * 4 + r] = a10[i][j];
      out[2 * 256 + l * 4 + r] = a11[i][j];
    }
}

/* Gram partial of a chunk: hjo_gram_waves(nr) sub-chunks, one per wave,
 * added in wave order. */
void hjo_gram_partial(const double *G0, const double *G1, unsigned ld,
                      unsigned r0, unsigned nr, double out[768])
{
  const unsigned nw = hjo_gram_waves(nr), sub = nr / nw;
  gram_wave(G0, G1, ld, r0, sub, out);
  for (unsigned w = 1u; w < nw; ++w) {
    double t[768];
    gram_wave(G0, G1, ld, r0 + w * sub, sub, t);
    for (unsigned e = 0u; e < 768u; ++e)
      out[e] = out[e] + t[e];
  }
}

#define F32(A, r, c) (A)[(c) * 32u + (r)]

/* Inner sweep(s) on R (32x32) accumulating W; reference
 * share/device_code_nosort.hpp:4-222 (sort == 0) and
 * share/device_code_cdsort_KISS.hpp:4-260 + share/device_code_cdsort.hpp:25-167
 * (sort == 1).  hyp: the block pair mixes J signs (local npos).
 * swapped_out (may be NULL): whether a pivot pair was swapped without a
 * rotation (the sort rule); W stays the identity exactly when no pivot pair
 * was rotated or swapped. */
void hjo_l0(double *R, double *W, int sort, int hyp, int definite,
               unsigned npos, const uint8_t *strat0, unsigned s0steps,
               unsigned nswp, unsigned *blk_s_out, unsigned *blk_b_out,
               int *swapped_out)
{
  unsigned blk_s = 0u, blk_b = 0u;
  int swapped = 0;
  for (unsigned swp = 0u; swp < nswp; ++swp) {
    unsigned swp_s = 0u, swp_b = 0u;
    for (unsigned step = 0u; step < s0steps; ++step) {
      for (unsigned y = 0u; y < 16u; ++y) {
        const unsigned p = strat0[(step * 16u + y) * 2u + 0u];
        const unsigned q = strat0[(step * 16u + y) * 2u + 1u];
        double Gp[32], Gq[32], Vp[32], Vq[32], v[32];
        for (unsigned x = 0u; x < 32u; ++x) {
          Gp[x] = F32(R, x, p); Gq[x] = F32(R, x, q);
          Vp[x] = F32(W, x, p); Vq[x] = F32(W, x, q);
        }
        for (unsigned x = 0u; x < 32u; ++x) v[x] = fma(Gp[x], Gp[x], 0.0);
        double Dp = hjo_sum32(v);
        for (unsigned x = 0u; x < 32u; ++x) v[x] = fma(Gq[x], Gq[x], 0.0);
        double Dq = hjo_sum32(v);
        for (unsigned x = 0u; x < 32u; ++x) v[x] = fma(Gp[x], Gq[x], 0.0);
        const double Apq = hjo_sum32(v);

        const double Dp_ = sqrt(Dp), Dq_ = sqrt(Dq), Apq_ = fabs(Apq), Dpq_ = Dp_ * Dq_;
        const int transf_s = !(Apq_ < (Dpq_ * HJO_MYTOL));
        swp_s += (unsigned)transf_s;

        /* local definiteness of the pivot pair (hyperbolic case) */
        int def = definite;
        if (hyp) {
          if ((p < npos) && (q < npos)) def = +1;
          else if ((p >= npos) && (q >= npos)) def = -1;
          else def = 0;
        }

        int transf_b = 0;
        if (!sort) {
          /* share/device_code_nosort.hpp:59-81 and :165-196 */
          if (transf_s) {
            double c, t, t_;
            if (!hyp || def) { transf_b = hjo_rot_t(Apq, Dp, Dq, &c, &t); t_ = -t; }
            else { transf_b = hjo_rot_h(Apq, Dp, Dq, &c, &t); t_ = t; }
            for (unsigned x = 0u; x < 32u; ++x) {
              if (transf_b) {
                F32(R, x, p) = c * fma(t_, Gq[x], Gp[x]);
                F32(R, x, q) = c * fma(t, Gp[x], Gq[x]);
                F32(W, x, p) = c * fma(t_, Vq[x], Vp[x]);
                F32(W, x, q) = c * fma(t, Vp[x], Vq[x]);
              }
              else {
                F32(R, x, p) = fma(t_, Gq[x], Gp[x]);
                F32(R, x, q) = fma(t, Gp[x], Gq[x]);
                F32(W, x, p) = fma(t_, Vq[x], Vp[x]);
                F32(W, x, q) = fma(t, Vp[x], Vq[x]);
              }
            }
          }
        }
        else {
          /* share/device_code_cdsort_KISS.hpp:226-274 (posd), :355-403 (negd),
           * share/device_code_cdsort.hpp:88-141 (mixed pair) */
          if (transf_s) {
            double c, t, t_;
            if (def) {
              transf_b = hjo_rot_t(Apq, Dp, Dq, &c, &t);
              t_ = -t;
              Dp = fma(t, Apq, Dp);
              Dq = fma(t_, Apq, Dq);
            }
            else {
              transf_b = hjo_rot_h(Apq, Dp, Dq, &c, &t);
              t_ = t;
            }
            const int keep = (!def) || ((def > 0) && (Dp >= Dq)) || ((def < 0) && (Dp <= Dq));
            const unsigned P = keep ? p : q, Q = keep ? q : p;
            for (unsigned x = 0u; x < 32u; ++x) {
              if (transf_b) {
                F32(R, x, P) = c * fma(t_, Gq[x], Gp[x]);
                F32(R, x, Q) = c * fma(t, Gp[x], Gq[x]);
                F32(W, x, P) = c * fma(t_, Vq[x], Vp[x]);
                F32(W, x, Q) = c * fma(t, Vp[x], Vq[x]);
              }
              else {
                F32(R, x, P) = fma(t_, Gq[x], Gp[x]);
                F32(R, x, Q) = fma(t, Gp[x], Gq[x]);
                F32(W, x, P) = fma(t_, Vq[x], Vp[x]);
                F32(W, x, Q) = fma(t, Vp[x], Vq[x]);
              }
            }
          }
          else if (((def > 0) && (Dp < Dq)) || ((def < 0) && (Dq < Dp))) {
            for (unsigned x = 0u; x < 32u; ++x) {
              F32(R, x, p) = Gq[x]; F32(R, x, q) = Gp[x];
              F32(W, x, p) = Vq[x]; F32(W, x, q) = Vp[x];
            }
            swapped = 1;
          }
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
  if (swapped_out)
    *swapped_out = swapped;
}

/* Level-1 block-pair body: share/device_code_cdsort_accumV.hpp:35-94 and its
 * siblings; factorisation share/device_code_common_Cholesky.hpp:58-171. */
unsigned hjo_pair(const double *partials, unsigned nchunk, int sort,
                  int hyp, int definite, unsigned npos,
                  const uint8_t *strat0, unsigned s0steps, unsigned nswp,
                  double W[1024], unsigned *big, int *upd, double Rout[1024])
{
  double P[768];
  memcpy(P, partials, sizeof(P));
  for (unsigned c = 1u; c < nchunk; ++c)
    for (unsigned e = 0u; e < 768u; ++e)
      P[e] = P[e] + partials[(size_t)c * 768u + e];
  double A[1024];
  memset(A, 0, sizeof(A));
  for (unsigned l = 0u; l < 64u; ++l)
    for (unsigned r = 0u; r < 4u; ++r) {
      const unsigned i = 4u * r + l / 16u, j = l % 16u;
      F32(A, i, j) = P[0 * 256 + l * 4 + r];
      F32(A, 16u + i, j) = P[1 * 256 + l * 4 + r];
      F32(A, 16u + i, 16u + j) = P[2 * 256 + l * 4 + r];
    }
  /* dCholesky32 (USE_RSQRT): cdiv with x*rsqrt(Akk) below, sqrt on the
   * diagonal; cmod(j,k) = fma(-A(x,k), A(j,k), A(x,j)). */
  for (unsigned k = 0u; k < 32u; ++k) {
    const double Akk = F32(A, k, k);
    const double rk = hjo_rsqrt(Akk);
    for (unsigned x = k + 1u; x < 32u; ++x)
      F32(A, x, k) = F32(A, x, k) * rk;
    F32(A, k, k) = sqrt(Akk);
    for (unsigned j = k + 1u; j < 32u; ++j)
      for (unsigned x = j; x < 32u; ++x)
        F32(A, x, j) = fma(-F32(A, x, k), F32(A, j, k), F32(A, x, j));
  }
  double R[1024];
  for (unsigned y = 0u; y < 32u; ++y)
    for (unsigned x = 0u; x < 32u; ++x)
      F32(R, y, x) = (x >= y) ? F32(A, x, y) : 0.0;
  for (unsigned y = 0u; y < 32u; ++y)
    for (unsigned x = 0u; x < 32u; ++x)
      F32(W, x, y) = (x == y) ? 1.0 : 0.0;
  unsigned s = 0u, b = 0u;
  int swapped = 0;
  hjo_l0(R, W, sort, hyp, definite, npos, strat0, s0steps, nswp, &s, &b, &swapped);
  if (Rout)
    memcpy(Rout, R, sizeof(R));
  *big = b;
  /* the block update is left out exactly when W is the identity: no pivot
   * pair rotated or (sorted variants) swapped; multiplying by it would only
   * turn negative zeros positive (this build's rule: the reference always
   * multiplies in its sorted variants) */
  *upd = (s != 0u) || swapped;
  return s;
}

/* share/device_code_common.hpp:146-185 (dMultAV), device summation order */
void hjo_update(double *G0, double *G1, unsigned ld, unsigned nr,
                const double W[1024])
{
  for (unsigned r = 0u; r < nr; ++r) {
    double g[32], o[32];
    for (unsigned k = 0u; k < 16u; ++k) {
      g[k] = G0[(size_t)k * ld + r];
      g[16u + k] = G1[(size_t)k * ld + r];
    }
    for (unsigned j = 0u; j < 32u; ++j) {
      double acc = 0.0;
      for (unsigned k = 0u; k < 32u; ++k)
        acc = fma(g[k], F32(W, k, j), acc);
      o[j] = acc;
    }
    for (unsigned k = 0u; k < 16u; ++k) {
      G0[(size_t)k * ld + r] = o[k];
      G1[(size_t)k * ld + r] = o[16u + k];
    }
  }
}

/* share/device_code_common.hpp:187-219 (dInitD), device summation order */
void hjo_initD(double *G, unsigned ld, unsigned nrow, unsigned ncol,
               unsigned nplus, double *D)
{
  for (unsigned j = 0u; j < ncol; ++j) {
    double *const g = G + (size_t)j * ld;
    double lanes[64];
    for (unsigned l = 0u; l < 64u; ++l) {
      double x = 0.0;
      for (unsigned r = l; r < nrow; r += 64u)
        x = fma(g[r], g[r], x);
      lanes[l] = x;
    }
    const double s = sum64(lanes);
    const double x = sqrt(s);
    for (unsigned r = 0u; r < nrow; ++r)
      g[r] = g[r] / x;
    D[j] = (j < nplus) ? s : -s;
  }
}

/* local J structure of a block pair, share/device_code_nosort_accumV.hpp:53-76 */
static void local_npos(unsigned nplus, unsigned iBlk, unsigned jBlk,
                       unsigned *npos, int *definite)
{
  const unsigned ifc0 = iBlk * 16u, ifc1 = jBlk * 16u;
  if (nplus <= ifc0) { *npos = 0u; *definite = -1; }
  else if (nplus <= (ifc0 + 16u)) { *npos = nplus - ifc0; *definite = 0; }
  else if (nplus <= ifc1) { *npos = 16u; *definite = 0; }
  else if (nplus <= (ifc1 + 16u)) { *npos = 16u + (nplus - ifc1); *definite = 0; }
  else { *npos = 32u; *definite = +1; }
}

static unsigned g_max_sweeps = HJO_NSWEEP;
static int g_single_wave_gram = 0;
static int g_fact = HJO_FACT_CHOLESKY;

/* the factorization of the block pairs (the reference's build options
 * USE_QR / USE_DRMAC, share/device_code_common.hpp:78-85,140-144) */
int hjo_set_factorization(int f)
{
  const int old = g_fact;
  if (f >= HJO_FACT_CHOLESKY && f <= HJO_FACT_QR_DRMAC)
    g_fact = f;
  return old;
}

/* Gram order of the batched launch: one chunk, one wave over all rows. */
int hjo_set_single_wave_gram(int on)
{
  const int old = g_single_wave_gram;
  g_single_wave_gram = on ? 1 : 0;
  return old;
}
unsigned hjo_set_max_sweeps(unsigned n)
{
  const unsigned old = g_max_sweeps;
  if (n >= 1u && n <= 256u)
    g_max_sweeps = n;
  return old;
}

/* optional per-visit trace: [sweep][step][pair] -> (rotations, big rotations) */
static unsigned *g_trace = NULL;
static size_t g_trace_cap = 0;
void hjo_set_trace(unsigned *buf, size_t cap_entries)
{
  g_trace = buf;
  g_trace_cap = cap_entries;
}

/* 1/HypJacL2.cu:13-202 */
int hjo_hypJacL2(unsigned routine, unsigned nrow, unsigned ncol, unsigned nplus,
                 double *G, unsigned ldG, double *V, unsigned ldV, double *D,
                 unsigned *glbSwp, unsigned long long *glb_s,
                 unsigned long long *glb_b,
                 const uint8_t *strat0, unsigned s0steps,
                 const uint16_t *strat1, unsigned s1steps,
                 int nthreads, long max_steps,
                 unsigned *rot_s_trace, unsigned *rot_b_trace)
{
  if (routine >= 16u) return -1;
  const int cdsort = (routine & HJO_CDSORT_1) != 0;
  const int full_svd = (routine & HJO_FULL_SVD) != 0;
  const int blk_ori = (routine & HJO_BLK_ORI) != 0;
  if (!nrow || (nrow % 64u)) return -2;
  if (!ncol || (ncol > nrow) || (ncol % 32u)) return -3;
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

  const unsigned nswp = blk_ori ? 1u : HJO_NSWEEP;
  const unsigned np = ncol / 32u;
  const unsigned nchunk = g_single_wave_gram ? 1u : hjo_gram_nchunk(nrow, np);
  const unsigned CH = nrow / nchunk;
  const int definite = nplus ? ((nplus < ncol) ? 0 : +1) : -1;

  if (full_svd)
    for (unsigned j = 0u; j < ncol; ++j)
      for (unsigned i = 0u; i < ncol; ++i)
        V[(size_t)j * ldV + i] = (i == j) ? 1.0 : 0.0;

  double *const part = (double*)malloc(sizeof(double) * 768u * nchunk * np);
  unsigned *const ps = (unsigned*)malloc(sizeof(unsigned) * 2u * np);
  if (!part || !ps) { free(part); free(ps); return -100; }

#ifdef _OPENMP
  if (nthreads > 0)
    omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif

  *glb_s = 0ull;
  *glb_b = 0ull;
  unsigned blk_swp = 0u;
  long steps_done = 0;
  int stop = 0;
  while (blk_swp < g_max_sweeps && !stop) {
    unsigned long long cvg = 0ull;
    for (unsigned step = 0u; step < s1steps; ++step) {
      if (max_steps > 0 && steps_done >= max_steps) { stop = 1; break; }
      ++steps_done;
#pragma omp parallel for schedule(dynamic, 1)
      for (unsigned pi = 0u; pi < np; ++pi) {
        const unsigned iBlk = strat1[((size_t)step * np + pi) * 2u + 0u];
        const unsigned jBlk = strat1[((size_t)step * np + pi) * 2u + 1u];
        double *const G0 = G + (size_t)iBlk * 16u * ldG;
        double *const G1 = G + (size_t)jBlk * 16u * ldG;
        double *const pp = part + (size_t)pi * 768u * nchunk;
        int hyp = 0, def = definite;
        unsigned npos = 0u;
        if (!definite) {
          local_npos(nplus, iBlk, jBlk, &npos, &def);
          hyp = (def == 0);
        }
        double W[1024];
        unsigned b = 0u, s = 0u;
        int upd = 0;
        if (g_fact != HJO_FACT_CHOLESKY) {
          s = hjo_pair_qr(G0, G1, ldG, nrow, cdsort, hyp, def, npos,
                          g_fact == HJO_FACT_QR_DRMAC, strat0, s0steps, nswp, W, &b, &upd);
        }
        else {
          if (g_single_wave_gram)
            gram_wave(G0, G1, ldG, 0u, nrow, pp);
          else
            for (unsigned c = 0u; c < nchunk; ++c)
              hjo_gram_partial(G0, G1, ldG, c * CH, CH, pp + (size_t)c * 768u);
          s = hjo_pair(pp, nchunk, cdsort, hyp, def, npos, strat0,
                       s0steps, nswp, W, &b, &upd, NULL);
        }
        ps[2u * pi] = s;
        ps[2u * pi + 1u] = b;
        if (upd) {
          hjo_update(G0, G1, ldG, nrow, W);
          if (full_svd)
            hjo_update(V + (size_t)iBlk * 16u * ldV, V + (size_t)jBlk * 16u * ldV, ldV, ncol, W);
        }
      }
      if (g_trace) {
        const size_t base = ((size_t)blk_swp * s1steps + step) * np;
        for (unsigned pi = 0u; pi < np; ++pi)
          if (base + pi < g_trace_cap) {
            g_trace[2u * (base + pi)] = ps[2u * pi];
            g_trace[2u * (base + pi) + 1u] = ps[2u * pi + 1u];
          }
      }
      /* counter semantics of share/device_code_nosort.hpp:94-103 */
      for (unsigned pi = 0u; pi < np; ++pi) {
        const unsigned s = ps[2u * pi], b = ps[2u * pi + 1u];
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
    if (rot_s_trace) rot_s_trace[blk_swp - 1u] = cvg_s;
    if (rot_b_trace) rot_b_trace[blk_swp - 1u] = cvg_b;
    if (!cvg_b || stop)
      break;
  }
  *glbSwp = blk_swp;
  hjo_initD(G, ldG, nrow, ncol, nplus, D);
  free(part);
  free(ps);
  return 0;
}
