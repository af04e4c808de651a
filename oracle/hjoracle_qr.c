/* hjoracle_qr.c -- CPU FP64 restatement of the reference's optional
 * factorization and sorting variants: the QR path (USE_QR,
 * share/device_code_common_QR.hpp:4-94 with dQR32 of
 * share/device_code_common_Kepler.hpp:326-371 and dGivens of
 * share/device_code_common_rotate.hpp:120-155) and the Drmac sorted level-0
 * (USE_DRMAC, share/device_code_cdsort_Drmac.hpp:4-316 with dNRM2_32 / dCOSA
 * of share/device_code_common_Kepler.hpp:230-324 and the scaling helpers of
 * share/device_code_common_scale.hpp).  TEST INFRASTRUCTURE ONLY (the
 * checker).
 *
 * Arithmetic order: the reference's, step by step; every 32-element dot
 * product or sum of squares is the xor butterfly 16,8,4,2,1 (hjo_sum32),
 * maxima and minima likewise.  Where the reference leaves the contraction of
 * a*b + c*d to its compiler (the Givens applications of dPeelOff), the
 * products and the sum are rounded separately, as the gfx950 kernels do. */
#include "hjoracle.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <string.h>

#define F32(A, r, c) (A)[(c) * 32u + (r)]

/* MU_, NU_, MU_EPS, NU_8 of share/HypJacL.hpp:53-86 */
static double mu_(void) { return 1.4916681462400414314624091158735E-154; }
static double nu_(void) { return 2.370187977027293260705088939015E+153; }
static double mu_eps(void) { return 1.0020841800044863889980540256751E-292; }
static double nu_8(void) { return 2.24711641857789464E+307; }

static double max32(const double v[32])
{
  double s[32], u[32];
  for (unsigned a = 0u; a < 32u; ++a) s[a] = fabs(v[a]);
  for (unsigned off = 16u; off; off >>= 1u) {
    for (unsigned a = 0u; a < 32u; ++a) u[a] = fmax(s[a], s[a ^ off]);
    memcpy(s, u, sizeof(s));
  }
  return s[0];
}

static double min32_nz(const double v[32])
{
  double s[32], u[32];
  for (unsigned a = 0u; a < 32u; ++a) s[a] = (v[a] == 0.0) ? DBL_MAX : fabs(v[a]);
  for (unsigned off = 16u; off; off >>= 1u) {
    for (unsigned a = 0u; a < 32u; ++a) u[a] = fmin(s[a], s[a ^ off]);
    memcpy(s, u, sizeof(s));
  }
  return s[0];
}

/* share/device_code_common_scale.hpp:4-56 */
static int find_scl(double x, double t)
{
  int x_n, t_n;
  const double x_m = frexp(x, &x_n);
  const double t_m = frexp(t, &t_n);
  return (x <= t) ? ((t_n - x_n) + (x_m < t_m)) : ((t_n - x_n) - (x_m > t_m));
}

static int find_scl2(double x)
{
  const int scl = ilogb(x);
  return scl + (scl & 1);
}

static void normalize2(int *scl, double *ssq)
{
  const int s = find_scl2(*ssq);
  *scl += s;
  *ssq = scalbn(*ssq, -s);
}

static int cmp_scl(int s1, int s2, double x1, double x2)
{
  if (s1 < s2) return -1;
  if (s2 < s1) return 1;
  if (x1 < x2) return -1;
  if (x2 < x1) return 1;
  return 0;
}

static void sum_scl(int s1, int s2, double x1, double x2, int *s, double *x)
{
  if (cmp_scl(s1, s2, x1, x2) <= 0) {
    *s = s2;
    *x = scalbn(x1, s1 - s2) + x2;
  }
  else {
    *s = s1;
    *x = scalbn(x2, s2 - s1) + x1;
  }
}

/* sums of squares of the entries in one magnitude band
 * (share/device_code_common_Kepler.hpp:182-201) */
static double ssq_band(const double x[32], int band, int scl)
{
  double v[32];
  for (unsigned a = 0u; a < 32u; ++a) {
    const double x_ = fabs(x[a]);
    double y = 0.0;
    if (band == 2) y = (x_ > nu_()) ? scalbn(x_, scl) : 0.0;
    else if (band == 1) y = ((x_ >= mu_()) && (x_ <= nu_())) ? x_ : 0.0;
    else y = (x_ < mu_()) ? scalbn(x_, scl) : 0.0;
    v[a] = y * y;
  }
  return hjo_sum32(v);
}

/* dSSQ32, share/device_code_common_Kepler.hpp:203-260: ssq * 2^scl is the
 * sum of squares, computed without overflow or harmful underflow */
static int ssq32(double *ssq, const double x[32])
{
  const double M = max32(x);
  if (M == 0.0) {
    *ssq = 0.0;
    return INT_MIN;
  }
  const double m = min32_nz(x);
  const int lo = (m < mu_()), hi = (M > nu_());
  int sc0, sc1, sc2, scl;
  double sq0, sq1, sq2, s;
  if (hi) {
    sc2 = find_scl(M, nu_());
    sq2 = ssq_band(x, 2, sc2);
    sc2 *= -2;
    if (m > nu_()) {
      scl = sc2;
      s = sq2;
    }
    else {
      sc1 = 0;
      sq1 = ssq_band(x, 1, 0);
      if (sq1 > 0.0) {
        normalize2(&sc1, &sq1);
        normalize2(&sc2, &sq2);
        sum_scl(sc1, sc2, sq1, sq2, &scl, &s);
      }
      else {
        scl = sc2;
        s = sq2;
      }
    }
  }
  else if (lo) {
    sc0 = find_scl(m, mu_());
    sq0 = ssq_band(x, 0, sc0);
    sc0 *= -2;
    if (M < mu_()) {
      scl = sc0;
      s = sq0;
    }
    else {
      sc1 = 0;
      sq1 = ssq_band(x, 1, 0);
      if (sq1 > 0.0) {
        normalize2(&sc1, &sq1);
        normalize2(&sc0, &sq0);
        sum_scl(sc0, sc1, sq0, sq1, &scl, &s);
      }
      else {
        scl = sc0;
        s = sq0;
      }
    }
  }
  else {
    double v[32];
    for (unsigned a = 0u; a < 32u; ++a) v[a] = x[a] * x[a];
    scl = 0;
    s = hjo_sum32(v);
  }
  *ssq = s;
  return scl;
}

/* dNRM2_32, share/device_code_common_Kepler.hpp:269-275 */
double hjo_nrm2_32(const double x[32])
{
  double ssq;
  const int scl = ssq32(&ssq, x) / 2;
  return scalbn(sqrt(ssq), scl);
}

/* dCOSA_ (Ap >= Aq), share/device_code_common_Kepler.hpp:277-303 */
static double cosa_(double Ap, double Aq, const double x[32], const double y[32])
{
  int big, small;
  double z[32];
  if (Aq >= 1.0) {
    big = ((Ap * Aq) >= DBL_MAX);
    small = 0;
  }
  else {
    big = 0;
    small = ((Ap * Aq) <= mu_eps());
  }
  if (big) {
    for (unsigned a = 0u; a < 32u; ++a)
      z[a] = (fabs(x[a]) >= fabs(y[a])) ? ((x[a] / Ap) * y[a]) : ((y[a] / Ap) * x[a]);
    return hjo_sum32(z) / Aq;
  }
  if (small) {
    for (unsigned a = 0u; a < 32u; ++a)
      z[a] = (fabs(x[a]) >= fabs(y[a])) ? (x[a] * (y[a] / Aq)) : (y[a] * (x[a] / Aq));
    return hjo_sum32(z) / Ap;
  }
  for (unsigned a = 0u; a < 32u; ++a) z[a] = x[a] * y[a];
  return (hjo_sum32(z) / Aq) / Ap;
}

/* dCOSA, share/device_code_common_Kepler.hpp:305-311 */
double hjo_cosa(double Ap, double Aq, const double x[32], const double y[32])
{
  return (Ap >= Aq) ? cosa_(Ap, Aq, x, y) : cosa_(Aq, Ap, y, x);
}

/* dGivens, share/device_code_common_rotate.hpp:120-155:
 * [c s; -s c] [f; g] = [r; 0], r >= 0 */
void hjo_givens(double f, double g, double *c, double *s, double *r)
{
  const double f_ = fabs(f), g_ = fabs(g);
  if (f_ >= g_) {
    if (g_ == 0.0) {
      *c = copysign(1.0, f);
      *s = 0.0;
      *r = f_;
    }
    else {
      const double g_f = g_ / f_;
      *r = f_ * sqrt(fma(g_f, g_f, 1.0));
      *c = f / *r;
      *s = g / *r;
    }
  }
  else {
    if (f_ == 0.0) {
      *c = 0.0;
      *s = copysign(1.0, g);
      *r = g_;
    }
    else {
      const double f_g = f_ / g_;
      *r = g_ * sqrt(fma(f_g, f_g, 1.0));
      *c = f / *r;
      *s = g / *r;
    }
  }
}

/* dQR32, share/device_code_common_Kepler.hpp:326-371: Householder QR of a
 * 32x32 column-major block in place (the R factor; entries below the
 * diagonal of the reduced columns set to zero) */
void hjo_qr32(double A[1024])
{
  for (unsigned k = 0u; k < 31u; ++k) {
    double Ak[32];
    for (unsigned x = 0u; x < 32u; ++x) Ak[x] = (x >= k) ? F32(A, x, k) : 0.0;
    const double nrm2 = hjo_nrm2_32(Ak);
    if (!(nrm2 > 0.0))
      continue;
    const double alpha = F32(A, k, k);
    const double beta = copysign(nrm2, alpha);
    const double alpha_beta = alpha + beta;
    const double tau = alpha_beta / beta;
    double v[32];
    for (unsigned x = 0u; x < 32u; ++x)
      v[x] = (x == k) ? 1.0 : ((x > k) ? (Ak[x] / alpha_beta) : 0.0);
    for (unsigned y = k + 1u; y < 32u; ++y) {
      double pr[32];
      for (unsigned x = 0u; x < 32u; ++x) pr[x] = v[x] * F32(A, x, y);
      const double dp = hjo_sum32(pr);
      const double tdp = -(tau * dp);
      for (unsigned x = 0u; x < 32u; ++x) F32(A, x, y) = fma(tdp, v[x], F32(A, x, y));
    }
    for (unsigned x = k; x < 32u; ++x) F32(A, x, k) = (x == k) ? -beta : 0.0;
  }
}

/* dPeelOff, share/device_code_common_QR.hpp:4-66: R0 <- the R factor of
 * [R0; R1] for upper triangular R0, R1, by a wavefront of Givens rotations:
 * at step k, row x of R0 meets row x-k of R1 (x >= k), the rotation taken
 * from R0(x,x), R1(x-k,x) and applied to columns x..31 */
void hjo_peeloff(double R0[1024], double R1[1024])
{
  for (unsigned k = 0u; k < 32u; ++k) {
    double c[32], s[32], r[32];
    for (unsigned x = k; x < 32u; ++x)
      hjo_givens(F32(R0, x, x), F32(R1, x - k, x), &c[x], &s[x], &r[x]);
    for (unsigned col = k; col < 32u; ++col)
      for (unsigned x = k; x <= col; ++x) {
        const unsigned xk = x - k;
        if (x < col) {
          const double a = F32(R0, x, col), b = F32(R1, xk, col);
          const double ca = c[x] * a, sb = s[x] * b, cb = c[x] * b, sa = s[x] * a;
          F32(R0, x, col) = ca + sb;
          F32(R1, xk, col) = cb - sa;
        }
        else {
          F32(R0, x, col) = r[x];
          F32(R1, xk, col) = 0.0;
        }
      }
  }
}

/* dFactorize (QR), share/device_code_common_QR.hpp:68-92: R of the nrow x 32
 * column pair [G0 G1] (16 + 16 columns, leading dimension ld) by 32-row
 * panels in order */
void hjo_factorize_qr(const double *G0, const double *G1, unsigned ld, unsigned nrow,
                      double R[1024])
{
  for (unsigned y = 0u; y < 16u; ++y)
    for (unsigned x = 0u; x < 32u; ++x) {
      F32(R, x, y) = G0[(size_t)y * ld + x];
      F32(R, x, y + 16u) = G1[(size_t)y * ld + x];
    }
  hjo_qr32(R);
  double P[1024];
  for (unsigned i = 32u; i < nrow; i += 32u) {
    for (unsigned y = 0u; y < 16u; ++y)
      for (unsigned x = 0u; x < 32u; ++x) {
        F32(P, x, y) = G0[(size_t)y * ld + i + x];
        F32(P, x, y + 16u) = G1[(size_t)y * ld + i + x];
      }
    hjo_qr32(P);
    hjo_peeloff(R, P);
  }
}

/* Drmac's sorted definite level-0, share/device_code_cdsort_Drmac.hpp:4-316
 * (posd: larger norm first; negd: smaller first).  R, W as in hjo_l0. */
void hjo_l0_drmac(double *R, double *W, int definite, const uint8_t *strat0,
                  unsigned s0steps, unsigned nswp, unsigned *blk_s_out,
                  unsigned *blk_b_out)
{
  const int pos = (definite >= 0);
  unsigned blk_s = 0u, blk_b = 0u;
  for (unsigned swp = 0u; swp < nswp; ++swp) {
    unsigned swp_s = 0u, swp_b = 0u;
    for (unsigned step = 0u; step < s0steps; ++step) {
      for (unsigned y = 0u; y < 16u; ++y) {
        const unsigned p = strat0[(step * 16u + y) * 2u + 0u];
        const unsigned q = strat0[(step * 16u + y) * 2u + 1u];
        double Gp[32], Gq[32], Vp[32], Vq[32];
        for (unsigned x = 0u; x < 32u; ++x) {
          Gp[x] = F32(R, x, p); Gq[x] = F32(R, x, q);
          Vp[x] = F32(W, x, p); Vq[x] = F32(W, x, q);
        }
        const double Ap = hjo_nrm2_32(Gp), Aq = hjo_nrm2_32(Gq);
        const double CosA = hjo_cosa(Ap, Aq, Gp, Gq);
        const int transf_s = (fabs(CosA) >= HJO_MYTOL);
        int transf_b = 0;
        if (transf_s) {
          const double Aq_p = Aq / Ap, Ap_q = Ap / Aq;
          const double ct2 = (Aq_p - Ap_q) / scalbn(CosA, 1);  /* tCot2 */
          if (ct2 <= nu_8()) {
            const double CT2 = fabs(ct2);
            double c, t;
            if (CT2 < HJO_SQRT_HEPS) {
              const double cot = CT2 + 1.0;
              t = copysign(1.0 / cot, ct2);
              c = hjo_rsqrt(fma(t, t, 1.0));
              transf_b = 1;
            }
            else if (CT2 >= HJO_SQRT_2_HEPS) {
              const double cot = scalbn(CT2, 1);
              c = 1.0;
              t = copysign(1.0 / cot, ct2);
            }
            else {
              const double cot = CT2 + sqrt(fma(CT2, CT2, 1.0));
              t = copysign(1.0 / cot, ct2);
              c = hjo_rsqrt(fma(t, t, 1.0));
              transf_b = (c < 1.0);
            }
            const double t_ = -t, tCosA = t * CosA, tCosA_ = -tCosA;
            const double Ap_ = Ap * sqrt(fma(tCosA_, Aq_p, 1.0));
            const double Aq_ = Aq * sqrt(fma(tCosA, Ap_q, 1.0));
            const int keep = pos ? (Ap_ >= Aq_) : (Ap_ <= Aq_);
            const unsigned P = keep ? p : q, Q = keep ? q : p;
            for (unsigned x = 0u; x < 32u; ++x) {
              const double rp = fma(t_, Gq[x], Gp[x]), rq = fma(t, Gp[x], Gq[x]);
              const double wp = fma(t_, Vq[x], Vp[x]), wq = fma(t, Vp[x], Vq[x]);
              F32(R, x, P) = transf_b ? c * rp : rp;
              F32(R, x, Q) = transf_b ? c * rq : rq;
              F32(W, x, P) = transf_b ? c * wp : wp;
              F32(W, x, Q) = transf_b ? c * wq : wq;
            }
          }
          else if (pos ? (Ap >= Aq) : (Ap <= Aq)) {
            /* |ct2| beyond NU_8: Gq made orthogonal to Gp directly */
            const double t_ = CosA, t = -t_;
            for (unsigned x = 0u; x < 32u; ++x) {
              const double gq = Aq * fma(t, Gp[x] / Ap, Gq[x] / Aq);
              const double wp = fma(t_, Vq[x], Vp[x]), wq = fma(t, Vp[x], Vq[x]);
              F32(R, x, q) = gq;
              F32(W, x, p) = wp;
              F32(W, x, q) = wq;
            }
          }
          else {
            const double t = CosA, t_ = -t;
            for (unsigned x = 0u; x < 32u; ++x) {
              const double gq = Ap * fma(t_, Gq[x] / Aq, Gp[x] / Ap);
              const double wq = fma(t_, Vq[x], Vp[x]), wp = fma(t, Vp[x], Vq[x]);
              F32(R, x, q) = gq;
              F32(R, x, p) = Gq[x];
              F32(W, x, q) = wq;
              F32(W, x, p) = wp;
            }
          }
        }
        else if (pos ? (Ap < Aq) : (Aq < Ap)) {
          for (unsigned x = 0u; x < 32u; ++x) {
            F32(R, x, p) = Gq[x]; F32(R, x, q) = Gp[x];
            F32(W, x, p) = Vq[x]; F32(W, x, q) = Vp[x];
          }
        }
        swp_s += (unsigned)transf_s;
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
}

/* one block pair of the QR path: R by hjo_factorize_qr, then the level-0
 * sweeps (Drmac's for sorted definite blocks when drmac, else the
 * reference's usual ones) */
unsigned hjo_pair_qr(const double *G0, const double *G1, unsigned ld, unsigned nrow,
                     int sort, int hyp, int definite, unsigned npos, int drmac,
                     const uint8_t *strat0, unsigned s0steps, unsigned nswp,
                     double W[1024], unsigned *big, int *upd)
{
  double R[1024];
  hjo_factorize_qr(G0, G1, ld, nrow, R);
  for (unsigned y = 0u; y < 32u; ++y)
    for (unsigned x = 0u; x < 32u; ++x)
      F32(W, x, y) = (x == y) ? 1.0 : 0.0;
  unsigned s = 0u, b = 0u;
  if (drmac && sort && !hyp)
    hjo_l0_drmac(R, W, definite, strat0, s0steps, nswp, &s, &b);
  else
    hjo_l0(R, W, sort, hyp, definite, npos, strat0, s0steps, nswp, &s, &b, NULL);
  *big = b;
  *upd = sort ? 1 : (s != 0u);
  return s;
}
