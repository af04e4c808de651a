// hj_qr.hip -- the QR factorization path of a block-pair visit and Drmac's
// sorted level-0 for gfx950 (the reference's build options USE_QR and
// USE_DRMAC, share/device_code_common.hpp:78-85,140-144).
//
// One workgroup of 512 lanes per block pair, in the reference's layout:
// lane x = tid % 32 is a row of the 32-row working blocks, y = tid / 32 a
// column pair (columns y and y + 16); each half of a 64-lane wave is one
// y-group, so the 32-element butterflies (16,8,4,2,1) stay inside a
// half-wave (ds_swizzle).  The visit:
//   factorize: R of the nrow x 32 column pair [G_i G_j] by 32-row panels in
//     order -- Householder QR of the first panel (dQR32,
//     share/device_code_common_Kepler.hpp:326-371), then per further panel
//     its Householder QR merged into R by a wavefront of Givens rotations
//     (dPeelOff / dFactorize, share/device_code_common_QR.hpp:4-94);
//   level-0: the reference's inner sweeps on R (8 bits of the dot products
//     identical to the Cholesky path's), or for sorted definite blocks with
//     DRMAC, Drmac's sweeps with overflow-safe norms and cosines
//     (share/device_code_cdsort_Drmac.hpp:4-316);
//   W to HBM; the block update is k_update's.
// This path is latency-bound (about 190 workgroup barriers per 32 rows of
// the pair); it exists, as in the reference, as an accuracy option for
// extremely graded column pairs, not for speed.
#include <climits>
#include <cfloat>

#include "hj_device.hpp"
#include "hj_kernels.hpp"

namespace hj {
namespace {

constexpr unsigned QLD = 33u;  // padded leading dimension of the LDS blocks
#define QA(A, r, c) (A)[(c) * QLD + (r)]

// MU_, NU_, MU_EPS, NU_8 of share/HypJacL.hpp:53-86
constexpr double QR_MU = 1.4916681462400414314624091158735E-154;
constexpr double QR_NU = 2.370187977027293260705088939015E+153;
constexpr double QR_MU_EPS = 1.0020841800044863889980540256751E-292;
constexpr double QR_NU_8 = 2.24711641857789464E+307;

// max / min of |x| over the 32 lanes of a half-wave (the butterfly of
// dMax32 / dMin32, share/device_code_common_Kepler.hpp:76-164)
__device__ __forceinline__ double max32(double x)
{
  double v = fabs(x);
  v = fmax(v, swz_xor<16>(v));
  v = fmax(v, swz_xor<8>(v));
  v = fmax(v, swz_xor<4>(v));
  v = fmax(v, swz_xor<2>(v));
  v = fmax(v, swz_xor<1>(v));
  return v;
}

__device__ __forceinline__ double min32_nz(double x)
{
  double v = (x == 0.0) ? DBL_MAX : fabs(x);
  v = fmin(v, swz_xor<16>(v));
  v = fmin(v, swz_xor<8>(v));
  v = fmin(v, swz_xor<4>(v));
  v = fmin(v, swz_xor<2>(v));
  v = fmin(v, swz_xor<1>(v));
  return v;
}

// share/device_code_common_scale.hpp:4-56
__device__ __forceinline__ int find_scl(double x, double t)
{
  int x_n, t_n;
  const double x_m = frexp(x, &x_n);
  const double t_m = frexp(t, &t_n);
  return (x <= t) ? ((t_n - x_n) + (x_m < t_m)) : ((t_n - x_n) - (x_m > t_m));
}

__device__ __forceinline__ void normalize2(int &scl, double &ssq)
{
  int s = ilogb(ssq);
  s = s + (s & 1);
  scl += s;
  ssq = scalbn(ssq, -s);
}

__device__ __forceinline__ void sum_scl(int s1, int s2, double x1, double x2, int &s, double &x)
{
  const int cmp = (s1 < s2) ? -1 : ((s2 < s1) ? 1 : ((x1 < x2) ? -1 : ((x2 < x1) ? 1 : 0)));
  if (cmp <= 0) {
    s = s2;
    x = scalbn(x1, s1 - s2) + x2;
  }
  else {
    s = s1;
    x = scalbn(x2, s2 - s1) + x1;
  }
}

// band 2: |x| > NU (scaled), 1: MU <= |x| <= NU, 0: |x| < MU (scaled)
__device__ __forceinline__ double ssq_band(double x, int band, int scl)
{
  const double x_ = fabs(x);
  double y;
  if (band == 2) y = (x_ > QR_NU) ? scalbn(x_, scl) : 0.0;
  else if (band == 1) y = ((x_ >= QR_MU) && (x_ <= QR_NU)) ? x_ : 0.0;
  else y = (x_ < QR_MU) ? scalbn(x_, scl) : 0.0;
  return sum32(y * y);
}

// dSSQ32, share/device_code_common_Kepler.hpp:203-260 (uniform per half-wave)
__device__ int ssq32(double &ssq, double x)
{
  const double M = max32(x);
  if (M == 0.0) {
    ssq = 0.0;
    return INT_MIN;
  }
  const double m = min32_nz(x);
  const bool lo = (m < QR_MU), hi = (M > QR_NU);
  int scl;
  if (hi) {
    int sc2 = find_scl(M, QR_NU);
    double sq2 = ssq_band(x, 2, sc2);
    sc2 *= -2;
    if (m > QR_NU) {
      scl = sc2;
      ssq = sq2;
    }
    else {
      int sc1 = 0;
      double sq1 = ssq_band(x, 1, 0);
      if (sq1 > 0.0) {
        normalize2(sc1, sq1);
        normalize2(sc2, sq2);
        sum_scl(sc1, sc2, sq1, sq2, scl, ssq);
      }
      else {
        scl = sc2;
        ssq = sq2;
      }
    }
  }
  else if (lo) {
    int sc0 = find_scl(m, QR_MU);
    double sq0 = ssq_band(x, 0, sc0);
    sc0 *= -2;
    if (M < QR_MU) {
      scl = sc0;
      ssq = sq0;
    }
    else {
      int sc1 = 0;
      double sq1 = ssq_band(x, 1, 0);
      if (sq1 > 0.0) {
        normalize2(sc1, sq1);
        normalize2(sc0, sq0);
        sum_scl(sc0, sc1, sq0, sq1, scl, ssq);
      }
      else {
        scl = sc0;
        ssq = sq0;
      }
    }
  }
  else {
    scl = 0;
    ssq = sum32(x * x);
  }
  return scl;
}

// dNRM2_32, share/device_code_common_Kepler.hpp:269-275
__device__ __forceinline__ double nrm2_32(double x)
{
  double ssq;
  const int scl = ssq32(ssq, x) / 2;
  return scalbn(sqrt(ssq), scl);
}

// dCOSA_ for Ap >= Aq, share/device_code_common_Kepler.hpp:277-303
__device__ __forceinline__ double cosa_(double Ap, double Aq, double x, double y)
{
  bool big, small;
  if (Aq >= 1.0) {
    big = ((Ap * Aq) >= DBL_MAX);
    small = false;
  }
  else {
    big = false;
    small = ((Ap * Aq) <= QR_MU_EPS);
  }
  if (big) {
    const double z = (fabs(x) >= fabs(y)) ? ((x / Ap) * y) : ((y / Ap) * x);
    return sum32(z) / Aq;
  }
  if (small) {
    const double z = (fabs(x) >= fabs(y)) ? (x * (y / Aq)) : (y * (x / Aq));
    return sum32(z) / Ap;
  }
  return (sum32(x * y) / Aq) / Ap;
}

__device__ __forceinline__ double cosa(double Ap, double Aq, double x, double y)
{
  return (Ap >= Aq) ? cosa_(Ap, Aq, x, y) : cosa_(Aq, Ap, y, x);
}

// dGivens, share/device_code_common_rotate.hpp:120-155
__device__ __forceinline__ void givens(double f, double g, double &c, double &s, double &r)
{
  const double f_ = fabs(f), g_ = fabs(g);
  if (f_ >= g_) {
    if (g_ == 0.0) {
      c = copysign(1.0, f);
      s = 0.0;
      r = f_;
    }
    else {
      const double g_f = g_ / f_;
      r = f_ * sqrt(__fma_rn(g_f, g_f, 1.0));
      c = f / r;
      s = g / r;
    }
  }
  else {
    if (f_ == 0.0) {
      c = 0.0;
      s = copysign(1.0, g);
      r = g_;
    }
    else {
      const double f_g = f_ / g_;
      r = g_ * sqrt(__fma_rn(f_g, f_g, 1.0));
      c = f / r;
      s = g / r;
    }
  }
}

// one Householder application to column col at step k (reflector v with
// v_k = 1, v_x = A(x,k) / (alpha + beta) below, scale tau)
__device__ __forceinline__ void hh_apply(double *A, unsigned x, unsigned k, unsigned col,
                                         double Axk_, double alpha_beta, double tau)
{
  const double Axy = QA(A, x, col);
  const double Vxk = (x == k) ? 1.0 : ((x > k) ? (Axk_ / alpha_beta) : 0.0);
  const double dp = sum32(Vxk * Axy);
  const double tdp = -(tau * dp);
  QA(A, x, col) = __fma_rn(tdp, Vxk, Axy);
}

// dQR32: Householder QR of the 32x32 block A in LDS, in place
__device__ void qr32(double *A, unsigned x, unsigned y0)
{
  const unsigned y1 = y0 + 16u;
#pragma unroll 1
  for (unsigned k = 0u; k < 31u; ++k) {
    double nrm2 = 0.0, Axk = 0.0, Axk_ = 0.0, beta = 0.0, alpha_beta = 0.0, tau = 0.0;
    const unsigned my0 = k + y0;
    if (my0 < 32u) {
      Axk = QA(A, x, k);
      Axk_ = (x >= k) ? Axk : 0.0;
      nrm2 = nrm2_32(Axk_);
      if (nrm2 > 0.0) {
        const double alpha = QA(A, k, k);
        beta = copysign(nrm2, alpha);
        alpha_beta = alpha + beta;
        tau = alpha_beta / beta;
      }
    }
    __syncthreads();
    if (nrm2 > 0.0) {
      if (my0 == k)
        QA(A, x, k) = (x == k) ? -beta : ((x > k) ? 0.0 : Axk);
      else
        hh_apply(A, x, k, my0, Axk_, alpha_beta, tau);
      const unsigned my1 = k + y1;
      if (my1 < 32u)
        hh_apply(A, x, k, my1, Axk_, alpha_beta, tau);
    }
    __syncthreads();
  }
}

// dPeelOff: R0 <- R of [R0; R1] by Givens rotations, row x of R0 meeting row
// x - k of R1 at step k
__device__ void peeloff(double *R0, double *R1, unsigned x, unsigned y0)
{
  const unsigned y1 = y0 + 16u;
#pragma unroll 1
  for (unsigned k = 0u; k < 32u; ++k) {
    const unsigned my0 = k + y0, my1 = k + y1;
    const bool active = (x >= k) && (x <= my1);
    const unsigned xk = x - k;
    double c = 0.0, s = 0.0, r = 0.0, a_ = 0.0, b_ = 0.0;
    bool store = false;
    if ((my0 < 32u) && active) {
      givens(QA(R0, x, x), QA(R1, xk, x), c, s, r);
      if (x < my0) {
        const double a = QA(R0, x, my0), b = QA(R1, xk, my0);
        const double ca = c * a, sb = s * b, cb = c * b, sa = s * a;
        a_ = ca + sb;
        b_ = cb - sa;
        store = true;
      }
      else if (x == my0) {
        a_ = r;
        b_ = 0.0;
        store = true;
      }
    }
    __syncthreads();
    if (store) {
      QA(R0, x, my0) = a_;
      QA(R1, xk, my0) = b_;
      store = false;
    }
    __syncthreads();
    if ((my1 < 32u) && active) {
      if (x < my1) {
        const double a = QA(R0, x, my1), b = QA(R1, xk, my1);
        const double ca = c * a, sb = s * b, cb = c * b, sa = s * a;
        a_ = ca + sb;
        b_ = cb - sa;
        store = true;
      }
      else if (x == my1) {
        a_ = r;
        b_ = 0.0;
        store = true;
      }
    }
    __syncthreads();
    if (store) {
      QA(R0, x, my1) = a_;
      QA(R1, xk, my1) = b_;
    }
    __syncthreads();
  }
}

// sum over the 16 y-groups of a per-group count (lane x == 0 of each group
// holds it); every lane gets the total
__device__ __forceinline__ unsigned group_total(unsigned v, unsigned x, unsigned y, unsigned *buf)
{
  if (x == 0u) buf[y] = v;
  __syncthreads();
  unsigned t = 0u;
#pragma unroll
  for (unsigned i = 0u; i < 16u; ++i) t += buf[i];
  __syncthreads();
  return t;
}

// The reference's level-0 sweeps (share/device_code_nosort.hpp:4-222,
// share/device_code_cdsort_KISS.hpp:4-260, share/device_code_cdsort.hpp:25-167)
// with 32 lanes per pivot pair: the same operations as the Cholesky path's
// 8-lane version (the 32-element butterfly is the same 16,8,4,2,1 order).
template <bool SORT>
__device__ void l0_plain(double *R, double *W, const uint8_t *S0, unsigned s0steps, unsigned nswp,
                         int def, bool hyp, unsigned npos, unsigned x, unsigned y,
                         unsigned *cbuf, unsigned &out_s, unsigned &out_b)
{
  unsigned blk_s = 0u, blk_b = 0u;
  for (unsigned swp = 0u; swp < nswp; ++swp) {
    unsigned my_s = 0u, my_b = 0u;
    for (unsigned step = 0u; step < s0steps; ++step) {
      const unsigned p = S0[(step * 16u + y) * 2u], q = S0[(step * 16u + y) * 2u + 1u];
      const double Gp = QA(R, x, p), Gq = QA(R, x, q), Vp = QA(W, x, p), Vq = QA(W, x, q);
      double Dp = sum32(__fma_rn(Gp, Gp, 0.0));
      double Dq = sum32(__fma_rn(Gq, Gq, 0.0));
      const double Apq = sum32(__fma_rn(Gp, Gq, 0.0));
      const double Dp_ = sqrt(Dp), Dq_ = sqrt(Dq), Apq_ = fabs(Apq), Dpq_ = Dp_ * Dq_;
      const int transf_s = !(Apq_ < (Dpq_ * MYTOL));
      int pdef = def;
      if (hyp) {
        if ((p < npos) && (q < npos)) pdef = +1;
        else if ((p >= npos) && (q >= npos)) pdef = -1;
        else pdef = 0;
      }
      int transf_b = 0;
      if (transf_s) {
        double c, t;
        const bool trig = SORT ? (pdef != 0) : (!hyp || pdef);
        transf_b = dev_rot(trig, Apq, Dp, Dq, c, t);
        const double t_ = trig ? -t : t;
        if (SORT && trig) {
          Dp = __fma_rn(t, Apq, Dp);
          Dq = __fma_rn(t_, Apq, Dq);
        }
        bool keep = true;
        if (SORT)
          keep = (!pdef) || ((pdef > 0) && (Dp >= Dq)) || ((pdef < 0) && (Dp <= Dq));
        const unsigned P = keep ? p : q, Q = keep ? q : p;
        const double rp = __fma_rn(t_, Gq, Gp), rq = __fma_rn(t, Gp, Gq);
        const double wp = __fma_rn(t_, Vq, Vp), wq = __fma_rn(t, Vp, Vq);
        QA(R, x, P) = transf_b ? c * rp : rp;
        QA(R, x, Q) = transf_b ? c * rq : rq;
        QA(W, x, P) = transf_b ? c * wp : wp;
        QA(W, x, Q) = transf_b ? c * wq : wq;
      }
      else if (SORT && (((pdef > 0) && (Dp < Dq)) || ((pdef < 0) && (Dq < Dp)))) {
        QA(R, x, p) = Gq;
        QA(R, x, q) = Gp;
        QA(W, x, p) = Vq;
        QA(W, x, q) = Vp;
      }
      my_s += (unsigned)transf_s;
      my_b += (unsigned)transf_b;
      __syncthreads();
    }
    const unsigned swp_s = group_total(my_s, x, y, cbuf);
    const unsigned swp_b = group_total(my_b, x, y, cbuf + 16);
    if (swp_s) {
      blk_s += swp_s;
      blk_b += swp_b;
    }
    else
      break;
  }
  out_s = blk_s;
  out_b = blk_b;
}

// Drmac's sorted definite level-0 (posd for definite > 0, negd for < 0),
// share/device_code_cdsort_Drmac.hpp:4-316
__device__ void l0_drmac(double *R, double *W, const uint8_t *S0, unsigned s0steps, unsigned nswp,
                         int definite, unsigned x, unsigned y, unsigned *cbuf, unsigned &out_s,
                         unsigned &out_b)
{
  const bool pos = (definite >= 0);
  unsigned blk_s = 0u, blk_b = 0u;
  for (unsigned swp = 0u; swp < nswp; ++swp) {
    unsigned my_s = 0u, my_b = 0u;
    for (unsigned step = 0u; step < s0steps; ++step) {
      const unsigned p = S0[(step * 16u + y) * 2u], q = S0[(step * 16u + y) * 2u + 1u];
      const double Gp = QA(R, x, p), Gq = QA(R, x, q), Vp = QA(W, x, p), Vq = QA(W, x, q);
      const double Ap = nrm2_32(Gp), Aq = nrm2_32(Gq);
      const double CosA = cosa(Ap, Aq, Gp, Gq);
      const int transf_s = (fabs(CosA) >= MYTOL);
      int transf_b = 0;
      if (transf_s) {
        const double Aq_p = Aq / Ap, Ap_q = Ap / Aq;
        const double ct2 = (Aq_p - Ap_q) / scalbn(CosA, 1);  // tCot2
        if (ct2 <= QR_NU_8) {
          const double CT2 = fabs(ct2);
          double c, t;
          if (CT2 < SQRT_HEPS) {
            const double cot = CT2 + 1.0;
            t = copysign(1.0 / cot, ct2);
            c = dev_rsqrt(__fma_rn(t, t, 1.0));
            transf_b = 1;
          }
          else if (CT2 >= SQRT_2_HEPS) {
            const double cot = scalbn(CT2, 1);
            c = 1.0;
            t = copysign(1.0 / cot, ct2);
          }
          else {
            const double cot = CT2 + sqrt(__fma_rn(CT2, CT2, 1.0));
            t = copysign(1.0 / cot, ct2);
            c = dev_rsqrt(__fma_rn(t, t, 1.0));
            transf_b = (c < 1.0);
          }
          const double t_ = -t, tCosA = t * CosA, tCosA_ = -tCosA;
          const double Ap_ = Ap * sqrt(__fma_rn(tCosA_, Aq_p, 1.0));
          const double Aq_ = Aq * sqrt(__fma_rn(tCosA, Ap_q, 1.0));
          const bool keep = pos ? (Ap_ >= Aq_) : (Ap_ <= Aq_);
          const unsigned P = keep ? p : q, Q = keep ? q : p;
          const double rp = __fma_rn(t_, Gq, Gp), rq = __fma_rn(t, Gp, Gq);
          const double wp = __fma_rn(t_, Vq, Vp), wq = __fma_rn(t, Vp, Vq);
          QA(R, x, P) = transf_b ? c * rp : rp;
          QA(R, x, Q) = transf_b ? c * rq : rq;
          QA(W, x, P) = transf_b ? c * wp : wp;
          QA(W, x, Q) = transf_b ? c * wq : wq;
        }
        else if (pos ? (Ap >= Aq) : (Ap <= Aq)) {
          const double t_ = CosA, t = -t_;
          QA(R, x, q) = Aq * __fma_rn(t, Gp / Ap, Gq / Aq);
          QA(W, x, p) = __fma_rn(t_, Vq, Vp);
          QA(W, x, q) = __fma_rn(t, Vp, Vq);
        }
        else {
          const double t = CosA, t_ = -t;
          QA(R, x, q) = Ap * __fma_rn(t_, Gq / Aq, Gp / Ap);
          QA(R, x, p) = Gq;
          QA(W, x, q) = __fma_rn(t_, Vq, Vp);
          QA(W, x, p) = __fma_rn(t, Vp, Vq);
        }
      }
      else if (pos ? (Ap < Aq) : (Aq < Ap)) {
        QA(R, x, p) = Gq;
        QA(R, x, q) = Gp;
        QA(W, x, p) = Vq;
        QA(W, x, q) = Vp;
      }
      my_s += (unsigned)transf_s;
      my_b += (unsigned)transf_b;
      __syncthreads();
    }
    const unsigned swp_s = group_total(my_s, x, y, cbuf);
    const unsigned swp_b = group_total(my_b, x, y, cbuf + 16);
    if (swp_s) {
      blk_s += swp_s;
      blk_b += swp_b;
    }
    else
      break;
  }
  out_s = blk_s;
  out_b = blk_b;
}

// One block pair per workgroup: the QR factorization, the level-0 sweeps,
// the convergence counter, the update flag, W.
template <bool SORT, bool DRMAC>
__global__ __launch_bounds__(512) void k_pair_qr(const double *__restrict__ G, unsigned ldG,
                                                 unsigned nrow, const uint16_t *__restrict__ pairs,
                                                 double *__restrict__ Wout,
                                                 unsigned *__restrict__ flags, RunCtl *__restrict__ ctl,
                                                 const uint8_t *__restrict__ strat0, unsigned s0steps,
                                                 unsigned nswp, int definite, unsigned nplus)
{
  __shared__ double A[32u * QLD];
  __shared__ double P[32u * QLD];
  __shared__ __attribute__((aligned(16))) uint8_t S0[STRAT0_TAB_BYTES];
  __shared__ unsigned cbuf[32];
  const unsigned tid = threadIdx.x, x = tid & 31u, y = tid >> 5;
  const unsigned pi = blockIdx.x;
  const unsigned iBlk = pairs[2u * pi], jBlk = pairs[2u * pi + 1u];
  if (tid < STRAT0_TAB_BYTES / 16u)
    reinterpret_cast<uint4 *>(S0)[tid] = reinterpret_cast<const uint4 *>(strat0)[tid];

  // local J structure, share/device_code_nosort_accumV.hpp:53-76
  bool hyp = false;
  int def = definite;
  unsigned npos = 0u;
  if (!definite) {
    const unsigned ifc0 = iBlk * 16u, ifc1 = jBlk * 16u;
    if (nplus <= ifc0) { npos = 0u; def = -1; }
    else if (nplus <= ifc0 + 16u) { npos = nplus - ifc0; def = 0; }
    else if (nplus <= ifc1) { npos = 16u; def = 0; }
    else if (nplus <= ifc1 + 16u) { npos = 16u + (nplus - ifc1); def = 0; }
    else { npos = 32u; def = +1; }
    hyp = (def == 0);
  }

  const double *const G0 = G + (size_t)(16u * iBlk + y) * ldG;
  const double *const G1 = G + (size_t)(16u * jBlk + y) * ldG;
  QA(A, x, y) = G0[x];
  QA(A, x, y + 16u) = G1[x];
  __syncthreads();
  qr32(A, x, y);
#pragma unroll 1
  for (unsigned i = 32u; i < nrow; i += 32u) {
    QA(P, x, y) = G0[i + x];
    QA(P, x, y + 16u) = G1[i + x];
    __syncthreads();
    qr32(P, x, y);
    peeloff(A, P, x, y);
  }
  QA(P, x, y) = (x == y) ? 1.0 : 0.0;
  QA(P, x, y + 16u) = (x == y + 16u) ? 1.0 : 0.0;
  __syncthreads();

  unsigned blk_s = 0u, blk_b = 0u;
  if (DRMAC && SORT && !hyp)
    l0_drmac(A, P, S0, s0steps, nswp, def, x, y, cbuf, blk_s, blk_b);
  else
    l0_plain<SORT>(A, P, S0, s0steps, nswp, def, hyp, npos, x, y, cbuf, blk_s, blk_b);

  if (tid == 0u) {
    // convergence counter, share/device_code_nosort.hpp:94-103
    unsigned long long *const cvg = &ctl->cvg;
    if (blk_s) {
      if (blk_b)
        atomicAdd(cvg, (((unsigned long long)blk_b) << 32) | blk_s);
      else
        atomicAdd(reinterpret_cast<unsigned *>(cvg), blk_s);
    }
    flags[pi] = SORT ? 1u : (blk_s ? 1u : 0u);
  }
  double *const wo = Wout + (size_t)pi * 1024u;
  wo[y * 32u + x] = QA(P, x, y);
  wo[(y + 16u) * 32u + x] = QA(P, x, y + 16u);
}

}  // namespace

void launch_pair_qr(hipStream_t s, bool sort, bool drmac, const double *G, unsigned ldG,
                    unsigned nrow, const uint16_t *pairs, unsigned np, double *W, unsigned *flags,
                    RunCtl *ctl, const uint8_t *strat0, unsigned s0steps, unsigned nswp,
                    int definite, unsigned nplus)
{
  if (sort && drmac)
    hipLaunchKernelGGL((k_pair_qr<true, true>), dim3(np), dim3(512), 0, s, G, ldG, nrow, pairs, W,
                       flags, ctl, strat0, s0steps, nswp, definite, nplus);
  else if (sort)
    hipLaunchKernelGGL((k_pair_qr<true, false>), dim3(np), dim3(512), 0, s, G, ldG, nrow, pairs, W,
                       flags, ctl, strat0, s0steps, nswp, definite, nplus);
  else
    hipLaunchKernelGGL((k_pair_qr<false, false>), dim3(np), dim3(512), 0, s, G, ldG, nrow, pairs, W,
                       flags, ctl, strat0, s0steps, nswp, definite, nplus);
}

}  // namespace hj
