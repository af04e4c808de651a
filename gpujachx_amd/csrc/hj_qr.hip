// hj_qr.hip -- the QR factorization of a block-pair visit and Drmac's sorted
// level-0 (the reference's build options USE_QR and USE_DRMAC,
// share/device_code_common.hpp:78-85,140-144), one wave per block pair.
//
// Layout.  A 64-lane wave owns one pair.  Its LDS holds two 32x32 blocks,
// R (the running triangular factor) and P (the panel being reduced; W once
// the factorization is done), column-major with leading dimension 33: the
// L0A layout of the Cholesky path, so that path's wave_level0 runs on them
// unchanged.  Lanes form 16 quads; lane r of quad q holds rows r, r+4, ...,
// r+28 of whatever column the quad works on, so a 32-row reduction is the
// butterfly 16,8,4,2,1 with levels 16, 8, 4 inside the lane and 2, 1 by
// quad DPP (sum4x8) -- the bits of the reference's 32-lane dSum32.  A single
// wave needs no workgroup barrier: its LDS accesses complete in program
// order, and wave_lds_sync() only keeps the compiler from moving them from
// one phase into the next.
//
// The visit:
//  * factorization by 32-row panels in order (dFactorize,
//    share/device_code_common_QR.hpp:68-92): Householder QR of the first
//    panel into R; for each further panel, its Householder QR into P, then
//    R <- R of [R; P] by a wavefront of Givens rotations (dPeelOff, :4-66).
//    A panel's global loads are issued before the previous panel's
//    reduction, so they arrive while it runs.
//      Householder step k (dQR32, share/device_code_common_Kepler.hpp:326-371):
//      every quad forms the overflow-safe norm of column k itself (no
//      broadcast), then quad q reflects columns k+1+q and k+17+q.
//      Givens step k: lanes 0..31 each form the rotation of one row x >= k
//      (with row x-k of P) into LDS; quad q then rotates columns k+q and
//      k+16+q.
//  * level-0: Drmac's sweeps for sorted definite blocks when asked
//    (share/device_code_cdsort_Drmac.hpp:4-316), quad q taking pivot pair q
//    of each step; otherwise the Cholesky path's wave_level0.
//  * counters, the update flag and W to HBM; the block update is k_update's.
// This path is latency-bound; like the reference's, it is an accuracy option
// for extremely graded column pairs, not a fast path.
#include <climits>
#include <cfloat>

#include "hj_device.hpp"
#include "hj_kernels.hpp"

namespace hj {
namespace {

// Magnitude limits of the overflow- and underflow-safe sums of squares
// (share/HypJacL.hpp:53-86).  Squares of entries in [SSQ_SMALL, SSQ_BIG]
// and 32-term sums of them stay normal and finite.
constexpr double SSQ_SMALL = 0x1p-511;              // sqrt of the smallest normal
constexpr double SSQ_BIG = 0x1.6a09e667f3bcap+509;  // a little below sqrt(DBL_MAX / 32)
constexpr double PROD_SMALL = 0x1p-970;             // smallest normal / epsilon
constexpr double COT_LIMIT = 0x1.fffffffffffffp+1020;  // DBL_MAX / 8

// ---------------------------------------------------------------------------
// Reductions over the 32 rows of one column held by a quad, v[j] = row r + 4j.

__device__ __forceinline__ double quad_max_abs(const double (&v)[8])
{
  double m = fabs(v[0]);
#pragma unroll
  for (unsigned j = 1u; j < 8u; ++j) m = fmax(m, fabs(v[j]));
  m = fmax(m, lane_xor2(m));
  return fmax(m, lane_xor1(m));
}

// smallest nonzero magnitude (DBL_MAX when every entry is zero)
__device__ __forceinline__ double quad_min_abs_nz(const double (&v)[8])
{
  double m = DBL_MAX;
#pragma unroll
  for (unsigned j = 0u; j < 8u; ++j) m = fmin(m, (v[j] == 0.0) ? DBL_MAX : fabs(v[j]));
  m = fmin(m, lane_xor2(m));
  return fmin(m, lane_xor1(m));
}

// The power of two that moves x next to the limit t: for x <= t the least s
// with x 2^s >= t, for x > t the greatest s with x 2^s <= t.
__device__ __forceinline__ int scale_to(double x, double t)
{
  int ex, et;
  const double mx = frexp(x, &ex), mt = frexp(t, &et);
  const int s = et - ex;
  return (x <= t) ? s + (mx < mt ? 1 : 0) : s - (mx > mt ? 1 : 0);
}

// A scaled value f 2^e with f brought to an even exponent's power of two
// removed (f in [1, 4)), so that the square root rescales exactly.
__device__ __forceinline__ void even_normalize(int &e, double &f)
{
  int s = ilogb(f);
  s += s & 1;
  e += s;
  f = scalbn(f, -s);
}

// (e, f) = f1 2^e1 + f2 2^e2: the smaller term, by exponent then fraction,
// rescaled to the larger one's exponent.
__device__ __forceinline__ void add_scaled(int e1, double f1, int e2, double f2, int &e, double &f)
{
  const bool first_smaller = (e1 != e2) ? (e1 < e2) : !(f2 < f1);
  if (first_smaller) {
    e = e2;
    f = scalbn(f1, e1 - e2) + f2;
  }
  else {
    e = e1;
    f = scalbn(f2, e2 - e1) + f1;
  }
}

// Sum of squares of the entries in one magnitude band, each scaled by 2^e
// first (outer bands) or taken as is (middle band).
enum Band { BAND_LOW = 0, BAND_MID = 1, BAND_HIGH = 2 };

__device__ __forceinline__ double band_ssq(const double (&v)[8], Band b, int e)
{
  double sq[8];
#pragma unroll
  for (unsigned j = 0u; j < 8u; ++j) {
    const double a = fabs(v[j]);
    double y;
    if (b == BAND_HIGH) y = (a > SSQ_BIG) ? scalbn(a, e) : 0.0;
    else if (b == BAND_MID) y = ((a >= SSQ_SMALL) && (a <= SSQ_BIG)) ? a : 0.0;
    else y = (a < SSQ_SMALL) ? scalbn(a, e) : 0.0;
    sq[j] = y * y;
  }
  return sum4x8(sq);
}

// The sum of squares of the column as f 2^e (returned), without overflow
// or harmful underflow: entries above SSQ_BIG or below SSQ_SMALL are summed
// in a band of their own, scaled toward the limit, and the band sums are
// combined by add_scaled (dSSQ32, share/device_code_common_Kepler.hpp:203-260).
// e is INT_MIN for a zero column.
__device__ int quad_ssq(const double (&v)[8], double &f)
{
  const double hi = quad_max_abs(v);
  if (hi == 0.0) {
    f = 0.0;
    return INT_MIN;
  }
  const double lo = quad_min_abs_nz(v);
  if (hi > SSQ_BIG) {
    int eh = scale_to(hi, SSQ_BIG);
    double fh = band_ssq(v, BAND_HIGH, eh);
    eh = -2 * eh;
    if (lo > SSQ_BIG) {
      f = fh;
      return eh;
    }
    int em = 0;
    double fm = band_ssq(v, BAND_MID, 0);
    if (!(fm > 0.0)) {
      f = fh;
      return eh;
    }
    even_normalize(em, fm);
    even_normalize(eh, fh);
    int e;
    add_scaled(em, fm, eh, fh, e, f);
    return e;
  }
  if (lo < SSQ_SMALL) {
    int el = scale_to(lo, SSQ_SMALL);
    double fl = band_ssq(v, BAND_LOW, el);
    el = -2 * el;
    if (hi < SSQ_SMALL) {
      f = fl;
      return el;
    }
    int em = 0;
    double fm = band_ssq(v, BAND_MID, 0);
    if (!(fm > 0.0)) {
      f = fl;
      return el;
    }
    even_normalize(em, fm);
    even_normalize(el, fl);
    int e;
    add_scaled(el, fl, em, fm, e, f);
    return e;
  }
  double sq[8];
#pragma unroll
  for (unsigned j = 0u; j < 8u; ++j) sq[j] = v[j] * v[j];
  f = sum4x8(sq);
  return 0;
}

// ||column||_2 (dNRM2_32, share/device_code_common_Kepler.hpp:269-275)
__device__ __forceinline__ double quad_nrm2(const double (&v)[8])
{
  double f;
  const int e = quad_ssq(v, f);
  return scalbn(sqrt(f), e / 2);
}

// x^T y / (nx ny) for nx >= ny, the division placed so that neither the
// product nor the quotient leaves the normal range
// (dCOSA_, share/device_code_common_Kepler.hpp:277-303)
__device__ __forceinline__ double quad_cos_ordered(double nx, double ny, const double (&x)[8],
                                                   const double (&y)[8])
{
  const bool huge = (ny >= 1.0) && ((nx * ny) >= DBL_MAX);
  const bool tiny = !(ny >= 1.0) && ((nx * ny) <= PROD_SMALL);
  double z[8];
  if (huge) {
#pragma unroll
    for (unsigned j = 0u; j < 8u; ++j)
      z[j] = (fabs(x[j]) >= fabs(y[j])) ? ((x[j] / nx) * y[j]) : ((y[j] / nx) * x[j]);
    return sum4x8(z) / ny;
  }
  if (tiny) {
#pragma unroll
    for (unsigned j = 0u; j < 8u; ++j)
      z[j] = (fabs(x[j]) >= fabs(y[j])) ? (x[j] * (y[j] / ny)) : (y[j] * (x[j] / ny));
    return sum4x8(z) / nx;
  }
#pragma unroll
  for (unsigned j = 0u; j < 8u; ++j) z[j] = x[j] * y[j];
  return (sum4x8(z) / ny) / nx;
}

__device__ __forceinline__ double quad_cos(double np, double nq, const double (&p)[8],
                                           const double (&q)[8])
{
  return (np >= nq) ? quad_cos_ordered(np, nq, p, q) : quad_cos_ordered(nq, np, q, p);
}

// ---------------------------------------------------------------------------
// Factorization.

// [c s; -s c] [f; g] = [r; 0] with r >= 0, r formed as the larger magnitude
// times sqrt(1 + (smaller / larger)^2) (dGivens,
// share/device_code_common_rotate.hpp:120-155)
__device__ __forceinline__ void givens(double f, double g, double &c, double &s, double &r)
{
  const double af = fabs(f), ag = fabs(g);
  const bool fbig = af >= ag;
  const double big = fbig ? af : ag, small = fbig ? ag : af;
  if (small == 0.0) {
    r = big;
    c = fbig ? copysign(1.0, f) : 0.0;
    s = fbig ? 0.0 : copysign(1.0, g);
    return;
  }
  const double q = small / big;
  r = big * sqrt(__fma_rn(q, q, 1.0));
  c = f / r;
  s = g / r;
}

// Householder QR of the 32x32 block A in place: for k = 0..30 the reflector
// I - tau v v^T with v_k = 1, v_x = A(x,k) / (alpha + beta) below the
// diagonal, beta = sign(alpha) ||A(k:31,k)||, applied to columns k+1..31 as
// A(:,c) += (-(tau v^T A(:,c))) v; column k becomes (-beta, 0, ..., 0).
__device__ void householder32(double *A, unsigned q, unsigned r)
{
#pragma unroll 1
  for (unsigned k = 0u; k < 31u; ++k) {
    double v[8];
#pragma unroll
    for (unsigned j = 0u; j < 8u; ++j) {
      const unsigned x = r + 4u * j;
      const double a = L0A(A, x, k);
      v[j] = (x >= k) ? a : 0.0;
    }
    const double nrm = quad_nrm2(v);
    if (!(nrm > 0.0))
      continue;  // the same in every lane: every quad holds column k
    const double alpha = L0A(A, k, k);
    // column k has been read by every lane before any lane writes it
    wave_lds_sync();
    const double beta = copysign(nrm, alpha), ab = alpha + beta, tau = ab / beta;
#pragma unroll
    for (unsigned j = 0u; j < 8u; ++j) {
      const unsigned x = r + 4u * j;
      v[j] = (x == k) ? 1.0 : ((x > k) ? v[j] / ab : 0.0);
    }
    for (unsigned c = k + 1u + q; c < 32u; c += 16u) {
      double a[8], pr[8];
#pragma unroll
      for (unsigned j = 0u; j < 8u; ++j) {
        a[j] = L0A(A, r + 4u * j, c);
        pr[j] = v[j] * a[j];
      }
      const double tdp = -(tau * sum4x8(pr));
#pragma unroll
      for (unsigned j = 0u; j < 8u; ++j)
        L0A(A, r + 4u * j, c) = __fma_rn(tdp, v[j], a[j]);
    }
    if (q == 15u) {
#pragma unroll
      for (unsigned j = 0u; j < 8u; ++j) {
        const unsigned x = r + 4u * j;
        if (x >= k) L0A(A, x, k) = (x == k) ? -beta : 0.0;
      }
    }
    wave_lds_sync();
  }
}

// R0 <- the R factor of [R0; R1] for upper triangular R0, R1: at step k,
// row x >= k of R0 meets row x-k of R1 in the rotation that zeroes R1(x-k, x)
// against R0(x, x), applied to columns x..31 (products and sums rounded
// separately).  gv: 96 doubles of LDS for the step's c, s, r.
__device__ void merge_rows(double *R0, double *R1, double *gv, unsigned l, unsigned q, unsigned r)
{
#pragma unroll 1
  for (unsigned k = 0u; k < 32u; ++k) {
    if ((l < 32u) && (l >= k)) {
      double c, s, rr;
      givens(L0A(R0, l, l), L0A(R1, l - k, l), c, s, rr);
      gv[l] = c;
      gv[32u + l] = s;
      gv[64u + l] = rr;
    }
    wave_lds_sync();
    for (unsigned col = k + q; col < 32u; col += 16u) {
#pragma unroll
      for (unsigned j = 0u; j < 8u; ++j) {
        const unsigned x = r + 4u * j;
        if ((x < k) || (x > col)) continue;
        if (x < col) {
          const double c = gv[x], s = gv[32u + x];
          const double a = L0A(R0, x, col), b = L0A(R1, x - k, col);
          const double ca = c * a, sb = s * b, cb = c * b, sa = s * a;
          L0A(R0, x, col) = ca + sb;
          L0A(R1, x - k, col) = cb - sa;
        }
        else {
          L0A(R0, x, col) = gv[64u + x];
          L0A(R1, x - k, col) = 0.0;
        }
      }
    }
    wave_lds_sync();
  }
}

// ---------------------------------------------------------------------------
// Drmac's sorted definite level-0 (posd for definite >= 0: the larger norm
// first; negd: the smaller first), share/device_code_cdsort_Drmac.hpp:4-316.
// The cosine of the pivot columns and their norms are formed overflow-safe;
// the rotation comes from cot(2 phi) = (nq/np - np/nq) / (2 cos), and where
// that exceeds COT_LIMIT the smaller column is made orthogonal to the larger
// one directly.  Quad q takes pivot pair q of a step.
__device__ void drmac_level0(double *R, double *W, const uint8_t *S0, unsigned s0steps,
                             unsigned nswp, int definite, unsigned q, unsigned r,
                             unsigned &out_s, unsigned &out_b)
{
  const bool pos = (definite >= 0);
  unsigned blk_s = 0u, blk_b = 0u;
  for (unsigned swp = 0u; swp < nswp; ++swp) {
    unsigned my_s = 0u, my_b = 0u;
    for (unsigned step = 0u; step < s0steps; ++step) {
      const unsigned p = S0[(step * 16u + q) * 2u], o = S0[(step * 16u + q) * 2u + 1u];
      double Gp[8], Go[8], Vp[8], Vo[8];
#pragma unroll
      for (unsigned j = 0u; j < 8u; ++j) {
        Gp[j] = L0A(R, r + 4u * j, p);
        Go[j] = L0A(R, r + 4u * j, o);
        Vp[j] = L0A(W, r + 4u * j, p);
        Vo[j] = L0A(W, r + 4u * j, o);
      }
      const double Ap = quad_nrm2(Gp), Ao = quad_nrm2(Go);
      const double cs = quad_cos(Ap, Ao, Gp, Go);
      const int rot = (fabs(cs) >= MYTOL);
      int big = 0;
      if (rot) {
        const double o_p = Ao / Ap, p_o = Ap / Ao;
        const double ct2 = (o_p - p_o) / scalbn(cs, 1);
        if (ct2 <= COT_LIMIT) {
          const double act = fabs(ct2);
          double c, t;
          if (act < SQRT_HEPS) {
            t = copysign(1.0 / (act + 1.0), ct2);
            c = dev_rsqrt(__fma_rn(t, t, 1.0));
            big = 1;
          }
          else if (act >= SQRT_2_HEPS) {
            t = copysign(1.0 / scalbn(act, 1), ct2);
            c = 1.0;
          }
          else {
            t = copysign(1.0 / (act + sqrt(__fma_rn(act, act, 1.0))), ct2);
            c = dev_rsqrt(__fma_rn(t, t, 1.0));
            big = (c < 1.0);
          }
          const double tc = t * cs;
          // the norms after the rotation decide which column goes first
          const double Ap2 = Ap * sqrt(__fma_rn(-tc, o_p, 1.0));
          const double Ao2 = Ao * sqrt(__fma_rn(tc, p_o, 1.0));
          const bool keep = pos ? (Ap2 >= Ao2) : (Ap2 <= Ao2);
          const unsigned P1 = keep ? p : o, P2 = keep ? o : p;
#pragma unroll
          for (unsigned j = 0u; j < 8u; ++j) {
            const double rp = __fma_rn(-t, Go[j], Gp[j]), ro = __fma_rn(t, Gp[j], Go[j]);
            const double wp = __fma_rn(-t, Vo[j], Vp[j]), wo = __fma_rn(t, Vp[j], Vo[j]);
            L0A(R, r + 4u * j, P1) = big ? c * rp : rp;
            L0A(R, r + 4u * j, P2) = big ? c * ro : ro;
            L0A(W, r + 4u * j, P1) = big ? c * wp : wp;
            L0A(W, r + 4u * j, P2) = big ? c * wo : wo;
          }
        }
        else if (pos ? (Ap >= Ao) : (Ap <= Ao)) {
          // column o loses its component along p; p stays first
#pragma unroll
          for (unsigned j = 0u; j < 8u; ++j) {
            L0A(R, r + 4u * j, o) = Ao * __fma_rn(-cs, Gp[j] / Ap, Go[j] / Ao);
            L0A(W, r + 4u * j, p) = __fma_rn(cs, Vo[j], Vp[j]);
            L0A(W, r + 4u * j, o) = __fma_rn(-cs, Vp[j], Vo[j]);
          }
        }
        else {
          // column p loses its component along o, and the two change places
#pragma unroll
          for (unsigned j = 0u; j < 8u; ++j) {
            L0A(R, r + 4u * j, o) = Ap * __fma_rn(-cs, Go[j] / Ao, Gp[j] / Ap);
            L0A(R, r + 4u * j, p) = Go[j];
            L0A(W, r + 4u * j, o) = __fma_rn(-cs, Vo[j], Vp[j]);
            L0A(W, r + 4u * j, p) = __fma_rn(cs, Vp[j], Vo[j]);
          }
        }
      }
      else if (pos ? (Ap < Ao) : (Ao < Ap)) {
#pragma unroll
        for (unsigned j = 0u; j < 8u; ++j) {
          L0A(R, r + 4u * j, p) = Go[j];
          L0A(R, r + 4u * j, o) = Gp[j];
          L0A(W, r + 4u * j, p) = Vo[j];
          L0A(W, r + 4u * j, o) = Vp[j];
        }
      }
      my_s += (unsigned)rot;
      my_b += (unsigned)big;
      wave_lds_sync();
    }
    // the four lanes of a quad hold its counts
    const unsigned swp_s = wave_sum_u(my_s) >> 2, swp_b = wave_sum_u(my_b) >> 2;
    if (!swp_s)
      break;
    blk_s += swp_s;
    blk_b += swp_b;
  }
  out_s = blk_s;
  out_b = blk_b;
}

// ---------------------------------------------------------------------------

// J of the pair's 32 columns (J is +1 on the first nplus columns of G):
// npos of them are positive, and they come first; definite is +1 / -1 when
// all / none are, else 0 (share/device_code_nosort_accumV.hpp:53-76, whose
// case nplus = 16 bj + 16 stays mixed with npos = 32).
__device__ __forceinline__ void pair_signs(int definite, unsigned nplus, unsigned bi, unsigned bj,
                                           int &def, bool &hyp, unsigned &npos)
{
  def = definite;
  hyp = false;
  npos = 0u;
  if (definite)
    return;
  if (nplus <= 16u * bi) {
    def = -1;
  }
  else if (nplus > 16u * bj + 16u) {
    def = +1;
    npos = 32u;
  }
  else {
    def = 0;
    hyp = true;
    npos = min(nplus - 16u * bi, 16u) + ((nplus > 16u * bj) ? nplus - 16u * bj : 0u);
  }
}

// One block pair per 64-lane workgroup.
template <bool SORT, bool DRMAC>
__global__ __launch_bounds__(64) void k_pair_qr(const double *__restrict__ G, unsigned ldG,
                                                unsigned nrow, const uint16_t *__restrict__ pairs,
                                                double *__restrict__ Wout,
                                                unsigned *__restrict__ flags, RunCtl *__restrict__ ctl,
                                                const uint8_t *__restrict__ strat0, unsigned s0steps,
                                                unsigned nswp, int definite, unsigned nplus)
{
  __shared__ double R[32u * L0LD];
  __shared__ double P[32u * L0LD];
  __shared__ double gv[96];
  __shared__ __attribute__((aligned(16))) uint8_t S0[STRAT0_TAB_BYTES];
  const unsigned l = threadIdx.x, q = l >> 2, r = l & 3u;
  const unsigned pi = blockIdx.x;
  const unsigned bi = pairs[2u * pi], bj = pairs[2u * pi + 1u];
  static_assert(STRAT0_TAB_BYTES == 64u * sizeof(uint4), "one 16-byte piece per lane");
  reinterpret_cast<uint4 *>(S0)[l] = reinterpret_cast<const uint4 *>(strat0)[l];

  int def;
  bool hyp;
  unsigned npos;
  pair_signs(definite, nplus, bi, bj, def, hyp, npos);

  // panel loads: lane l takes row l % 32 of the pair's columns 2i + l / 32
  // (i < 8 in block bi, i >= 8 in block bj)
  const unsigned row = l & 31u, half = l >> 5;
  const double *const c0 = G + (size_t)(16u * bi + half) * ldG + row;
  const double *const c1 = G + (size_t)(16u * bj + half) * ldG + row;
  auto load_panel = [&](double (&v)[16], unsigned i0) {
#pragma unroll
    for (unsigned i = 0u; i < 16u; ++i)
      v[i] = ((i < 8u) ? c0 : c1)[(size_t)(2u * (i & 7u)) * ldG + i0];
  };
  double pf[16];
  load_panel(pf, 0u);
#pragma unroll
  for (unsigned i = 0u; i < 16u; ++i) L0A(R, row, 2u * i + half) = pf[i];
  if (nrow > 32u)
    load_panel(pf, 32u);
  wave_lds_sync();
  householder32(R, q, r);
#pragma unroll 1
  for (unsigned i0 = 32u; i0 < nrow; i0 += 32u) {
#pragma unroll
    for (unsigned i = 0u; i < 16u; ++i) L0A(P, row, 2u * i + half) = pf[i];
    if (i0 + 32u < nrow)
      load_panel(pf, i0 + 32u);
    wave_lds_sync();
    householder32(P, q, r);
    merge_rows(R, P, gv, l, q, r);
  }
  // W = I
#pragma unroll
  for (unsigned i = 0u; i < 16u; ++i) {
    const unsigned c = 2u * i + half;
    L0A(P, row, c) = (row == c) ? 1.0 : 0.0;
  }
  wave_lds_sync();

  unsigned blk_s = 0u, blk_b = 0u;
  if (DRMAC && SORT && !hyp) {
    drmac_level0(R, P, S0, s0steps, nswp, def, q, r, blk_s, blk_b);
  }
  else {
    double unused_cos;
    bool unused_swap;
    wave_level0<SORT, false>(R, P, S0, s0steps, nswp, def, hyp, npos, l, blk_s, blk_b, unused_swap,
                             unused_cos);
  }
  wave_lds_sync();

  if (l == 0u) {
    // the run's counter: big rotations in the high word, all in the low one;
    // without a big rotation only the low word is added to, with no carry
    // out of it (share/device_code_nosort.hpp:94-103)
    if (blk_s && blk_b)
      atomicAdd(&ctl->cvg, (((unsigned long long)blk_b) << 32) | blk_s);
    else if (blk_s)
      atomicAdd(reinterpret_cast<unsigned *>(&ctl->cvg), blk_s);
    flags[pi] = SORT ? 1u : (blk_s ? 1u : 0u);
  }
  double *const wo = Wout + (size_t)pi * 1024u;
#pragma unroll
  for (unsigned i = 0u; i < 16u; ++i) {
    const unsigned c = 2u * i + half;
    wo[c * 32u + row] = L0A(P, row, c);
  }
}

}  // namespace

void launch_pair_qr(hipStream_t s, bool sort, bool drmac, const double *G, unsigned ldG,
                    unsigned nrow, const uint16_t *pairs, unsigned np, double *W, unsigned *flags,
                    RunCtl *ctl, const uint8_t *strat0, unsigned s0steps, unsigned nswp,
                    int definite, unsigned nplus)
{
  if (sort && drmac)
    hipLaunchKernelGGL((k_pair_qr<true, true>), dim3(np), dim3(64), 0, s, G, ldG, nrow, pairs, W,
                       flags, ctl, strat0, s0steps, nswp, definite, nplus);
  else if (sort)
    hipLaunchKernelGGL((k_pair_qr<true, false>), dim3(np), dim3(64), 0, s, G, ldG, nrow, pairs, W,
                       flags, ctl, strat0, s0steps, nswp, definite, nplus);
  else
    hipLaunchKernelGGL((k_pair_qr<false, false>), dim3(np), dim3(64), 0, s, G, ldG, nrow, pairs, W,
                       flags, ctl, strat0, s0steps, nswp, definite, nplus);
}

}  // namespace hj
