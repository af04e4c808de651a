// hj_device.hpp -- device-side scalar arithmetic of the method (gfx950).
// Every operation is an IEEE-754 double operation (fma, +, *, /, sqrt are
// correctly rounded on gfx950), written out explicitly; the translation unit
// is compiled with -ffp-contract=off so no other fusion happens.
#ifndef HJ_DEVICE_HPP
#define HJ_DEVICE_HPP

#include <hip/hip_runtime.h>
#include <cstdint>

namespace hj {

// constants of the method (share/HypJacL.hpp:11-51)
constexpr double SQRT_HEPS = 0x1.6a09e667f3bcdp-27;  // sqrt(2^-53)
constexpr double SQRT_2_HEPS = 0x1p27;               // sqrt(2 / 2^-53)
constexpr double MYTOL = 0x1.6a09e667f3bcdp-51;      // sqrt(32) * 2^-53
constexpr double COSHI_FIX = 0x1.279a74590331cp+0;   // 2/sqrt(3), correctly rounded
constexpr double TANH1_FIX = 0.5;

// x held in a register at this point: the operations producing it are
// issued here, on every path, rather than moved into a later branch (where
// they would run after, not beside, the other work of the step)
__device__ __forceinline__ double pin(double x)
{
  asm volatile("" : "+v"(x));
  return x;
}

__device__ __forceinline__ uint64_t d2u(double x) { return (uint64_t)__double_as_longlong(x); }
__device__ __forceinline__ double u2d(uint64_t u) { return __longlong_as_double((long long)u); }

// m*(y+d)^2 - 1 with exact products; see dev_rsqrt.
__device__ __forceinline__ double rs_resid(double m, double y, double d)
{
  const double h = y * y, l = __fma_rn(y, y, -h);
  const double p = m * h, pe = __fma_rn(m, h, -p);
  const double q = m * l;
  const double my = m * y;
  double T = (p - 1.0) + pe;
  T = T + q;
  T = T + (2.0 * d) * my;
  T = T + (m * d) * d;
  return T;
}

// Correctly rounded 1/sqrt(m) for m in [1,4): the hardware reciprocal square
// root refined by two Newton steps y += y(1 - m y^2)/2 lands within an ulp
// (the hardware seed has more than 20 correct bits, so the Newton error is
// far below the final rounding); y then moves to a neighbour when the
// midpoint between them lies on the other side of 1/sqrt(m), judged by the
// sign of m*mid^2 - 1 with fma-exact products.  Straight-line code.
__device__ __forceinline__ double rsqrt_core(double m)
{
  double y = __builtin_amdgcn_rsq(m);
  double h = m * y;
  y = __fma_rn(0.5 * y, __fma_rn(-h, y, 1.0), y);
  h = m * y;
  y = __fma_rn(0.5 * y, __fma_rn(-h, y, 1.0), y);
  const double up = u2d(d2u(y) + 1u), dn = u2d(d2u(y) - 1u);
  const double Tu = rs_resid(m, y, (up - y) * 0.5);
  const double Td = rs_resid(m, y, -((y - dn) * 0.5));
  return (Tu < 0.0) ? up : ((Td > 0.0) ? dn : y);
}

// Correctly rounded 1/sqrt(x).  Replaces my_drsqrt_rn of
// share/device_code_common.hpp:89-133 (whose contract is the correctly
// rounded result).  Own method: x = m * 4^k with m in [1,4), rsqrt_core(m),
// exact rescaling; IEEE special values as the reference returns them.
__device__ __forceinline__ double dev_rsqrt(double x)
{
  // special inputs resolved by selects (no branches: the callers' code stays
  // straight-line for the scheduler); the main path also runs on them, its
  // value discarded.  x = mt * 2^e with mt in [0.5, 1) (v_frexp_*, subnormal
  // x included), then x = m * 4^k with m = mt * 2^(2 - e mod 2) in [1, 4)
  const bool special = !((x > 0.0) && (x < __builtin_inf()));  // not positive finite
  const int e = __builtin_amdgcn_frexp_exp(x);
  const double mt = __builtin_amdgcn_frexp_mant(x);
  const int odd = e & 1;
  const double m = __builtin_amdgcn_ldexp(mt, 2 - odd);
  const int k = (e - 2 + odd) >> 1;  // exact: e - 2 + odd is even
  const double y = rsqrt_core(m);
  const double normal = __builtin_amdgcn_ldexp(y, -k);  // exact: the result is a normal number
  const uint64_t b0 = d2u(x);
  double sp = 0.0;                                                   // +inf
  sp = (x < 0.0) ? u2d(0xfff8000000000000ull) : sp;                  // negative: NaN
  sp = (x == 0.0) ? u2d((b0 & 0x8000000000000000ull) | 0x7ff0000000000000ull) : sp;  // +-0: +-inf
  sp = (x != x) ? x + x : sp;                                        // NaN
  return special ? sp : normal;
}

// 2x2 rotation of share/device_code_common_rotate.hpp:189-237 (USE_RSQRT),
// trigonometric (trig) or hyperbolic, evaluated without divergent branches:
// every path's operations are issued and the taken branch's results are
// selected, so the values equal the branch formulas bit for bit.  The square
// root and the reciprocal square root are skipped by a wave-uniform branch
// when no active lane of the wave takes a branch that uses them (a nearly
// converged wave: every pivot pair's |zeta| at or above sqrt(2/eps), c = 1).
//   trig: zeta = (Dq-Dp)/(2Apq); |zeta| < sqrt(eps/2): t = 1/(|zeta|+1),
//         |zeta| >= sqrt(2/eps): t = 1/(2|zeta|), c = 1, else
//         t = 1/(|zeta| + sqrt(zeta^2+1)); c = rsqrt(1+t^2); big = c < 1.
//   hyp:  zeta = -(Dq+Dp)/(2Apq); |zeta| >= sqrt(2/eps): t = 1/(2|zeta|),
//         c = 1; |zeta| > 1: t = 1/(|zeta| + sqrt(zeta^2-1)),
//         c = rsqrt(1-t^2), big = c > 1; else c = 2/sqrt(3), t = 1/2.
__device__ __forceinline__ int dev_rot(bool trig, double Apq, double Dp, double Dq, double &c,
                                       double &t)
{
  const double h = trig ? (Dq - Dp) : (Dq + Dp);
  const double qz = pin(h / (2.0 * Apq));
  const double ct2 = trig ? qz : -qz;
  const double CT2 = fabs(ct2);
  const bool large = CT2 >= SQRT_2_HEPS;
  const bool small = trig && (CT2 < SQRT_HEPS);
  const bool fix = !trig && !large && !(CT2 > 1.0);
  double sq = 0.0;
  if (__ballot(!large && !small && !fix) != 0ull)
    sq = pin(sqrt(__fma_rn(CT2, CT2, trig ? 1.0 : -1.0)));
  const double cot = large ? 2.0 * CT2 : (small ? CT2 + 1.0 : CT2 + sq);
  const double rcot = pin(1.0 / cot);  // (a statement of its own: a pin inside ?: runs on one arm)
  const double tt = fix ? TANH1_FIX : copysign(rcot, ct2);
  double cr = 1.0;
  if (__ballot(!large && !fix) != 0ull)
    cr = pin(dev_rsqrt(__fma_rn(trig ? tt : -tt, tt, 1.0)));
  c = large ? 1.0 : (fix ? COSHI_FIX : cr);
  t = tt;
  return large ? 0 : ((small || fix) ? 1 : (trig ? (cr < 1.0) : (cr > 1.0)));
}

__device__ __forceinline__ int dev_rot_t(double Apq, double Dp, double Dq, double &c, double &t)
{
  return dev_rot(true, Apq, Dp, Dq, c, t);
}

__device__ __forceinline__ int dev_rot_h(double Apq, double Dp, double Dq, double &c, double &t)
{
  return dev_rot(false, Apq, Dp, Dq, c, t);
}

// xor-swizzle of a double inside each 32-lane half of the wave
// (ds_swizzle bit mode: and 0x1f, or 0, xor M).
template <int M>
__device__ __forceinline__ double swz_xor(double v)
{
  const int lo = __builtin_amdgcn_ds_swizzle(__double2loint(v), (M << 10) | 0x1F);
  const int hi = __builtin_amdgcn_ds_swizzle(__double2hiint(v), (M << 10) | 0x1F);
  return __hiloint2double(hi, lo);
}

// Sum over the 32 lanes of a half wave by the xor butterfly 16,8,4,2,1
// (the order of dSum32, share/device_code_common_Kepler.hpp:26-74).  Every
// lane ends with the same bits (IEEE addition commutes).
__device__ __forceinline__ double sum32(double v)
{
  v = v + swz_xor<16>(v);
  v = v + swz_xor<8>(v);
  v = v + swz_xor<4>(v);
  v = v + swz_xor<2>(v);
  v = v + swz_xor<1>(v);
  return v;
}

// Lane exchanges inside groups of 8 lanes by DPP (no LDS traffic):
// xor 1 and xor 2 are quad permutations, xor 4 is the half-row mirror
// (lane i <- 7-i) followed by the quad reversal (i <- 3-i).
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v)
{
  // every lane reads a lane of its own row (quad / half-row permutations):
  // no lane keeps an old value, so none is set up
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_xor4(double v) { return dpp_mov<0x1B>(dpp_mov<0x141>(v)); }
__device__ __forceinline__ double lane_xor2(double v) { return dpp_mov<0x4E>(v); }
__device__ __forceinline__ double lane_xor1(double v) { return dpp_mov<0xB1>(v); }

// Sum of the 32 rows of one pivot pair held by 8 lanes, lane r holding rows
// r, r+8, r+16, r+24 as v0..v3: the xor butterfly 16,8,4,2,1 of the 32-lane
// reduction (levels 16 and 8 inside the lane, 4,2,1 across lanes).
__device__ __forceinline__ double sum8x4(double v0, double v1, double v2, double v3)
{
  double s = (v0 + v2) + (v1 + v3);
  s = s + lane_xor4(s);
  s = s + lane_xor2(s);
  s = s + lane_xor1(s);
  return s;
}

__device__ __forceinline__ double readlane_d(double v, int lane)
{
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// A double in LDS: helpers that take generic pointers to LDS (a caller's
// dynamic LDS region, a function not inlined) access them through this type,
// so the accesses are LDS instructions (ds_*) rather than flat ones.
typedef __attribute__((address_space(3))) double lds_f64;

// LDS of one wave's pair state: the 32x32 matrices R and W column-major with
// leading dimension 33
constexpr unsigned L0LD = 33u;
#define L0A(A, r, c) (A)[(c) * L0LD + (r)]

// the LDS accesses of a wave before this point complete before any after it
__device__ __forceinline__ void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// In-place 32x32 Cholesky by one wave, share/device_code_common_Cholesky.hpp:58-158
// (USE_RSQRT): column k is scaled by rsqrt(A(k,k)), the diagonal set to
// sqrt(A(k,k)), then A(x,j) = fma(-A(x,k), A(j,k), A(x,j)) for the rows
// x >= j > k.  A: the Gram's lower triangle in LDS (A[c * LD + r] = A(r,c),
// c <= r); on return A holds R = L^T (zeros below the diagonal).  col: 32
// doubles of LDS for the column broadcast.
//
// Lane l keeps row x = l % 32 of the columns c = 2i + l/32 (i = 0..15), so
// one instruction updates an even column in one half of the wave and an odd
// one in the other: 16 registers and half the update instructions of a
// row-per-lane layout.  The updates run on every row; above the diagonal
// they produce values that no later step reads (the pivot is read on the
// diagonal, the broadcast column below it) and that are not stored.  The
// next pivot A(k+1,k+1) - L(k+1,k)^2 is formed from two readlanes as soon
// as column k is scaled, so its rsqrt runs beside the column's updates.
// check: whether every factor entry is finite and the diagonal positive
// (wave-uniform).
template <unsigned LD>
__device__ __forceinline__ bool chol32_wave(double *A_, double *col_, unsigned l, bool check)
{
  lds_f64 *const A = (lds_f64 *)A_;
  lds_f64 *const col = (lds_f64 *)col_;
  const unsigned x = l & 31u, h = l >> 5;
  double b[16];
#pragma unroll
  for (unsigned i = 0u; i < 16u; ++i) {
    const unsigned c = 2u * i + h;
    b[i] = (c <= x) ? A[c * LD + x] : 0.0;
  }
  double akk = readlane_d(b[0], 0);
  double rk = dev_rsqrt(akk), dk = sqrt(akk);
#pragma unroll
  for (unsigned k = 0u; k < 32u; ++k) {
    const unsigned ik = k >> 1, hk = k & 1u;
    const bool own = (h == hk);
    if (own && x > k) b[ik] = b[ik] * rk;
    else if (own && x == k) b[ik] = dk;
    if (k + 1u < 32u) {
      const unsigned k1 = k + 1u;
      if (own) col[x] = b[ik];
      const double lk = readlane_d(b[ik], hk * 32u + k1);
      const double d1 = readlane_d(b[k1 >> 1], (k1 & 1u) * 32u + k1);
      akk = __fma_rn(-lk, lk, d1);
      const double rn = dev_rsqrt(akk), dn = sqrt(akk);
      // the other half reads what this half stored: without the fence the
      // compiler may issue a lane's read of col[x] before the (conditional,
      // other-lane) store, as no store of its own precedes it
      wave_lds_sync();
      const double lx = col[x];
#pragma unroll
      for (unsigned i = k1 >> 1; i < 16u; ++i) {
        // column 2i + h; for i = k1/2 with k1 odd, only half 1's column is
        // beyond k
        if ((k1 & 1u) && i == (k1 >> 1) && h == 0u) continue;
        b[i] = __fma_rn(-lx, col[2u * i + h], b[i]);
      }
      rk = rn;
      dk = dn;
      wave_lds_sync();
    }
  }
  bool ok = true;
  if (check) {
#pragma unroll
    for (unsigned i = 0u; i < 16u; ++i) {
      const unsigned c = 2u * i + h;
      if (c <= x) ok = ok && (fabs(b[i]) <= 0x1.fffffffffffffp+1023);
      if (c == x) ok = ok && (b[i] > 0.0);
    }
    ok = (__ballot(ok) == ~0ull);
  }
#pragma unroll
  for (unsigned i = 0u; i < 16u; ++i) {
    const unsigned c = 2u * i + h;
    A[x * LD + c] = (c <= x) ? b[i] : 0.0;
  }
  return ok;
}

// Sum of the 32 rows of one pivot pair held by 4 lanes, lane r holding rows
// r + 4k as v[k]: the xor butterfly 16,8,4,2,1 of the reference's 32-lane
// reduction (levels 16, 8 and 4 inside the lane, 2 and 1 across lanes) --
// the bits of sum8x4 on the same rows.
__device__ __forceinline__ double sum4x8(const double *v)
{
  const double a0 = v[0] + v[4], a1 = v[1] + v[5], a2 = v[2] + v[6], a3 = v[3] + v[7];
  double s = (a0 + a2) + (a1 + a3);
  s = s + lane_xor2(s);
  s = s + lane_xor1(s);
  return s;
}

__device__ __forceinline__ unsigned wave_sum_u(unsigned v)
{
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1)
    v += (unsigned)__shfl_xor((int)v, o, 64);
  return v;
}

// Level-0 sweeps of one block pair by one wave (the inner Jacobi of
// share/device_code_nosort.hpp:4-222 for SORT false,
// share/device_code_cdsort_KISS.hpp:4-260 and share/device_code_cdsort.hpp:25-167
// for SORT true).  R: the Cholesky factor, W: the identity on entry, both in
// the L0A layout.  Slot y = l/4 takes pivot pair y of a step, lane r = l%4 of
// the slot holds rows r + 4k of the two columns; the pivot pairs of a step
// touch disjoint columns, so one wave-level LDS sync per step suffices.
// Returns the rotation counts of all its sweeps (the reference's counting:
// a sweep without a rotation ends the loop and adds nothing) in every lane,
// whether a pivot pair was swapped by the sort rule (W is the identity
// exactly when nothing was rotated or swapped), and with COS the largest
// |A_pq| / sqrt(A_pp A_qq) over the pivots of the first sweep.
template <bool SORT, bool COS>
__device__ __forceinline__ void wave_level0(double *R_, double *W_, const uint8_t *S0,
                                            unsigned s0steps, unsigned nswp, int def, bool hyp,
                                            unsigned npos, unsigned l, unsigned &out_s,
                                            unsigned &out_b, bool &out_swap, double &out_cos)
{
  lds_f64 *const R = (lds_f64 *)R_;
  lds_f64 *const W = (lds_f64 *)W_;
  const unsigned y = l >> 2, r = l & 3u;
  unsigned blk_s = 0u, blk_b = 0u;
  bool swapped = false;
  double my_cos = 0.0;
  for (unsigned swp = 0u; swp < nswp; ++swp) {
    unsigned my_s = 0u, my_b = 0u;
    for (unsigned step = 0u; step < s0steps; ++step) {
      const unsigned p = S0[(step * 16u + y) * 2u], q = S0[(step * 16u + y) * 2u + 1u];
      double Gp[8], Gq[8];
#pragma unroll
      for (unsigned k = 0u; k < 8u; ++k) {
        Gp[k] = L0A(R, r + 4u * k, p);
        Gq[k] = L0A(R, r + 4u * k, q);
      }
      double Dp, Dq, Apq;
      {
        double t[8];
#pragma unroll
        for (unsigned k = 0u; k < 8u; ++k)
          t[k] = __fma_rn(Gp[k], Gp[k], 0.0);
        Dp = sum4x8(t);
#pragma unroll
        for (unsigned k = 0u; k < 8u; ++k)
          t[k] = __fma_rn(Gq[k], Gq[k], 0.0);
        Dq = sum4x8(t);
#pragma unroll
        for (unsigned k = 0u; k < 8u; ++k)
          t[k] = __fma_rn(Gp[k], Gq[k], 0.0);
        Apq = sum4x8(t);
      }
      const double Dp_ = sqrt(Dp), Dq_ = sqrt(Dq), Apq_ = fabs(Apq), Dpq_ = Dp_ * Dq_;
      const int transf_s = !(Apq_ < (Dpq_ * MYTOL));
      if (COS && swp == 0u)
        my_cos = fmax(my_cos, (Dpq_ > 0.0) ? (Apq_ / Dpq_) : 0.0);
      int pdef = def;
      if (hyp) {
        if ((p < npos) && (q < npos)) pdef = +1;
        else if ((p >= npos) && (q >= npos)) pdef = -1;
        else pdef = 0;
      }
      int transf_b = 0;
      // W's columns p, q (the same lane reads and writes the same entries
      // within a step, so they may be read this late)
      double Vp[8], Vq[8];
#pragma unroll
      for (unsigned k = 0u; k < 8u; ++k) {
        Vp[k] = L0A(W, r + 4u * k, p);
        Vq[k] = L0A(W, r + 4u * k, q);
      }
      if (transf_s) {
        double c, t, t_;
        bool trig;
        if (!SORT) trig = !hyp || pdef;
        else trig = (pdef != 0);
        transf_b = dev_rot(trig, Apq, Dp, Dq, c, t);
        t_ = trig ? -t : t;
        if (SORT && trig) {
          Dp = __fma_rn(t, Apq, Dp);
          Dq = __fma_rn(t_, Apq, Dq);
        }
        bool keep = true;
        if (SORT)
          keep = (!pdef) || ((pdef > 0) && (Dp >= Dq)) || ((pdef < 0) && (Dp <= Dq));
        const unsigned PP = keep ? p : q, QQ = keep ? q : p;
#pragma unroll
        for (unsigned k = 0u; k < 8u; ++k) {
          // big rotation: scale by c; otherwise c == 1 exactly (dev_rot), and
          // c * x is x
          const double rp = __fma_rn(t_, Gq[k], Gp[k]), rq = __fma_rn(t, Gp[k], Gq[k]);
          const double wp = __fma_rn(t_, Vq[k], Vp[k]), wq = __fma_rn(t, Vp[k], Vq[k]);
          L0A(R, r + 4u * k, PP) = c * rp;
          L0A(R, r + 4u * k, QQ) = c * rq;
          L0A(W, r + 4u * k, PP) = c * wp;
          L0A(W, r + 4u * k, QQ) = c * wq;
        }
      }
      else if (SORT && (((pdef > 0) && (Dp < Dq)) || ((pdef < 0) && (Dq < Dp)))) {
#pragma unroll
        for (unsigned k = 0u; k < 8u; ++k) {
          L0A(R, r + 4u * k, p) = Gq[k];
          L0A(R, r + 4u * k, q) = Gp[k];
          L0A(W, r + 4u * k, p) = Vq[k];
          L0A(W, r + 4u * k, q) = Vp[k];
        }
        swapped = true;
      }
      my_s += (unsigned)transf_s;
      my_b += (unsigned)transf_b;
      wave_lds_sync();
    }
    // every lane of a slot holds the slot's counts: sum over the wave / 4
    const unsigned swp_s = wave_sum_u(my_s) >> 2, swp_b = wave_sum_u(my_b) >> 2;
    if (swp_s) {
      blk_s += swp_s;
      blk_b += swp_b;
    }
    else
      break;
  }
  out_s = blk_s;
  out_b = blk_b;
  out_swap = __ballot(swapped) != 0ull;
  if (COS) {
#pragma unroll
    for (int o = 32; o >= 4; o >>= 1)
      my_cos = fmax(my_cos, __shfl_xor(my_cos, o, 64));
  }
  out_cos = my_cos;
}

}  // namespace hj

#endif
