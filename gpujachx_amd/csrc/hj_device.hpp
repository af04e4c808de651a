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

// Correctly rounded 1/sqrt(x).  Replaces my_drsqrt_rn of
// share/device_code_common.hpp:89-133 (whose contract is the correctly
// rounded result).  Own method: scale x to m in [1,4) by an even power of
// two, take y = RN(1/RN(sqrt(m))) (within 2 ulps), then move to a neighbour
// while the midpoint between y and it lies on the other side of 1/sqrt(m),
// judged by the sign of m*mid^2 - 1 computed with fma-exact products.
__device__ __forceinline__ double dev_rsqrt(double x)
{
  uint64_t b = d2u(x);
  if (x != x) return x + x;
  if (x == 0.0) return u2d((b & 0x8000000000000000ull) | 0x7ff0000000000000ull);
  if (x < 0.0) return u2d(0xfff8000000000000ull);
  if (b == 0x7ff0000000000000ull) return 0.0;
  int kadj = 0;
  if (!(b >> 52)) {
    x = x * 0x1p108;
    b = d2u(x);
    kadj = 54;
  }
  const int e = (int)(b >> 52) - 1023;
  const int k = (e >= 0) ? (e / 2) : -((1 - e) / 2);
  const double m = u2d((b & 0x000fffffffffffffull) | ((uint64_t)(1023 + e - 2 * k) << 52));
  double y = 1.0 / sqrt(m);
  for (int it = 0; it < 4; ++it) {
    const double up = u2d(d2u(y) + 1u);
    if (rs_resid(m, y, (up - y) * 0.5) < 0.0) y = up;
    else break;
  }
  for (int it = 0; it < 4; ++it) {
    const double dn = u2d(d2u(y) - 1u);
    if (rs_resid(m, y, -((y - dn) * 0.5)) > 0.0) y = dn;
    else break;
  }
  const int s = kadj - k;
  return y * u2d((uint64_t)(1023 + s) << 52);
}

// Trigonometric 2x2 rotation, share/device_code_common_rotate.hpp:189-213.
__device__ __forceinline__ int dev_rot_t(double Apq, double Dp, double Dq, double &c, double &t)
{
  const double h = Dq - Dp;
  const double ct2 = h / (2.0 * Apq);
  const double CT2 = fabs(ct2);
  if (CT2 < SQRT_HEPS) {
    const double cot = CT2 + 1.0;
    t = copysign(1.0 / cot, ct2);
    c = dev_rsqrt(__fma_rn(t, t, 1.0));
    return 1;
  }
  else if (CT2 >= SQRT_2_HEPS) {
    const double cot = 2.0 * CT2;
    c = 1.0;
    t = copysign(1.0 / cot, ct2);
    return 0;
  }
  else {
    const double cot = CT2 + sqrt(__fma_rn(CT2, CT2, 1.0));
    t = copysign(1.0 / cot, ct2);
    c = dev_rsqrt(__fma_rn(t, t, 1.0));
    return (c < 1.0);
  }
}

// Hyperbolic 2x2 rotation, share/device_code_common_rotate.hpp:214-237.
__device__ __forceinline__ int dev_rot_h(double Apq, double Dp, double Dq, double &c, double &t)
{
  const double h = Dq + Dp;
  const double ct2 = -(h / (2.0 * Apq));
  const double CT2 = fabs(ct2);
  if (CT2 >= SQRT_2_HEPS) {
    const double cot = 2.0 * CT2;
    c = 1.0;
    t = copysign(1.0 / cot, ct2);
    return 0;
  }
  else if (CT2 > 1.0) {
    const double cot = CT2 + sqrt(__fma_rn(CT2, CT2, -1.0));
    t = copysign(1.0 / cot, ct2);
    c = dev_rsqrt(__fma_rn(-t, t, 1.0));
    return (c > 1.0);
  }
  else {
    c = COSHI_FIX;
    t = TANH1_FIX;
    return 1;
  }
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

}  // namespace hj

#endif
