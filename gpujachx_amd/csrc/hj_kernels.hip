// hj_kernels.hip -- level-1 kernels of the blocked one-sided (hyperbolic)
// Jacobi SVD for gfx950 (MI355X).
//
// One level-1 step (all block pairs of one STRAT1 step) is three launches:
//   k_gram   : partial 32x32 Gram of each block pair over a row chunk, on
//              FP64 MFMA 16x16x4 (one wave per (pair, chunk));
//   k_pair   : per pair, sum of the partials, 32x32 Cholesky, the inner
//              Jacobi sweep(s) on R in LDS accumulating W (one 512-thread
//              workgroup per pair, one pivot pair per 32-lane half wave);
//   k_update : [G_i G_j] <- [G_i G_j] W and [V_i V_j] <- [V_i V_j] W on FP64
//              MFMA (one wave per (pair, row chunk)).
// The reference fuses the three into one CUDA block per pair
// (share/device_code_cdsort_accumV.hpp:4-94 and siblings); splitting them
// lets every step fill the 256 CUs whatever the number of pairs.
#include "hj_device.hpp"
#include "hj_kernels.hpp"

typedef double d4 __attribute__((ext_vector_type(4)));

namespace hj {

__device__ __forceinline__ d4 mfma(double a, double b, d4 c)
{
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// Gram partials.  Wave (pair pi, chunk c) covers rows [c*CH, (c+1)*CH).
// Lane l owns column l%16 of both blocks; per 64-row group it loads rows
// 16*(l/16)+t, t=0..15 (128 contiguous bytes per column per lane), and MFMA
// step t takes, for k = l/16, row 16k+t of the group.  Tiles: 00 = block i
// with itself, 10 = block j with block i, 11 = block j with itself.
__global__ __launch_bounds__(64) void k_gram(const double *__restrict__ G, unsigned ldG,
                                             const uint16_t *__restrict__ pairs, unsigned nchunk,
                                             unsigned CH, double *__restrict__ part)
{
  const unsigned l = threadIdx.x;
  const unsigned pi = blockIdx.x / nchunk, c = blockIdx.x - pi * nchunk;
  const unsigned iBlk = pairs[2u * pi], jBlk = pairs[2u * pi + 1u];
  const double *const g0 = G + (size_t)(16u * iBlk + (l & 15u)) * ldG + (size_t)c * CH + 16u * (l >> 4);
  const double *const g1 = G + (size_t)(16u * jBlk + (l & 15u)) * ldG + (size_t)c * CH + 16u * (l >> 4);
  d4 a00 = {0.0, 0.0, 0.0, 0.0}, a10 = a00, a11 = a00;
  for (unsigned g = 0u; g < CH; g += 64u) {
    double x0[16], x1[16];
    const double2 *const p0 = reinterpret_cast<const double2 *>(g0 + g);
    const double2 *const p1 = reinterpret_cast<const double2 *>(g1 + g);
#pragma unroll
    for (unsigned t = 0u; t < 8u; ++t) {
      const double2 v0 = p0[t], v1 = p1[t];
      x0[2u * t] = v0.x;
      x0[2u * t + 1u] = v0.y;
      x1[2u * t] = v1.x;
      x1[2u * t + 1u] = v1.y;
    }
#pragma unroll
    for (unsigned t = 0u; t < 16u; ++t) {
      a00 = mfma(x0[t], x0[t], a00);
      a10 = mfma(x1[t], x0[t], a10);
      a11 = mfma(x1[t], x1[t], a11);
    }
  }
  double *const o = part + (size_t)blockIdx.x * 768u + l * 4u;
#pragma unroll
  for (unsigned r = 0u; r < 4u; ++r) {
    o[r] = a00[r];
    o[256u + r] = a10[r];
    o[512u + r] = a11[r];
  }
}

// ---------------------------------------------------------------------------
// Block pair: Gram -> Cholesky -> inner Jacobi -> W.
// Thread t: x = t%32 (row of R and W), y = t/32 (pivot-pair slot 0..15).

#define F32(A, r, c) (A)[(c) * 32u + (r)]

template <bool SORT>
__global__ __launch_bounds__(512) void k_pair(const uint16_t *__restrict__ pairs,
                                              const double *__restrict__ part, unsigned nchunk,
                                              double *__restrict__ Wout, unsigned *__restrict__ flags,
                                              unsigned long long *__restrict__ cvg,
                                              unsigned *__restrict__ pstat,
                                              const uint8_t *__restrict__ strat0, unsigned s0steps,
                                              unsigned nswp, int definite, unsigned nplus)
{
  __shared__ double R[1024];
  __shared__ double W[1024];
  __shared__ uint8_t S0[STRAT0_TAB_BYTES];
  __shared__ unsigned cnt_s[16], cnt_b[16];

  const unsigned tid = threadIdx.x, x = tid & 31u, y = tid >> 5;
  const unsigned pi = blockIdx.x;
  const unsigned iBlk = pairs[2u * pi], jBlk = pairs[2u * pi + 1u];

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

  for (unsigned e = tid; e < s0steps * 32u; e += 512u)
    S0[e] = strat0[e];

  // sum of the chunk partials in chunk order
  const double *const pp = part + (size_t)pi * nchunk * 768u;
  for (unsigned e = tid; e < 768u; e += 512u) {
    double s = pp[e];
    for (unsigned c = 1u; c < nchunk; ++c)
      s = s + pp[(size_t)c * 768u + e];
    const unsigned tile = e >> 8, l = (e >> 2) & 63u, r = e & 3u;
    const unsigned i = 4u * r + (l >> 4), j = l & 15u;
    const unsigned row = (tile == 0u) ? i : (16u + i);
    const unsigned col = (tile == 2u) ? (16u + j) : j;
    F32(R, row, col) = s;
  }
  F32(W, x, y) = (x == y) ? 1.0 : 0.0;
  F32(W, x, y + 16u) = (x == y + 16u) ? 1.0 : 0.0;
  __syncthreads();

  // Cholesky, share/device_code_common_Cholesky.hpp:58-158 (USE_RSQRT):
  // column k scaled by rsqrt(A(k,k)), diagonal sqrt(A(k,k)), then
  // A(x,j) = fma(-A(x,k), A(j,k), A(x,j)) for x >= j > k.
  for (unsigned k = 0u; k < 32u; ++k) {
    if (tid < 32u && x >= k) {
      const double Akk = F32(R, k, k);
      const double Axk = F32(R, x, k);
      F32(R, x, k) = (x > k) ? (Axk * dev_rsqrt(Akk)) : sqrt(Akk);
    }
    __syncthreads();
    const unsigned j0 = k + 1u + y, j1 = j0 + 16u;
    if (j0 < 32u && x >= j0)
      F32(R, x, j0) = __fma_rn(-F32(R, x, k), F32(R, j0, k), F32(R, x, j0));
    if (j1 < 32u && x >= j1)
      F32(R, x, j1) = __fma_rn(-F32(R, x, k), F32(R, j1, k), F32(R, x, j1));
    __syncthreads();
  }
  {
    // R = L^T with zeros below the diagonal
    const double a0 = (x >= y) ? F32(R, x, y) : 0.0;
    const double a1 = (x >= y + 16u) ? F32(R, x, y + 16u) : 0.0;
    __syncthreads();
    F32(R, y, x) = a0;
    F32(R, y + 16u, x) = a1;
    __syncthreads();
  }

  // Inner Jacobi: share/device_code_nosort.hpp:4-222 (SORT false),
  // share/device_code_cdsort_KISS.hpp:4-260 and share/device_code_cdsort.hpp:25-167
  // (SORT true).  One barrier per step: the pivot pairs of a step touch
  // disjoint columns; rotation counts are kept per slot and summed per sweep.
  unsigned blk_s = 0u, blk_b = 0u;
  for (unsigned swp = 0u; swp < nswp; ++swp) {
    unsigned my_s = 0u, my_b = 0u;
    for (unsigned step = 0u; step < s0steps; ++step) {
      const unsigned p = S0[(step * 16u + y) * 2u], q = S0[(step * 16u + y) * 2u + 1u];
      const double Gp = F32(R, x, p), Gq = F32(R, x, q);
      const double Vp = F32(W, x, p), Vq = F32(W, x, q);
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
      if (!SORT) {
        if (transf_s) {
          double c, t, t_;
          if (!hyp || pdef) { transf_b = dev_rot_t(Apq, Dp, Dq, c, t); t_ = -t; }
          else { transf_b = dev_rot_h(Apq, Dp, Dq, c, t); t_ = t; }
          if (transf_b) {
            F32(R, x, p) = c * __fma_rn(t_, Gq, Gp);
            F32(R, x, q) = c * __fma_rn(t, Gp, Gq);
            F32(W, x, p) = c * __fma_rn(t_, Vq, Vp);
            F32(W, x, q) = c * __fma_rn(t, Vp, Vq);
          }
          else {
            F32(R, x, p) = __fma_rn(t_, Gq, Gp);
            F32(R, x, q) = __fma_rn(t, Gp, Gq);
            F32(W, x, p) = __fma_rn(t_, Vq, Vp);
            F32(W, x, q) = __fma_rn(t, Vp, Vq);
          }
        }
      }
      else {
        if (transf_s) {
          double c, t, t_;
          if (pdef) {
            transf_b = dev_rot_t(Apq, Dp, Dq, c, t);
            t_ = -t;
            Dp = __fma_rn(t, Apq, Dp);
            Dq = __fma_rn(t_, Apq, Dq);
          }
          else {
            transf_b = dev_rot_h(Apq, Dp, Dq, c, t);
            t_ = t;
          }
          const bool keep = (!pdef) || ((pdef > 0) && (Dp >= Dq)) || ((pdef < 0) && (Dp <= Dq));
          const unsigned P = keep ? p : q, Q = keep ? q : p;
          if (transf_b) {
            F32(R, x, P) = c * __fma_rn(t_, Gq, Gp);
            F32(R, x, Q) = c * __fma_rn(t, Gp, Gq);
            F32(W, x, P) = c * __fma_rn(t_, Vq, Vp);
            F32(W, x, Q) = c * __fma_rn(t, Vp, Vq);
          }
          else {
            F32(R, x, P) = __fma_rn(t_, Gq, Gp);
            F32(R, x, Q) = __fma_rn(t, Gp, Gq);
            F32(W, x, P) = __fma_rn(t_, Vq, Vp);
            F32(W, x, Q) = __fma_rn(t, Vp, Vq);
          }
        }
        else if (((pdef > 0) && (Dp < Dq)) || ((pdef < 0) && (Dq < Dp))) {
          F32(R, x, p) = Gq;
          F32(R, x, q) = Gp;
          F32(W, x, p) = Vq;
          F32(W, x, q) = Vp;
        }
      }
      my_s += (unsigned)transf_s;
      my_b += (unsigned)transf_b;
      __syncthreads();
    }
    if (x == 0u) {
      cnt_s[y] = my_s;
      cnt_b[y] = my_b;
    }
    __syncthreads();
    unsigned swp_s = 0u, swp_b = 0u;
    for (unsigned i = 0u; i < 16u; ++i) {
      swp_s += cnt_s[i];
      swp_b += cnt_b[i];
    }
    __syncthreads();
    if (swp_s) {
      blk_s += swp_s;
      blk_b += swp_b;
    }
    else
      break;
  }

  // convergence counter, share/device_code_nosort.hpp:94-103
  if (tid == 0u) {
    if (blk_s) {
      if (blk_b)
        atomicAdd(cvg, (((unsigned long long)blk_b) << 32) | blk_s);
      else
        atomicAdd(reinterpret_cast<unsigned *>(cvg), blk_s);
    }
    flags[pi] = SORT ? 1u : (blk_s ? 1u : 0u);
    pstat[2u * pi] = blk_s;
    pstat[2u * pi + 1u] = blk_b;
  }
  double *const wo = Wout + (size_t)pi * 1024u;
  wo[tid] = W[tid];
  wo[tid + 512u] = W[tid + 512u];
}

// ---------------------------------------------------------------------------
// Block update on MFMA: out^T = W^T [G_i G_j]^T per 16-row slab.
// A operand (W^T, 16 output columns x 4 k): lane l gives W(4u + l/16, 16h + l%16);
// B operand (G^T, 4 k x 16 rows): lane l gives G(r + l%16, column 4u + l/16);
// D: lane l, reg q holds output column 16h + 4q + l/16 of row r + l%16.
// The chain over u = 0..7, k = 0..3 runs over the 32 columns in order.
__global__ __launch_bounds__(64) void k_update(double *__restrict__ G, unsigned ldG, unsigned nrowG,
                                               unsigned chG, unsigned nchG,
                                               double *__restrict__ V, unsigned ldV, unsigned nrowV,
                                               unsigned chV, unsigned nchV,
                                               const uint16_t *__restrict__ pairs,
                                               const double *__restrict__ Wbuf,
                                               const unsigned *__restrict__ flags)
{
  const unsigned l = threadIdx.x;
  const unsigned per = nchG + nchV;
  const unsigned pi = blockIdx.x / per, c = blockIdx.x - pi * per;
  if (!flags[pi])
    return;
  double *M;
  unsigned ld, r0, r1;
  if (c < nchG) {
    M = G; ld = ldG; r0 = c * chG; r1 = min(r0 + chG, nrowG);
  }
  else {
    M = V; ld = ldV; r0 = (c - nchG) * chV; r1 = min(r0 + chV, nrowV);
  }
  const unsigned iBlk = pairs[2u * pi], jBlk = pairs[2u * pi + 1u];
  const double *const w = Wbuf + (size_t)pi * 1024u;
  double a0[8], a1[8];
#pragma unroll
  for (unsigned u = 0u; u < 8u; ++u) {
    a0[u] = F32(w, 4u * u + (l >> 4), (l & 15u));
    a1[u] = F32(w, 4u * u + (l >> 4), 16u + (l & 15u));
  }
  // column of the pair read by this lane for instruction u: 4u + l/16
  double *cp[8];
#pragma unroll
  for (unsigned u = 0u; u < 8u; ++u) {
    const unsigned cc = 4u * u + (l >> 4);
    const unsigned gc = (cc < 16u) ? (16u * iBlk + cc) : (16u * jBlk + cc - 16u);
    cp[u] = M + (size_t)gc * ld + (l & 15u);
  }
  // output columns for reg q: 4q + l/16 (+16 for the second tile)
  double *op0[4], *op1[4];
#pragma unroll
  for (unsigned q = 0u; q < 4u; ++q) {
    const unsigned cc = 4u * q + (l >> 4);
    op0[q] = M + (size_t)(16u * iBlk + cc) * ld + (l & 15u);
    op1[q] = M + (size_t)(16u * jBlk + cc) * ld + (l & 15u);
  }
  for (unsigned r = r0; r < r1; r += 16u) {
    double b[8];
#pragma unroll
    for (unsigned u = 0u; u < 8u; ++u)
      b[u] = cp[u][r];
    d4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = acc0;
#pragma unroll
    for (unsigned u = 0u; u < 8u; ++u) {
      acc0 = mfma(a0[u], b[u], acc0);
      acc1 = mfma(a1[u], b[u], acc1);
    }
#pragma unroll
    for (unsigned q = 0u; q < 4u; ++q) {
      op0[q][r] = acc0[q];
      op1[q][r] = acc1[q];
    }
  }
}

// ---------------------------------------------------------------------------
// Column norms, G -> U, D = +-||g||^2 (share/device_code_common.hpp:187-219).
// One wave per column: lane l sums rows l, l+64, ... by fma, then an xor
// butterfly over 64 lanes (32,16,8,4,2,1).
__global__ __launch_bounds__(256) void k_initD(double *__restrict__ G, double *__restrict__ D,
                                               unsigned ifc, unsigned nRow, unsigned nRank,
                                               unsigned nPlus, unsigned ldG)
{
  const unsigned wid = threadIdx.x >> 6, lid = threadIdx.x & 63u;
  const unsigned cix = blockIdx.x * 4u + wid;
  if (cix >= nRank)
    return;
  double *const g = G + (size_t)cix * ldG;
  double x = 0.0;
  for (unsigned r = lid; r < nRow; r += 64u) {
    const double v = g[r];
    x = __fma_rn(v, v, x);
  }
  x = x + __shfl_xor(x, 32);
  x = x + __shfl_xor(x, 16);
  x = x + __shfl_xor(x, 8);
  x = x + __shfl_xor(x, 4);
  x = x + __shfl_xor(x, 2);
  x = x + __shfl_xor(x, 1);
  const double s = sqrt(x);
  for (unsigned r = lid; r < nRow; r += 64u)
    g[r] = g[r] / s;
  if (!lid)
    D[cix] = ((ifc + cix) < nPlus) ? x : -x;
}

// V(ifc + c, c) = 1 for c < nRank, zero elsewhere in those columns
// (share/device_code_common.hpp:221-230 after the memset of
// 1/HypJacL2.cu:96-97), written in full so no memset is needed.
__global__ __launch_bounds__(256) void k_initV(double *__restrict__ V, unsigned ifc, unsigned nRank,
                                               unsigned ldV)
{
  const size_t n = (size_t)nRank * ldV;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x) {
    const size_t col = e / ldV, row = e - col * ldV;
    V[e] = (row == ifc + col) ? 1.0 : 0.0;
  }
}

__global__ void k_zero_u64(unsigned long long *p) { *p = 0ull; }

// Scalar self-test kernel.
__global__ void k_dbg_scalar(int kind, int n, const double *a, const double *b, const double *c,
                             double *oc, double *ot, int *ob)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double cc = 0.0, tt = 0.0;
  int big = 0;
  if (kind == 0) cc = dev_rsqrt(a[i]);
  else if (kind == 1) big = dev_rot_t(a[i], b[i], c[i], cc, tt);
  else big = dev_rot_h(a[i], b[i], c[i], cc, tt);
  oc[i] = cc;
  ot[i] = tt;
  ob[i] = big;
}

// ---------------------------------------------------------------------------
// Launchers

void launch_gram(hipStream_t s, const double *G, unsigned ldG, const uint16_t *pairs, unsigned np,
                 unsigned nchunk, unsigned CH, double *part)
{
  hipLaunchKernelGGL(k_gram, dim3(np * nchunk), dim3(64), 0, s, G, ldG, pairs, nchunk, CH, part);
}

void launch_pair(hipStream_t s, bool sort, const uint16_t *pairs, unsigned np, const double *part,
                 unsigned nchunk, double *W, unsigned *flags, unsigned long long *cvg,
                 unsigned *pstat, const uint8_t *strat0, unsigned s0steps, unsigned nswp,
                 int definite, unsigned nplus)
{
  if (sort)
    hipLaunchKernelGGL(k_pair<true>, dim3(np), dim3(512), 0, s, pairs, part, nchunk, W, flags, cvg,
                       pstat, strat0, s0steps, nswp, definite, nplus);
  else
    hipLaunchKernelGGL(k_pair<false>, dim3(np), dim3(512), 0, s, pairs, part, nchunk, W, flags,
                       cvg, pstat, strat0, s0steps, nswp, definite, nplus);
}

void launch_update(hipStream_t s, double *G, unsigned ldG, unsigned nrowG, unsigned chG,
                   double *V, unsigned ldV, unsigned nrowV, unsigned chV, const uint16_t *pairs,
                   unsigned np, const double *W, const unsigned *flags)
{
  const unsigned nchG = (nrowG + chG - 1u) / chG;
  const unsigned nchV = V ? (nrowV + chV - 1u) / chV : 0u;
  hipLaunchKernelGGL(k_update, dim3(np * (nchG + nchV)), dim3(64), 0, s, G, ldG, nrowG, chG, nchG,
                     V, ldV, nrowV, chV, nchV, pairs, W, flags);
}

void launch_initD(hipStream_t s, double *G, double *D, unsigned ifc, unsigned nRow, unsigned nRank,
                  unsigned nPlus, unsigned ldG)
{
  hipLaunchKernelGGL(k_initD, dim3((nRank + 3u) / 4u), dim3(256), 0, s, G, D, ifc, nRow, nRank,
                     nPlus, ldG);
}

void launch_initV(hipStream_t s, double *V, unsigned ifc, unsigned nRank, unsigned ldV)
{
  const size_t n = (size_t)nRank * ldV;
  size_t blocks = (n + 255u) / 256u;
  if (blocks > 65536u) blocks = 65536u;
  if (!blocks) blocks = 1u;
  hipLaunchKernelGGL(k_initV, dim3((unsigned)blocks), dim3(256), 0, s, V, ifc, nRank, ldV);
}

void launch_zero_u64(hipStream_t s, unsigned long long *p)
{
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, p);
}

void launch_dbg_scalar(hipStream_t s, int kind, int n, const double *a, const double *b,
                       const double *c, double *oc, double *ot, int *ob)
{
  hipLaunchKernelGGL(k_dbg_scalar, dim3((n + 255) / 256), dim3(256), 0, s, kind, n, a, b, c, oc, ot,
                     ob);
}

}  // namespace hj
