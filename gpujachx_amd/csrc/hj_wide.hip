// hj_wide.hip -- the wide-block form: blocks of 32 columns, so a level-1
// visit treats a 64-column block pair (64x64 Gram, Cholesky and level-0 on
// R, 64x64 W).  A sweep moves half the HBM bytes of the reference's
// 16-column blocks (share/HypJacL.hpp:112-116) for the same FP64 flops.
// Step by step: Gram partials (k_gram_w), visits (k_pair_w), block update
// (k_update_w), each in the arithmetic of the 16-column kernels written for
// 64 columns:
//   - Gram: the 10 16x16 tiles of the pair's lower triangle on FP64 MFMA
//     16x16x4, per tile the chain over the rows of a wave's sub-chunk in
//     16-row slabs, the waves added in order, the chunks in order by the
//     visit;
//   - Cholesky (share/device_code_common_Cholesky.hpp:58-158, USE_RSQRT) at
//     64 pivots, one row per lane, the pivots on one wave and the column
//     updates on the others;
//   - level-0: 32 pivot pairs a step, 16 lanes a pair, lane r holding rows
//     r + 16i (i < 4) of both columns of R and W; dot products by the xor
//     butterfly 32,16 in the lane, 8,4,2,1 across lanes; the convergence
//     test at tolerance sqrt(64) 2^-53;
//   - update: out(r, j) = MFMA chain over k = 0..63.
// With the learned pair-skip a step's plan marks skipped pairs: no Gram, no
// visit (its record keeps the last executed visit), no update.
#include "../../include/hypjac.h"
#include "hj_kernels.hpp"
#include "hj_device.hpp"
#include "hj_fused.hpp"
#include "hj_util.hpp"

namespace hj {

namespace {

constexpr unsigned WTILES = 10u;             // tiles of a 64-column pair's lower triangle
constexpr unsigned WPART = WTILES * 256u;    // doubles per Gram partial
constexpr unsigned LDRW = 65u;               // padded column of R and W in LDS
constexpr double WTOL = 0x1p-50;             // sqrt(64) * 2^-53
#define RW(A, r, c) (A)[(c) * LDRW + (r)]

// tile tt -> (ti, tj), tt = ti (ti + 1) / 2 + tj
__device__ __forceinline__ unsigned tile_i(unsigned tt)
{
  return (tt >= 6u) ? 3u : ((tt >= 3u) ? 2u : ((tt >= 1u) ? 1u : 0u));
}

// Gram partials of the pairs of one step: workgroup (pair, row chunk), its
// waves taking equal row sub-chunks, wave 0 adding the others' tiles in
// wave order; partial of (pair, chunk) at part[(pi * nchunk + c) * WPART],
// tile tt at + 256 tt, element (4r + l/16, l%16) at + 4l + r.  A slab (16
// rows x 64 columns) is read 128 bytes of a column per lane group (lane l:
// row l%16 of column 4u + l/16) with the next slab's loads in flight, and
// turned into the MFMA operand layout (lane l, register r of group t: row
// 4r + l/16 of column 16t + l%16) through the wave's LDS buffer.
constexpr unsigned LDTW = 17u;  // padded column of a wave's slab buffer

struct GramLdsW {
  union {
    d4 red[WTILES][64];  // one wave's tiles at a time
    double T[GRAM_WAVES][64u * LDTW];
  };
};

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_gram_w(const double *__restrict__ G, unsigned ldG,
                                                const uint16_t *__restrict__ pairs,
                                                unsigned nchunk, unsigned CH,
                                                double *__restrict__ part,
                                                const uint8_t *__restrict__ plan)
{
  __shared__ GramLdsW S;
  const unsigned l = threadIdx.x & 63u, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const unsigned pi = blockIdx.x / nchunk, c = blockIdx.x - pi * nchunk;
  if (plan && plan[pi] == PLAN_SKIP)
    return;  // a skipped visit needs no Gram
  const unsigned a = pairs[2u * pi], b = pairs[2u * pi + 1u];
  const unsigned SUB = CH / nw;
  const size_t r0 = (size_t)c * CH + (size_t)w * SUB + (l & 15u);
  // column 4u + l/16 of the pair: block a for u < 8, block b after
  const double *const ga = G + (size_t)(32u * a + (l >> 4)) * ldG + r0;
  const double *const gb = G + (size_t)(32u * b + (l >> 4)) * ldG + r0;
  lds_f64 *const T = (lds_f64 *)S.T[w];
  d4 acc[WTILES];
#pragma unroll
  for (unsigned tt = 0u; tt < WTILES; ++tt)
    acc[tt] = d4{0.0, 0.0, 0.0, 0.0};
  double xn[16];
#pragma unroll
  for (unsigned u = 0u; u < 8u; ++u) {
    xn[u] = ga[(size_t)(4u * u) * ldG];
    xn[8u + u] = gb[(size_t)(4u * u) * ldG];
  }
  for (unsigned s = 0u; s < SUB; s += 16u) {
#pragma unroll
    for (unsigned u = 0u; u < 16u; ++u)
      T[(4u * u + (l >> 4)) * LDTW + (l & 15u)] = xn[u];
    if (s + 16u < SUB) {
#pragma unroll
      for (unsigned u = 0u; u < 8u; ++u) {
        xn[u] = ga[(size_t)(4u * u) * ldG + s + 16u];
        xn[8u + u] = gb[(size_t)(4u * u) * ldG + s + 16u];
      }
    }
    wave_lds_sync();
    d4 X[4];
#pragma unroll
    for (unsigned t = 0u; t < 4u; ++t)
#pragma unroll
      for (unsigned r = 0u; r < 4u; ++r)
        X[t][r] = T[(16u * t + (l & 15u)) * LDTW + 4u * r + (l >> 4)];
    wave_lds_sync();
#pragma unroll
    for (unsigned r = 0u; r < 4u; ++r) {
      unsigned tt = 0u;
#pragma unroll
      for (unsigned ti = 0u; ti < 4u; ++ti)
#pragma unroll
        for (unsigned tj = 0u; tj <= ti; ++tj, ++tt)
          acc[tt] = mfma(X[ti][r], X[tj][r], acc[tt]);
    }
  }
  // wave 0 adds the other waves' tiles in wave order, one wave's at a time
  // through the (former slab) buffer
  for (unsigned u = 1u; u < nw; ++u) {
    __syncthreads();
    if (w == u) {
#pragma unroll
      for (unsigned tt = 0u; tt < WTILES; ++tt)
        S.red[tt][l] = acc[tt];
    }
    __syncthreads();
    if (w == 0u) {
#pragma unroll
      for (unsigned tt = 0u; tt < WTILES; ++tt)
        acc[tt] = acc[tt] + S.red[tt][l];
    }
  }
  if (w == 0u) {
    double *const o = part + (size_t)blockIdx.x * WPART;
#pragma unroll
    for (unsigned tt = 0u; tt < WTILES; ++tt)
      *reinterpret_cast<d4 *>(o + tt * 256u + 4u * l) = acc[tt];
  }
}

// Sum of the 64 rows of one pivot pair held by 16 lanes, lane r holding
// rows r + 16i as v[i]: the xor butterfly 32,16 inside the lane, 8 (a
// rotation by 8 inside the 16-lane DPP row), 4,2,1 across lanes.
__device__ __forceinline__ double lane_xor8(double v) { return dpp_mov<0x128>(v); }

__device__ __forceinline__ double sum16x4(const double *v)
{
  double s = (v[0] + v[2]) + (v[1] + v[3]);
  s = s + lane_xor8(s);
  s = s + lane_xor4(s);
  s = s + lane_xor2(s);
  s = s + lane_xor1(s);
  return s;
}

constexpr unsigned VT = 512u;  // threads of a visit: 32 pivot pairs x 16 lanes

// In-place 64x64 Cholesky by the workgroup's eight waves, in LDS (lane x
// of every wave standing for row x of the lower triangle):
// share/device_code_common_Cholesky.hpp:58-158 (USE_RSQRT).  Per pivot k,
// wave 0 scales column k (rsqrt(A(k,k)) below the diagonal, sqrt(A(k,k)) on
// it), copies it to col and forms the next pivot A(k+1,k+1) - L(k+1,k)^2;
// after a barrier it computes that pivot's rsqrt and sqrt while waves 1-7
// apply A(x,j) = fma(-A(x,k), A(j,k), A(x,j)) to the columns j > k in
// batches of 8 (a batch's loads, then its fmas, then its stores; rows above
// j produce values never read).  Then R = L^T, zeros below the diagonal.
__device__ __forceinline__ void chol64_block(double *A_, double *col_, unsigned tid)
{
  lds_f64 *const A = (lds_f64 *)A_;
  lds_f64 *const col = (lds_f64 *)col_;
  const unsigned x = tid & 63u, w = tid >> 6;
  double akk = 0.0, rk = 0.0, dk = 0.0;
  if (w == 0u) {
    akk = readlane_d(A[x], 0);
    rk = dev_rsqrt(akk);
    dk = sqrt(akk);
  }
  for (unsigned k = 0u; k < 64u; ++k) {
    const unsigned k1 = k + 1u;
    if (w == 0u) {
      double v = A[k * LDRW + x];
      v = (x > k) ? v * rk : ((x == k) ? dk : v);
      A[k * LDRW + x] = v;
      col[x] = v;
      if (k1 < 64u) {
        const double lk = readlane_d(v, k1);
        const double d1 = readlane_d(A[k1 * LDRW + x], k1);
        akk = __fma_rn(-lk, lk, d1);
      }
    }
    __syncthreads();
    if (k1 < 64u) {
      if (w == 0u) {
        rk = dev_rsqrt(akk);
        dk = sqrt(akk);
      }
      else {
        const double v = col[x];
        for (unsigned jb = (k1 & ~7u) + 8u * (w - 1u); jb < 64u; jb += 8u * (VT / 64u - 1u)) {
          double av[8], cv[8];
#pragma unroll
          for (unsigned i = 0u; i < 8u; ++i) {
            av[i] = A[(jb + i) * LDRW + x];
            cv[i] = col[jb + i];
          }
#pragma unroll
          for (unsigned i = 0u; i < 8u; ++i)
            if (jb + i >= k1) av[i] = __fma_rn(-v, cv[i], av[i]);
#pragma unroll
          for (unsigned i = 0u; i < 8u; ++i)
            A[(jb + i) * LDRW + x] = av[i];
        }
      }
    }
    __syncthreads();
  }
  // R(r, c) = L(c, r) for r <= c: all reads, a barrier, all writes
  double t[4096u / VT];
#pragma unroll
  for (unsigned i = 0u; i < 4096u / VT; ++i) {
    const unsigned e = tid + VT * i, r = e & 63u, c = e >> 6;
    t[i] = (r <= c) ? A[r * LDRW + c] : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (unsigned i = 0u; i < 4096u / VT; ++i) {
    const unsigned e = tid + VT * i, r = e & 63u, c = e >> 6;
    A[c * LDRW + r] = t[i];
  }
}

constexpr unsigned S0W_BYTES = 64u * 32u * 2u;  // level-0 table of order 64 (<= 64 steps)

struct PairLdsW {
  double R[64u * LDRW];
  double W[64u * LDRW];
  double col[64];
  unsigned cnt_s[32], cnt_b[32];
  float cmax[32];
  float bnorm[2];
  int swp;
  uint8_t S0[S0W_BYTES];
};

// One 64-column block pair per workgroup of 512 threads: chunk partials ->
// R (lower), Cholesky, level-0 sweeps on R accumulating W, counters, W and
// the update flag out.  The arithmetic of pair_visit at 64 columns.
template <bool SORT>
__global__ __launch_bounds__(512) void k_pair_w(const uint16_t *__restrict__ pairs,
                                                const double *__restrict__ part, unsigned nchunk,
                                                double *__restrict__ Wout,
                                                unsigned *__restrict__ flags,
                                                RunCtl *__restrict__ ctl,
                                                const uint8_t *__restrict__ strat0,
                                                unsigned s0steps, unsigned nswp, int definite,
                                                unsigned nplus, const uint8_t *__restrict__ plan,
                                                VisitRec *__restrict__ rec)
{
  __shared__ PairLdsW L;
  const unsigned tid = threadIdx.x, pi = blockIdx.x;
  unsigned long long *const cvg = &ctl->cvg;
  if (plan && plan[pi] == PLAN_SKIP) {
    // skipped visit: no update, the record keeps the last executed visit
    if (tid == 0u) {
      flags[pi] = 0u;
      rec[pi].skips += 1u;
      atomicAdd(&ctl->skipped, 1u);
    }
    return;
  }
  double *const R = L.R;
  double *const W = L.W;
  const unsigned iBlk = pairs[2u * pi], jBlk = pairs[2u * pi + 1u];
  // local J structure, share/device_code_nosort_accumV.hpp:53-76
  bool hyp = false;
  int def = definite;
  unsigned npos = 0u;
  if (!definite) {
    const unsigned ifc0 = iBlk * 32u, ifc1 = jBlk * 32u;
    if (nplus <= ifc0) { npos = 0u; def = -1; }
    else if (nplus <= ifc0 + 32u) { npos = nplus - ifc0; def = 0; }
    else if (nplus <= ifc1) { npos = 32u; def = 0; }
    else if (nplus <= ifc1 + 32u) { npos = 32u + (nplus - ifc1); def = 0; }
    else { npos = 64u; def = +1; }
    hyp = (def == 0);
  }
  if (tid == 0u)
    L.swp = 0;
  for (unsigned e = tid; e < 2u * 32u * s0steps; e += VT)
    L.S0[e] = strat0[e];
  {
    const double *const pp = part + (size_t)pi * nchunk * WPART;
    // two entries a thread at a time, all their chunk loads in flight
    // together; the chunks summed in chunk order
#pragma unroll 1
    for (unsigned q = 0u; q < WPART / VT; ++q) {
      double v[1][MAX_GRAM_CHUNKS];
#pragma unroll
      for (unsigned h = 0u; h < 1u; ++h)
#pragma unroll
        for (unsigned c = 0u; c < MAX_GRAM_CHUNKS; ++c)
          if (c < nchunk) v[h][c] = pp[(size_t)c * WPART + tid + VT * (q + h)];
#pragma unroll
      for (unsigned h = 0u; h < 1u; ++h) {
        const unsigned e = tid + VT * (q + h);
        double s = v[h][0];
#pragma unroll
        for (unsigned c = 1u; c < MAX_GRAM_CHUNKS; ++c)
          if (c < nchunk) s = s + v[h][c];
        const unsigned tt = e >> 8, ti = tile_i(tt), tj = tt - ti * (ti + 1u) / 2u;
        const unsigned l = (e >> 2) & 63u, r = e & 3u;
        RW(R, 16u * ti + 4u * r + (l >> 4), 16u * tj + (l & 15u)) = s;
      }
    }
  }
  for (unsigned e = tid; e < 4096u; e += VT)
    RW(W, e & 63u, e >> 6) = ((e & 63u) == (e >> 6)) ? 1.0 : 0.0;
  __syncthreads();
  if (rec && tid < 2u) {
    // squared Frobenius norms of the two blocks, for the visit record
    double sn = 0.0;
    for (unsigned k = 0u; k < 32u; ++k)
      sn = sn + RW(R, 32u * tid + k, 32u * tid + k);
    L.bnorm[tid] = (float)log2(sn);
  }
  __syncthreads();
  chol64_block(R, L.col, tid);
  __syncthreads();

  const unsigned y = tid >> 4, r = tid & 15u;
  unsigned blk_s = 0u, blk_b = 0u;
  bool swapped = false;
  double my_cos = 0.0;  // max |Apq| / sqrt(Dp Dq) of the slot's pivots, first inner sweep
  for (unsigned swp = 0u; swp < nswp; ++swp) {
    unsigned my_s = 0u, my_b = 0u;
    for (unsigned step = 0u; step < s0steps; ++step) {
      const unsigned p = L.S0[(step * 32u + y) * 2u], q = L.S0[(step * 32u + y) * 2u + 1u];
      double Gp[4], Gq[4], Vp[4], Vq[4], t0[4], t1[4], t2[4];
#pragma unroll
      for (unsigned i = 0u; i < 4u; ++i) {
        Gp[i] = RW(R, r + 16u * i, p);
        Gq[i] = RW(R, r + 16u * i, q);
        Vp[i] = RW(W, r + 16u * i, p);
        Vq[i] = RW(W, r + 16u * i, q);
      }
#pragma unroll
      for (unsigned i = 0u; i < 4u; ++i) {
        t0[i] = __fma_rn(Gp[i], Gp[i], 0.0);
        t1[i] = __fma_rn(Gq[i], Gq[i], 0.0);
        t2[i] = __fma_rn(Gp[i], Gq[i], 0.0);
      }
      double Dp = sum16x4(t0), Dq = sum16x4(t1);
      const double Apq = sum16x4(t2);
      const double Dp_ = sqrt(Dp), Dq_ = sqrt(Dq), Apq_ = fabs(Apq), Dpq_ = Dp_ * Dq_;
      const int transf_s = !(Apq_ < (Dpq_ * WTOL));
      if (rec && swp == 0u && Dpq_ > 0.0)
        my_cos = fmax(my_cos, Apq_ / Dpq_);
      int pdef = def;
      if (hyp) {
        if ((p < npos) && (q < npos)) pdef = +1;
        else if ((p >= npos) && (q >= npos)) pdef = -1;
        else pdef = 0;
      }
      bool trig;
      if (!SORT) trig = !hyp || pdef;
      else trig = (pdef != 0);
      double c, t;
      const int rot_b = dev_rot(trig, Apq, Dp, Dq, c, t);
      int transf_b = 0;
      if (transf_s) {
        transf_b = rot_b;
        const double t_ = trig ? -t : t;
        if (SORT && trig) {
          Dp = __fma_rn(t, Apq, Dp);
          Dq = __fma_rn(t_, Apq, Dq);
        }
        bool keep = true;
        if (SORT)
          keep = (!pdef) || ((pdef > 0) && (Dp >= Dq)) || ((pdef < 0) && (Dp <= Dq));
        const unsigned P = keep ? p : q, Q = keep ? q : p;
#pragma unroll
        for (unsigned i = 0u; i < 4u; ++i) {
          const double rp = __fma_rn(t_, Gq[i], Gp[i]), rq = __fma_rn(t, Gp[i], Gq[i]);
          const double wp = __fma_rn(t_, Vq[i], Vp[i]), wq = __fma_rn(t, Vp[i], Vq[i]);
          RW(R, r + 16u * i, P) = c * rp;
          RW(R, r + 16u * i, Q) = c * rq;
          RW(W, r + 16u * i, P) = c * wp;
          RW(W, r + 16u * i, Q) = c * wq;
        }
      }
      else if (SORT && (((pdef > 0) && (Dp < Dq)) || ((pdef < 0) && (Dq < Dp)))) {
#pragma unroll
        for (unsigned i = 0u; i < 4u; ++i) {
          RW(R, r + 16u * i, p) = Gq[i];
          RW(R, r + 16u * i, q) = Gp[i];
          RW(W, r + 16u * i, p) = Vq[i];
          RW(W, r + 16u * i, q) = Vp[i];
        }
        swapped = true;
      }
      my_s += (unsigned)transf_s;
      my_b += (unsigned)transf_b;
      __syncthreads();
    }
    if (r == 0u) {
      L.cnt_s[y] = my_s;
      L.cnt_b[y] = my_b;
    }
    __syncthreads();
    unsigned swp_s = 0u, swp_b = 0u;
    for (unsigned i = 0u; i < 32u; ++i) {
      swp_s += L.cnt_s[i];
      swp_b += L.cnt_b[i];
    }
    __syncthreads();
    if (!swp_s)
      break;
    blk_s += swp_s;
    blk_b += swp_b;
  }
  if (swapped)
    L.swp = 1;
  if (r == 0u)
    L.cmax[y] = (float)my_cos;
  __syncthreads();
  if (tid == 0u) {
    // convergence counter, share/device_code_nosort.hpp:94-103
    if (blk_s) {
      if (blk_b)
        atomicAdd(cvg, (((unsigned long long)blk_b) << 32) | blk_s);
      else
        atomicAdd(reinterpret_cast<unsigned *>(cvg), blk_s);
    }
    flags[pi] = (blk_s || L.swp) ? 1u : 0u;
    if (rec) {
      // the record in the units of a 32-column pair's inner sweep (496 pivot
      // pairs): this pair's counts scaled from its 2016
      float cm = 0.0f;
      for (unsigned i = 0u; i < 32u; ++i)
        cm = fmaxf(cm, L.cmax[i]);
      VisitRec v;
      v.rot_s = (min(blk_s, 2016u) * 496u + 1008u) / 2016u;
      v.rot_b = (min(blk_b, 2016u) * 496u + 1008u) / 2016u;
      v.maxcos = cm;
      v.l2ni = L.bnorm[0];
      v.l2nj = L.bnorm[1];
      v.sweep = ctl->sweep;
      v.skips = 0u;
      v.flags = 0u;
      rec[pi] = v;
    }
  }
  double *const wo = Wout + (size_t)pi * 4096u;
  for (unsigned e = tid; e < 4096u; e += VT)
    wo[e] = RW(W, e & 63u, e >> 6);
}

// Block update M[:, pair] <- M[:, pair] W of the pairs of one step: a
// workgroup per (pair, 512-row chunk), W in LDS (row-major, padded), each
// wave 16-row slabs of the chunk with the next slab's loads in flight.  Per
// slab the MFMA chains run over k = 0..63 in order: A = W^T (lane l:
// W(4u + l/16, 16h + l%16)), B = the slab's transpose (lane l:
// M(row + l%16, column 4u + l/16)); lane l, register q of tile h holds
// output column 16h + 4q + l/16 of row row + l%16.  Pairs whose W is the
// identity (flag 0) are left alone.
constexpr unsigned UPW_ROWS = 512u;
constexpr unsigned LDWW = 65u;

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_update_w(
    double *__restrict__ M, unsigned ldM, unsigned nrow, const uint16_t *__restrict__ pairs,
    const double *__restrict__ Wbuf, const unsigned *__restrict__ flags, unsigned nchunk)
{
  __shared__ double Wl[64u * LDWW];
  const unsigned pi = blockIdx.x / nchunk, c = blockIdx.x - pi * nchunk;
  if (!flags[pi])
    return;
  const unsigned l = threadIdx.x & 63u, w = threadIdx.x >> 6;
  const unsigned a = pairs[2u * pi], b = pairs[2u * pi + 1u];
  const double *const Wp = Wbuf + (size_t)pi * 4096u;  // W(k, j) at 64 j + k
  for (unsigned e = threadIdx.x; e < 4096u; e += blockDim.x)
    Wl[(e & 63u) * LDWW + (e >> 6)] = Wp[e];  // Wl[k * LDWW + j] = W(k, j)
  __syncthreads();
  const unsigned rbeg = c * UPW_ROWS, rend = min(rbeg + UPW_ROWS, nrow);
  // column 4u + l/16 of the pair: block a for u < 8, block b after
  double *const ca = M + (size_t)(32u * a + (l >> 4)) * ldM + (l & 15u);
  double *const cb = M + (size_t)(32u * b + (l >> 4)) * ldM + (l & 15u);
  const double *const wl = Wl + (l >> 4) * LDWW + (l & 15u);
  unsigned row = rbeg + 16u * w;
  double xn[16];
  if (row < rend) {
#pragma unroll
    for (unsigned u = 0u; u < 8u; ++u) {
      xn[u] = ca[(size_t)(4u * u) * ldM + row];
      xn[8u + u] = cb[(size_t)(4u * u) * ldM + row];
    }
  }
  for (; row < rend; row += 64u) {
    double x[16];
#pragma unroll
    for (unsigned u = 0u; u < 16u; ++u)
      x[u] = xn[u];
    if (row + 64u < rend) {
#pragma unroll
      for (unsigned u = 0u; u < 8u; ++u) {
        xn[u] = ca[(size_t)(4u * u) * ldM + row + 64u];
        xn[8u + u] = cb[(size_t)(4u * u) * ldM + row + 64u];
      }
    }
    d4 acc[4];
#pragma unroll
    for (unsigned h = 0u; h < 4u; ++h)
      acc[h] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (unsigned u = 0u; u < 16u; ++u) {
      const double *const wr = wl + 4u * u * LDWW;
#pragma unroll
      for (unsigned h = 0u; h < 4u; ++h)
        acc[h] = mfma(wr[16u * h], x[u], acc[h]);
    }
#pragma unroll
    for (unsigned h = 0u; h < 4u; ++h)
#pragma unroll
      for (unsigned q = 0u; q < 4u; ++q)
        ((h < 2u) ? ca : cb)[(size_t)(16u * (h & 1u) + 4u * q) * ldM + row] = acc[h][q];
  }
}

}  // namespace

unsigned wide_part_doubles() { return WPART; }
unsigned wide_update_chunks(unsigned nrow) { return (nrow + UPW_ROWS - 1u) / UPW_ROWS; }

void launch_gram_w(hipStream_t s, const double *G, unsigned ldG, const uint16_t *pairs,
                   unsigned np, unsigned nchunk, unsigned CH, double *part, const uint8_t *plan)
{
  const unsigned nw = gram_waves(CH);
  hipLaunchKernelGGL(k_gram_w, dim3(np * nchunk), dim3(64u * nw), 0, s, G, ldG, pairs, nchunk, CH,
                     part, plan);
}

void launch_pair_w(hipStream_t s, bool sort, const uint16_t *pairs, unsigned np,
                   const double *part, unsigned nchunk, double *W, unsigned *flags,
                   RunCtl *ctl, const uint8_t *strat0, unsigned s0steps,
                   unsigned nswp, int definite, unsigned nplus, const uint8_t *plan,
                   VisitRec *rec)
{
  if (sort)
    hipLaunchKernelGGL(k_pair_w<true>, dim3(np), dim3(VT), 0, s, pairs, part, nchunk, W, flags,
                       ctl, strat0, s0steps, nswp, definite, nplus, plan, rec);
  else
    hipLaunchKernelGGL(k_pair_w<false>, dim3(np), dim3(VT), 0, s, pairs, part, nchunk, W, flags,
                       ctl, strat0, s0steps, nswp, definite, nplus, plan, rec);
}

void launch_update_w(hipStream_t s, double *M, unsigned ldM, unsigned nrow,
                     const uint16_t *pairs, unsigned np, const double *W,
                     const unsigned *flags)
{
  const unsigned nch = wide_update_chunks(nrow);
  hipLaunchKernelGGL(k_update_w, dim3(np * nch), dim3(256), 0, s, M, ldM, nrow, pairs, W, flags,
                     nch);
}

}  // namespace hj
