// hj_batch.hip -- many independent (H)SVDs of uneven sizes in one persistent
// launch.
//
// Every workgroup takes whole problems from an atomic queue (longest first)
// and runs the blocked one-sided Jacobi method on each: level-2 sweeps in
// strategy order, one level-1 step at a time; the block pairs of a step are
// spread over the workgroup's waves, one wave doing everything for its pair
// (FP64 MFMA Gram over all rows, 32x32 Cholesky, the level-0 inner sweeps on
// R in the wave's LDS region, FP64 MFMA update of G and V), and a workgroup
// barrier separates the steps.  Workgroups never wait on each other, so the
// grid need not be resident.  The arithmetic of a problem is that of the
// single-problem path with the Gram of a pair formed by one wave over all
// rows in order (the oracle's single-wave Gram order), so each problem is
// checked bit for bit against the oracle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/hypjac.h"
#include "hj_device.hpp"
#include "hj_fused.hpp"
#include "hj_kernels.hpp"
#include "hj_strat.hpp"
#include "hj_util.hpp"

typedef double d4 __attribute__((ext_vector_type(4)));

namespace hj {
namespace {

constexpr unsigned BLDR = L0LD;                       // padded leading dimension (the level-0 layout)
// doubles per wave: R and W (the fused pass: the W of both pairs, LDW rows),
// and the 16x16 transpose buffer of the fused pass's stores
constexpr unsigned WAVE_LDS = 2u * 32u * BLDR + 16u * LDT;
static_assert(BLDR == LDW, "the fused pass keeps its W copies in the R / W regions");
constexpr unsigned BATCH_MAX_WAVES = 8u;
// waves per SIMD the batch kernel is compiled for (register budget 512/n)
#ifndef BATCH_WAVES_PER_SIMD
#define BATCH_WAVES_PER_SIMD 2
#endif

struct BProb {
  double *G, *V, *D;
  unsigned nrow, ncol, nplus, ldG, ldV, routine;
  unsigned s1off;  // first pair of this problem's strategy table in the pool
  unsigned steps, np;
  unsigned qoff;  // first closed quad of this problem's table in the quad pool, or NO_QUADS
};
constexpr unsigned NO_QUADS = 0xFFFFFFFFu;

struct BOut {
  unsigned sweeps, pad;
  unsigned long long rot_s, rot_b;
  unsigned long long t0, t1;  // wall clock (100 MHz) at start and end
};

__device__ __forceinline__ d4 mfma64(double a, double b, d4 c)
{
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

#define BR(A, r, c) (A)[(c) * BLDR + (r)]

// Gram of the pair (columns 16 iBlk.., 16 jBlk..) over all nrow rows by one
// wave: 16-row slabs in order, MFMA step j taking row 4j + k in k-slot k
// (the order of the single-problem Gram with one chunk and one wave); result
// into R (both 16x16 diagonal tiles and the lower off-diagonal tile).
__device__ void wave_gram(const double *__restrict__ G, unsigned ldG, unsigned nrow, unsigned iBlk,
                          unsigned jBlk, double *R, unsigned l)
{
  const unsigned r0 = l >> 4;
  const double *const g0 = G + (size_t)(16u * iBlk + (l & 15u)) * ldG + r0;
  const double *const g1 = G + (size_t)(16u * jBlk + (l & 15u)) * ldG + r0;
  d4 a00 = {0.0, 0.0, 0.0, 0.0}, a10 = a00, a11 = a00;
  for (unsigned g = 0u; g < nrow; g += 32u) {
    double x0[8], x1[8];
#pragma unroll
    for (unsigned t = 0u; t < 8u; ++t) {
      x0[t] = g0[g + 4u * t];
      x1[t] = g1[g + 4u * t];
    }
#pragma unroll
    for (unsigned t = 0u; t < 8u; ++t) {
      a00 = mfma64(x0[t], x0[t], a00);
      a10 = mfma64(x1[t], x0[t], a10);
      a11 = mfma64(x1[t], x1[t], a11);
    }
  }
  // lane l, register r holds element (4r + l/16, l%16) of each tile
#pragma unroll
  for (unsigned r = 0u; r < 4u; ++r) {
    const unsigned i = 4u * r + (l >> 4), j = l & 15u;
    BR(R, i, j) = a00[r];
    BR(R, 16u + i, j) = a10[r];
    BR(R, 16u + i, 16u + j) = a11[r];
  }
}

// [M_i M_j] <- [M_i M_j] W over rows [0, nrow) by one wave, 16-row slabs
// (the arithmetic of k_update: per element an fma chain over k = 0..31).
__device__ void wave_update(double *__restrict__ M, unsigned ld, unsigned nrow, unsigned iBlk,
                            unsigned jBlk, const double *W, unsigned l, unsigned wld)
{
  // W(k, j) at W[j * wld + k] (the LDS copy, wld = BLDR, or the scratch, 32)
  double a0[8], a1[8];
#pragma unroll
  for (unsigned u = 0u; u < 8u; ++u) {
    a0[u] = W[(l & 15u) * wld + 4u * u + (l >> 4)];
    a1[u] = W[(16u + (l & 15u)) * wld + 4u * u + (l >> 4)];
  }
  double *cp[8];
#pragma unroll
  for (unsigned u = 0u; u < 8u; ++u) {
    const unsigned cc = 4u * u + (l >> 4);
    const unsigned gc = (cc < 16u) ? (16u * iBlk + cc) : (16u * jBlk + cc - 16u);
    cp[u] = M + (size_t)gc * ld + (l & 15u);
  }
  double *op0[4], *op1[4];
#pragma unroll
  for (unsigned q = 0u; q < 4u; ++q) {
    const unsigned cc = 4u * q + (l >> 4);
    op0[q] = M + (size_t)(16u * iBlk + cc) * ld + (l & 15u);
    op1[q] = M + (size_t)(16u * jBlk + cc) * ld + (l & 15u);
  }
  double nb[8];
#pragma unroll
  for (unsigned u = 0u; u < 8u; ++u)
    nb[u] = cp[u][0];
  for (unsigned r = 0u; r < nrow; r += 16u) {
    double b[8];
#pragma unroll
    for (unsigned u = 0u; u < 8u; ++u)
      b[u] = nb[u];
    if (r + 16u < nrow) {
#pragma unroll
      for (unsigned u = 0u; u < 8u; ++u)
        nb[u] = cp[u][r + 16u];
    }
    d4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = acc0;
#pragma unroll
    for (unsigned u = 0u; u < 8u; ++u) {
      acc0 = mfma64(a0[u], b[u], acc0);
      acc1 = mfma64(a1[u], b[u], acc1);
    }
#pragma unroll
    for (unsigned q = 0u; q < 4u; ++q) {
      op0[q][r] = acc0[q];
      op1[q][r] = acc1[q];
    }
  }
}

// The visit of one block pair by one wave on the Gram in R (BR layout, the
// lower triangle): Cholesky, level-0 sweep(s), W in W.  Returns the rotation
// counts of the pair (s, b) in every lane, and whether W differs from the
// identity (a rotation or a swap: the update is left out otherwise).
template <bool SORT>
__device__ __noinline__ void wave_visit(const BProb &P, unsigned iBlk, unsigned jBlk, double *R, double *W,
                           const uint8_t *S0, unsigned s0steps, unsigned nswp, int definite,
                           unsigned l, unsigned &out_s, unsigned &out_b, bool &out_upd)
{
  // local J structure, share/device_code_nosort_accumV.hpp:53-76
  bool hyp = false;
  int def = definite;
  unsigned npos = 0u;
  if (!definite) {
    const unsigned ifc0 = iBlk * 16u, ifc1 = jBlk * 16u, nplus = P.nplus;
    if (nplus <= ifc0) { npos = 0u; def = -1; }
    else if (nplus <= ifc0 + 16u) { npos = nplus - ifc0; def = 0; }
    else if (nplus <= ifc1) { npos = 16u; def = 0; }
    else if (nplus <= ifc1 + 16u) { npos = 16u + (nplus - ifc1); def = 0; }
    else { npos = 32u; def = +1; }
    hyp = (def == 0);
  }

  // Cholesky, share/device_code_common_Cholesky.hpp:58-158 (USE_RSQRT)
  (void)chol32_wave<BLDR>(R, W, l, false);  // W is free until the identity is written below
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (unsigned e = 0u; e < 16u; ++e) {
    const unsigned k = l + 64u * e;
    BR(W, k & 31u, k >> 5) = ((k & 31u) == (k >> 5)) ? 1.0 : 0.0;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

  // level-0 sweeps (slot y = l/4 of the 16 pivot pairs, lane l%4 of the
  // slot holding rows l%4 + 4k)
  unsigned blk_s = 0u, blk_b = 0u;
  bool swapped = false;
  double cos_unused = 0.0;
  wave_level0<SORT, false>(R, W, S0, s0steps, nswp, def, hyp, npos, l, blk_s, blk_b, swapped,
                           cos_unused);
  out_s = blk_s;
  out_b = blk_b;
  out_upd = (blk_s != 0u) || swapped;
}

// One block pair by one wave: Gram, visit, update (the unfused step).
template <bool SORT>
__device__ void wave_pair(const BProb &P, unsigned iBlk, unsigned jBlk, double *R, double *W,
                          const uint8_t *S0, unsigned s0steps, unsigned nswp, int definite,
                          unsigned l, unsigned &out_s, unsigned &out_b)
{
  wave_gram(P.G, P.ldG, P.nrow, iBlk, jBlk, R, l);
  wave_lds_sync();
  bool upd = false;
  wave_visit<SORT>(P, iBlk, jBlk, R, W, S0, s0steps, nswp, definite, l, out_s, out_b, upd);
  if (upd) {
    wave_update(P.G, P.ldG, P.nrow, iBlk, jBlk, W, l, BLDR);
    if (P.routine & HYPJAC_FULL_SVD)
      wave_update(P.V, P.ldV, P.ncol, iBlk, jBlk, W, l, BLDR);
  }
}

// Gram tiles of a pair (00, 10, 11; lane l, register r at l*4 + r, the MFMA
// output layout) from the fused pass's scratch into R, as wave_gram leaves
// them
__device__ __forceinline__ void gram_to_R(const double *__restrict__ g, double *R, unsigned l)
{
#pragma unroll
  for (unsigned r = 0u; r < 4u; ++r) {
    const unsigned i = 4u * r + (l >> 4), j = l & 15u;
    BR(R, i, j) = g[l * 4u + r];
    BR(R, 16u + i, j) = g[256u + l * 4u + r];
    BR(R, 16u + i, 16u + j) = g[512u + l * 4u + r];
  }
}

// The fused transition of a closed quad by one wave over all rows: the
// update of both pairs of step s (W0, W1 in LDS, W(k, j) at k*LDW + j) and
// the Gram of both pairs of step s+1 from the updated values in registers,
// in the single-wave Gram's order (16-row slabs ascending), to g0, g1.
__device__ void wave_upgram(double *__restrict__ G, unsigned ldG, unsigned nrow, unsigned A0,
                            unsigned B0, unsigned A1, unsigned B1, bool u0, bool u1, const Quad &Q,
                            const double *W0, const double *W1, double *T, unsigned l,
                            double *__restrict__ g0, double *__restrict__ g1)
{
  d4 z = {0.0, 0.0, 0.0, 0.0};
  d4 g00 = z, g10 = z, g11 = z, h00 = z, h10 = z, h11 = z;
  double xa[8], xb[8];
  slab_load(G, ldG, 0u, A0, B0, u0, l, xa);
  slab_load(G, ldG, 0u, A1, B1, u1, l, xb);
  for (unsigned r = 0u; r < nrow; r += 16u) {
    double ca[8], cb[8];
#pragma unroll
    for (unsigned u = 0u; u < 8u; ++u) {
      ca[u] = xa[u];
      cb[u] = xb[u];
    }
    if (r + 16u < nrow) {
      slab_load(G, ldG, r + 16u, A0, B0, u0, l, xa);
      slab_load(G, ldG, r + 16u, A1, B1, u1, l, xb);
    }
    d4 t0, t1, t2, t3;
    slab_compute(G, ldG, r, A0, B0, u0, W0, l, ca, t0, t1, T);
    slab_compute(G, ldG, r, A1, B1, u1, W1, l, cb, t2, t3, T);
    {
      const d4 X = sel4(Q.x0, t0, t1, t2, t3), Y = sel4(Q.y0, t0, t1, t2, t3);
      gram_fp64_slab(X, Y, g00, g10, g11);
    }
    {
      const d4 X = sel4(Q.x1, t0, t1, t2, t3), Y = sel4(Q.y1, t0, t1, t2, t3);
      gram_fp64_slab(X, Y, h00, h10, h11);
    }
  }
#pragma unroll
  for (unsigned r = 0u; r < 4u; ++r) {
    g0[l * 4u + r] = g00[r];
    g0[256u + l * 4u + r] = g10[r];
    g0[512u + l * 4u + r] = g11[r];
    g1[l * 4u + r] = h00[r];
    g1[256u + l * 4u + r] = h10[r];
    g1[512u + l * 4u + r] = h11[r];
  }
}

__global__ void k_batch_reset(unsigned *queue)
{
  if (threadIdx.x == 0u) *queue = 0u;
}

__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(BATCH_WAVES_PER_SIMD))) void k_batch(const BProb *__restrict__ probs,
                                               const unsigned *__restrict__ order, unsigned nprob,
                                               unsigned *__restrict__ queue,
                                               const uint16_t *__restrict__ s1pool,
                                               const uint8_t *__restrict__ strat0, unsigned s0steps,
                                               unsigned max_sweeps, BOut *__restrict__ out,
                                               const Quad *__restrict__ qpool,
                                               double *__restrict__ scr,
                                               unsigned *__restrict__ sflags, unsigned np_max)
{
  extern __shared__ double lds[];
  __shared__ uint8_t S0[1024];
  __shared__ unsigned s_task;
  __shared__ unsigned cnt[2u * BATCH_MAX_WAVES];
  const unsigned l = threadIdx.x & 63u, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  double *const R = lds + (size_t)wv * WAVE_LDS;
  double *const W = R + 32u * BLDR;
  double *const T = W + 32u * BLDR;
  // this workgroup's scratch of the fused steps: the Grams of the next
  // step's pairs, the W and update flags of the running step's pairs
  double *const sg = scr + (size_t)blockIdx.x * np_max * (768u + 1024u);
  double *const sw = sg + (size_t)np_max * 768u;
  unsigned *const sf = sflags + (size_t)blockIdx.x * np_max;
  for (unsigned e = threadIdx.x; e < s0steps * 32u; e += blockDim.x)
    S0[e] = strat0[e];
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0u)
      s_task = atomicAdd(queue, 1u);
    __syncthreads();
    const unsigned task = s_task;
    if (task >= nprob)
      break;
    const unsigned pid = order[task];
    const BProb P = probs[pid];
    unsigned long long t0 = 0ull;
    if (threadIdx.x == 0u)
      t0 = wall_clock64();
    const bool sort = (P.routine & HYPJAC_CDSORT_1) != 0u;
    const bool full = (P.routine & HYPJAC_FULL_SVD) != 0u;
    const unsigned nswp = (P.routine & HYPJAC_BLK_ORI) ? 1u : HYPJAC_NSWEEP;
    const int definite = P.nplus ? ((P.nplus < P.ncol) ? 0 : +1) : -1;
    if (full) {
      // V <- I (columns 0..ncol-1, rows 0..ncol-1)
      const size_t tot = (size_t)P.ncol * P.ncol;
      for (size_t e = threadIdx.x; e < tot; e += blockDim.x) {
        const unsigned i = (unsigned)(e % P.ncol), j = (unsigned)(e / P.ncol);
        P.V[(size_t)j * P.ldV + i] = (i == j) ? 1.0 : 0.0;
      }
    }
    unsigned long long tot_s = 0ull, tot_b = 0ull;
    unsigned sweeps = 0u;
    const uint16_t *const tab = s1pool + 2u * (size_t)P.s1off;
    while (sweeps < max_sweeps) {
      unsigned sw_s = 0u, sw_b = 0u;
      for (unsigned st = 0u; st < P.steps; ++st) {
        __syncthreads();
        const bool fused = (P.qoff != NO_QUADS) && (st + 1u < P.steps);
        if (!fused && (P.qoff == NO_QUADS || st == 0u)) {
          // unfused step: each wave's pairs, Gram -> visit -> update
          for (unsigned pi = wv; pi < P.np; pi += nw) {
            const unsigned iBlk = tab[2u * (st * P.np + pi)], jBlk = tab[2u * (st * P.np + pi) + 1u];
            unsigned s = 0u, b = 0u;
            if (sort)
              wave_pair<true>(P, iBlk, jBlk, R, W, S0, s0steps, nswp, definite, l, s, b);
            else
              wave_pair<false>(P, iBlk, jBlk, R, W, S0, s0steps, nswp, definite, l, s, b);
            sw_s += s;
            sw_b += b;
          }
          continue;
        }
        // fused steps: the visits (Gram of step 0 from G, later ones from
        // the previous transition's pass), then one pass per closed quad
        // updating both pairs and forming the next step's Grams; the last
        // step of a sweep updates after each visit
        for (unsigned pi = wv; pi < P.np; pi += nw) {
          const unsigned iBlk = tab[2u * (st * P.np + pi)], jBlk = tab[2u * (st * P.np + pi) + 1u];
          if (st == 0u)
            wave_gram(P.G, P.ldG, P.nrow, iBlk, jBlk, R, l);
          else
            gram_to_R(sg + (size_t)pi * 768u, R, l);
          wave_lds_sync();
          unsigned s = 0u, b = 0u;
          bool upd = false;
          if (sort)
            wave_visit<true>(P, iBlk, jBlk, R, W, S0, s0steps, nswp, definite, l, s, b, upd);
          else
            wave_visit<false>(P, iBlk, jBlk, R, W, S0, s0steps, nswp, definite, l, s, b, upd);
          sw_s += s;
          sw_b += b;
          if (!fused) {
            if (upd) {
              wave_update(P.G, P.ldG, P.nrow, iBlk, jBlk, W, l, BLDR);
              if (full)
                wave_update(P.V, P.ldV, P.ncol, iBlk, jBlk, W, l, BLDR);
            }
          }
          else {
            double *const w = sw + (size_t)pi * 1024u;
#pragma unroll
            for (unsigned e = 0u; e < 16u; ++e) {
              const unsigned k = l + 64u * e;
              w[k] = BR(W, k & 31u, k >> 5);
            }
            if (l == 0u)
              sf[pi] = upd ? 1u : 0u;
          }
          wave_lds_sync();
        }
        if (!fused)
          continue;
        __syncthreads();
        const unsigned nq = P.np / 2u;
        const Quad *const Qs = qpool + P.qoff + (size_t)st * nq;
        for (unsigned qi = wv; qi < nq; qi += nw) {
          const Quad Q = Qs[qi];
          const unsigned A0 = tab[2u * (st * P.np + Q.p0)], B0 = tab[2u * (st * P.np + Q.p0) + 1u];
          const unsigned A1 = tab[2u * (st * P.np + Q.p1)], B1 = tab[2u * (st * P.np + Q.p1) + 1u];
          const bool u0 = sf[Q.p0] != 0u, u1 = sf[Q.p1] != 0u;
          const double *const w0 = sw + (size_t)Q.p0 * 1024u;
          const double *const w1 = sw + (size_t)Q.p1 * 1024u;
          // W(k, j) of both pairs into this wave's R / W regions (k * LDW + j)
#pragma unroll
          for (unsigned e = 0u; e < 16u; ++e) {
            const unsigned k = l + 64u * e;
            R[(k & 31u) * LDW + (k >> 5)] = w0[k];
            W[(k & 31u) * LDW + (k >> 5)] = w1[k];
          }
          wave_lds_sync();
          wave_upgram(P.G, P.ldG, P.nrow, A0, B0, A1, B1, u0, u1, Q, R, W, T, l,
                      sg + (size_t)Q.n0 * 768u, sg + (size_t)Q.n1 * 768u);
          if (full) {
            if (u0)
              wave_update(P.V, P.ldV, P.ncol, A0, B0, w0, l, 32u);
            if (u1)
              wave_update(P.V, P.ldV, P.ncol, A1, B1, w1, l, 32u);
          }
          wave_lds_sync();
        }
      }
      if (l == 0u) {
        cnt[2u * wv] = sw_s;
        cnt[2u * wv + 1u] = sw_b;
      }
      __syncthreads();
      unsigned cs = 0u, cb = 0u;
      for (unsigned v = 0u; v < nw; ++v) {
        cs += cnt[2u * v];
        cb += cnt[2u * v + 1u];
      }
      ++sweeps;
      tot_s += cs;
      tot_b += cb;
      __syncthreads();
      if (!cb)
        break;
    }
    // initD: one wave per column, the arithmetic of k_initD
    for (unsigned cix = wv; cix < P.ncol; cix += nw) {
      double *const g = P.G + (size_t)cix * P.ldG;
      double x = 0.0;
      for (unsigned r = l; r < P.nrow; r += 64u) {
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
      for (unsigned r = l; r < P.nrow; r += 64u)
        g[r] = g[r] / s;
      if (!l)
        P.D[cix] = (cix < P.nplus) ? x : -x;
    }
    if (threadIdx.x == 0u) {
      BOut o;
      o.sweeps = sweeps;
      o.pad = 0u;
      o.rot_s = tot_s;
      o.rot_b = tot_b;
      o.t0 = t0;
      o.t1 = wall_clock64();
      out[pid] = o;
    }
  }
}

// strategy families the batch may use for an order n1 (built-in generators)
bool family_for(unsigned strategy, unsigned n1, std::string &fam)
{
  const char *pref = (strategy == HJ_BATCH_STRAT_MMSTEP) ? "mmstep"
                     : (strategy == HJ_BATCH_STRAT_BRENTL) ? "BrentL" : "cycwor";
  if (can_generate(pref, n1)) { fam = pref; return true; }
  if (can_generate("BrentL", n1)) { fam = "BrentL"; return true; }
  if (can_generate("mmstep", n1)) { fam = "mmstep"; return true; }
  return false;
}

struct Plan {
  unsigned n = 0u;
  hj_batch_cfg cfg;
  BProb *d_probs = nullptr;
  unsigned *d_order = nullptr, *d_queue = nullptr;
  uint16_t *d_pool = nullptr;
  Quad *d_quads = nullptr;       // closed quads of the fusable tables
  double *d_scr = nullptr;       // per workgroup: Grams and W of the fused steps
  unsigned *d_sflags = nullptr;  // per workgroup: update flags of the fused steps
  unsigned np_max = 1u;
  uint8_t *d_s0 = nullptr;
  unsigned s0steps = 0u;
  BOut *d_out = nullptr;
  std::vector<BOut> h_out;
  std::vector<BProb> h_probs;
  hipGraphExec_t gexec = nullptr;
  hipStream_t cap = nullptr;
  unsigned grid = 0u;
  size_t lds = 0u;
  double clock_mhz = 100.0;
};

double cost_of(const hj_batch_desc &d)
{
  const double nv = (d.routine & HYPJAC_FULL_SVD) ? d.ncol : 0.0;
  return (double)d.ncol * d.ncol * (2.0 * d.nrow + nv);
}

}  // namespace
}  // namespace hj

using namespace hj;

extern "C" void hj_batch_default_cfg(hj_batch_cfg *cfg)
{
  if (!cfg) return;
  cfg->waves = 4u;
  cfg->wg_per_cu = 2u;
  cfg->strategy = HJ_BATCH_STRAT_CYCWOR;
  cfg->max_sweeps = HYPJAC_NSWEEP;
  cfg->longest_first = 1u;
}

static int check_desc(const hj_batch_desc &d)
{
  if (d.routine >= 16u) return -1;
  if (!d.nrow || (d.nrow % 64u)) return -2;
  if (!d.ncol || (d.ncol > d.nrow) || (d.ncol % 32u) || d.ncol > 16u * STRAT1_MAX_STEPS) return -3;
  if (d.nplus > d.ncol) return -4;
  if (!d.G) return -5;
  if (d.ldG < d.nrow || (d.ldG & 1u) || (reinterpret_cast<uintptr_t>(d.G) & 15u)) return -6;
  if (d.routine & HYPJAC_FULL_SVD) {
    if (!d.V) return -7;
    if (d.ldV < d.ncol || (d.ldV & 1u)) return -8;
  }
  if (!d.D) return -9;
  return 0;
}

extern "C" void hj_batch_free(void *plan);

extern "C" void *hj_batch_plan(const hj_batch_desc *probs, unsigned n, const hj_batch_cfg *cfg_in,
                               int *err)
{
  int e = 0;
  if (!probs || !n) e = -20;
  hj_batch_cfg cfg;
  hj_batch_default_cfg(&cfg);
  if (cfg_in) cfg = *cfg_in;
  if (!e && (cfg.waves < 1u || cfg.waves > BATCH_MAX_WAVES || !cfg.wg_per_cu || !cfg.max_sweeps))
    e = -21;
  for (unsigned i = 0u; !e && i < n; ++i)
    e = check_desc(probs[i]);
  if (e) {
    if (err) *err = e;
    return nullptr;
  }
  Plan *P = new Plan;
  P->n = n;
  P->cfg = cfg;
  // strategy tables: one per distinct (family, n1)
  std::vector<uint16_t> pool;
  std::vector<Quad> qpool;
  std::vector<std::pair<std::string, unsigned>> keys;
  std::vector<unsigned> key_off, key_steps, key_qoff;
  const bool nofuse = getenv("HJ_BATCH_NOFUSE") != nullptr;
  P->h_probs.resize(n);
  for (unsigned i = 0u; i < n; ++i) {
    const hj_batch_desc &d = probs[i];
    const unsigned n1 = d.ncol / 16u;
    std::string fam;
    if (!family_for(cfg.strategy, n1, fam))
      HJ_DIE("hj_batch_plan: no strategy of this order");
    unsigned k = 0u;
    while (k < keys.size() && !(keys[k].first == fam && keys[k].second == n1)) ++k;
    if (k == keys.size()) {
      Strategy t;
      if (!generate(fam.c_str(), n1, t) || t.pairs != n1 / 2u)
        HJ_DIE("hj_batch_plan: cannot generate a strategy table");
      keys.push_back({fam, n1});
      key_off.push_back((unsigned)(pool.size() / 2u));
      key_steps.push_back(t.steps);
      // fused steps when every transition of the table has closed quads
      std::vector<uint16_t> flat(t.tab.begin(), t.tab.end());
      std::vector<Quad> q;
      std::vector<char> closed;
      bool all = !nofuse && closed_quads(flat.data(), t.steps, t.pairs, q, &closed);
      for (unsigned st = 0u; all && st + 1u < t.steps; ++st)
        all = closed[st] != 0;
      key_qoff.push_back(all ? (unsigned)qpool.size() : NO_QUADS);
      if (all) qpool.insert(qpool.end(), q.begin(), q.end());
      pool.insert(pool.end(), t.tab.begin(), t.tab.end());
    }
    BProb b;
    b.G = d.G;
    b.V = (d.routine & HYPJAC_FULL_SVD) ? d.V : nullptr;
    b.D = d.D;
    b.nrow = d.nrow;
    b.ncol = d.ncol;
    b.nplus = d.nplus;
    b.ldG = d.ldG;
    b.ldV = d.ldV;
    b.routine = d.routine;
    b.s1off = key_off[k];
    b.steps = key_steps[k];
    b.np = n1 / 2u;
    b.qoff = key_qoff[k];
    P->np_max = std::max(P->np_max, b.np);
    P->h_probs[i] = b;
  }
  std::vector<unsigned> order(n);
  std::iota(order.begin(), order.end(), 0u);
  if (cfg.longest_first)
    std::stable_sort(order.begin(), order.end(),
                     [&](unsigned a, unsigned b) { return cost_of(probs[a]) > cost_of(probs[b]); });
  Strategy s0;
  if (!generate("cycwor", 32u, s0) || s0.pairs != 16u || s0.steps > 32u)
    HJ_DIE("hj_batch_plan: cannot generate cycwor 32");
  std::vector<uint8_t> s0b(1024u, 0u);
  for (size_t i = 0u; i < s0.tab.size() && i < 1024u; ++i) s0b[i] = (uint8_t)s0.tab[i];
  P->s0steps = s0.steps;

  HIP_CALL(hipMalloc(&P->d_probs, sizeof(BProb) * n));
  HIP_CALL(hipMalloc(&P->d_order, sizeof(unsigned) * n));
  HIP_CALL(hipMalloc(&P->d_queue, sizeof(unsigned)));
  HIP_CALL(hipMalloc(&P->d_pool, sizeof(uint16_t) * std::max<size_t>(pool.size(), 2u)));
  HIP_CALL(hipMalloc(&P->d_s0, 1024u));
  HIP_CALL(hipMalloc(&P->d_out, sizeof(BOut) * n));
  HIP_CALL(hipMemcpy(P->d_probs, P->h_probs.data(), sizeof(BProb) * n, hipMemcpyHostToDevice));
  HIP_CALL(hipMemcpy(P->d_order, order.data(), sizeof(unsigned) * n, hipMemcpyHostToDevice));
  HIP_CALL(hipMemcpy(P->d_pool, pool.data(), sizeof(uint16_t) * pool.size(), hipMemcpyHostToDevice));
  HIP_CALL(hipMemcpy(P->d_s0, s0b.data(), 1024u, hipMemcpyHostToDevice));
  HIP_CALL(hipMalloc(&P->d_quads, sizeof(Quad) * std::max<size_t>(qpool.size(), 1u)));
  if (!qpool.empty())
    HIP_CALL(hipMemcpy(P->d_quads, qpool.data(), sizeof(Quad) * qpool.size(), hipMemcpyHostToDevice));
  P->h_out.resize(n);

  int dev = 0;
  HIP_CALL(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIP_CALL(hipGetDeviceProperties(&prop, dev));
  P->grid = cfg.wg_per_cu * (unsigned)prop.multiProcessorCount;
  P->lds = sizeof(double) * WAVE_LDS * cfg.waves;
  const size_t lds_max = std::max(prop.sharedMemPerBlock, prop.sharedMemPerBlockOptin);
  if (P->lds + 1100u > lds_max) {
    // LDS of a workgroup beyond the device limit (workgroups beyond the
    // per-CU capacity are not an error: they wait for a free slot)
    hj_batch_free(P);
    if (err) *err = -22;
    return nullptr;
  }
  HIP_CALL(hipFuncSetAttribute(reinterpret_cast<const void *>(k_batch),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)P->lds));
  HIP_CALL(hipMalloc(&P->d_scr, sizeof(double) * (768u + 1024u) * P->np_max * P->grid));
  HIP_CALL(hipMalloc(&P->d_sflags, sizeof(unsigned) * P->np_max * P->grid));
  int rate_khz = 0;
  if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && rate_khz > 0)
    P->clock_mhz = rate_khz / 1000.0;

  // the batch as one graph: queue reset (a kernel, no memset node) and the
  // persistent launch
  HIP_CALL(hipStreamCreateWithFlags(&P->cap, hipStreamNonBlocking));
  hipGraph_t graph;
  HIP_CALL(hipStreamBeginCapture(P->cap, hipStreamCaptureModeThreadLocal));
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, P->cap, P->d_queue);
  hipLaunchKernelGGL(k_batch, dim3(P->grid), dim3(64u * cfg.waves), P->lds, P->cap, P->d_probs,
                     P->d_order, n, P->d_queue, P->d_pool, P->d_s0, P->s0steps, cfg.max_sweeps,
                     P->d_out, P->d_quads, P->d_scr, P->d_sflags, P->np_max);
  HIP_CALL(hipStreamEndCapture(P->cap, &graph));
  HIP_CALL(hipGraphInstantiate(&P->gexec, graph, nullptr, nullptr, 0));
  HIP_CALL(hipGraphDestroy(graph));
  if (err) *err = 0;
  return P;
}

extern "C" int hj_batch_run(void *plan, void *stream)
{
  Plan *P = static_cast<Plan *>(plan);
  if (!P) return -1;
  hipStream_t caller = static_cast<hipStream_t>(stream);
  hipEvent_t ev;
  HIP_CALL(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  HIP_CALL(hipEventRecord(ev, caller));
  HIP_CALL(hipStreamWaitEvent(P->cap, ev, 0));
  HIP_CALL(hipGraphLaunch(P->gexec, P->cap));
  HIP_CALL(hipEventRecord(ev, P->cap));
  HIP_CALL(hipStreamWaitEvent(caller, ev, 0));
  HIP_CALL(hipStreamSynchronize(P->cap));
  HIP_CALL(hipEventDestroy(ev));
  HIP_CALL(hipMemcpy(P->h_out.data(), P->d_out, sizeof(BOut) * P->n, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int hj_batch_results(void *plan, unsigned *sweeps, unsigned long long *rot_s,
                                unsigned long long *rot_b, double *start_us, double *end_us)
{
  Plan *P = static_cast<Plan *>(plan);
  if (!P) return -1;
  unsigned long long tmin = ~0ull;
  for (const BOut &o : P->h_out) tmin = std::min(tmin, o.t0);
  for (unsigned i = 0u; i < P->n; ++i) {
    const BOut &o = P->h_out[i];
    if (sweeps) sweeps[i] = o.sweeps;
    if (rot_s) rot_s[i] = o.rot_s;
    if (rot_b) rot_b[i] = o.rot_b;
    if (start_us) start_us[i] = (double)(o.t0 - tmin) / P->clock_mhz;
    if (end_us) end_us[i] = (double)(o.t1 - tmin) / P->clock_mhz;
  }
  return 0;
}

extern "C" void hj_batch_free(void *plan)
{
  Plan *P = static_cast<Plan *>(plan);
  if (!P) return;
  if (P->gexec) HIP_CALL(hipGraphExecDestroy(P->gexec));
  if (P->cap) HIP_CALL(hipStreamDestroy(P->cap));
  for (void *p : {(void *)P->d_probs, (void *)P->d_order, (void *)P->d_queue, (void *)P->d_pool,
                  (void *)P->d_s0, (void *)P->d_out, (void *)P->d_quads, (void *)P->d_scr,
                  (void *)P->d_sflags})
    if (p) HIP_CALL(hipFree(p));
  delete P;
}

// Autotuner: times every candidate configuration on a size-stratified sample
// of the batch (copies of its matrices, so the caller's data are untouched)
// and returns the fastest.
extern "C" int hj_batch_autotune(const hj_batch_desc *probs, unsigned n, unsigned sample,
                                 hj_batch_cfg *best, double *best_rate)
{
  if (!probs || !n || !best) return -1;
  for (unsigned i = 0u; i < n; ++i)
    if (check_desc(probs[i])) return check_desc(probs[i]);
  if (!sample || sample > n) sample = n;
  std::vector<unsigned> idx(n);
  std::iota(idx.begin(), idx.end(), 0u);
  std::stable_sort(idx.begin(), idx.end(),
                   [&](unsigned a, unsigned b) { return cost_of(probs[a]) > cost_of(probs[b]); });
  std::vector<hj_batch_desc> sub;
  const double stride = (double)n / sample;
  for (unsigned k = 0u; k < sample; ++k) sub.push_back(probs[idx[(unsigned)(k * stride)]]);
  // private copies: G (and V, D) of the sample, plus pristine G
  std::vector<double *> pristine(sub.size());
  std::vector<hj_batch_desc> work = sub;
  for (size_t k = 0u; k < sub.size(); ++k) {
    const size_t gb = sizeof(double) * (size_t)sub[k].ldG * sub[k].ncol;
    HIP_CALL(hipMalloc(&pristine[k], gb));
    HIP_CALL(hipMemcpy(pristine[k], sub[k].G, gb, hipMemcpyDeviceToDevice));
    HIP_CALL(hipMalloc(&work[k].G, gb));
    HIP_CALL(hipMalloc(&work[k].D, sizeof(double) * sub[k].ncol));
    if (sub[k].routine & HYPJAC_FULL_SVD)
      HIP_CALL(hipMalloc(&work[k].V, sizeof(double) * (size_t)sub[k].ldV * sub[k].ncol));
  }
  hj_batch_cfg base;
  hj_batch_default_cfg(&base);
  base.max_sweeps = best->max_sweeps ? best->max_sweeps : base.max_sweeps;
  const unsigned waves[] = {2u, 4u, 8u};
  const unsigned wgs[] = {1u, 2u, 4u};
  const unsigned strats[] = {HJ_BATCH_STRAT_CYCWOR, HJ_BATCH_STRAT_BRENTL, HJ_BATCH_STRAT_MMSTEP};
  double best_t = 1e300;
  hj_batch_cfg bc = base;
  hipEvent_t e0, e1;
  HIP_CALL(hipEventCreate(&e0));
  HIP_CALL(hipEventCreate(&e1));
  hipStream_t s;
  HIP_CALL(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  for (unsigned st : strats)
    for (unsigned wv : waves)
      for (unsigned wg : wgs) {
        if (wv * wg > 8u) continue;  // LDS: at most 8 wave regions per CU
        hj_batch_cfg c = base;
        c.waves = wv;
        c.wg_per_cu = wg;
        c.strategy = st;
        int err = 0;
        void *pl = hj_batch_plan(work.data(), (unsigned)work.size(), &c, &err);
        if (!pl) continue;
        for (size_t k = 0u; k < sub.size(); ++k)
          HIP_CALL(hipMemcpyAsync(work[k].G, pristine[k],
                                  sizeof(double) * (size_t)sub[k].ldG * sub[k].ncol,
                                  hipMemcpyDeviceToDevice, s));
        HIP_CALL(hipEventRecord(e0, s));
        hj_batch_run(pl, s);
        HIP_CALL(hipEventRecord(e1, s));
        HIP_CALL(hipEventSynchronize(e1));
        float ms = 0.0f;
        HIP_CALL(hipEventElapsedTime(&ms, e0, e1));
        hj_batch_free(pl);
        if (ms < best_t) {
          best_t = ms;
          bc = c;
        }
      }
  HIP_CALL(hipStreamDestroy(s));
  HIP_CALL(hipEventDestroy(e0));
  HIP_CALL(hipEventDestroy(e1));
  for (size_t k = 0u; k < sub.size(); ++k) {
    HIP_CALL(hipFree(pristine[k]));
    HIP_CALL(hipFree(work[k].G));
    HIP_CALL(hipFree(work[k].D));
    if (work[k].V && (sub[k].routine & HYPJAC_FULL_SVD)) HIP_CALL(hipFree(work[k].V));
  }
  *best = bc;
  if (best_rate) *best_rate = sub.size() / (best_t * 1e-3);
  return 0;
}
