// hj_kernels.hip -- level-1 kernels of the blocked one-sided (hyperbolic)
// Jacobi SVD for gfx950 (MI355X).
//
// One level-1 step (all block pairs of one STRAT1 step) is three launches:
//   k_gram   : partial 32x32 Gram of each block pair over a row chunk, on
//              FP64 MFMA 16x16x4 (one wave per (pair, chunk));
//   k_pair   : per pair, sum of the partials, 32x32 Cholesky, the inner
//              Jacobi sweep(s) on R in LDS accumulating W (one 128-thread
//              workgroup per pair, 8 lanes per pivot pair);
//   k_update : [G_i G_j] <- [G_i G_j] W and [V_i V_j] <- [V_i V_j] W on FP64
//              MFMA (one wave per (pair, row chunk)).
// The reference fuses the three into one CUDA block per pair
// (share/device_code_cdsort_accumV.hpp:4-94 and siblings); splitting them
// lets every step fill the 256 CUs whatever the number of pairs.
#include "hj_device.hpp"
#include "hj_fused.hpp"
#include "hj_kernels.hpp"

typedef double d4 __attribute__((ext_vector_type(4)));

// Output rows of the 16x16 MFMA results (lane l, register r; the column is
// l % 16), measured on gfx950: FP64 16x16x4 holds row 4r + l/16; FP32
// 16x16x32 BF16 holds row 4(l/16) + r (and its operands take k = 8(l/16) + t
// from slot t of lane l).
#define F32_OUT_ROW(l, r) (4u * ((l) >> 4) + (r))

namespace hj {


// ---------------------------------------------------------------------------
// Gram partials.  Workgroup (pair pi, chunk c) covers rows [c*CH, (c+1)*CH)
// with nw waves; wave w takes the sub-chunk of CH/nw rows starting at
// c*CH + w*CH/nw, and the nw wave results are added in wave order through
// LDS.  Within a wave, lane l owns column l%16 of both blocks and k-slot
// k = l/16; MFMA step j of a 16-row slab takes row 4j + k -- the rows a lane
// holds after the block update's MFMA, so the fused update feeds the same
// sequence from registers.  Tiles: 00 = block i with itself, 10 = block j
// with block i, 11 = block j with itself.
__global__ __launch_bounds__(256, 2) void k_gram(const double *__restrict__ G, unsigned ldG,
                                              const uint16_t *__restrict__ pairs, unsigned nchunk,
                                              unsigned CH, double *__restrict__ part,
                                              const uint8_t *__restrict__ skip)
{
  __shared__ d4 red[3][GRAM_WAVES - 1][64];
  const unsigned l = threadIdx.x & 63u, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const unsigned pi = blockIdx.x / nchunk, c = blockIdx.x - pi * nchunk;
  if (skip && skip[pi])
    return;
  const unsigned iBlk = pairs[2u * pi], jBlk = pairs[2u * pi + 1u];
  const unsigned SUB = CH / nw;
  const size_t r0 = (size_t)c * CH + (size_t)w * SUB + (l >> 4);
  const double *const g0 = G + (size_t)(16u * iBlk + (l & 15u)) * ldG + r0;
  const double *const g1 = G + (size_t)(16u * jBlk + (l & 15u)) * ldG + r0;
  d4 a00 = {0.0, 0.0, 0.0, 0.0}, a10 = a00, a11 = a00;
  // two 16-row slabs per iteration (value t: row 4t + k of the 32 rows),
  // the next two in flight
  double n0[8], n1[8];
#pragma unroll
  for (unsigned t = 0u; t < 8u; ++t) {
    n0[t] = g0[4u * t];
    n1[t] = g1[4u * t];
  }
  for (unsigned g = 0u; g < SUB; g += 32u) {
    double x0[8], x1[8];
#pragma unroll
    for (unsigned t = 0u; t < 8u; ++t) {
      x0[t] = n0[t];
      x1[t] = n1[t];
    }
    if (g + 32u < SUB) {
#pragma unroll
      for (unsigned t = 0u; t < 8u; ++t) {
        n0[t] = g0[g + 32u + 4u * t];
        n1[t] = g1[g + 32u + 4u * t];
      }
    }
#pragma unroll
    for (unsigned t = 0u; t < 8u; ++t) {
      a00 = mfma(x0[t], x0[t], a00);
      a10 = mfma(x1[t], x0[t], a10);
      a11 = mfma(x1[t], x1[t], a11);
    }
  }
  if (w) {
    red[0][w - 1u][l] = a00;
    red[1][w - 1u][l] = a10;
    red[2][w - 1u][l] = a11;
  }
  __syncthreads();
  if (w == 0u) {
    for (unsigned v = 1u; v < nw; ++v) {
      a00 = a00 + red[0][v - 1u][l];
      a10 = a10 + red[1][v - 1u][l];
      a11 = a11 + red[2][v - 1u][l];
    }
    double *const o = part + (size_t)blockIdx.x * 768u + l * 4u;
#pragma unroll
    for (unsigned r = 0u; r < 4u; ++r) {
      o[r] = a00[r];
      o[256u + r] = a10[r];
      o[512u + r] = a11[r];
    }
  }
}

// ---------------------------------------------------------------------------
// Split-precision Gram partials (same grid, chunks, waves and output layout
// as k_gram) for the pairs whose plan byte is PLAN_SPLIT.
//
// Per 16-row slab a wave scales each column by a power of two (2^-e with
// 2^(e-1) <= max |x| < 2^e over the column's 16 rows, so the scaled values
// lie in (-1, 1) whatever the FP64 range and however the column scales of a
// block differ), rounds to FP32 and splits every value into three bfloat16
// pieces x = h + m + l (h = bf16(x), m = bf16(x - h), l = bf16(x - h - m):
// the FP32 value exactly, up to the last piece's rounding).  The products of
// a tile run on BF16 MFMA 16x16x16 with FP32 accumulation in three magnitude
// classes, hh, hm + mh, hl + lh + mm (the dropped terms ml, lm, ll are below
// 2^-24 relative), and after each slab the classes are added smallest first
// in FP64, unscaled exactly by 2^(e_row + e_col) and added to the FP64
// accumulators.  Relative error of an entry, against sqrt(A_ii A_jj): about
// 2^-23 per slab, FP64 across slabs.  Rows: lane l holds rows 4t + l/16 of
// column l%16 in slot t, as the FP64 Gram and the block update hold them.
typedef __attribute__((ext_vector_type(4))) short s4;
typedef __attribute__((ext_vector_type(4))) float f4;

// row of a 16x16 FP32 MFMA result held by lane l in register r
__device__ __forceinline__ unsigned f32_out_row(unsigned l, unsigned r)
{
  return F32_OUT_ROW(l, r);
}

// BF16 MFMA 16x16x16: lane l gives 4 values of column l%16 of each operand
// (its rows 4t + l/16 of the slab, slot t); any one-to-one assignment of
// rows to k serves a Gram, the same for both operands
__device__ __forceinline__ f4 mfma_bf(const s4 &a, const s4 &b, f4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// biased FP64 exponent field of x (0 for zero / subnormal)
__device__ __forceinline__ int exp_field(double x)
{
  return (__double2hiint(x) >> 20) & 0x7ff;
}

// max over the four lanes l, l^16, l^32, l^48 (the lanes of one column)
__device__ __forceinline__ int column_max_int(int v)
{
  v = max(v, __shfl_xor(v, 16, 64));
  v = max(v, __shfl_xor(v, 32, 64));
  return v;
}

__device__ __forceinline__ short bf16_of(float v)
{
  return __builtin_bit_cast(short, (__bf16)v);
}

__device__ __forceinline__ float float_of_bf16(short b)
{
  return __uint_as_float(((unsigned)(unsigned short)b) << 16);
}

// three-way bf16 split of 4 FP64 values scaled by 2^-e
__device__ __forceinline__ void split3(const d4 &x, int e, s4 &h, s4 &m, s4 &lo)
{
#pragma unroll
  for (unsigned t = 0u; t < 4u; ++t) {
    const float v = (float)ldexp(x[t], -e);
    const short a = bf16_of(v);
    const float r1 = v - float_of_bf16(a);
    const short b = bf16_of(r1);
    const float r2 = r1 - float_of_bf16(b);
    h[t] = a;
    m[t] = b;
    lo[t] = bf16_of(r2);
  }
}

struct Cls {
  f4 c0, c1, c2;  // hh, hm + mh, hl + lh + mm
};

__device__ __forceinline__ void tile_products(Cls &C, const s4 &ha, const s4 &ma, const s4 &la,
                                              const s4 &hb, const s4 &mb, const s4 &lb)
{
  C.c0 = mfma_bf(ha, hb, C.c0);
  C.c1 = mfma_bf(ha, mb, C.c1);
  C.c1 = mfma_bf(ma, hb, C.c1);
  C.c2 = mfma_bf(ha, lb, C.c2);
  C.c2 = mfma_bf(la, hb, C.c2);
  C.c2 = mfma_bf(ma, mb, C.c2);
}

// adds the classes of a tile to its FP64 accumulator, element r (row
// 4(l/16) + r, column l%16) unscaled by 2^(erow[r] + ecol)
__device__ __forceinline__ void fold(d4 &acc, const Cls &C, const int *erow, int ecol)
{
#pragma unroll
  for (unsigned r = 0u; r < 4u; ++r) {
    double t = (double)C.c2[r] + (double)C.c1[r];
    t = t + (double)C.c0[r];
    acc[r] = acc[r] + ldexp(t, erow[r] + ecol);
  }
}

// split Gram of one 16-row slab: X, Y hold the lane's rows 4t + l/16 of
// column l%16 of the pair's first and second block
__device__ __forceinline__ void gram_split_slab(const d4 &X, const d4 &Y, unsigned l, d4 &a00,
                                                d4 &a10, d4 &a11)
{
  int e0 = 0, e1 = 0;
#pragma unroll
  for (unsigned t = 0u; t < 4u; ++t) {
    e0 = max(e0, exp_field(X[t]));
    e1 = max(e1, exp_field(Y[t]));
  }
  e0 = column_max_int(e0);
  e1 = column_max_int(e1);
  e0 = e0 ? (e0 - 1022) : 0;
  e1 = e1 ? (e1 - 1022) : 0;
  // exponents of the columns behind the rows this lane's results hold
  int r0e[4], r1e[4];
#pragma unroll
  for (unsigned r = 0u; r < 4u; ++r) {
    r0e[r] = __shfl(e0, (int)f32_out_row(l, r), 64);
    r1e[r] = __shfl(e1, (int)f32_out_row(l, r), 64);
  }
  s4 h0, m0, q0, h1, m1, q1;
  split3(X, e0, h0, m0, q0);
  split3(Y, e1, h1, m1, q1);
  const f4 z = {0.0f, 0.0f, 0.0f, 0.0f};
  Cls C00 = {z, z, z}, C10 = {z, z, z}, C11 = {z, z, z};
  tile_products(C00, h0, m0, q0, h0, m0, q0);
  tile_products(C10, h1, m1, q1, h0, m0, q0);
  tile_products(C11, h1, m1, q1, h1, m1, q1);
  fold(a00, C00, r0e, e0);
  fold(a10, C10, r1e, e0);
  fold(a11, C11, r1e, e1);
}

__global__ __launch_bounds__(256, 2) void k_gram_split(const double *__restrict__ G, unsigned ldG,
                                                       const uint16_t *__restrict__ pairs,
                                                       unsigned nchunk, unsigned CH,
                                                       double *__restrict__ part,
                                                       const uint8_t *__restrict__ plan)
{
  __shared__ d4 red[3][GRAM_WAVES - 1][64];
  const unsigned l = threadIdx.x & 63u, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const unsigned pi = blockIdx.x / nchunk, c = blockIdx.x - pi * nchunk;
  if (plan[pi] != PLAN_SPLIT)
    return;
  const unsigned iBlk = pairs[2u * pi], jBlk = pairs[2u * pi + 1u];
  const unsigned SUB = CH / nw;
  const size_t r0 = (size_t)c * CH + (size_t)w * SUB + (l >> 4);
  const double *const g0 = G + (size_t)(16u * iBlk + (l & 15u)) * ldG + r0;
  const double *const g1 = G + (size_t)(16u * jBlk + (l & 15u)) * ldG + r0;
  d4 a00 = {0.0, 0.0, 0.0, 0.0}, a10 = a00, a11 = a00;
  double n0[8], n1[8];
#pragma unroll
  for (unsigned t = 0u; t < 8u; ++t) {
    n0[t] = g0[4u * t];
    n1[t] = g1[4u * t];
  }
  for (unsigned g = 0u; g < SUB; g += 32u) {
    d4 X0, X1, Y0, Y1;
#pragma unroll
    for (unsigned t = 0u; t < 4u; ++t) {
      X0[t] = n0[t];
      X1[t] = n0[4u + t];
      Y0[t] = n1[t];
      Y1[t] = n1[4u + t];
    }
    if (g + 32u < SUB) {
#pragma unroll
      for (unsigned t = 0u; t < 8u; ++t) {
        n0[t] = g0[g + 32u + 4u * t];
        n1[t] = g1[g + 32u + 4u * t];
      }
    }
    gram_split_slab(X0, Y0, l, a00, a10, a11);
    gram_split_slab(X1, Y1, l, a00, a10, a11);
  }
  if (w) {
    red[0][w - 1u][l] = a00;
    red[1][w - 1u][l] = a10;
    red[2][w - 1u][l] = a11;
  }
  __syncthreads();
  if (w == 0u) {
    for (unsigned v = 1u; v < nw; ++v) {
      a00 = a00 + red[0][v - 1u][l];
      a10 = a10 + red[1][v - 1u][l];
      a11 = a11 + red[2][v - 1u][l];
    }
    // to the FP64 MFMA layout of the partials: element (i, j) at lane
    // j + 16 (i % 4), register i / 4
    double *const o = part + (size_t)blockIdx.x * 768u;
    const unsigned j = l & 15u;
#pragma unroll
    for (unsigned r = 0u; r < 4u; ++r) {
      const unsigned i = f32_out_row(l, r);
      const unsigned idx = (j + 16u * (i & 3u)) * 4u + (i >> 2);
      o[idx] = a00[r];
      o[256u + idx] = a10[r];
      o[512u + idx] = a11[r];
    }
  }
}

// ---------------------------------------------------------------------------
// Block pair: Gram -> Cholesky -> inner Jacobi -> W, two waves per pair.
// R and W live in LDS, column-major with a padded leading dimension.
// Cholesky: wave 0, lane x (x < 32) keeps row x of the Gram in registers;
// the pivot column is broadcast with readlane, no barriers.
// Inner Jacobi: pivot-pair slot y = t/8 (0..15); lane r = t%8 of the slot
// keeps rows r, r+8, r+16, r+24 of the two columns of R and of W.

#define LDR 33u
#define RA(A, r, c) (A)[(c) * LDR + (r)]
#define F32(A, r, c) (A)[(c) * 32u + (r)]

// LDS of one pair visit (128 threads)
struct PairLds {
  double R[32u * LDR];
  double W[32u * LDR];
  double cholcol[32];
  unsigned cnt_s[16], cnt_b[16];
  float cmax[16];
  float bnorm[2];
  int chol_ok;
  int go;     // the visit still runs inner sweeps
  int swp;    // a pivot pair was swapped by the sort rule
};

// One pair visit by the 128 threads tid = 0..127 of a workgroup that may
// run two visits side by side (256 threads, one PairLds each): both halves
// execute the same barriers, a half without work (no pair, a skipped visit,
// a failed split Cholesky, converged inner sweeps) passes them idle.
// go_other: the other half's go flag, or null when the workgroup holds one
// visit.  plan: the visit's plan (PLAN_FP64 without a plan).
template <bool SORT>
__device__ __forceinline__ void pair_visit(unsigned tid, bool present, unsigned pi, PairLds &L,
                                           const int *go_other, const uint16_t *__restrict__ pairs,
                                           const double *__restrict__ part, unsigned nchunk,
                                           double *__restrict__ Wout, unsigned *__restrict__ flags,
                                           RunCtl *__restrict__ ctl, VisitRec *__restrict__ rec,
                                           uint8_t plan, const uint8_t *S0, unsigned s0steps,
                                           unsigned nswp, int definite, unsigned nplus)
{
  double *const R = L.R;
  double *const W = L.W;
  if (tid == 0u)
    L.swp = 0;
  if (present && plan == PLAN_SKIP && tid == 0u) {
    // skipped visit: no update, the record keeps the last executed visit
    flags[pi] = 0u;
    rec[pi].skips += 1u;
    atomicAdd(&ctl->skipped, 1u);
  }
  bool live = present && plan != PLAN_SKIP;
  unsigned iBlk = 0u, jBlk = 0u;
  if (live) {
    iBlk = pairs[2u * pi];
    jBlk = pairs[2u * pi + 1u];
  }

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

  if (live) {
    // sum of the chunk partials in chunk order: the six entries of a thread
    // with all their chunk loads in flight together (one memory round trip)
    const double *const pp = part + (size_t)pi * nchunk * 768u;
    {
      const unsigned e0 = tid;
      double v[6][MAX_GRAM_CHUNKS];
#pragma unroll
      for (unsigned q = 0u; q < 6u; ++q)
#pragma unroll
        for (unsigned c = 0u; c < MAX_GRAM_CHUNKS; ++c)
          if (c < nchunk) v[q][c] = pp[(size_t)c * 768u + e0 + 128u * q];
#pragma unroll
      for (unsigned q = 0u; q < 6u; ++q) {
        const unsigned e = e0 + 128u * q;
        const unsigned tile = e >> 8;
        double s = v[q][0];
#pragma unroll
        for (unsigned c = 1u; c < MAX_GRAM_CHUNKS; ++c)
          if (c < nchunk) s = s + v[q][c];
        const unsigned l = (e >> 2) & 63u, r = e & 3u;
        const unsigned i = 4u * r + (l >> 4), j = l & 15u;
        const unsigned row = (tile == 0u) ? i : (16u + i);
        const unsigned col = (tile == 2u) ? (16u + j) : j;
        RA(R, row, col) = s;
      }
    }
    for (unsigned e = tid; e < 1024u; e += 128u)
      RA(W, e & 31u, e >> 5) = ((e & 31u) == (e >> 5)) ? 1.0 : 0.0;
  }
  __syncthreads();
  if (live && tid < 2u) {
    // squared Frobenius norms of the two blocks, for the visit record
    double sn = 0.0;
    for (unsigned k = 0u; k < 16u; ++k)
      sn = sn + RA(R, 16u * tid + k, 16u * tid + k);
    L.bnorm[tid] = (float)log2(sn);
  }

  // Cholesky, share/device_code_common_Cholesky.hpp:58-158 (USE_RSQRT):
  // column k scaled by rsqrt(A(k,k)), diagonal sqrt(A(k,k)), then
  // A(x,j) = fma(-A(x,k), A(j,k), A(x,j)) for x >= j > k; R = L^T.
  if (live && tid < 64u) {
    // a Gram from the split products may miss positive definiteness (nearly
    // dependent columns): then no update, FP64 at the next visit
    const bool ok = chol32_wave<LDR>(R, L.cholcol, tid, plan == PLAN_SPLIT);
    if (plan == PLAN_SPLIT && tid == 0u) L.chol_ok = ok ? 1 : 0;
  }
  __syncthreads();
  if (live && plan == PLAN_SPLIT && !L.chol_ok) {
    if (tid == 0u) {
      flags[pi] = 0u;
      VisitRec v;
      v.rot_s = 0u;
      v.rot_b = 0u;
      v.maxcos = 1.0f;
      v.l2ni = L.bnorm[0];
      v.l2nj = L.bnorm[1];
      v.sweep = ctl->sweep;
      v.skips = 0u;
      v.flags = REC_SPLIT | REC_SPLIT_FAIL;
      rec[pi] = v;
      atomicAdd(&ctl->split, 1u);
    }
    live = false;
  }

  // Inner Jacobi: share/device_code_nosort.hpp:4-222 (SORT false),
  // share/device_code_cdsort_KISS.hpp:4-260 and share/device_code_cdsort.hpp:25-167
  // (SORT true).  One barrier per step: the pivot pairs of a step touch
  // disjoint columns; rotation counts are kept per slot and summed per sweep.
  const unsigned y = tid >> 3, r = tid & 7u;
  unsigned blk_s = 0u, blk_b = 0u;
  bool swapped = false;
  double my_cos = 0.0;
  // |Apq| / (|p| |q|) of the previous step, divided during the next step's
  // loads (off the step's chain of dependent operations)
  double cos_n = 0.0, cos_d = 0.0;
  bool sweeping = live;
  for (unsigned swp = 0u; swp < nswp; ++swp) {
    // the halves go on while either still sweeps
    if (tid == 0u) L.go = sweeping ? 1 : 0;
    __syncthreads();
    if (!L.go && !(go_other && *go_other))
      break;
    unsigned my_s = 0u, my_b = 0u;
    for (unsigned step = 0u; step < s0steps; ++step) {
      if (sweeping) {
      const unsigned p = S0[(step * 16u + y) * 2u], q = S0[(step * 16u + y) * 2u + 1u];
      double Gp[4], Gq[4], Vp[4], Vq[4];
#pragma unroll
      for (unsigned i = 0u; i < 4u; ++i) {
        Gp[i] = RA(R, r + 8u * i, p);
        Gq[i] = RA(R, r + 8u * i, q);
        Vp[i] = RA(W, r + 8u * i, p);
        Vq[i] = RA(W, r + 8u * i, q);
      }
      if (swp == 0u) {
        const double cr = pin(cos_n / cos_d);
        my_cos = fmax(my_cos, (cos_d > 0.0) ? cr : 0.0);
      }
      double Dp = sum8x4(__fma_rn(Gp[0], Gp[0], 0.0), __fma_rn(Gp[1], Gp[1], 0.0),
                         __fma_rn(Gp[2], Gp[2], 0.0), __fma_rn(Gp[3], Gp[3], 0.0));
      double Dq = sum8x4(__fma_rn(Gq[0], Gq[0], 0.0), __fma_rn(Gq[1], Gq[1], 0.0),
                         __fma_rn(Gq[2], Gq[2], 0.0), __fma_rn(Gq[3], Gq[3], 0.0));
      const double Apq = sum8x4(__fma_rn(Gp[0], Gq[0], 0.0), __fma_rn(Gp[1], Gq[1], 0.0),
                                __fma_rn(Gp[2], Gq[2], 0.0), __fma_rn(Gp[3], Gq[3], 0.0));
      const double Dp_ = sqrt(Dp), Dq_ = sqrt(Dq), Apq_ = fabs(Apq), Dpq_ = Dp_ * Dq_;
      const int transf_s = !(Apq_ < (Dpq_ * MYTOL));
      cos_n = Apq_;
      cos_d = Dpq_;
      int pdef = def;
      if (hyp) {
        if ((p < npos) && (q < npos)) pdef = +1;
        else if ((p >= npos) && (q >= npos)) pdef = -1;
        else pdef = 0;
      }
      bool trig;
      if (!SORT) trig = !hyp || pdef;
      else trig = (pdef != 0);
      // the rotation is computed beside the norms and the test, on every
      // lane; only a transforming pair uses it
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
          // big rotation: scale by c; otherwise c == 1 exactly (dev_rot), and
          // c * x is x
          const double rp = __fma_rn(t_, Gq[i], Gp[i]), rq = __fma_rn(t, Gp[i], Gq[i]);
          const double wp = __fma_rn(t_, Vq[i], Vp[i]), wq = __fma_rn(t, Vp[i], Vq[i]);
          RA(R, r + 8u * i, P) = c * rp;
          RA(R, r + 8u * i, Q) = c * rq;
          RA(W, r + 8u * i, P) = c * wp;
          RA(W, r + 8u * i, Q) = c * wq;
        }
      }
      else if (SORT && (((pdef > 0) && (Dp < Dq)) || ((pdef < 0) && (Dq < Dp)))) {
#pragma unroll
        for (unsigned i = 0u; i < 4u; ++i) {
          RA(R, r + 8u * i, p) = Gq[i];
          RA(R, r + 8u * i, q) = Gp[i];
          RA(W, r + 8u * i, p) = Vq[i];
          RA(W, r + 8u * i, q) = Vp[i];
        }
        swapped = true;
      }
      my_s += (unsigned)transf_s;
      my_b += (unsigned)transf_b;
      }
      __syncthreads();
    }
    if (sweeping && swp == 0u) {
      const double cr = pin(cos_n / cos_d);
      my_cos = fmax(my_cos, (cos_d > 0.0) ? cr : 0.0);
      cos_d = 0.0;
    }
    if (sweeping && r == 0u) {
      L.cnt_s[y] = my_s;
      L.cnt_b[y] = my_b;
    }
    __syncthreads();
    if (sweeping) {
      unsigned swp_s = 0u, swp_b = 0u;
      for (unsigned i = 0u; i < 16u; ++i) {
        swp_s += L.cnt_s[i];
        swp_b += L.cnt_b[i];
      }
      if (swp_s) {
        blk_s += swp_s;
        blk_b += swp_b;
      }
      else
        sweeping = false;
    }
    __syncthreads();
  }

  if (live && r == 0u)
    L.cmax[y] = (float)my_cos;
  if (live && swapped)
    L.swp = 1;
  __syncthreads();
  if (!live)
    return;
  // convergence counter, share/device_code_nosort.hpp:94-103
  if (tid == 0u) {
    unsigned long long *const cvg = &ctl->cvg;
    if (blk_s) {
      if (blk_b)
        atomicAdd(cvg, (((unsigned long long)blk_b) << 32) | blk_s);
      else
        atomicAdd(reinterpret_cast<unsigned *>(cvg), blk_s);
    }
    // the block update is left out exactly when W is the identity (no
    // rotation, no swap)
    flags[pi] = (blk_s || L.swp) ? 1u : 0u;
    float cm = 0.0f;
    for (unsigned i = 0u; i < 16u; ++i)
      cm = fmaxf(cm, L.cmax[i]);
    VisitRec v;
    v.rot_s = blk_s;
    v.rot_b = blk_b;
    v.maxcos = cm;
    v.l2ni = L.bnorm[0];
    v.l2nj = L.bnorm[1];
    v.sweep = ctl->sweep;
    v.skips = 0u;
    v.flags = (plan == PLAN_SPLIT) ? REC_SPLIT : 0u;
    rec[pi] = v;
    if (plan == PLAN_SPLIT)
      atomicAdd(&ctl->split, 1u);
  }
  double *const wo = Wout + (size_t)pi * 1024u;
  for (unsigned e = tid; e < 1024u; e += 128u)
    wo[e] = RA(W, e & 31u, e >> 5);
}

// The plan a visit runs: a visit planned split whose Gram partials come from
// the FP64 Gram (a fused transition's that forms FP64 Grams only) is an FP64
// visit -- it is counted and checked as one.
__device__ __forceinline__ uint8_t visit_plan(uint8_t plan, bool split_src)
{
  return (plan == PLAN_SPLIT && !split_src) ? PLAN_FP64 : plan;
}

// split_src: the partials of this launch's pairs came from k_gram followed by
// k_gram_split (a standalone Gram launch), so split-planned pairs hold split
// Grams
template <bool SORT>
__global__ __launch_bounds__(128) void k_pair(const uint16_t *__restrict__ pairs,
                                              const double *__restrict__ part, unsigned nchunk,
                                              double *__restrict__ Wout, unsigned *__restrict__ flags,
                                              RunCtl *__restrict__ ctl, VisitRec *__restrict__ rec,
                                              const uint8_t *__restrict__ skip, int split_src,
                                              const uint8_t *__restrict__ strat0, unsigned s0steps,
                                              unsigned nswp, int definite, unsigned nplus)
{
  __shared__ PairLds L;
  __shared__ __attribute__((aligned(16))) uint8_t S0[STRAT0_TAB_BYTES];
  const unsigned tid = threadIdx.x, pi = blockIdx.x;
  const uint8_t plan = visit_plan(skip ? skip[pi] : PLAN_FP64, split_src != 0);
  if (plan == PLAN_SKIP) {
    // skipped visit: no update, the record keeps the last executed visit
    if (tid == 0u) {
      flags[pi] = 0u;
      rec[pi].skips += 1u;
      atomicAdd(&ctl->skipped, 1u);
    }
    return;
  }
  // the level-0 table (STRAT0_TAB_BYTES, 16-byte aligned) in one pass of
  // 16-byte loads
  if (tid < STRAT0_TAB_BYTES / 16u)
    reinterpret_cast<uint4 *>(S0)[tid] = reinterpret_cast<const uint4 *>(strat0)[tid];
  pair_visit<SORT>(tid, true, pi, L, nullptr, pairs, part, nchunk, Wout, flags, ctl, rec, plan, S0,
                   s0steps, nswp, definite, nplus);
}

// ---------------------------------------------------------------------------
// Block update on MFMA: out^T = W^T [G_i G_j]^T per 16-row slab.
// A operand (W^T, 16 output columns x 4 k): lane l gives W(4u + l/16, 16h + l%16);
// B operand (G^T, 4 k x 16 rows): lane l gives G(r + l%16, column 4u + l/16);
// D: lane l, reg q holds output column 16h + 4q + l/16 of row r + l%16.
// The chain over u = 0..7, k = 0..3 runs over the 32 columns in order.
// rows [r0, r1) of the two 16-column blocks iBlk, jBlk of M (leading
// dimension ld) times the pair's 32x32 W, by one wave (lane l)
__device__ __forceinline__ void update_rows(double *__restrict__ M, unsigned ld, unsigned r0,
                                            unsigned r1, unsigned iBlk, unsigned jBlk,
                                            const double *__restrict__ w, unsigned l)
{
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
  double nb[8];
#pragma unroll
  for (unsigned u = 0u; u < 8u; ++u)
    nb[u] = cp[u][r0];
  for (unsigned r = r0; r < r1; r += 16u) {
    double b[8];
#pragma unroll
    for (unsigned u = 0u; u < 8u; ++u)
      b[u] = nb[u];
    if (r + 16u < r1) {
#pragma unroll
      for (unsigned u = 0u; u < 8u; ++u)
        nb[u] = cp[u][r + 16u];
    }
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
  update_rows(M, ld, r0, r1, pairs[2u * pi], pairs[2u * pi + 1u], Wbuf + (size_t)pi * 1024u, l);
}

// ---------------------------------------------------------------------------
// Column norms, G -> U, D = +-||g||^2 (share/device_code_common.hpp:187-219).
// One wave per column: lane l sums rows l, l+64, ... by fma, then an xor
// butterfly over 64 lanes (32,16,8,4,2,1).
// ---------------------------------------------------------------------------
// Fused step transition: the block update of step s for the two pairs of a
// closed quad and the Gram partials of the two pairs of step s+1 they form,
// from the updated values in registers (one HBM read and one write of the
// four blocks instead of the update's read + write and the next Gram's
// read).  Grid: nquads x nchunk workgroups of nw waves, the Gram's chunks
// and waves (gram_waves), so the partials are bit for bit those of k_gram /
// k_gram_split on the updated matrix.
//
// Update orientation D = G W (A = G, B = W): lane l, register r of the
// result holds row 4r + l/16 of the 16-row slab, column l%16 of the tile --
// exactly the Gram's operand; the products and their chain over k = 0..31
// are those of k_update (operands swapped: the same fma results).

// stores a wave-reduced Gram partial: FP64 tiles at l*4 + r as k_gram does,
// split tiles (FP32 MFMA layout) the way k_gram_split stores them
__device__ __forceinline__ void store_partial(double *o, unsigned l, bool split, const d4 &a00,
                                              const d4 &a10, const d4 &a11)
{
#pragma unroll
  for (unsigned r = 0u; r < 4u; ++r) {
    unsigned idx = l * 4u + r;
    if (split) {
      const unsigned i = f32_out_row(l, r);
      idx = ((l & 15u) + 16u * (i & 3u)) * 4u + (i >> 2);
    }
    o[idx] = a00[r];
    o[256u + idx] = a10[r];
    o[512u + idx] = a11[r];
  }
}

// relaxed device-scope load (a vector load, coherent across the XCDs)
__device__ __forceinline__ unsigned ld_relaxed(const unsigned *p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// LDS of one fused workgroup: wave partials, W of both pairs, the waves'
// transpose buffers
struct UpgLds {
  d4 red[2][3][GRAM_WAVES - 1][64];
  double Wl[2][32u * LDW];
  double Tb[GRAM_WAVES][16u * LDT];
};

// One (quad, row chunk) of the fused transition by the workgroup's first nw
// waves (the others only pass the barriers).  Every non-skipped next pair
// gets its Gram from the updated slabs in registers: the split Gram (the
// slab arithmetic and partial layout of k_gram_split) when fsplit is set and
// the pair is planned split, else the FP64 Gram.
__device__ __forceinline__ void upgram_item(UpgLds &S, double *__restrict__ G, unsigned ldG,
                                            const uint16_t *__restrict__ pairs,
                                            const double *__restrict__ Wbuf,
                                            const unsigned *__restrict__ flags, const Quad &Q,
                                            unsigned c, unsigned nchunk, unsigned CH,
                                            double *__restrict__ part,
                                            const uint8_t *__restrict__ plan_next, unsigned nw,
                                            bool fsplit)
{
  const unsigned l = threadIdx.x & 63u, w = threadIdx.x >> 6;
  const unsigned A0 = pairs[2u * Q.p0], B0 = pairs[2u * Q.p0 + 1u];
  const unsigned A1 = pairs[2u * Q.p1], B1 = pairs[2u * Q.p1 + 1u];
  // flags written by another launch's pair visits: vector loads, never
  // through the scalar cache
  const bool u0 = ld_relaxed(flags + Q.p0) != 0u, u1 = ld_relaxed(flags + Q.p1) != 0u;
  uint8_t pn0 = plan_next ? plan_next[Q.n0] : PLAN_FP64;
  uint8_t pn1 = plan_next ? plan_next[Q.n1] : PLAN_FP64;
  if (!u0 && !u1 && pn0 == PLAN_SKIP && pn1 == PLAN_SKIP)
    return;  // nothing to update, no Gram wanted
  // blocks of a pair without an update are read only for a next pair's Gram
  const bool g0 = pn0 != PLAN_SKIP, g1 = pn1 != PLAN_SKIP;
  const bool nt0 = (g0 && (Q.x0 == 0u || Q.y0 == 0u)) || (g1 && (Q.x1 == 0u || Q.y1 == 0u));
  const bool nt1 = (g0 && (Q.x0 == 1u || Q.y0 == 1u)) || (g1 && (Q.x1 == 1u || Q.y1 == 1u));
  const bool nt2 = (g0 && (Q.x0 == 2u || Q.y0 == 2u)) || (g1 && (Q.x1 == 2u || Q.y1 == 2u));
  const bool nt3 = (g0 && (Q.x0 == 3u || Q.y0 == 3u)) || (g1 && (Q.x1 == 3u || Q.y1 == 3u));
  const bool sp0 = fsplit && pn0 == PLAN_SPLIT, sp1 = fsplit && pn1 == PLAN_SPLIT;
  // W(k, j) of both pairs into LDS (row-major, padded)
  for (unsigned e = threadIdx.x; e < 2048u; e += blockDim.x) {
    const unsigned pp = e >> 10, k = (e >> 5) & 31u, j = e & 31u;
    S.Wl[pp][k * LDW + j] = F32(Wbuf + (size_t)(pp ? Q.p1 : Q.p0) * 1024u, k, j);
  }
  __syncthreads();
  const unsigned SUB = CH / nw;
  const size_t rbeg = (size_t)c * CH + (size_t)w * SUB;
  double *const Tw = S.Tb[w];
  d4 z = {0.0, 0.0, 0.0, 0.0};
  d4 g00 = z, g10 = z, g11 = z, h00 = z, h10 = z, h11 = z;
  if (w < nw) {
    double xa[8], xb[8];
#if UPGRAM_PREFETCH
    slab_load(G, ldG, rbeg, A0, B0, u0, l, xa, nt0, nt1);
    slab_load(G, ldG, rbeg, A1, B1, u1, l, xb, nt2, nt3);
#endif
#if UPGRAM_PREFETCH >= 2
    // a second slab in flight
    double ya[8], yb[8];
    if (16u < SUB) {
      slab_load(G, ldG, rbeg + 16u, A0, B0, u0, l, ya, nt0, nt1);
      slab_load(G, ldG, rbeg + 16u, A1, B1, u1, l, yb, nt2, nt3);
    }
#endif
    for (unsigned g = 0u; g < SUB; g += 16u) {
      d4 t0, t1, t2, t3;
#if UPGRAM_PREFETCH
      double ca[8], cb[8];
#pragma unroll
      for (unsigned u = 0u; u < 8u; ++u) {
        ca[u] = xa[u];
        cb[u] = xb[u];
      }
#if UPGRAM_PREFETCH >= 2
#pragma unroll
      for (unsigned u = 0u; u < 8u; ++u) {
        xa[u] = ya[u];
        xb[u] = yb[u];
      }
      if (g + 32u < SUB) {
        slab_load(G, ldG, rbeg + g + 32u, A0, B0, u0, l, ya, nt0, nt1);
        slab_load(G, ldG, rbeg + g + 32u, A1, B1, u1, l, yb, nt2, nt3);
      }
#else
      if (g + 16u < SUB) {
        slab_load(G, ldG, rbeg + g + 16u, A0, B0, u0, l, xa, nt0, nt1);
        slab_load(G, ldG, rbeg + g + 16u, A1, B1, u1, l, xb, nt2, nt3);
      }
#endif
      slab_compute(G, ldG, rbeg + g, A0, B0, u0, S.Wl[0], l, ca, t0, t1, Tw);
      slab_compute(G, ldG, rbeg + g, A1, B1, u1, S.Wl[1], l, cb, t2, t3, Tw);
#else
      slab_load(G, ldG, rbeg + g, A0, B0, u0, l, xa, nt0, nt1);
      slab_load(G, ldG, rbeg + g, A1, B1, u1, l, xb, nt2, nt3);
      slab_compute(G, ldG, rbeg + g, A0, B0, u0, S.Wl[0], l, xa, t0, t1, Tw);
      slab_compute(G, ldG, rbeg + g, A1, B1, u1, S.Wl[1], l, xb, t2, t3, Tw);
#endif
      if (pn0 != PLAN_SKIP) {
        const d4 X = sel4(Q.x0, t0, t1, t2, t3), Y = sel4(Q.y0, t0, t1, t2, t3);
        if (sp0)
          gram_split_slab(X, Y, l, g00, g10, g11);
        else
          gram_fp64_slab(X, Y, g00, g10, g11);
      }
      if (pn1 != PLAN_SKIP) {
        const d4 X = sel4(Q.x1, t0, t1, t2, t3), Y = sel4(Q.y1, t0, t1, t2, t3);
        if (sp1)
          gram_split_slab(X, Y, l, h00, h10, h11);
        else
          gram_fp64_slab(X, Y, h00, h10, h11);
      }
    }
    if (w && w < nw) {
      S.red[0][0][w - 1u][l] = g00;
      S.red[0][1][w - 1u][l] = g10;
      S.red[0][2][w - 1u][l] = g11;
      S.red[1][0][w - 1u][l] = h00;
      S.red[1][1][w - 1u][l] = h10;
      S.red[1][2][w - 1u][l] = h11;
    }
  }
  __syncthreads();
  if (w == 0u) {
    for (unsigned v = 1u; v < nw; ++v) {
      g00 = g00 + S.red[0][0][v - 1u][l];
      g10 = g10 + S.red[0][1][v - 1u][l];
      g11 = g11 + S.red[0][2][v - 1u][l];
      h00 = h00 + S.red[1][0][v - 1u][l];
      h10 = h10 + S.red[1][1][v - 1u][l];
      h11 = h11 + S.red[1][2][v - 1u][l];
    }
    if (pn0 != PLAN_SKIP)
      store_partial(part + ((size_t)Q.n0 * nchunk + c) * 768u, l, sp0, g00, g10, g11);
    if (pn1 != PLAN_SKIP)
      store_partial(part + ((size_t)Q.n1 * nchunk + c) * 768u, l, sp1, h00, h10, h11);
  }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(UPGRAM_WAVES)))
void k_upgram(double *__restrict__ G, unsigned ldG, const uint16_t *__restrict__ pairs,
              const double *__restrict__ Wbuf, const unsigned *__restrict__ flags,
              const Quad *__restrict__ quads, unsigned nchunk, unsigned CH,
              double *__restrict__ part, const uint8_t *__restrict__ plan_next, int fsplit)
{
  __shared__ UpgLds S;
  const unsigned qi = blockIdx.x / nchunk, c = blockIdx.x - qi * nchunk;
  const Quad Q = quads[qi];
  upgram_item(S, G, ldG, pairs, Wbuf, flags, Q, c, nchunk, CH, part, plan_next, blockDim.x >> 6,
              fsplit != 0);
}

// ---------------------------------------------------------------------------
// Staged sweep: the fused transition of half the quads of step s runs in the
// same launch as the visits of the pairs of another step whose inputs the
// previous launch completed, so the visits' latency hides under the fused
// items' HBM stream.  Workgroups [0, npw) hold two visits each; then come
// the niw item workgroups, j taking items j, j + niw, ..., and, with V,
// the V workgroups (one (pair, row chunk) update of V per wave).  The
// launch boundaries order everything else; the arithmetic is that of
// k_pair / k_upgram / k_update.
// LDS of a staged workgroup: a fused item's, or two pair visits'
union StageLds {
  UpgLds upg;
  PairLds pair[2];
};

template <bool SORT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(UPGRAM_WAVES)))
void k_stage(StageArgs A)
{
  __shared__ StageLds S;
  __shared__ __attribute__((aligned(16))) uint8_t S0[STRAT0_TAB_BYTES];
  const unsigned tid = threadIdx.x, b = blockIdx.x;
  if (b < A.npw) {
    if (tid < STRAT0_TAB_BYTES / 16u)
      reinterpret_cast<uint4 *>(S0)[tid] = reinterpret_cast<const uint4 *>(A.strat0)[tid];
    const unsigned half = tid >> 7, ht = tid & 127u, k = 2u * b + half;
    const bool present = k < A.npl;
    const unsigned pi = present ? A.plist[k] : 0u;
    // every Gram a staged launch's visits read comes from fused items (split
    // for split-planned pairs exactly when they form split Grams)
    const uint8_t plan =
        visit_plan((present && A.pplan) ? A.pplan[pi] : PLAN_FP64, A.fsplit != 0);
    pair_visit<SORT>(ht, present, pi, S.pair[half], &S.pair[half ^ 1u].go, A.ppairs, A.ppart,
                     A.nchunk, A.pW, A.pflags, A.ctl, A.prec, plan, S0, A.s0steps, A.nswp,
                     A.definite, A.nplus);
  }
  else if (b - A.npw >= A.niw) {
    // V items: one (pair, row chunk) per wave
    const unsigned j = b - A.npw - A.niw;
    const unsigned v = 4u * j + (tid >> 6);
    if (v < A.nvitems) {
      const unsigned pi = v / A.nchV, c = v - pi * A.nchV;
      if (A.vflags[pi]) {
        const unsigned r0 = c * A.chV;
        update_rows(A.V, A.ldV, r0, min(r0 + A.chV, A.nrowV), A.vpairs[2u * pi],
                    A.vpairs[2u * pi + 1u], A.vW + (size_t)pi * 1024u, tid & 63u);
      }
    }
  }
  else {
    // item workgroup j takes items j, j + stride, ...: stride niw, or with
    // the learned skip (items with work first) one item each while they fit
    // the slots beside the visits, else two each
    unsigned nitems = A.nquads * A.nchunk;
    unsigned stride = A.niw;
    if (A.qact) {
      nitems = min(nitems, ld_relaxed(A.qact) * A.nchunk);
      stride = (nitems <= A.islots) ? nitems : (nitems + 1u) / 2u;
    }
    const unsigned j = b - A.npw;
    for (unsigned i = j; i < nitems && j < stride; i += stride) {
      const unsigned qi = i / A.nchunk, c = i - qi * A.nchunk;
      // a copy in registers: through a reference into global memory its
      // byte fields would be reloaded after every store of the item
      const Quad Q = A.quads[A.qord ? A.qord[qi] : qi];
      upgram_item(S.upg, A.G, A.ldG, A.upairs, A.uW, A.uflags, Q, c, A.nchunk, A.CH, A.upart,
                  A.uplan, A.nw, A.fsplit != 0);
      __syncthreads();
    }
  }
}

// The visit lists of a staged sweep in the order the learned skip makes
// cheapest: per transition t and half (A: entries [0, h2), B: the rest) of
// the pair list, the pairs of step t+1 that run before the skipped ones
// (stable), so that the visit workgroups holding only skipped visits end at
// once and free their slots for fused items.  One workgroup of 256 per
// (transition, half), over the half in chunks of 256 entries.
__global__ __launch_bounds__(256) void k_order_visits(const uint16_t *__restrict__ plist,
                                                      uint16_t *__restrict__ out,
                                                      const uint8_t *__restrict__ plan,
                                                      unsigned np, unsigned h2)
{
  __shared__ unsigned run[256];
  __shared__ unsigned total;
  const unsigned t = blockIdx.x >> 1, half = blockIdx.x & 1u, tid = threadIdx.x;
  const unsigned off = t * np + (half ? h2 : 0u), n = half ? np - h2 : h2;
  const uint8_t *const pl = plan + (size_t)(t + 1u) * np;
  if (tid == 0u)
    total = 0u;
  __syncthreads();
  // running visits of the half
  for (unsigned c0 = 0u; c0 < n; c0 += 256u) {
    const unsigned i = c0 + tid;
    if (i < n && pl[plist[off + i]] != PLAN_SKIP)
      atomicAdd(&total, 1u);
  }
  __syncthreads();
  const unsigned nrun = total;
  unsigned base = 0u;  // running visits in the chunks before this one
  for (unsigned c0 = 0u; c0 < n; c0 += 256u) {
    const unsigned i = c0 + tid;
    const unsigned pi = (i < n) ? plist[off + i] : 0u;
    const unsigned f = (i < n && pl[pi] != PLAN_SKIP) ? 1u : 0u;
    // inclusive prefix sum of f over the chunk (Hillis-Steele in LDS)
    run[tid] = f;
    __syncthreads();
    for (unsigned d = 1u; d < 256u; d <<= 1) {
      const unsigned v = (tid >= d) ? run[tid - d] : 0u;
      __syncthreads();
      run[tid] += v;
      __syncthreads();
    }
    const unsigned before = base + run[tid] - f;  // running entries before i
    if (i < n)
      out[off + (f ? before : nrun + (i - before))] = (uint16_t)pi;
    base += run[255];
    __syncthreads();  // run[] is rewritten by the next chunk
  }
}

// The fused items of a staged sweep with the learned skip, heavy quads
// first: per transition t and half, the quads ordered (stably) by the work
// their items do -- an update per running pair of step t, a Gram per running
// pair of step t+1 -- so the items that return at once come last and fill
// the slots the heavy ones leave.  One workgroup of 64 per (transition,
// half); a half holds at most 64 ... nq/2 quads, looped in chunks of 64.
__global__ __launch_bounds__(64) void k_order_quads(const Quad *__restrict__ quads,
                                                    uint16_t *__restrict__ out,
                                                    unsigned *__restrict__ act,
                                                    const uint8_t *__restrict__ plan,
                                                    unsigned np, unsigned h)
{
  const unsigned t = blockIdx.x >> 1, half = blockIdx.x & 1u, l = threadIdx.x;
  const unsigned nq = np / 2u, q0 = half ? h : 0u, n = half ? nq - h : h;
  const Quad *const qs = quads + (size_t)t * nq;
  const uint8_t *const p0 = plan + (size_t)t * np, *const p1 = plan + (size_t)(t + 1u) * np;
  uint16_t *const o = out + (size_t)t * nq + q0;
  // counting sort by weight 4..0 (stable: chunks in order, lanes in order)
  unsigned start[5];
  {
    unsigned cnt[5] = {0u, 0u, 0u, 0u, 0u};
    for (unsigned c0 = 0u; c0 < n; c0 += 64u) {
      unsigned w = 5u;
      if (c0 + l < n) {
        const Quad Q = qs[q0 + c0 + l];
        w = 4u - ((p0[Q.p0] != PLAN_SKIP) + (p0[Q.p1] != PLAN_SKIP) + (p1[Q.n0] != PLAN_SKIP) +
                  (p1[Q.n1] != PLAN_SKIP));
      }
#pragma unroll
      for (unsigned k = 0u; k < 5u; ++k)
        cnt[k] += (unsigned)__popcll(__ballot(w == k));
    }
    unsigned acc = 0u;
#pragma unroll
    for (unsigned k = 0u; k < 5u; ++k) {
      start[k] = acc;
      acc += cnt[k];
    }
    // quads with any work (weight < 4), a vector store
    if (l == 0u)
      act[blockIdx.x] = n - cnt[4];
  }
  const unsigned long long below = (l == 0u) ? 0ull : (~0ull >> (64u - l));
  for (unsigned c0 = 0u; c0 < n; c0 += 64u) {
    unsigned w = 5u;
    if (c0 + l < n) {
      const Quad Q = qs[q0 + c0 + l];
      w = 4u - ((p0[Q.p0] != PLAN_SKIP) + (p0[Q.p1] != PLAN_SKIP) + (p1[Q.n0] != PLAN_SKIP) +
                (p1[Q.n1] != PLAN_SKIP));
    }
#pragma unroll
    for (unsigned k = 0u; k < 5u; ++k) {
      const unsigned long long m = __ballot(w == k);
      if (w == k)
        o[start[k] + (unsigned)__popcll(m & below)] = (uint16_t)(q0 + c0 + l);
      start[k] += (unsigned)__popcll(m);
    }
  }
}

void launch_order_quads(hipStream_t s, const Quad *quads, uint16_t *out, unsigned *act,
                        const uint8_t *plan, unsigned np, unsigned steps, unsigned h)
{
  hipLaunchKernelGGL(k_order_quads, dim3(2u * (steps - 1u)), dim3(64), 0, s, quads, out, act, plan,
                     np, h);
}

void launch_order_visits(hipStream_t s, const uint16_t *plist, uint16_t *out, const uint8_t *plan,
                         unsigned np, unsigned steps, unsigned h2)
{
  hipLaunchKernelGGL(k_order_visits, dim3(2u * (steps - 1u)), dim3(256), 0, s, plist, out, plan, np,
                     h2);
}

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

// Start of a level-2 sweep: keep the previous counter, clear the running
// one, advance the sweep index (a kernel, so no memset node in the graph).
// Start of a level-2 sweep: the previous sweep's counter becomes prev_cvg,
// the counters restart, and the forms' switches for this sweep follow from
// the previous sweep: a sweep after one that had skipped or split visits and
// no big rotation certifies (every visit, FP64); otherwise skipping is on
// when the form is and the previous sweep's big-rotation fraction reached
// min_frac, the split Gram when the form is.
__global__ void k_sweep_begin(RunCtl *ctl)
{
  if (threadIdx.x == 0u) {
    const unsigned prev_b = (unsigned)(ctl->cvg >> 32);
    const bool first = (ctl->sweep == 0u);
    const bool certify = !first && !prev_b && (ctl->skipped || ctl->split);
    const float frac = (float)prev_b / ctl->max_rot;
    ctl->skip_on = (ctl->learned && !first && !certify && frac >= ctl->min_frac) ? 1u : 0u;
    ctl->split_on = (ctl->split_form && !certify) ? 1u : 0u;
    ctl->prev_cvg = ctl->cvg;
    ctl->cvg = 0ull;
    ctl->sweep += 1u;
    ctl->skipped = 0u;
    ctl->split = 0u;
    ctl->proposed = 0u;
  }
}

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
                 unsigned nchunk, unsigned CH, double *part, const uint8_t *skip)
{
  const unsigned nw = gram_waves(CH);
  hipLaunchKernelGGL(k_gram, dim3(np * nchunk), dim3(64u * nw), 0, s, G, ldG, pairs, nchunk, CH, part,
                     skip);
}

void launch_upgram(hipStream_t s, double *G, unsigned ldG, const uint16_t *pairs, const double *W,
                   const unsigned *flags, const Quad *quads, unsigned nquads, unsigned nchunk,
                   unsigned CH, double *part, const uint8_t *plan_next, bool fsplit)
{
  const unsigned nw = gram_waves(CH);
  hipLaunchKernelGGL(k_upgram, dim3(nquads * nchunk), dim3(64u * nw), 0, s, G, ldG, pairs, W, flags,
                     quads, nchunk, CH, part, plan_next, fsplit ? 1 : 0);
}

void launch_stage(hipStream_t s, bool sort, const StageArgs &a)
{
  const unsigned nb = a.npw + a.niw + a.nvw;
  if (!nb)
    return;
  if (sort)
    hipLaunchKernelGGL(k_stage<true>, dim3(nb), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(k_stage<false>, dim3(nb), dim3(256), 0, s, a);
}

void launch_gram_split(hipStream_t s, const double *G, unsigned ldG, const uint16_t *pairs,
                       unsigned np, unsigned nchunk, unsigned CH, double *part, const uint8_t *plan)
{
  const unsigned nw = gram_waves(CH);
  hipLaunchKernelGGL(k_gram_split, dim3(np * nchunk), dim3(64u * nw), 0, s, G, ldG, pairs, nchunk, CH,
                     part, plan);
}

void launch_pair(hipStream_t s, bool sort, const uint16_t *pairs, unsigned np, const double *part,
                 unsigned nchunk, double *W, unsigned *flags, RunCtl *ctl, VisitRec *rec,
                 const uint8_t *skip, bool split_src, const uint8_t *strat0, unsigned s0steps,
                 unsigned nswp, int definite, unsigned nplus)
{
  const int ss = split_src ? 1 : 0;
  if (sort)
    hipLaunchKernelGGL(k_pair<true>, dim3(np), dim3(128), 0, s, pairs, part, nchunk, W, flags, ctl,
                       rec, skip, ss, strat0, s0steps, nswp, definite, nplus);
  else
    hipLaunchKernelGGL(k_pair<false>, dim3(np), dim3(128), 0, s, pairs, part, nchunk, W, flags, ctl,
                       rec, skip, ss, strat0, s0steps, nswp, definite, nplus);
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

void launch_sweep_begin(hipStream_t s, RunCtl *ctl)
{
  hipLaunchKernelGGL(k_sweep_begin, dim3(1), dim3(64), 0, s, ctl);
}

void launch_dbg_scalar(hipStream_t s, int kind, int n, const double *a, const double *b,
                       const double *c, double *oc, double *ot, int *ob)
{
  hipLaunchKernelGGL(k_dbg_scalar, dim3((n + 255) / 256), dim3(256), 0, s, kind, n, a, b, c, oc, ot,
                     ob);
}

}  // namespace hj
