// hj_fused.hpp -- wave-level pieces of the fused step transition (the block
// update of a closed quad's two pairs feeding the next step's Gram from
// registers), shared by k_upgram / k_stage (hj_kernels.hip) and the batched
// launch (hj_batch.hip).
//
// Update orientation D = G W (A = G, B = W): lane l, register r of the
// result holds row 4r + l/16 of the 16-row slab, column l%16 of the tile --
// exactly the Gram's operand; the products and their chain over k = 0..31
// are those of k_update (operands swapped: the same fma results).
#ifndef HJ_FUSED_HPP
#define HJ_FUSED_HPP

#include "hj_device.hpp"

typedef double d4 __attribute__((ext_vector_type(4)));

namespace hj {

__device__ __forceinline__ d4 mfma(double a, double b, d4 c)
{
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// tile i (uniform) of four, selected by value
__device__ __forceinline__ d4 sel4(unsigned i, d4 a, d4 b, d4 c, d4 d)
{
  d4 o;
#pragma unroll
  for (unsigned r = 0u; r < 4u; ++r) {
    const double ab = (i & 1u) ? b[r] : a[r];
    const double cd = (i & 1u) ? d[r] : c[r];
    o[r] = (i & 2u) ? cd : ab;
  }
  return o;
}

#define LDW 33u  // padded row length of the W copies in LDS
// waves per SIMD the fused kernel is compiled for (register budget 512/n)
#ifndef UPGRAM_WAVES
#define UPGRAM_WAVES 2
#endif
#ifndef UPGRAM_PREFETCH
#define UPGRAM_PREFETCH 1
#endif
// streaming (non-temporal) loads and stores of G in the fused kernel
#ifndef UPGRAM_NT
#define UPGRAM_NT 1
#endif
#if UPGRAM_NT
#define UG_LOAD(p) __builtin_nontemporal_load(p)
#define UG_STORE(v, p) __builtin_nontemporal_store(v, p)
#else
#define UG_LOAD(p) (*(p))
#define UG_STORE(v, p) (*(p) = (v))
#endif

// One 16-row slab of a pair, in the update's operand layout (lane l: row
// l%16 of pair column 4u + l/16, x[u], u = 0..7; 128 contiguous bytes of a
// column per instruction).  A pair without an update is read the same way
// and turned into the Gram layout in LDS (slab_compute); la / lb false: no
// next pair's Gram wants block a / b, nothing to read.
__device__ __forceinline__ void slab_load(const double *__restrict__ G, unsigned ldG, size_t row,
                                          unsigned a, unsigned b, bool upd, unsigned l, double *x,
                                          bool la = true, bool lb = true)
{
  if (upd) {
#pragma unroll
    for (unsigned u = 0u; u < 8u; ++u) {
      const unsigned cc = 4u * u + (l >> 4);
      const unsigned gc = (cc < 16u) ? (16u * a + cc) : (16u * b + cc - 16u);
      x[u] = UG_LOAD(G + (size_t)gc * ldG + row + (l & 15u));
    }
  }
  else {
    // one branch per block (uniform), the loads of a block back to back
    const double *const ia = G + (size_t)(16u * a + (l >> 4)) * ldG + row + (l & 15u);
    const double *const ib = G + (size_t)(16u * b + (l >> 4)) * ldG + row + (l & 15u);
    if (la) {
#pragma unroll
      for (unsigned u = 0u; u < 4u; ++u)
        x[u] = UG_LOAD(ia + (size_t)(4u * u) * ldG);
    }
    else {
#pragma unroll
      for (unsigned u = 0u; u < 4u; ++u)
        x[u] = 0.0;
    }
    if (lb) {
#pragma unroll
      for (unsigned u = 0u; u < 4u; ++u)
        x[4u + u] = UG_LOAD(ib + (size_t)(4u * u) * ldG);
    }
    else {
#pragma unroll
      for (unsigned u = 0u; u < 4u; ++u)
        x[4u + u] = 0.0;
    }
  }
}

#ifndef UPGRAM_LDS_STORE
#define UPGRAM_LDS_STORE 1
#endif
#define LDT 17u  // padded column length of a wave's 16x16 transpose buffer

// stores a 16x16 tile held in the Gram layout (lane l, register r: row
// 4r + l/16, column l%16) to columns 16 blk .. of G, rows row .. row+15: via
// the wave's LDS buffer T the lanes write 16 consecutive rows (128 bytes) of
// one column per instruction
__device__ __forceinline__ void store_tile(double *__restrict__ G, unsigned ldG, size_t row,
                                           unsigned blk, const d4 &t, unsigned l, double *T)
{
#if UPGRAM_LDS_STORE
#pragma unroll
  for (unsigned r = 0u; r < 4u; ++r)
    T[(l & 15u) * LDT + 4u * r + (l >> 4)] = t[r];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  double v[4];
#pragma unroll
  for (unsigned q = 0u; q < 4u; ++q)
    v[q] = T[(4u * q + (l >> 4)) * LDT + (l & 15u)];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (unsigned q = 0u; q < 4u; ++q)
    UG_STORE(v[q], G + (size_t)(16u * blk + 4u * q + (l >> 4)) * ldG + row + (l & 15u));
#else
  double *const o = G + (size_t)(16u * blk + (l & 15u)) * ldG + row + (l >> 4);
#pragma unroll
  for (unsigned r = 0u; r < 4u; ++r)
    o[4u * r] = t[r];
#endif
}

// [G_a G_b] W of the slab, W(k, j) = Wl[k * LDW + j] in LDS, written back;
// t0, t1: the slab of blocks a, b in the Gram layout
__device__ __forceinline__ void slab_compute(double *__restrict__ G, unsigned ldG, size_t row,
                                             unsigned a, unsigned b, bool upd, const double *Wl,
                                             unsigned l, const double *x, d4 &t0, d4 &t1, double *T)
{
  if (upd) {
    d4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = acc0;
#pragma unroll
    for (unsigned u = 0u; u < 8u; ++u) {
      const double *const wr = Wl + (4u * u + (l >> 4)) * LDW + (l & 15u);
      acc0 = mfma(x[u], wr[0], acc0);
      acc1 = mfma(x[u], wr[16], acc1);
    }
    store_tile(G, ldG, row, a, acc0, l, T);
    store_tile(G, ldG, row, b, acc1, l, T);
    t0 = acc0;
    t1 = acc1;
  }
  else {
    // no update: the slab to the Gram layout (lane l, register r: row
    // 4r + l/16 of column l%16), one block at a time through T
#pragma unroll
    for (unsigned h = 0u; h < 2u; ++h) {
#pragma unroll
      for (unsigned u = 0u; u < 4u; ++u)
        T[(4u * u + (l >> 4)) * LDT + (l & 15u)] = x[4u * h + u];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      d4 t;
#pragma unroll
      for (unsigned r = 0u; r < 4u; ++r)
        t[r] = T[(l & 15u) * LDT + 4u * r + (l >> 4)];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (h)
        t1 = t;
      else
        t0 = t;
    }
  }
}

__device__ __forceinline__ void gram_fp64_slab(const d4 &X, const d4 &Y, d4 &a00, d4 &a10, d4 &a11)
{
#pragma unroll
  for (unsigned r = 0u; r < 4u; ++r) {
    a00 = mfma(X[r], X[r], a00);
    a10 = mfma(Y[r], X[r], a10);
    a11 = mfma(Y[r], Y[r], a11);
  }
}

}  // namespace hj

#endif
