// hj_kernels.hpp -- launchers of the gfx950 kernels (hj_kernels.hip).
#ifndef HJ_KERNELS_HPP
#define HJ_KERNELS_HPP

#include <hip/hip_runtime.h>
#include <cstdint>

#define STRAT0_TAB_BYTES 1024u  // 32 steps x 16 pairs x 2 indices

namespace hj {

void launch_gram(hipStream_t s, const double *G, unsigned ldG, const uint16_t *pairs, unsigned np,
                 unsigned nchunk, unsigned CH, double *part);
void launch_pair(hipStream_t s, bool sort, const uint16_t *pairs, unsigned np, const double *part,
                 unsigned nchunk, double *W, unsigned *flags, unsigned long long *cvg,
                 unsigned *pstat, const uint8_t *strat0, unsigned s0steps, unsigned nswp,
                 int definite, unsigned nplus);
void launch_update(hipStream_t s, double *G, unsigned ldG, unsigned nrowG, unsigned chG,
                   double *V, unsigned ldV, unsigned nrowV, unsigned chV, const uint16_t *pairs,
                   unsigned np, const double *W, const unsigned *flags);
void launch_initD(hipStream_t s, double *G, double *D, unsigned ifc, unsigned nRow, unsigned nRank,
                  unsigned nPlus, unsigned ldG);
void launch_initV(hipStream_t s, double *V, unsigned ifc, unsigned nRank, unsigned ldV);
void launch_zero_u64(hipStream_t s, unsigned long long *p);
void launch_dbg_scalar(hipStream_t s, int kind, int n, const double *a, const double *b,
                       const double *c, double *oc, double *ot, int *ob);

}  // namespace hj

#endif
