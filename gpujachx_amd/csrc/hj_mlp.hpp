// hj_mlp.hpp -- sizes and parameter layout of the pair-skip MLP.
#ifndef HJ_MLP_HPP
#define HJ_MLP_HPP

#include <hip/hip_runtime.h>
#include <cstdint>

#include "hj_kernels.hpp"

namespace hj {

constexpr unsigned MLP_F = 12u;  // features of a visit
constexpr unsigned MLP_H = 64u;  // hidden units per layer
// parameter vector: W1[F][H], b1[H], W2[H][H], b2[H], W3[H], b3
constexpr unsigned MLP_OFF_W1 = 0u;
constexpr unsigned MLP_OFF_B1 = MLP_OFF_W1 + MLP_F * MLP_H;
constexpr unsigned MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr unsigned MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr unsigned MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr unsigned MLP_OFF_B3 = MLP_OFF_W3 + MLP_H;
constexpr unsigned MLP_NPARAM = MLP_OFF_B3 + 1u;

// bw: the block width of the pairs (16, or 32 for 64-column pairs)
void launch_plan(hipStream_t s, const VisitRec *rec, unsigned nkeys, const uint16_t *strat1,
                 unsigned nplus, float log2n1, float max_rot, RunCtl *ctl,
                 const uint16_t *model, bool learned, uint8_t *plan, unsigned bw = 16u);

}  // namespace hj

#endif
