// hj_kernels.hpp -- launchers of the gfx950 kernels (hj_kernels.hip).
#ifndef HJ_KERNELS_HPP
#define HJ_KERNELS_HPP

#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

#define STRAT0_TAB_BYTES 1024u  // 32 steps x 16 pairs x 2 indices
#define MAX_GRAM_CHUNKS 16u     // upper bound of gram_nchunk()
#define GRAM_WAVES 4u           // waves per Gram chunk (upper bound)

namespace hj {

// Per-visit record of one block pair (indexed by step * npairs + slot, the
// pair's position in the level-1 strategy), written by k_pair at every
// executed visit and read by the skip predictor.
struct VisitRec {
  unsigned rot_s;   // rotations at the last executed visit
  unsigned rot_b;   // big rotations at the last executed visit
  float maxcos;     // max |A_pq| / sqrt(A_pp A_qq) over the pivots of its first inner sweep
  float l2ni;       // log2 ||G_i||_F^2 of the first block (from the Gram)
  float l2nj;       // log2 ||G_j||_F^2 of the second block
  unsigned sweep;   // level-2 sweep (1-based) of the last executed visit, 0 = never
  unsigned skips;   // visits skipped since then
  unsigned flags;   // REC_* bits of the last executed visit
};

// VisitRec.flags
constexpr unsigned REC_SPLIT = 1u;       // the visit used the split-precision Gram
constexpr unsigned REC_SPLIT_FAIL = 2u;  // ... whose Cholesky failed: no update, FP64 next

// visit plan of the running sweep (one byte per (step, pair))
constexpr uint8_t PLAN_FP64 = 0u;   // FP64 Gram
constexpr uint8_t PLAN_SKIP = 1u;   // skipped
constexpr uint8_t PLAN_SPLIT = 2u;  // split-precision Gram
// proposals of k_plan, settled by k_plan_gate before the steps run
constexpr uint8_t PLAN_PROP_SKIP_FP64 = 3u;   // skip, else FP64
constexpr uint8_t PLAN_PROP_SKIP_SPLIT = 4u;  // skip, else split

// Device-side control words of one level-2 run.
struct RunCtl {
  unsigned long long cvg;       // rotation counter of the running sweep (big << 32 | all)
  unsigned long long prev_cvg;  // counter of the previous sweep
  unsigned sweep;               // running sweep, 1-based
  unsigned skip_on;             // skipping allowed in this sweep
  float threshold;              // skip when P(big rotation) < threshold
  unsigned max_skips;           // at most this many consecutive skipped visits
  unsigned skipped;             // visits skipped in the running sweep
  unsigned split_on;            // split-precision Gram allowed in this sweep
  float tau;                    // ... for pairs whose last max |cos| exceeds tau
  unsigned split;               // visits with the split Gram in the running sweep
  // policy of the forms, set once per run; k_sweep_begin derives skip_on /
  // split_on of every sweep from it and the previous sweep's counters
  unsigned learned;             // skipping form on
  unsigned split_form;          // split form on
  float min_frac;               // skip only after a sweep with at least this big-rotation fraction
  float min_skip_frac;          // ... and only if the model proposes at least this fraction of skips
  float max_rot;                // most big rotations a sweep can have (visits x 496)
  unsigned proposed;            // skips the model proposed for the running sweep
};

// waves sharing one Gram chunk of CH rows: each takes CH/nw rows, a multiple
// of 64 (the wave results are added in wave order)
inline unsigned gram_waves(unsigned CH)
{
  unsigned nw = GRAM_WAVES;
  while (nw > 1u && (CH % (64u * nw)))
    nw >>= 1u;
  return nw;
}

void launch_gram(hipStream_t s, const double *G, unsigned ldG, const uint16_t *pairs, unsigned np,
                 unsigned nchunk, unsigned CH, double *part, const uint8_t *skip);
// A closed quad of a step transition s -> s+1: two pairs of step s whose
// four blocks form two pairs of step s+1.  Tiles 0..3 are the first and
// second block of pair p0, then of pair p1.
struct Quad {
  uint16_t p0, p1;          // slots of the current pairs in step s
  uint16_t n0, n1;          // slots of the next pairs in step s+1
  uint8_t x0, y0, x1, y1;   // tiles of the first / second block of n0 and n1
};

bool closed_quads(const uint16_t *tab, unsigned steps, unsigned np, std::vector<Quad> &out,
                  std::vector<char> *fused);

void launch_upgram(hipStream_t s, double *G, unsigned ldG, const uint16_t *pairs, const double *W,
                   const unsigned *flags, const Quad *quads, unsigned nquads, unsigned nchunk,
                   unsigned CH, double *part, const uint8_t *plan_next, bool fsplit);

// One launch of a staged sweep: the visits of a list of pairs of one step
// (two per workgroup) beside the fused transition of a range of closed quads
// of another (nchunk items per quad) and, when V is accumulated, the update
// of V by the W of a step whose visits are complete.  The parts touch
// disjoint data: the visits read Gram partials and write W / flags of their
// step's buffers, the fused items read the W / flags of theirs and write G
// and the partials of the next step's, the V items read W / flags and write
// V.
struct StageArgs {
  double *G;
  unsigned ldG;
  unsigned npl, npw;       // pair visits, their workgroups (ceil(npl / 2))
  const uint16_t *plist;   // [npl] pair slots of the visits' step
  const uint16_t *ppairs;  // level-1 pairs of the visits' step
  const double *ppart;     // Gram partials the visits read
  double *pW;              // W the visits write
  unsigned *pflags;        // update flags the visits write
  VisitRec *prec;          // records of the visits' step
  const uint8_t *pplan;    // plans of the visits' step, or null
  const uint16_t *upairs;  // level-1 pairs of the fused items' step
  const double *uW;        // W of that step
  const unsigned *uflags;  // update flags of that step
  const Quad *quads;       // first quad of the range (with qord: of the transition)
  const uint16_t *qord;    // the range's quads in item order, or null (in index order)
  const unsigned *qact;    // with qord: how many of them (first in that order) have work, or null
  unsigned islots;         // with qact: item workgroups that fit beside the visits at once
  unsigned nquads;
  unsigned nchunk, CH, nw;
  unsigned niw;            // item workgroups (1..nquads*nchunk; 0 without items)
  double *upart;           // Gram partials of the next step
  const uint8_t *uplan;    // plans of the next step, or null
  int fsplit;              // items form split Grams for split-planned pairs
                           // (and the visits' Grams were formed that way)
  // update of V by the W of one step: (pair, row chunk) items, one per wave
  double *V;
  unsigned ldV, nrowV, chV, nchV;
  unsigned nvitems, nvw;   // items (pairs x nchV), their workgroups (ceil(nvitems / 4))
  const uint16_t *vpairs;  // level-1 pairs of that step
  const double *vW;        // its W
  const unsigned *vflags;  // its update flags
  RunCtl *ctl;
  const uint8_t *strat0;
  unsigned s0steps, nswp;
  int definite;
  unsigned nplus;
};

void launch_stage(hipStream_t s, bool sort, const StageArgs &a);
// quads of a staged sweep's transitions, heaviest items first (learned skip)
void launch_order_quads(hipStream_t s, const Quad *quads, uint16_t *out, unsigned *act,
                        const uint8_t *plan, unsigned np, unsigned steps, unsigned h);
// visit lists of a staged sweep, running pairs first (learned skip)
void launch_order_visits(hipStream_t s, const uint16_t *plist, uint16_t *out, const uint8_t *plan,
                         unsigned np, unsigned steps, unsigned h2);

// QR-path visit (hj_qr.hip): factorization of the pair read from G, level-0
// sweeps (Drmac's for sorted definite blocks when drmac), W and the flags
void launch_pair_qr(hipStream_t s, bool sort, bool drmac, const double *G, unsigned ldG,
                    unsigned nrow, const uint16_t *pairs, unsigned np, double *W, unsigned *flags,
                    RunCtl *ctl, const uint8_t *strat0, unsigned s0steps, unsigned nswp,
                    int definite, unsigned nplus);


void launch_gram_split(hipStream_t s, const double *G, unsigned ldG, const uint16_t *pairs,
                       unsigned np, unsigned nchunk, unsigned CH, double *part, const uint8_t *plan);
// split_src: the partials come from a standalone Gram launch, so pairs
// planned split hold split Grams (else every pair is an FP64 visit)
void launch_pair(hipStream_t s, bool sort, const uint16_t *pairs, unsigned np, const double *part,
                 unsigned nchunk, double *W, unsigned *flags, RunCtl *ctl, VisitRec *rec,
                 const uint8_t *skip, bool split_src, const uint8_t *strat0, unsigned s0steps,
                 unsigned nswp, int definite, unsigned nplus);
void launch_update(hipStream_t s, double *G, unsigned ldG, unsigned nrowG, unsigned chG,
                   double *V, unsigned ldV, unsigned nrowV, unsigned chV, const uint16_t *pairs,
                   unsigned np, const double *W, const unsigned *flags);
void launch_initD(hipStream_t s, double *G, double *D, unsigned ifc, unsigned nRow, unsigned nRank,
                  unsigned nPlus, unsigned ldG);
void launch_initV(hipStream_t s, double *V, unsigned ifc, unsigned nRank, unsigned ldV);

// The wide-block form (hj_wide.hip): blocks of 32 columns, 64-column pairs;
// Gram partials of wide_part_doubles() doubles per (pair, row chunk), W of
// 64x64 per pair, the update in wide_update_chunks(nrow) workgroups a pair.
unsigned wide_part_doubles();
unsigned wide_update_chunks(unsigned nrow);
// plan (null: every visit runs): the step's plans, skipped pairs get no Gram
// and no visit; rec (null: none): the visits' records
void launch_gram_w(hipStream_t s, const double *G, unsigned ldG, const uint16_t *pairs,
                   unsigned np, unsigned nchunk, unsigned CH, double *part, const uint8_t *plan);
void launch_pair_w(hipStream_t s, bool sort, const uint16_t *pairs, unsigned np,
                   const double *part, unsigned nchunk, double *W, unsigned *flags,
                   RunCtl *ctl, const uint8_t *strat0, unsigned s0steps,
                   unsigned nswp, int definite, unsigned nplus, const uint8_t *plan,
                   VisitRec *rec);
void launch_update_w(hipStream_t s, double *M, unsigned ldM, unsigned nrow,
                     const uint16_t *pairs, unsigned np, const double *W,
                     const unsigned *flags);
void launch_sweep_begin(hipStream_t s, RunCtl *ctl);
void launch_dbg_scalar(hipStream_t s, int kind, int n, const double *a, const double *b,
                       const double *c, double *oc, double *ot, int *ob);


}  // namespace hj

#endif
