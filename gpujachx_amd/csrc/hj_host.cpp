// hj_host.cpp -- level-2 driver (hypJacL2), the device-resident variant, the
// thin level-1 layer and device configuration, on the HIP runtime.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <array>
#include <vector>

#include "../../include/hypjac.h"
#include "hj_kernels.hpp"
#include "hj_mlp.hpp"
#include "hj_strat.hpp"
#include "hj_util.hpp"

namespace hj {

#define HJ_MAX_SWEEP_CAP 256u
struct Stats {
  unsigned nswp = 0u;
  unsigned rot_s[HJ_MAX_SWEEP_CAP] = {0u};
  unsigned rot_b[HJ_MAX_SWEEP_CAP] = {0u};
  double sec[HJ_MAX_SWEEP_CAP] = {0.0};
  unsigned skip[HJ_MAX_SWEEP_CAP] = {0u};
  unsigned split[HJ_MAX_SWEEP_CAP] = {0u};
  // kinds 0-4: per-launch events (profiling); 5: graph-launched sweeps
  double kms[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  unsigned long long kcnt[6] = {0ull, 0ull, 0ull, 0ull, 0ull, 0ull};
  unsigned long long skipped = 0ull, visits = 0ull, splits = 0ull;
};
static Stats g_stats;
static int g_profile = 0;
static long g_step_limit = 0;
static unsigned g_max_sweeps = HYPJAC_NSWEEP;  // level-2 sweep cap
// block width: 16 (the reference's) or 32 (64-column pairs, plain FP64)
static unsigned g_width = 16u;
// forms and the pair-skip model
static unsigned g_form = 0u;
static bool g_model_ready = false;
static uint16_t g_model_bf16[MLP_NPARAM];
static float g_skip_threshold = 0.05f;
static unsigned g_skip_max = 2u;
static float g_split_tau = 1e-3f;  // split Gram while the pair's last max |cos| > tau
// the fused transitions form the split Gram of split-planned pairs (else
// only the standalone Gram launches do, the fused ones FP64).  Off by
// default: at N = 8192 (bench matrix class, two seeds) the learned+split
// SVD took 0-5 % longer with it and the HSVD (NPLUS = N/2) 6-7 % (a sweep
// more): the splits' VALU work and registers cost the fused item more than
// the BF16 MFMAs save over the FP64 Gram's.
static int g_fused_split = 0;
// skipping only while the previous sweep's big rotations are at least this
// fraction of the most a sweep can have (the last sweeps run unskipped)
static float g_skip_min_frac = 0.01f;
// runs that mix J signs skip only above this multiple of the floor (their
// skips cost more extra sweeps: N = 8192 and 16384, measured), per block
// width (the big-rotation fraction of a 64-column pair's sweep is counted
// against 2016 pivot pairs)
static float g_skip_floor_mixed = 3.0f;
static float g_skip_floor_mixed_w32 = 6.0f;
// ... and only in sweeps where the model proposes at least this fraction of
// all visits
static float g_skip_min_propose = 0.1f;
// ... and only in runs with at least this many visits per sweep (small runs
// converge in few sweeps; one sweep added costs more than skipping saves)
static unsigned g_skip_min_visits = 32000u;
// per-sweep visit records copied to the host (trace collection)
// factorization of the block pairs: HJ_FACT_CHOLESKY (Gram + Cholesky),
// HJ_FACT_QR, HJ_FACT_QR_DRMAC
static int g_fact = HJ_FACT_CHOLESKY;
// staged sweeps (k_stage): 0 off, 1 where a step's block updates move at
// least STAGE_MIN_BYTES, 2 wherever they apply
static int g_stage = 1;
// Two staged launches cost about 2 max(Th, Tv) + 22 us per step against
// 2 Th + Tv step by step (Th: half the step's block updates at the HBM
// rate, Tv ~ 70 us: a visit's latency): they pay once Th exceeds about
// 46 us, i.e. 4.8e8 bytes read and written per step (N = 8192: 1.07e9;
// M = 4096, N = 2048 with V: 2.0e8, measured 86.5 -> 104 ms staged).
constexpr double STAGE_MIN_BYTES = 4.8e8;
static int g_last_stage = 0;
static void *g_trace_buf = nullptr;
static unsigned g_trace_sweeps = 0u;

static bool quiet()
{
  const char *q = getenv("HJ_QUIET");
  return q && *q && *q != '0';
}

unsigned gram_nchunk(unsigned nrow, unsigned npairs)
{
  unsigned target = npairs ? (2048u / npairs) : 1u;
  if (target > 16u) target = 16u;
  if (!target) target = 1u;
  const unsigned g = nrow / 64u;
  unsigned nc = 1u;
  while ((nc * 2u) <= target && g && !(g % (nc * 2u)))
    nc *= 2u;
  return nc;
}

// rows per wave of k_update: at least ~8192 waves per step, 64..1024 rows
static unsigned update_rows(unsigned nrow, unsigned nV, unsigned np)
{
  const unsigned long long total = (unsigned long long)np * (nrow + nV);
  unsigned ch = 64u;
  while (ch < 1024u && total / ch > 8192ull)
    ch *= 2u;
  return ch;
}

// The closed quads of every transition t -> t+1 of a level-1 table (flat
// [steps][np][2]): for each pair of step t, the pair holding its first
// block's next partner; the transition is closed when the two pairs' four
// blocks form two pairs of step t+1.  out: [steps-1][np/2] (entries of
// open transitions undefined); fused[t] = transition t is closed.  Returns
// whether any transition is closed.
bool closed_quads(const uint16_t *tab, unsigned steps, unsigned np, std::vector<Quad> &out,
                  std::vector<char> *fused)
{
  if (fused) fused->assign(steps, 0);
  if (np < 2u || (np & 1u) || steps < 2u) return false;
  out.assign((size_t)(steps - 1u) * (np / 2u), Quad());
  const unsigned nb = 2u * np;
  std::vector<int> slotA(nb), slotB(nb), partB(nb);
  bool any = false;
  auto T = [&](unsigned t, unsigned p, unsigned k) { return (unsigned)tab[((size_t)t * np + p) * 2u + k]; };
  for (unsigned t = 0u; t + 1u < steps; ++t) {
    for (unsigned p = 0u; p < np; ++p) {
      slotA[T(t, p, 0)] = (int)p;
      slotA[T(t, p, 1)] = (int)p;
      const unsigned x = T(t + 1u, p, 0), y = T(t + 1u, p, 1);
      slotB[x] = (int)p;
      slotB[y] = (int)p;
      partB[x] = (int)y;
      partB[y] = (int)x;
    }
    std::vector<char> done(np, 0);
    unsigned nq = 0u;
    bool ok = true;
    for (unsigned p = 0u; p < np && ok; ++p) {
      if (done[p]) continue;
      const unsigned a = T(t, p, 0), b = T(t, p, 1);
      const unsigned p1 = (unsigned)slotA[partB[a]];
      if (p1 == p || done[p1]) { ok = false; break; }
      const unsigned c = T(t, p1, 0), d = T(t, p1, 1);
      const unsigned pb = (unsigned)partB[b];
      if (!((pb == c || pb == d) && pb != (unsigned)partB[a])) { ok = false; break; }
      auto tile = [&](unsigned blk) -> uint8_t {
        return (uint8_t)((blk == a) ? 0u : (blk == b) ? 1u : (blk == c) ? 2u : 3u);
      };
      Quad q;
      q.p0 = (uint16_t)p;
      q.p1 = (uint16_t)p1;
      q.n0 = (uint16_t)slotB[a];
      q.n1 = (uint16_t)slotB[b];
      q.x0 = tile(T(t + 1u, q.n0, 0));
      q.y0 = tile(T(t + 1u, q.n0, 1));
      q.x1 = tile(T(t + 1u, q.n1, 0));
      q.y1 = tile(T(t + 1u, q.n1, 1));
      done[p] = done[p1] = 1;
      out[(size_t)t * (np / 2u) + nq++] = q;
    }
    const bool closed = ok && nq == np / 2u;
    if (fused) (*fused)[t] = closed ? 1 : 0;
    any = any || closed;
  }
  return any;
}

// Device workspace of one level-2 run.
struct Work {
  unsigned np = 0u, steps = 0u, s0steps = 0u, nchunk = 0u, CH = 0u, chG = 0u, chV = 0u;
  uint16_t *d_s1 = nullptr;
  uint8_t *d_s0 = nullptr;
  double *d_part = nullptr, *d_W = nullptr;
  unsigned *d_flags = nullptr;
  RunCtl *d_ctl = nullptr, *h_ctl = nullptr;
  VisitRec *d_rec = nullptr;   // [steps][np] last executed visit of each pair
  uint8_t *d_skip = nullptr;   // [steps][np] skip mask of the running sweep
  uint16_t *d_model = nullptr; // bf16 MLP parameters
  Quad *d_quads = nullptr;     // [steps - 1][np / 2] closed quads of each step transition
  std::vector<char> fuse;      // transition s -> s+1 runs as one fused update + Gram
  std::vector<Quad> hquads;
  // staged sweep: per transition the next pairs of quads [0, nq/2) then of
  // the rest; transitions whose first-half quads need a pair of the previous
  // transition's second half
  bool stage_ok = false;
  uint16_t *d_plist = nullptr;
  uint16_t *d_plist_run = nullptr;  // d_plist reordered per sweep (learned skip)
  uint16_t *d_qord = nullptr;       // per transition, the quads in item order (learned skip)
  unsigned *d_qact = nullptr;       // per transition and half, its quads with any work (learned skip)
  std::vector<char> cross;

  size_t keys() const { return (size_t)np * steps; }
  // step parity buffers: the Gram partials a step's visits read, the W and
  // update flags they write (parity 0 is what the step-by-step path uses)
  double *part(unsigned par) const { return d_part + (size_t)par * 768u * nchunk * np; }
  double *W(unsigned par) const { return d_W + (size_t)par * 1024u * np; }
  unsigned *flags(unsigned par) const { return d_flags + (size_t)par * np; }

  void alloc(unsigned nrow, unsigned ncol, bool withV)
  {
    np = ncol / 32u;
    steps = STRAT1_STEPS;
    s0steps = STRAT0_STEPS;
    nchunk = gram_nchunk(nrow, np);
    CH = nrow / nchunk;
    chG = update_rows(nrow, withV ? ncol : 0u, np);
    chV = chG;
    HIP_CALL(hipMalloc(&d_s1, sizeof(uint16_t) * 2u * np * steps));
    HIP_CALL(hipMalloc(&d_s0, STRAT0_TAB_BYTES));
    HIP_CALL(hipMalloc(&d_part, sizeof(double) * 768u * nchunk * np * 2u));
    HIP_CALL(hipMalloc(&d_W, sizeof(double) * 1024u * np * 2u));
    HIP_CALL(hipMalloc(&d_flags, sizeof(unsigned) * np * 2u));
    HIP_CALL(hipMalloc(&d_ctl, sizeof(RunCtl)));
    HIP_CALL(hipHostMalloc(&h_ctl, sizeof(RunCtl), hipHostMallocDefault));
    HIP_CALL(hipMalloc(&d_rec, sizeof(VisitRec) * keys()));
    HIP_CALL(hipMalloc(&d_skip, keys()));
    HIP_CALL(hipMalloc(&d_model, sizeof(uint16_t) * MLP_NPARAM));
    std::vector<uint16_t> s1((size_t)2u * np * steps);
    for (unsigned s = 0u; s < steps; ++s)
      for (unsigned p = 0u; p < np; ++p) {
        s1[((size_t)s * np + p) * 2u] = strat1[s][p][0];
        s1[((size_t)s * np + p) * 2u + 1u] = strat1[s][p][1];
      }
    HIP_CALL(hipMemcpy(d_s1, s1.data(), sizeof(uint16_t) * s1.size(), hipMemcpyHostToDevice));
    std::vector<uint8_t> s0(STRAT0_TAB_BYTES, 0u);
    for (unsigned s = 0u; s < s0steps; ++s)
      for (unsigned p = 0u; p < 16u; ++p) {
        s0[(s * 16u + p) * 2u] = strat0[s][p][0];
        s0[(s * 16u + p) * 2u + 1u] = strat0[s][p][1];
      }
    HIP_CALL(hipMemcpy(d_s0, s0.data(), STRAT0_TAB_BYTES, hipMemcpyHostToDevice));
    // records start as "never visited", no pair skipped, counters zero
    std::vector<VisitRec> r(keys());
    memset(r.data(), 0, sizeof(VisitRec) * r.size());
    HIP_CALL(hipMemcpy(d_rec, r.data(), sizeof(VisitRec) * r.size(), hipMemcpyHostToDevice));
    std::vector<uint8_t> z(keys(), 0u);
    HIP_CALL(hipMemcpy(d_skip, z.data(), keys(), hipMemcpyHostToDevice));
    memset(h_ctl, 0, sizeof(RunCtl));
    HIP_CALL(hipMemcpy(d_ctl, h_ctl, sizeof(RunCtl), hipMemcpyHostToDevice));
    build_quads();
  }


  // A transition s -> s+1 is fused when every pair of step s, with the pair
  // holding its first block's next partner, forms a closed quad.
  void build_quads()
  {
    fuse.assign(steps, 0);
    if (getenv("HJ_NOFUSE")) return;
    std::vector<uint16_t> tab((size_t)steps * np * 2u);
    for (unsigned t = 0u; t < steps; ++t)
      for (unsigned p = 0u; p < np; ++p) {
        tab[((size_t)t * np + p) * 2u] = strat1[t][p][0];
        tab[((size_t)t * np + p) * 2u + 1u] = strat1[t][p][1];
      }
    std::vector<Quad> all;
    if (!closed_quads(tab.data(), steps, np, all, &fuse)) return;
    HIP_CALL(hipMalloc(&d_quads, sizeof(Quad) * all.size()));
    HIP_CALL(hipMemcpy(d_quads, all.data(), sizeof(Quad) * all.size(), hipMemcpyHostToDevice));
    hquads.swap(all);
  }

  // The launch plan of a staged sweep (every transition fused, at least two
  // quads per transition).  Quads [0, h) of a transition form its first
  // half, h = nq/2.
  bool build_stage()
  {
    stage_ok = false;
    if (hquads.empty() || steps < 2u || np < 4u) return false;
    for (unsigned t = 0u; t + 1u < steps; ++t)
      if (!fuse[t]) return false;
    const unsigned nq = np / 2u, h = nq / 2u;
    std::vector<uint16_t> pl((size_t)(steps - 1u) * np);
    std::vector<char> inB(np);
    cross.assign(steps, 0);
    for (unsigned t = 0u; t + 1u < steps; ++t) {
      const Quad *q = hquads.data() + (size_t)t * nq;
      for (unsigned i = 0u; i < nq; ++i) {
        pl[(size_t)t * np + 2u * i] = q[i].n0;
        pl[(size_t)t * np + 2u * i + 1u] = q[i].n1;
      }
      if (t) {
        const Quad *pq = hquads.data() + (size_t)(t - 1u) * nq;
        std::fill(inB.begin(), inB.end(), 0);
        for (unsigned i = h; i < nq; ++i)
          inB[pq[i].n0] = inB[pq[i].n1] = 1;
        for (unsigned i = 0u; i < h; ++i)
          if (inB[q[i].p0] || inB[q[i].p1])
            cross[t] = 1;
      }
    }
    HIP_CALL(hipMalloc(&d_plist, sizeof(uint16_t) * pl.size()));
    HIP_CALL(hipMemcpy(d_plist, pl.data(), sizeof(uint16_t) * pl.size(), hipMemcpyHostToDevice));
    HIP_CALL(hipMalloc(&d_plist_run, sizeof(uint16_t) * pl.size()));
    HIP_CALL(hipMalloc(&d_qord, sizeof(uint16_t) * (size_t)(steps - 1u) * (np / 2u)));
    HIP_CALL(hipMalloc(&d_qact, sizeof(unsigned) * 2u * (size_t)(steps - 1u)));
    stage_ok = true;
    return true;
  }

  void release()
  {
    HIP_CALL(hipFree(d_s1));
    HIP_CALL(hipFree(d_s0));
    HIP_CALL(hipFree(d_part));
    HIP_CALL(hipFree(d_W));
    HIP_CALL(hipFree(d_flags));
    HIP_CALL(hipFree(d_ctl));
    HIP_CALL(hipHostFree(h_ctl));
    HIP_CALL(hipFree(d_rec));
    HIP_CALL(hipFree(d_skip));
    HIP_CALL(hipFree(d_model));
    if (d_quads) HIP_CALL(hipFree(d_quads));
    if (d_plist) HIP_CALL(hipFree(d_plist));
    if (d_plist_run) HIP_CALL(hipFree(d_plist_run));
    if (d_qord) HIP_CALL(hipFree(d_qord));
    if (d_qact) HIP_CALL(hipFree(d_qact));
  }
};

static void check_tables(unsigned ncol)
{
  if (STRAT0_PAIRS != 16u || !STRAT0_STEPS)
    HJ_DIE("init_strats: level-0 strategy must have order 32");
  if (2u * STRAT1_PAIRS != ncol / HYPJACL1_NCOLB || !STRAT1_STEPS)
    HJ_DIE("init_strats: level-1 strategy order differs from ncol/16");
}

// One level-1 step: Gram partials, pair kernels, block update.
static void issue_step(hipStream_t s, const Work &w, unsigned step, bool sort, unsigned nswp,
                       int definite, unsigned nplus, double *dG, unsigned lddG, unsigned nrow,
                       double *dV, unsigned lddV, unsigned ncol, hipEvent_t *ev, bool planned,
                       bool split)
{
  const size_t off = (size_t)step * w.np;
  const uint16_t *pairs = w.d_s1 + off * 2u;
  const uint8_t *skip = planned ? w.d_skip + off : nullptr;
  if (ev) HIP_CALL(hipEventRecord(ev[0], s));
  if (g_fact != HJ_FACT_CHOLESKY) {
    // QR path: the pair kernel reads the pair from G itself (no Gram)
    if (ev) HIP_CALL(hipEventRecord(ev[1], s));
    launch_pair_qr(s, sort, g_fact == HJ_FACT_QR_DRMAC, dG, lddG, nrow, pairs, w.np, w.d_W,
                   w.d_flags, w.d_ctl, w.d_s0, w.s0steps, nswp, definite, nplus);
    if (ev) HIP_CALL(hipEventRecord(ev[2], s));
    launch_update(s, dG, lddG, nrow, w.chG, dV, lddV, ncol, w.chV, pairs, w.np, w.d_W, w.d_flags);
    if (ev) HIP_CALL(hipEventRecord(ev[3], s));
    return;
  }
  launch_gram(s, dG, lddG, pairs, w.np, w.nchunk, w.CH, w.d_part, skip);
  if (split)
    launch_gram_split(s, dG, lddG, pairs, w.np, w.nchunk, w.CH, w.d_part, skip);
  if (ev) HIP_CALL(hipEventRecord(ev[1], s));
  launch_pair(s, sort, pairs, w.np, w.d_part, w.nchunk, w.d_W, w.d_flags, w.d_ctl, w.d_rec + off,
              skip, split, w.d_s0, w.s0steps, nswp, definite, nplus);
  if (ev) HIP_CALL(hipEventRecord(ev[2], s));
  launch_update(s, dG, lddG, nrow, w.chG, dV, lddV, ncol, w.chV, pairs, w.np, w.d_W, w.d_flags);
  if (ev) HIP_CALL(hipEventRecord(ev[3], s));
}

// One level-2 sweep: a standalone Gram where the previous transition was not
// fused (always for step 0), the pair kernels, and for each fused
// transition the update of step s together with the Gram of step s+1.
// Profiling events per step: before Gram, before pair, before the update
// (or the fused update + Gram), after.
static unsigned issue_sweep(hipStream_t s, const Work &w, bool sort, unsigned nswp, int definite,
                            unsigned nplus, double *dG, unsigned lddG, unsigned nrow, double *dV,
                            unsigned lddV, unsigned ncol, hipEvent_t *evs, long *steps_left,
                            bool planned, bool split, std::vector<char> *gram_launched)
{
  bool have_gram = false;
  unsigned issued = 0u;
  for (unsigned st = 0u; st < w.steps; ++st) {
    if (steps_left) {
      if (*steps_left <= 0) return issued;
      --*steps_left;
    }
    ++issued;
    if (g_fact != HJ_FACT_CHOLESKY) {
      if (gram_launched) (*gram_launched)[st] = 0;
      issue_step(s, w, st, sort, nswp, definite, nplus, dG, lddG, nrow, dV, lddV, ncol,
                 evs ? evs + (size_t)4u * st : nullptr, false, false);
      continue;
    }
    if (gram_launched) (*gram_launched)[st] = have_gram ? 0 : 1;
    hipEvent_t *ev = evs ? evs + (size_t)4u * st : nullptr;
    const size_t off = (size_t)st * w.np;
    const uint16_t *pairs = w.d_s1 + off * 2u;
    const uint8_t *plan = planned ? w.d_skip + off : nullptr;
    if (ev) HIP_CALL(hipEventRecord(ev[0], s));
    if (!have_gram) {
      launch_gram(s, dG, lddG, pairs, w.np, w.nchunk, w.CH, w.d_part, plan);
      if (split)
        launch_gram_split(s, dG, lddG, pairs, w.np, w.nchunk, w.CH, w.d_part, plan);
    }
    if (ev) HIP_CALL(hipEventRecord(ev[1], s));
    // a split-planned visit holds a split Gram after a standalone Gram
    // launch, and after a fused transition that forms split Grams
    launch_pair(s, sort, pairs, w.np, w.d_part, w.nchunk, w.d_W, w.d_flags, w.d_ctl, w.d_rec + off,
                plan, split && (!have_gram || g_fused_split), w.d_s0, w.s0steps, nswp, definite,
                nplus);
    if (ev) HIP_CALL(hipEventRecord(ev[2], s));
    const bool fused = (st + 1u < w.steps) && w.fuse[st] &&
                       !(steps_left && *steps_left <= 0);
    if (fused) {
      launch_upgram(s, dG, lddG, pairs, w.d_W, w.d_flags, w.d_quads + (size_t)st * (w.np / 2u),
                    w.np / 2u, w.nchunk, w.CH, w.d_part,
                    planned ? w.d_skip + off + w.np : nullptr, split && g_fused_split);
      if (dV)
        launch_update(s, dG, lddG, 0u, w.chG, dV, lddV, ncol, w.chV, pairs, w.np, w.d_W, w.d_flags);
    }
    else
      launch_update(s, dG, lddG, nrow, w.chG, dV, lddV, ncol, w.chV, pairs, w.np, w.d_W, w.d_flags);
    if (ev) HIP_CALL(hipEventRecord(ev[3], s));
    have_gram = fused;
  }
  return issued;
}

// One level-2 sweep as staged launches.  With A_t / B_t the first / second
// half of the quads of transition t -> t+1 and P_A(t+1) / P_B(t+1) the pairs
// of step t+1 they form:
//   Gram of step 0, visits of step 0;
//   for t = 0 .. S-2:
//     X_t: fused transition of A_t   + visits of P_B(t) (t >= 1)
//     Y_t: fused transition of B_t   + visits of P_A(t+1)
//   visits of P_B(S-1), update of step S-1.
// A launch's visits need only earlier launches (the Gram partials of their
// pairs); a fused item needs the W of its two pairs: those of A_t come from
// P_A(t), visited in Y_{t-1}, unless the quads of A_t take a pair of P_B(t)
// (cross[t]): then X_t runs as two launches, visits first.  Step parity
// buffers keep a launch's visits (step t+1) and fused items (step t) apart.
// Profiling events: one before the first launch and one after each launch.
// Fused items per item workgroup: two where every item runs (the item
// workgroups and the visits' then all start at once, one round of at most
// two per CU: plain N = 8192 4.99 -> 4.88 s; filling every free slot
// instead, some workgroups taking two items, is slower: 140 against 124 us
// per launch).  With the learned skip the launch holds one workgroup per
// item, and the items with work, ordered first on the device, are taken one
// per workgroup while they fit the slots beside the visits, else two per
// workgroup (a static pairing of partly skipped items in index order is
// uneven: 4.45 s against 4.19 s with one each; one each where every item
// runs leaves a second round: 152 against 125 us per launch).
static unsigned cu_count()
{
  static int cached_dev = -1;
  static unsigned cached = 0u;
  int dev = 0;
  HIP_CALL(hipGetDevice(&dev));
  if (dev != cached_dev) {
    int n = 0;
    HIP_CALL(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    cached = n > 0 ? (unsigned)n : 1u;
    cached_dev = dev;
  }
  return cached;
}

static unsigned issue_sweep_staged(hipStream_t s, const Work &w, bool sort, unsigned nswp,
                                   int definite, unsigned nplus, double *dG, unsigned lddG,
                                   unsigned nrow, double *dV, unsigned lddV, unsigned ncol,
                                   hipEvent_t *ev, bool planned, bool split, bool skipping)
{
  const unsigned np = w.np, nq = np / 2u, h = nq / 2u, S = w.steps;
  const uint8_t *plan = planned ? w.d_skip : nullptr;
  const unsigned ncu = cu_count();
  unsigned nev = 0u;
  auto mark = [&]() {
    if (ev) HIP_CALL(hipEventRecord(ev[nev], s));
    ++nev;
  };
  StageArgs base{};
  base.G = dG;
  base.ldG = lddG;
  base.nchunk = w.nchunk;
  base.CH = w.CH;
  base.nw = gram_waves(w.CH);
  base.ctl = w.d_ctl;
  base.strat0 = w.d_s0;
  base.s0steps = w.s0steps;
  base.nswp = nswp;
  base.definite = definite;
  base.nplus = nplus;
  base.fsplit = (split && g_fused_split) ? 1 : 0;
  auto visits = [&](StageArgs &x, unsigned st, const uint16_t *list, unsigned n) {
    x.npl = n;
    x.npw = (n + 1u) / 2u;
    x.plist = list;
    x.ppairs = w.d_s1 + (size_t)st * np * 2u;
    x.ppart = w.part(st & 1u);
    x.pW = w.W(st & 1u);
    x.pflags = w.flags(st & 1u);
    x.prec = w.d_rec + (size_t)st * np;
    x.pplan = plan ? plan + (size_t)st * np : nullptr;
  };
  // V by the W of step st (all its visits done: those of P_A(st) in Y_{st-1},
  // of P_B(st) in X_st), in Y_st
  auto vupdate = [&](StageArgs &x, unsigned st) {
    if (!dV)
      return;
    x.V = dV;
    x.ldV = lddV;
    x.nrowV = ncol;
    x.chV = w.chV;
    x.nchV = (ncol + w.chV - 1u) / w.chV;
    x.nvitems = np * x.nchV;
    x.nvw = (x.nvitems + 3u) / 4u;
    x.vpairs = w.d_s1 + (size_t)st * np * 2u;
    x.vW = w.W(st & 1u);
    x.vflags = w.flags(st & 1u);
  };
  auto fused = [&](StageArgs &x, unsigned st, unsigned q0, unsigned n) {
    x.upairs = w.d_s1 + (size_t)st * np * 2u;
    x.uW = w.W(st & 1u);
    x.uflags = w.flags(st & 1u);
    x.quads = skipping ? w.d_quads + (size_t)st * nq : w.d_quads + (size_t)st * nq + q0;
    x.qord = skipping ? w.d_qord + (size_t)st * nq + q0 : nullptr;
    x.nquads = n;
    x.niw = skipping ? n * w.nchunk : (n * w.nchunk + 1u) / 2u;
    // with the learned skip the items with work come first and the item
    // workgroups take one each while they fit the slots beside the visits
    // (two resident workgroups per CU), else two each
    x.qact = skipping ? w.d_qact + 2u * st + (q0 ? 1u : 0u) : nullptr;
    x.islots = (2u * ncu > x.npw) ? 2u * ncu - x.npw : 1u;
    x.upart = w.part((st + 1u) & 1u);
    x.uplan = plan ? plan + (size_t)(st + 1u) * np : nullptr;
  };
  auto stage = [&](StageArgs &x) { launch_stage(s, sort, x); };
  // with the learned skip, each half-list's running visits first
  const uint16_t *const plist = skipping ? w.d_plist_run : w.d_plist;
  mark();
  if (skipping) {
    launch_order_visits(s, w.d_plist, w.d_plist_run, w.d_skip, np, S, 2u * h);
    launch_order_quads(s, w.d_quads, w.d_qord, w.d_qact, w.d_skip, np, S, h);
  }
  launch_gram(s, dG, lddG, w.d_s1, np, w.nchunk, w.CH, w.part(0u), plan);
  if (split)
    launch_gram_split(s, dG, lddG, w.d_s1, np, w.nchunk, w.CH, w.part(0u), plan);
  launch_pair(s, sort, w.d_s1, np, w.part(0u), w.nchunk, w.W(0u), w.flags(0u), w.d_ctl, w.d_rec,
              plan, split, w.d_s0, w.s0steps, nswp, definite, nplus);
  mark();
  for (unsigned t = 0u; t + 1u < S; ++t) {
    StageArgs x = base;
    if (t) {
      visits(x, t, plist + (size_t)(t - 1u) * np + 2u * h, np - 2u * h);
      if (w.cross[t]) {
        stage(x);
        mark();
        x = base;
      }
    }
    fused(x, t, 0u, h);
    stage(x);
    mark();
    StageArgs y = base;
    visits(y, t + 1u, plist + (size_t)t * np, 2u * h);
    fused(y, t, h, nq - h);
    vupdate(y, t);
    stage(y);
    mark();
  }
  StageArgs z = base;
  visits(z, S - 1u, plist + (size_t)(S - 2u) * np + 2u * h, np - 2u * h);
  stage(z);
  launch_update(s, dG, lddG, nrow, w.chG, dV, lddV, ncol, w.chV, w.d_s1 + (size_t)(S - 1u) * np * 2u,
                np, w.W((S - 1u) & 1u), w.flags((S - 1u) & 1u));
  mark();
  return nev;
}

// Sweeps + initD on device data (the body of 1/HypJacL2.cu:104-178).
//
// With the learned pair-skip form, every sweep after the first starts with
// the predictor, which marks the visits whose model probability of a big
// rotation is below the threshold; those visits cost nothing.  A sweep with
// skipped visits never ends the run: when it finds no big rotation, the next
// sweep runs unskipped (the certification sweep), and only an unskipped sweep
// without a big rotation ends it -- the reference's test (1/HypJacL2.cu:167).
static void level2_core(unsigned routine, unsigned nrow, unsigned ncol, unsigned nplus, double *dG,
                        unsigned lddG, double *dV, unsigned lddV, double *dD, unsigned *glbSwp,
                        unsigned long long *glb_s, unsigned long long *glb_b, hipStream_t caller)
{
  // Work runs on a stream of its own (capturable even when the caller uses
  // the null stream), ordered after the caller's prior work.
  hipStream_t s = nullptr;
  HIP_CALL(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  {
    hipEvent_t e;
    HIP_CALL(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIP_CALL(hipEventRecord(e, caller));
    HIP_CALL(hipStreamWaitEvent(s, e, 0));
    HIP_CALL(hipEventDestroy(e));
  }
  const bool cdsort = (routine & HYPJAC_CDSORT_1) != 0u;
  const bool full_svd = (routine & HYPJAC_FULL_SVD) != 0u;
  const bool blk_ori = (routine & HYPJAC_BLK_ORI) != 0u;
  const unsigned nswp = blk_ori ? 1u : HYPJAC_NSWEEP;
  check_tables(ncol);

  Work w;
  w.alloc(nrow, ncol, full_svd);
  const double step_bytes = 512.0 * (ncol / 32u) * ((double)nrow + (full_svd ? ncol : 0u));
  const bool staged = g_stage && (g_stage == 2 || step_bytes >= STAGE_MIN_BYTES) &&
                      g_step_limit <= 0 && g_fact == HJ_FACT_CHOLESKY && w.build_stage();
  g_last_stage = staged ? 1 : 0;
  if (full_svd)
    launch_initV(s, dV, 0u, ncol, lddV);
  const int definite = nplus ? ((nplus < ncol) ? 0 : +1) : -1;
  double *const V = full_svd ? dV : nullptr;
  // the forms act on the Gram of the Cholesky path; the QR path runs plain
  const bool qr = g_fact != HJ_FACT_CHOLESKY;
  const bool learned = !qr && (g_form & HJ_FORM_LEARNED_SKIP) && g_model_ready &&
                       (size_t)(ncol / 32u) * STRAT1_STEPS >= g_skip_min_visits;
  const bool split = !qr && (g_form & HJ_FORM_SPLIT_GRAM) != 0u;
  const bool planned = learned || split;
  if (learned)
    HIP_CALL(hipMemcpy(w.d_model, g_model_bf16, sizeof(uint16_t) * MLP_NPARAM, hipMemcpyHostToDevice));
  const float log2n1 = (float)std::log2((double)(ncol / 16u));
  const float max_rot = (float)w.keys() * 496.0f;

  g_stats = Stats();
  *glb_s = 0ull;
  *glb_b = 0ull;
  {
    // the forms' policy, once per run (k_sweep_begin applies it every sweep)
    RunCtl &c = *w.h_ctl;
    c.learned = learned ? 1u : 0u;
    c.split_form = split ? 1u : 0u;
    c.threshold = g_skip_threshold;
    c.max_skips = g_skip_max;
    c.tau = g_split_tau;
    c.min_frac = g_skip_min_frac * ((definite == 0) ? g_skip_floor_mixed : 1.0f);
    c.min_skip_frac = g_skip_min_propose;
    c.max_rot = max_rot;
    HIP_CALL(hipMemcpyAsync(w.d_ctl, w.h_ctl, sizeof(RunCtl), hipMemcpyHostToDevice, s));
    HIP_CALL(hipStreamSynchronize(s));
  }
  const bool eager = g_profile || (g_step_limit > 0) || getenv("HJ_NOGRAPH");
  hipGraphExec_t gexec = nullptr;
  // graph path: events around each sweep's graph launch, and the launches
  // of the sweep's dominant kind it holds (staged launches; else the steps'
  // update-class launches)
  hipEvent_t gev[2] = {nullptr, nullptr};
  unsigned long long graph_launches = 0ull;
  std::vector<hipEvent_t> evs;
  if (g_profile) {
    evs.resize((size_t)4u * w.steps + 8u);
    for (auto &e : evs) HIP_CALL(hipEventCreate(&e));
  }
  auto sweep_head = [&](hipStream_t st) {
    launch_sweep_begin(st, w.d_ctl);
    if (planned)
      launch_plan(st, w.d_rec, (unsigned)w.keys(), w.d_s1, nplus, log2n1, max_rot, w.d_ctl, w.d_model,
                  learned, w.d_skip);
  };
  if (!eager) {
    // one sweep as a graph: start-of-sweep kernel (counters, no memset
    // node), the predictor, all steps, the control block to pinned memory
    hipGraph_t graph;
    HIP_CALL(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    sweep_head(s);
    if (staged)
      graph_launches = issue_sweep_staged(s, w, cdsort, nswp, definite, nplus, dG, lddG, nrow, V,
                                          lddV, ncol, nullptr, planned, split, learned) - 2u;
    else
      graph_launches = issue_sweep(s, w, cdsort, nswp, definite, nplus, dG, lddG, nrow, V, lddV,
                                   ncol, nullptr, nullptr, planned, split, nullptr);
    HIP_CALL(hipMemcpyAsync(w.h_ctl, w.d_ctl, sizeof(RunCtl), hipMemcpyDeviceToHost, s));
    HIP_CALL(hipStreamEndCapture(s, &graph));
    HIP_CALL(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    HIP_CALL(hipGraphDestroy(graph));
    HIP_CALL(hipEventCreate(&gev[0]));
    HIP_CALL(hipEventCreate(&gev[1]));
  }

  long long swp_tim = 0;
  stopwatch_reset(swp_tim);
  unsigned blk_swp = 0u;
  long steps_done = 0;
  bool stop = false;
  unsigned issued = 0u;
  std::vector<char> gram_launched(w.steps, 0);
  unsigned nstage_ev = 0u;
  g_stats.skipped = 0ull;
  while (blk_swp < g_max_sweeps) {
    if (eager) {
      sweep_head(s);
      long left = (g_step_limit > 0) ? (g_step_limit - steps_done) : (long)w.steps;
      if (staged) {
        nstage_ev = issue_sweep_staged(s, w, cdsort, nswp, definite, nplus, dG, lddG, nrow, V,
                                       lddV, ncol, g_profile ? evs.data() : nullptr, planned,
                                       split, learned);
        issued = w.steps;
      }
      else
        issued = issue_sweep(s, w, cdsort, nswp, definite, nplus, dG, lddG, nrow, V, lddV, ncol,
                             g_profile ? evs.data() : nullptr, &left, planned, split,
                             &gram_launched);
      steps_done += issued;
      if (g_step_limit > 0 && steps_done >= g_step_limit)
        stop = true;
      HIP_CALL(hipMemcpyAsync(w.h_ctl, w.d_ctl, sizeof(RunCtl), hipMemcpyDeviceToHost, s));
    }
    else {
      HIP_CALL(hipEventRecord(gev[0], s));
      HIP_CALL(hipGraphLaunch(gexec, s));
      HIP_CALL(hipEventRecord(gev[1], s));
    }
    HIP_CALL(hipStreamSynchronize(s));
    if (!eager) {
      float ms = 0.0f;
      HIP_CALL(hipEventElapsedTime(&ms, gev[0], gev[1]));
      g_stats.kms[5] += ms;
      g_stats.kcnt[5] += graph_launches;
    }
    if (g_profile && staged) {
      // kind 4: every launch after step 0's Gram and visits
      for (unsigned e = 1u; e + 1u < nstage_ev; ++e) {
        float ms = 0.0f;
        HIP_CALL(hipEventElapsedTime(&ms, evs[e], evs[e + 1u]));
        g_stats.kms[4] += ms;
        g_stats.kcnt[4] += 1ull;
      }
    }
    else if (g_profile)
      for (unsigned st = 0u; st < issued; ++st)
        for (int k = 0; k < 3; ++k) {
          if (k == 0 && !gram_launched[st]) continue;
          float ms = 0.0f;
          HIP_CALL(hipEventElapsedTime(&ms, evs[(size_t)4u * st + k], evs[(size_t)4u * st + k + 1]));
          g_stats.kms[k] += ms;
          g_stats.kcnt[k] += 1ull;
        }
    ++blk_swp;
    const unsigned long long cvg = w.h_ctl->cvg;
    const unsigned skipped = w.h_ctl->skipped;
    const unsigned nsplit = w.h_ctl->split;
    const unsigned cvg_s = static_cast<unsigned>(cvg);
    const unsigned cvg_b = static_cast<unsigned>(cvg >> 32u);
    *glb_s += cvg_s;
    *glb_b += cvg_b;
    const double tim_s = stopwatch_lap(swp_tim) * TS2S;
    if (blk_swp <= HJ_MAX_SWEEP_CAP) {
      g_stats.rot_s[blk_swp - 1u] = cvg_s;
      g_stats.rot_b[blk_swp - 1u] = cvg_b;
      g_stats.sec[blk_swp - 1u] = tim_s;
      g_stats.skip[blk_swp - 1u] = skipped;
      g_stats.split[blk_swp - 1u] = nsplit;
    }
    g_stats.splits += nsplit;
    g_stats.nswp = blk_swp;
    g_stats.skipped += skipped;
    g_stats.visits += w.keys() - skipped;
    if (g_trace_buf && blk_swp <= g_trace_sweeps) {
      HIP_CALL(hipMemcpy(static_cast<char *>(g_trace_buf) + (size_t)(blk_swp - 1u) * w.keys() * sizeof(VisitRec),
                         w.d_rec, w.keys() * sizeof(VisitRec), hipMemcpyDeviceToHost));
    }
    if (!quiet()) {
      (void)fprintf(stdout, "BLK_SWP(%2u), ROT_S(%10u), ROT_B(%10u), TIME(%#12.6f s)\n", blk_swp,
                    cvg_s, cvg_b, tim_s);
      (void)fflush(stdout);
    }
    if (stop)
      break;
    // no big rotation ends the run only in an all-FP64, unskipped sweep;
    // otherwise the next sweep certifies (k_sweep_begin switches the forms
    // off for it)
    if (!cvg_b && !skipped && !nsplit)
      break;
  }
  if (gexec) HIP_CALL(hipGraphExecDestroy(gexec));
  for (hipEvent_t e : gev)
    if (e) HIP_CALL(hipEventDestroy(e));
  for (auto &e : evs) HIP_CALL(hipEventDestroy(e));

  *glbSwp = blk_swp;
  launch_initD(s, dG, dD, 0u, nrow, ncol, nplus, lddG);
  HIP_CALL(hipStreamSynchronize(s));
  HIP_CALL(hipStreamDestroy(s));
  w.release();
}

// The level-2 run with blocks of 32 columns (hj_wide.hip): the strategy
// families loaded by init_strats at the orders this width needs (64 for
// level 0, ncol/32 for level 1, from the library's generators), then per
// sweep every step's Gram partials, visits and block updates (G, and V with
// ALG & 2), and the reference's stopping test (1/HypJacL2.cu:167).  With
// the learned pair-skip form the sweep starts with the plan of its visits
// (the 16-column form's model and run-level policy, the visit records in
// its units) and ends the run only when unskipped; the split Gram and the
// QR factorization act at width 16 only.
static void level2_wide(unsigned routine, unsigned nrow, unsigned ncol, unsigned nplus, double *dG,
                        unsigned lddG, double *dV, unsigned lddV, double *dD, unsigned *glbSwp,
                        unsigned long long *glb_s, unsigned long long *glb_b, hipStream_t caller)
{
  hipStream_t s = nullptr;
  HIP_CALL(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  {
    hipEvent_t e;
    HIP_CALL(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIP_CALL(hipEventRecord(e, caller));
    HIP_CALL(hipStreamWaitEvent(s, e, 0));
    HIP_CALL(hipEventDestroy(e));
  }
  const bool cdsort = (routine & HYPJAC_CDSORT_1) != 0u;
  const bool full_svd = (routine & HYPJAC_FULL_SVD) != 0u;
  const bool blk_ori = (routine & HYPJAC_BLK_ORI) != 0u;
  const unsigned nswp = blk_ori ? 1u : HYPJAC_NSWEEP;
  double *const V = full_svd ? dV : nullptr;
  Strategy t0, t1;
  const char *const f0 = (STRAT0 >= 1u && STRAT0 <= 6u) ? kFamilies[STRAT0 - 1u] : nullptr;
  const char *const f1 = (STRAT1 >= 1u && STRAT1 <= 6u) ? kFamilies[STRAT1 - 1u] : nullptr;
  if (!f0 || !generate(f0, 64u, t0) || t0.steps > 64u)
    HJ_DIE("32-column blocks: no level-0 strategy of order 64 in the loaded family");
  if (!f1 || !generate(f1, ncol / 32u, t1))
    HJ_DIE("32-column blocks: no level-1 strategy of order ncol/32 in the loaded family");
  const unsigned np = ncol / 64u, steps = t1.steps;
  const unsigned nchunk = gram_nchunk(nrow, np), CH = nrow / nchunk;
  std::vector<uint8_t> s0(t0.tab.begin(), t0.tab.end());
  double *d_part = nullptr, *d_W = nullptr;
  unsigned *d_flags = nullptr;
  RunCtl *d_ctl = nullptr, *h_ctl = nullptr;
  VisitRec *d_rec = nullptr;
  uint8_t *d_plan = nullptr;
  uint16_t *d_model = nullptr;
  uint8_t *d_s0 = nullptr;
  uint16_t *d_s1 = nullptr;
  HIP_CALL(hipMalloc(&d_part, sizeof(double) * (size_t)np * nchunk * wide_part_doubles()));
  HIP_CALL(hipMalloc(&d_W, sizeof(double) * (size_t)np * 4096u));
  HIP_CALL(hipMalloc(&d_flags, sizeof(unsigned) * np));
  // the learned pair-skip at this width: the model and run-level policy of
  // the 16-column form, the records in its units
  const size_t nkeys = (size_t)np * steps;
  const bool learned = (g_form & HJ_FORM_LEARNED_SKIP) && g_model_ready &&
                       nkeys >= g_skip_min_visits;
  HIP_CALL(hipMalloc(&d_ctl, sizeof(RunCtl)));
  HIP_CALL(hipHostMalloc(&h_ctl, sizeof(RunCtl), hipHostMallocDefault));
  memset(h_ctl, 0, sizeof(RunCtl));
  const float max_rot = (float)nkeys * 2016.0f;
  if (learned) {
    HIP_CALL(hipMalloc(&d_rec, sizeof(VisitRec) * nkeys));
    HIP_CALL(hipMemsetAsync(d_rec, 0, sizeof(VisitRec) * nkeys, s));
    HIP_CALL(hipMalloc(&d_plan, nkeys));
    HIP_CALL(hipMalloc(&d_model, sizeof(uint16_t) * MLP_NPARAM));
    HIP_CALL(hipMemcpyAsync(d_model, g_model_bf16, sizeof(uint16_t) * MLP_NPARAM,
                            hipMemcpyHostToDevice, s));
    h_ctl->learned = 1u;
    h_ctl->threshold = g_skip_threshold;
    h_ctl->max_skips = g_skip_max;
    h_ctl->tau = g_split_tau;
    h_ctl->min_frac = g_skip_min_frac * ((nplus && nplus < ncol) ? g_skip_floor_mixed_w32 : 1.0f);
    h_ctl->min_skip_frac = g_skip_min_propose;
  }
  h_ctl->max_rot = max_rot;
  HIP_CALL(hipMemcpyAsync(d_ctl, h_ctl, sizeof(RunCtl), hipMemcpyHostToDevice, s));
  HIP_CALL(hipStreamSynchronize(s));
  const float log2n1 = (float)std::log2((double)(ncol / 32u));
  HIP_CALL(hipMalloc(&d_s0, s0.size()));
  HIP_CALL(hipMalloc(&d_s1, sizeof(uint16_t) * t1.tab.size()));
  HIP_CALL(hipMemcpyAsync(d_s0, s0.data(), s0.size(), hipMemcpyHostToDevice, s));
  HIP_CALL(hipMemcpyAsync(d_s1, t1.tab.data(), sizeof(uint16_t) * t1.tab.size(),
                          hipMemcpyHostToDevice, s));
  const int definite = nplus ? ((nplus < ncol) ? 0 : +1) : -1;
  if (V)
    launch_initV(s, V, 0u, ncol, lddV);
  std::vector<hipEvent_t> evs;
  if (g_profile) {
    evs.resize((size_t)4u * steps);
    for (auto &e : evs) HIP_CALL(hipEventCreate(&e));
  }
  g_stats = Stats();
  *glb_s = 0ull;
  *glb_b = 0ull;
  unsigned blk_swp = 0u;
  long long swp_tim = 0;
  stopwatch_reset(swp_tim);
  // one sweep as a graph (start-of-sweep kernel, plan, every step, the
  // control block to pinned memory), launched once a sweep; with profiling
  // the launches go one by one with events around them
  hipGraphExec_t gexec = nullptr;
  auto sweep = [&]() {
    // counters, the sweep's switches (k_sweep_begin), the plan of its visits
    launch_sweep_begin(s, d_ctl);
    if (learned)
      launch_plan(s, d_rec, (unsigned)nkeys, d_s1, nplus, log2n1, max_rot, d_ctl, d_model, true,
                  d_plan, 32u);
    for (unsigned st = 0u; st < steps; ++st) {
      const uint16_t *const pairs = d_s1 + (size_t)st * np * 2u;
      const uint8_t *const plan = learned ? d_plan + (size_t)st * np : nullptr;
      hipEvent_t *const ev = g_profile ? evs.data() + (size_t)4u * st : nullptr;
      if (ev) HIP_CALL(hipEventRecord(ev[0], s));
      launch_gram_w(s, dG, lddG, pairs, np, nchunk, CH, d_part, plan);
      if (ev) HIP_CALL(hipEventRecord(ev[1], s));
      launch_pair_w(s, cdsort, pairs, np, d_part, nchunk, d_W, d_flags, d_ctl, d_s0, t0.steps,
                    nswp, definite, nplus, plan, learned ? d_rec + (size_t)st * np : nullptr);
      if (ev) HIP_CALL(hipEventRecord(ev[2], s));
      launch_update_w(s, dG, lddG, nrow, pairs, np, d_W, d_flags);
      if (V)
        launch_update_w(s, V, lddV, ncol, pairs, np, d_W, d_flags);
      if (ev) HIP_CALL(hipEventRecord(ev[3], s));
    }
    HIP_CALL(hipMemcpyAsync(h_ctl, d_ctl, sizeof(RunCtl), hipMemcpyDeviceToHost, s));
  };
  if (!g_profile && !getenv("HJ_NOGRAPH")) {
    hipGraph_t graph;
    HIP_CALL(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    sweep();
    HIP_CALL(hipStreamEndCapture(s, &graph));
    HIP_CALL(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    HIP_CALL(hipGraphDestroy(graph));
  }
  while (blk_swp < g_max_sweeps) {
    if (gexec)
      HIP_CALL(hipGraphLaunch(gexec, s));
    else
      sweep();
    HIP_CALL(hipStreamSynchronize(s));
    if (g_profile)
      for (unsigned st = 0u; st < steps; ++st)
        for (int k = 0; k < 3; ++k) {
          float ms = 0.0f;
          HIP_CALL(hipEventElapsedTime(&ms, evs[(size_t)4u * st + k], evs[(size_t)4u * st + k + 1]));
          g_stats.kms[k] += ms;
          g_stats.kcnt[k] += 1ull;
        }
    ++blk_swp;
    const unsigned cvg_s = static_cast<unsigned>(h_ctl->cvg);
    const unsigned cvg_b = static_cast<unsigned>(h_ctl->cvg >> 32u);
    const unsigned skipped = h_ctl->skipped;
    *glb_s += cvg_s;
    *glb_b += cvg_b;
    const double tim_s = stopwatch_lap(swp_tim) * TS2S;
    if (blk_swp <= HJ_MAX_SWEEP_CAP) {
      g_stats.rot_s[blk_swp - 1u] = cvg_s;
      g_stats.rot_b[blk_swp - 1u] = cvg_b;
      g_stats.sec[blk_swp - 1u] = tim_s;
      g_stats.skip[blk_swp - 1u] = skipped;
    }
    g_stats.nswp = blk_swp;
    g_stats.skipped += skipped;
    g_stats.visits += nkeys - skipped;
    if (!quiet()) {
      (void)fprintf(stdout, "BLK_SWP(%2u), ROT_S(%10u), ROT_B(%10u), TIME(%#12.6f s)\n", blk_swp,
                    cvg_s, cvg_b, tim_s);
      (void)fflush(stdout);
    }
    // no big rotation ends the run only in an unskipped sweep; otherwise
    // the next sweep certifies (k_sweep_begin turns skipping off for it)
    if (!cvg_b && !skipped)
      break;
  }
  if (gexec) HIP_CALL(hipGraphExecDestroy(gexec));
  for (auto &e : evs) HIP_CALL(hipEventDestroy(e));
  *glbSwp = blk_swp;
  launch_initD(s, dG, dD, 0u, nrow, ncol, nplus, lddG);
  HIP_CALL(hipStreamSynchronize(s));
  HIP_CALL(hipStreamDestroy(s));
  HIP_CALL(hipFree(d_part));
  HIP_CALL(hipFree(d_W));
  HIP_CALL(hipFree(d_flags));
  HIP_CALL(hipFree(d_ctl));
  HIP_CALL(hipHostFree(h_ctl));
  if (d_rec) HIP_CALL(hipFree(d_rec));
  if (d_plan) HIP_CALL(hipFree(d_plan));
  if (d_model) HIP_CALL(hipFree(d_model));
  HIP_CALL(hipFree(d_s0));
  HIP_CALL(hipFree(d_s1));
}

// whether the families loaded by init_strats have the tables the 32-column
// blocks need (orders 64 and ncol/32)
static bool wide_tables_ok(unsigned ncol)
{
  if (STRAT0 < 1u || STRAT0 > 6u || STRAT1 < 1u || STRAT1 > 6u)
    return false;
  return can_generate(kFamilies[STRAT0 - 1u], 64u) && family_steps(kFamilies[STRAT0 - 1u], 64u) <= 64u &&
         can_generate(kFamilies[STRAT1 - 1u], ncol / 32u);
}

// the level-2 run at the block width in force
static void level2_run(unsigned routine, unsigned nrow, unsigned ncol, unsigned nplus, double *dG,
                       unsigned lddG, double *dV, unsigned lddV, double *dD, unsigned *glbSwp,
                       unsigned long long *glb_s, unsigned long long *glb_b, hipStream_t caller)
{
  if (g_width == 32u)
    level2_wide(routine, nrow, ncol, nplus, dG, lddG, dV, lddV, dD, glbSwp, glb_s, glb_b, caller);
  else
    level2_core(routine, nrow, ncol, nplus, dG, lddG, dV, lddV, dD, glbSwp, glb_s, glb_b, caller);
}

// argument checks of 1/HypJacL2.cu:39-77
// (with blocks of `width` columns ncol must be a multiple of 2 width)
static int check_args(unsigned routine, unsigned nrow, unsigned ncol, unsigned nplus,
                      const void *G, unsigned ldG, const void *V, unsigned ldV, const void *D,
                      const void *glbSwp, const void *glb_s, const void *glb_b,
                      unsigned width = 16u)
{
  if (routine >= 16u) return -1;
  const bool full_svd = (routine & HYPJAC_FULL_SVD) != 0u;
  if (!nrow || (nrow % 64u)) return -2;
  if (!ncol || (ncol > nrow) || (ncol % (2u * width))) return -3;
  if (nplus > ncol) return -4;
  if (!G) return -5;
  if (ldG < nrow) return -6;
  if (full_svd) {
    if (!V) return -7;
    if (ldV < ncol) return -8;
  }
  if (!D) return -9;
  if (!glbSwp) return -10;
  if (!glb_s) return -11;
  if (!glb_b) return -12;
  return 0;
}

// Thin level-1 layer state (share/device_code_globals.hpp:4-20).
struct Thin {
  double *G = nullptr, *V = nullptr;
  unsigned long long *cvg = nullptr;
  unsigned nRow = 0u, nCol = 0u, ldG = 0u, ldV = 0u, nSwp = 0u;
  Work w;
  bool ready = false;
};
static Thin g_thin;

// The rotation counter of the thin layer is the caller's 8-byte device
// word: the step's count (added by the pair kernels to the workspace's
// control block) is added to it after the step.
__global__ void k_thin_accum(hj::RunCtl *ctl, unsigned long long *cvg)
{
  if (threadIdx.x == 0u) {
    *cvg += ctl->cvg;
    ctl->cvg = 0ull;
  }
}

static void thin_step(unsigned step, bool sort, bool withV, int definite, unsigned nplus)
{
  if (!g_thin.ready) HJ_DIE("initSymbols not called");
  if (step >= g_thin.w.steps) HJ_DIE("step out of range");
  issue_step(nullptr, g_thin.w, step, sort, g_thin.nSwp, definite, nplus, g_thin.G, g_thin.ldG,
             g_thin.nRow, withV ? g_thin.V : nullptr, g_thin.ldV, g_thin.nCol, nullptr, false, false);
  hipLaunchKernelGGL(k_thin_accum, dim3(1), dim3(64), 0, nullptr, g_thin.w.d_ctl, g_thin.cvg);
  HIP_CALL(hipGetLastError());
}

}  // namespace hj

using namespace hj;

extern "C" int hypJacL2(const unsigned routine, const unsigned nrow, const unsigned ncol,
                        const unsigned nplus, double *const hG, const unsigned ldhG,
                        double *const hV, const unsigned ldhV, double *const hD,
                        unsigned *const glbSwp, unsigned long long *const glb_s,
                        unsigned long long *const glb_b, double *const timing)
{
  long long timers[4] = {0ll, 0ll, 0ll, 0ll};
  stopwatch_reset(timers[0]);
  const int rc = check_args(routine, nrow, ncol, nplus, hG, ldhG, hV, ldhV, hD, glbSwp, glb_s, glb_b,
                            g_width);
  if (rc)
    return rc;
  if (g_width == 32u && !wide_tables_ok(ncol))
    return -3;
  stopwatch_reset(timers[3]);
  const bool full_svd = (routine & HYPJAC_FULL_SVD) != 0u;
  const unsigned lddG = nrow, lddV = full_svd ? ncol : 0u;
  double *dG = nullptr, *dV = nullptr, *dD = nullptr;
  HIP_CALL(hipMalloc(&dG, sizeof(double) * (size_t)lddG * ncol));
  if (full_svd)
    HIP_CALL(hipMalloc(&dV, sizeof(double) * (size_t)lddV * ncol));
  HIP_CALL(hipMalloc(&dD, sizeof(double) * ncol));
  hipStream_t s = nullptr;
  HIP_CALL(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  HIP_CALL(hipMemcpy2DAsync(dG, lddG * sizeof(double), hG, ldhG * sizeof(double),
                            nrow * sizeof(double), ncol, hipMemcpyHostToDevice, s));
  HIP_CALL(hipStreamSynchronize(s));
  timers[1] = stopwatch_lap(timers[3]);

  level2_run(routine, nrow, ncol, nplus, dG, lddG, dV, lddV, dD, glbSwp, glb_s, glb_b, s);
  timers[2] = stopwatch_lap(timers[3]);

  HIP_CALL(hipMemcpy2DAsync(hG, ldhG * sizeof(double), dG, lddG * sizeof(double),
                            nrow * sizeof(double), ncol, hipMemcpyDeviceToHost, s));
  if (full_svd)
    HIP_CALL(hipMemcpy2DAsync(hV, ldhV * sizeof(double), dV, lddV * sizeof(double),
                              ncol * sizeof(double), ncol, hipMemcpyDeviceToHost, s));
  HIP_CALL(hipMemcpyAsync(hD, dD, ncol * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CALL(hipStreamSynchronize(s));
  HIP_CALL(hipStreamDestroy(s));
  HIP_CALL(hipFree(dD));
  if (dV) HIP_CALL(hipFree(dV));
  HIP_CALL(hipFree(dG));
  timers[3] = stopwatch_lap(timers[3]);
  timers[0] = stopwatch_lap(timers[0]);
  if (timing)
    for (unsigned i = 0u; i < 4u; ++i)
      timing[i] = timers[i] * TS2S;
  return 0;
}

extern "C" int hypJacL2_dev(unsigned routine, unsigned nrow, unsigned ncol, unsigned nplus,
                            double *dG, unsigned lddG, double *dV, unsigned lddV, double *dD,
                            unsigned *glbSwp, unsigned long long *glb_s,
                            unsigned long long *glb_b, double *timing, void *stream)
{
  long long t0 = 0, t1 = 0;
  stopwatch_reset(t0);
  const int rc = check_args(routine, nrow, ncol, nplus, dG, lddG, dV, lddV, dD, glbSwp, glb_s, glb_b,
                            g_width);
  if (rc)
    return rc;
  if (g_width == 32u && !wide_tables_ok(ncol))
    return -3;
  const bool full_svd = (routine & HYPJAC_FULL_SVD) != 0u;
  if ((lddG & 1u) || (reinterpret_cast<uintptr_t>(dG) & 15u)) return -13;
  if (full_svd && ((lddV & 1u) || (reinterpret_cast<uintptr_t>(dV) & 15u))) return -13;
  hipStream_t s = static_cast<hipStream_t>(stream);
  stopwatch_reset(t1);
  level2_run(routine, nrow, ncol, nplus, dG, lddG, full_svd ? dV : nullptr, lddV, dD, glbSwp,
              glb_s, glb_b, s);
  const long long tc = stopwatch_lap(t1);
  const long long tw = stopwatch_lap(t0);
  if (timing) {
    timing[0] = tw * TS2S;
    timing[1] = 0.0;
    timing[2] = tc * TS2S;
    timing[3] = 0.0;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Allocation-free level-2 run: every work buffer is carved from `scratch`,
// a fixed device array of the library (each GPU holds its own copy until the
// process ends).  In host code the array's name only stands for the symbol;
// its device address is looked up once per device (hipGetSymbolAddress, with
// that device current) and only that address is handed to kernels and
// copies.  Calls that overlap on one GPU share the array and race.

constexpr size_t NOALLOC_BYTES = (size_t)8u << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];
static void *g_scratch_ptr[NOALLOC_MAX_DEVICES];

namespace {

__global__ void k_zero_words(unsigned *p, size_t n)
{
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

// the scratch array's device address on the current device
unsigned char *scratch_here()
{
  int dev = 0;
  HIP_CALL(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
    HJ_DIE("hypJacL2_dev_noalloc: device number beyond the scratch table");
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    HIP_CALL(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = p;
  }
  return static_cast<unsigned char *>(g_scratch_ptr[dev]);
}

size_t align256(size_t b) { return (b + 255u) & ~(size_t)255u; }

// Byte offsets of the work buffers of a step-by-step level-2 run in the
// scratch array; total() is what the run needs.
struct ScratchPlan {
  size_t ctl, s0, s1, flags, W, part, rec, end;
  ScratchPlan(unsigned np, unsigned steps, unsigned nchunk)
  {
    ctl = 0u;
    s0 = align256(ctl + sizeof(RunCtl));
    s1 = align256(s0 + STRAT0_TAB_BYTES);
    flags = align256(s1 + sizeof(uint16_t) * 2u * np * steps);
    W = align256(flags + sizeof(unsigned) * np);
    part = align256(W + sizeof(double) * 1024u * np);
    rec = align256(part + sizeof(double) * 768u * nchunk * np);
    end = align256(rec + sizeof(VisitRec) * (size_t)np * steps);
  }
};

}  // namespace

// The sweeps of level2_core in the plain FP64 form, each level-1 step
// launched as Gram, pair visits and block update on the caller's stream,
// with the control block read back after every sweep.
static void level2_noalloc(unsigned routine, unsigned nrow, unsigned ncol, unsigned nplus,
                           double *dG, unsigned lddG, double *dV, unsigned lddV, double *dD,
                           unsigned *glbSwp, unsigned long long *glb_s,
                           unsigned long long *glb_b, hipStream_t s, unsigned char *base,
                           const ScratchPlan &sp)
{
  const bool cdsort = (routine & HYPJAC_CDSORT_1) != 0u;
  const bool full_svd = (routine & HYPJAC_FULL_SVD) != 0u;
  const unsigned nswp = (routine & HYPJAC_BLK_ORI) ? 1u : HYPJAC_NSWEEP;
  Work w;  // views into the scratch array; nothing is allocated or freed
  w.np = ncol / 32u;
  w.steps = STRAT1_STEPS;
  w.s0steps = STRAT0_STEPS;
  w.nchunk = gram_nchunk(nrow, w.np);
  w.CH = nrow / w.nchunk;
  w.chG = update_rows(nrow, full_svd ? ncol : 0u, w.np);
  w.chV = w.chG;
  w.d_ctl = reinterpret_cast<RunCtl *>(base + sp.ctl);
  w.d_s0 = base + sp.s0;
  w.d_s1 = reinterpret_cast<uint16_t *>(base + sp.s1);
  w.d_flags = reinterpret_cast<unsigned *>(base + sp.flags);
  w.d_W = reinterpret_cast<double *>(base + sp.W);
  w.d_part = reinterpret_cast<double *>(base + sp.part);
  w.d_rec = reinterpret_cast<VisitRec *>(base + sp.rec);
  // tables straight from the host arrays (strat1 rows hold STRAT1_MAX_PAIRS
  // pairs; the run's np of each row are taken)
  HIP_CALL(hipMemcpyAsync(w.d_s0, &strat0[0][0][0], STRAT0_TAB_BYTES, hipMemcpyHostToDevice, s));
  HIP_CALL(hipMemcpy2DAsync(w.d_s1, sizeof(uint16_t) * 2u * w.np, &strat1[0][0][0],
                            sizeof(uint16_t) * 2u * STRAT1_MAX_PAIRS, sizeof(uint16_t) * 2u * w.np,
                            w.steps, hipMemcpyHostToDevice, s));
  // the control block and the visit records start at zero (the array keeps
  // the last call's contents)
  hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, s, reinterpret_cast<unsigned *>(w.d_ctl),
                     sizeof(RunCtl) / sizeof(unsigned));
  const size_t nrec = sizeof(VisitRec) * w.keys() / sizeof(unsigned);
  hipLaunchKernelGGL(k_zero_words, dim3((unsigned)std::min<size_t>((nrec + 255u) / 256u, 1024u)),
                     dim3(256), 0, s, reinterpret_cast<unsigned *>(w.d_rec), nrec);
  HIP_CALL(hipGetLastError());
  if (full_svd)
    launch_initV(s, dV, 0u, ncol, lddV);
  const int definite = nplus ? ((nplus < ncol) ? 0 : +1) : -1;
  double *const V = full_svd ? dV : nullptr;

  g_stats = Stats();
  *glb_s = 0ull;
  *glb_b = 0ull;
  long long swp_tim = 0;
  stopwatch_reset(swp_tim);
  unsigned blk_swp = 0u;
  RunCtl h;
  while (blk_swp < g_max_sweeps) {
    launch_sweep_begin(s, w.d_ctl);
    for (unsigned st = 0u; st < w.steps; ++st)
      issue_step(s, w, st, cdsort, nswp, definite, nplus, dG, lddG, nrow, V, lddV, ncol, nullptr,
                 false, false);
    HIP_CALL(hipMemcpyAsync(&h, w.d_ctl, sizeof(RunCtl), hipMemcpyDeviceToHost, s));
    HIP_CALL(hipStreamSynchronize(s));
    ++blk_swp;
    const unsigned cvg_s = static_cast<unsigned>(h.cvg);
    const unsigned cvg_b = static_cast<unsigned>(h.cvg >> 32u);
    *glb_s += cvg_s;
    *glb_b += cvg_b;
    const double tim_s = stopwatch_lap(swp_tim) * TS2S;
    if (blk_swp <= HJ_MAX_SWEEP_CAP) {
      g_stats.rot_s[blk_swp - 1u] = cvg_s;
      g_stats.rot_b[blk_swp - 1u] = cvg_b;
      g_stats.sec[blk_swp - 1u] = tim_s;
    }
    g_stats.nswp = blk_swp;
    g_stats.visits += w.keys();
    if (!quiet()) {
      (void)fprintf(stdout, "BLK_SWP(%2u), ROT_S(%10u), ROT_B(%10u), TIME(%#12.6f s)\n", blk_swp,
                    cvg_s, cvg_b, tim_s);
      (void)fflush(stdout);
    }
    if (!cvg_b)
      break;
  }
  *glbSwp = blk_swp;
  launch_initD(s, dG, dD, 0u, nrow, ncol, nplus, lddG);
  HIP_CALL(hipStreamSynchronize(s));
  g_last_stage = 0;
}

extern "C" int hypJacL2_dev_noalloc(unsigned routine, unsigned nrow, unsigned ncol, unsigned nplus,
                                    double *dG, unsigned lddG, double *dV, unsigned lddV,
                                    double *dD, unsigned *glbSwp, unsigned long long *glb_s,
                                    unsigned long long *glb_b, double *timing, void *stream)
{
  long long t0 = 0, t1 = 0;
  stopwatch_reset(t0);
  const int rc = check_args(routine, nrow, ncol, nplus, dG, lddG, dV, lddV, dD, glbSwp, glb_s, glb_b);
  if (rc)
    return rc;
  const bool full_svd = (routine & HYPJAC_FULL_SVD) != 0u;
  if ((lddG & 1u) || (reinterpret_cast<uintptr_t>(dG) & 15u)) return -13;
  if (full_svd && ((lddV & 1u) || (reinterpret_cast<uintptr_t>(dV) & 15u))) return -13;
  check_tables(ncol);
  const unsigned np = ncol / 32u;
  const ScratchPlan sp(np, STRAT1_STEPS, gram_nchunk(nrow, np));
  if (sp.end > NOALLOC_BYTES) {
    if (!quiet())
      (void)fprintf(stderr, "hypJacL2_dev_noalloc: ncol %u needs %zu bytes of work buffers, "
                            "more than the %zu of the scratch array\n", ncol, sp.end, NOALLOC_BYTES);
    return HJ_ERR_SCRATCH;
  }
  unsigned char *const base = scratch_here();
  stopwatch_reset(t1);
  level2_noalloc(routine, nrow, ncol, nplus, dG, lddG, full_svd ? dV : nullptr, lddV, dD, glbSwp,
                 glb_s, glb_b, static_cast<hipStream_t>(stream), base, sp);
  const long long tc = stopwatch_lap(t1);
  const long long tw = stopwatch_lap(t0);
  if (timing) {
    timing[0] = tw * TS2S;
    timing[1] = 0.0;
    timing[2] = tc * TS2S;
    timing[3] = 0.0;
  }
  return 0;
}

extern "C" void hj_last_sweeps(unsigned *nswp, unsigned *rot_s, unsigned *rot_b, double *sec)
{
  if (nswp) *nswp = g_stats.nswp;
  for (unsigned i = 0u; i < g_stats.nswp; ++i) {
    if (rot_s) rot_s[i] = g_stats.rot_s[i];
    if (rot_b) rot_b[i] = g_stats.rot_b[i];
    if (sec) sec[i] = g_stats.sec[i];
  }
}

extern "C" void hj_set_profile(int on) { g_profile = on; }

extern "C" void hj_set_form(unsigned form) { g_form = form; }

extern "C" int hj_set_skip_model(const float *params, unsigned nparam, float threshold,
                                 unsigned max_skips)
{
  if (!params || nparam != MLP_NPARAM) return -1;
  for (unsigned i = 0u; i < MLP_NPARAM; ++i) {
    // bfloat16, round to nearest even
    uint32_t u;
    memcpy(&u, &params[i], 4);
    if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
    g_model_bf16[i] = (uint16_t)(u >> 16);
  }
  g_skip_threshold = threshold;
  g_skip_max = max_skips;
  g_model_ready = true;
  return 0;
}

extern "C" void hj_set_visit_trace(void *host_buf, unsigned max_sweeps)
{
  g_trace_buf = host_buf;
  g_trace_sweeps = host_buf ? max_sweeps : 0u;
}

extern "C" void hj_last_skips(unsigned *skip, unsigned long long *skipped, unsigned long long *visits)
{
  for (unsigned i = 0u; i < g_stats.nswp && i < HJ_MAX_SWEEP_CAP; ++i)
    if (skip) skip[i] = g_stats.skip[i];
  if (skipped) *skipped = g_stats.skipped;
  if (visits) *visits = g_stats.visits;
}

extern "C" void hj_set_split_tau(float tau) { g_split_tau = tau; }
extern "C" int hj_set_fused_split(int on)
{
  const int old = g_fused_split;
  g_fused_split = on ? 1 : 0;
  return old;
}

extern "C" void hj_set_skip_min_frac(float frac) { g_skip_min_frac = frac; }

// The floor of runs that mix J signs, as a multiple of hj_set_skip_min_frac's
extern "C" void hj_set_skip_floor_mixed(float factor)
{
  (g_width == 32u ? g_skip_floor_mixed_w32 : g_skip_floor_mixed) = factor;
}

extern "C" void hj_set_skip_min_propose(float frac) { g_skip_min_propose = frac; }

extern "C" void hj_set_skip_min_visits(unsigned visits) { g_skip_min_visits = visits; }

extern "C" void hj_last_splits(unsigned *split, unsigned long long *total)
{
  for (unsigned i = 0u; i < g_stats.nswp && i < HJ_MAX_SWEEP_CAP; ++i)
    if (split) split[i] = g_stats.split[i];
  if (total) *total = g_stats.splits;
}

// The 32x32 Gram of one 32-column block pair (columns 0..31 of a column-major
// nrow x 32 host matrix, leading dimension nrow) from the FP64 (split = 0) or
// the split-precision (split = 1) Gram kernel, partials summed in chunk
// order; both triangles are filled.
extern "C" int hj_dbg_gram(int split, unsigned nrow, const double *hG, double *hA)
{
  if (!hG || !hA || !nrow || (nrow % 64u)) return -1;
  const unsigned nchunk = gram_nchunk(nrow, 1u), CH = nrow / nchunk;
  double *dG = nullptr, *dP = nullptr;
  uint16_t *dpairs = nullptr;
  uint8_t *dplan = nullptr;
  HIP_CALL(hipMalloc(&dG, sizeof(double) * 32u * nrow));
  HIP_CALL(hipMalloc(&dP, sizeof(double) * 768u * nchunk));
  HIP_CALL(hipMalloc(&dpairs, 2u * sizeof(uint16_t)));
  HIP_CALL(hipMalloc(&dplan, 1u));
  const uint16_t pr[2] = {0u, 1u};
  const uint8_t pl = split ? PLAN_SPLIT : PLAN_FP64;
  HIP_CALL(hipMemcpy(dG, hG, sizeof(double) * 32u * nrow, hipMemcpyHostToDevice));
  HIP_CALL(hipMemcpy(dpairs, pr, sizeof(pr), hipMemcpyHostToDevice));
  HIP_CALL(hipMemcpy(dplan, &pl, 1u, hipMemcpyHostToDevice));
  if (split)
    launch_gram_split(nullptr, dG, nrow, dpairs, 1u, nchunk, CH, dP, dplan);
  else
    launch_gram(nullptr, dG, nrow, dpairs, 1u, nchunk, CH, dP, dplan);
  HIP_CALL(hipGetLastError());
  HIP_CALL(hipDeviceSynchronize());
  std::vector<double> P((size_t)768u * nchunk);
  HIP_CALL(hipMemcpy(P.data(), dP, sizeof(double) * P.size(), hipMemcpyDeviceToHost));
  for (unsigned e = 0u; e < 768u; ++e) {
    double v = P[e];
    for (unsigned c = 1u; c < nchunk; ++c) v = v + P[(size_t)c * 768u + e];
    const unsigned tile = e >> 8, l = (e >> 2) & 63u, r = e & 3u;
    const unsigned i = 4u * r + (l >> 4), j = l & 15u;
    const unsigned row = (tile == 0u) ? i : (16u + i);
    const unsigned col = (tile == 2u) ? (16u + j) : j;
    hA[(size_t)col * 32u + row] = v;
    if (tile == 1u) hA[(size_t)row * 32u + col] = v;
  }
  HIP_CALL(hipFree(dG));
  HIP_CALL(hipFree(dP));
  HIP_CALL(hipFree(dpairs));
  HIP_CALL(hipFree(dplan));
  return 0;
}

extern "C" double hj_kernel_ms(int kind, unsigned long long *launches)
{
  if (kind < 0 || kind > 5) return -1.0;
  if (launches) *launches = g_stats.kcnt[kind];
  return g_stats.kms[kind];
}

extern "C" unsigned hj_gram_nchunk(unsigned nrow, unsigned npairs) { return gram_nchunk(nrow, npairs); }

extern "C" void hj_set_step_limit(long steps) { g_step_limit = steps; }

extern "C" int hj_set_staged(int mode)
{
  const int prev = g_stage;
  g_stage = (mode < 0) ? 0 : (mode > 2) ? 2 : mode;
  return prev;
}

extern "C" int hj_last_staged(void) { return g_last_stage; }

extern "C" int hj_set_factorization(int f)
{
  const int prev = g_fact;
  if (f == HJ_FACT_CHOLESKY || f == HJ_FACT_QR || f == HJ_FACT_QR_DRMAC)
    g_fact = f;
  return prev;
}

extern "C" unsigned hj_set_block_width(unsigned width)
{
  const unsigned old = g_width;
  if (width == 16u || width == 32u)
    g_width = width;
  return old;
}

extern "C" unsigned hj_set_max_sweeps(unsigned n)
{
  const unsigned old = g_max_sweeps;
  if (n >= 1u && n <= HJ_MAX_SWEEP_CAP)
    g_max_sweeps = n;
  return old;
}

extern "C" int hj_dbg_scalar(int kind, int n, const double *a, const double *b, const double *c,
                             double *out_c, double *out_t, int *out_big)
{
  if (n <= 0) return 0;
  double *da, *db, *dc, *oc, *ot;
  int *ob;
  const size_t B = sizeof(double) * (size_t)n;
  HIP_CALL(hipMalloc(&da, B));
  HIP_CALL(hipMalloc(&db, B));
  HIP_CALL(hipMalloc(&dc, B));
  HIP_CALL(hipMalloc(&oc, B));
  HIP_CALL(hipMalloc(&ot, B));
  HIP_CALL(hipMalloc(&ob, sizeof(int) * (size_t)n));
  HIP_CALL(hipMemcpy(da, a, B, hipMemcpyHostToDevice));
  HIP_CALL(hipMemcpy(db, b ? b : a, B, hipMemcpyHostToDevice));
  HIP_CALL(hipMemcpy(dc, c ? c : a, B, hipMemcpyHostToDevice));
  launch_dbg_scalar(nullptr, kind, n, da, db, dc, oc, ot, ob);
  HIP_CALL(hipGetLastError());
  HIP_CALL(hipDeviceSynchronize());
  HIP_CALL(hipMemcpy(out_c, oc, B, hipMemcpyDeviceToHost));
  HIP_CALL(hipMemcpy(out_t, ot, B, hipMemcpyDeviceToHost));
  HIP_CALL(hipMemcpy(out_big, ob, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost));
  HIP_CALL(hipFree(da));
  HIP_CALL(hipFree(db));
  HIP_CALL(hipFree(dc));
  HIP_CALL(hipFree(oc));
  HIP_CALL(hipFree(ot));
  HIP_CALL(hipFree(ob));
  return 0;
}

extern "C" const char *hj_version(void) { return "gpujachx_amd 0.1 (gfx950)"; }

// ---------------------------------------------------------------------------
// Device configuration, share/cuda_helper.cu:5-54.

extern "C" int configureGPUex(const int dev, const unsigned maxShMemB)
{
  if (dev < 0) HJ_DIE("device number negative");
  (void)maxShMemB;  // LDS is 160 KiB per CU on gfx950; no cache split to choose
  HIP_CALL(hipSetDevice(dev));
  hipDeviceProp_t p;
  HIP_CALL(hipGetDeviceProperties(&p, dev));
  if (p.warpSize != 64) {
    char msg[256];
    (void)snprintf(msg, sizeof(msg), "HIP device %d has %d lanes per wave, must be 64", dev, p.warpSize);
    HJ_DIE(msg);
  }
  if (strncmp(p.gcnArchName, "gfx9", 4)) {
    char msg[256];
    (void)snprintf(msg, sizeof(msg), "HIP device %d is %s, not a gfx9 (CDNA) part", dev, p.gcnArchName);
    HJ_DIE(msg);
  }
  (void)scratch_here();  // the allocation-free entry's array on this device
  return p.major * 10 + p.minor;
}

extern "C" int configureGPU(const int dev) { return configureGPUex(dev, 16384u); }

// ---------------------------------------------------------------------------
// Thin level-1 layer, share/device_code.cu:15-118.

extern "C" void initSymbols(double *const G, double *const V, volatile unsigned long long *const cvg,
                            const unsigned nRow, const unsigned nCol, const unsigned ldG,
                            const unsigned ldV, const unsigned nSwp)
{
  if (g_thin.ready) {
    g_thin.w.release();
    g_thin.ready = false;
  }
  check_tables(nCol);
  g_thin.G = G;
  g_thin.V = V;
  g_thin.cvg = const_cast<unsigned long long *>(cvg);
  g_thin.nRow = nRow;
  g_thin.nCol = nCol;
  g_thin.ldG = ldG;
  g_thin.ldV = ldV;
  g_thin.nSwp = nSwp;
  g_thin.w.alloc(nRow, nCol, V != nullptr);
  g_thin.ready = true;
}

extern "C" void defJacL1(const unsigned step, const int definite) { thin_step(step, false, false, definite ? definite : 1, 0u); }
extern "C" void hypJacL1(const unsigned step, const unsigned npos) { thin_step(step, false, false, 0, npos); }
extern "C" void defJacL1v(const unsigned step, const int definite) { thin_step(step, false, true, definite ? definite : 1, 0u); }
extern "C" void hypJacL1v(const unsigned step, const unsigned npos) { thin_step(step, false, true, 0, npos); }
extern "C" void defJacL1s(const unsigned step, const int definite) { thin_step(step, true, false, definite ? definite : 1, 0u); }
extern "C" void hypJacL1s(const unsigned step, const unsigned npos) { thin_step(step, true, false, 0, npos); }
extern "C" void defJacL1sv(const unsigned step, const int definite) { thin_step(step, true, true, definite ? definite : 1, 0u); }
extern "C" void hypJacL1sv(const unsigned step, const unsigned npos) { thin_step(step, true, true, 0, npos); }

extern "C" void initD(double *const G, double *const D, const unsigned ifc, const unsigned nRow,
                      const unsigned nRank, const unsigned nPlus, const unsigned ldG)
{
  launch_initD(nullptr, G, D, ifc, nRow, nRank, nPlus, ldG);
  HIP_CALL(hipGetLastError());
}

extern "C" void initV(double *const V, const unsigned ifc, const unsigned nRank, const unsigned ldV)
{
  launch_initV(nullptr, V, ifc, nRank, ldV);
  HIP_CALL(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Built-in strategy tables (no strat.so needed).

extern "C" int hj_strat_generate(const char *fam, unsigned n, unsigned short *out, unsigned *steps)
{
  Strategy st;
  if (!generate(fam, n, st))
    return -1;
  if (steps) *steps = st.steps;
  if (out) memcpy(out, st.tab.data(), sizeof(uint16_t) * st.tab.size());
  return 0;
}

extern "C" void hj_init_strats_builtin(const char *snp0, unsigned n0, const char *snp1, unsigned n1)
{
  Strategy a, b;
  if (!snp0 || !snp1) HJ_DIE("strategy name empty");
  if (n0 != 2u * STRAT0_MAX_PAIRS) HJ_DIE("n0 must be 32");
  if (!n1 || (n1 & 1u) || n1 > 2u * STRAT1_MAX_PAIRS) HJ_DIE("n1 not positive even <= 1024");
  if (!generate(snp0, n0, a)) HJ_DIE("NO SUCH STRATEGY for SNP0 & n0");
  if (!generate(snp1, n1, b)) HJ_DIE("NO SUCH STRATEGY for SNP1 & n1");
  STRAT0 = family_id(snp0);
  STRAT0_STEPS = a.steps;
  STRAT0_PAIRS = a.pairs;
  memset(strat0, 0, sizeof(strat0));
  for (unsigned s = 0u; s < a.steps; ++s)
    for (unsigned p = 0u; p < a.pairs; ++p)
      for (unsigned i = 0u; i < 2u; ++i)
        strat0[s][p][i] = (unsigned char)a.tab[((size_t)s * a.pairs + p) * 2u + i];
  STRAT1 = family_id(snp1);
  STRAT1_STEPS = b.steps;
  STRAT1_PAIRS = b.pairs;
  memset(strat1, 0, sizeof(strat1));
  for (unsigned s = 0u; s < b.steps; ++s)
    for (unsigned p = 0u; p < b.pairs; ++p)
      for (unsigned i = 0u; i < 2u; ++i)
        strat1[s][p][i] = b.tab[((size_t)s * b.pairs + p) * 2u + i];
}
