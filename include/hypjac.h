/* hypjac.h -- C ABI of the MI355X (gfx950) blocked one-sided (hyperbolic)
 * Jacobi SVD library (libhypjac.so).
 *
 * The first block mirrors the entry points a GPUJACHx caller binds; each one
 * names the reference declaration it replaces.  Semantics (argument codes,
 * ALG bits, the stdout lines, fail-stop on runtime errors) are the same.
 * The second block holds additions of this library (device-resident inputs,
 * explicit tables, statistics, the split-precision and learned-skip forms,
 * the batched launch).  No torch or HIP types appear in any signature.
 */
#ifndef HYPJAC_H
#define HYPJAC_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------- */
/* Reference surface                                                        */
/* ---------------------------------------------------------------------- */

/* ALG bits (share/HypJacL.hpp:88-110) */
#define HYPJAC_BLK_ORI  8u  /* one inner sweep per block pair (else <= 30) */
#define HYPJAC_ACCUMV_2 4u  /* defined by the reference, never read */
#define HYPJAC_FULL_SVD 2u  /* accumulate V */
#define HYPJAC_CDSORT_1 1u  /* sort inside the 32-column blocks */

/* strategy IDs (share/HypJacL.hpp:118-152) */
#define STRAT_MMSTEP 1u
#define STRAT_BRENTL 2u
#define STRAT_COLCYC 3u
#define STRAT_CYCLOC 4u
#define STRAT_ROWCYC 5u
#define STRAT_CYCWOR 6u
#define STRAT_BLKREC 7u

#define HYPJACL1_NCOLB   16u   /* columns per block (share/HypJacL.hpp:112-116) */
#define HYPJAC_NSWEEP    30u   /* sweep cap per level (share/HypJacL.hpp:7-9) */
#define STRAT0_MAX_PAIRS 16u
#define STRAT0_MAX_STEPS 32u
#define STRAT1_MAX_PAIRS 512u
#define STRAT1_MAX_STEPS 1024u
#define HYPJAC_IDADIM_SIZE 4u  /* share/hdf5_helper.hpp:9-13 */

/* Level-2 driver.  Replaces hypJacL2 of 1/HypJacL2.hpp:6-27 (definition
 * 1/HypJacL2.cu:13-202): same arguments, same return codes
 * (-1 routine >= 16, -2 nrow, -3 ncol, -4 nplus, -5 hG, -6 ldhG, -7 hV,
 * -8 ldhV, -9 hD, -10..-12 counters), host buffers in, host buffers out,
 * device memory allocated and freed inside.  hG returns U (unit columns),
 * hD = j_i * sigma_i^2, hV (if routine & 2) the accumulated V.  Requires a
 * prior init_strats().  Prints one BLK_SWP line per sweep to stdout. */
int hypJacL2(const unsigned routine, const unsigned nrow, const unsigned ncol,
             const unsigned nplus, double *const hG, const unsigned ldhG,
             double *const hV, const unsigned ldhV, double *const hD,
             unsigned *const glbSwp, unsigned long long *const glb_s,
             unsigned long long *const glb_b, double *const timing);

/* Strategy loader.  Replaces init_strats of share/HypJacL.hpp:206
 * (definition share/HypJacL.cu:11-121): dlopen(sdy), dlsym("<snp>%05u"),
 * copy into the process-wide tables below; fail-stop (message + exit) on
 * any error, as the reference does. */
void init_strats(const char *const sdy, const char *const snp0,
                 const unsigned n0, const char *const snp1, const unsigned n1);

/* Process-wide strategy state, share/HypJacL.hpp:200-204. */
extern unsigned STRAT0, STRAT0_STEPS, STRAT0_PAIRS;
extern unsigned STRAT1, STRAT1_STEPS, STRAT1_PAIRS;
extern unsigned char strat0[STRAT0_MAX_STEPS][STRAT0_MAX_PAIRS][2];
extern unsigned short strat1[STRAT1_MAX_STEPS][STRAT1_MAX_PAIRS][2];

/* Device selection.  Replaces configureGPU/configureGPUex of
 * share/cuda_helper.hpp:29-30 (share/cuda_helper.cu:5-54).  Returns
 * 10*major+minor of the device (95 for gfx950); dies if the device is not a
 * 64-lane gfx9 part. */
int configureGPU(const int dev);
int configureGPUex(const int dev, const unsigned maxShMemB);

/* Strategy plugin helpers, share/my_utils.hpp:30-33 (share/my_utils.cu:30-61). */
void *strat_open(const char *const sdy);
int strat_close(void *const h);
const void *strat_ptr(void *const h, const char *const snp, const unsigned n);
int fexist(const char *const fn);

/* Thin level-1 layer on a process-wide context, share/device_code.hpp:6-21
 * (share/device_code.cu:15-118).  initSymbols binds device G/V, the
 * convergence counter (an 8-byte word in device memory, not host-mapped
 * memory) and the sizes; each *JacL1* call launches one level-1 step
 * (Gram, pair, update) for all pairs of STRAT1 step `step` on the default
 * stream. */
void initSymbols(double *const G, double *const V,
                 volatile unsigned long long *const cvg, const unsigned nRow,
                 const unsigned nCol, const unsigned ldG, const unsigned ldV,
                 const unsigned nSwp);
void defJacL1(const unsigned step, const int definite);
void hypJacL1(const unsigned step, const unsigned npos);
void defJacL1v(const unsigned step, const int definite);
void hypJacL1v(const unsigned step, const unsigned npos);
void defJacL1s(const unsigned step, const int definite);
void hypJacL1s(const unsigned step, const unsigned npos);
void defJacL1sv(const unsigned step, const int definite);
void hypJacL1sv(const unsigned step, const unsigned npos);
void initD(double *const G, double *const D, const unsigned ifc,
           const unsigned nRow, const unsigned nRank, const unsigned nPlus,
           const unsigned ldG);
void initV(double *const V, const unsigned ifc, const unsigned nRank,
           const unsigned ldV);

/* ---------------------------------------------------------------------- */
/* Additions of this library                                                */
/* ---------------------------------------------------------------------- */

/* hypJacL2 on device-resident data: dG (lddG x ncol), dV (lddV x ncol, if
 * routine & 2), dD (ncol) are device pointers; lddG, lddV must be even and
 * the pointers 16-byte aligned.  Work is issued on `stream` (a hipStream_t,
 * NULL = default stream); returns after the result is complete.  Same
 * return codes as hypJacL2 (-5/-7/-9 for NULL pointers, -13 alignment). */
int hypJacL2_dev(unsigned routine, unsigned nrow, unsigned ncol,
                 unsigned nplus, double *dG, unsigned lddG, double *dV,
                 unsigned lddV, double *dD, unsigned *glbSwp,
                 unsigned long long *glb_s, unsigned long long *glb_b,
                 double *timing, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (serving loops, harnesses that forbid hipMalloc, code that must not
 * reach the allocator): hypJacL2_dev's computation with every work buffer
 * taken from a fixed 8 MiB device array of the library (one per GPU, held
 * until the process ends; its address is looked up by configureGPU or at the
 * first call on a device), launched step by step on `stream` in the plain
 * FP64 form (the forms of hj_set_form are not used).  Allocates no device or
 * pinned memory, creates no stream, event or graph.  Results equal
 * hypJacL2_dev's plain form bit for bit.  Same return codes as
 * hypJacL2_dev, and HJ_ERR_SCRATCH when the run's work buffers exceed the
 * array (ncol <= 2048 always fits, at any nrow).  Calls that overlap on one
 * GPU share the array and race; calls in sequence are safe. */
#define HJ_ERR_SCRATCH (-14)
int hypJacL2_dev_noalloc(unsigned routine, unsigned nrow, unsigned ncol,
                         unsigned nplus, double *dG, unsigned lddG, double *dV,
                         unsigned lddV, double *dD, unsigned *glbSwp,
                         unsigned long long *glb_s, unsigned long long *glb_b,
                         double *timing, void *stream);

/* Forms of the method (hj_set_form, bit set; 0 = the reference's method). */
#define HJ_FORM_LEARNED_SKIP 1u  /* skip visits the pair-skip model predicts
                                    rotation-free; unskipped certification */
#define HJ_FORM_SPLIT_GRAM   2u  /* split-precision BF16 MFMA Gram while far
                                    from convergence, FP64 afterwards */
void hj_set_form(unsigned form);

/* Parameters of the pair-skip MLP (hj_mlp_nparam() floats, layout W1[F][H],
 * b1[H], W2[H][H], b2[H], W3[H], b3 with F = hj_mlp_nfeat(), H = 64),
 * stored as bfloat16; a visit is skipped when the predicted probability of
 * a big rotation is below `threshold`, at most `max_skips` visits in a row.
 * Returns -1 on a size mismatch. */
int hj_set_skip_model(const float *params, unsigned nparam, float threshold,
                      unsigned max_skips);
/* Skipping only in sweeps whose previous sweep had at least this fraction
 * of the possible big rotations (default 0.01): the final sweeps of a run go
 * unskipped.  And only in sweeps where the model proposes at least
 * min_propose of all visits (default 0.1): a few skips cannot pay for the
 * sweeps they may add. */
void hj_set_skip_min_frac(float frac);
/* runs whose J mixes signs (0 < NPLUS < N) use factor x that floor; the
 * factor of the block width in force (hj_set_block_width): default 3 at
 * width 16, 6 at width 32 */
void hj_set_skip_floor_mixed(float factor);
void hj_set_skip_min_propose(float min_propose);
/* Skipping only in runs with at least this many block-pair visits per sweep
 * (default 32000, i.e. N >= 4096 with 32-column pairs). */
void hj_set_skip_min_visits(unsigned visits);
unsigned hj_mlp_nparam(void);
unsigned hj_mlp_nfeat(void);
unsigned hj_visit_record_size(void);

/* Visit records (hj_visit_record_size() bytes each, steps*npairs per sweep)
 * of the following calls are copied to host_buf after every sweep, for up to
 * max_sweeps sweeps (NULL switches it off): the training traces. */
void hj_set_visit_trace(void *host_buf, unsigned max_sweeps);

/* Split-precision Gram (HJ_FORM_SPLIT_GRAM): a pair's Gram is formed from
 * three-way BF16 splits on BF16 MFMA with FP32 accumulation and FP64
 * recombination while its last max |cos| exceeds tau (default 1e-3) and in
 * the first sweep; the rotations, the Cholesky, the updates and the
 * convergence test stay FP64, and a run ends only on a sweep that is all
 * FP64 (and unskipped) without a big rotation. */
void hj_set_split_tau(float tau);
/* 1: the fused block update + next-step Gram forms the split Gram of the
 * pairs planned split, as the standalone Gram launch does; 0 (default): only
 * the standalone Gram launches (step 0 of a sweep) do, the fused ones form
 * FP64 Grams.  Returns the previous setting. */
int hj_set_fused_split(int on);
/* Split-Gram visits per sweep of the last call and their total. */
void hj_last_splits(unsigned *split, unsigned long long *total);
/* Self-test: the 32x32 Gram of columns 0..31 of a column-major nrow x 32
 * host matrix from the FP64 (split = 0) or split (split = 1) Gram kernel. */
int hj_dbg_gram(int split, unsigned nrow, const double *hG, double *hA);

/* ---- Many independent (H)SVDs in one persistent launch ---------------- */

/* One problem: the arguments of hypJacL2_dev (device pointers; G and V are
 * overwritten by U and V, D receives j_i sigma_i^2). */
typedef struct hj_batch_desc {
  unsigned routine, nrow, ncol, nplus, ldG, ldV;
  double *G, *V, *D;
} hj_batch_desc;

#define HJ_BATCH_STRAT_CYCWOR 0u  /* cycwor where it exists, else BrentL */
#define HJ_BATCH_STRAT_BRENTL 1u
#define HJ_BATCH_STRAT_MMSTEP 2u

/* Launch shape and schedule (hj_batch_autotune picks them). */
typedef struct hj_batch_cfg {
  unsigned waves;          /* waves per workgroup, 1..8 (one block pair each) */
  unsigned wg_per_cu;      /* workgroups per CU in the persistent grid */
  unsigned strategy;       /* HJ_BATCH_STRAT_* for the level-1 tables */
  unsigned max_sweeps;     /* sweep cap per problem (reference: 30) */
  unsigned longest_first;  /* queue order: largest M*N^2 first (else input order) */
} hj_batch_cfg;

void hj_batch_default_cfg(hj_batch_cfg *cfg);
/* Builds the plan of a batch: strategy tables, problem descriptors and the
 * batch captured as one hipGraph (queue reset + the persistent kernel).
 * Returns NULL and *err < 0 on an invalid problem (the hypJacL2 codes of
 * the first bad problem) or configuration. */
void *hj_batch_plan(const hj_batch_desc *probs, unsigned n,
                    const hj_batch_cfg *cfg, int *err);
/* Runs the plan (ordered after `stream`'s prior work) and waits for it. */
int hj_batch_run(void *plan, void *stream);
/* Per problem: sweeps, rotation totals, device start/end time in us since
 * the first problem started (NULL outputs are skipped). */
int hj_batch_results(void *plan, unsigned *sweeps, unsigned long long *rot_s,
                     unsigned long long *rot_b, double *start_us,
                     double *end_us);
void hj_batch_free(void *plan);
/* Times every candidate launch shape and strategy on `sample` problems of
 * the batch (copies; the inputs are untouched); best->max_sweeps is kept. */
int hj_batch_autotune(const hj_batch_desc *probs, unsigned n, unsigned sample,
                      hj_batch_cfg *best, double *best_rate);

/* Skipped visits per sweep of the last call, their total, executed visits. */
void hj_last_skips(unsigned *skip, unsigned long long *skipped,
                   unsigned long long *visits);

/* Features of the next visit computed from records (the same device code
 * the predictor uses); X: nkeys x hj_mlp_nfeat() floats. */
int hj_skip_features(const void *rec, unsigned nkeys,
                     const unsigned short *strat1, unsigned nplus, unsigned n1,
                     unsigned sweep_now, float prev_frac, float *X);

/* MLP on the GPU (own HIP kernels; BF16 weights, FP32 accumulation):
 * full-batch loss and gradient (weighted binary cross-entropy, positives
 * weighted by wpos), prediction, and Adam training (epoch losses out). */
int hj_mlp_grad(const float *X, const float *y, unsigned n,
                const float *params, float wpos, float *grads, float *loss);
int hj_mlp_predict(const float *X, unsigned n, const float *params,
                   float *prob);
int hj_mlp_train(const float *X, const float *y, unsigned n, unsigned epochs,
                 unsigned batch, float lr, float wpos, unsigned seed,
                 float *params, float *epoch_loss);

/* Per-sweep statistics of the last hypJacL2/hypJacL2_dev call in this
 * process: rot_s[i], rot_b[i] for sweeps i < *nswp (arrays as long as the
 * sweep cap). */
void hj_last_sweeps(unsigned *nswp, unsigned *rot_s, unsigned *rot_b,
                    double *sweep_seconds);

/* Kernel timing of the last call (HIP events on the work stream).  With
 * hj_set_profile(1) (graphs are not used while profiling), per launch:
 * kind 0 = Gram, 1 = pair, 2 = update, 3 = unused,
 * 4 = the launches of a staged sweep after step 0's visits (k_stage, and
 * the last step's update).  Without profiling, kind 5 = the graph-launched
 * sweeps: their GPU time (events around each graph launch) and the
 * launches of each sweep's dominant kind they hold (staged launches, else
 * one update-class launch per step).  Returns
 * total ms and launch count. */
void hj_set_profile(int on);
double hj_kernel_ms(int kind, unsigned long long *launches);

/* Number of row chunks of the Gram of one block pair (for nrow rows and
 * npairs pairs per step); the oracle reproduces the same split. */
unsigned hj_gram_nchunk(unsigned nrow, unsigned npairs);

/* Level-2 sweep cap (the reference's compile-time HYPJAC_NSWEEP, 30 by
 * default), 1..256; returns the previous cap.  Inner sweeps (ALG < 8) keep
 * the reference's cap of 30. */
/* Block width of the following hypJacL2 / hypJacL2_dev calls: 16 (default,
 * the reference's HYPJACL1_NCOLB, share/HypJacL.hpp:112-116) or 32
 * (64-column block pairs: a 64x64 Gram, Cholesky and level-0 per visit,
 * half the HBM bytes per sweep).  At 32 the plain FP64 method and the
 * learned pair-skip run; the split Gram bit of the form and the QR / Drmac
 * factorization are ignored (Cholesky, FP64 Grams).  The strategy families
 * loaded by init_strats are generated at orders 64 and ncol/32: -3 when
 * ncol is not a multiple of 64 or a family has no table of such an order.
 * hypJacL2_dev_noalloc and the level-1 layer keep width 16.  Other values
 * are ignored.  Returns the previous width. */
unsigned hj_set_block_width(unsigned width);
unsigned hj_set_max_sweeps(unsigned n);

/* Test limit: stop after `steps` level-1 steps in total (0 = no limit). */
void hj_set_step_limit(long steps);

/* Staged sweeps (no counterpart in the reference; same role as its
 * per-step launches, 1/HypJacL2.cu:131-148): each step transition runs as
 * two launches, each the fused block update + next Gram of half the step's
 * closed quads beside the pair visits whose Gram the previous launch
 * completed (and, with V, the update of V), so the visits' latency hides
 * under the block updates' HBM stream.  The results are bit for bit those
 * of the step-by-step launches.  mode 0: off; 1 (default): where a step's
 * block updates read and write at least 4.8e8 bytes (512 (nrow + nV)
 * ncol/32; below that the visits' latency is longer than half a step's
 * updates and step-by-step launches are faster); 2: wherever they apply.
 * They apply when every step transition forms closed quads, ncol >= 128, no step limit is set and the factorization is
 * Cholesky.  Returns the previous mode; hj_last_staged() tells whether the
 * last hypJacL2 / hypJacL2_dev call used them. */
int hj_set_staged(int mode);
int hj_last_staged(void);

/* Factorization of the block pairs, at run time where the reference chooses
 * at build time (share/device_code_common.hpp:78-85,140-144):
 * HJ_FACT_CHOLESKY -- Gram and Cholesky (the reference's default build,
 *   share/device_code_common_Cholesky.hpp:4-171);
 * HJ_FACT_QR -- R by 32-row panels, Householder QR of each and Givens merges
 *   (-DUSE_QR: share/device_code_common_QR.hpp:4-94);
 * HJ_FACT_QR_DRMAC -- the same with Drmac's overflow-safe sorted level-0
 *   for sorted definite blocks (-DUSE_DRMAC:
 *   share/device_code_cdsort_Drmac.hpp:4-316).
 * The learned and split forms act on the Cholesky path only.  Returns the
 * previous setting (an unknown value leaves it unchanged). */
#define HJ_FACT_CHOLESKY 0
#define HJ_FACT_QR       1
#define HJ_FACT_QR_DRMAC 2
int hj_set_factorization(int f);

/* Device self-test of the scalar kernels: for n inputs, kind 0 = rsqrt(a),
 * 1 = trigonometric rotation (a=Apq,b=Dp,c=Dq), 2 = hyperbolic rotation.
 * Outputs c, t, big per input.  Returns 0 on success. */
int hj_dbg_scalar(int kind, int n, const double *a, const double *b,
                  const double *c, double *out_c, double *out_t, int *out_big);

/* Built-in strategy tables: generate family `fam` of order n into
 * out[steps][n/2][2] (may be NULL to query *steps); -1 if this library has
 * no construction for (fam, n).  hj_init_strats_builtin fills the same
 * process-wide tables as init_strats without a strat.so. */
int hj_strat_generate(const char *fam, unsigned n, unsigned short *out,
                      unsigned *steps);
void hj_init_strats_builtin(const char *snp0, unsigned n0, const char *snp1,
                            unsigned n1);

/* Version string of the library build. */
const char *hj_version(void);

#ifdef __cplusplus
}
#endif

#endif /* HYPJAC_H */
