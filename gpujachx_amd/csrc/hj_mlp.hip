// hj_mlp.hip -- learned pair-skip: features of a block-pair visit, the MLP
// (12 inputs, two hidden layers of 64 ReLU units, one logit; BF16 weights,
// FP32 accumulation), its inference kernel that writes the skip mask of the
// next level-2 sweep, and the training kernels (forward, backward, Adam).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "hj_kernels.hpp"
#include "hj_mlp.hpp"
#include "hj_util.hpp"

namespace hj {

// round to the nearest bfloat16 (ties to even) and back to float
__host__ __device__ inline float bf16_round(float x)
{
  uint32_t u;
  memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) == 0x7f800000u)
    return x;
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float r;
  memcpy(&r, &u, 4);
  return r;
}

__host__ __device__ inline uint16_t bf16_bits(float x)
{
  const float r = bf16_round(x);
  uint32_t u;
  memcpy(&u, &r, 4);
  return (uint16_t)(u >> 16);
}

__device__ inline float bf16_to_float(uint16_t b)
{
  return __uint_as_float(((uint32_t)b) << 16);
}

// Features of the next visit of a block pair from its record (the last
// executed visit) and run-level context.  Identical for training and
// inference.
__device__ void visit_features(const VisitRec &v, unsigned sweep_now, float prev_frac, float jmix,
                               float log2n1, float *f)
{
  f[0] = fminf((float)v.rot_b, 496.0f) / 496.0f;
  f[1] = fminf((float)v.rot_s, 496.0f) / 496.0f;
  f[2] = log10f(fmaxf(v.maxcos, 1e-30f)) / 16.0f;
  f[3] = (v.l2ni - v.l2nj) / 64.0f;
  f[4] = fabsf(v.l2ni - v.l2nj) / 64.0f;
  f[5] = (float)sweep_now / 32.0f;
  f[6] = (float)(sweep_now - v.sweep) / 4.0f;
  f[7] = (float)v.skips / 4.0f;
  f[8] = prev_frac;
  f[9] = log10f(fmaxf(prev_frac, 1e-12f)) / 12.0f;
  f[10] = jmix;
  f[11] = log2n1 / 10.0f;
}

// fraction of negative-J columns among the 2 bw columns of blocks (bi, bj)
// of bw columns each
__device__ inline float pair_jmix(unsigned bi, unsigned bj, unsigned nplus, unsigned bw)
{
  auto neg = [&](unsigned b) {
    const unsigned lo = bw * b, hi = lo + bw;
    if (nplus <= lo) return bw;
    if (nplus >= hi) return 0u;
    return hi - nplus;
  };
  return (float)(neg(bi) + neg(bj)) / (float)(2u * bw);
}

// MLP forward on one feature vector with float weights in w (layout of
// hj_mlp.hpp); returns the logit of P(visit has a big rotation).
__device__ float mlp_logit(const float *w, const float *f)
{
  float h1[MLP_H], h2;
  for (int j = 0; j < MLP_H; ++j) {
    float a = w[MLP_OFF_B1 + j];
    for (int i = 0; i < MLP_F; ++i)
      a = fmaf(f[i], w[MLP_OFF_W1 + i * MLP_H + j], a);
    h1[j] = fmaxf(a, 0.0f);
  }
  float z = w[MLP_OFF_B3];
  for (int j = 0; j < MLP_H; ++j) {
    float a = w[MLP_OFF_B2 + j];
    for (int i = 0; i < MLP_H; ++i)
      a = fmaf(h1[i], w[MLP_OFF_W2 + i * MLP_H + j], a);
    h2 = fmaxf(a, 0.0f);
    z = fmaf(h2, w[MLP_OFF_W3 + j], z);
  }
  return z;
}

// Plan of the running sweep, one thread per (step, slot): PLAN_SKIP when
// skipping is on and the model's probability of a big rotation is below the
// threshold (at most max_skips visits in a row); otherwise PLAN_SPLIT when
// the split Gram is on and the pair is far from convergence (never visited,
// or its last max |cos| above tau and its last split Cholesky did not fail);
// otherwise PLAN_FP64.
__global__ __launch_bounds__(256) void k_plan(const VisitRec *__restrict__ rec, unsigned nkeys,
                                              const uint16_t *__restrict__ strat1, unsigned nplus,
                                              float log2n1, float max_rot,
                                              RunCtl *__restrict__ ctl,
                                              const uint16_t *__restrict__ model, int learned,
                                              uint8_t *__restrict__ plan, unsigned bw)
{
  __shared__ float w[MLP_NPARAM];
  if (learned) {
    for (unsigned e = threadIdx.x; e < MLP_NPARAM; e += blockDim.x)
      w[e] = bf16_to_float(model[e]);
    __syncthreads();
  }
  const unsigned key = blockIdx.x * blockDim.x + threadIdx.x;
  if (key >= nkeys)
    return;
  const VisitRec v = rec[key];
  bool skip = false;
  if (learned && ctl->skip_on && v.sweep != 0u && v.skips < ctl->max_skips) {
    const float prev = (float)(unsigned)(ctl->prev_cvg >> 32) / max_rot;
    float f[MLP_F];
    visit_features(v, ctl->sweep, prev, pair_jmix(strat1[2u * key], strat1[2u * key + 1u], nplus, bw),
                   log2n1, f);
    const float z = mlp_logit(w, f);
    const float p = 1.0f / (1.0f + expf(-z));
    skip = (p < ctl->threshold);
  }
  const bool split = ctl->split_on &&
                     (v.sweep == 0u || (v.maxcos > ctl->tau && !(v.flags & REC_SPLIT_FAIL)));
  uint8_t s = split ? PLAN_SPLIT : PLAN_FP64;
  if (skip)
    s = split ? PLAN_PROP_SKIP_SPLIT : PLAN_PROP_SKIP_FP64;
  // proposals counted per wave
  const unsigned long long b = __ballot(skip);
  if ((threadIdx.x & 63u) == 0u && b)
    atomicAdd(&ctl->proposed, (unsigned)__popcll(b));
  plan[key] = s;
}

// Settles the proposals: the sweep skips the proposed visits only if they
// are at least min_skip_frac of all (a few skips cannot pay for the sweeps
// they may add); otherwise they run, FP64 or split as planned.
__global__ __launch_bounds__(256) void k_plan_gate(unsigned nkeys, const RunCtl *__restrict__ ctl,
                                                   uint8_t *__restrict__ plan)
{
  const unsigned key = blockIdx.x * blockDim.x + threadIdx.x;
  if (key >= nkeys)
    return;
  const uint8_t s = plan[key];
  if (s < PLAN_PROP_SKIP_FP64)
    return;
  const bool go = (float)ctl->proposed >= ctl->min_skip_frac * (float)nkeys;
  plan[key] = go ? PLAN_SKIP : ((s == PLAN_PROP_SKIP_SPLIT) ? PLAN_SPLIT : PLAN_FP64);
}

// Features of many records (training-set construction).
__global__ void k_features(const VisitRec *__restrict__ rec, unsigned nkeys,
                           const uint16_t *__restrict__ strat1, unsigned nplus, float log2n1,
                           unsigned sweep_now, float prev_frac, float *__restrict__ X)
{
  const unsigned key = blockIdx.x * blockDim.x + threadIdx.x;
  if (key >= nkeys)
    return;
  visit_features(rec[key], sweep_now, prev_frac,
                 pair_jmix(strat1[2u * key], strat1[2u * key + 1u], nplus, 16u), log2n1,
                 X + (size_t)key * MLP_F);
}

// ---------------------------------------------------------------------------
// Training.  Forward keeps the pre-activations; weights are the bf16-rounded
// master weights; every accumulation is FP32.

__global__ void k_round_params(const float *__restrict__ p, float *__restrict__ q, unsigned n)
{
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) q[i] = bf16_round(p[i]);
}

// layer: Y[b][j] = bias[j] + sum_i X[b][i] W[i][j], optional ReLU copy
__global__ void k_dense_fwd(const float *__restrict__ X, unsigned B, unsigned In, unsigned Out,
                            const float *__restrict__ W, const float *__restrict__ bias,
                            const unsigned *__restrict__ idx, float *__restrict__ Z,
                            float *__restrict__ A)
{
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * Out) return;
  const unsigned b = t / Out, j = t - b * Out;
  const float *x = X + (size_t)(idx ? idx[b] : b) * In;
  float a = bias[j];
  for (unsigned i = 0u; i < In; ++i)
    a = fmaf(x[i], W[i * Out + j], a);
  Z[t] = a;
  if (A) A[t] = fmaxf(a, 0.0f);
}

// d logit for weighted binary cross-entropy, mean over the batch
__global__ void k_bce_grad(const float *__restrict__ z, const float *__restrict__ y,
                           const unsigned *__restrict__ idx, unsigned B, float wpos,
                           float *__restrict__ dz, float *__restrict__ loss_terms)
{
  const unsigned b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const float yy = y[idx ? idx[b] : b];
  const float wt = (yy > 0.5f) ? wpos : 1.0f;
  const float zz = z[b];
  const float p = 1.0f / (1.0f + expf(-zz));
  dz[b] = wt * (p - yy) / (float)B;
  // softplus(z) - y z, stable
  const float sp = fmaxf(zz, 0.0f) + log1pf(expf(-fabsf(zz)));
  loss_terms[b] = wt * (sp - yy * zz);
}

// dX[b][i] = sum_j dY[b][j] W[i][j], then times ReLU'(Zprev) if given
__global__ void k_dense_bwd_data(const float *__restrict__ dY, unsigned B, unsigned In, unsigned Out,
                                 const float *__restrict__ W, const float *__restrict__ Zprev,
                                 float *__restrict__ dX)
{
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * In) return;
  const unsigned b = t / In, i = t - b * In;
  float a = 0.0f;
  for (unsigned j = 0u; j < Out; ++j)
    a = fmaf(dY[(size_t)b * Out + j], W[i * Out + j], a);
  if (Zprev) a = (Zprev[t] > 0.0f) ? a : 0.0f;
  dX[t] = a;
}

// gW[i][j] = sum_b X[b][i] dY[b][j]; gb[j] = sum_b dY[b][j] (one thread per
// weight, the batch in order: deterministic)
__global__ void k_dense_bwd_w(const float *__restrict__ X, const unsigned *__restrict__ idx,
                              const float *__restrict__ dY, unsigned B, unsigned In, unsigned Out,
                              float *__restrict__ gW, float *__restrict__ gb)
{
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < In * Out) {
    const unsigned i = t / Out, j = t - i * Out;
    float a = 0.0f;
    for (unsigned b = 0u; b < B; ++b)
      a = fmaf(X[(size_t)(idx ? idx[b] : b) * In + i], dY[(size_t)b * Out + j], a);
    gW[t] = a;
  }
  else if (t < In * Out + Out) {
    const unsigned j = t - In * Out;
    float a = 0.0f;
    for (unsigned b = 0u; b < B; ++b)
      a += dY[(size_t)b * Out + j];
    gb[j] = a;
  }
}

__global__ void k_adam(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, unsigned n, float lr, float b1, float b2,
                       float eps, float c1, float c2)
{
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  m[i] = b1 * m[i] + (1.0f - b1) * gi;
  v[i] = b2 * v[i] + (1.0f - b2) * gi * gi;
  const float mh = m[i] / c1, vh = v[i] / c2;
  p[i] -= lr * mh / (sqrtf(vh) + eps);
}

void launch_plan(hipStream_t s, const VisitRec *rec, unsigned nkeys, const uint16_t *strat1,
                 unsigned nplus, float log2n1, float max_rot, RunCtl *ctl,
                 const uint16_t *model, bool learned, uint8_t *plan, unsigned bw)
{
  hipLaunchKernelGGL(k_plan, dim3((nkeys + 255u) / 256u), dim3(256), 0, s, rec, nkeys, strat1, nplus,
                     log2n1, max_rot, ctl, model, learned ? 1 : 0, plan, bw);
  if (learned)
    hipLaunchKernelGGL(k_plan_gate, dim3((nkeys + 255u) / 256u), dim3(256), 0, s, nkeys,
                       (const RunCtl *)ctl, plan);
}

// Device buffers of one training / gradient evaluation.
struct MlpBufs {
  unsigned B = 0u;
  float *Z1, *A1, *Z2, *A2, *z3, *dz3, *dZ2, *dZ1, *lt, *g, *q;
  void alloc(unsigned b)
  {
    B = b;
    HIP_CALL(hipMalloc(&Z1, sizeof(float) * B * MLP_H));
    HIP_CALL(hipMalloc(&A1, sizeof(float) * B * MLP_H));
    HIP_CALL(hipMalloc(&Z2, sizeof(float) * B * MLP_H));
    HIP_CALL(hipMalloc(&A2, sizeof(float) * B * MLP_H));
    HIP_CALL(hipMalloc(&z3, sizeof(float) * B));
    HIP_CALL(hipMalloc(&dz3, sizeof(float) * B));
    HIP_CALL(hipMalloc(&dZ2, sizeof(float) * B * MLP_H));
    HIP_CALL(hipMalloc(&dZ1, sizeof(float) * B * MLP_H));
    HIP_CALL(hipMalloc(&lt, sizeof(float) * B));
    HIP_CALL(hipMalloc(&g, sizeof(float) * MLP_NPARAM));
    HIP_CALL(hipMalloc(&q, sizeof(float) * MLP_NPARAM));
  }
  void release()
  {
    for (float *p : {Z1, A1, Z2, A2, z3, dz3, dZ2, dZ1, lt, g, q}) HIP_CALL(hipFree(p));
  }
};

static inline unsigned nblk(unsigned n) { return (n + 255u) / 256u; }

// forward + backward of one batch (rows idx[0..B) of X, or the first B rows);
// gradients into bufs.g; returns nothing (loss terms in bufs.lt)
static void mlp_fwd_bwd(hipStream_t s, const float *X, const float *y, const unsigned *idx, unsigned B,
                        const float *params, float wpos, MlpBufs &m)
{
  const float *q = m.q;
  hipLaunchKernelGGL(k_round_params, dim3(nblk(MLP_NPARAM)), dim3(256), 0, s, params, m.q, MLP_NPARAM);
  hipLaunchKernelGGL(k_dense_fwd, dim3(nblk(B * MLP_H)), dim3(256), 0, s, X, B, MLP_F, MLP_H,
                     q + MLP_OFF_W1, q + MLP_OFF_B1, idx, m.Z1, m.A1);
  hipLaunchKernelGGL(k_dense_fwd, dim3(nblk(B * MLP_H)), dim3(256), 0, s, m.A1, B, MLP_H, MLP_H,
                     q + MLP_OFF_W2, q + MLP_OFF_B2, nullptr, m.Z2, m.A2);
  hipLaunchKernelGGL(k_dense_fwd, dim3(nblk(B)), dim3(256), 0, s, m.A2, B, MLP_H, 1u, q + MLP_OFF_W3,
                     q + MLP_OFF_B3, nullptr, m.z3, nullptr);
  hipLaunchKernelGGL(k_bce_grad, dim3(nblk(B)), dim3(256), 0, s, m.z3, y, idx, B, wpos, m.dz3, m.lt);
  // last layer
  hipLaunchKernelGGL(k_dense_bwd_w, dim3(nblk(MLP_H + 1u)), dim3(256), 0, s, m.A2, nullptr, m.dz3, B,
                     MLP_H, 1u, m.g + MLP_OFF_W3, m.g + MLP_OFF_B3);
  hipLaunchKernelGGL(k_dense_bwd_data, dim3(nblk(B * MLP_H)), dim3(256), 0, s, m.dz3, B, MLP_H, 1u,
                     q + MLP_OFF_W3, m.Z2, m.dZ2);
  hipLaunchKernelGGL(k_dense_bwd_w, dim3(nblk(MLP_H * MLP_H + MLP_H)), dim3(256), 0, s, m.A1, nullptr,
                     m.dZ2, B, MLP_H, MLP_H, m.g + MLP_OFF_W2, m.g + MLP_OFF_B2);
  hipLaunchKernelGGL(k_dense_bwd_data, dim3(nblk(B * MLP_H)), dim3(256), 0, s, m.dZ2, B, MLP_H, MLP_H,
                     q + MLP_OFF_W2, m.Z1, m.dZ1);
  hipLaunchKernelGGL(k_dense_bwd_w, dim3(nblk(MLP_F * MLP_H + MLP_H)), dim3(256), 0, s, X, idx, m.dZ1,
                     B, MLP_F, MLP_H, m.g + MLP_OFF_W1, m.g + MLP_OFF_B1);
}

}  // namespace hj

using namespace hj;

extern "C" unsigned hj_mlp_nparam(void) { return MLP_NPARAM; }
extern "C" unsigned hj_mlp_nfeat(void) { return MLP_F; }

extern "C" int hj_mlp_grad(const float *X, const float *y, unsigned n, const float *params, float wpos,
                           float *grads, float *loss)
{
  if (!n) return -1;
  float *dX, *dy, *dp;
  HIP_CALL(hipMalloc(&dX, sizeof(float) * n * MLP_F));
  HIP_CALL(hipMalloc(&dy, sizeof(float) * n));
  HIP_CALL(hipMalloc(&dp, sizeof(float) * MLP_NPARAM));
  HIP_CALL(hipMemcpy(dX, X, sizeof(float) * n * MLP_F, hipMemcpyHostToDevice));
  HIP_CALL(hipMemcpy(dy, y, sizeof(float) * n, hipMemcpyHostToDevice));
  HIP_CALL(hipMemcpy(dp, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice));
  MlpBufs m;
  m.alloc(n);
  mlp_fwd_bwd(nullptr, dX, dy, nullptr, n, dp, wpos, m);
  HIP_CALL(hipGetLastError());
  HIP_CALL(hipDeviceSynchronize());
  HIP_CALL(hipMemcpy(grads, m.g, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost));
  std::vector<float> lt(n);
  HIP_CALL(hipMemcpy(lt.data(), m.lt, sizeof(float) * n, hipMemcpyDeviceToHost));
  double L = 0.0;
  for (float v : lt) L += v;
  if (loss) *loss = (float)(L / n);
  m.release();
  HIP_CALL(hipFree(dX));
  HIP_CALL(hipFree(dy));
  HIP_CALL(hipFree(dp));
  return 0;
}

extern "C" int hj_mlp_predict(const float *X, unsigned n, const float *params, float *prob)
{
  if (!n) return -1;
  float *dX, *dp;
  HIP_CALL(hipMalloc(&dX, sizeof(float) * n * MLP_F));
  HIP_CALL(hipMalloc(&dp, sizeof(float) * MLP_NPARAM));
  HIP_CALL(hipMemcpy(dX, X, sizeof(float) * n * MLP_F, hipMemcpyHostToDevice));
  HIP_CALL(hipMemcpy(dp, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice));
  MlpBufs m;
  m.alloc(n);
  const float *q = m.q;
  hipLaunchKernelGGL(k_round_params, dim3(nblk(MLP_NPARAM)), dim3(256), 0, 0, dp, m.q, MLP_NPARAM);
  hipLaunchKernelGGL(k_dense_fwd, dim3(nblk(n * MLP_H)), dim3(256), 0, 0, dX, n, MLP_F, MLP_H,
                     q + MLP_OFF_W1, q + MLP_OFF_B1, nullptr, m.Z1, m.A1);
  hipLaunchKernelGGL(k_dense_fwd, dim3(nblk(n * MLP_H)), dim3(256), 0, 0, m.A1, n, MLP_H, MLP_H,
                     q + MLP_OFF_W2, q + MLP_OFF_B2, nullptr, m.Z2, m.A2);
  hipLaunchKernelGGL(k_dense_fwd, dim3(nblk(n)), dim3(256), 0, 0, m.A2, n, MLP_H, 1u, q + MLP_OFF_W3,
                     q + MLP_OFF_B3, nullptr, m.z3, nullptr);
  HIP_CALL(hipGetLastError());
  HIP_CALL(hipDeviceSynchronize());
  std::vector<float> z(n);
  HIP_CALL(hipMemcpy(z.data(), m.z3, sizeof(float) * n, hipMemcpyDeviceToHost));
  for (unsigned i = 0u; i < n; ++i) prob[i] = 1.0f / (1.0f + std::exp(-z[i]));
  m.release();
  HIP_CALL(hipFree(dX));
  HIP_CALL(hipFree(dp));
  return 0;
}

extern "C" int hj_mlp_train(const float *X, const float *y, unsigned n, unsigned epochs,
                            unsigned batch, float lr, float wpos, unsigned seed, float *params,
                            float *epoch_loss)
{
  if (!n || !batch) return -1;
  if (batch > n) batch = n;
  float *dX, *dy, *dp, *mm, *vv;
  unsigned *didx;
  HIP_CALL(hipMalloc(&dX, sizeof(float) * n * MLP_F));
  HIP_CALL(hipMalloc(&dy, sizeof(float) * n));
  HIP_CALL(hipMalloc(&dp, sizeof(float) * MLP_NPARAM));
  HIP_CALL(hipMalloc(&mm, sizeof(float) * MLP_NPARAM));
  HIP_CALL(hipMalloc(&vv, sizeof(float) * MLP_NPARAM));
  HIP_CALL(hipMalloc(&didx, sizeof(unsigned) * n));
  HIP_CALL(hipMemcpy(dX, X, sizeof(float) * n * MLP_F, hipMemcpyHostToDevice));
  HIP_CALL(hipMemcpy(dy, y, sizeof(float) * n, hipMemcpyHostToDevice));
  HIP_CALL(hipMemcpy(dp, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice));
  HIP_CALL(hipMemset(mm, 0, sizeof(float) * MLP_NPARAM));
  HIP_CALL(hipMemset(vv, 0, sizeof(float) * MLP_NPARAM));
  MlpBufs m;
  m.alloc(batch);
  std::vector<unsigned> perm(n);
  uint64_t st = 0x9e3779b97f4a7c15ull ^ seed;
  auto rnd = [&]() {
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    return st;
  };
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  unsigned long long t = 0ull;
  std::vector<float> lt(batch);
  for (unsigned ep = 0u; ep < epochs; ++ep) {
    for (unsigned i = 0u; i < n; ++i) perm[i] = i;
    for (unsigned i = n - 1u; i > 0u; --i) {
      const unsigned j = (unsigned)(rnd() % (i + 1u));
      std::swap(perm[i], perm[j]);
    }
    HIP_CALL(hipMemcpy(didx, perm.data(), sizeof(unsigned) * n, hipMemcpyHostToDevice));
    double L = 0.0;
    unsigned nb = 0u;
    for (unsigned off = 0u; off + batch <= n; off += batch) {
      mlp_fwd_bwd(nullptr, dX, dy, didx + off, batch, dp, wpos, m);
      ++t;
      const float c1 = 1.0f - std::pow(b1, (float)t), c2 = 1.0f - std::pow(b2, (float)t);
      hipLaunchKernelGGL(k_adam, dim3(nblk(MLP_NPARAM)), dim3(256), 0, 0, dp, mm, vv, m.g,
                         MLP_NPARAM, lr, b1, b2, eps, c1, c2);
      if (off == 0u || off + 2u * batch > n) {
        HIP_CALL(hipMemcpy(lt.data(), m.lt, sizeof(float) * batch, hipMemcpyDeviceToHost));
        double s = 0.0;
        for (float v : lt) s += v;
        L += s / batch;
        ++nb;
      }
    }
    HIP_CALL(hipGetLastError());
    if (epoch_loss) epoch_loss[ep] = (float)(L / (nb ? nb : 1u));
  }
  HIP_CALL(hipDeviceSynchronize());
  HIP_CALL(hipMemcpy(params, dp, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost));
  m.release();
  for (void *p : {(void *)dX, (void *)dy, (void *)dp, (void *)mm, (void *)vv, (void *)didx})
    HIP_CALL(hipFree(p));
  return 0;
}

extern "C" int hj_skip_features(const void *rec, unsigned nkeys, const unsigned short *strat1,
                                unsigned nplus, unsigned n1, unsigned sweep_now, float prev_frac,
                                float *X)
{
  if (!nkeys) return -1;
  VisitRec *dr;
  uint16_t *ds;
  float *dX;
  HIP_CALL(hipMalloc(&dr, sizeof(VisitRec) * nkeys));
  HIP_CALL(hipMalloc(&ds, sizeof(uint16_t) * 2u * nkeys));
  HIP_CALL(hipMalloc(&dX, sizeof(float) * MLP_F * nkeys));
  HIP_CALL(hipMemcpy(dr, rec, sizeof(VisitRec) * nkeys, hipMemcpyHostToDevice));
  HIP_CALL(hipMemcpy(ds, strat1, sizeof(uint16_t) * 2u * nkeys, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_features, dim3(nblk(nkeys)), dim3(256), 0, 0, dr, nkeys, ds, nplus,
                     (float)std::log2((double)n1), sweep_now, prev_frac, dX);
  HIP_CALL(hipGetLastError());
  HIP_CALL(hipDeviceSynchronize());
  HIP_CALL(hipMemcpy(X, dX, sizeof(float) * MLP_F * nkeys, hipMemcpyDeviceToHost));
  HIP_CALL(hipFree(dr));
  HIP_CALL(hipFree(ds));
  HIP_CALL(hipFree(dX));
  return 0;
}

extern "C" unsigned hj_visit_record_size(void) { return (unsigned)sizeof(VisitRec); }
