// spray_surrogate.hip -- training data, training and inference of the
// learned surrogate of the substep rates (surrogate.h).
//
//  * surr_sample_kernel: samples the hull with a counter-based generator and
//    evaluates the plain FP64 substep (substep_rates, the code of the FP64
//    update kernel; this unit is compiled without contraction like that one)
//    to produce (inputs, targets) pairs.
//  * mlp_fwd_kernel / mlp_fwdbwd_kernel / mlp_dw_kernel / mlp_adam_kernel:
//    FP32 forward pass, per-sample back-propagation, deterministic
//    batch reduction of the weight gradients (fixed chunking, FP64 sums in a
//    fixed order) and the Adam update.
//  * surr_infer_kernel: the bfloat16 MFMA inference of surrogate_mlp.h on
//    a flat array of inputs.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "spray_ctx.h"
#include "surrogate_mlp.h"

namespace pelemp {

namespace {

// ---- counter-based uniform numbers in (0, 1) ----------------------------
__device__ inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ inline double uniform01(uint64_t seed, long long i, int j) {
  const uint64_t r = splitmix64(seed ^ splitmix64(static_cast<uint64_t>(i) * 16u + static_cast<uint64_t>(j)));
  return (static_cast<double>(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

template <int NF>
struct SampleArgs {
  FuelConst<double, NF> fc;
  pelemp_surrogate_norm hull;
  long long n;
  uint64_t seed;
  float* x;
  float* t;
  float* w;
};

// Air: O2 0.233, N2 0.767 by mass (species 0 and 1 of every mechanism).
template <class M, int NF>
__global__ void __launch_bounds__(256) surr_sample_kernel(SampleArgs<NF> a) {
  constexpr int NS = M::NS;
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const pelemp_surrogate_norm& H = a.hull;
  const double u0 = uniform01(a.seed, i, 0), u1 = uniform01(a.seed, i, 1), u2 = uniform01(a.seed, i, 2);
  const double u3 = uniform01(a.seed, i, 3), u4 = uniform01(a.seed, i, 4), u5 = uniform01(a.seed, i, 5);
  const double u6 = uniform01(a.seed, i, 6);
  const double ld0 = rm::log(static_cast<double>(H.d_min)), ld1 = rm::log(static_cast<double>(H.d_max));
  const double dia = rm::exp(ld0 + u0 * (ld1 - ld0));
  const double Tg = H.Tg_min + u1 * (static_cast<double>(H.Tg_max) - H.Tg_min);
  const double lp0 = rm::log(static_cast<double>(H.p_min)), lp1 = rm::log(static_cast<double>(H.p_max));
  const double p = rm::exp(lp0 + u2 * (lp1 - lp0));
  const double urel = H.urel_max * u3 * u3;
  const double Yfg = H.Yfg_max * u4 * u4 * u4;
  double Yl[NF];
  if constexpr (NF == 1) {
    Yl[0] = 1.0;
  } else {
    Yl[0] = u5;
    Yl[1] = 1.0 - u5;
#pragma unroll
    for (int k = 2; k < NF; ++k) Yl[k] = 0.0;
  }
  GasAt<double, NS> g;
#pragma unroll
  for (int n = 0; n < NS; ++n) g.Y[n] = 0.0;
  g.Y[0] = 0.233 * (1.0 - Yfg);
  g.Y[1] = 0.767 * (1.0 - Yfg);
#pragma unroll
  for (int k = 0; k < NF; ++k) g.Y[fuel_species<NS, NF>(k)] = Yfg * Yl[k];
  double s = 0.0;
#pragma unroll
  for (int n = 0; n < NS; ++n) s += g.Y[n] / M::W[n];
  g.T = Tg;
  g.rho = p / (kRU * s * Tg);
  g.vel[0] = urel;
  g.vel[1] = 0.0;
  g.vel[2] = 0.0;
  gas_define<M, double>(g);
  double cB[NF];
  calc_boil_t<NF, double>(a.fc, g.p, cB);
  double Tb = 0.0, rinv = 0.0, mwinv = 0.0;
#pragma unroll
  for (int k = 0; k < NF; ++k) Tb += Yl[k] * cB[k];
  const double Thi = H.Td_frac_max * Tb;
  const double Td = H.Td_min + u6 * (Thi - H.Td_min);
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    rinv += Yl[k] / rhoL<double, NF>(a.fc, rmin(Td, cB[k]), k);
    mwinv += Yl[k] / a.fc.mw_fuel[k];
  }
  const double rho_part = 1.0 / rinv;
  const double pmass = (kPIc / 6.) * rho_part * (dia * dia * dia);
  const double vel0[3] = {0.0, 0.0, 0.0};
  SubRates<M, NF, double> o;
  substep_rates<M, NF, double>(g, a.fc, Td, dia, pmass, rho_part, Yl, vel0, 0.0, cB, o);
  VaporLite<double, NF> v;
  vapor_lite<M, NF, double>(a.fc, g, Td, 1.0 / mwinv, Yl, cB, v);
  float x[kSurrIn];
  surr_features<double>(dia, Td, urel, g.T, g.p, v.Ysk_f, Yl[0], NF > 1 ? Yl[1] : 0.0, x);
  bool ok = v.ok && Thi > H.Td_min && o.rhoD > 0. && o.lamNu > 0. && o.mu_skin > 0. &&
            surr_in_hull<double>(H, dia, Td, Tb, urel, g.T, g.p, v.Yfg, v.B_M);
  const double t0 = ok ? rm::log(o.rhoD) : 0.0;
  const double t1 = ok ? rm::log(o.lamNu) : 0.0;
  const double t2 = ok ? rm::log(o.mu_skin) : 0.0;
  ok = ok && (t0 == t0) && (t1 == t1) && (t2 == t2);
#pragma unroll
  for (int q = 0; q < kSurrIn; ++q) a.x[i * kSurrIn + q] = x[q];
  a.t[i * 3 + 0] = ok ? static_cast<float>(t0) : 0.f;
  a.t[i * 3 + 1] = ok ? static_cast<float>(t1) : 0.f;
  a.t[i * 3 + 2] = ok ? static_cast<float>(t2) : 0.f;
  a.w[i] = ok ? 1.f : 0.f;
}

// ---- FP32 training kernels -------------------------------------------------
constexpr int kMlpThreads = 128;
constexpr int kMaxBatch = 65536;
constexpr int kDwChunk = 256;  // samples per partial sum of the weight gradient

struct MlpNorm {
  float c[kSurrIn], s[kSurrIn], mean[kSurrOut], inv_std[kSurrOut];
};

inline MlpNorm make_mlp_norm(const pelemp_surrogate_norm& n) {
  MlpNorm m;
  for (int q = 0; q < kSurrIn; ++q) {
    m.c[q] = n.in_center[q];
    m.s[q] = n.in_scale[q];
  }
  for (int q = 0; q < kSurrOut; ++q) {
    m.mean[q] = n.out_mean[q];
    m.inv_std[q] = 1.f / n.out_std[q];
  }
  return m;
}

__device__ inline float qbf(float v, bool quant) {
  return quant ? static_cast<float>(static_cast<__bf16>(v)) : v;
}

// parameters into LDS; weights rounded to bfloat16 when quant
__device__ inline void stage_params(float* P, const float* __restrict__ src, bool quant) {
  for (int i = threadIdx.x; i < kSurrParams; i += blockDim.x) {
    float v = src[i];
    bool is_w = false;
#pragma unroll
    for (int L = 0; L < kSurrLin; ++L)
      if (i >= surr_w_off(L) && i < surr_b_off(L)) is_w = true;
    P[i] = (quant && is_w) ? qbf(v, true) : v;
  }
  __syncthreads();
}

// Forward pass of one sample; if ws_h / ws_d are given, stores each linear
// layer's input (ws_h) and the activation's derivative at each hidden layer
// (ws_d).
__device__ inline void mlp_forward_one(const float* P, const float* xn, bool quant, float* y, float* ws_h0,
                                       float* ws_h, float* ws_d, long long j, int nb) {
  float hin[kSurrHid];
#pragma unroll
  for (int i = 0; i < kSurrIn; ++i) {
    hin[i] = qbf(xn[i], quant);
    if (ws_h0) ws_h0[j * kSurrIn + i] = hin[i];
  }
  int nin = kSurrIn;
  for (int L = 0; L < 4; ++L) {
    float hout[kSurrHid];
    const float* W = P + surr_w_off(L);
    const float* b = P + surr_b_off(L);
#pragma unroll 4
    for (int o = 0; o < kSurrHid; ++o) {
      float z = b[o];
      if (nin == kSurrIn) {
#pragma unroll
        for (int i = 0; i < kSurrIn; ++i) z += W[o * kSurrIn + i] * hin[i];
      } else {
#pragma unroll
        for (int i = 0; i < kSurrHid; ++i) z += W[o * kSurrHid + i] * hin[i];
      }
      // hard swish z clamp(z/6 + 1/2, 0, 1), derivative 0 / z/3 + 1/2 / 1
      const float a = surr_act(z);
      if (ws_d)
        ws_d[(static_cast<long long>(L) * nb + j) * kSurrHid + o] =
            (z <= -3.f) ? 0.f : ((z >= 3.f) ? 1.f : z * (1.f / 3.f) + 0.5f);
      hout[o] = qbf(a, quant);
    }
#pragma unroll
    for (int o = 0; o < kSurrHid; ++o) {
      hin[o] = hout[o];
      if (ws_h) ws_h[(static_cast<long long>(L) * nb + j) * kSurrHid + o] = hout[o];
    }
    nin = kSurrHid;
  }
  const float* W = P + surr_w_off(4);
  const float* b = P + surr_b_off(4);
#pragma unroll
  for (int q = 0; q < kSurrOut; ++q) {
    float z = b[q];
#pragma unroll
    for (int i = 0; i < kSurrHid; ++i) z += W[q * kSurrHid + i] * hin[i];
    y[q] = z;
  }
}

__global__ void __launch_bounds__(kMlpThreads) mlp_fwd_kernel(const float* __restrict__ params, MlpNorm nm,
                                                               int quant, long long n, const float* __restrict__ x,
                                                               float* __restrict__ y) {
  __shared__ float P[kSurrParams];
  stage_params(P, params, quant != 0);
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float xn[kSurrIn], out[kSurrOut];
#pragma unroll
  for (int q = 0; q < kSurrIn; ++q) xn[q] = (x[i * kSurrIn + q] - nm.c[q]) * nm.s[q];
  mlp_forward_one(P, xn, quant != 0, out, nullptr, nullptr, nullptr, 0, 0);
#pragma unroll
  for (int q = 0; q < kSurrOut; ++q) y[i * kSurrOut + q] = out[q];
}

// Batch of nb samples: sample j of the batch is data row (a*(b0+j) + c) mod n.
struct BatchMap {
  long long n, a, c, b0;
};
__device__ inline long long batch_row(const BatchMap& m, int j) {
  // a, c, b0 + j < n < 2^31: the product fits in 64 bits
  const unsigned long long k = static_cast<unsigned long long>(m.b0 + j) % static_cast<unsigned long long>(m.n);
  return static_cast<long long>((static_cast<unsigned long long>(m.a) * k + static_cast<unsigned long long>(m.c)) %
                                static_cast<unsigned long long>(m.n));
}

struct Workspace {
  float* h0;      // [nb][8]    input of layer 0
  float* h;       // [4][nb][64] input of layers 1..4
  float* d;       // [4][nb][64] act' then delta of hidden layers 0..3
  float* dout;    // [nb][3]    delta of the output layer
  double* loss;   // [nb]
  double* part;   // [nchunk][kSurrParams] partial gradient sums
};

__global__ void __launch_bounds__(kMlpThreads) mlp_fwdbwd_kernel(const float* __restrict__ params, MlpNorm nm,
                                                                  int quant, BatchMap bm, int nb,
                                                                  const float* __restrict__ x,
                                                                  const float* __restrict__ t,
                                                                  const float* __restrict__ w, Workspace ws) {
  __shared__ float P[kSurrParams];
  stage_params(P, params, quant != 0);
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nb) return;
  const long long r = batch_row(bm, j);
  float xn[kSurrIn], y[kSurrOut];
#pragma unroll
  for (int q = 0; q < kSurrIn; ++q) xn[q] = (x[r * kSurrIn + q] - nm.c[q]) * nm.s[q];
  mlp_forward_one(P, xn, quant != 0, y, ws.h0, ws.h, ws.d, j, nb);
  const float wj = w ? w[r] : 1.f;
  double l = 0.;
  float dl[kSurrOut];
#pragma unroll
  for (int q = 0; q < kSurrOut; ++q) {
    const float tn = (t[r * kSurrOut + q] - nm.mean[q]) * nm.inv_std[q];
    const float e = y[q] - tn;
    l += static_cast<double>(wj) * static_cast<double>(e) * static_cast<double>(e);
    dl[q] = 2.f * wj * e / static_cast<float>(nb);
    ws.dout[static_cast<long long>(j) * kSurrOut + q] = dl[q];
  }
  ws.loss[j] = l / nb;
  // back-propagation: g = W^T delta, delta_L = g * act'(z_L)
  float g[kSurrHid];
  {
    const float* W = P + surr_w_off(4);
#pragma unroll
    for (int i = 0; i < kSurrHid; ++i) {
      float s = 0.f;
#pragma unroll
      for (int q = 0; q < kSurrOut; ++q) s += W[q * kSurrHid + i] * dl[q];
      g[i] = s;
    }
  }
  for (int L = 3; L >= 0; --L) {
    float* dptr = ws.d + (static_cast<long long>(L) * nb + j) * kSurrHid;
    float dl_h[kSurrHid];
#pragma unroll
    for (int o = 0; o < kSurrHid; ++o) {
      dl_h[o] = g[o] * dptr[o];
      dptr[o] = dl_h[o];
    }
    if (L > 0) {
      const float* W = P + surr_w_off(L);
#pragma unroll
      for (int i = 0; i < kSurrHid; ++i) g[i] = 0.f;
#pragma unroll 4
      for (int o = 0; o < kSurrHid; ++o) {
        const float dv = dl_h[o];
#pragma unroll
        for (int i = 0; i < kSurrHid; ++i) g[i] += W[o * kSurrHid + i] * dv;
      }
    }
  }
}

// Partial sums of the weight gradients over chunks of kDwChunk samples:
// part[chunk][p] (FP64, samples summed in order).
__global__ void __launch_bounds__(256) mlp_dw_kernel(Workspace ws, int nb) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int chunk = blockIdx.y;
  if (p >= kSurrParams) return;
  int L = 0;
#pragma unroll
  for (int q = 1; q < kSurrLin; ++q)
    if (p >= surr_w_off(q)) L = q;
  const int nin = surr_fan_in(L), nout = surr_fan_out(L);
  const bool is_b = p >= surr_b_off(L);
  const int o = is_b ? (p - surr_b_off(L)) : (p - surr_w_off(L)) / nin;
  const int i = is_b ? 0 : (p - surr_w_off(L)) % nin;
  const float* dl = (L == 4) ? ws.dout : ws.d + static_cast<long long>(L) * nb * kSurrHid;
  const float* hi = (L == 0) ? ws.h0 : ws.h + static_cast<long long>(L - 1) * nb * kSurrHid;
  const int j0 = chunk * kDwChunk;
  const int j1 = (j0 + kDwChunk < nb) ? j0 + kDwChunk : nb;
  double s = 0.;
  if (is_b) {
    for (int j = j0; j < j1; ++j) s += static_cast<double>(dl[static_cast<long long>(j) * nout + o]);
  } else {
    for (int j = j0; j < j1; ++j)
      s += static_cast<double>(dl[static_cast<long long>(j) * nout + o]) *
           static_cast<double>(hi[static_cast<long long>(j) * nin + i]);
  }
  ws.part[static_cast<long long>(chunk) * kSurrParams + p] = s;
}

// Gradient = ordered sum of the chunk partials; optional Adam step.
__global__ void __launch_bounds__(256) mlp_adam_kernel(float* __restrict__ params, float* __restrict__ grad_out,
                                                       const double* __restrict__ part, int nchunk,
                                                       float* __restrict__ m1, float* __restrict__ m2, int do_step,
                                                       float lr_t, float b1, float b2, float eps_t) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kSurrParams) return;
  double s = 0.;
  for (int c = 0; c < nchunk; ++c) s += part[static_cast<long long>(c) * kSurrParams + p];
  const float g = static_cast<float>(s);
  if (grad_out) grad_out[p] = g;
  if (do_step) {
    const float a = b1 * m1[p] + (1.f - b1) * g;
    const float b = b2 * m2[p] + (1.f - b2) * g * g;
    m1[p] = a;
    m2[p] = b;
    params[p] -= lr_t * a / (sqrtf(b) + eps_t);
  }
}

// Ordered sum of the per-sample losses of a batch into *acc.
__global__ void __launch_bounds__(256) loss_sum_kernel(const double* __restrict__ loss, int nb, double* acc) {
  __shared__ double sh[256];
  double s = 0.;
  for (int j = threadIdx.x; j < nb; j += 256) s += loss[j];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (static_cast<int>(threadIdx.x) < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) *acc += sh[0];
}

// ---- MFMA inference ---------------------------------------------------------
__global__ void __launch_bounds__(256) surr_infer_kernel(const SurrWeights* __restrict__ Wg, MlpNorm nm, long long n,
                                                         const float* __restrict__ x, float* __restrict__ y) {
  __shared__ SurrWeights W;
  surr_stage(W, Wg, threadIdx.x, blockDim.x);
  __syncthreads();
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  float xn[kSurrIn], out[kSurrOut];
#pragma unroll
  for (int q = 0; q < kSurrIn; ++q) xn[q] = (i < n) ? (x[i * kSurrIn + q] - nm.c[q]) * nm.s[q] : 0.f;
  surr_mlp(W, xn, out);  // all lanes take part
  if (i < n) {
#pragma unroll
    for (int q = 0; q < kSurrOut; ++q) y[i * kSurrOut + q] = out[q];
  }
}

// One MFMA on row-major A (32x16) and B (16x32) with the assumed layout.
__global__ void __launch_bounds__(64) mfma_probe_kernel(const uint16_t* __restrict__ A, const uint16_t* __restrict__ B,
                                                        float* __restrict__ D) {
  const int l = threadIdx.x;
  const int h = l >> 5, c = l & 31;
  bf16x8_t a, b;
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    a[v] = __builtin_bit_cast(__bf16, A[c * 16 + 8 * h + v]);
    b[v] = __builtin_bit_cast(__bf16, B[(8 * h + v) * 32 + c]);
  }
  f32x16_t acc = {};
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
#pragma unroll
  for (int r = 0; r < 16; ++r) D[(8 * (r / 4) + 4 * h + (r % 4)) * 32 + c] = acc[r];
}

int hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    set_last_error(std::string(what) + ": " + hipGetErrorString(e));
    return PELEMP_E_HIP;
  }
  return PELEMP_OK;
}

template <class M, int NF>
int launch_sample(const pelemp_spray_data& sd, const pelemp_surrogate_norm& hull, long long n, uint64_t seed,
                  float* x, float* t, float* w, hipStream_t s) {
  std::vector<double> mw;
  for (int k = 0; k < M::NS; ++k) mw.push_back(M::W[k]);
  SampleArgs<NF> a;
  a.fc = make_fuel_const<double, NF>(sd, mw.data());
  a.hull = hull;
  a.n = n;
  a.seed = seed;
  a.x = x;
  a.t = t;
  a.w = w;
  const long long nb = (n + 255) / 256;
  hipLaunchKernelGGL((surr_sample_kernel<M, NF>), dim3(static_cast<unsigned>(nb)), dim3(256), 0, s, a);
  return hip_check(hipGetLastError(), "sample launch");
}

// device allocations freed on scope exit
struct DevScope {
  std::vector<void*> p;
  ~DevScope() {
    for (void* q : p) (void)hipFree(q);
  }
  template <class T>
  int get(T** out, size_t bytes) {
    void* q = nullptr;
    if (hipMalloc(&q, bytes) != hipSuccess) {
      set_last_error("device allocation failed");
      return PELEMP_E_NOMEM;
    }
    p.push_back(q);
    *out = static_cast<T*>(q);
    return PELEMP_OK;
  }
};

int alloc_ws(DevScope& ds, Workspace& ws, int nb) {
  int st = PELEMP_OK;
  const int nchunk = (nb + kDwChunk - 1) / kDwChunk;
  st |= ds.get(&ws.h0, sizeof(float) * static_cast<size_t>(nb) * kSurrIn);
  st |= ds.get(&ws.h, sizeof(float) * 4ull * nb * kSurrHid);
  st |= ds.get(&ws.d, sizeof(float) * 4ull * nb * kSurrHid);
  st |= ds.get(&ws.dout, sizeof(float) * static_cast<size_t>(nb) * kSurrOut);
  st |= ds.get(&ws.loss, sizeof(double) * static_cast<size_t>(nb));
  st |= ds.get(&ws.part, sizeof(double) * static_cast<size_t>(nchunk) * kSurrParams);
  return st;
}

int launch_fwdbwd(const float* params, const MlpNorm& nm, int quant, const BatchMap& bm, int nb, const float* x,
                  const float* t, const float* w, const Workspace& ws, hipStream_t s) {
  hipLaunchKernelGGL(mlp_fwdbwd_kernel, dim3((nb + kMlpThreads - 1) / kMlpThreads), dim3(kMlpThreads), 0, s, params,
                     nm, quant, bm, nb, x, t, w, ws);
  const int nchunk = (nb + kDwChunk - 1) / kDwChunk;
  hipLaunchKernelGGL(mlp_dw_kernel, dim3((kSurrParams + 255) / 256, nchunk), dim3(256), 0, s, ws, nb);
  return hip_check(hipGetLastError(), "mlp launch");
}

// FP64 words [p, p + n) set to zero (a kernel rather than a memset, so the
// call can sit in a captured graph).
__global__ void __launch_bounds__(64) zero_f64_kernel(double* __restrict__ p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0.;
}

}  // namespace

}  // namespace pelemp

using namespace pelemp;

// ---- workspace of the allocation-free loss/gradient call --------------------
// One statically sized workspace per device, for callers that may not call
// hipMalloc (graph capture, tight loops): the Workspace of a batch of at most
// kNoAllocMaxBatch samples plus the loss accumulator, each array aligned to
// 256 bytes.  Every call on a device shares that device's copy and leaves
// its contents behind; calls that overlap on one device race.
namespace {
constexpr int kNoAllocMaxBatch = PELEMP_MLP_NOALLOC_MAX_BATCH;
constexpr size_t kNoAllocAlign = 256;
constexpr size_t noalloc_round(size_t b) { return (b + kNoAllocAlign - 1) / kNoAllocAlign * kNoAllocAlign; }
constexpr size_t kNoAllocChunks = (kNoAllocMaxBatch + pelemp::kDwChunk - 1) / pelemp::kDwChunk;
constexpr size_t kNoAllocOffH = noalloc_round(sizeof(float) * kNoAllocMaxBatch * pelemp::kSurrIn);
constexpr size_t kNoAllocOffD = kNoAllocOffH + noalloc_round(sizeof(float) * 4 * kNoAllocMaxBatch * pelemp::kSurrHid);
constexpr size_t kNoAllocOffDout =
    kNoAllocOffD + noalloc_round(sizeof(float) * 4 * kNoAllocMaxBatch * pelemp::kSurrHid);
constexpr size_t kNoAllocOffLoss =
    kNoAllocOffDout + noalloc_round(sizeof(float) * kNoAllocMaxBatch * pelemp::kSurrOut);
constexpr size_t kNoAllocOffPart = kNoAllocOffLoss + noalloc_round(sizeof(double) * kNoAllocMaxBatch);
constexpr size_t kNoAllocOffAcc =
    kNoAllocOffPart + noalloc_round(sizeof(double) * kNoAllocChunks * pelemp::kSurrParams);
constexpr size_t kNoAllocBytes = kNoAllocOffAcc + noalloc_round(sizeof(double));
static_assert(kNoAllocBytes == PELEMP_MLP_NOALLOC_SCRATCH_BYTES, "header and workspace size disagree");
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kNoAllocMaxDevices] = {};

// Device address of the current device's workspace: looked up once per
// device (with that device current) and kept; the array's own name is a
// host-side stand-in and is never passed to a kernel.
int noalloc_scratch(unsigned char** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_check(e, "hipGetDevice");
  if (dev < 0 || dev >= kNoAllocMaxDevices) {
    set_last_error("device index beyond the workspace table");
    return PELEMP_E_ARG;
  }
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_check(e, "hipGetSymbolAddress(scratch)");
    if (!p) {
      set_last_error("hipGetSymbolAddress(scratch) returned no address");
      return PELEMP_E_HIP;
    }
    g_noalloc_ptr[dev] = p;
  }
  *out = static_cast<unsigned char*>(g_noalloc_ptr[dev]);
  return PELEMP_OK;
}
}  // namespace

extern "C" {

int pelemp_surrogate_sample(int32_t mech, const pelemp_spray_data* sd, int32_t apply_latent_shift,
                            const pelemp_surrogate_norm* hull, int64_t n, uint64_t seed, float* x, float* t, float* w,
                            void* stream) {
  if (!sd || !hull || n <= 0 || !x || !t || !w) {
    set_last_error("invalid argument");
    return PELEMP_E_ARG;
  }
  pelemp_spray_data d = *sd;
  const int ns = mech_ns(mech);
  if (ns < 0) {
    set_last_error("unknown mechanism");
    return PELEMP_E_ARG;
  }
  if (apply_latent_shift) {
    std::vector<double> hh(ns);
    mech_h(mech, sd->ref_T, hh.data());
    for (int i = 0; i < d.nfuel; ++i) d.latent[i] = sd->ref_latent[i] - hh[sd->indx[i]];
  }
  for (int i = 0; i < d.nfuel; ++i)
    if (d.indx[i] != ns - d.nfuel + i) {
      set_last_error("fuel species must be the last nfuel species of the mechanism");
      return PELEMP_E_BAD_CONFIG;
    }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (mech == PELEMP_MECH_DODECANE_AIR && d.nfuel == 1)
    return launch_sample<Mech_dodecane_air, 1>(d, *hull, n, seed, x, t, w, s);
  if (mech == PELEMP_MECH_DECANE_DODECANE_AIR && d.nfuel == 2)
    return launch_sample<Mech_decane_dodecane_air, 2>(d, *hull, n, seed, x, t, w, s);
  set_last_error("the surrogate supports one or two liquid species (mechanisms 0 and 1)");
  return PELEMP_E_BAD_CONFIG;
}

int pelemp_mlp_forward(const float* params, const pelemp_surrogate_norm* norm, int32_t quant, int64_t n,
                       const float* x, float* y, void* stream) {
  if (!params || !norm || n <= 0 || !x || !y) {
    set_last_error("invalid argument");
    return PELEMP_E_ARG;
  }
  const MlpNorm nm = make_mlp_norm(*norm);
  const long long nb = (n + kMlpThreads - 1) / kMlpThreads;
  hipLaunchKernelGGL(mlp_fwd_kernel, dim3(static_cast<unsigned>(nb)), dim3(kMlpThreads), 0,
                     static_cast<hipStream_t>(stream), params, nm, quant, static_cast<long long>(n), x, y);
  return hip_check(hipGetLastError(), "mlp forward launch");
}

int pelemp_mlp_loss_grad(const float* params, const pelemp_surrogate_norm* norm, int32_t quant, int64_t n,
                         const float* x, const float* t, const float* w, float* grad, double* loss, void* stream) {
  if (!params || !norm || n <= 0 || n > kMaxBatch || !x || !t || !grad || !loss) {
    set_last_error("invalid argument");
    return PELEMP_E_ARG;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = static_cast<int>(n);
  DevScope ds;
  Workspace ws;
  double* acc = nullptr;
  int st = alloc_ws(ds, ws, nb);
  st |= ds.get(&acc, sizeof(double));
  if (st != PELEMP_OK) return PELEMP_E_NOMEM;
  const MlpNorm nm = make_mlp_norm(*norm);
  BatchMap bm{n, 1, 0, 0};
  if ((st = hip_check(hipMemsetAsync(acc, 0, sizeof(double), s), "memset")) != PELEMP_OK) return st;
  if ((st = launch_fwdbwd(params, nm, quant, bm, nb, x, t, w, ws, s)) != PELEMP_OK) return st;
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((kSurrParams + 255) / 256), dim3(256), 0, s, nullptr, grad, ws.part,
                     (nb + kDwChunk - 1) / kDwChunk, nullptr, nullptr, 0, 0.f, 0.f, 0.f, 0.f);
  hipLaunchKernelGGL(loss_sum_kernel, dim3(1), dim3(256), 0, s, ws.loss, nb, acc);
  if ((st = hip_check(hipGetLastError(), "mlp launch")) != PELEMP_OK) return st;
  if ((st = hip_check(hipMemcpyAsync(loss, acc, sizeof(double), hipMemcpyDeviceToHost, s), "copy")) != PELEMP_OK)
    return st;
  return hip_check(hipStreamSynchronize(s), "sync");
}

int pelemp_mlp_loss_grad_noalloc(const float* params, const pelemp_surrogate_norm* norm, int32_t quant, int64_t n,
                                 const float* x, const float* t, const float* w, float* grad, double* loss,
                                 void* stream) {
  if (!params || !norm || n <= 0 || !x || !t || !grad || !loss) {
    set_last_error("invalid argument");
    return PELEMP_E_ARG;
  }
  if (n > kNoAllocMaxBatch) {
    set_last_error("pelemp_mlp_loss_grad_noalloc: batch of " + std::to_string(n) + " samples exceeds the " +
                   std::to_string(kNoAllocMaxBatch) + "-sample workspace");
    return PELEMP_E_ARG;
  }
  unsigned char* base = nullptr;
  int st = noalloc_scratch(&base);
  if (st != PELEMP_OK) return st;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = static_cast<int>(n);
  // the arrays are laid out for kNoAllocMaxBatch samples; a smaller batch
  // uses the front of each (the kernels index with nb)
  Workspace ws;
  ws.h0 = reinterpret_cast<float*>(base);
  ws.h = reinterpret_cast<float*>(base + kNoAllocOffH);
  ws.d = reinterpret_cast<float*>(base + kNoAllocOffD);
  ws.dout = reinterpret_cast<float*>(base + kNoAllocOffDout);
  ws.loss = reinterpret_cast<double*>(base + kNoAllocOffLoss);
  ws.part = reinterpret_cast<double*>(base + kNoAllocOffPart);
  double* acc = reinterpret_cast<double*>(base + kNoAllocOffAcc);
  const MlpNorm nm = make_mlp_norm(*norm);
  BatchMap bm{n, 1, 0, 0};
  // the accumulator keeps the previous call's loss: clear it first
  hipLaunchKernelGGL(zero_f64_kernel, dim3(1), dim3(64), 0, s, acc, 1);
  if ((st = hip_check(hipGetLastError(), "zero launch")) != PELEMP_OK) return st;
  if ((st = launch_fwdbwd(params, nm, quant, bm, nb, x, t, w, ws, s)) != PELEMP_OK) return st;
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((kSurrParams + 255) / 256), dim3(256), 0, s, nullptr, grad, ws.part,
                     (nb + kDwChunk - 1) / kDwChunk, nullptr, nullptr, 0, 0.f, 0.f, 0.f, 0.f);
  hipLaunchKernelGGL(loss_sum_kernel, dim3(1), dim3(256), 0, s, ws.loss, nb, acc);
  if ((st = hip_check(hipGetLastError(), "mlp launch")) != PELEMP_OK) return st;
  if ((st = hip_check(hipMemcpyAsync(loss, acc, sizeof(double), hipMemcpyDeviceToHost, s), "copy")) != PELEMP_OK)
    return st;
  return hip_check(hipStreamSynchronize(s), "sync");
}

int pelemp_mlp_train(float* params, const pelemp_surrogate_norm* norm, int64_t n, const float* x, const float* t,
                     const float* w, const pelemp_mlp_train_cfg* cfg, double* epoch_loss, void* stream) {
  if (!params || !norm || !cfg || n <= 0 || n >= (1LL << 31) || !x || !t || cfg->epochs <= 0 || cfg->batch <= 0 ||
      cfg->batch > kMaxBatch || cfg->batch > n) {
    set_last_error("invalid argument");
    return PELEMP_E_ARG;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = cfg->batch;
  const int nchunk = (nb + kDwChunk - 1) / kDwChunk;
  DevScope ds;
  Workspace ws;
  float *m1 = nullptr, *m2 = nullptr;
  double* acc = nullptr;
  int st = alloc_ws(ds, ws, nb);
  st |= ds.get(&m1, sizeof(float) * kSurrParams);
  st |= ds.get(&m2, sizeof(float) * kSurrParams);
  st |= ds.get(&acc, sizeof(double) * static_cast<size_t>(cfg->epochs));
  if (st != PELEMP_OK) return PELEMP_E_NOMEM;
  st |= hip_check(hipMemsetAsync(m1, 0, sizeof(float) * kSurrParams, s), "memset");
  st |= hip_check(hipMemsetAsync(m2, 0, sizeof(float) * kSurrParams, s), "memset");
  st |= hip_check(hipMemsetAsync(acc, 0, sizeof(double) * cfg->epochs, s), "memset");
  if (st != PELEMP_OK) return PELEMP_E_HIP;
  const MlpNorm nm = make_mlp_norm(*norm);
  const long long nbatch = n / nb;  // the tail of each epoch's order is left out
  const long long total = nbatch * cfg->epochs;
  uint64_t rs = cfg->seed * 0x9E3779B97F4A7C15ull + 12345u;
  auto next = [&]() {
    rs ^= rs << 13;
    rs ^= rs >> 7;
    rs ^= rs << 17;
    return rs;
  };
  long long step = 0;
  for (int e = 0; e < cfg->epochs; ++e) {
    // a random multiplier coprime with n and a random offset: a new order each epoch
    long long a = 1;
    for (int tries = 0; tries < 64; ++tries) {
      const long long cand = static_cast<long long>(next() % static_cast<uint64_t>(n));
      if (cand > 0 && std::gcd(cand, static_cast<long long>(n)) == 1) {
        a = cand;
        break;
      }
    }
    const long long c = static_cast<long long>(next() % static_cast<uint64_t>(n));
    for (long long b = 0; b < nbatch; ++b) {
      ++step;
      const double frac = static_cast<double>(step - 1) / static_cast<double>(total > 1 ? total - 1 : 1);
      const double lr = cfg->lr_final + 0.5 * (cfg->lr - cfg->lr_final) * (1.0 + std::cos(kPIc * frac));
      const double bc1 = 1.0 - std::pow(cfg->beta1, static_cast<double>(step));
      const double bc2 = 1.0 - std::pow(cfg->beta2, static_cast<double>(step));
      const float lr_t = static_cast<float>(lr * std::sqrt(bc2) / bc1);
      const float eps_t = static_cast<float>(cfg->eps * std::sqrt(bc2));
      BatchMap bm{n, a, c, b * nb};
      if ((st = launch_fwdbwd(params, nm, cfg->quant, bm, nb, x, t, w, ws, s)) != PELEMP_OK) return st;
      hipLaunchKernelGGL(mlp_adam_kernel, dim3((kSurrParams + 255) / 256), dim3(256), 0, s, params, nullptr, ws.part,
                         nchunk, m1, m2, 1, lr_t, static_cast<float>(cfg->beta1), static_cast<float>(cfg->beta2),
                         eps_t);
      hipLaunchKernelGGL(loss_sum_kernel, dim3(1), dim3(256), 0, s, ws.loss, nb, acc + e);
      if ((st = hip_check(hipGetLastError(), "train launch")) != PELEMP_OK) return st;
    }
    if (cfg->verbose) {
      double le = 0.;
      if ((st = hip_check(hipMemcpyAsync(&le, acc + e, sizeof(double), hipMemcpyDeviceToHost, s), "copy")) !=
          PELEMP_OK)
        return st;
      if ((st = hip_check(hipStreamSynchronize(s), "sync")) != PELEMP_OK) return st;
      std::fprintf(stderr, "# epoch %d/%d: loss %.4e\n", e + 1, cfg->epochs, le / static_cast<double>(nbatch));
      std::fflush(stderr);
    }
  }
  std::vector<double> h(cfg->epochs);
  if ((st = hip_check(hipMemcpyAsync(h.data(), acc, sizeof(double) * cfg->epochs, hipMemcpyDeviceToHost, s),
                      "copy")) != PELEMP_OK)
    return st;
  if ((st = hip_check(hipStreamSynchronize(s), "sync")) != PELEMP_OK) return st;
  if (epoch_loss)
    for (int e = 0; e < cfg->epochs; ++e) epoch_loss[e] = h[e] / static_cast<double>(nbatch);
  return PELEMP_OK;
}

int pelemp_mlp_infer_bf16(const float* params_host, const pelemp_surrogate_norm* norm, int64_t n, const float* x,
                          float* y, void* stream) {
  if (!params_host || !norm || n <= 0 || !x || !y) {
    set_last_error("invalid argument");
    return PELEMP_E_ARG;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  std::vector<SurrWeights> host(1);
  surr_pack(params_host, host[0]);
  DevScope ds;
  SurrWeights* dW = nullptr;
  int st = ds.get(&dW, sizeof(SurrWeights));
  if (st != PELEMP_OK) return st;
  if ((st = hip_check(hipMemcpyAsync(dW, host.data(), sizeof(SurrWeights), hipMemcpyHostToDevice, s), "copy")) !=
      PELEMP_OK)
    return st;
  const MlpNorm nm = make_mlp_norm(*norm);
  const long long nb = (n + 255) / 256;
  hipLaunchKernelGGL(surr_infer_kernel, dim3(static_cast<unsigned>(nb)), dim3(256), 0, s, dW, nm,
                     static_cast<long long>(n), x, y);
  if ((st = hip_check(hipGetLastError(), "infer launch")) != PELEMP_OK) return st;
  // the packed weights are freed on return: wait for the kernel
  return hip_check(hipStreamSynchronize(s), "sync");
}

int pelemp_mfma_probe(const uint16_t* a_host, const uint16_t* b_host, float* d_host) {
  if (!a_host || !b_host || !d_host) {
    set_last_error("invalid argument");
    return PELEMP_E_ARG;
  }
  DevScope ds;
  uint16_t *A = nullptr, *B = nullptr;
  float* D = nullptr;
  int st = ds.get(&A, 2 * 32 * 16);
  st |= ds.get(&B, 2 * 16 * 32);
  st |= ds.get(&D, 4 * 32 * 32);
  if (st != PELEMP_OK) return PELEMP_E_NOMEM;
  st |= hip_check(hipMemcpy(A, a_host, 2 * 32 * 16, hipMemcpyHostToDevice), "copy");
  st |= hip_check(hipMemcpy(B, b_host, 2 * 16 * 32, hipMemcpyHostToDevice), "copy");
  if (st != PELEMP_OK) return PELEMP_E_HIP;
  hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, 0, A, B, D);
  if ((st = hip_check(hipGetLastError(), "probe launch")) != PELEMP_OK) return st;
  return hip_check(hipMemcpy(d_host, D, 4 * 32 * 32, hipMemcpyDeviceToHost), "copy");
}

}  // extern "C"
