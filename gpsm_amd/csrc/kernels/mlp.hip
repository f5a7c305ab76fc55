// Learned join-order planner network on gfx950: MLP 32 -> 128 -> 128 -> 1
// (ReLU), BF16 weights, FP32 accumulation, trained with Adam.  No reference
// counterpart (the reference orders joins with a fixed heuristic,
// match/Match.cu:1599-1661).
//
// One wave handles 32 samples of the forward pass and of the backward pass
// down to the deltas; the two weight gradients that are GEMMs over the
// batch (W2's and W1's) run in a second kernel as split-K GEMMs from
// operands the first writes feature-major, so that each workgroup adds its
// partial gradients to memory once instead of once per 32 samples.  Both
// hidden layers, the backward product d2 W2 and the weight gradients run on
// v_mfma_f32_32x32x16_bf16; the output layer (a 128-wide dot product per
// sample) and the bias gradients are FP32 FMAs.  BF16 rounding points (round to nearest even): inputs, weights,
// hidden activations h1 (operand of layer 2), and the deltas d2, d1 where
// they are GEMM operands.  Accumulation order differs from a sequential sum,
// so results agree with an exact evaluation to FP32 rounding.
//
// 32x32x16 MFMA operand layout (64 lanes, l = lane):
//   A (32 x 16): lane l holds row l%32, k = 8*(l/32) + 0..7
//   B (16 x 32): lane l holds col l%32, k = 8*(l/32) + 0..7
//   C (32 x 32): lane l, register r holds row 8*(r/4) + 4*(l/32) + r%4, col l%32
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "api_internal.h"
#include "rng.h"
#include "wave.h"

using namespace gpsm;

namespace {

constexpr int F = GPSM_FEAT, H = GPSM_HIDDEN;
constexpr int OW1 = 0, OB1 = H * F, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + H;
constexpr int NP = OB3 + 1;
constexpr int XS = F + 8;  // LDS row strides (bf16), padded against bank conflicts
constexpr int HS = H + 8;
static_assert(XS % 8 == 0 && HS % 8 == 0 && (32 * XS) % 8 == 0 && (32 * HS) % 8 == 0, "16-byte aligned operand runs");

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t bf_bits(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ __bf16 to_bf(float f) { return __builtin_bit_cast(__bf16, bf_bits(f)); }
__device__ __forceinline__ float from_bf(__bf16 h) {
  return __builtin_bit_cast(float, (uint32_t)__builtin_bit_cast(uint16_t, h) << 16);
}
__device__ __forceinline__ float round_bf(float f) { return from_bf(to_bf(f)); }

__device__ __forceinline__ f32x16 mfma(const bf16x8& a, const bf16x8& b, const f32x16& c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ uint32_t c_row(uint32_t r, uint32_t lane) { return 8 * (r / 4) + 4 * (lane / 32) + (r % 4); }

// Eight consecutive BF16 values (an MFMA operand's k run) in one 16-byte
// read: every use is 16-byte aligned (row strides XS, HS, F and H are
// multiples of 8 values, k offsets multiples of 8).
__device__ __forceinline__ bf16x8 ld8(const __bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }

struct Weights {
  const __bf16* w1;   // [H][F]
  const __bf16* w2;   // [H][H]
  const __bf16* w2t;  // [H][H] transposed
  const float* master;
  const float* norm;  // input standardisation: mu[F], inv_sd[F], applied in FP32 before BF16 rounding
};

struct alignas(16) TileLds {  // 16-byte aligned rows for ld8 (each member a multiple of 16 bytes)
  __bf16 xs[32 * XS];
  __bf16 h1[32 * HS];
  __bf16 d[32 * HS];
  float y[32];
};

// Forward of 32 samples; returns z2 (pre-activation, bias added) in regs.
__device__ void forward_tile(const Weights& W, const float* __restrict__ x, uint32_t b0, uint32_t B, TileLds& L,
                             f32x16 z2[4]) {
  const uint32_t lane = lane_id();
  // inputs -> bf16 in LDS (zeros past B)
  for (uint32_t i = lane; i < 32 * F; i += WAVE) {
    const uint32_t s = i / F, j = i % F;
    const float v = (b0 + s < B) ? (x[(size_t)(b0 + s) * F + j] - W.norm[j]) * W.norm[F + j] : 0.0f;
    L.xs[s * XS + j] = to_bf(v);
  }
  wave_sync();
  const uint32_t rl = lane % 32, kh = 8 * (lane / 32);
  // layer 1: z1 = x W1^T + b1, h1 = bf16(relu(z1))
  for (int t = 0; t < H / 32; ++t) {
    f32x16 acc = {};
    for (int s = 0; s < F / 16; ++s)
      acc = mfma(ld8(&L.xs[rl * XS + 16 * s + kh]), ld8(&W.w1[(32 * t + rl) * F + 16 * s + kh]), acc);
    const uint32_t colh = 32 * t + rl;
    const float bias = W.master[OB1 + colh];
    for (int r = 0; r < 16; ++r) L.h1[c_row(r, lane) * HS + colh] = to_bf(fmaxf(acc[r] + bias, 0.0f));
  }
  wave_sync();
  // layer 2: z2 = h1 W2^T + b2
  for (int t = 0; t < H / 32; ++t) {
    f32x16 acc = {};
    for (int s = 0; s < H / 16; ++s)
      acc = mfma(ld8(&L.h1[rl * HS + 16 * s + kh]), ld8(&W.w2[(32 * t + rl) * H + 16 * s + kh]), acc);
    const float bias = W.master[OB2 + 32 * t + rl];
    for (int r = 0; r < 16; ++r) acc[r] += bias;
    z2[t] = acc;
  }
  // layer 3: y = relu(z2) . bf16(W3) + b3; per-lane partial sums over its
  // 32 columns, then a butterfly over the 32 lanes sharing those rows
  float part[16];
  for (int r = 0; r < 16; ++r) part[r] = 0.0f;
  for (int t = 0; t < H / 32; ++t) {
    const float w3 = round_bf(W.master[OW3 + 32 * t + rl]);
    for (int r = 0; r < 16; ++r) part[r] += fmaxf(z2[t][r], 0.0f) * w3;
  }
  for (int r = 0; r < 16; ++r)
    for (int d = 1; d < 32; d <<= 1) part[r] += __shfl_xor(part[r], d, WAVE);
  if (rl == 0)
    for (int r = 0; r < 16; ++r) L.y[c_row(r, lane)] = part[r] + W.master[OB3];
  wave_sync();
}

__global__ void __launch_bounds__(64) mlp_forward_kernel(Weights W, const float* __restrict__ x, uint32_t B,
                                                         float* __restrict__ y) {
  __shared__ TileLds L;
  const uint32_t b0 = blockIdx.x * 32;
  f32x16 z2[4];
  forward_tile(W, x, b0, B, L, z2);
  const uint32_t lane = lane_id();
  if (lane < 32 && b0 + lane < B) y[b0 + lane] = L.y[lane];
}

// The operands of the weight-gradient GEMMs over the batch padded to Bp
// samples (a multiple of 32): d2 (output-layer deltas, H features), h1
// (hidden activations, H), d1 (first-layer deltas, H) and x (standardised
// inputs, F), all BF16, one after the other in act, each in blocks of 16
// samples: value (feature i, sample b) of an n-feature matrix at
// opnd(n, i, b) = ((b / 16) * n + i) * 16 + b % 16.  One MFMA operand (32
// features x 16 samples) is then 1 KB contiguous, one 16-byte load per lane.
__device__ __forceinline__ size_t opnd(uint32_t n, uint32_t i, uint32_t b) {
  return ((size_t)(b >> 4) * n + i) * 16 + (b & 15);
}
__device__ __forceinline__ void put_rows(__bf16* __restrict__ dst, uint32_t b0, const __bf16* src, uint32_t stride,
                                         uint32_t nrows) {
  // feature i's 32 samples of the tile: four runs of 8, one 16-byte store each
  for (uint32_t i = lane_id(); i < nrows; i += WAVE)
    for (uint32_t c = 0; c < 4; ++c) {
      bf16x8 v;
      for (uint32_t j = 0; j < 8; ++j) v[j] = src[(8 * c + j) * stride + i];
      *reinterpret_cast<bf16x8*>(dst + opnd(nrows, i, b0 + 8 * c)) = v;
    }
}

// Forward + backward of 32 samples of the loss 0.5*mean((y-t)^2): the loss,
// the bias gradients and W3's are added atomically into grad (FP32,
// parameter order); the operands of W2's and W1's gradients are written to
// act for mlp_wgrad_kernel.  (A wave looping over several tiles with these
// sums in registers needs 256 VGPRs, one wave per SIMD, and was slower.)
__global__ void __launch_bounds__(64) mlp_grad_kernel(Weights W, const float* __restrict__ x,
                                                      const float* __restrict__ tgt, uint32_t B, uint32_t Bp,
                                                      float* __restrict__ grad, float* __restrict__ loss,
                                                      __bf16* __restrict__ act) {
  __shared__ TileLds L;
  __shared__ float dys[32];
  const uint32_t lane = lane_id(), rl = lane % 32, kh = 8 * (lane / 32);
  const uint32_t b0 = blockIdx.x * 32;
  f32x16 z2[4];
  forward_tile(W, x, b0, B, L, z2);
  // dL/dy and the loss
  float lpart = 0.0f;
  if (lane < 32) {
    float dy = 0.0f;
    if (b0 + lane < B) {
      const float e = L.y[lane] - tgt[b0 + lane];
      dy = e / (float)B;
      lpart = 0.5f * e * e / (float)B;
    }
    dys[lane] = dy;
  }
  wave_sync();
  lpart = wave_sum_f(lpart);
  float gb3 = lane < 32 ? dys[lane] : 0.0f;
  gb3 = wave_sum_f(gb3);
  if (lane == 0) {
    atomicAdd(loss, lpart);
    atomicAdd(&grad[OB3], gb3);
  }
  // output layer: gW3 = sum_b dy h2; d2 = [z2 > 0] dy bf16(W3); gb2 = sum_b d2
  for (int t = 0; t < H / 32; ++t) {
    const uint32_t colh = 32 * t + rl;
    const float w3 = round_bf(W.master[OW3 + colh]);
    float gw3 = 0.0f, gb2 = 0.0f;
    for (int r = 0; r < 16; ++r) {
      const uint32_t row = c_row(r, lane);
      const float dy = dys[row];
      const float z = z2[t][r];
      gw3 += dy * fmaxf(z, 0.0f);
      const float d2 = z > 0.0f ? dy * w3 : 0.0f;
      gb2 += d2;
      L.d[row * HS + colh] = to_bf(d2);
    }
    gw3 += __shfl_xor(gw3, 32, WAVE);
    gb2 += __shfl_xor(gb2, 32, WAVE);
    if (lane < 32) {
      atomicAdd(&grad[OW3 + colh], gw3);
      atomicAdd(&grad[OB2 + colh], gb2);
    }
  }
  wave_sync();
  // the operands of gW2 = d2^T h1 and gW1 = d1^T x (mlp_wgrad_kernel)
  put_rows(act, b0, L.d, HS, H);
  put_rows(act + (size_t)H * Bp, b0, L.h1, HS, H);
  put_rows(act + (size_t)3 * H * Bp, b0, L.xs, XS, F);
  // dh1[b][j] = sum_i d2[b][i] W2[i][j];  d1 = [h1 > 0] dh1
  f32x16 dh[4];
  for (int t = 0; t < H / 32; ++t) {
    f32x16 acc = {};
    for (int s = 0; s < H / 16; ++s)
      acc = mfma(ld8(&L.d[rl * HS + 16 * s + kh]), ld8(&W.w2t[(32 * t + rl) * H + 16 * s + kh]), acc);
    dh[t] = acc;
  }
  wave_sync();
  for (int t = 0; t < H / 32; ++t) {
    const uint32_t colh = 32 * t + rl;
    float gb1 = 0.0f;
    for (int r = 0; r < 16; ++r) {
      const uint32_t row = c_row(r, lane);
      const float d1 = from_bf(L.h1[row * HS + colh]) > 0.0f ? dh[t][r] : 0.0f;
      gb1 += d1;
      L.d[row * HS + colh] = to_bf(d1);
    }
    gb1 += __shfl_xor(gb1, 32, WAVE);
    if (lane < 32) atomicAdd(&grad[OB1 + colh], gb1);
  }
  wave_sync();
  put_rows(act + (size_t)2 * H * Bp, b0, L.d, HS, H);
}

// Weight gradients from the operands mlp_grad_kernel wrote: gW2[i][j] = sum_b d2[b][i] h1[b][j] and gW1[i][j] =
// sum_b d1[b][i] x[b][j] over the samples [k0, k1) of this workgroup (split
// K): wave w owns rows 32w..32w+31 of both, i.e. four 32x32 tiles of gW2
// and one of gW1, in MFMA accumulators over the chunk, 16 samples per step
// (each operand one 16-byte load per lane), then added to grad once.
__global__ void __launch_bounds__(256) mlp_wgrad_kernel(const __bf16* __restrict__ act, uint32_t Bp, uint32_t KC,
                                                        float* __restrict__ grad) {
  const uint32_t lane = lane_id(), w = threadIdx.x / WAVE, rl = lane % 32, kh = 8 * (lane / 32);
  const uint32_t k0 = blockIdx.x * KC, k1 = min(Bp, k0 + KC);
  const __bf16* d2t = act;
  const __bf16* h1t = act + (size_t)H * Bp;
  const __bf16* d1t = act + (size_t)2 * H * Bp;
  const __bf16* xt = act + (size_t)3 * H * Bp;
  f32x16 acc2[4] = {}, acc1 = {};
  for (uint32_t k = k0; k < k1; k += 16) {
    const bf16x8 a2 = ld8(d2t + opnd(H, 32 * w + rl, k + kh));
    const bf16x8 a1 = ld8(d1t + opnd(H, 32 * w + rl, k + kh));
    const bf16x8 bx = ld8(xt + opnd(F, rl, k + kh));
    bf16x8 bh[4];
    for (int jt = 0; jt < 4; ++jt) bh[jt] = ld8(h1t + opnd(H, 32 * jt + rl, k + kh));
    for (int jt = 0; jt < 4; ++jt) acc2[jt] = mfma(a2, bh[jt], acc2[jt]);
    acc1 = mfma(a1, bx, acc1);
  }
  for (int jt = 0; jt < 4; ++jt)
    for (int r = 0; r < 16; ++r) atomicAdd(&grad[OW2 + (32 * w + c_row(r, lane)) * H + 32 * jt + rl], acc2[jt][r]);
  for (int r = 0; r < 16; ++r) atomicAdd(&grad[OW1 + (32 * w + c_row(r, lane)) * F + rl], acc1[r]);
}

// BF16 copies of the weights (and W2 transposed) from the FP32 masters.
__global__ void refresh_kernel(const float* __restrict__ m, __bf16* __restrict__ w1, __bf16* __restrict__ w2,
                               __bf16* __restrict__ w2t) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < H * F) w1[i] = to_bf(m[OW1 + i]);
  if (i < H * H) {
    const uint32_t r = i / H, c = i % H;
    const __bf16 v = to_bf(m[OW2 + i]);
    w2[i] = v;
    w2t[c * H + r] = v;
  }
}

__global__ void adam_kernel(float* __restrict__ p, float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, float lr, float c1, float c2) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NP) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  g[i] = 0.0f;
}

}  // namespace

struct gpsm_mlp {
  gpsm_ctx* ctx = nullptr;
  float* p = nullptr;  // FP32 masters [NP]
  float* g = nullptr;  // gradients
  float* m = nullptr;
  float* v = nullptr;
  __bf16* w1 = nullptr;
  __bf16* w2 = nullptr;
  __bf16* w2t = nullptr;
  float* loss = nullptr;
  float* norm = nullptr;  // mu[F], inv_sd[F]
  DevBuf x, t, y;
  DevBuf act;  // the weight-gradient GEMMs' operands (mlp_grad_kernel)
  HostBuf hx, hy;  // pinned staging of forward's features and outputs
  int step = 0;
};

namespace {

Weights weights_of(const gpsm_mlp* mm) { return Weights{mm->w1, mm->w2, mm->w2t, mm->p, mm->norm}; }

int refresh(gpsm_mlp* mm) {
  hipLaunchKernelGGL(refresh_kernel, dim3((H * H + 255) / 256), dim3(256), 0, mm->ctx->stream, mm->p, mm->w1, mm->w2,
                     mm->w2t);
  return hipGetLastError() == hipSuccess ? GPSM_OK : GPSM_EDEVICE;
}

int stage(gpsm_mlp* mm, const float* x, const float* t, uint32_t B) {
  hipStream_t s = mm->ctx->stream;
  if (mm->x.ensure(sizeof(float) * (size_t)B * F) || mm->y.ensure(sizeof(float) * B) ||
      (t && mm->t.ensure(sizeof(float) * B)))
    return GPSM_ENOMEM;
  if (hipMemcpyAsync(mm->x.ptr, x, sizeof(float) * (size_t)B * F, hipMemcpyHostToDevice, s) != hipSuccess)
    return GPSM_EDEVICE;
  if (t && hipMemcpyAsync(mm->t.ptr, t, sizeof(float) * B, hipMemcpyHostToDevice, s) != hipSuccess)
    return GPSM_EDEVICE;
  return GPSM_OK;
}

int run_grad(gpsm_mlp* mm, uint32_t B) {
  hipStream_t s = mm->ctx->stream;
  const uint32_t Bp = (B + 31) / 32 * 32;
  if (mm->act.ensure(sizeof(__bf16) * (size_t)(3 * H + F) * Bp)) return GPSM_ENOMEM;
  __bf16* act = reinterpret_cast<__bf16*>(mm->act.ptr);
  hipMemsetAsync(mm->loss, 0, sizeof(float), s);
  hipLaunchKernelGGL(mlp_grad_kernel, dim3(Bp / 32), dim3(64), 0, s, weights_of(mm),
                     reinterpret_cast<const float*>(mm->x.ptr), reinterpret_cast<const float*>(mm->t.ptr), B, Bp,
                     mm->g, mm->loss, act);
  // split K: at most 64 workgroups of at least 256 samples (a multiple of 16)
  const uint32_t chunks = std::min<uint32_t>(64, std::max<uint32_t>(1, Bp / 256));
  const uint32_t KC = ((Bp + chunks - 1) / chunks + 15) / 16 * 16;
  hipLaunchKernelGGL(mlp_wgrad_kernel, dim3((Bp + KC - 1) / KC), dim3(256), 0, s, act, Bp, KC, mm->g);
  return hipGetLastError() == hipSuccess ? GPSM_OK : GPSM_EDEVICE;
}

}  // namespace

extern "C" {

uint64_t gpsm_mlp_num_params(void) { return NP; }

int gpsm_mlp_create(gpsm_ctx* c, uint64_t seed, gpsm_mlp** out) {
  if (!c || !out) return GPSM_EINVAL;
  *out = nullptr;
  hipSetDevice(c->device);
  gpsm_mlp* mm = new gpsm_mlp();
  mm->ctx = c;
  if (hipMalloc(&mm->p, sizeof(float) * NP) || hipMalloc(&mm->g, sizeof(float) * NP) ||
      hipMalloc(&mm->m, sizeof(float) * NP) || hipMalloc(&mm->v, sizeof(float) * NP) ||
      hipMalloc(&mm->w1, sizeof(__bf16) * H * F) || hipMalloc(&mm->w2, sizeof(__bf16) * H * H) ||
      hipMalloc(&mm->w2t, sizeof(__bf16) * H * H) || hipMalloc(&mm->loss, sizeof(float)) ||
      hipMalloc(&mm->norm, sizeof(float) * 2 * F)) {
    gpsm_mlp_free(mm);
    return GPSM_ENOMEM;
  }
  // He-style uniform initialisation from a SplitMix64 stream
  std::vector<float> p(NP, 0.0f);
  SplitMix64 rng(seed, 99);
  auto fill = [&](int off, int n, int fan_in) {
    const double a = std::sqrt(6.0 / fan_in);
    for (int i = 0; i < n; ++i) p[off + i] = (float)((2.0 * rng.uniform() - 1.0) * a);
  };
  fill(OW1, H * F, F);
  fill(OW2, H * H, H);
  fill(OW3, H, H);
  int rc = gpsm_mlp_set_input_norm(mm, nullptr, nullptr);  // identity
  if (!rc) rc = gpsm_mlp_set_params(mm, p.data());
  if (rc) {
    gpsm_mlp_free(mm);
    return rc;
  }
  *out = mm;
  return GPSM_OK;
}

void gpsm_mlp_free(gpsm_mlp* mm) {
  if (!mm) return;
  if (mm->ctx && mm->ctx->planner == mm) mm->ctx->planner = nullptr;
  hipFree(mm->p);
  hipFree(mm->g);
  hipFree(mm->m);
  hipFree(mm->v);
  hipFree(mm->w1);
  hipFree(mm->w2);
  hipFree(mm->w2t);
  hipFree(mm->loss);
  hipFree(mm->norm);
  mm->x.release();
  mm->hx.release();
  mm->hy.release();
  mm->t.release();
  mm->y.release();
  mm->act.release();
  delete mm;
}

int gpsm_mlp_set_params(gpsm_mlp* mm, const float* in) {
  if (!mm || !in) return GPSM_EINVAL;
  hipStream_t s = mm->ctx->stream;
  if (hipMemcpyAsync(mm->p, in, sizeof(float) * NP, hipMemcpyHostToDevice, s) ||
      hipMemsetAsync(mm->g, 0, sizeof(float) * NP, s) || hipMemsetAsync(mm->m, 0, sizeof(float) * NP, s) ||
      hipMemsetAsync(mm->v, 0, sizeof(float) * NP, s))
    return GPSM_EDEVICE;
  mm->step = 0;
  int rc = refresh(mm);
  if (rc) return rc;
  return hipStreamSynchronize(s) == hipSuccess ? GPSM_OK : GPSM_EDEVICE;
}

int gpsm_mlp_get_params(gpsm_mlp* mm, float* out) {
  if (!mm || !out) return GPSM_EINVAL;
  if (hipMemcpyAsync(out, mm->p, sizeof(float) * NP, hipMemcpyDeviceToHost, mm->ctx->stream) ||
      hipStreamSynchronize(mm->ctx->stream))
    return GPSM_EDEVICE;
  return GPSM_OK;
}

// The planner calls this once per planning round with a few thousand rows:
// features and outputs go through pinned staging so that the copies queue
// on the stream with the kernel and the round costs one synchronisation.
int gpsm_mlp_forward(gpsm_mlp* mm, const float* x, uint32_t B, float* y) {
  if (!mm || (B && (!x || !y))) return GPSM_EINVAL;
  if (!B) return GPSM_OK;
  const size_t xb = sizeof(float) * (size_t)B * F, yb = sizeof(float) * B;
  if (mm->x.ensure(xb) || mm->y.ensure(yb) || mm->hx.ensure(xb) || mm->hy.ensure(yb)) return GPSM_ENOMEM;
  hipStream_t s = mm->ctx->stream;
  std::memcpy(mm->hx.ptr, x, xb);
  if (hipMemcpyAsync(mm->x.ptr, mm->hx.ptr, xb, hipMemcpyHostToDevice, s) != hipSuccess) return GPSM_EDEVICE;
  hipLaunchKernelGGL(mlp_forward_kernel, dim3((B + 31) / 32), dim3(64), 0, s, weights_of(mm),
                     reinterpret_cast<const float*>(mm->x.ptr), B, reinterpret_cast<float*>(mm->y.ptr));
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(mm->hy.ptr, mm->y.ptr, yb, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return GPSM_EDEVICE;
  std::memcpy(y, mm->hy.ptr, yb);
  return GPSM_OK;
}

int gpsm_mlp_grad(gpsm_mlp* mm, const float* x, const float* t, uint32_t B, float* loss, float* grad) {
  if (!mm || !x || !t || !B || !grad) return GPSM_EINVAL;
  int rc = stage(mm, x, t, B);
  if (rc) return rc;
  hipStream_t s = mm->ctx->stream;
  hipMemsetAsync(mm->g, 0, sizeof(float) * NP, s);
  if ((rc = run_grad(mm, B))) return rc;
  float l = 0;
  if (hipMemcpyAsync(grad, mm->g, sizeof(float) * NP, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(&l, mm->loss, sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return GPSM_EDEVICE;
  hipMemsetAsync(mm->g, 0, sizeof(float) * NP, s);
  if (loss) *loss = l;
  return GPSM_OK;
}

int gpsm_mlp_train_step(gpsm_mlp* mm, const float* x, const float* t, uint32_t B, float lr, float* loss) {
  if (!mm || !x || !t || !B) return GPSM_EINVAL;
  int rc = stage(mm, x, t, B);
  if (rc) return rc;
  if ((rc = run_grad(mm, B))) return rc;
  hipStream_t s = mm->ctx->stream;
  ++mm->step;
  const float c1 = 1.0f - std::pow(0.9f, (float)mm->step), c2 = 1.0f - std::pow(0.999f, (float)mm->step);
  hipLaunchKernelGGL(adam_kernel, dim3((NP + 255) / 256), dim3(256), 0, s, mm->p, mm->g, mm->m, mm->v, lr, c1, c2);
  if ((rc = refresh(mm))) return rc;
  float l = 0;
  if (hipMemcpyAsync(&l, mm->loss, sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return GPSM_EDEVICE;
  if (loss) *loss = l;
  return GPSM_OK;
}

int gpsm_mlp_set_input_norm(gpsm_mlp* mm, const float* mu, const float* inv_sd) {
  if (!mm) return GPSM_EINVAL;
  float h[2 * F];
  for (int j = 0; j < F; ++j) {
    h[j] = mu ? mu[j] : 0.0f;
    h[F + j] = inv_sd ? inv_sd[j] : 1.0f;
  }
  if (hipMemcpyAsync(mm->norm, h, sizeof(h), hipMemcpyHostToDevice, mm->ctx->stream) != hipSuccess ||
      hipStreamSynchronize(mm->ctx->stream) != hipSuccess)
    return GPSM_EDEVICE;
  return GPSM_OK;
}

int gpsm_mlp_get_input_norm(gpsm_mlp* mm, float* mu, float* inv_sd) {
  if (!mm || !mu || !inv_sd) return GPSM_EINVAL;
  float h[2 * F];
  if (hipMemcpyAsync(h, mm->norm, sizeof(h), hipMemcpyDeviceToHost, mm->ctx->stream) != hipSuccess ||
      hipStreamSynchronize(mm->ctx->stream) != hipSuccess)
    return GPSM_EDEVICE;
  for (int j = 0; j < F; ++j) {
    mu[j] = h[j];
    inv_sd[j] = h[F + j];
  }
  return GPSM_OK;
}

int gpsm_set_planner(gpsm_ctx* c, gpsm_mlp* mm) {
  if (!c) return GPSM_EINVAL;
  c->planner = mm;
  return GPSM_OK;
}

}  // extern "C"
