// train.hip -- differentiable render step: loss, hand-written backward
// (compositing, MLPs on MFMA, FP32 atomic scatter into the hash-table
// gradient) and Adam.
//
// The reference trains nothing (its external renderer ships a trained
// snapshot); this is the training leg of the same render path.  Per step:
//   train_rays       rays from (pose, pixel) drawn with splitmix64, targets
//                    gathered from the target images
//   march_emit       shared with the renderer (occupied samples)
//   encode_parts     shared, FP32 table, XCD-partitioned levels
//   mlp_train<fwd>   density + colour MLPs on v_mfma_f32_16x16x4_f32
//   composite        shared (per-ray colour)
//   train_loss       L2 loss, dL/dC per ray
//   train_dcomposite per ray: T_i, suffix sums -> dL/dlog(sigma), dL/dpre(rgb)
//   mlp_train<bwd>   recomputes the activations of a 16-sample tile (nothing
//                    per-sample is stored), back-propagates with transposed
//                    weights on MFMA, and forms the weight gradient as an MFMA
//                    over samples (tile transposed through LDS), summed over
//                    the wave's tiles in MFMA accumulators, then reduced in
//                    LDS and flushed once per workgroup
//   scatter_slices   dL/dfeature into the table gradient (trilinear weights,
//                    8 corners, 2 features), one level slice per workgroup
//                    accumulated in LDS
//   adam             fused update, zeroes the gradient for the next step
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "device_common.h"

namespace nes {

__device__ __forceinline__ float4v mfma32(float a, float b, float4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Column of W3 fed by row k of the layer-3 input block [out2 (16) | SH (16)]:
// out2 row 0 is log-sigma (no weight), rows 1..15 are inputs 16..30, SH rows
// are inputs 0..15.
__device__ __forceinline__ int w3_col(int k) { return k < 16 ? (k == 0 ? -1 : 15 + k) : k - 16; }

// Layer-1 input row q (block b = q / 16) holds feature part_feature(q/4, q%4).
__device__ __forceinline__ int w1_col(int q) { return part_feature(q >> 2, q & 3); }

constexpr int kTrainWaves = 4;

template <int V>
struct IC {
  static constexpr int value = V;
};
// Weight-gradient blocks (16 x 16) per layer and their first index in the
// per-wave accumulator array: layer 5 1x4, layer 4 4x4, layer 3 4x2,
// layer 2 1x4, layer 1 4x2.
constexpr int kWgradBlocks = 40;

struct TrainMlpArgs {
  const float4 *pos;
  const float4 *feat;       // [8][stride]
  const uint32_t *n_ptr;
  uint32_t stride;
  const RayData *rays;
  const float *weights;     // 9344
  float4 *out;              // fwd: (sigma, r, g, b)
  const float4 *dout;       // bwd: (dL/dlogsigma, dL/dpre_r, _g, _b)
  float4 *dfeat;            // bwd: [8][stride] records of 2 float4: (position, 0), dL/dfeature
  float *grad_w;            // bwd: 9344
};

// LDS copy of the weights: row strides S1..S5 words, matrices at P1..P5.
// The backward pads the rows to odd strides (33 words for W1, 65 for W2, W4
// and W5; W3's 31 already is odd), so that the 16 lanes of an MFMA A-operand
// read that step through 16 rows do not all fall in one LDS bank (the step
// 2.23 -> 2.20 ms); the forward keeps the packed layout, which measured
// 1 % faster there.
template <bool kPad>
struct WLayout {
  static constexpr int S1 = kPad ? 33 : 32, S2 = kPad ? 65 : 64, S3 = 31, S4 = kPad ? 65 : 64,
                       S5 = kPad ? 65 : 64;
  static constexpr int P1 = 0, P2 = P1 + 64 * S1, P3 = P2 + 16 * S2, P4 = P3 + 64 * S3,
                       P5 = P4 + 64 * S4, kSize = P5 + 3 * S5;
  // LDS index of weight i of the packed layout
  static __device__ __forceinline__ int index(int i) {
    if (i < kW2) return P1 + (i - kW1) / 32 * S1 + (i - kW1) % 32;
    if (i < kW3) return P2 + (i - kW2) / 64 * S2 + (i - kW2) % 64;
    if (i < kW4) return P3 + (i - kW3);
    if (i < kW5) return P4 + (i - kW4) / 64 * S4 + (i - kW4) % 64;
    return P5 + (i - kW5) / 64 * S5 + (i - kW5) % 64;
  }
};
static_assert(kW2 - kW1 == 64 * 32 && kW3 - kW2 == 16 * 64 && kW4 - kW3 == 64 * 31 &&
                  kW5 - kW4 == 64 * 64 && kWeights - kW5 == 3 * 64,
              "weight matrix shapes");
static_assert(WLayout<false>::kSize == kWeights, "the packed layout");

template <bool kBwd>
__global__ __launch_bounds__(256) void mlp_train(TrainMlpArgs a) {
  using WL = WLayout<kBwd>;
  constexpr int kS1 = WL::S1, kS2 = WL::S2, kS3 = WL::S3, kS4 = WL::S4, kS5 = WL::S5;
  constexpr int kP1 = WL::P1, kP2 = WL::P2, kP3 = WL::P3, kP4 = WL::P4, kP5 = WL::P5;
  __shared__ float W[WL::kSize];
  __shared__ float dW[kBwd ? kWeights : 1];
  __shared__ float tr_d[kBwd ? kTrainWaves : 1][64][17];
  __shared__ float tr_x[kBwd ? kTrainWaves : 1][64][17];
  for (int i = (int)threadIdx.x; i < kWeights; i += (int)blockDim.x) {
    W[WL::index(i)] = a.weights[i];
    if (kBwd) dW[i] = 0.0f;
  }
  __syncthreads();
  const uint32_t n = min(*a.n_ptr, a.stride);
  const int lane = lane_id(), wave = (int)(threadIdx.x >> 6);
  const int j = lane & 15, g = lane >> 4;
  const float4v zero = {0.0f, 0.0f, 0.0f, 0.0f};
  const uint32_t n_tiles = (n + 15) / 16;
  // dW blocks summed over all of this wave's tiles in MFMA accumulators
  float4v wacc[kBwd ? kWgradBlocks : 1];
#pragma unroll
  for (int i = 0; i < (kBwd ? kWgradBlocks : 1); ++i) wacc[i] = zero;
  for (uint32_t tile = blockIdx.x * kTrainWaves + wave; tile < n_tiles;
       tile += gridDim.x * kTrainWaves) {
    const uint32_t s = tile * 16 + (uint32_t)j;
    const uint32_t sc = min(s, n - 1);
    const bool valid = s < n;
    // ---------------- forward (recomputed in the backward)
    float xf[2][4];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const float4 v = a.feat[(size_t)(4 * b + g) * a.stride + sc];
      xf[b][0] = v.x; xf[b][1] = v.y; xf[b][2] = v.z; xf[b][3] = v.w;
    }
    const float4 ps = a.pos[sc];
    const RayData ray = load_ray(a.rays, __float_as_uint(ps.w));
    float shv[16];
    sh4(ray.d, shv);
    float shq[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float v = shv[r];
#pragma unroll
      for (int c = 1; c < 4; ++c)
        if (g == c) v = shv[4 * c + r];
      shq[r] = v;
    }
    float4v d1[4], d3[4], d4[4], d2 = zero, d5 = zero;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      float4v acc = zero;
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          acc = mfma32(W[kP1 + (16 * m + j) * kS1 + w1_col(16 * b + 4 * g + r)], xf[b][r], acc);
      d1[m] = acc;
    }
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        d2 = mfma32(W[kP2 + j * kS2 + 16 * m + 4 * g + r], fmaxf(d1[m][r], 0.0f), d2);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      float4v acc = zero;
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int col = w3_col(16 * b + 4 * g + r);
          const float wv = col < 0 ? 0.0f : W[kP3 + (16 * m + j) * kS3 + col];
          acc = mfma32(wv, b == 0 ? d2[r] : shq[r], acc);
        }
      d3[m] = acc;
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      float4v acc = zero;
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          acc = mfma32(W[kP4 + (16 * m + j) * kS4 + 16 * b + 4 * g + r], fmaxf(d3[b][r], 0.0f), acc);
      d4[m] = acc;
    }
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        d5 = mfma32(j < 3 ? W[kP5 + j * kS5 + 16 * b + 4 * g + r] : 0.0f, fmaxf(d4[b][r], 0.0f), d5);
    if (!kBwd) {
      if (g == 0 && valid)
        a.out[s] = make_float4(expf(d2[0]), sigmoidf_(d5[0]), sigmoidf_(d5[1]), sigmoidf_(d5[2]));
      continue;
    }
    // ---------------- backward
    float4 go = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (g == 0 && valid) go = a.dout[s];
    const float4v e5 = {go.y, go.z, go.w, 0.0f};  // rows 0..2 of block 0 (g == 0)
    float4v e4[4], e3[4], e1[4], e2 = zero;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      float4v acc = zero;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int o = 4 * g + r;
        acc = mfma32(o < 3 ? W[kP5 + o * kS5 + 16 * m + j] : 0.0f, e5[r], acc);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] = d4[m][r] > 0.0f ? acc[r] : 0.0f;
      e4[m] = acc;
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      float4v acc = zero;
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          acc = mfma32(W[kP4 + (16 * b + 4 * g + r) * kS4 + 16 * m + j], e4[b][r], acc);
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] = d3[m][r] > 0.0f ? acc[r] : 0.0f;
      e3[m] = acc;
    }
    {
      const int col = w3_col(j);  // row j of the out2 block
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          e2 = mfma32(col < 0 ? 0.0f : W[kP3 + (16 * b + 4 * g + r) * kS3 + col], e3[b][r], e2);
      if (g == 0) e2[0] = go.x;  // log-sigma row
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      float4v acc = zero;
#pragma unroll
      for (int r = 0; r < 4; ++r) acc = mfma32(W[kP2 + (4 * g + r) * kS2 + 16 * m + j], e2[r], acc);
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] = d1[m][r] > 0.0f ? acc[r] : 0.0f;
      e1[m] = acc;
    }
#pragma unroll
    for (int bq = 0; bq < 2; ++bq) {
      float4v acc = zero;
      const int col = w1_col(16 * bq + j);
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc = mfma32(W[kP1 + (16 * b + 4 * g + r) * kS1 + col], e1[b][r], acc);
      // rows 16bq + 4g + r = part 4bq + g, component r
      // the sample's position beside its gradient: the scatter reads one 32 B record
      if (valid) {
        float4 *rec = a.dfeat + 2 * ((size_t)(4 * bq + g) * a.stride + s);
        rec[0] = make_float4(ps.x, ps.y, ps.z, 0.0f);
        rec[1] = make_float4(acc[0], acc[1], acc[2], acc[3]);
      }
    }
    // ---------------- weight gradient: dW_l += sum_s delta_l x_l^T, the
    // tile's delta and x transposed through LDS, accumulated in wacc
    float(*td)[17] = tr_d[wave];
    float(*tx)[17] = tr_x[wave];
    auto put = [&](float(*t)[17], int b, const float4v &v) {
#pragma unroll
      for (int r = 0; r < 4; ++r) t[16 * b + 4 * g + r][j] = valid ? v[r] : 0.0f;
    };
    // OB x IB blocks starting at accumulator BASE
    auto wgrad = [&](auto ob, auto ib, auto base) {
      constexpr int OB = decltype(ob)::value, IB = decltype(ib)::value, BASE = decltype(base)::value;
      wave_lds_fence();
#pragma unroll
      for (int mo = 0; mo < OB; ++mo)
#pragma unroll
        for (int mk = 0; mk < IB; ++mk) {
          float4v acc = wacc[BASE + mo * IB + mk];
#pragma unroll
          for (int t = 0; t < 4; ++t) acc = mfma32(td[16 * mo + j][4 * t + g], tx[16 * mk + j][4 * t + g], acc);
          wacc[BASE + mo * IB + mk] = acc;
        }
      wave_lds_fence();
    };
    // layer 5: delta rows 0..2, x = h4
    put(td, 0, e5);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      float4v h = {fmaxf(d4[b][0], 0.0f), fmaxf(d4[b][1], 0.0f), fmaxf(d4[b][2], 0.0f), fmaxf(d4[b][3], 0.0f)};
      put(tx, b, h);
    }
    wgrad(IC<1>{}, IC<4>{}, IC<0>{});
    // layer 4: delta e4, x = h3
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      put(td, b, e4[b]);
      float4v h = {fmaxf(d3[b][0], 0.0f), fmaxf(d3[b][1], 0.0f), fmaxf(d3[b][2], 0.0f), fmaxf(d3[b][3], 0.0f)};
      put(tx, b, h);
    }
    wgrad(IC<4>{}, IC<4>{}, IC<4>{});
    // layer 3: delta e3, x = [out2 | SH]
#pragma unroll
    for (int b = 0; b < 4; ++b) put(td, b, e3[b]);
    put(tx, 0, d2);
    {
      const float4v shv4 = {shq[0], shq[1], shq[2], shq[3]};
      put(tx, 1, shv4);
    }
    wgrad(IC<4>{}, IC<2>{}, IC<20>{});
    // layer 2: delta e2, x = h1
    put(td, 0, e2);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      float4v h = {fmaxf(d1[b][0], 0.0f), fmaxf(d1[b][1], 0.0f), fmaxf(d1[b][2], 0.0f), fmaxf(d1[b][3], 0.0f)};
      put(tx, b, h);
    }
    wgrad(IC<1>{}, IC<4>{}, IC<28>{});
    // layer 1: delta e1, x = features
#pragma unroll
    for (int b = 0; b < 4; ++b) put(td, b, e1[b]);
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const float4v x4 = {xf[b][0], xf[b][1], xf[b][2], xf[b][3]};
      put(tx, b, x4);
    }
    wgrad(IC<4>{}, IC<2>{}, IC<32>{});
  }
  if (kBwd) {
    // this wave's sums into the workgroup's LDS dW, dest(o, k) = dW index or
    // -1.  Within a wave every (o, k) is a different index, so the waves add
    // in turn with plain loads and stores (an FP32 atomic add in LDS costs
    // about 3 cycles per lane).
    auto flush = [&](auto ob, auto ib, auto base, auto dest) {
      constexpr int OB = decltype(ob)::value, IB = decltype(ib)::value, BASE = decltype(base)::value;
#pragma unroll
      for (int mo = 0; mo < OB; ++mo)
#pragma unroll
        for (int mk = 0; mk < IB; ++mk)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int idx = dest(16 * mo + 4 * g + r, 16 * mk + j);
            if (idx >= 0) dW[idx] += wacc[BASE + mo * IB + mk][r];
          }
    };
    for (int turn = 0; turn < kTrainWaves; ++turn) {
      if (wave == turn) {
        flush(IC<1>{}, IC<4>{}, IC<0>{}, [](int o, int k) { return o < 3 ? kW5 + o * 64 + k : -1; });
        flush(IC<4>{}, IC<4>{}, IC<4>{}, [](int o, int k) { return kW4 + o * 64 + k; });
        flush(IC<4>{}, IC<2>{}, IC<20>{}, [](int o, int k) {
          const int col = w3_col(k);
          return col < 0 ? -1 : kW3 + o * 31 + col;
        });
        flush(IC<1>{}, IC<4>{}, IC<28>{}, [](int o, int k) { return kW2 + o * 64 + k; });
        flush(IC<4>{}, IC<2>{}, IC<32>{}, [](int o, int k) { return kW1 + o * 32 + w1_col(k); });
      }
      __syncthreads();
    }
    for (int i = (int)threadIdx.x; i < kWeights; i += (int)blockDim.x)
      if (dW[i] != 0.0f) unsafeAtomicAdd(a.grad_w + i, dW[i]);
  }
}

}  // namespace nes

// FP32 forward over the work area's sample list (the FP32 render path and the
// training forward): density + colour MLPs on v_mfma_f32_16x16x4_f32.
int nes::launch_mlp_f32(const float *weights, RenderWork &w, const uint32_t *n_ptr,
                        hipStream_t stream) {
  TrainMlpArgs ma{};
  ma.pos = w.sample_pos.as<float4>();
  ma.feat = w.feat.as<float4>();
  ma.n_ptr = n_ptr;  // the pass's sample count
  ma.stride = w.capacity;
  ma.rays = w.rays.as<RayData>();
  ma.weights = weights;
  ma.out = w.sample_out.as<float4>();
  mlp_train<false><<<256 * 2, 256, 0, stream>>>(ma);
  NES_HIP_TRY(hipGetLastError());
  return NES_OK;
}

namespace nes {

// ---------------------------------------------------------------- rays
struct TrainRayArgs {
  const ViewDesc *poses;
  uint32_t n_poses, n_rays;
  uint64_t seed, step;
  const float *targets;  // [pose][H][W][3]
  RayData *rays;
  RayAcc *acc;
  float *target;         // [ray][3]
  const float *ray_o;    // explicit rays (optional)
  const float *ray_d;
};

__device__ __forceinline__ uint64_t mix64_d(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void train_rays(TrainRayArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_rays) return;
  RayData r;
  if (a.ray_o) {
    for (int k = 0; k < 3; ++k) {
      r.o[k] = a.ray_o[3 * i + k];
      r.d[k] = a.ray_d[3 * i + k];
    }
  } else {
    // same stream layout as the host sampler: stream 1000 + step, index i
    const uint64_t key = mix64_d(a.seed ^ ((1000ull + a.step) * 0xD1B54A32D192ED03ull));
    const uint64_t u = mix64_d(key + i);
    const uint32_t pose = (uint32_t)((u & 0xffffffffull) % a.n_poses);
    const ViewDesc &v = a.poses[pose];
    const uint32_t pix = (uint32_t)((u >> 32) % ((uint64_t)v.width * v.height));
    make_ray(v, pix % v.width, pix / v.width, r.o, r.d);
    const float *t = a.targets + ((size_t)pose * v.width * v.height + pix) * 3;
    a.target[3 * i + 0] = t[0];
    a.target[3 * i + 1] = t[1];
    a.target[3 * i + 2] = t[2];
  }
  r.n_steps = ray_box(r.o, r.d, &r.t_start);
  a.rays[i] = r;
  RayAcc c;
  c.T = 1.0f; c.r = c.g = c.b = 0.0f; c.depth = 0.0f; c.n_samples = 0;
  c.done = r.n_steps == 0;
  c.pad = i;
  a.acc[i] = c;
}

// Loss = mean over rays and channels of (C - target)^2; dC = dL/dC.
// Sum of the values of one workgroup's lanes (blockDim.x == kLossThreads), in
// a fixed order; the total is valid in thread 0.
constexpr int kLossThreads = 256;
__device__ __forceinline__ float block_sum(float l) {
  __shared__ float wsum[kLossThreads / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
  if (lane_id() == 0) wsum[threadIdx.x >> 6] = l;
  __syncthreads();
  float t = 0.0f;
  if (threadIdx.x == 0)
#pragma unroll
    for (int w = 0; w < kLossThreads / 64; ++w) t += wsum[w];
  return t;
}

// dL/dC per ray and one partial loss per workgroup; loss_sum adds the partials
// in order, so the loss is the same on every run (one float atomic per
// wavefront into a single address took 55 us for 2^18 rays, these two 10 us).
__global__ __launch_bounds__(kLossThreads) void train_loss(const float *rgb, const float *target,
                                                           uint32_t n, float *dC, float *partial) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (i < n) {
    const float norm = 1.0f / (3.0f * (float)n);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float e = rgb[3 * i + c] - target[3 * i + c];
      l += e * e * norm;
      dC[3 * i + c] = 2.0f * e * norm;
    }
  }
  l = block_sum(l);
  if (threadIdx.x == 0) partial[blockIdx.x] = l;
}

__global__ __launch_bounds__(kLossThreads) void loss_sum(const float *partial, uint32_t n_part,
                                                         float *loss) {
  float l = 0.0f;
  for (uint32_t i = threadIdx.x; i < n_part; i += kLossThreads) l += partial[i];
  l = block_sum(l);
  if (threadIdx.x == 0) *loss = l;
}

// Per ray: walk the composited samples again and form the gradients with
// respect to each sample's log-density and colour pre-activations.  With
// w_i = T_i a_i, g = dL/dC and S_i = sum_{k>i} w_k (c_k . g):
//   dL/dsigma_i = delta (T_{i+1} (c_i . g) - S_i),   dL/dc_i = w_i g
// The suffix sums are accumulated back to front (a forward pass parks
// T_{i+1} (c_i . g) and w_i (c_i . g) in the output slots), so no late sample
// takes its S_i as the difference of two nearly equal totals.
__global__ __launch_bounds__(256) void train_dcomposite(
    uint32_t n_slots, const RayData *__restrict__ rays, const RayAcc *__restrict__ acc,
    const uint32_t *__restrict__ counts, const uint32_t *__restrict__ offsets,
    const float4 *__restrict__ sout, const float *__restrict__ dC, float4 *__restrict__ dout) {
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = lane_id();
  const bool valid = slot < n_slots;
  uint32_t cnt = 0, base = 0, used = 0;
  float g[3] = {0.0f, 0.0f, 0.0f};
  if (valid) {
    cnt = counts[slot] & kCountMask;
    base = offsets[slot];
    if (cnt) {
      const RayAcc a = acc[slot];
      used = min(cnt, a.n_samples);
      g[0] = dC[3 * a.pad + 0];
      g[1] = dC[3 * a.pad + 1];
      g[2] = dC[3 * a.pad + 2];
    }
  }
  base = __shfl(base, 0);  // lane 0 is valid whenever any lane is
  const uint32_t mx = wave_max(cnt);
  const uint64_t below = lanes_below(lane);
  // front to back over the ray's samples (rank-major positions)
  float T = 1.0f;
  uint32_t before = 0;
  for (uint32_t i = 0; i < mx; ++i) {  // wavefront-uniform
    const uint64_t has = __ballot(cnt > i);
    const uint32_t at = base + before + (uint32_t)__popcll(has & below);
    before += (uint32_t)__popcll(has);
    if (i < used) {
      const float4 s = sout[at];
      const float sd = s.x * kDelta;
      const float w = T * -expm1f(-sd);
      const float Tn = T * expf(-sd);
      const float cg = s.y * g[0] + s.z * g[1] + s.w * g[2];
      dout[at] = make_float4(Tn * cg, w * cg, w, 0.0f);
      T = Tn;
    } else if (i < cnt) {
      dout[at] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // after termination
    }
  }
  // back to front: suffix sums S_i and the final gradients
  float S = 0.0f;
  for (uint32_t i = mx; i-- > 0;) {
    const uint64_t has = __ballot(cnt > i);
    before -= (uint32_t)__popcll(has);
    const uint32_t at = base + before + (uint32_t)__popcll(has & below);
    if (i < used) {
      const float4 p = dout[at];
      const float4 s = sout[at];
      const float dsigma = kDelta * (p.x - S);
      S += p.y;
      const float w = p.z;
      dout[at] = make_float4(dsigma * s.x, w * g[0] * s.y * (1.0f - s.y),
                             w * g[1] * s.z * (1.0f - s.z), w * g[2] * s.w * (1.0f - s.w));
    }
  }
  (void)rays;
}

// dL/dfeature -> table gradient, one slice of one level per workgroup.
//
// A float atomic to global memory costs one memory request, and the chip
// retires about 21 G of them per second however they are spread: added with
// global atomics (4 lanes per sample so that the adds of a (y, z) corner pair
// leave as one request), the 16 levels of a 2^18-ray step took 2.47 ms, 92 %
// of that rate.  Instead, every workgroup owns kSliceEntries consecutive
// entries of one level, sums their gradient in LDS with FP32 atomic adds and
// adds the slice into the table gradient once at the end.  The LDS adds are
// compare-and-swap loops (lds_add_f32): this GPU's native FP32 atomic add in
// LDS costs about 3 cycles per lane, the loop about a tenth of that
// (nes_device_lds_atomic_rate: 2.0e11 against 1.8e12 adds per second).
//
// The samples are first sorted into one bucket per (level, slice) that their
// corners touch (a counting sort: bucket_count, bucket_scan, bucket_place),
// so each slice's workgroup reads only the samples it needs.  A hashed
// level's corner pair (y, z) lies in slice (y P1 ^ z P2) >> 14 (x < 2^14
// leaves those bits alone); a dense level's pairs in at most two neighbouring
// slices.
//
// The hashed levels' slices (32 each) take their whole bucket and own their
// entries outright, so their final add is a plain load and store; the dense
// levels, whose few slices every sample hits, split their buckets over
// kDenseGroups workgroups, each adding its partial slice with global atomics.
// Workgroup b runs on XCD b % 8; a hashed slice runs on the XCD named by its
// level's part (levels p and 15 - p on XCD p), whose L2 holds that feature
// plane, and the lists are then evened out from their tails.
constexpr uint32_t kSliceEntries = 16384;  // 128 KB of FP32 feature pairs
constexpr int kSliceShift = 14;
constexpr int kSliceThreads = 1024;
constexpr int kSliceUnroll = 8;  // bucket entries per thread and loop iteration
constexpr uint32_t kSliceTasksPerXcd = 128;
constexpr uint32_t kDenseGroups[kDenseLevels] = {64, 64, 32, 16, 8};
static_assert(kSliceEntries == 1u << kSliceShift, "slice size");
static_assert(kLevelRes[kLevels - 1] < (int)kSliceEntries,
              "x and x + 1 differ below the slice bits of a hash");

__host__ __device__ constexpr uint32_t level_slices(int l) {
  return (kLevelOffset[l + 1] - kLevelOffset[l] + kSliceEntries - 1u) >> kSliceShift;
}
__host__ __device__ constexpr uint32_t bucket_base(int l) {
  return l == 0 ? 0u : bucket_base(l - 1) + level_slices(l - 1);
}
constexpr uint32_t kBuckets = bucket_base(kLevels);  // 374
constexpr int kBucketThreads = 256;
constexpr int kBucketBlocks = 512;
constexpr int kBucketScanThreads = 1024;
static_assert(kBuckets <= (uint32_t)kBucketScanThreads, "one bucket per scan thread");
static_assert((kLevelRes[kDenseLevels - 1] + 1) * (kLevelRes[kDenseLevels - 1] + 2) + 1 < (int)kSliceEntries,
              "a dense level's corners span at most two slices");
// bucket entries per sample: four per hashed level, at most two per dense one
constexpr uint32_t kEntriesPerSample = 4 * (kLevels - kDenseLevels) + 2 * kDenseLevels;
// Bucket offsets are 32-bit and entries hold sample << 2: at most this many
// samples per step (about 79 M; 2^18 rays make about 0.86 M).
constexpr uint32_t kMaxBucketSamples = 0xffffffffu / kEntriesPerSample;
static_assert(kMaxBucketSamples < (1u << 30), "entries hold sample << 2");

struct SliceTask {
  uint32_t level;           // kLevels or more marks an empty slot
  uint32_t slice;           // entries [slice * kSliceEntries, + count) of the level
  uint32_t count;
  uint32_t split, n_split;  // entries [c * split / n_split, c * (split + 1) / n_split) of the bucket
};

std::vector<SliceTask> slice_tasks() {
  std::vector<std::vector<SliceTask>> xcd(kParts);
  for (int l = 0; l < kLevels; ++l) {
    const uint32_t size = kLevelOffset[l + 1] - kLevelOffset[l];
    const uint32_t groups = l < kDenseLevels ? kDenseGroups[l] : 1u;
    const int home = min(l, kLevels - 1 - l);
    for (uint32_t k = 0; k < level_slices(l); ++k)
      for (uint32_t g = 0; g < groups; ++g)
        xcd[home].push_back(SliceTask{(uint32_t)l, k, std::min(kSliceEntries, size - k * kSliceEntries),
                                      g, groups});
  }
  // hashed slices first on every XCD (they run longest), then even out the lists
  for (auto &v : xcd)
    std::stable_sort(v.begin(), v.end(), [](const SliceTask &x, const SliceTask &y) {
      return (x.level >= (uint32_t)kDenseLevels) > (y.level >= (uint32_t)kDenseLevels);
    });
  // estimated cost: a hashed slice walks its whole bucket, a dense task a
  // group's share (measured: about twice as long)
  auto cost = [](const std::vector<SliceTask> &v) {
    uint32_t c = 0;
    for (const SliceTask &x : v) c += x.level >= (uint32_t)kDenseLevels ? 2u : 1u;
    return c;
  };
  for (;;) {
    auto lo = std::min_element(xcd.begin(), xcd.end(), [&](auto &x, auto &y) { return cost(x) < cost(y); });
    auto hi = std::max_element(xcd.begin(), xcd.end(), [&](auto &x, auto &y) { return cost(x) < cost(y); });
    if (hi->empty()) break;
    const uint32_t moved = hi->back().level >= (uint32_t)kDenseLevels ? 2u : 1u;
    if (cost(*hi) < cost(*lo) + 2 * moved) break;
    lo->push_back(hi->back());
    hi->pop_back();
  }
  std::vector<SliceTask> t(kParts * kSliceTasksPerXcd, SliceTask{2 * kLevels, 0, 0, 0, 0});
  for (int p = 0; p < kParts; ++p) {
    if (xcd[p].size() > kSliceTasksPerXcd) return {};
    std::copy(xcd[p].begin(), xcd[p].end(), t.begin() + p * kSliceTasksPerXcd);
  }
  return t;
}

// Every bucket entry of the workgroup's samples, in a fixed order: f(bucket,
// entry).  An entry is sample << 2 | pair: a hashed level's four (y, z)
// corner pairs enter the buckets of their slices separately; a dense level's
// pairs span at most two neighbouring slices (the x corner of pair 0 is the
// smallest entry, the x + 1 corner of pair 3 the largest, fewer than
// kSliceEntries apart), and the sample enters each of them once, pair 0.
template <class F>
__device__ __forceinline__ void for_each_bucket(const float4 *__restrict__ pos, uint32_t n, F &&f) {
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const float4 p = pos[s];
#pragma unroll
    for (int l = 0; l < kLevels; ++l) {
      const float scale = (float)kLevelRes[l];
      const uint32_t y = (uint32_t)fminf(floorf(p.y * scale), scale - 1.0f);
      const uint32_t z = (uint32_t)fminf(floorf(p.z * scale), scale - 1.0f);
      if (l < kDenseLevels) {
        const uint32_t n1 = (uint32_t)kLevelRes[l] + 1u;
        const uint32_t x = (uint32_t)fminf(floorf(p.x * scale), scale - 1.0f);
        const uint32_t e = x + y * n1 + z * n1 * n1;
        const uint32_t lo = e >> kSliceShift, hi = (e + n1 * n1 + n1 + 1u) >> kSliceShift;
        f(bucket_base(l) + lo, s << 2);
        if (hi != lo) f(bucket_base(l) + hi, s << 2);
      } else {
        const uint32_t hy0 = y * kPrime1, hz0 = z * kPrime2;
        const uint32_t hy[2] = {hy0, hy0 + kPrime1}, hz[2] = {hz0, hz0 + kPrime2};
#pragma unroll
        for (int c = 0; c < 4; ++c)
          f(bucket_base(l) + (((hy[c & 1] ^ hz[c >> 1]) & ((1u << kLog2T) - 1u)) >> kSliceShift),
            (s << 2) | (uint32_t)c);
      }
    }
  }
}

// Bucket sizes: per workgroup in LDS, then into count[kBuckets] (zeroed before).
__global__ __launch_bounds__(kBucketThreads) void bucket_count(const float4 *__restrict__ pos,
                                                              const uint32_t *__restrict__ n_ptr,
                                                              uint32_t stride, uint32_t *count) {
  __shared__ uint32_t hist[kBuckets];
  for (uint32_t b = threadIdx.x; b < kBuckets; b += blockDim.x) hist[b] = 0;
  __syncthreads();
  for_each_bucket(pos, min(*n_ptr, stride), [&](uint32_t b, uint32_t) { atomicAdd(&hist[b], 1u); });
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < kBuckets; b += blockDim.x)
    if (hist[b]) atomicAdd(&count[b], hist[b]);
}

// Exclusive prefix sums of the bucket sizes: off[kBuckets + 1], and the
// placement cursor of every bucket set to its start.
__global__ __launch_bounds__(kBucketScanThreads) void bucket_scan(const uint32_t *count,
                                                                 uint32_t *off, uint32_t *cursor) {
  __shared__ uint32_t wsum[kBucketScanThreads / 64];
  const uint32_t b = threadIdx.x;
  const uint32_t c = b < kBuckets ? count[b] : 0u;
  const int lane = lane_id();
  uint32_t inc = c;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(inc, d);
    if (lane >= d) inc += o;
  }
  if (lane == 63) wsum[b >> 6] = inc;
  __syncthreads();
  uint32_t ex = inc - c;
  for (uint32_t w = 0; w < (b >> 6); ++w) ex += wsum[w];
  if (b < kBuckets) {
    off[b] = ex;
    cursor[b] = ex;
  }
  if (b == kBuckets - 1) off[kBuckets] = ex + c;
}

// The samples into their buckets: each workgroup counts its entries again,
// reserves a range of every bucket with one global atomic, then writes.
__global__ __launch_bounds__(kBucketThreads) void bucket_place(const float4 *__restrict__ pos,
                                                              const uint32_t *__restrict__ n_ptr,
                                                              uint32_t stride, uint32_t *cursor,
                                                              uint32_t *__restrict__ items) {
  __shared__ uint32_t hist[kBuckets];
  for (uint32_t b = threadIdx.x; b < kBuckets; b += blockDim.x) hist[b] = 0;
  __syncthreads();
  const uint32_t n = min(*n_ptr, stride);
  for_each_bucket(pos, n, [&](uint32_t b, uint32_t) { atomicAdd(&hist[b], 1u); });
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < kBuckets; b += blockDim.x)
    hist[b] = hist[b] ? atomicAdd(&cursor[b], hist[b]) : 0u;
  __syncthreads();
  for_each_bucket(pos, n, [&](uint32_t b, uint32_t e) { items[atomicAdd(&hist[b], 1u)] = e; });
}

// One bucket entry of level L (its sample's feature gradient df and position
// p; pair c on a hashed level, every pair on a dense one) into the LDS slice,
// entries [first, first + count) of the level: the corners that fall in it.
template <int L>
__device__ __forceinline__ void slice_entry(const float4 &df, const float4 &p, uint32_t c,
                                            uint32_t first, uint32_t count, float *acc) {
  constexpr float kScale = (float)kLevelRes[L];
  constexpr bool kDense = L < kDenseLevels;
  constexpr bool kHi = L >= kParts;  // second level of its part: (z, w) of the plane
  constexpr uint32_t kN1 = (uint32_t)kLevelRes[L] + 1u;
  const float d0 = kHi ? df.z : df.x, d1 = kHi ? df.w : df.y;
  if (d0 == 0.0f && d1 == 0.0f) return;
  const float xc[3] = {p.x, p.y, p.z};
  uint32_t i0[3];
  float fr[3];
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const float pp = xc[ax] * kScale;
    const float fl = fminf(floorf(pp), kScale - 1.0f);
    i0[ax] = (uint32_t)fl;
    fr[ax] = pp - fl;
  }
  auto add_pair = [&](uint32_t cy, uint32_t cz) {
    const uint32_t y = i0[1] + cy, z = i0[2] + cz;
    uint32_t e0, e1;
    if (kDense) {
      e0 = i0[0] + y * kN1 + z * (kN1 * kN1) - first;  // wraps above count outside the slice
      e1 = e0 + 1u;
    } else {
      const uint32_t h = (y * kPrime1) ^ (z * kPrime2);
      e0 = ((i0[0] ^ h) & ((1u << kLog2T) - 1u)) - first;
      e1 = (((i0[0] + 1u) ^ h) & ((1u << kLog2T) - 1u)) - first;
    }
    const float wyz = (cy ? fr[1] : 1.0f - fr[1]) * (cz ? fr[2] : 1.0f - fr[2]);
    const float w0 = (1.0f - fr[0]) * wyz, w1 = fr[0] * wyz;
    if (e0 < count) lds_add_f32x2(&acc[2 * e0], w0 * d0, w0 * d1);
    if (e1 < count) lds_add_f32x2(&acc[2 * e1], w1 * d0, w1 * d1);
  };
  if (kDense) {
#pragma unroll
    for (int k = 0; k < 4; ++k) add_pair(k & 1, k >> 1);
  } else {
    add_pair(c & 1u, c >> 1);
  }
}

// A dense level's bucket entry on four lanes (lane bit 1 the x corner, bit 0
// the feature): one LDS add per corner pair and lane, so one instruction holds
// the adds of 16 samples, which collide on few entries at these coarse levels.
template <int L>
__device__ __forceinline__ void slice_dense_lane(const float4 &df, const float4 &p, uint32_t xb,
                                                 uint32_t ft, uint32_t first, uint32_t count,
                                                 float *acc) {
  constexpr float kScale = (float)kLevelRes[L];
  constexpr uint32_t kN1 = (uint32_t)kLevelRes[L] + 1u;
  const float d = ft ? df.y : df.x;  // dense levels are the first level of their part
  if (df.x == 0.0f && df.y == 0.0f) return;
  const float xc[3] = {p.x, p.y, p.z};
  uint32_t i0[3];
  float fr[3];
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const float pp = xc[ax] * kScale;
    const float fl = fminf(floorf(pp), kScale - 1.0f);
    i0[ax] = (uint32_t)fl;
    fr[ax] = pp - fl;
  }
  const uint32_t x = i0[0] + xb;
  const float wx = xb ? fr[0] : 1.0f - fr[0];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const uint32_t y = i0[1] + (c & 1), z = i0[2] + (c >> 1);
    const uint32_t e = x + y * kN1 + z * (kN1 * kN1) - first;  // wraps above count outside the slice
    const float w = wx * ((c & 1) ? fr[1] : 1.0f - fr[1]) * ((c >> 1) ? fr[2] : 1.0f - fr[2]);
    if (e < count) lds_add_f32(&acc[2 * e + ft], w * d);
  }
}

// Bucket entries [lo, hi): kSliceUnroll entries per thread (hashed levels) or
// per four lanes (dense levels) in flight at once.
template <int L>
__device__ __forceinline__ void slice_walk(const uint32_t *__restrict__ items,
                                           const float4 *__restrict__ rec_part, uint32_t lo,
                                           uint32_t hi, uint32_t first, uint32_t count,
                                           float *acc) {
  constexpr bool kDense = L < kDenseLevels;
  constexpr uint32_t kLanes = kDense ? 4u : 1u;  // lanes per entry
  constexpr uint32_t kPerPass = kSliceThreads / kLanes;
  const uint32_t me = threadIdx.x / kLanes;
  for (uint32_t base = lo; base < hi; base += kPerPass * kSliceUnroll) {
    float4 df[kSliceUnroll], p[kSliceUnroll];
    uint32_t ent[kSliceUnroll];
#pragma unroll
    for (int u = 0; u < kSliceUnroll; ++u) {
      const uint32_t i = base + u * kPerPass + me;
      ent[u] = i < hi ? items[i] : 0u;
      const float4 *rec = rec_part + 2 * (size_t)(ent[u] >> 2);  // (position, 0), gradient
      p[u] = i < hi ? rec[0] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      df[u] = i < hi ? rec[1] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
#pragma unroll
    for (int u = 0; u < kSliceUnroll; ++u) {
      if (kDense)
        slice_dense_lane<L>(df[u], p[u], (threadIdx.x >> 1) & 1u, threadIdx.x & 1u, first, count, acc);
      else
        slice_entry<L>(df[u], p[u], ent[u] & 3u, first, count, acc);
    }
  }
}

__global__ __launch_bounds__(kSliceThreads) void scatter_slices(
    const float4 *__restrict__ dfeat,
    const uint32_t *__restrict__ items, const uint32_t *__restrict__ off, uint32_t stride,
    const SliceTask *__restrict__ tasks, float *grad_table) {
  __shared__ __align__(16) float acc[2 * kSliceEntries];  // entry pairs for 64-bit compare-and-swap
  const SliceTask tk = tasks[(blockIdx.x % kParts) * kSliceTasksPerXcd + blockIdx.x / kParts];
  if (tk.level >= (uint32_t)kLevels) return;  // empty slot (block-uniform)
  const int level = (int)tk.level;
  const int part = min(level, kLevels - 1 - level);
  const float4 *__restrict__ rec_part = dfeat + 2 * (size_t)part * stride;
  float *grad_level = grad_table + 2 * (size_t)kDevLevelOffset[0];
  uint32_t bucket = tk.slice;
#pragma unroll
  for (int l = 1; l < kLevels; ++l)
    if (level == l) {
      grad_level = grad_table + 2 * (size_t)kDevLevelOffset[l];
      bucket = bucket_base(l) + tk.slice;
    }
  for (uint32_t i = threadIdx.x; i < 2 * tk.count; i += kSliceThreads) acc[i] = 0.0f;
  __syncthreads();
  const uint32_t b0 = off[bucket], cnt = off[bucket + 1] - b0;
  const uint32_t lo = b0 + (uint32_t)((uint64_t)cnt * tk.split / tk.n_split);
  const uint32_t hi = b0 + (uint32_t)((uint64_t)cnt * (tk.split + 1) / tk.n_split);
  const uint32_t first = tk.slice * kSliceEntries;
#define NES_SLICE_CASE(L)                                                  \
  case L:                                                                  \
    slice_walk<L>(items, rec_part, lo, hi, first, tk.count, acc);          \
    break;
  switch (level) {
    NES_SLICE_CASE(0) NES_SLICE_CASE(1) NES_SLICE_CASE(2) NES_SLICE_CASE(3)
    NES_SLICE_CASE(4) NES_SLICE_CASE(5) NES_SLICE_CASE(6) NES_SLICE_CASE(7)
    NES_SLICE_CASE(8) NES_SLICE_CASE(9) NES_SLICE_CASE(10) NES_SLICE_CASE(11)
    NES_SLICE_CASE(12) NES_SLICE_CASE(13) NES_SLICE_CASE(14) NES_SLICE_CASE(15)
  }
#undef NES_SLICE_CASE
  __syncthreads();
  float *g = grad_level + 2 * (size_t)first;
  for (uint32_t i = threadIdx.x; i < 2 * tk.count; i += kSliceThreads) {
    const float v = acc[i];
    if (v == 0.0f) continue;
    if (tk.n_split == 1)
      g[i] += v;  // the only writer of this entry
    else
      unsafeAtomicAdd(g + i, v);
  }
}

__global__ void adam_kernel(float *p, float *grad, float *m, float *v, size_t n, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const float gi = grad[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    grad[i] = 0.0f;
  }
}

__global__ void zero_f32(float *p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.0f;
}

}  // namespace nes

using namespace nes;

struct nes_trainer {
  nes_train_config cfg{};
  size_t n_table = (size_t)kDevTableEntries * 2, n_params = 0;  // device layout
  float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
  uint32_t *occ = nullptr;
  SliceTask *tasks = nullptr;  // kParts * kSliceTasksPerXcd
  ViewDesc *poses = nullptr;
  float *targets = nullptr;
  float *target = nullptr, *loss = nullptr, *loss_part = nullptr;
  DevBuf dout, dfeat, items, xo, xd, dCb;
  uint32_t *bucket_count = nullptr, *bucket_off = nullptr, *bucket_cursor = nullptr;
  float *dC = nullptr;
  RenderWork work;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  uint64_t step = 0;
  std::mutex mu;
};

namespace nes {

static int trainer_prepare(nes_trainer *t, uint32_t n_rays) {
  RenderWork &w = t->work;
  w.n_slots = n_rays;
  w.n_views = 0;
  w.out_px = n_rays;
  NES_TRY(w.rays.reserve(sizeof(RayData) * n_rays));
  NES_TRY(w.acc.reserve(sizeof(RayAcc) * n_rays));
  NES_TRY(w.counts.reserve(sizeof(uint32_t) * n_rays));
  NES_TRY(w.offsets.reserve(sizeof(uint32_t) * n_rays));
  NES_TRY(w.out.reserve(out_layout(n_rays).total));
  NES_TRY(w.total.reserve(kTotBytes));
  NES_TRY(w.views.reserve(sizeof(ViewDesc)));
  if (!w.total_host) NES_HIP_TRY(hipHostMalloc((void **)&w.total_host, kTotBytes));
  if (w.capacity < (size_t)n_rays * 4 || w.capacity_precision != NES_PRECISION_FP32)
    NES_TRY(render_reserve_samples(w, NES_PRECISION_FP32, std::max<size_t>(w.capacity, (size_t)n_rays * 4)));
  return NES_OK;
}

static int trainer_grow(nes_trainer *t, size_t cap) {
  NES_TRY(render_reserve_samples(t->work, NES_PRECISION_FP32, cap));
  NES_TRY(t->dout.reserve(sizeof(float4) * (cap + 1)));
  NES_TRY(t->dfeat.reserve(2 * sizeof(float4) * kParts * (cap + 1)));
  // a step's samples stay below kMaxBucketSamples (checked in trainer_forward)
  NES_TRY(t->items.reserve(sizeof(uint32_t) * kEntriesPerSample *
                           std::min<size_t>(cap + 1, kMaxBucketSamples)));
  return NES_OK;
}

// Forward of n_rays prepared by train_rays: samples, features, MLP outputs and
// per-ray colour.  Re-runs with larger buffers when the samples overflow.
static int trainer_forward(nes_trainer *t, uint32_t n_rays, const TrainRayArgs &ra) {
  RenderWork &w = t->work;
  hipStream_t s = t->stream;
  NES_TRY(trainer_grow(t, std::max<size_t>(w.capacity, (size_t)n_rays * 4)));
  for (int attempt = 0; attempt < 2; ++attempt) {
    train_rays<<<(n_rays + 255) / 256, 256, 0, s>>>(ra);
    NES_HIP_TRY(hipGetLastError());
    NES_TRY(launch_march_emit(t->occ, w, 0, kMaxSteps, s));
    NES_HIP_TRY(hipMemcpyAsync(w.total_host, w.total.ptr, 4, hipMemcpyDeviceToHost, s));
    NES_HIP_TRY(hipStreamSynchronize(s));
    const uint32_t n = w.total_host[0];
    if (n >= kMaxBucketSamples) {
      set_last_error("training step has " + std::to_string(n) +
                     " samples, more than the gradient buckets index (" +
                     std::to_string(kMaxBucketSamples) + ")");
      return NES_ERANGE;
    }
    if (n > w.capacity) {
      NES_TRY(trainer_grow(t, (size_t)n + n / 8 + 1024));
      continue;
    }
    NES_TRY(launch_encode_f32(t->params, w, s));
    NES_TRY(launch_mlp_f32(t->params + t->n_table, w, w.total.as<uint32_t>(), s));
    NES_TRY(launch_composite(w, true, s));
    return NES_OK;
  }
  set_last_error("training sample buffers could not be grown");
  return NES_ENOMEM;
}

// Backward from dC (per ray, already on the device) into t->grads.
static int trainer_backward(nes_trainer *t, uint32_t n_rays) {
  RenderWork &w = t->work;
  hipStream_t s = t->stream;
  train_dcomposite<<<(n_rays + 255) / 256, 256, 0, s>>>(
      n_rays, w.rays.as<RayData>(), w.acc.as<RayAcc>(), w.counts.as<uint32_t>(),
      w.offsets.as<uint32_t>(), w.sample_out.as<float4>(), t->dC, t->dout.as<float4>());
  NES_HIP_TRY(hipGetLastError());
  TrainMlpArgs ma{};
  ma.pos = w.sample_pos.as<float4>();
  ma.feat = w.feat.as<float4>();
  ma.n_ptr = w.total.as<uint32_t>();
  ma.stride = w.capacity;
  ma.rays = w.rays.as<RayData>();
  ma.weights = t->params + t->n_table;
  ma.dout = t->dout.as<float4>();
  ma.dfeat = t->dfeat.as<float4>();
  ma.grad_w = t->grads + t->n_table;
  mlp_train<true><<<256, 256, 0, s>>>(ma);
  NES_HIP_TRY(hipGetLastError());
  zero_f32<<<1, 1024, 0, s>>>(reinterpret_cast<float *>(t->bucket_count), kBuckets);
  bucket_count<<<kBucketBlocks, kBucketThreads, 0, s>>>(w.sample_pos.as<float4>(), w.total.as<uint32_t>(),
                                                        w.capacity, t->bucket_count);
  bucket_scan<<<1, kBucketScanThreads, 0, s>>>(t->bucket_count, t->bucket_off, t->bucket_cursor);
  bucket_place<<<kBucketBlocks, kBucketThreads, 0, s>>>(w.sample_pos.as<float4>(), w.total.as<uint32_t>(),
                                                        w.capacity, t->bucket_cursor, t->items.as<uint32_t>());
  scatter_slices<<<kParts * kSliceTasksPerXcd, kSliceThreads, 0, s>>>(
      t->dfeat.as<float4>(), t->items.as<uint32_t>(), t->bucket_off,
      w.capacity, t->tasks, t->grads);
  NES_HIP_TRY(hipGetLastError());
  return NES_OK;
}

static TrainRayArgs step_rays(nes_trainer *t, uint64_t step) {
  TrainRayArgs ra{};
  ra.poses = t->poses;
  ra.n_poses = t->cfg.n_poses;
  ra.n_rays = t->cfg.rays_per_batch;
  ra.seed = t->cfg.seed;
  ra.step = step;
  ra.targets = t->targets;
  ra.rays = t->work.rays.as<RayData>();
  ra.acc = t->work.acc.as<RayAcc>();
  ra.target = t->target;
  return ra;
}

// Loss and gradients of one step's batch into t->grads (zeroed first).
static int trainer_batch(nes_trainer *t, uint64_t step, float *loss_out) {
  const uint32_t n = t->cfg.rays_per_batch;
  hipStream_t s = t->stream;
  NES_TRY(trainer_prepare(t, n));
  NES_TRY(trainer_forward(t, n, step_rays(t, step)));
  const OutLayout L = out_layout(t->work.out_px);
  const float *rgbf = reinterpret_cast<const float *>(t->work.out.as<uint8_t>() + L.rgbf);
  const uint32_t n_part = (n + kLossThreads - 1) / kLossThreads;
  train_loss<<<n_part, kLossThreads, 0, s>>>(rgbf, t->target, n, t->dC, t->loss_part);
  loss_sum<<<1, kLossThreads, 0, s>>>(t->loss_part, n_part, t->loss);
  NES_HIP_TRY(hipGetLastError());
  NES_TRY(trainer_backward(t, n));
  if (loss_out) {
    NES_HIP_TRY(hipMemcpyAsync(loss_out, t->loss, 4, hipMemcpyDeviceToHost, s));
    NES_HIP_TRY(hipStreamSynchronize(s));
  }
  return NES_OK;
}

}  // namespace nes

extern "C" {

int nes_trainer_create(const nes_scene_desc *init, const nes_train_config *cfg,
                       const float *pose_c2w, const float *targets, nes_trainer **out) {
  if (!init || !init->table || !init->weights || !init->occupancy || !cfg || !out) return NES_EINVAL;
  NES_TRY(ensure_device());
  auto t = std::make_unique<nes_trainer>();
  t->cfg = *cfg;
  if (!t->cfg.rays_per_batch) t->cfg.rays_per_batch = 1u << 18;
  if (!t->cfg.lr) t->cfg.lr = 1e-2f;
  if (!t->cfg.beta1) t->cfg.beta1 = 0.9f;
  if (!t->cfg.beta2) t->cfg.beta2 = 0.99f;
  if (!t->cfg.eps) t->cfg.eps = 1e-15f;
  if (!t->cfg.fov_y_deg) t->cfg.fov_y_deg = 60.0f;
  if (t->cfg.rays_per_batch > kMaxSlots) return NES_ERANGE;
  const bool have_targets = pose_c2w && targets && t->cfg.n_poses && t->cfg.width && t->cfg.height;
  auto build = [&]() -> int {
    t->n_params = t->n_table + kWeights;
    NES_HIP_TRY(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking));
    NES_HIP_TRY(hipEventCreate(&t->ev0));
    NES_HIP_TRY(hipEventCreate(&t->ev1));
    for (float **p : {&t->params, &t->grads, &t->m, &t->v})
      NES_HIP_TRY(hipMalloc(p, t->n_params * sizeof(float)));
    NES_HIP_TRY(hipMalloc(&t->occ, NES_OCC_WORDS * sizeof(uint32_t)));
    const std::vector<SliceTask> tasks = slice_tasks();
    if (tasks.empty()) {
      set_last_error("gradient slices exceed the workgroups per XCD");
      return NES_EIO;
    }
    NES_HIP_TRY(hipMalloc(&t->tasks, tasks.size() * sizeof(SliceTask)));
    NES_HIP_TRY(hipMalloc(&t->bucket_count, kBuckets * sizeof(uint32_t)));
    NES_HIP_TRY(hipMalloc(&t->bucket_off, (kBuckets + 1) * sizeof(uint32_t)));
    NES_HIP_TRY(hipMalloc(&t->bucket_cursor, kBuckets * sizeof(uint32_t)));
    NES_HIP_TRY(hipMemcpy(t->tasks, tasks.data(), tasks.size() * sizeof(SliceTask),
                          hipMemcpyHostToDevice));
    NES_HIP_TRY(hipMalloc(&t->loss, 64));
    NES_HIP_TRY(hipMalloc(&t->loss_part, (kMaxSlots / kLossThreads) * sizeof(float)));
    const size_t nr = t->cfg.rays_per_batch;
    NES_HIP_TRY(hipMalloc(&t->target, nr * 3 * sizeof(float)));
    NES_TRY(t->dCb.reserve(nr * 3 * sizeof(float)));
    t->dC = t->dCb.as<float>();
    // the padding between levels stays zero (gradients and moments there stay 0)
    NES_HIP_TRY(hipMemset(t->params, 0, t->n_params * sizeof(float)));
    NES_TRY(table_copy(t->params, true, init->table, false, sizeof(float), hipMemcpyHostToDevice, nullptr));
    NES_HIP_TRY(hipMemcpy(t->params + t->n_table, init->weights, kWeights * sizeof(float),
                          hipMemcpyHostToDevice));
    NES_HIP_TRY(hipMemcpy(t->occ, init->occupancy, NES_OCC_WORDS * sizeof(uint32_t),
                          hipMemcpyHostToDevice));
    for (float *p : {t->grads, t->m, t->v})
      zero_f32<<<1024, 256, 0, t->stream>>>(p, t->n_params);
    if (have_targets) {
      std::vector<ViewDesc> pv(t->cfg.n_poses);
      for (uint32_t i = 0; i < t->cfg.n_poses; ++i) {
        ViewDesc &v = pv[i];
        std::memset(&v, 0, sizeof(v));
        std::memcpy(v.c2w, pose_c2w + 12 * i, sizeof(v.c2w));
        v.focal = (float)(0.5 * (double)t->cfg.height /
                          std::tan(0.5 * (double)t->cfg.fov_y_deg * 3.14159265358979323846 / 180.0));
        v.width = t->cfg.width;
        v.height = t->cfg.height;
      }
      const size_t tb = (size_t)t->cfg.n_poses * t->cfg.width * t->cfg.height * 3 * sizeof(float);
      NES_HIP_TRY(hipMalloc(&t->poses, pv.size() * sizeof(ViewDesc)));
      NES_HIP_TRY(hipMalloc(&t->targets, tb));
      NES_HIP_TRY(hipMemcpy(t->poses, pv.data(), pv.size() * sizeof(ViewDesc), hipMemcpyHostToDevice));
      NES_HIP_TRY(hipMemcpy(t->targets, targets, tb, hipMemcpyHostToDevice));
    }
    NES_HIP_TRY(hipStreamSynchronize(t->stream));
    return NES_OK;
  };
  if (const int rc = build(); rc != NES_OK) {
    nes_trainer_destroy(t.release());  // what was allocated before the failure
    return rc;
  }
  *out = t.release();
  return NES_OK;
}

int nes_trainer_destroy(nes_trainer *t) {
  if (!t) return NES_EINVAL;
  if (t->stream) hipStreamSynchronize(t->stream);
  t->dCb.release();
  for (float *p : {t->params, t->grads, t->m, t->v, t->targets, t->target, t->loss, t->loss_part})
    if (p) hipFree(p);
  if (t->occ) hipFree(t->occ);
  if (t->tasks) hipFree(t->tasks);
  if (t->poses) hipFree(t->poses);
  t->dout.release();
  t->dfeat.release();
  t->items.release();
  for (uint32_t *p : {t->bucket_count, t->bucket_off, t->bucket_cursor})
    if (p) hipFree(p);
  t->xo.release();
  t->xd.release();
  render_work_free(t->work);
  if (t->ev0) hipEventDestroy(t->ev0);
  if (t->ev1) hipEventDestroy(t->ev1);
  if (t->stream) hipStreamDestroy(t->stream);
  delete t;
  return NES_OK;
}

int nes_trainer_step(nes_trainer *t, float *loss, nes_train_stats *stats) {
  if (!t || !t->poses) return NES_EINVAL;
  std::lock_guard<std::mutex> lk(t->mu);
  hipStream_t s = t->stream;
  NES_HIP_TRY(hipEventRecord(t->ev0, s));
  float l = 0.0f;
  NES_TRY(trainer_batch(t, t->step, nullptr));
  t->step++;
  const float c1 = 1.0f - std::pow(t->cfg.beta1, (float)t->step);
  const float c2 = 1.0f - std::pow(t->cfg.beta2, (float)t->step);
  adam_kernel<<<4096, 256, 0, s>>>(t->params, t->grads, t->m, t->v, t->n_params, t->cfg.lr,
                                   t->cfg.beta1, t->cfg.beta2, t->cfg.eps, c1, c2);
  NES_HIP_TRY(hipGetLastError());
  NES_HIP_TRY(hipMemcpyAsync(&l, t->loss, 4, hipMemcpyDeviceToHost, s));
  NES_HIP_TRY(hipEventRecord(t->ev1, s));
  NES_HIP_TRY(hipEventSynchronize(t->ev1));
  if (loss) *loss = l;
  if (stats) {
    float ms = 0.0f;
    hipEventElapsedTime(&ms, t->ev0, t->ev1);
    stats->ms = ms;
    stats->rays = t->cfg.rays_per_batch;
    stats->samples = t->work.total_host[0];
    stats->step = t->step;
  }
  return NES_OK;
}

int nes_trainer_batch_grad(nes_trainer *t, uint64_t step, float *loss, float *grad_table,
                           float *grad_weights) {
  if (!t || !t->poses) return NES_EINVAL;
  std::lock_guard<std::mutex> lk(t->mu);
  hipStream_t s = t->stream;
  zero_f32<<<1024, 256, 0, s>>>(t->grads, t->n_params);
  NES_TRY(trainer_batch(t, step, loss));
  if (grad_table)
    NES_TRY(table_copy(grad_table, false, t->grads, true, sizeof(float), hipMemcpyDeviceToHost, s));
  if (grad_weights)
    NES_HIP_TRY(hipMemcpyAsync(grad_weights, t->grads + t->n_table, kWeights * sizeof(float),
                               hipMemcpyDeviceToHost, s));
  zero_f32<<<1024, 256, 0, s>>>(t->grads, t->n_params);
  NES_HIP_TRY(hipStreamSynchronize(s));
  return NES_OK;
}

int nes_trainer_get_params(nes_trainer *t, float *table, float *weights) {
  if (!t) return NES_EINVAL;
  std::lock_guard<std::mutex> lk(t->mu);
  NES_HIP_TRY(hipStreamSynchronize(t->stream));
  if (table) NES_TRY(table_copy(table, false, t->params, true, sizeof(float), hipMemcpyDeviceToHost, nullptr));
  if (weights)
    NES_HIP_TRY(hipMemcpy(weights, t->params + t->n_table, kWeights * sizeof(float), hipMemcpyDeviceToHost));
  return NES_OK;
}

int nes_trainer_set_params(nes_trainer *t, const float *table, const float *weights) {
  if (!t) return NES_EINVAL;
  std::lock_guard<std::mutex> lk(t->mu);
  NES_HIP_TRY(hipStreamSynchronize(t->stream));
  if (table) NES_TRY(table_copy(t->params, true, table, false, sizeof(float), hipMemcpyHostToDevice, nullptr));
  if (weights)
    NES_HIP_TRY(hipMemcpy(t->params + t->n_table, weights, kWeights * sizeof(float), hipMemcpyHostToDevice));
  return NES_OK;
}

// Forward of explicit rays (n x 3 origins, unit directions) into the work
// area, leaving their per-ray colour in its outputs.
static int trainer_forward_rays(nes_trainer *t, const float *ray_o, const float *ray_d,
                                uint32_t n) {
  hipStream_t s = t->stream;
  NES_TRY(trainer_prepare(t, n));
  NES_TRY(t->xo.reserve(12ull * n));
  NES_TRY(t->xd.reserve(12ull * n));
  NES_HIP_TRY(hipMemcpyAsync(t->xo.ptr, ray_o, 12ull * n, hipMemcpyHostToDevice, s));
  NES_HIP_TRY(hipMemcpyAsync(t->xd.ptr, ray_d, 12ull * n, hipMemcpyHostToDevice, s));
  TrainRayArgs ra{};
  ra.n_rays = n;
  ra.rays = t->work.rays.as<RayData>();
  ra.acc = t->work.acc.as<RayAcc>();
  ra.ray_o = t->xo.as<float>();
  ra.ray_d = t->xd.as<float>();
  return trainer_forward(t, n, ra);
}

int nes_trainer_forward(nes_trainer *t, const float *ray_o, const float *ray_d, uint32_t n,
                        float *rgb) {
  if (!t || !ray_o || !ray_d || !n || n > kMaxSlots) return NES_EINVAL;
  std::lock_guard<std::mutex> lk(t->mu);
  hipStream_t s = t->stream;
  NES_TRY(trainer_forward_rays(t, ray_o, ray_d, n));
  if (rgb) {
    const OutLayout L = out_layout(t->work.out_px);
    NES_HIP_TRY(hipMemcpyAsync(rgb, t->work.out.as<uint8_t>() + L.rgbf, 12ull * n,
                               hipMemcpyDeviceToHost, s));
  }
  NES_HIP_TRY(hipStreamSynchronize(s));
  return NES_OK;
}

// The backward takes the rays again and recomputes their forward (samples,
// features, MLP outputs, compositing state) before differentiating, so it
// depends on no state left by an earlier call: any number of forwards,
// steps or batch gradients may come between a forward and its backward.
int nes_trainer_backward(nes_trainer *t, const float *ray_o, const float *ray_d, uint32_t n,
                         const float *d_rgb, float *grad_table, float *grad_weights) {
  if (!t || !ray_o || !ray_d || !d_rgb || !n || n > kMaxSlots) return NES_EINVAL;
  std::lock_guard<std::mutex> lk(t->mu);
  hipStream_t s = t->stream;
  NES_TRY(trainer_forward_rays(t, ray_o, ray_d, n));
  NES_TRY(t->dCb.reserve(12ull * n));
  t->dC = t->dCb.as<float>();
  NES_HIP_TRY(hipMemcpyAsync(t->dC, d_rgb, 12ull * n, hipMemcpyHostToDevice, s));
  zero_f32<<<1024, 256, 0, s>>>(t->grads, t->n_params);
  NES_TRY(trainer_backward(t, n));
  if (grad_table)
    NES_TRY(table_copy(grad_table, false, t->grads, true, sizeof(float), hipMemcpyDeviceToHost, s));
  if (grad_weights)
    NES_HIP_TRY(hipMemcpyAsync(grad_weights, t->grads + t->n_table, kWeights * sizeof(float),
                               hipMemcpyDeviceToHost, s));
  zero_f32<<<1024, 256, 0, s>>>(t->grads, t->n_params);
  NES_HIP_TRY(hipStreamSynchronize(s));
  return NES_OK;
}

// ------------------------------------------------------------ checkpoint
// File layout (little-endian): "NESCKPT1", u32 version 1, u32 weights,
// u64 table floats, u64 step, then f32 parameters (table, weights), first
// moments and second moments, table floats + weights each; the table parts
// in the packed layout of the C-ABI (level after level, no padding).
namespace {
const char kCkptMagic[8] = {'N', 'E', 'S', 'C', 'K', 'P', 'T', '1'};
struct CkptHeader {
  char magic[8];
  uint32_t version, weights;
  uint64_t table, step;
};
}  // namespace

int nes_trainer_save(nes_trainer *t, const char *path) {
  if (!t || !path) return NES_EINVAL;
  std::lock_guard<std::mutex> lk(t->mu);
  NES_HIP_TRY(hipStreamSynchronize(t->stream));
  const size_t n_packed = (size_t)kTableEntries * 2 + kWeights;
  std::vector<float> dev(t->n_params), packed(n_packed);
  FILE *f = std::fopen(path, "wb");
  if (!f) {
    set_last_error(std::string("cannot create ") + path);
    return NES_EIO;
  }
  CkptHeader h;
  std::memcpy(h.magic, kCkptMagic, 8);
  h.version = 1;
  h.weights = kWeights;
  h.table = (uint64_t)kTableEntries * 2;
  h.step = t->step;
  bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1;
  for (float *src : {t->params, t->m, t->v}) {
    if (!ok) break;
    if (hipMemcpy(dev.data(), src, t->n_params * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) {
      ok = false;
      break;
    }
    table_copy(packed.data(), false, dev.data(), true, sizeof(float), hipMemcpyHostToHost, nullptr);
    std::memcpy(packed.data() + h.table, dev.data() + t->n_table, kWeights * sizeof(float));
    ok = std::fwrite(packed.data(), sizeof(float), n_packed, f) == n_packed;
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) {
    set_last_error(std::string("writing checkpoint ") + path + " failed");
    return NES_EIO;
  }
  return NES_OK;
}

int nes_trainer_load(nes_trainer *t, const char *path) {
  if (!t || !path) return NES_EINVAL;
  std::lock_guard<std::mutex> lk(t->mu);
  NES_HIP_TRY(hipStreamSynchronize(t->stream));
  FILE *f = std::fopen(path, "rb");
  if (!f) {
    set_last_error(std::string("cannot open ") + path);
    return NES_EIO;
  }
  CkptHeader h;
  const size_t n_packed = (size_t)kTableEntries * 2 + kWeights;
  if (std::fread(&h, sizeof(h), 1, f) != 1 || std::memcmp(h.magic, kCkptMagic, 8) != 0 ||
      h.version != 1 || h.weights != (uint32_t)kWeights || h.table != (uint64_t)kTableEntries * 2) {
    std::fclose(f);
    set_last_error("not a checkpoint of this scene representation");
    return NES_EINVAL;
  }
  // read everything before touching the trainer, so a short file changes nothing
  std::vector<float> buf(3 * n_packed);
  const bool ok = std::fread(buf.data(), sizeof(float), buf.size(), f) == buf.size();
  std::fclose(f);
  if (!ok) {
    set_last_error("checkpoint file is truncated");
    return NES_EINVAL;
  }
  std::vector<float> dev(t->n_params, 0.0f);  // level padding stays zero
  float *dst[3] = {t->params, t->m, t->v};
  for (int i = 0; i < 3; ++i) {
    const float *src = buf.data() + (size_t)i * n_packed;
    table_copy(dev.data(), true, src, false, sizeof(float), hipMemcpyHostToHost, nullptr);
    std::memcpy(dev.data() + t->n_table, src + h.table, kWeights * sizeof(float));
    NES_HIP_TRY(hipMemcpy(dst[i], dev.data(), t->n_params * sizeof(float), hipMemcpyHostToDevice));
  }
  t->step = h.step;
  return NES_OK;
}

}  // extern "C"
