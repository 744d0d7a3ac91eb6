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
//   scatter_parts    XCD-partitioned FP32 atomic adds of dL/dfeature into the
//                    table gradient (trilinear weights, 8 corners, 2 features)
//   adam             fused update, zeroes the gradient for the next step
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>

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
  float4 *dfeat;            // bwd: [8][stride]
  float *grad_w;            // bwd: 9344
};

template <bool kBwd>
__global__ __launch_bounds__(256) void mlp_train(TrainMlpArgs a) {
  __shared__ float W[kWeights];
  __shared__ float dW[kBwd ? kWeights : 1];
  __shared__ float tr_d[kBwd ? kTrainWaves : 1][64][17];
  __shared__ float tr_x[kBwd ? kTrainWaves : 1][64][17];
  for (int i = (int)threadIdx.x; i < kWeights; i += (int)blockDim.x) {
    W[i] = a.weights[i];
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
    const RayData ray = load_ray(a.rays, __float_as_uint(a.pos[sc].w));
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
          acc = mfma32(W[kW1 + (16 * m + j) * 32 + w1_col(16 * b + 4 * g + r)], xf[b][r], acc);
      d1[m] = acc;
    }
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        d2 = mfma32(W[kW2 + j * 64 + 16 * m + 4 * g + r], fmaxf(d1[m][r], 0.0f), d2);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      float4v acc = zero;
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int col = w3_col(16 * b + 4 * g + r);
          const float wv = col < 0 ? 0.0f : W[kW3 + (16 * m + j) * 31 + col];
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
          acc = mfma32(W[kW4 + (16 * m + j) * 64 + 16 * b + 4 * g + r], fmaxf(d3[b][r], 0.0f), acc);
      d4[m] = acc;
    }
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        d5 = mfma32(j < 3 ? W[kW5 + j * 64 + 16 * b + 4 * g + r] : 0.0f, fmaxf(d4[b][r], 0.0f), d5);
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
        acc = mfma32(o < 3 ? W[kW5 + o * 64 + 16 * m + j] : 0.0f, e5[r], acc);
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
          acc = mfma32(W[kW4 + (16 * b + 4 * g + r) * 64 + 16 * m + j], e4[b][r], acc);
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
          e2 = mfma32(col < 0 ? 0.0f : W[kW3 + (16 * b + 4 * g + r) * 31 + col], e3[b][r], e2);
      if (g == 0) e2[0] = go.x;  // log-sigma row
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      float4v acc = zero;
#pragma unroll
      for (int r = 0; r < 4; ++r) acc = mfma32(W[kW2 + (4 * g + r) * 64 + 16 * m + j], e2[r], acc);
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
        for (int r = 0; r < 4; ++r) acc = mfma32(W[kW1 + (16 * b + 4 * g + r) * 32 + col], e1[b][r], acc);
      // rows 16bq + 4g + r = part 4bq + g, component r
      if (valid) a.dfeat[(size_t)(4 * bq + g) * a.stride + s] = make_float4(acc[0], acc[1], acc[2], acc[3]);
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
    // this wave's sums into the workgroup's LDS dW, dest(o, k) = dW index or -1
    auto flush = [&](auto ob, auto ib, auto base, auto dest) {
      constexpr int OB = decltype(ob)::value, IB = decltype(ib)::value, BASE = decltype(base)::value;
#pragma unroll
      for (int mo = 0; mo < OB; ++mo)
#pragma unroll
        for (int mk = 0; mk < IB; ++mk)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int idx = dest(16 * mo + 4 * g + r, 16 * mk + j);
            if (idx >= 0) atomicAdd(&dW[idx], wacc[BASE + mo * IB + mk][r]);
          }
    };
    flush(IC<1>{}, IC<4>{}, IC<0>{}, [](int o, int k) { return o < 3 ? kW5 + o * 64 + k : -1; });
    flush(IC<4>{}, IC<4>{}, IC<4>{}, [](int o, int k) { return kW4 + o * 64 + k; });
    flush(IC<4>{}, IC<2>{}, IC<20>{}, [](int o, int k) {
      const int col = w3_col(k);
      return col < 0 ? -1 : kW3 + o * 31 + col;
    });
    flush(IC<1>{}, IC<4>{}, IC<28>{}, [](int o, int k) { return kW2 + o * 64 + k; });
    flush(IC<4>{}, IC<2>{}, IC<32>{}, [](int o, int k) { return kW1 + o * 32 + w1_col(k); });
    __syncthreads();
    for (int i = (int)threadIdx.x; i < kWeights; i += (int)blockDim.x)
      if (dW[i] != 0.0f) unsafeAtomicAdd(a.grad_w + i, dW[i]);
  }
}

}  // namespace nes

// FP32 forward over the work area's sample list (the FP32 render path and the
// training forward): density + colour MLPs on v_mfma_f32_16x16x4_f32.
int nes::launch_mlp_f32(const float *weights, RenderWork &w, hipStream_t stream) {
  TrainMlpArgs ma{};
  ma.pos = w.sample_pos.as<float4>();
  ma.feat = w.feat.as<float4>();
  ma.n_ptr = w.total.as<uint32_t>();
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
__global__ void train_loss(const float *rgb, const float *target, uint32_t n, float *dC,
                           float *loss) {
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
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
  if (lane_id() == 0 && l != 0.0f) atomicAdd(loss, l);
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

// dL/dfeature -> table gradient, XCD-partitioned like encode_parts.
//
// The device's float-atomic rate is set by memory requests (about 21 G/s for
// the whole chip), and the adds of one instruction that fall in one 64 B line
// leave the CU as one request.  So four lanes share a sample: lane bit 1 picks
// the x corner, bit 0 the feature.  The two x corners of a (y, z) pair are
// entries e and e + 1 on dense levels and differ only in bits 0-2 of the hash
// on hashed levels when x % 8 != 7, so the four adds of one instruction's quad
// land in one line 7 times in 8: about 4.5 requests per sample and level
// instead of 16.
//
// Levels 0 and 1 (4,913 and 12,167 dense entries) are hit by every sample on
// a few hundred lines, and adds to one line queue behind each other: alone
// they cost 2x and 1.5x any other level.  Parts 0 and 1 therefore sum their
// dense level in LDS (39 / 97 KB) and add the workgroup's totals once at the
// end, 16 lanes to a line; the grid is a fixed kScatterBlocks workgroups per
// part so those flushes stay few.  Levels 2 (29,791 entries, 238 KB) and 3
// (79,507, 636 KB) do not fit; parts 2 and 3 run their workgroups in groups
// of 2 and 4, each summing one slice of the level over all samples and
// taking the level's partner (13, 12) for every 2nd / 4th block of samples.
// The LDS array is sized for the largest slice, so every workgroup of the
// kernel holds one CU's LDS (measured on the 2^18-ray step: levels 0-1 in
// LDS 4.52 ms, only level 0 4.70, none 5.05, levels 0-2 4.22).
constexpr uint32_t kScatterBlocks = 256;
__host__ __device__ constexpr uint32_t dense_floats(int l) {
  return 2 * (kLevelOffset[l + 1] - kLevelOffset[l]);
}
__host__ __device__ constexpr uint32_t dense_slices(int l) { return l == 2 ? 2u : l == 3 ? 4u : 1u; }
__host__ __device__ constexpr uint32_t dense_slice_floats(int l) {
  return (dense_floats(l) + dense_slices(l) - 1) / dense_slices(l);
}
constexpr uint32_t kDenseLdsFloats = dense_slice_floats(3) > dense_slice_floats(2)
    ? (dense_slice_floats(3) > dense_slice_floats(1) ? dense_slice_floats(3) : dense_slice_floats(1))
    : (dense_slice_floats(2) > dense_slice_floats(1) ? dense_slice_floats(2) : dense_slice_floats(1));
static_assert(kDenseLdsFloats * 4 <= 160 * 1024, "dense-level slice exceeds the LDS");
static_assert(kScatterBlocks % 4 == 0, "workgroup groups of 4");

__global__ __launch_bounds__(256) void scatter_parts(const float4 *__restrict__ pos,
                                                     const float4 *__restrict__ dfeat,
                                                     const uint32_t *__restrict__ n_ptr,
                                                     uint32_t stride, float *grad_table,
                                                     uint32_t blocks_per_part) {
  constexpr uint32_t kPerBlock = 256 / 4;
  __shared__ float dense_sum[kDenseLdsFloats];
  const uint32_t n = min(*n_ptr, stride);
  const int part = (int)(blockIdx.x % kParts);
  const uint32_t grp = blockIdx.x / kParts;
  const PartLevels lp = part_levels(part);
  const uint32_t xb = (threadIdx.x >> 1) & 1u, ft = threadIdx.x & 1u;
  // block-uniform: the dense level of parts 0-3 is summed in LDS, in slices
  const bool in_lds = part <= 3;
  const uint32_t ns = in_lds ? dense_slices(part) : 1u;
  const uint32_t sub = grp % ns, g_idx = grp / ns, g_cnt = blocks_per_part / ns;
  const uint32_t slice = in_lds ? dense_slice_floats(part) : 0u;
  const uint32_t nfl = in_lds ? min(slice, dense_floats(part) - sub * slice) : 0u;
  const uint32_t off = in_lds ? 2u * kDevLevelOffset[part] + sub * slice : 0u;
  if (in_lds)
    for (uint32_t i = threadIdx.x; i < nfl; i += blockDim.x) dense_sum[i] = 0.0f;
  __syncthreads();
  for (uint32_t s = g_idx * kPerBlock + (threadIdx.x >> 2); s < n; s += g_cnt * kPerBlock) {
    const float4 p = pos[s];
    const float4 df = dfeat[(size_t)part * stride + s];
    const float xc[3] = {p.x, p.y, p.z};
    // sliced parts: the fine level of a block of samples goes to one of the group
    const bool fine_here = ns == 1 || (s / kPerBlock) % ns == sub;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      if (q == 1 && !fine_here) continue;
      const float d0 = q == 0 ? df.x : df.z, d1 = q == 0 ? df.y : df.w;
      if (d0 == 0.0f && d1 == 0.0f) continue;
      const float dv = ft ? d1 : d0;
      uint32_t i0[3];
      float fr[3];
#pragma unroll
      for (int ax = 0; ax < 3; ++ax) {
        const float pp = xc[ax] * lp.scale[q];
        const float fl = fminf(floorf(pp), lp.scale[q] - 1.0f);
        i0[ax] = (uint32_t)fl;
        fr[ax] = pp - fl;
      }
      const uint32_t x = i0[0] + xb;
      const float wx = xb ? fr[0] : 1.0f - fr[0];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const uint32_t y = i0[1] + (c & 1), z = i0[2] + (c >> 1);
        const uint32_t dense = x + y * lp.n1[q] + z * lp.n1sq[q];
        const uint32_t hash = (x ^ (y * kPrime1) ^ (z * kPrime2)) & ((1u << kLog2T) - 1u);
        const uint32_t e = lp.off[q] + (lp.dense[q] ? dense : hash);
        const float w = wx * ((c & 1) ? fr[1] : 1.0f - fr[1]) * ((c >> 1) ? fr[2] : 1.0f - fr[2]);
        if (in_lds && q == 0) {
          const uint32_t li = 2 * e + ft - off;  // wraps above nfl outside this half
          if (li < nfl) atomicAdd(&dense_sum[li], w * dv);
        } else {
          unsafeAtomicAdd(grad_table + 2 * (size_t)e + ft, w * dv);
        }
      }
    }
  }
  __syncthreads();
  if (in_lds)
    for (uint32_t i = threadIdx.x; i < nfl; i += blockDim.x) {
      const float v = dense_sum[i];
      if (v != 0.0f) unsafeAtomicAdd(grad_table + off + i, v);
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
  ViewDesc *poses = nullptr;
  float *targets = nullptr;
  float *target = nullptr, *loss = nullptr;
  DevBuf dout, dfeat, xo, xd, dCb;
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
  NES_TRY(t->dfeat.reserve(sizeof(float4) * kParts * (cap + 1)));
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
    if (n > w.capacity) {
      NES_TRY(trainer_grow(t, (size_t)n + n / 8 + 1024));
      continue;
    }
    NES_TRY(launch_encode_f32(t->params, w, s));
    NES_TRY(launch_mlp_f32(t->params + t->n_table, w, s));
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
  const OutLayout L = out_layout(w.out_px);
  (void)L;
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
  const uint32_t bpp = kScatterBlocks;
  scatter_parts<<<bpp * kParts, 256, 0, s>>>(w.sample_pos.as<float4>(), t->dfeat.as<float4>(),
                                             w.total.as<uint32_t>(), w.capacity, t->grads, bpp);
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
  zero_f32<<<1, 64, 0, s>>>(t->loss, 1);
  const OutLayout L = out_layout(t->work.out_px);
  const float *rgbf = reinterpret_cast<const float *>(t->work.out.as<uint8_t>() + L.rgbf);
  train_loss<<<(n + 255) / 256, 256, 0, s>>>(rgbf, t->target, n, t->dC, t->loss);
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
    NES_HIP_TRY(hipMalloc(&t->loss, 64));
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
  for (float *p : {t->params, t->grads, t->m, t->v, t->targets, t->target, t->loss})
    if (p) hipFree(p);
  if (t->occ) hipFree(t->occ);
  if (t->poses) hipFree(t->poses);
  t->dout.release();
  t->dfeat.release();
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
