// render.hip -- the render path the reference delegates to its external
// instant-ngp renderer (src/server.cpp:161-194 sends a FrameRequest and waits
// for RGB + depth): ray generation, occupancy-grid march, 16-level hash-grid
// encoding, density/colour MLPs and volume compositing.
//
// Pipeline per pass (a pass covers a window of march steps; one pass of 128
// steps is the default):
//   march_emit    one thread per ray slot: camera ray, box clip, occupied
//                 steps; a wave scan + one atomic per wavefront reserves a
//                 contiguous segment of the sample list on the device
//   encode_parts  16-level hash-grid encoding, XCD-partitioned: workgroups on
//                 XCD p encode levels p and 15-p, so each XCD's L2 caches two
//                 levels of the table
//   MLPs          density + colour MLPs over the dense sample list:
//                   FP32:  mlp_train<false> (train.hip), FP32 features on
//                          v_mfma_f32_16x16x4_f32
//                   mixed: mlp_mixed, FP16 features on
//                          v_mfma_f32_16x16x32_f16, FP32 accumulation
//   composite     one thread per ray: front-to-back alpha compositing in FP32,
//                 stops at T < 1e-4, writes RGB24 / depth / float RGB
// Both precisions share every kernel except the table type in encode_parts,
// the MLP kernel and the width of the MLP output, so their throughput ratio
// isolates those; both MLPs run on matrix cores.  The sample
// total stays on the device (no host round trip inside a render).
//
// Sample geometry (ray, box clip, step positions, occupancy cell) is defined
// in FP32 with explicit fused multiply-adds and correctly rounded division and
// square root, so every implementation visits the same samples.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "device_common.h"

namespace nes {

// ----------------------------------------------------------- march kernel
// One kernel for both march walks: count this ray's occupied steps, reserve
// the wavefront's range of the sample list (one atomic per wavefront on a
// device-side cursor), then write the records in rank-major order.  The
// sample total never leaves the device: later kernels read it from the
// cursor, and writes beyond `capacity` are dropped (the host re-renders with a
// larger buffer when the final total exceeds it).  counts[slot] receives the
// ray's sample count (and the step a later pass resumes at, see kCountMask),
// offsets[slot] the wavefront's range start.  A ray emits at most `cap`
// samples per pass; with `resume` it starts where its previous pass stopped.
struct MarchArgs {
  const ViewDesc *views;
  uint32_t n_views, n_slots;
  const uint32_t *occ;
  RayData *rays;
  RayAcc *acc;
  uint32_t *counts, *offsets, *samples;
  float4 *pos;
  uint32_t *cursor;  // the pass's sample cursor (a word of RenderWork::total)
  uint32_t capacity, cap;
  int k_begin, k_end, first, packed, resume;
  uint2 *ray_sh;
  uint2 *tinfo;
};

__global__ __launch_bounds__(256) void march_emit(MarchArgs m) {
  const ViewDesc *__restrict__ views = m.views;
  const uint32_t n_views = m.n_views, n_slots = m.n_slots;
  const uint32_t *__restrict__ occ = m.occ;
  RayData *rays = m.rays;
  RayAcc *acc = m.acc;
  uint32_t *counts = m.counts, *offsets = m.offsets, *samples = m.samples;
  float4 *__restrict__ pos = m.pos;
  const uint32_t capacity = m.capacity;
  const int packed = m.packed;
  uint2 *ray_sh = m.ray_sh, *tinfo = m.tinfo;
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = lane_id();
  uint32_t cnt = 0;
  uint32_t knext = kMaxSteps;           // where a later pass resumes this ray
  uint32_t bits[4] = {0u, 0u, 0u, 0u};  // occupied steps of this pass
  RayData r;
  bool active = false;
  if (slot < n_slots) {
    if (m.first) {
      uint32_t vi = 0;
      for (uint32_t i = 1; i < n_views; ++i)
        if (slot >= views[i].slot_base) vi = i;
      const ViewDesc &v = views[vi];
      const uint32_t local = slot - v.slot_base;
      const uint32_t tile = local >> 6, within = local & 63;
      const uint32_t px = (tile % v.tiles_x) * 8 + (within & 7);
      const uint32_t py = (tile / v.tiles_x) * 8 + (within >> 3);
      RayAcc a;
      a.T = 1.0f; a.r = a.g = a.b = 0.0f; a.depth = 0.0f; a.n_samples = 0;
      if (px < v.width && py < v.height) {
        make_ray(v, px, py, r.o, r.d);
        r.n_steps = ray_box(r.o, r.d, &r.t_start);
        a.pad = v.px_base + py * v.width + px;
      } else {
        r.o[0] = r.o[1] = r.o[2] = 0.0f;
        r.d[0] = r.d[1] = r.d[2] = 0.0f;
        r.t_start = 0.0f;
        r.n_steps = 0;
        a.pad = 0xffffffffu;
      }
      a.done = r.n_steps == 0;
      if (tinfo) {
        // single-pass render: compositing needs only t_start and the pixel
        tinfo[slot] = make_uint2(__float_as_uint(r.t_start), a.pad);
      } else {
        acc[slot] = a;
      }
      // the FP32 MLP reads the ray direction; the mixed one reads ray_sh
      if (!tinfo || !packed) rays[slot] = r;
      active = !a.done;
    } else {
      r = load_ray(rays, slot);
      active = !acc[slot].done;
    }
    if (active) {
      int k = m.k_begin;
      if (m.resume) k = max(k, (int)(counts[slot] >> 16));
      const int kend = min(m.k_end, (int)r.n_steps);
      knext = (uint32_t)max(kend, k);
      // four steps' occupancy words loaded at once (unconditionally: the
      // clamped cell index is inside the bitfield past kend too), then
      // taken in order
      constexpr int kStepUnroll = 4;
      bool capped = false;
      for (; k < kend && !capped; k += kStepUnroll) {
        uint32_t word[kStepUnroll], cell[kStepUnroll];
#pragma unroll
        for (int u = 0; u < kStepUnroll; ++u) {
          float x[3], t;
          step_pos(r, k + u, x, &t);
          cell[u] = occ_cell(x);
          word[u] = __ldg(occ + (cell[u] >> 5));
        }
#pragma unroll
        for (int u = 0; u < kStepUnroll; ++u) {
          const bool hit = k + u < kend && ((word[u] >> (cell[u] & 31u)) & 1u);
          if (hit && !capped) {
            bits[(k + u) >> 5] |= 1u << ((k + u) & 31);
            if (++cnt == m.cap) {
              knext = (uint32_t)(k + u) + 1u;
              capped = true;
            }
          }
        }
      }
    }
  }
  if (ray_sh && cnt) {
    // the colour MLP's SH input of this ray, in FP16 (mixed path); written by
    // every pass in which the ray has samples
    float sh[16];
    sh4(r.d, sh);
    uint32_t hw[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      half2v h;
      h[0] = (_Float16)sh[2 * k];
      h[1] = (_Float16)sh[2 * k + 1];
      hw[k] = __builtin_bit_cast(uint32_t, h);
    }
    uint4 *dst = reinterpret_cast<uint4 *>(ray_sh + (size_t)slot * 4);
    dst[0] = make_uint4(hw[0], hw[1], hw[2], hw[3]);
    dst[1] = make_uint4(hw[4], hw[5], hw[6], hw[7]);
  }
  const uint32_t wave_total = wave_sum(cnt);
  uint32_t base = 0;
  if (lane == 0 && wave_total) base = atomicAdd(m.cursor, wave_total);
  base = __shfl(base, 0);
  if (slot < n_slots) {
    counts[slot] = cnt | (knext << 16);
    offsets[slot] = base;
  }
  const uint32_t mx = wave_max(cnt);
  const uint64_t below = lanes_below(lane);
  uint32_t before = 0;  // samples of lower ranks
  int wd = 0;
  uint32_t b = bits[0];
  for (uint32_t i = 0; i < mx; ++i) {  // wavefront-uniform
    const uint64_t has = __ballot(cnt > i);
    if (cnt > i) {
      while (b == 0) {
        ++wd;
        b = wd == 1 ? bits[1] : wd == 2 ? bits[2] : bits[3];
      }
      const int k = 32 * wd + __builtin_ctz(b);
      b &= b - 1;
      const uint32_t at = base + before + (uint32_t)__popcll(has & below);
      if (at < capacity) {
        float x[3], t;
        step_pos(r, k, x, &t);
        const float xc[3] = {fminf(fmaxf(x[0], 0.0f), 1.0f), fminf(fmaxf(x[1], 0.0f), 1.0f),
                             fminf(fmaxf(x[2], 0.0f), 1.0f)};
        if (packed)
          reinterpret_cast<uint2 *>(pos)[at] = pack_pos21(xc);
        else
          pos[at] = make_float4(xc[0], xc[1], xc[2], __uint_as_float(slot));
        samples[at] = (slot << 7) | (uint32_t)k;
      }
    }
    before += (uint32_t)__popcll(has);
  }
}


constexpr int kEvalWaves = 4;

// ------------------------------------------------------ FP32 VALU eval
// The plain FP32 path: FP32 table features, one lane per sample, every MAC a
// VALU FMA with the weights in LDS (wave-uniform broadcasts).
template <int NI, int NO, bool kRelu>
__device__ __forceinline__ void dense_f32(const float *w, const float *in, float *out) {
#pragma unroll
  for (int o = 0; o < NO; ++o) {
    float acc = 0.0f;
#pragma unroll
    for (int i = 0; i < NI; ++i) acc = fmaf(w[o * NI + i], in[i], acc);
    out[o] = kRelu ? fmaxf(acc, 0.0f) : acc;
    // keep the scheduler from hoisting every weight read of the layer
    if ((o & 3) == 3) __asm__ volatile("" ::: "memory");
  }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void mlp_fp32_valu(
    const float4 *__restrict__ pos, const float4 *__restrict__ feat,
    const uint32_t *__restrict__ n_ptr, uint32_t stride, const RayData *__restrict__ rays,
    const float *__restrict__ Wg, float4 *__restrict__ out) {
  const uint32_t n = min(*n_ptr, stride);
  __shared__ __attribute__((aligned(16))) float W[kWeights];
  for (int i = (int)threadIdx.x; i < kWeights / 4; i += (int)blockDim.x)
    reinterpret_cast<float4 *>(W)[i] = reinterpret_cast<const float4 *>(Wg)[i];
  __syncthreads();
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float f[32];
#pragma unroll
    for (int p = 0; p < kParts; ++p) {
      const float4 v = feat[(size_t)p * stride + i];
      f[part_feature(p, 0)] = v.x;
      f[part_feature(p, 1)] = v.y;
      f[part_feature(p, 2)] = v.z;
      f[part_feature(p, 3)] = v.w;
    }
    const RayData r = load_ray(rays, __float_as_uint(pos[i].w));
    float h1[64], o2[16], in3[31], h3[64], h4[64], o5[3], sh[16];
    dense_f32<32, 64, true>(W + kW1, f, h1);
    dense_f32<64, 16, false>(W + kW2, h1, o2);
    sh4(r.d, sh);
#pragma unroll
    for (int k = 0; k < 16; ++k) in3[k] = sh[k];
#pragma unroll
    for (int k = 0; k < 15; ++k) in3[16 + k] = o2[1 + k];
    dense_f32<31, 64, true>(W + kW3, in3, h3);
    dense_f32<64, 64, true>(W + kW4, h3, h4);
    dense_f32<64, 3, false>(W + kW5, h4, o5);
    out[i] = make_float4(expf(o2[0]), sigmoidf_(o5[0]), sigmoidf_(o5[1]), sigmoidf_(o5[2]));
  }
}

// ----------------------------------------------------------- mixed eval
// MFMA tile: v_mfma_f32_16x16x32_f16, D[16x16] += A[16x32] B[32x16].
// Rows (M) are output neurons, columns (N) are 16 samples, K is the input.
// Lane l holds B column l%16 at K-slots (l/16, s = 0..7); D rows 4(l/16)+r.
// A neuron block's D therefore already sits where the next layer's B needs it
// once the K index of slot (g, s) is taken to be
//   s < 4 : block 2c   neuron 4g + s        s >= 4 : block 2c+1 neuron 4g + s-4
// which the host bakes into the weight fragments (build_mlp_fragments).  The
// level-split encode delivers layer 1's B operand straight from the gathers.
__device__ __forceinline__ float4v mfma(half8 a, half8 b, float4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// ReLU of eight FP32 accumulators as the next layer's FP16 B operand: round
// pairs to FP16 (v_cvt_pk_f16_f32, round to nearest even) and clamp at zero
// in FP16 (v_pk_max_f16); rounding is monotonic and keeps the sign, so this
// equals rounding the clamped values.
__device__ __forceinline__ half8 pack_relu(float4v lo, float4v hi) {
  typedef float float2v_ __attribute__((ext_vector_type(2)));
  const half2v z = {(_Float16)0.0f, (_Float16)0.0f};
  const half2v a = __builtin_elementwise_max(__builtin_convertvector((float2v_){lo[0], lo[1]}, half2v), z);
  const half2v b = __builtin_elementwise_max(__builtin_convertvector((float2v_){lo[2], lo[3]}, half2v), z);
  const half2v c = __builtin_elementwise_max(__builtin_convertvector((float2v_){hi[0], hi[1]}, half2v), z);
  const half2v d = __builtin_elementwise_max(__builtin_convertvector((float2v_){hi[2], hi[3]}, half2v), z);
  half8 r;
  r[0] = a[0]; r[1] = a[1]; r[2] = b[0]; r[3] = b[1];
  r[4] = c[0]; r[5] = c[1]; r[6] = d[0]; r[7] = d[1];
  return r;
}

// Mixed-path sigmoid: hardware exp and reciprocal (about 1 ulp each, far
// below the FP16 network's own rounding).
__device__ __forceinline__ float sigmoid_fast(float v) {
  return __builtin_amdgcn_rcpf(1.0f + __expf(-v));
}

// ------------------------------------------------- XCD-partitioned encode
// MI355X dispatches workgroup b to XCD b % 8, and each XCD has its own 4 MB
// L2.  Part p = b % 8 encodes only levels p and 15 - p for every sample, so
// an XCD's L2 holds one coarse level plus one (2 MB in FP16) hashed level and
// the corner gathers hit in L2 instead of streaming 64-byte lines from the
// Infinity Cache.  (In FP32 a hashed level alone is 4 MB.)  Features go to
// feat[p][sample] as {level p f0, f1, level 15-p f0, f1}.
template <class Tab, int P>
__device__ __forceinline__ void encode_part_loop(const typename PosIn<Tab>::type *__restrict__ pos,
                                                 uint32_t n, uint32_t stride, const Tab *table,
                                                 typename PartOut<Tab>::type *__restrict__ feat,
                                                 uint32_t grp, uint32_t blocks_per_part,
                                                 uint32_t *xch) {
  constexpr int LA = P, LB = 15 - P;
  const char *base_a = level_base<Tab, LA>(table), *base_b = level_base<Tab, LB>(table);
  typename PartOut<Tab>::type *dst = feat + (size_t)P * stride;
  for (uint32_t s = grp * blockDim.x + threadIdx.x; s < n; s += blocks_per_part * blockDim.x) {
    float xc[3];
    PosIn<Tab>::get(pos[s], xc);
    float2 fa, fb;
    encode_two<Tab, LA, LB>(base_a, base_b, xc, fa, fb, xch);
    dst[s] = PartOut<Tab>::pack(fa, fb);
  }
}

// Each part is its own instantiation, so level resolutions, offsets and the
// dense/hashed choice are compile-time constants.
template <class Tab>
__global__ __launch_bounds__(256) void encode_parts(const typename PosIn<Tab>::type *__restrict__ pos,
                                                    const uint32_t *__restrict__ n_ptr,
                                                    uint32_t stride, const Tab *__restrict__ table,
                                                    typename PartOut<Tab>::type *__restrict__ feat,
                                                    uint32_t blocks_per_part) {
  __shared__ uint32_t xch_all[4][256];  // per-wavefront gather exchange (FP16 hashed levels)
  uint32_t *xch = xch_all[threadIdx.x >> 6];
  const uint32_t n = min(*n_ptr, stride);
  const uint32_t grp = blockIdx.x / kParts;
  switch (blockIdx.x % kParts) {
    case 0: encode_part_loop<Tab, 0>(pos, n, stride, table, feat, grp, blocks_per_part, xch); break;
    case 1: encode_part_loop<Tab, 1>(pos, n, stride, table, feat, grp, blocks_per_part, xch); break;
    case 2: encode_part_loop<Tab, 2>(pos, n, stride, table, feat, grp, blocks_per_part, xch); break;
    case 3: encode_part_loop<Tab, 3>(pos, n, stride, table, feat, grp, blocks_per_part, xch); break;
    case 4: encode_part_loop<Tab, 4>(pos, n, stride, table, feat, grp, blocks_per_part, xch); break;
    case 5: encode_part_loop<Tab, 5>(pos, n, stride, table, feat, grp, blocks_per_part, xch); break;
    case 6: encode_part_loop<Tab, 6>(pos, n, stride, table, feat, grp, blocks_per_part, xch); break;
    default: encode_part_loop<Tab, 7>(pos, n, stride, table, feat, grp, blocks_per_part, xch); break;
  }
}

// FP16 MFMA MLPs: lane (g, j) of a 16-sample tile loads parts 2g and 2g+1 of
// sample j -- the layer-1 B operand with K-slot s of group g holding feature
// part_feature(2g + s/4, s%4) (baked into the W1 fragments).  Each wave keeps
// the 20 weight fragments in registers for all its tiles: re-reading them
// from LDS per tile moves 20 KB of LDS data per 16 samples, and the frame is
// power-limited, so that energy costs clock in every kernel.
__global__ __launch_bounds__(256) void mlp_mixed(const uint32_t *__restrict__ samples,
                                                 const uint2 *__restrict__ feat,
                                                 const uint32_t *__restrict__ n_ptr,
                                                 uint32_t stride,
                                                 const uint2 *__restrict__ ray_sh,
                                                 const half8 *__restrict__ wfrag,
                                                 uint2 *__restrict__ out) {
  const uint32_t n = min(*n_ptr, stride);
  __shared__ half8 wl[20 * 64];
  for (int i = (int)threadIdx.x; i < 20 * 64; i += (int)blockDim.x) wl[i] = wfrag[i];
  __syncthreads();
  const int lane = lane_id(), wave = (int)(threadIdx.x >> 6);
  const int j = lane & 15, g = lane >> 4;
  half8 wf[20];
#pragma unroll
  for (int i = 0; i < 20; ++i) wf[i] = wl[64 * i + lane];
  const float4v zero = {0.0f, 0.0f, 0.0f, 0.0f};
  const uint32_t n_tiles = (n + 15) / 16;
  const uint32_t tile_step = gridDim.x * kEvalWaves;
  uint32_t tile = blockIdx.x * kEvalWaves + wave;
  if (tile >= n_tiles) return;
  // Software pipeline: the sample record and features of the next tile are
  // loaded while this tile's MFMAs run, its SH (which needs the record) at
  // the end of the iteration.  The loads are unconditional (indices clamped
  // to the last sample) so that each wait covers only what it needs.
  uint32_t sc = min(tile * 16 + (uint32_t)j, n - 1);
  uint2 fa = feat[(size_t)(2 * g) * stride + sc];
  uint2 fb = feat[(size_t)(2 * g + 1) * stride + sc];
  uint2 shw = ray_sh[(size_t)(samples[sc] >> 7) * 4 + g];
  while (true) {
    const uint32_t s = tile * 16 + (uint32_t)j;
    const uint32_t next = tile + tile_step;
    const bool more = next < n_tiles;
    const uint32_t nsc = min(next * 16 + (uint32_t)j, n - 1);
    const uint32_t nrec = samples[nsc];
    const uint2 nfa = feat[(size_t)(2 * g) * stride + nsc];
    const uint2 nfb = feat[(size_t)(2 * g + 1) * stride + nsc];
    __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the MFMAs
    const uint32_t words[4] = {fa.x, fa.y, fb.x, fb.y};
    half8 b1;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const half2v h = __builtin_bit_cast(half2v, words[k]);
      b1[2 * k] = h.x;
      b1[2 * k + 1] = h.y;
    }
    float4v d1[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) d1[m] = mfma(wf[m], b1, zero);
    float4v d2 = mfma(wf[4], pack_relu(d1[0], d1[1]), zero);
    d2 = mfma(wf[5], pack_relu(d1[2], d1[3]), d2);
    const half2v sh01 = __builtin_bit_cast(half2v, shw.x), sh23 = __builtin_bit_cast(half2v, shw.y);
    half8 b3;
    b3[0] = (_Float16)d2[0]; b3[1] = (_Float16)d2[1];
    b3[2] = (_Float16)d2[2]; b3[3] = (_Float16)d2[3];
    b3[4] = sh01[0]; b3[5] = sh01[1]; b3[6] = sh23[0]; b3[7] = sh23[1];
    float4v d3[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) d3[m] = mfma(wf[6 + m], b3, zero);
    const half8 b4a = pack_relu(d3[0], d3[1]), b4b = pack_relu(d3[2], d3[3]);
    float4v d4[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      d4[m] = mfma(wf[10 + 2 * m], b4a, zero);
      d4[m] = mfma(wf[11 + 2 * m], b4b, d4[m]);
    }
    float4v d5 = mfma(wf[18], pack_relu(d4[0], d4[1]), zero);
    d5 = mfma(wf[19], pack_relu(d4[2], d4[3]), d5);
    if (g == 0 && s < n) {
      // (log sigma, r, g, b) as FP16; the FP32 composite takes the exp
      half2v lo, hi;
      lo.x = (_Float16)d2[0];
      lo.y = (_Float16)sigmoid_fast(d5[0]);
      hi.x = (_Float16)sigmoid_fast(d5[1]);
      hi.y = (_Float16)sigmoid_fast(d5[2]);
      out[s] = make_uint2(__builtin_bit_cast(uint32_t, lo), __builtin_bit_cast(uint32_t, hi));
    }
    if (!more) break;
    __builtin_amdgcn_sched_barrier(0);
    shw = ray_sh[(size_t)(nrec >> 7) * 4 + g];
    __builtin_amdgcn_sched_barrier(0);  // and the SH load at the end of the iteration
    tile = next;
    fa = nfa;
    fb = nfb;
  }
}

// ------------------------------------------------------------- composite
struct OutPtrs {
  uint8_t *rgb24;
  float *depth;
  float *rgbf;
  uint16_t *nsamp;
};

// Per-sample MLP output as the composite reads it: FP32 (sigma, r, g, b), or
// the mixed path's FP16 (log sigma, r, g, b).
__device__ __forceinline__ float4 sample_value(const float4 &v) { return v; }
__device__ __forceinline__ float4 sample_value(const uint2 &v) {
  const half2v lo = __builtin_bit_cast(half2v, v.x), hi = __builtin_bit_cast(half2v, v.y);
  return make_float4(__expf((float)lo.x), (float)lo.y, (float)hi.x, (float)hi.y);
}

// One lane per ray slot, walking the wavefront's rank-major sample range
// eight ranks at a time (positions from ballots, loads issued together, then
// composited in order).
template <class SampleOut>
__global__ __launch_bounds__(256) void composite(
    uint32_t n_slots, const RayData *__restrict__ rays, RayAcc *acc,
    const uint32_t *__restrict__ counts, const uint32_t *__restrict__ offsets,
    const uint32_t *__restrict__ samples, const SampleOut *__restrict__ sout,
    const uint32_t *__restrict__ n_ptr, uint32_t capacity, int keep, int last, OutPtrs o,
    const uint2 *__restrict__ tinfo, uint32_t *stats) {
  if (*n_ptr > capacity) return;  // overflowed pass: the host re-renders
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = lane_id();
  const bool valid = slot < n_slots;
  RayAcc a{};
  uint32_t cnt = 0, base = 0;
  float ts = 0.0f;
  if (valid) {
    if (tinfo) {
      // single-pass render: fresh state, t_start and pixel from the march
      const uint2 ti = tinfo[slot];
      a.T = 1.0f;
      a.pad = ti.y;
      ts = __uint_as_float(ti.x);
    } else {
      a = acc[slot];
    }
    cnt = counts[slot] & kCountMask;
    base = offsets[slot];
  }
  base = __shfl(base, 0);  // lane 0 is valid whenever any lane is
  const bool live = valid && !a.done && cnt;
  if (live && !tinfo) ts = load_ray(rays, slot).t_start;
  const uint32_t n_before = a.n_samples;
  const uint32_t mx = wave_max(cnt);
  const uint64_t below = lanes_below(lane);
  uint32_t before = 0;
  for (uint32_t i0 = 0; i0 < mx; i0 += 8) {  // wavefront-uniform
    if (__ballot(live && !a.done && i0 < cnt) == 0) break;
    SampleOut sb[8];
    uint32_t kb[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint32_t i = i0 + (uint32_t)u;
      const uint64_t has = __ballot(cnt > i);
      const uint32_t at = base + before + (uint32_t)__popcll(has & below);
      before += (uint32_t)__popcll(has);
      sb[u] = SampleOut{};
      kb[u] = 0u;
      if (live && cnt > i && !a.done) {
        sb[u] = sout[at];
        kb[u] = samples[at];
      }
    }
    if (live) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (i0 + (uint32_t)u >= cnt || a.done) break;
        const float4 sm = sample_value(sb[u]);
        const float t = step_t((int)(kb[u] & 127u), ts);
        const float sd = sm.x * kDelta;
        const float alpha = -expm1f(-sd);
        const float w = a.T * alpha;
        a.r = fmaf(w, sm.y, a.r);
        a.g = fmaf(w, sm.z, a.g);
        a.b = fmaf(w, sm.w, a.b);
        a.depth = fmaf(w, t, a.depth);
        a.T *= expf(-sd);
        a.n_samples++;
        if (a.T < kTStop) a.done = 1;
      }
    }
  }
  if (stats) {
    // Termination counters, touched only by wavefronts with a ray that
    // reached T < 1e-4 in this pass (none in a scene that never does), on
    // one of kTotTermSlots lines so that such wavefronts rarely share one.
    const bool term = live && a.done;
    const uint64_t tb = __ballot(term);
    if (tb) {
      const uint32_t here = a.n_samples - n_before;  // composited in this pass
      bool saved = false;
      if (term && !last) saved = (counts[slot] >> 16) < load_ray(rays, slot).n_steps;
      const uint32_t ts_sum = wave_sum(term ? a.n_samples : 0u);
      const uint32_t waste = wave_sum(term ? cnt - here : 0u);
      const uint32_t ns = (uint32_t)__popcll(__ballot(saved));
      if (lane == 0) {
        uint32_t *c = stats + kTotTermBase +
                      kTotTermStride * (((blockIdx.x * blockDim.x + threadIdx.x) >> 6) % kTotTermSlots);
        atomicAdd(c + 0, (uint32_t)__popcll(tb));
        atomicAdd(c + 1, ts_sum);
        if (waste) atomicAdd(c + 2, waste);
        if (ns) atomicAdd(c + 3, ns);
      }
    }
    // every pass adds its sample total once
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(stats + kTotEvaluated, *n_ptr);
  }
  if (!valid) return;
  // the state carries into the next pass; after the last pass only a caller
  // that reads it again (keep: the training backward) stores it
  if (live && (!last || keep)) acc[slot] = a;
  if (last && a.pad != 0xffffffffu) {
    const uint32_t p = a.pad;
    const float c[3] = {a.r, a.g, a.b};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float q = floorf(fmaf(c[k], 255.0f, 0.5f));
      o.rgb24[3 * (size_t)p + k] = (uint8_t)fminf(fmaxf(q, 0.0f), 255.0f);
      o.rgbf[3 * (size_t)p + k] = c[k];
    }
    o.depth[p] = a.depth;
    o.nsamp[p] = (uint16_t)a.n_samples;
  }
}

__global__ void zero_u32(uint32_t *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// FP32 expected depth -> the GRAY8 depth plane of the response.
__global__ void depth_gray8_kernel(const float *__restrict__ depth, size_t n, float scale,
                                   uint8_t *__restrict__ gray) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) gray[i] = depth_gray8(depth[i], scale);
}

// ----------------------------------------------------------------- host
void build_mlp_fragments(const float *w, _Float16 *frag) {
  // slot (g, s) of K-chunk c -> input neuron of a 64-wide hidden layer
  auto hid = [](int c, int g, int s) { return s < 4 ? 32 * c + 4 * g + s : 32 * c + 16 + 4 * g + (s - 4); };
  auto put = [&](int mf, int lane, int s, float v) { frag[((size_t)mf * 64 + lane) * 8 + s] = (_Float16)v; };
  for (int lane = 0; lane < 64; ++lane) {
    const int i = lane & 15, g = lane >> 4;
    for (int s = 0; s < 8; ++s) {
      // layer 1: 32 features -> 64; K-slot (g, s) holds feature
      // part_feature(2g + s/4, s%4) as delivered by encode_parts
      for (int m = 0; m < 4; ++m)
        put(m, lane, s, w[kW1 + (16 * m + i) * 32 + part_feature(2 * g + s / 4, s % 4)]);
      // layer 2: 64 -> 16, two K-chunks
      for (int c = 0; c < 2; ++c) put(4 + c, lane, s, w[kW2 + i * 64 + hid(c, g, s)]);
      // layer 3: [density outputs 4g..4g+3 | SH 4g..4g+3] -> 64; density
      // output 0 is log-sigma and gets no weight; outputs 1..15 are inputs
      // 16..30 of W3, SH are inputs 0..15
      for (int m = 0; m < 4; ++m) {
        float v;
        if (s < 4) {
          const int o2 = 4 * g + s;
          v = o2 == 0 ? 0.0f : w[kW3 + (16 * m + i) * 31 + 16 + (o2 - 1)];
        } else {
          v = w[kW3 + (16 * m + i) * 31 + 4 * g + (s - 4)];
        }
        put(6 + m, lane, s, v);
      }
      // layer 4: 64 -> 64
      for (int m = 0; m < 4; ++m)
        for (int c = 0; c < 2; ++c) put(10 + 2 * m + c, lane, s, w[kW4 + (16 * m + i) * 64 + hid(c, g, s)]);
      // layer 5: 64 -> 3 (rows 3..15 zero)
      for (int c = 0; c < 2; ++c) put(18 + c, lane, s, i < 3 ? w[kW5 + i * 64 + hid(c, g, s)] : 0.0f);
    }
  }
}

int scene_upload(SceneDev &s, const nes_scene_desc *desc) {
  if (!desc || !desc->table || !desc->weights || !desc->occupancy) return NES_EINVAL;
  const size_t nt = (size_t)kTableEntries * 2, ndev = (size_t)kDevTableEntries * 2;
  if (!s.table_f32) {
    NES_HIP_TRY(hipMalloc(&s.table_f32, ndev * sizeof(float)));
    NES_HIP_TRY(hipMalloc(&s.table_f16, ndev * sizeof(_Float16)));
    NES_HIP_TRY(hipMalloc(&s.weights_f32, kWeights * sizeof(float)));
    NES_HIP_TRY(hipMalloc(&s.wfrag, 20 * 64 * 8 * sizeof(_Float16)));
    NES_HIP_TRY(hipMalloc(&s.occ, NES_OCC_WORDS * sizeof(uint32_t)));
    // the padding between levels (device layout) is never read; keep it zero
    NES_HIP_TRY(hipMemset(s.table_f32, 0, ndev * sizeof(float)));
    NES_HIP_TRY(hipMemset(s.table_f16, 0, ndev * sizeof(_Float16)));
  }
  std::vector<_Float16> t16(nt);
  for (size_t i = 0; i < nt; ++i) t16[i] = (_Float16)desc->table[i];
  std::vector<_Float16> frag(20 * 64 * 8);
  build_mlp_fragments(desc->weights, frag.data());
  NES_TRY(table_copy(s.table_f32, true, desc->table, false, sizeof(float), hipMemcpyHostToDevice, nullptr));
  NES_TRY(table_copy(s.table_f16, true, t16.data(), false, sizeof(_Float16), hipMemcpyHostToDevice, nullptr));
  NES_HIP_TRY(hipMemcpy(s.weights_f32, desc->weights, kWeights * sizeof(float), hipMemcpyHostToDevice));
  NES_HIP_TRY(hipMemcpy(s.wfrag, frag.data(), frag.size() * sizeof(_Float16), hipMemcpyHostToDevice));
  NES_HIP_TRY(hipMemcpy(s.occ, desc->occupancy, NES_OCC_WORDS * sizeof(uint32_t), hipMemcpyHostToDevice));
  return NES_OK;
}

void scene_free(SceneDev &s) {
  hipFree(s.table_f32);
  hipFree(s.table_f16);
  hipFree(s.weights_f32);
  hipFree(s.wfrag);
  hipFree(s.occ);
  s = SceneDev();
}

void render_work_free(RenderWork &w) {
  for (DevBuf *b : {&w.views, &w.rays, &w.acc, &w.counts, &w.offsets, &w.ray_sh, &w.tinfo, &w.gray, &w.samples,
                    &w.sample_out, &w.sample_pos, &w.feat, &w.out, &w.total})
    b->release();
  if (w.total_host) hipHostFree(w.total_host);
  w.total_host = nullptr;
  if (w.ev_views) {
    hipEventSynchronize(w.ev_views);
    hipEventDestroy(w.ev_views);
    w.ev_views = nullptr;
  }
  if (w.views_pinned) hipHostFree(w.views_pinned);
  w.views_pinned = nullptr;
  w.views_pinned_cap = 0;
  for (hipEvent_t &e : w.ev)
    if (e) {
      hipEventDestroy(e);
      e = nullptr;
    }
}

static size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

OutLayout out_layout(size_t px) {
  OutLayout l;
  l.rgb24 = 0;
  l.depth = align_up(3 * px, 256);
  l.rgbf = align_up(l.depth + 4 * px, 256);
  l.nsamp = align_up(l.rgbf + 12 * px, 256);
  l.total = align_up(l.nsamp + 2 * px, 256);
  return l;
}

static OutPtrs out_ptrs(RenderWork &w) {
  const OutLayout L = out_layout(w.out_px);
  OutPtrs op;
  op.rgb24 = w.out.as<uint8_t>() + L.rgb24;
  op.depth = reinterpret_cast<float *>(w.out.as<uint8_t>() + L.depth);
  op.rgbf = reinterpret_cast<float *>(w.out.as<uint8_t>() + L.rgbf);
  op.nsamp = reinterpret_cast<uint16_t *>(w.out.as<uint8_t>() + L.nsamp);
  return op;
}

int render_setup(RenderWork &w, const nes_camera *views, int n_views, hipStream_t stream) {
  // Until this render completes, the work area describes no views: a fetch
  // after a failed setup or pass is refused instead of reading stale layouts.
  w.n_views = 0;
  w.out_px = 0;
  if (!views || n_views < 1) return NES_EINVAL;
  w.host_views.resize((size_t)n_views);
  uint64_t slots = 0, px = 0;
  for (int i = 0; i < n_views; ++i) {
    const nes_camera &c = views[i];
    if (c.width < 1 || c.height < 1 || c.width > 16384 || c.height > 16384 ||
        !(c.fov_y_deg > 0.0f && c.fov_y_deg < 180.0f)) {
      set_last_error("bad view size or field of view");
      return NES_EINVAL;
    }
    ViewDesc &v = w.host_views[(size_t)i];
    std::memcpy(v.c2w, c.c2w, sizeof(v.c2w));
    v.focal = (float)(0.5 * (double)c.height /
                      std::tan(0.5 * (double)c.fov_y_deg * 3.14159265358979323846 / 180.0));
    v.width = c.width;
    v.height = c.height;
    v.tiles_x = (c.width + 7) / 8;
    const uint32_t tiles_y = (c.height + 7) / 8;
    v.slot_base = (uint32_t)slots;
    v.slot_count = v.tiles_x * tiles_y * 64;
    v.px_base = (uint32_t)px;
    v.pad = 0;
    slots += v.slot_count;
    px += (uint64_t)c.width * c.height;
    if (slots > kMaxSlots) {
      set_last_error("render batch exceeds 2^25 ray slots");
      return NES_ERANGE;
    }
  }
  w.n_views = (uint32_t)n_views;
  w.n_slots = (uint32_t)slots;
  w.out_px = px;
  NES_TRY(w.views.reserve(sizeof(ViewDesc) * n_views));
  NES_TRY(w.rays.reserve(sizeof(RayData) * slots));
  NES_TRY(w.acc.reserve(sizeof(RayAcc) * slots));
  NES_TRY(w.counts.reserve(sizeof(uint32_t) * slots));
  NES_TRY(w.offsets.reserve(sizeof(uint32_t) * slots));
  NES_TRY(w.ray_sh.reserve(32 * slots));
  NES_TRY(w.tinfo.reserve(8 * slots));
  NES_TRY(w.out.reserve(out_layout(px).total));
  NES_TRY(w.total.reserve(kTotBytes));
  if (!w.total_host) NES_HIP_TRY(hipHostMalloc((void **)&w.total_host, kTotBytes));
  for (hipEvent_t &e : w.ev)
    if (!e) NES_HIP_TRY(hipEventCreate(&e));
  if (!w.ev_views) NES_HIP_TRY(hipEventCreateWithFlags(&w.ev_views, hipEventDisableTiming));
  NES_HIP_TRY(hipEventSynchronize(w.ev_views));  // the previous upload has read the staging
  if ((uint32_t)n_views > w.views_pinned_cap) {
    if (w.views_pinned) NES_HIP_TRY(hipHostFree(w.views_pinned));
    w.views_pinned = nullptr;
    w.views_pinned_cap = 0;
    const uint32_t cap = std::max<uint32_t>((uint32_t)n_views, kPinnedViews);
    NES_HIP_TRY(hipHostMalloc((void **)&w.views_pinned, sizeof(ViewDesc) * cap));
    w.views_pinned_cap = cap;
  }
  std::memcpy(w.views_pinned, w.host_views.data(), sizeof(ViewDesc) * n_views);
  NES_HIP_TRY(hipMemcpyAsync(w.views.ptr, w.views_pinned, sizeof(ViewDesc) * n_views,
                             hipMemcpyHostToDevice, stream));
  NES_HIP_TRY(hipEventRecord(w.ev_views, stream));
  return NES_OK;
}

int render_work_reserve(RenderWork &w, uint64_t max_slots, uint64_t max_px, uint32_t max_views,
                        int precision) {
  if (max_slots > kMaxSlots) return NES_ERANGE;
  NES_TRY(w.views.reserve(sizeof(ViewDesc) * std::max<uint32_t>(max_views, kPinnedViews)));
  NES_TRY(w.rays.reserve(sizeof(RayData) * max_slots));
  NES_TRY(w.acc.reserve(sizeof(RayAcc) * max_slots));
  NES_TRY(w.counts.reserve(sizeof(uint32_t) * max_slots));
  NES_TRY(w.offsets.reserve(sizeof(uint32_t) * max_slots));
  NES_TRY(w.ray_sh.reserve(32 * max_slots));
  NES_TRY(w.tinfo.reserve(8 * max_slots));
  NES_TRY(w.out.reserve(out_layout(max_px).total));
  NES_TRY(w.total.reserve(kTotBytes));
  return render_reserve_samples(w, precision, std::max<size_t>(max_slots * 4, (size_t)1 << 20));
}

int render_reserve_samples(RenderWork &w, int precision, size_t capacity) {
  if (capacity > 0xffffffffull) capacity = 0xffffffffull;
  const size_t fbytes = precision != NES_PRECISION_MIXED ? sizeof(float4) : sizeof(uint2);
  NES_TRY(w.samples.reserve(sizeof(uint32_t) * (capacity + 1)));
  NES_TRY(w.sample_out.reserve(sizeof(float4) * (capacity + 1)));
  NES_TRY(w.sample_pos.reserve(sizeof(float4) * (capacity + 1)));
  NES_TRY(w.feat.reserve(fbytes * kParts * (capacity + 1)));
  w.capacity = (uint32_t)capacity;
  w.capacity_precision = precision;
  return NES_OK;
}

static MarchArgs march_args(RenderWork &w, const uint32_t *occ, int k0, int k1, bool first,
                            uint32_t ray_cap, bool resume) {
  MarchArgs m;
  m.views = w.views.as<ViewDesc>();
  m.n_views = w.n_views;
  m.n_slots = w.n_slots;
  m.occ = occ;
  m.rays = w.rays.as<RayData>();
  m.acc = w.acc.as<RayAcc>();
  m.counts = w.counts.as<uint32_t>();
  m.offsets = w.offsets.as<uint32_t>();
  m.samples = w.samples.as<uint32_t>();
  m.pos = w.sample_pos.as<float4>();
  m.cursor = w.total.as<uint32_t>();
  m.capacity = w.capacity;
  m.cap = ray_cap;
  m.k_begin = k0;
  m.k_end = k1;
  m.first = first ? 1 : 0;
  m.packed = 0;
  m.resume = resume ? 1 : 0;
  m.ray_sh = nullptr;
  m.tinfo = nullptr;
  return m;
}

int render_launch_pass(const SceneDev &scene, RenderWork &w, int precision, int k0, int k1,
                       bool first, bool last, hipStream_t stream, hipEvent_t ev_eval_begin,
                       hipEvent_t ev_eval_end, hipEvent_t ev_encode_end, uint32_t ray_cap,
                       bool resume) {
  const unsigned blocks = (unsigned)((w.n_slots + 255) / 256);
  uint32_t *const total = w.total.as<uint32_t>();
  uint32_t *cursor = total + (first ? kTotCursor : kTotCursorLater);
  const uint32_t cap = w.capacity;
  // a render in one pass keeps no per-ray state between march and composite
  uint2 *tinfo = first && last ? w.tinfo.as<uint2>() : nullptr;
  // the frame counters start with the first pass, a later pass's cursor with
  // that pass
  zero_u32<<<(kTotWords + 255) / 256, 256, 0, stream>>>(first ? total : cursor, first ? kTotWords : 1);
  MarchArgs m = march_args(w, scene.occ, k0, k1, first, ray_cap, resume);
  m.cursor = cursor;
  m.packed = precision == NES_PRECISION_MIXED ? 1 : 0;
  m.ray_sh = precision == NES_PRECISION_MIXED ? w.ray_sh.as<uint2>() : nullptr;
  m.tinfo = tinfo;
  march_emit<<<blocks, 256, 0, stream>>>(m);
  NES_HIP_TRY(hipGetLastError());
  if (ev_eval_begin) NES_HIP_TRY(hipEventRecord(ev_eval_begin, stream));
  // fixed grids with grid-stride loops: the sample count stays on the device
  const uint32_t bpp = part_blocks(cap);
  const bool f32 = precision != NES_PRECISION_MIXED;
  if (f32) {
    encode_parts<float><<<bpp * kParts, 256, 0, stream>>>(w.sample_pos.as<float4>(), cursor, cap,
                                                           scene.table_f32, w.feat.as<float4>(),
                                                           bpp);
    NES_HIP_TRY(hipGetLastError());
    if (ev_encode_end) NES_HIP_TRY(hipEventRecord(ev_encode_end, stream));
    if (precision == NES_PRECISION_FP32_VALU)
      mlp_fp32_valu<<<256 * 4, 256, 0, stream>>>(w.sample_pos.as<float4>(), w.feat.as<float4>(),
                                                 cursor, cap, w.rays.as<RayData>(),
                                                 scene.weights_f32, w.sample_out.as<float4>());
    else
      NES_TRY(launch_mlp_f32(scene.weights_f32, w, cursor, stream));
  } else {
    encode_parts<_Float16><<<bpp * kParts, 256, 0, stream>>>(w.sample_pos.as<uint2>(), cursor,
                                                              cap, scene.table_f16,
                                                              w.feat.as<uint2>(), bpp);
    NES_HIP_TRY(hipGetLastError());
    if (ev_encode_end) NES_HIP_TRY(hipEventRecord(ev_encode_end, stream));
    mlp_mixed<<<256 * 8, 256, 0, stream>>>(w.samples.as<uint32_t>(), w.feat.as<uint2>(), cursor,
                                           cap, w.ray_sh.as<uint2>(),
                                           reinterpret_cast<const half8 *>(scene.wfrag),
                                           w.sample_out.as<uint2>());
  }
  NES_HIP_TRY(hipGetLastError());
  if (ev_eval_end) NES_HIP_TRY(hipEventRecord(ev_eval_end, stream));
  const OutPtrs optr = out_ptrs(w);
  uint32_t *stats = w.total.as<uint32_t>();
  if (f32)
    composite<float4><<<blocks, 256, 0, stream>>>(
        w.n_slots, w.rays.as<RayData>(), w.acc.as<RayAcc>(), w.counts.as<uint32_t>(),
        w.offsets.as<uint32_t>(), w.samples.as<uint32_t>(), w.sample_out.as<float4>(), cursor, cap,
        0, last ? 1 : 0, optr, tinfo, stats);
  else
    composite<uint2><<<blocks, 256, 0, stream>>>(
        w.n_slots, w.rays.as<RayData>(), w.acc.as<RayAcc>(), w.counts.as<uint32_t>(),
        w.offsets.as<uint32_t>(), w.samples.as<uint32_t>(), w.sample_out.as<uint2>(), cursor, cap,
        0, last ? 1 : 0, optr, tinfo, stats);
  NES_HIP_TRY(hipGetLastError());
  // one transfer per pass: its cursor, and after the last pass every word
  // from there on (a copy costs the same few microseconds at any of these
  // sizes)
  const size_t at = (size_t)(cursor - total);
  NES_HIP_TRY(hipMemcpyAsync(w.total_host + at, cursor, last ? 4 * (kTotWords - at) : 4,
                             hipMemcpyDeviceToHost, stream));
  return NES_OK;
}

int render_launch_frame(const SceneDev &scene, RenderWork &w, int precision, hipStream_t stream,
                        const hipEvent_t *ev) {
  w.total_host[1] = 0;  // no second pass unless launched below
  const uint32_t cap1 = w.first_pass_cap;
  auto rec = [&](int i) { return ev ? hipEventRecord(ev[i], stream) : hipSuccess; };
  NES_HIP_TRY(rec(0));
  if (cap1 >= (uint32_t)kMaxSteps) {
    w.frame_passes = 1;
    return render_launch_pass(scene, w, precision, 0, kMaxSteps, true, true, stream,
                              ev ? ev[1] : nullptr, ev ? ev[3] : nullptr, ev ? ev[2] : nullptr);
  }
  // two passes, no host round trip between them: the second resumes every
  // ray the first left above T = 1e-4 at the step after its last sample
  w.frame_passes = 2;
  NES_TRY(render_launch_pass(scene, w, precision, 0, kMaxSteps, true, false, stream,
                             ev ? ev[1] : nullptr, ev ? ev[3] : nullptr, ev ? ev[2] : nullptr,
                             cap1, false));
  NES_HIP_TRY(rec(4));
  return render_launch_pass(scene, w, precision, 0, kMaxSteps, false, true, stream,
                            ev ? ev[5] : nullptr, ev ? ev[7] : nullptr, ev ? ev[6] : nullptr,
                            kMaxSteps, true);
}

uint32_t render_frame_max_total(const RenderWork &w) {
  return std::max(w.total_host[0], w.frame_passes > 1 ? w.total_host[1] : 0u);
}

int render_launch_bounded(const SceneDev &scene, RenderWork &w, int precision,
                          hipStream_t stream) {
  if (w.capacity < w.n_slots) {
    set_last_error("sample buffers hold less than one sample per ray");
    return NES_ENOMEM;
  }
  // each pass emits at most n_slots * cap <= capacity samples: no overflow
  const uint32_t cap = std::min<uint32_t>(w.capacity / std::max<uint32_t>(w.n_slots, 1u),
                                          kMaxSteps);
  const int passes = (int)((kMaxSteps + cap - 1) / cap);
  w.total_host[1] = 0;
  w.frame_passes = (uint32_t)passes;
  for (int p = 0; p < passes; ++p)
    NES_TRY(render_launch_pass(scene, w, precision, 0, kMaxSteps, p == 0, p == passes - 1, stream,
                               nullptr, nullptr, nullptr, cap, p > 0));
  return NES_OK;
}

bool render_growth_fits(const RenderWork &w, int precision, size_t capacity) {
  if (!w.budget) return true;
  const size_t cur = w.samples.bytes + w.sample_out.bytes + w.sample_pos.bytes + w.feat.bytes;
  return work_device_bytes(w) - cur + sample_buffer_bytes(capacity, precision) <= w.budget;
}

size_t sample_buffer_bytes(size_t capacity, int precision) {
  const size_t c = capacity + 1, fb = precision != NES_PRECISION_MIXED ? sizeof(float4) : sizeof(uint2);
  size_t t = 0;
  // the buffers of render_reserve_samples, with DevBuf's growth slack
  for (size_t need : {sizeof(uint32_t) * c, sizeof(float4) * c, sizeof(float4) * c, fb * kParts * c})
    t += need + need / 8 + 256;
  return t;
}

// Early-termination policy for the next frame of this work area.  A single
// pass evaluates every occupied sample; when more than 5 % of them lie behind
// the point where T fell below 1e-4, the next frame runs two passes, the first
// giving each ray as many samples as the terminated rays needed on average
// (rounded up).  Multi-pass frames go back to one pass when the first pass
// stopped fewer than 1 % of the pixels with steps still ahead of them.
FrameCounts render_frame_finish(RenderWork &w) {
  FrameCounts c;
  c.passes = w.frame_passes;
  c.evaluated = w.total_host[2];
  for (int s = 0; s < kTotTermSlots; ++s) {
    const uint32_t *t = w.total_host + kTotTermBase + kTotTermStride * s;
    c.terminated += t[0];
    c.term_samples += t[1];
    c.wasted += t[2];
    c.saved_rays += t[3];
  }
  c.composited = c.evaluated - std::min(c.wasted, c.evaluated);
  // one pass: switch to two when more than 5 % of the evaluated samples lay
  // behind T < 1e-4; several passes: keep two while more than 1 % of the
  // pixels stopped early with steps left (the work the first pass saved)
  const bool wasteful = c.wasted * 20 > c.evaluated;
  const bool saving = c.saved_rays * 100 > w.out_px;
  if (c.terminated && ((c.passes == 1 && wasteful) || (c.passes > 1 && saving))) {
    const uint64_t m = (c.term_samples + c.terminated - 1) / c.terminated;
    w.first_pass_cap = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(m, 1), kMaxSteps - 1);
  } else {
    w.first_pass_cap = kMaxSteps;
  }
  return c;
}

// Single-stage launchers reused by the training step (rays and compositing
// state already initialised by the caller, one output index per ray slot).
int launch_march_emit(const uint32_t *occ, RenderWork &w, int k0, int k1, hipStream_t stream) {
  const unsigned blocks = (unsigned)((w.n_slots + 255) / 256);
  zero_u32<<<1, 64, 0, stream>>>(w.total.as<uint32_t>(), 2);
  march_emit<<<blocks, 256, 0, stream>>>(march_args(w, occ, k0, k1, false, kMaxSteps, false));
  NES_HIP_TRY(hipGetLastError());
  return NES_OK;
}

int launch_encode_f32(const float *table, RenderWork &w, hipStream_t stream) {
  const uint32_t bpp = part_blocks(w.capacity);
  encode_parts<float><<<bpp * kParts, 256, 0, stream>>>(w.sample_pos.as<float4>(),
                                                         w.total.as<uint32_t>(), w.capacity,
                                                         table, w.feat.as<float4>(), bpp);
  NES_HIP_TRY(hipGetLastError());
  return NES_OK;
}

int launch_composite(RenderWork &w, bool last, hipStream_t stream) {
  const unsigned blocks = (unsigned)((w.n_slots + 255) / 256);
  composite<float4><<<blocks, 256, 0, stream>>>(w.n_slots, w.rays.as<RayData>(), w.acc.as<RayAcc>(),
                                               w.counts.as<uint32_t>(), w.offsets.as<uint32_t>(),
                                               w.samples.as<uint32_t>(), w.sample_out.as<float4>(),
                                               w.total.as<uint32_t>(), w.capacity, 1, last ? 1 : 0,
                                               out_ptrs(w), nullptr, nullptr);
  NES_HIP_TRY(hipGetLastError());
  return NES_OK;
}

// Initial sample capacity: four samples per ray slot, or what the last render
// of this work area needed.
int render_prepare_capacity(RenderWork &w, int precision) {
  size_t want = std::max<size_t>((size_t)w.n_slots * 4, 1u << 20);
  if (w.capacity >= want && w.capacity_precision == precision) return NES_OK;
  want = std::max<size_t>(want, w.capacity);
  return render_reserve_samples(w, precision, want);
}

size_t work_device_bytes(const RenderWork &w) {
  size_t b = 0;
  for (const DevBuf *d : {&w.views, &w.rays, &w.acc, &w.counts, &w.offsets, &w.ray_sh, &w.tinfo, &w.gray, &w.samples,
                          &w.sample_out, &w.sample_pos, &w.feat, &w.out, &w.total})
    b += d->bytes;
  return b;
}

// Over budget: drop every buffer (they only grow) so that a later, smaller
// render starts from nothing, and report NES_ENOMEM.
static int check_budget(RenderWork &w) {
  if (w.budget && work_device_bytes(w) > w.budget) {
    for (DevBuf *d : {&w.views, &w.rays, &w.acc, &w.counts, &w.offsets, &w.ray_sh, &w.tinfo, &w.gray, &w.samples,
                      &w.sample_out, &w.sample_pos, &w.feat, &w.out, &w.total})
      d->release();
    w.capacity = 0;
    w.capacity_precision = -1;
    w.n_views = 0;
    set_last_error("render buffers exceed the context's HBM budget");
    return NES_ENOMEM;
  }
  return NES_OK;
}

static int render_enqueue_inner(const SceneDev &scene, RenderWork &w, int precision,
                                const nes_camera *views, int n_views, int steps_per_pass,
                                hipStream_t stream, nes_render_stats *stats);

int render_enqueue(const SceneDev &scene, RenderWork &w, int precision,
                   const nes_camera *views, int n_views, int steps_per_pass,
                   hipStream_t stream, nes_render_stats *stats) {
  const int rc = render_enqueue_inner(scene, w, precision, views, n_views, steps_per_pass, stream,
                                      stats);
  if (rc != NES_OK) {
    // a failed render leaves nothing to fetch
    w.n_views = 0;
    w.out_px = 0;
  }
  return rc;
}

static bool valid_precision(int p) {
  return p == NES_PRECISION_FP32 || p == NES_PRECISION_MIXED || p == NES_PRECISION_FP32_VALU;
}

// Stage times of one pass from its four events and the event that follows it.
static void add_pass_times(const hipEvent_t *ev, hipEvent_t next, double t[4]) {
  float a = 0, b = 0, c = 0, d = 0;
  hipEventElapsedTime(&a, ev[0], ev[1]);
  hipEventElapsedTime(&b, ev[1], ev[2]);
  hipEventElapsedTime(&c, ev[2], ev[3]);
  hipEventElapsedTime(&d, ev[3], next);
  t[0] += a;  // march
  t[1] += b;  // encode
  t[2] += c;  // MLPs
  t[3] += d;  // composite
}

static int render_enqueue_inner(const SceneDev &scene, RenderWork &w, int precision,
                                const nes_camera *views, int n_views, int steps_per_pass,
                                hipStream_t stream, nes_render_stats *stats) {
  if (!valid_precision(precision) || steps_per_pass < 0) return NES_EINVAL;
  NES_TRY(render_setup(w, views, n_views, stream));
  NES_TRY(render_prepare_capacity(w, precision));
  NES_TRY(check_budget(w));
  double t[4] = {0, 0, 0, 0};  // march, encode, MLPs, composite
  FrameCounts fc;
  hipEvent_t *ev = w.ev + 1, end = w.ev[9];
  NES_HIP_TRY(hipEventRecord(w.ev[0], stream));
  if (steps_per_pass == 0 || steps_per_pass >= kMaxSteps) {
    // whole frame: one pass, or two when the previous frame of this work area
    // evaluated samples behind terminated rays (steps_per_pass 0)
    if (steps_per_pass) w.first_pass_cap = kMaxSteps;
    bool bounded = false;
    for (int attempt = 0;; ++attempt) {
      if (bounded) {
        NES_HIP_TRY(hipEventRecord(ev[0], stream));
        NES_TRY(render_launch_bounded(scene, w, precision, stream));
      } else {
        NES_TRY(render_launch_frame(scene, w, precision, stream, ev));
      }
      NES_HIP_TRY(hipEventRecord(end, stream));
      NES_HIP_TRY(hipEventSynchronize(end));
      const uint32_t n = render_frame_max_total(w);
      if (bounded || n <= w.capacity) break;
      if (attempt > 0) {
        set_last_error("sample capacity still exceeded after regrowth");
        return NES_ENOMEM;
      }
      // grow when the budget allows it, else render within the buffers
      const size_t want = (size_t)n + n / 8 + 1024;
      if (render_growth_fits(w, precision, want))
        NES_TRY(render_reserve_samples(w, precision, want));
      else
        bounded = true;
    }
    if (bounded) {
      float a = 0;
      hipEventElapsedTime(&a, ev[0], end);
      t[1] += a;  // no per-stage events over bounded passes
    } else {
      add_pass_times(ev, w.frame_passes > 1 ? ev[4] : end, t);
      if (w.frame_passes > 1) add_pass_times(ev + 4, end, t);
    }
    fc = render_frame_finish(w);
  } else {
    // fixed windows of steps_per_pass march steps, one host round trip each
    const int spp = steps_per_pass;
    const int passes = (kMaxSteps + spp - 1) / spp;
    w.frame_passes = 1;
    for (int p = 0; p < passes; ++p) {
      const int k0 = p * spp, k1 = std::min(kMaxSteps, k0 + spp);
      uint32_t n = 0;
      for (int attempt = 0;; ++attempt) {
        NES_HIP_TRY(hipEventRecord(ev[0], stream));
        NES_TRY(render_launch_pass(scene, w, precision, k0, k1, p == 0, p == passes - 1, stream,
                                   ev[1], ev[3], ev[2]));
        NES_HIP_TRY(hipEventRecord(end, stream));
        NES_HIP_TRY(hipEventSynchronize(end));
        n = w.total_host[p == 0 ? 0 : 1];
        if (n <= w.capacity) break;
        if (attempt > 0) {
          set_last_error("sample capacity still exceeded after regrowth");
          return NES_ENOMEM;
        }
        const size_t want = (size_t)n + n / 8 + 1024;
        if (!render_growth_fits(w, precision, want)) {
          set_last_error("a pass's samples exceed the context's HBM budget");
          return NES_ENOMEM;
        }
        NES_TRY(render_reserve_samples(w, precision, want));
      }
      add_pass_times(ev, end, t);
    }
    w.frame_passes = (uint32_t)passes;
    fc = render_frame_finish(w);
    w.first_pass_cap = kMaxSteps;  // windows are the caller's choice, not learned from
  }
  float total = 0;
  hipEventElapsedTime(&total, w.ev[0], end);
  if (stats) {
    stats->rays = w.out_px;
    stats->samples = fc.evaluated;
    stats->ms_total = total;
    stats->ms_march = t[0];
    stats->ms_encode = t[1];
    stats->ms_mlp = t[2];
    stats->ms_eval = t[1] + t[2];
    stats->ms_composite = t[3];
    stats->composited = fc.composited;
    stats->terminated_rays = fc.terminated;
    stats->passes = fc.passes;
  }
  return NES_OK;
}

int render_fetch(RenderWork &w, int view, const nes_view_out *out, hipStream_t stream) {
  if (!out || view < 0 || (uint32_t)view >= w.n_views) return NES_EINVAL;
  const ViewDesc &v = w.host_views[(size_t)view];
  const size_t px = (size_t)v.width * v.height, base = v.px_base;
  if (base + px > w.out_px) {
    set_last_error("view lies outside the last render's outputs");
    return NES_EINVAL;
  }
  const OutLayout L = out_layout(w.out_px);
  const uint8_t *o = w.out.as<uint8_t>();
  if (out->rgb24)
    NES_HIP_TRY(hipMemcpyAsync(out->rgb24, o + L.rgb24 + 3 * base, 3 * px, hipMemcpyDeviceToHost, stream));
  if (out->depth)
    NES_HIP_TRY(hipMemcpyAsync(out->depth, o + L.depth + 4 * base, 4 * px, hipMemcpyDeviceToHost, stream));
  if (out->rgb_f32)
    NES_HIP_TRY(hipMemcpyAsync(out->rgb_f32, o + L.rgbf + 12 * base, 12 * px, hipMemcpyDeviceToHost, stream));
  if (out->n_samples)
    NES_HIP_TRY(hipMemcpyAsync(out->n_samples, o + L.nsamp + 2 * base, 2 * px, hipMemcpyDeviceToHost, stream));
  if (out->depth_gray8) {
    NES_TRY(w.gray.reserve(px));
    depth_gray8_kernel<<<(unsigned)((px + 255) / 256), 256, 0, stream>>>(
        reinterpret_cast<const float *>(o + L.depth) + base, px, depth_scale(out->depth_far),
        w.gray.as<uint8_t>());
    NES_HIP_TRY(hipGetLastError());
    NES_HIP_TRY(hipMemcpyAsync(out->depth_gray8, w.gray.ptr, px, hipMemcpyDeviceToHost, stream));
  }
  NES_HIP_TRY(hipStreamSynchronize(stream));
  return NES_OK;
}

std::shared_ptr<SceneDev> new_scene() {
  return std::shared_ptr<SceneDev>(new SceneDev(), [](SceneDev *p) {
    scene_free(*p);
    delete p;
  });
}

std::shared_ptr<SceneDev> ctx_scene(nes_ctx *ctx) {
  std::lock_guard<std::mutex> lk(ctx->scene_mu);
  return ctx->scene;
}

}  // namespace nes

using namespace nes;

extern "C" {

int nes_ctx_create(const nes_scene_desc *scene, uint64_t hbm_budget_bytes, nes_ctx **out) {
  if (!scene || !out) return NES_EINVAL;
  NES_TRY(ensure_device());
  nes_ctx *c = new nes_ctx();
  c->budget = hbm_budget_bytes ? hbm_budget_bytes : (16ull << 30);
  c->work.budget = c->budget;
  c->scene = new_scene();
  int rc = scene_upload(*c->scene, scene);
  if (rc == NES_OK && hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) rc = NES_EIO;
  if (rc != NES_OK) {
    delete c;
    return rc;
  }
  *out = c;
  return NES_OK;
}

int nes_ctx_destroy(nes_ctx *ctx) {
  if (!ctx) return NES_EINVAL;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    render_work_free(ctx->work);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
  }
  delete ctx;  // frees the scene copies (servers on it are destroyed first)
  return NES_OK;
}

int nes_ctx_update_scene(nes_ctx *ctx, const nes_scene_desc *scene) {
  if (!ctx || !scene || !scene->table || !scene->weights || !scene->occupancy) return NES_EINVAL;
  std::shared_ptr<SceneDev> next;
  {
    // reuse the previous copy once no render or batch holds it any more
    std::lock_guard<std::mutex> lk(ctx->scene_mu);
    if (ctx->spare && ctx->spare.use_count() == 1) next = ctx->spare;
    ctx->spare.reset();
  }
  if (!next) next = new_scene();
  // the copy being written is referenced by no launched work
  NES_TRY(scene_upload(*next, scene));
  std::lock_guard<std::mutex> lk(ctx->scene_mu);
  ctx->spare = ctx->scene;
  ctx->scene = next;
  return NES_OK;
}

int nes_render_device(nes_ctx *ctx, int precision, const nes_camera *views, int n_views,
                      int steps_per_pass, nes_render_stats *stats) {
  if (!ctx) return NES_EINVAL;
  std::lock_guard<std::mutex> lk(ctx->mu);
  const std::shared_ptr<SceneDev> sc = ctx_scene(ctx);
  return render_enqueue(*sc, ctx->work, precision, views, n_views, steps_per_pass, ctx->stream,
                        stats);
}

int nes_render_fetch(nes_ctx *ctx, int view, const nes_view_out *out) {
  if (!ctx) return NES_EINVAL;
  std::lock_guard<std::mutex> lk(ctx->mu);
  return render_fetch(ctx->work, view, out, ctx->stream);
}

int nes_render(nes_ctx *ctx, int precision, const nes_camera *views, int n_views,
               const nes_view_out *outs, int steps_per_pass, nes_render_stats *stats) {
  if (!ctx || !outs) return NES_EINVAL;
  std::lock_guard<std::mutex> lk(ctx->mu);
  const std::shared_ptr<SceneDev> sc = ctx_scene(ctx);
  NES_TRY(render_enqueue(*sc, ctx->work, precision, views, n_views, steps_per_pass, ctx->stream,
                         stats));
  for (int i = 0; i < n_views; ++i) NES_TRY(render_fetch(ctx->work, i, &outs[i], ctx->stream));
  return NES_OK;
}

}  // extern "C"
