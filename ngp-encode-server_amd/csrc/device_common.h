// device_common.h -- device helpers shared by the render and training
// kernels: FP32 sample geometry, SH basis, table access and the
// XCD-partitioned level layout.
#pragma once

#include "render.h"

namespace nes {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef _Float16 half2v __attribute__((ext_vector_type(2)));
typedef float float4v __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------- geometry
__device__ __forceinline__ void make_ray(const ViewDesc &v, uint32_t px,
                                         uint32_t py, float o[3], float d[3]) {
#pragma clang fp contract(off)
  const float u = (((float)px + 0.5f) - 0.5f * (float)v.width) / v.focal;
  const float w = (((float)py + 0.5f) - 0.5f * (float)v.height) / v.focal;
  float dw[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    dw[i] = fmaf(v.c2w[4 * i + 0], u, fmaf(v.c2w[4 * i + 1], w, v.c2w[4 * i + 2]));
  const float n = sqrtf(fmaf(dw[0], dw[0], fmaf(dw[1], dw[1], dw[2] * dw[2])));
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    d[i] = dw[i] / n;
    o[i] = v.c2w[4 * i + 3];
  }
}

__device__ __forceinline__ float step_t(int k, float t_start) {
  return fmaf((float)k + 0.5f, kDelta, t_start);
}

// Returns the number of candidate steps (<= 128) inside [0,1]^3.
__device__ __forceinline__ uint32_t ray_box(const float o[3], const float d[3],
                                            float *t_start) {
#pragma clang fp contract(off)
  float tn = -INFINITY, tf = INFINITY;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (d[a] == 0.0f) {
      if (o[a] < 0.0f || o[a] > 1.0f) return 0;
      continue;
    }
    const float t0 = (0.0f - o[a]) / d[a];
    const float t1 = (1.0f - o[a]) / d[a];
    const float lo = t0 < t1 ? t0 : t1;
    const float hi = t0 < t1 ? t1 : t0;
    tn = lo > tn ? lo : tn;
    tf = hi < tf ? hi : tf;
  }
  const float ts = tn > 0.0f ? tn : 0.0f;
  *t_start = ts;
  if (!(tf > ts)) return 0;
  uint32_t n = 0;
  while (n < (uint32_t)kMaxSteps && step_t((int)n, ts) < tf) ++n;
  return n;
}

// Occupancy cell of a position (clamped to the grid, so any position gives
// an index inside the bitfield).
__device__ __forceinline__ uint32_t occ_cell(const float x[3]) {
  int c[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int ci = (int)floorf(x[a] * 128.0f);
    c[a] = min(max(ci, 0), kOccRes - 1);
  }
  return (uint32_t)(c[0] + kOccRes * (c[1] + kOccRes * c[2]));
}

__device__ __forceinline__ void step_pos(const RayData &r, int k, float x[3],
                                         float *t) {
  const float tt = step_t(k, r.t_start);
#pragma unroll
  for (int a = 0; a < 3; ++a) x[a] = fmaf(tt, r.d[a], r.o[a]);
  *t = tt;
}

__device__ __forceinline__ RayData load_ray(const RayData *rays, uint32_t slot) {
  const float4 *p = reinterpret_cast<const float4 *>(rays + slot);
  const float4 a = p[0], b = p[1];
  RayData r;
  r.o[0] = a.x; r.o[1] = a.y; r.o[2] = a.z;
  r.d[0] = a.w; r.d[1] = b.x; r.d[2] = b.y;
  r.t_start = b.z;
  r.n_steps = __float_as_uint(b.w);
  return r;
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

// ------------------------------------------------- rank-major sample list
// The samples of one wavefront's 64 ray slots (one 8x8 pixel tile) occupy
// one contiguous range of the sample list, reserved with a single atomic.
// Inside it they are ordered by rank: first the first occupied sample of
// every ray that has one (in lane order), then every second sample, and so
// on.  Rank i of lane l sits at
//   base + sum_{r<i} |{lanes with count > r}| + |{lanes < l with count > i}|.
// Neighbouring lanes of the encode and MLP kernels thus see samples of
// neighbouring pixels at similar depths (their table gathers share cache
// lines), and a ray's samples are still visited in march order.  Every
// kernel that walks the list by ray runs one lane per slot with
// slot = 64 * wavefront + lane and keeps all 64 lanes in the loop.
__device__ __forceinline__ uint64_t lanes_below(int lane) {
  return lane ? (~0ull >> (64 - lane)) : 0ull;
}

__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o));
  return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}


// --------------------------------------------------------------- encoding
__device__ __forceinline__ void sh4(const float d[3], float sh[16]) {
  const float x = d[0], y = d[1], z = d[2];
  const float xx = x * x, yy = y * y, zz = z * z;
  sh[0] = 0.28209479177387814f;
  sh[1] = -0.48860251190291987f * y;
  sh[2] = 0.48860251190291987f * z;
  sh[3] = -0.48860251190291987f * x;
  sh[4] = 1.0925484305920792f * x * y;
  sh[5] = -1.0925484305920792f * y * z;
  sh[6] = 0.94617469575755997f * zz - 0.31539156525251999f;
  sh[7] = -1.0925484305920792f * x * z;
  sh[8] = 0.54627421529603959f * xx - 0.54627421529603959f * yy;
  sh[9] = 0.59004358992664352f * y * (-3.0f * xx + yy);
  sh[10] = 2.8906114426405538f * x * y * z;
  sh[11] = 0.45704579946446572f * y * (1.0f - 5.0f * zz);
  sh[12] = 0.3731763325901154f * z * (5.0f * zz - 3.0f);
  sh[13] = 0.45704579946446572f * x * (1.0f - 5.0f * zz);
  sh[14] = 1.4453057213202769f * z * (xx - yy);
  sh[15] = 0.59004358992664352f * x * (-xx + 3.0f * yy);
}

// One table entry (2 features) as stored: a float2, or two FP16 in a word.
template <class Tab> struct RawFeat;
template <> struct RawFeat<float> {
  typedef float2 type;
  static __device__ __forceinline__ float2 unpack(float2 v) { return v; }
};
template <> struct RawFeat<_Float16> {
  typedef uint32_t type;
  static __device__ __forceinline__ float2 unpack(uint32_t bits) {
    const half2v h = __builtin_bit_cast(half2v, bits);
    return make_float2((float)h.x, (float)h.y);
  }
};

__device__ __forceinline__ float sigmoidf_(float v) { return 1.0f / (1.0f + expf(-v)); }

// Orders a wavefront's own LDS writes before its later LDS reads of other
// lanes' data (LDS ops of one wavefront complete in order).
// FP32 atomic add into LDS as a compare-and-swap loop: the add is IEEE FP32
// of the value found, retried when another lane changed it in between.  On
// gfx950 the native FP32 LDS atomic add costs about 3 cycles per lane; a
// 32-bit compare-and-swap costs a small fraction of that.
__device__ __forceinline__ void lds_add_f32(float *addr, float v) {
  uint32_t *w = reinterpret_cast<uint32_t *>(addr);
  uint32_t seen = *w;
  for (;;) {
    const uint32_t prev = atomicCAS(w, seen, __float_as_uint(__uint_as_float(seen) + v));
    if (prev == seen) break;
    seen = prev;
  }
}

// Both FP32 features of one entry (8-byte aligned) in one 64-bit
// compare-and-swap loop: one LDS operation where lds_add_f32 needs two.
__device__ __forceinline__ void lds_add_f32x2(float *addr, float a, float b) {
  unsigned long long *w = reinterpret_cast<unsigned long long *>(addr);
  unsigned long long seen = *w;
  for (;;) {
    const float lo = __uint_as_float((uint32_t)seen) + a;
    const float hi = __uint_as_float((uint32_t)(seen >> 32)) + b;
    const unsigned long long want =
        (unsigned long long)__float_as_uint(lo) | ((unsigned long long)__float_as_uint(hi) << 32);
    const unsigned long long prev = atomicCAS(w, seen, want);
    if (prev == seen) break;
    seen = prev;
  }
}

__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}


// ------------------------------------------------- XCD-partitioned levels
constexpr int kParts = 8;

struct PartLevels {
  float scale[2];
  uint32_t n1[2], n1sq[2], off[2];
  bool dense[2];
};

// Levels of part p (runtime form, used by the training scatter).
__device__ __forceinline__ PartLevels part_levels(int part) {
  PartLevels p;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int l = q == 0 ? part : 15 - part;
    int res = 0;
    uint32_t off = 0;
#pragma unroll
    for (int c = 0; c < 16; ++c)
      if (l == c) {
        res = kLevelRes[c];
        off = kDevLevelOffset[c];
      }
    p.scale[q] = (float)res;
    p.n1[q] = (uint32_t)res + 1u;
    p.n1sq[q] = p.n1[q] * p.n1[q];
    p.off[q] = off;
    p.dense[q] = l < kDenseLevels;
  }
  return p;
}

// Compile-time constants of level L for table type Tab (FP16: 4-byte
// entries, FP32: 8-byte entries; both in the device layout, each level
// 32-entry aligned).
template <class Tab, int L>
struct LevelC {
  static constexpr bool kF16 = std::is_same<Tab, _Float16>::value;
  static constexpr float kScale = (float)kLevelRes[L];
  static constexpr uint32_t kN1 = (uint32_t)kLevelRes[L] + 1u;
  static constexpr bool kDense = L < kDenseLevels;
  static constexpr uint32_t kOff = kDevLevelOffset[L];
  static constexpr int kShift = kF16 ? 2 : 3;  // log2 bytes per entry
};

// Byte offsets (within level L) of the 8 corners around xc and the
// trilinear fractions.  Corner c = (x + (c & 1), y + (c >> 1 & 1),
// z + (c >> 2)); the arithmetic is the reference's index (dense: x + y N1 +
// z N1^2; hashed: (x ^ y P1 ^ z P2) mod 2^19) carried out directly in byte
// units: x < 2^19, so ((x ^ h) & M) << s == (x << s) ^ ((h << s) & (M << s)),
// and (y P1) << s == y (P1 << s) mod 2^32.
template <class Tab, int L>
__device__ __forceinline__ void level_corners(const float xc[3], uint32_t boff[8], float fr[3]) {
  using C = LevelC<Tab, L>;
  constexpr int S = C::kShift;
  uint32_t i0[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float pp = xc[a] * C::kScale;
    const float fl = fminf(floorf(pp), C::kScale - 1.0f);
    i0[a] = (uint32_t)fl;
    fr[a] = pp - fl;
  }
  const uint32_t x0 = i0[0] << S, x1 = x0 + (1u << S);
  if constexpr (C::kDense) {
    constexpr uint32_t dy = C::kN1 << S, dz = (C::kN1 * C::kN1) << S;
    const uint32_t b = x0 + __umul24(i0[1], dy) + __umul24(i0[2], dz);
#pragma unroll
    for (int c = 0; c < 8; ++c)
      boff[c] = b + ((c & 1) ? (1u << S) : 0u) + ((c & 2) ? dy : 0u) + ((c & 4) ? dz : 0u);
  } else {
    constexpr uint32_t m = ((1u << kLog2T) - 1u) << S;
    constexpr uint32_t p1 = kPrime1 << S, p2 = kPrime2 << S;
    const uint32_t hy0 = i0[1] * p1, hy1 = hy0 + p1, hz0 = i0[2] * p2, hz1 = hz0 + p2;
    const uint32_t h[4] = {(hy0 ^ hz0) & m, (hy1 ^ hz0) & m, (hy0 ^ hz1) & m, (hy1 ^ hz1) & m};
#pragma unroll
    for (int c = 0; c < 8; ++c) boff[c] = ((c & 1) ? x1 : x0) ^ h[c >> 1];
  }
}

template <class Tab>
__device__ __forceinline__ typename RawFeat<Tab>::type load_entry(const char *level_base,
                                                                  uint32_t boff) {
  return __ldg(reinterpret_cast<const typename RawFeat<Tab>::type *>(level_base + boff));
}

// Entries e and e + 1 with one load (8 B FP16 / 16 B FP32 at the alignment
// of one entry: gfx950 serves unaligned multi-dword loads).
typedef uint2 __attribute__((aligned(4))) EntryPairF16;
typedef float4 __attribute__((aligned(8))) EntryPairF32;
__device__ __forceinline__ void load_pair(const char *level_base, uint32_t boff, uint32_t &a,
                                          uint32_t &b) {
  const EntryPairF16 p = *reinterpret_cast<const EntryPairF16 *>(level_base + boff);
  a = p.x;
  b = p.y;
}
__device__ __forceinline__ void load_pair(const char *level_base, uint32_t boff, float2 &a,
                                          float2 &b) {
  const EntryPairF32 p = *reinterpret_cast<const EntryPairF32 *>(level_base + boff);
  a = make_float2(p.x, p.y);
  b = make_float2(p.z, p.w);
}

__device__ __forceinline__ uint32_t pick4(const uint4 &q, uint32_t i) {
  const uint32_t lo = (i & 1u) ? q.y : q.x, hi = (i & 1u) ? q.w : q.z;
  return (i & 2u) ? hi : lo;
}

// The 8 corner entries of level L.  The gathers are bound by the vector
// L1's address rate: one address per lane and instruction, whatever the
// width up to 16 B, and a lane masked off still takes its slot.
// * Dense levels: corners c and c + 1 are the adjacent entries e and e + 1,
//   one load per pair.
// * Hashed FP16 levels: corners c and c + 1 of a (y, z) pair are entries h
//   and h ^ (x ^ (x + 1)), in one aligned 16 B chunk (the level starts on a
//   128 B boundary) unless x % 4 == 3 -- the same for all four pairs of a
//   lane.  One chunk load per pair serves both corners; the second corners
//   of the lanes with x % 4 == 3 (a quarter of them, four loads each) are
//   loaded by the whole wavefront, packed into as few instructions as their
//   count needs, through the wavefront's 256-word LDS exchange `xch`.  The
//   finest levels (kChunkedMaxLevel and up) keep one gather per corner:
//   there almost every corner misses L1, and the 16 B loads made the XCDs
//   encoding them slower (measured per XCD).
constexpr int kChunkedMaxLevel = 13;
template <class Tab, int L>
__device__ __forceinline__ void load_corners(const char *base, const uint32_t boff[8],
                                             typename RawFeat<Tab>::type v[8], uint32_t *xch) {
  if constexpr (LevelC<Tab, L>::kDense) {
#pragma unroll
    for (int c = 0; c < 8; c += 2) load_pair(base, boff[c], v[c], v[c + 1]);
  } else if constexpr (LevelC<Tab, L>::kF16 && L <= kChunkedMaxLevel) {
    const bool same = ((boff[0] ^ boff[1]) >> 4) == 0u;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const uint32_t b0 = boff[2 * p], b1 = boff[2 * p + 1];
      const uint4 q = __ldg(reinterpret_cast<const uint4 *>(base + (b0 & ~15u)));
      v[2 * p] = pick4(q, (b0 >> 2) & 3u);
      v[2 * p + 1] = pick4(q, (b1 >> 2) & 3u);
    }
    const uint64_t needy = __ballot(!same);
    if (needy) {  // wavefront-uniform
      const int lane = lane_id();
      const uint64_t below = lanes_below(lane);
      const uint64_t act = __ballot(true);
      const uint32_t r = (uint32_t)__popcll(needy & below);
      const uint32_t slots = 4u * (uint32_t)__popcll(needy);
      const uint32_t me = (uint32_t)__popcll(act & below), n_act = (uint32_t)__popcll(act);
      if (!same) {
#pragma unroll
        for (int p = 0; p < 4; ++p) xch[4 * r + p] = boff[2 * p + 1];
      }
      wave_lds_fence();
      for (uint32_t t = me; t < slots; t += n_act) xch[t] = load_entry<Tab>(base, xch[t]);
      wave_lds_fence();
      if (!same) {
#pragma unroll
        for (int p = 0; p < 4; ++p) v[2 * p + 1] = xch[4 * r + p];
      }
      wave_lds_fence();  // the exchange is free for the next level
    }
  } else {
#pragma unroll
    for (int c = 0; c < 8; ++c) v[c] = load_entry<Tab>(base, boff[c]);
  }
}

__device__ __forceinline__ float2 trilinear(const float fr[3], const float2 fv[8]) {
  const float gx = 1.0f - fr[0], gy = 1.0f - fr[1], gz = 1.0f - fr[2];
  float a0 = 0.0f, a1 = 0.0f;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const float w = ((c & 1) ? fr[0] : gx) * (((c >> 1) & 1) ? fr[1] : gy) *
                    (((c >> 2) & 1) ? fr[2] : gz);
    a0 = fmaf(w, fv[c].x, a0);
    a1 = fmaf(w, fv[c].y, a1);
  }
  return make_float2(a0, a1);
}

// Features of levels LA and LB at one position; all 16 gathers are issued
// before the first is used.  Each level's base is a wave-uniform pointer, so
// a gather is one 32-bit VALU offset plus a scalar-based load.
template <class Tab, int LA, int LB>
__device__ __forceinline__ void encode_two(const char *base_a, const char *base_b,
                                           const float xc[3], float2 &fa, float2 &fb,
                                           uint32_t *xch) {
  uint32_t oa[8], ob[8];
  float fra[3], frb[3];
  level_corners<Tab, LA>(xc, oa, fra);
  level_corners<Tab, LB>(xc, ob, frb);
  typename RawFeat<Tab>::type va[8], vb[8];
  load_corners<Tab, LA>(base_a, oa, va, xch);
  load_corners<Tab, LB>(base_b, ob, vb, xch);
  float2 fva[8], fvb[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    fva[c] = RawFeat<Tab>::unpack(va[c]);
    fvb[c] = RawFeat<Tab>::unpack(vb[c]);
  }
  fa = trilinear(fra, fva);
  fb = trilinear(frb, fvb);
}

template <class Tab, int L>
__device__ __forceinline__ const char *level_base(const Tab *table) {
  return reinterpret_cast<const char *>(table) + ((size_t)LevelC<Tab, L>::kOff << LevelC<Tab, L>::kShift);
}

// Sample positions as the encode reads them.  The FP32 path keeps the exact
// float4 (x, y, z, slot).  The mixed path rounds x, y, z in [0, 1] to
// multiples of 2^-21 and packs them 21 bits each into 8 bytes: every XCD part
// reads every position, so this halves the encode's position stream; the
// rounding (at most 2^-22) moves a level-15 interpolation weight by at most
// 2048 * 2^-22 = 4.9e-4, below the FP16 table's own rounding.
__device__ __forceinline__ uint2 pack_pos21(const float x[3]) {
  uint32_t q[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) q[a] = min(__float2uint_rn(x[a] * 2097152.0f), 2097151u);
  return make_uint2(q[0] | (q[1] << 21), (q[1] >> 11) | (q[2] << 10));
}

template <class Tab> struct PosIn;
template <> struct PosIn<float> {
  typedef float4 type;
  static __device__ __forceinline__ void get(const float4 &p, float x[3]) {
    x[0] = p.x;
    x[1] = p.y;
    x[2] = p.z;
  }
};
template <> struct PosIn<_Float16> {
  typedef uint2 type;
  static __device__ __forceinline__ void get(const uint2 &p, float x[3]) {
    x[0] = (float)(p.x & 0x1FFFFFu) * 0x1p-21f;
    x[1] = (float)((p.x >> 21) | ((p.y & 0x3FFu) << 11)) * 0x1p-21f;
    x[2] = (float)(p.y >> 10) * 0x1p-21f;
  }
};

template <class Tab> struct PartOut;
template <> struct PartOut<float> {
  typedef float4 type;
  static __device__ __forceinline__ float4 pack(float2 a, float2 b) { return make_float4(a.x, a.y, b.x, b.y); }
};
template <> struct PartOut<_Float16> {
  typedef uint2 type;
  static __device__ __forceinline__ uint2 pack(float2 a, float2 b) {
    half2v ha, hb;
    ha.x = (_Float16)a.x; ha.y = (_Float16)a.y;
    hb.x = (_Float16)b.x; hb.y = (_Float16)b.y;
    return make_uint2(__builtin_bit_cast(uint32_t, ha), __builtin_bit_cast(uint32_t, hb));
  }
};

// Feature index of part p, component c (the order encode_parts writes).
__host__ __device__ constexpr int part_feature(int p, int c) {
  return c < 2 ? 2 * p + c : 2 * (15 - p) + (c - 2);
}


}  // namespace nes
