// nes_internal.h -- shared definitions of libnes_render (host + device).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/nes_render.h"

namespace nes {

// ---------------------------------------------------------------- status
#define NES_HIP_TRY(expr)                                                    \
  do {                                                                       \
    hipError_t _e = (expr);                                                  \
    if (_e != hipSuccess) {                                                  \
      nes::set_last_error(std::string(#expr) + ": " + hipGetErrorString(_e)); \
      return _e == hipErrorOutOfMemory ? NES_ENOMEM : NES_EIO;               \
    }                                                                        \
  } while (0)

#define NES_TRY(expr)          \
  do {                         \
    int _rc = (expr);          \
    if (_rc != NES_OK) return _rc; \
  } while (0)

void set_last_error(const std::string &msg);
int ensure_device();  // selects device 0 once, checks it is gfx950

// ------------------------------------------------------------- constants
constexpr int kLevels = 16;
constexpr int kLog2T = 19;
constexpr uint32_t kTableEntries = 6098925u;
constexpr int kWeights = 9344;
constexpr int kOccRes = 128;
constexpr int kMaxSteps = 128;
// (float)(sqrt(3) / 128): the march step; every kernel and the host agree on it.
constexpr float kDelta = 0.0135316469f;
constexpr float kTStop = 1e-4f;
constexpr uint32_t kPrime1 = 2654435761u;
constexpr uint32_t kPrime2 = 805459861u;

// Level resolutions N_l = floor(16 * 2^(7l/15)).
constexpr int kLevelRes[kLevels] = {16,  22,  30,  42,  58,   80,   111,  153,
                                    212, 294, 406, 561, 776, 1072, 1482, 2048};
// Dense levels (N+1)^3 <= 2^19 are levels 0..4; entry offsets of each level.
constexpr int kDenseLevels = 5;
constexpr uint32_t kLevelOffset[kLevels + 1] = {
    0u,       4913u,    17080u,   46871u,   126378u,  331757u,
    856045u,  1380333u, 1904621u, 2428909u, 2953197u, 3477485u,
    4001773u, 4526061u, 5050349u, 5574637u, 6098925u};

// Device layout of the table: the same levels with each one starting on a
// 32-entry boundary (128 B of FP16 entries, 256 B of FP32).  Packed level
// after level, every hashed level starts 5 entries past a 64 B line, so the
// pair of x corners (hash indices h and h ^ (x ^ (x + 1))) straddles two
// lines 22 % of the time instead of 6 % (FP32 scatter: 5.5 requests per
// sample and level instead of 4.5).  Host buffers keep the packed layout;
// every copy between the two goes level by level (table_copy).
constexpr uint32_t align32(uint32_t x) { return (x + 31u) & ~31u; }
constexpr uint32_t dev_level_offset(int l) {
  return l == 0 ? 0u
                : align32(dev_level_offset(l - 1) + (kLevelOffset[l] - kLevelOffset[l - 1]));
}
constexpr uint32_t kDevLevelOffset[kLevels + 1] = {
    dev_level_offset(0),  dev_level_offset(1),  dev_level_offset(2),  dev_level_offset(3),
    dev_level_offset(4),  dev_level_offset(5),  dev_level_offset(6),  dev_level_offset(7),
    dev_level_offset(8),  dev_level_offset(9),  dev_level_offset(10), dev_level_offset(11),
    dev_level_offset(12), dev_level_offset(13), dev_level_offset(14), dev_level_offset(15),
    dev_level_offset(16)};
constexpr uint32_t kDevTableEntries = kDevLevelOffset[kLevels];
static_assert(kDevTableEntries == 6099008u, "device table layout");

// Copy a table of 2 features per entry (feat_bytes each: 4 FP32, 2 FP16)
// between the packed layout and the device layout, level by level.
// dst_dev / src_dev say which side is in the device layout.  With a null
// stream every level is a synchronous hipMemcpy (host memory may be
// released on return); otherwise the copies are queued on `stream` and the
// caller keeps host memory alive until the stream has passed them.  Host to
// host copies use memcpy.
int table_copy(void *dst, bool dst_dev, const void *src, bool src_dev, size_t feat_bytes,
               hipMemcpyKind kind, hipStream_t stream);

// MLP layout (row-major [out][in] per layer).
constexpr int kW1 = 0, kW2 = 2048, kW3 = 3072, kW4 = 5056, kW5 = 9152;

// ------------------------------------------------------------------- RNG
inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t rand_u64(uint64_t seed, uint64_t stream, uint64_t i) {
  return mix64(mix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) + i);
}
inline double rand_uniform(uint64_t seed, uint64_t stream, uint64_t i) {
  return (double)(rand_u64(seed, stream, i) >> 11) * 0x1.0p-53;
}

// ------------------------------------------------------------ GRAY8 depth
// The depth plane of a RenderedFrame (proto/nes.proto:24) is GRAY8:
// g = clamp(floor(fma(d, 255 / far, 0.5)), 0, 255) in FP32 (NaN -> 0).
inline float depth_scale(float far) {
  return 255.0f / (far > 0.0f ? far : NES_DEPTH_FAR_DEFAULT);
}
__host__ __device__ inline uint8_t depth_gray8(float d, float scale) {
  const float q = floorf(fmaf(d, scale, 0.5f));
  return (uint8_t)fminf(fmaxf(q, 0.0f), 255.0f);
}

// ------------------------------------------------------------ frame prep
struct VFilter {
  int taps = 0;                 // filter size
  std::vector<int16_t> coef;    // [chroma_rows][taps]
  std::vector<int32_t> pos;     // [chroma_rows]
};
// libswscale initFilter() for the vertical chroma filter of an
// RGB24 WxH -> YUV420P WxH context (bicubic default scaler).
VFilter build_vchroma_filter(int h);

// Device copy of a VFilter, cached per frame height.
struct DevVFilter {
  int taps = 0;
  int16_t *coef = nullptr;
  int32_t *pos = nullptr;
};
int get_dev_vfilter(int h, DevVFilter *out);

hipStream_t library_stream();

int launch_frameprep(const uint8_t *rgb_a, const float *depth_a,
                     const uint8_t *rgb_b, const float *depth_b, int w, int h,
                     int n_frames, size_t rgb_frame_bytes,
                     size_t depth_frame_bytes, uint8_t *y, int y_stride,
                     size_t y_frame_bytes, uint8_t *u, int u_stride,
                     size_t u_frame_bytes, uint8_t *v, int v_stride,
                     size_t v_frame_bytes, int sws_mode, hipStream_t stream);

// The depth stream of one view straight from FP32 depth: quantise to GRAY8
// (scale = 255 / far; the plane goes to `gray` when not null) and convert
// GRAY8 -> YUV420P as nes_gray8_to_yuv420p does.  `depth` is w*h, tight.
int launch_depth_yuv(const float *depth, int w, int h, float scale, uint8_t *gray, uint8_t *y,
                     int y_stride, uint8_t *u, int u_stride, uint8_t *v, int v_stride,
                     int sws_mode, hipStream_t stream);

// Grow-only device scratch buffer.
struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  int reserve(size_t need);
  void release();
  template <class T>
  T *as() const { return static_cast<T *>(ptr); }
};

}  // namespace nes
