// frameprep.hip -- nearer-depth composite + RGB24 -> YUV420P (K5).
//
// Replaces RenderedFrame::convert_frame (include/base/video/
// rendered_frame.h:24-33) -> SwsContextManager (src/base/video/
// type_managers.cc:143-155): sws_getContext(W, H, RGB24, W, H, YUV420P,
// flags = 0) + one whole-frame sws_scale.  With flags 0 libswscale 5.5 picks
// its default bicubic scaler and, for an RGB source, the generic path:
//   luma   Y14 = (RY r + GY g + BY b + (16.5 << 15)) >> 9      (rgb24ToY_c)
//          Y15 = 2 * Y14 (1-tap unscaled hScale16To15, shift 13)
//          Y   = clip((Y15 + 64) >> 7)                          (yuv2plane1_8_c)
//   chroma U14 = (RU (r0+r1) + GU (g0+g1) + BU (b0+b1) + (256.03 << 15)) >> 10
//          on horizontally paired pixels (rgb24ToUV_half_c), U15 = 2 * U14,
//          then an 8-tap 2:1 bicubic vertical filter (initFilter, 12-bit
//          coefficients, clamp-to-edge borders) whose arithmetic depends on
//          the build: x86-64 uses yuv2yuvX_sse3 (pmulhw taps) for every output
//          row but the last two luma rows, which use yuv2planeX_8_c.
//
// Kernel layout: one workgroup per (256-pixel strip, 16 chroma rows) tile of
// one frame.  Phase 1 streams the tile's source rows (plus the filter's
// 6-row halo) once from HBM, composites A/B, writes luma and parks the 15-bit
// chroma intermediates in LDS; phase 2 runs the vertical chroma filter out of
// LDS.  Inputs are read once (+~19% halo), outputs written once.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>

#include "nes_internal.h"

namespace nes {

// ------------------------------------------------------------------ host
// utils.c initFilter() of libswscale 5.5 restated for the bicubic scaler
// (param B = 0, C = 0.6), followed by the x86 filter alignment (2 for a
// vertical filter), the border fix-up and the error-diffusion normalisation.
static int init_filter(int x_inc, int src_w, int dst_w, int filter_align,
                       int one, int src_pos, int dst_pos,
                       std::vector<int16_t> &coef, std::vector<int32_t> &fpos) {
  const int ratio = src_w / dst_w;
  int lg = 0;
  for (unsigned r = (unsigned)(ratio > 0 ? ratio : 1); r >>= 1;) ++lg;
  const int64_t fone = 1ll << (54 - (lg < 8 ? lg : 8));
  std::vector<int64_t> filt;
  std::vector<int32_t> p((size_t)dst_w + 3, 0);
  int size;
  if (std::llabs((long long)x_inc - 0x10000) < 10 && src_pos == dst_pos) {
    size = 1;
    filt.assign((size_t)dst_w, fone);
    for (int i = 0; i < dst_w; ++i) p[i] = i;
  } else {
    size = x_inc <= (1 << 16) ? 5 : 1 + (4 * src_w + dst_w - 1) / dst_w;
    size = std::min(size, src_w - 2);
    size = std::max(size, 1);
    filt.assign((size_t)dst_w * size, 0);
    int64_t xdst = ((dst_pos * (int64_t)x_inc) >> 7) - ((src_pos * 0x10000ll) >> 7);
    const int64_t B = 0, C = (int64_t)(0.6 * (1 << 24));
    for (int i = 0; i < dst_w; ++i) {
      int xx = (int)((xdst - (size - 2) * (1ll << 16)) / (1 << 17));
      p[i] = xx;
      for (int j = 0; j < size; ++j, ++xx) {
        int64_t d = std::llabs((int64_t)xx * (1 << 17) - xdst) << 13;
        if (x_inc > (1 << 16)) d = d * dst_w / src_w;
        int64_t c = 0;
        if (d < (1ll << 31)) {
          int64_t dd = (d * d) >> 30, ddd = (dd * d) >> 30;
          if (d < (1ll << 30))
            c = (12 * (1 << 24) - 9 * B - 6 * C) * ddd +
                (-18 * (1 << 24) + 12 * B + 6 * C) * dd +
                (6 * (1 << 24) - 2 * B) * (1 << 30);
          else
            c = (-B - 6 * C) * ddd + (6 * B + 30 * C) * dd +
                (-12 * B - 48 * C) * d + (8 * B + 24 * C) * (1 << 30);
        }
        c /= (1ll << 54) / fone;
        filt[(size_t)i * size + j] = c;
      }
      xdst += 2 * (int64_t)x_inc;
    }
  }
  // Reduce: drop negligible leading taps (shifting the window right) and
  // measure the widest remaining support.
  const double cutoff = 0.002 * (double)fone;
  int min_size = 0;
  for (int i = dst_w - 1; i >= 0; --i) {
    int64_t *f = &filt[(size_t)i * size];
    int keep = size;
    int64_t acc = 0;
    for (int j = 0; j < size; ++j) {
      acc += std::llabs(f[0]);
      if ((double)acc > cutoff) break;
      if (i < dst_w - 1 && p[i] >= p[i + 1]) break;
      for (int k = 1; k < size; ++k) f[k - 1] = f[k];
      f[size - 1] = 0;
      p[i]++;
    }
    acc = 0;
    for (int j = size - 1; j > 0; --j) {
      acc += std::llabs(f[j]);
      if ((double)acc > cutoff) break;
      keep--;
    }
    min_size = std::max(min_size, keep);
  }
  if (min_size == 1 && filter_align == 2) filter_align = 1;
  const int out = (min_size + filter_align - 1) & ~(filter_align - 1);
  std::vector<int64_t> fr((size_t)dst_w * out, 0);
  for (int i = 0; i < dst_w; ++i)
    for (int j = 0; j < out && j < size; ++j) fr[(size_t)i * out + j] = filt[(size_t)i * size + j];
  // Borders: fold taps that fall outside the source onto the edge sample.
  for (int i = 0; i < dst_w; ++i) {
    int64_t *f = &fr[(size_t)i * out];
    if (p[i] < 0) {
      for (int j = 1; j < out; ++j) {
        int left = std::max(j + p[i], 0);
        f[left] += f[j];
        f[j] = 0;
      }
      p[i] = 0;
    }
    if (p[i] + out > src_w) {
      int shift = p[i] + std::min(out - src_w, 0);
      int64_t acc = 0;
      for (int j = out - 1; j >= 0; --j)
        if (p[i] + j >= src_w) {
          acc += f[j];
          f[j] = 0;
        }
      for (int j = out - 1; j >= 0; --j) f[j] = j < shift ? 0 : f[j - shift];
      p[i] -= shift;
      f[src_w - 1 - p[i]] += acc;
    }
  }
  coef.assign((size_t)dst_w * out, 0);
  fpos.assign((size_t)dst_w, 0);
  for (int i = 0; i < dst_w; ++i) {
    int64_t sum = 0, err = 0;
    for (int j = 0; j < out; ++j) sum += fr[(size_t)i * out + j];
    sum = (sum + one / 2) / one;
    if (!sum) sum = 1;
    for (int j = 0; j < out; ++j) {
      int64_t v = fr[(size_t)i * out + j] + err;
      int64_t iv = (v >= 0 ? v + (sum >> 1) : v - (sum >> 1)) / sum;
      coef[(size_t)i * out + j] = (int16_t)iv;
      err = v - iv * sum;
    }
    fpos[i] = p[i];
  }
  return out;
}

VFilter build_vchroma_filter(int h) {
  // Source chroma is full height (RGB input), destination half height.  Both
  // chroma positions default to -513, i.e. centred: get_local_pos() = 128.
  const int dst_h = (h + 1) >> 1;
  const int inc = (int)((((int64_t)h << 16) + (dst_h >> 1)) / dst_h);
  VFilter f;
  f.taps = init_filter(inc, h, dst_h, 2, 1 << 12, 128, 128, f.coef, f.pos);
  return f;
}

namespace {
std::mutex g_filter_mu;
std::map<int, DevVFilter> g_filters;
}  // namespace

static bool filter_fits_tiles(const VFilter &f, int h);

int get_dev_vfilter(int h, DevVFilter *out) {
  std::lock_guard<std::mutex> lk(g_filter_mu);
  auto it = g_filters.find(h);
  if (it != g_filters.end()) {
    *out = it->second;
    return NES_OK;
  }
  VFilter f = build_vchroma_filter(h);
  if (f.taps < 1 || f.taps > 16 || !filter_fits_tiles(f, h)) {
    set_last_error("unsupported vertical filter size");
    return NES_ERANGE;
  }
  DevVFilter d;
  d.taps = f.taps;
  NES_HIP_TRY(hipMalloc(&d.coef, f.coef.size() * sizeof(int16_t)));
  NES_HIP_TRY(hipMalloc(&d.pos, f.pos.size() * sizeof(int32_t)));
  NES_HIP_TRY(hipMemcpy(d.coef, f.coef.data(), f.coef.size() * sizeof(int16_t),
                        hipMemcpyHostToDevice));
  NES_HIP_TRY(hipMemcpy(d.pos, f.pos.data(), f.pos.size() * sizeof(int32_t),
                        hipMemcpyHostToDevice));
  g_filters[h] = d;
  *out = d;
  return NES_OK;
}

// ---------------------------------------------------------------- device
constexpr int kFpQuads = 64;               // 4-pixel quads per tile row
// 16 chroma rows per tile: a tile reads its 32 source rows plus the filter's
// 6-row halo (the most any height needs; filter_fits_tiles checks it), and
// the 19.5 KB of LDS let 8 workgroups (32 waves) share a CU.  Measured on
// 256 720p frames: 32 rows 0.83 ms, 16 rows 0.74 ms, 8 rows 0.75 ms -- the
// occupancy gained outweighs the extra halo reads.
constexpr int kFpCRows = 16;
constexpr int kFpRows = 2 * kFpCRows + 6;   // LDS source rows
constexpr int kFpThreads = 256;

// Source rows one tile of kFpCRows chroma rows reads (the kernel's lo..hi):
// every tile of a filter must fit the kFpRows rows of the LDS tile.
static bool filter_fits_tiles(const VFilter &f, int h) {
  const int ch = h >> 1;
  for (int k0 = 0; k0 < ch; k0 += kFpCRows) {
    const int k1 = std::min(k0 + kFpCRows, ch);
    const int lo = std::min(f.pos[k0], 2 * k0);
    const int hi = std::min(std::max(f.pos[k1 - 1] + f.taps - 1, 2 * k1 - 1), h - 1);
    if (lo < 0 || hi - lo + 1 > kFpRows) return false;
  }
  return true;
}

struct FrameprepArgs {
  const uint8_t *rgb_a;
  const float *depth_a;
  const uint8_t *rgb_b;
  const float *depth_b;
  uint8_t *y;
  uint8_t *u;
  uint8_t *v;
  const int16_t *coef;
  const int32_t *pos;
  size_t rgb_fb, depth_fb, y_fb, u_fb, v_fb;  // per-frame byte offsets
  int w, h, y_stride, u_stride, v_stride, taps, mode;
};

__device__ __forceinline__ int y_of(int r, int g, int b) {
  int y14 = (8414 * r + 16519 * g + 3208 * b + 524544) >> 9;
  int y15 = min(2 * y14, 32767);
  return min(max((y15 + 64) >> 7, 0), 255);
}
__device__ __forceinline__ int u15_of(int rs, int gs, int bs) {
  return min(2 * ((-4865 * rs - 9528 * gs + 14392 * bs + 8389120) >> 10), 32767);
}
__device__ __forceinline__ int v15_of(int rs, int gs, int bs) {
  return min(2 * ((14392 * rs - 12061 * gs - 2332 * bs + 8389120) >> 10), 32767);
}
__device__ __forceinline__ bool take_b(float da, float db) {
  float ka = isnan(da) ? INFINITY : da;
  float kb = isnan(db) ? INFINITY : db;
  return !(ka <= kb);
}

template <bool kFast, bool kTwo>
__global__ __launch_bounds__(kFpThreads) void frameprep_kernel(FrameprepArgs a) {
  __shared__ int16_t su[kFpRows][2 * kFpQuads];
  __shared__ int16_t sv[kFpRows][2 * kFpQuads];

  const int f = blockIdx.z;
  const uint8_t *rgb_a = a.rgb_a + f * a.rgb_fb;
  const float *dep_a = a.depth_a ? (const float *)((const uint8_t *)a.depth_a + f * a.depth_fb) : nullptr;
  const uint8_t *rgb_b = kTwo ? a.rgb_b + f * a.rgb_fb : nullptr;
  const float *dep_b = kTwo ? (const float *)((const uint8_t *)a.depth_b + f * a.depth_fb) : nullptr;
  uint8_t *yp = a.y + f * a.y_fb;
  uint8_t *up = a.u + f * a.u_fb;
  uint8_t *vp = a.v + f * a.v_fb;

  const int ch = a.h >> 1, cw = a.w >> 1;
  const int k0 = blockIdx.y * kFpCRows;
  const int k1 = min(k0 + kFpCRows, ch);
  const int lo = min(a.pos[k0], 2 * k0);
  const int hi = min(max(a.pos[k1 - 1] + a.taps - 1, 2 * k1 - 1), a.h - 1);
  const int nrows = hi - lo + 1;
  const int q_base = blockIdx.x * kFpQuads;

  // Phase 1: composite, luma out, chroma intermediates into LDS.
  for (int it = threadIdx.x; it < nrows * kFpQuads; it += kFpThreads) {
    const int rr = it / kFpQuads, q = it % kFpQuads;
    const int r = lo + rr;
    const int x0 = 4 * (q_base + q);
    if (x0 >= a.w) continue;
    const int np = min(4, a.w - x0);  // 4, or 2 at a ragged right edge
    int px[4][3];
    const size_t rowp = (size_t)r * a.w + x0;
    if (kFast) {
      const uint32_t *ra = (const uint32_t *)(rgb_a + 3 * rowp);
      uint32_t wa[3] = {ra[0], ra[1], ra[2]};
      uint32_t wb[3] = {0, 0, 0};
      float4 da = make_float4(0, 0, 0, 0), db = da;
      if (kTwo) {
        const uint32_t *rb = (const uint32_t *)(rgb_b + 3 * rowp);
        wb[0] = rb[0]; wb[1] = rb[1]; wb[2] = rb[2];
        da = *(const float4 *)(dep_a + rowp);
        db = *(const float4 *)(dep_b + rowp);
      }
      const float dav[4] = {da.x, da.y, da.z, da.w}, dbv[4] = {db.x, db.y, db.z, db.w};
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        const bool useb = kTwo && take_b(dav[p], dbv[p]);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const int byte = 3 * p + c;
          const uint32_t word = useb ? wb[byte >> 2] : wa[byte >> 2];
          px[p][c] = (word >> (8 * (byte & 3))) & 0xff;
        }
      }
    } else {
      for (int p = 0; p < 4; ++p) {
        if (p >= np) { px[p][0] = px[p][1] = px[p][2] = 0; continue; }
        const uint8_t *src = rgb_a + 3 * (rowp + p);
        if (kTwo && take_b(dep_a[rowp + p], dep_b[rowp + p])) src = rgb_b + 3 * (rowp + p);
        px[p][0] = src[0]; px[p][1] = src[1]; px[p][2] = src[2];
      }
    }
    if (r >= 2 * k0 && r < 2 * k1) {
      uint8_t *yr = yp + (size_t)r * a.y_stride + x0;
      if (kFast) {
        uint32_t packed = 0;
#pragma unroll
        for (int p = 0; p < 4; ++p) packed |= (uint32_t)y_of(px[p][0], px[p][1], px[p][2]) << (8 * p);
        *(uint32_t *)yr = packed;
      } else {
        for (int p = 0; p < np; ++p) yr[p] = (uint8_t)y_of(px[p][0], px[p][1], px[p][2]);
      }
    }
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      if (2 * e >= np) break;
      const int rs = px[2 * e][0] + px[2 * e + 1][0];
      const int gs = px[2 * e][1] + px[2 * e + 1][1];
      const int bs = px[2 * e][2] + px[2 * e + 1][2];
      su[rr][2 * q + e] = (int16_t)u15_of(rs, gs, bs);
      sv[rr][2 * q + e] = (int16_t)v15_of(rs, gs, bs);
    }
  }
  __syncthreads();

  // Phase 2: vertical chroma filter out of LDS.
  const int c_base = 2 * q_base;
  for (int it = threadIdx.x; it < (k1 - k0) * 2 * kFpQuads; it += kFpThreads) {
    const int k = k0 + it / (2 * kFpQuads);
    const int cl = it % (2 * kFpQuads);
    const int c = c_base + cl;
    if (c >= cw) continue;
    const int16_t *cf = a.coef + (size_t)k * a.taps;
    const int p0 = a.pos[k] - lo;
    const bool x86 = a.mode == NES_SWS_X86 && 2 * k < a.h - 2;
    int ou, ov;
    if (x86) {
      int16_t acu = (int16_t)((64 + 8 * (a.taps - 1)) >> 4), acv = acu;
      for (int j = 0; j < a.taps; ++j) {
        const int row = min(p0 + j, nrows - 1);
        const int cj = cf[j];
        acu = (int16_t)(acu + (int16_t)((cj * (int)su[row][cl]) >> 16));
        acv = (int16_t)(acv + (int16_t)((cj * (int)sv[row][cl]) >> 16));
      }
      ou = acu >> 3;
      ov = acv >> 3;
    } else {
      int vu = 64 << 12, vv = 64 << 12;
      for (int j = 0; j < a.taps; ++j) {
        const int row = min(p0 + j, nrows - 1);
        vu += cf[j] * (int)su[row][cl];
        vv += cf[j] * (int)sv[row][cl];
      }
      ou = vu >> 19;
      ov = vv >> 19;
    }
    up[(size_t)k * a.u_stride + c] = (uint8_t)min(max(ou, 0), 255);
    vp[(size_t)k * a.v_stride + c] = (uint8_t)min(max(ov, 0), 255);
  }
}

static int frameprep_check(const uint8_t *rgb_a, const float *depth_a, const uint8_t *rgb_b,
                           const float *depth_b, int w, int h, int n_frames, const uint8_t *y,
                           int y_stride, const uint8_t *u, int u_stride, const uint8_t *v,
                           int v_stride, int sws_mode) {
  if (w < 2 || h < 2 || (w & 1) || (h & 1) || n_frames < 1 || !rgb_a || !y || !u || !v) {
    set_last_error("frame prep needs even w, h >= 2 and non-null buffers");
    return NES_EINVAL;
  }
  if ((rgb_b == nullptr) != (depth_b == nullptr) || (rgb_b && !depth_a)) {
    set_last_error("composite needs both depths");
    return NES_EINVAL;
  }
  if (y_stride < w || u_stride < w / 2 || v_stride < w / 2) return NES_EINVAL;
  if (sws_mode != NES_SWS_X86 && sws_mode != NES_SWS_C) return NES_EINVAL;
  if (n_frames > 65535) return NES_ERANGE;
  return NES_OK;
}

// The frame-prep launch with the vertical chroma filter `flt` (device table).
static int launch_frameprep_filter(const uint8_t *rgb_a, const float *depth_a,
                                   const uint8_t *rgb_b, const float *depth_b, int w, int h,
                                   int n_frames, size_t rgb_fb, size_t depth_fb, uint8_t *y,
                                   int y_stride, size_t y_fb, uint8_t *u, int u_stride,
                                   size_t u_fb, uint8_t *v, int v_stride, size_t v_fb,
                                   int sws_mode, const DevVFilter &flt, hipStream_t stream);

int launch_frameprep(const uint8_t *rgb_a, const float *depth_a,
                     const uint8_t *rgb_b, const float *depth_b, int w, int h,
                     int n_frames, size_t rgb_fb, size_t depth_fb, uint8_t *y,
                     int y_stride, size_t y_fb, uint8_t *u, int u_stride,
                     size_t u_fb, uint8_t *v, int v_stride, size_t v_fb,
                     int sws_mode, hipStream_t stream) {
  NES_TRY(frameprep_check(rgb_a, depth_a, rgb_b, depth_b, w, h, n_frames, y, y_stride, u,
                          u_stride, v, v_stride, sws_mode));
  DevVFilter flt;
  NES_TRY(get_dev_vfilter(h, &flt));
  return launch_frameprep_filter(rgb_a, depth_a, rgb_b, depth_b, w, h, n_frames, rgb_fb,
                                 depth_fb, y, y_stride, y_fb, u, u_stride, u_fb, v, v_stride,
                                 v_fb, sws_mode, flt, stream);
}

static int launch_frameprep_filter(const uint8_t *rgb_a, const float *depth_a,
                                   const uint8_t *rgb_b, const float *depth_b, int w, int h,
                                   int n_frames, size_t rgb_fb, size_t depth_fb, uint8_t *y,
                                   int y_stride, size_t y_fb, uint8_t *u, int u_stride,
                                   size_t u_fb, uint8_t *v, int v_stride, size_t v_fb,
                                   int sws_mode, const DevVFilter &flt, hipStream_t stream) {
  FrameprepArgs a;
  a.rgb_a = rgb_a; a.depth_a = depth_a; a.rgb_b = rgb_b; a.depth_b = depth_b;
  a.y = y; a.u = u; a.v = v; a.coef = flt.coef; a.pos = flt.pos;
  a.rgb_fb = rgb_fb; a.depth_fb = depth_fb; a.y_fb = y_fb; a.u_fb = u_fb; a.v_fb = v_fb;
  a.w = w; a.h = h; a.y_stride = y_stride; a.u_stride = u_stride; a.v_stride = v_stride;
  a.taps = flt.taps; a.mode = sws_mode;
  auto al = [](const void *p, size_t n) { return ((uintptr_t)p % n) == 0; };
  bool fast = (w % 4 == 0) && al(rgb_a, 4) && al(y, 4) && (y_stride % 4 == 0) &&
              (rgb_fb % 4 == 0) && (y_fb % 4 == 0);
  if (rgb_b) fast = fast && al(rgb_b, 4) && al(depth_a, 16) && al(depth_b, 16) && (depth_fb % 16 == 0);
  const int ch = h / 2;
  dim3 grid((unsigned)((w / 4 + kFpQuads - 1) / kFpQuads + ((w % 4) ? 1 : 0)),
            (unsigned)((ch + kFpCRows - 1) / kFpCRows), (unsigned)n_frames);
  // quads cover ceil(w/4) columns; the ragged quad (w % 4 == 2) is handled by
  // the per-pixel path
  grid.x = (unsigned)(((w + 3) / 4 + kFpQuads - 1) / kFpQuads);
  if (fast) {
    if (rgb_b) frameprep_kernel<true, true><<<grid, kFpThreads, 0, stream>>>(a);
    else frameprep_kernel<true, false><<<grid, kFpThreads, 0, stream>>>(a);
  } else {
    if (rgb_b) frameprep_kernel<false, true><<<grid, kFpThreads, 0, stream>>>(a);
    else frameprep_kernel<false, false><<<grid, kFpThreads, 0, stream>>>(a);
  }
  NES_HIP_TRY(hipGetLastError());
  return NES_OK;
}

// GRAY8 full range -> YUV420P limited range (the depth stream).  Luma:
// hScale8To15 (g << 7) then lumRangeFromJpeg_c; chroma planes of a gray
// source hold the constant 1 << 14 before the vertical filter.
// kDepth: the source is FP32 expected depth, quantised to the GRAY8 depth
// plane on the fly (depth_gray8, `scale` = 255 / far); that plane is also
// written to `gray_out` when it is not null.
template <bool kDepth>
__global__ void gray_kernel(const uint8_t *g, int g_stride, int w, int h,
                            uint8_t *y, int y_stride, uint8_t *u, int u_stride,
                            uint8_t *v, int v_stride, const int16_t *coef,
                            const int32_t *pos, int taps, int mode, const float *depth,
                            float scale, uint8_t *gray_out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (x < w) {
    int gv;
    if (kDepth) {
      gv = depth_gray8(depth[(size_t)r * w + x], scale);
      if (gray_out) gray_out[(size_t)r * w + x] = (uint8_t)gv;
    } else {
      gv = g[(size_t)r * g_stride + x];
    }
    int v15 = gv << 7;
    v15 = (v15 * 14071 + 33561947) >> 14;
    y[(size_t)r * y_stride + x] = (uint8_t)min(max((v15 + 64) >> 7, 0), 255);
  }
  if ((r & 1) == 0 && x < (w >> 1)) {
    const int k = r >> 1;
    const int16_t *cf = coef + (size_t)k * taps;
    int out;
    if (mode == NES_SWS_X86 && 2 * k < h - 2) {
      int16_t acc = (int16_t)((64 + 8 * (taps - 1)) >> 4);
      for (int j = 0; j < taps; ++j) acc = (int16_t)(acc + (int16_t)(((int)cf[j] << 14) >> 16));
      out = acc >> 3;
    } else {
      int val = 64 << 12;
      for (int j = 0; j < taps; ++j) val += (int)cf[j] << 14;
      out = val >> 19;
    }
    out = min(max(out, 0), 255);
    u[(size_t)k * u_stride + x] = (uint8_t)out;
    v[(size_t)k * v_stride + x] = (uint8_t)out;
  }
  (void)pos;
}

int launch_depth_yuv(const float *depth, int w, int h, float scale, uint8_t *gray, uint8_t *y,
                     int y_stride, uint8_t *u, int u_stride, uint8_t *v, int v_stride,
                     int sws_mode, hipStream_t stream) {
  if (!depth || !y || !u || !v || w < 2 || h < 2 || (w & 1) || (h & 1)) return NES_EINVAL;
  if (y_stride < w || u_stride < w / 2 || v_stride < w / 2) return NES_EINVAL;
  if (sws_mode != NES_SWS_X86 && sws_mode != NES_SWS_C) return NES_EINVAL;
  DevVFilter flt;
  NES_TRY(get_dev_vfilter(h, &flt));
  dim3 grid((unsigned)((w + 255) / 256), (unsigned)h);
  gray_kernel<true><<<grid, 256, 0, stream>>>(nullptr, 0, w, h, y, y_stride, u, u_stride, v,
                                              v_stride, flt.coef, flt.pos, flt.taps, sws_mode,
                                              depth, scale, gray);
  NES_HIP_TRY(hipGetLastError());
  return NES_OK;
}

namespace {
std::mutex g_host_mu;
DevBuf g_in, g_out;
}  // namespace

// ------------------------------------------------ allocation-free variant
// nes_frameprep_device_static keeps its one scratch buffer -- the vertical
// chroma filter of the frame height (coefficients, then row positions) -- in
// this fixed device array instead of the per-height device cache, so a call
// allocates no device memory.  Each GPU holds its own copy until the process
// ends; the host refers to it only through the address hipGetSymbolAddress
// returns for the current device.
constexpr int kStaticMaxHeight = 8192;
constexpr int kStaticMaxTaps = 16;
constexpr size_t kStaticCoefBytes = (size_t)(kStaticMaxHeight / 2) * kStaticMaxTaps * sizeof(int16_t);
constexpr size_t kStaticScratchBytes = kStaticCoefBytes + (size_t)(kStaticMaxHeight / 2) * sizeof(int32_t);
static __device__ __align__(256) uint8_t scratch[kStaticScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_static_mu;
void *g_scratch_dev[kMaxDevices] = {};        // device address of `scratch` per device
std::map<int, VFilter> g_static_filters;       // host filter tables by height
}  // namespace

// The current device's address of `scratch`, looked up on the first call on
// each device with that device current.
static int static_scratch(void **out) {
  int dev = 0;
  NES_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return NES_ENODEV;
  std::lock_guard<std::mutex> lk(g_static_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_last_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return NES_EIO;
    }
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return NES_OK;
}

// Host filter table of height h (built once per height; host memory only).
static int static_filter(int h, const VFilter **out) {
  std::lock_guard<std::mutex> lk(g_static_mu);
  auto it = g_static_filters.find(h);
  if (it == g_static_filters.end()) {
    VFilter f = build_vchroma_filter(h);
    if (f.taps < 1 || f.taps > kStaticMaxTaps || !filter_fits_tiles(f, h)) {
      set_last_error("unsupported vertical filter size");
      return NES_ERANGE;
    }
    it = g_static_filters.emplace(h, std::move(f)).first;
  }
  *out = &it->second;
  return NES_OK;
}

}  // namespace nes

using namespace nes;

extern "C" {

int nes_frameprep(const uint8_t *rgb_a, const float *depth_a,
                  const uint8_t *rgb_b, const float *depth_b, int w, int h,
                  int rgb_stride, uint8_t *y, int y_stride, uint8_t *u,
                  int u_stride, uint8_t *v, int v_stride, int sws_mode) {
  if (!rgb_a || !y || !u || !v || w < 2 || h < 2 || (w & 1) || (h & 1)) return NES_EINVAL;
  if (rgb_stride < 3 * w || y_stride < w || u_stride < w / 2 || v_stride < w / 2) return NES_EINVAL;
  if ((rgb_b == nullptr) != (depth_b == nullptr) || (rgb_b && !depth_a)) return NES_EINVAL;
  NES_TRY(ensure_device());
  std::lock_guard<std::mutex> lk(g_host_mu);
  const size_t px = (size_t)w * h;
  const size_t rgb_bytes = 3 * px, dep_bytes = 4 * px;
  const size_t in_bytes = rgb_b ? 2 * rgb_bytes + 2 * dep_bytes + 64 : rgb_bytes + 64;
  const size_t cw = (size_t)w / 2, chh = (size_t)h / 2;
  const size_t out_bytes = px + 2 * cw * chh + 64;
  NES_TRY(g_in.reserve(in_bytes));
  NES_TRY(g_out.reserve(out_bytes));
  hipStream_t s = library_stream();
  uint8_t *d_rgb_a = g_in.as<uint8_t>();
  // depth buffers are placed 16-byte aligned after the colour planes
  size_t off_rgb_b = (rgb_bytes + 15) & ~size_t(15);
  size_t off_da = (off_rgb_b + rgb_bytes + 15) & ~size_t(15);
  size_t off_db = off_da + dep_bytes;
  uint8_t *d_rgb_b = rgb_b ? d_rgb_a + off_rgb_b : nullptr;
  float *d_da = rgb_b ? (float *)(d_rgb_a + off_da) : nullptr;
  float *d_db = rgb_b ? (float *)(d_rgb_a + off_db) : nullptr;
  NES_HIP_TRY(hipMemcpy2DAsync(d_rgb_a, 3 * w, rgb_a, rgb_stride, 3 * w, h,
                               hipMemcpyHostToDevice, s));
  if (rgb_b) {
    NES_HIP_TRY(hipMemcpy2DAsync(d_rgb_b, 3 * w, rgb_b, rgb_stride, 3 * w, h,
                                 hipMemcpyHostToDevice, s));
    NES_HIP_TRY(hipMemcpyAsync(d_da, depth_a, dep_bytes, hipMemcpyHostToDevice, s));
    NES_HIP_TRY(hipMemcpyAsync(d_db, depth_b, dep_bytes, hipMemcpyHostToDevice, s));
  }
  uint8_t *d_y = g_out.as<uint8_t>();
  uint8_t *d_u = d_y + px;
  uint8_t *d_v = d_u + cw * chh;
  NES_TRY(launch_frameprep(d_rgb_a, d_da, d_rgb_b, d_db, w, h, 1, 0, 0, d_y, w, 0,
                           d_u, (int)cw, 0, d_v, (int)cw, 0, sws_mode, s));
  NES_HIP_TRY(hipMemcpy2DAsync(y, y_stride, d_y, w, w, h, hipMemcpyDeviceToHost, s));
  NES_HIP_TRY(hipMemcpy2DAsync(u, u_stride, d_u, cw, cw, chh, hipMemcpyDeviceToHost, s));
  NES_HIP_TRY(hipMemcpy2DAsync(v, v_stride, d_v, cw, cw, chh, hipMemcpyDeviceToHost, s));
  NES_HIP_TRY(hipStreamSynchronize(s));
  return NES_OK;
}

int nes_frameprep_device(const uint8_t *rgb_a, const float *depth_a,
                         const uint8_t *rgb_b, const float *depth_b, int w,
                         int h, int n_frames, uint8_t *y, int y_stride,
                         uint8_t *u, int u_stride, uint8_t *v, int v_stride,
                         int sws_mode, float *elapsed_ms) {
  NES_TRY(ensure_device());
  hipStream_t s = library_stream();
  const size_t px = (size_t)w * h;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (elapsed_ms) {
    NES_HIP_TRY(hipEventCreate(&e0));
    NES_HIP_TRY(hipEventCreate(&e1));
    NES_HIP_TRY(hipEventRecord(e0, s));
  }
  int rc = launch_frameprep(rgb_a, depth_a, rgb_b, depth_b, w, h, n_frames, 3 * px,
                            4 * px, y, y_stride, (size_t)y_stride * h, u, u_stride,
                            (size_t)u_stride * (h / 2), v, v_stride,
                            (size_t)v_stride * (h / 2), sws_mode, s);
  if (elapsed_ms) {
    hipEventRecord(e1, s);
    hipEventSynchronize(e1);
    hipEventElapsedTime(elapsed_ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
  }
  if (rc != NES_OK) return rc;
  NES_HIP_TRY(hipStreamSynchronize(s));
  return NES_OK;
}

int nes_frameprep_device_static(const uint8_t *rgb_a, const float *depth_a,
                                const uint8_t *rgb_b, const float *depth_b, int w, int h,
                                int n_frames, uint8_t *y, int y_stride, uint8_t *u,
                                int u_stride, uint8_t *v, int v_stride, int sws_mode,
                                void *stream) {
  NES_TRY(ensure_device());
  NES_TRY(frameprep_check(rgb_a, depth_a, rgb_b, depth_b, w, h, n_frames, y, y_stride, u,
                          u_stride, v, v_stride, sws_mode));
  if (h > kStaticMaxHeight) {
    set_last_error("nes_frameprep_device_static: frame height above 8192 (its fixed scratch "
                   "holds the chroma filter of at most 4096 rows)");
    return NES_ERANGE;
  }
  void *sp = nullptr;
  NES_TRY(static_scratch(&sp));
  const VFilter *f = nullptr;
  NES_TRY(static_filter(h, &f));
  hipStream_t s = static_cast<hipStream_t>(stream);
  DevVFilter flt;
  flt.taps = f->taps;
  flt.coef = static_cast<int16_t *>(sp);
  flt.pos = reinterpret_cast<int32_t *>(static_cast<uint8_t *>(sp) + kStaticCoefBytes);
  // the table of this call's height replaces whatever the last call left
  NES_HIP_TRY(hipMemcpyAsync(flt.coef, f->coef.data(), f->coef.size() * sizeof(int16_t),
                             hipMemcpyHostToDevice, s));
  NES_HIP_TRY(hipMemcpyAsync(flt.pos, f->pos.data(), f->pos.size() * sizeof(int32_t),
                             hipMemcpyHostToDevice, s));
  const size_t px = (size_t)w * h;
  NES_TRY(launch_frameprep_filter(rgb_a, depth_a, rgb_b, depth_b, w, h, n_frames, 3 * px, 4 * px,
                                  y, y_stride, (size_t)y_stride * h, u, u_stride,
                                  (size_t)u_stride * (h / 2), v, v_stride,
                                  (size_t)v_stride * (h / 2), sws_mode, flt, s));
  if (!stream) NES_HIP_TRY(hipStreamSynchronize(s));
  return NES_OK;
}

int nes_gray8_to_yuv420p(const uint8_t *gray, int w, int h, int g_stride,
                         uint8_t *y, int y_stride, uint8_t *u, int u_stride,
                         uint8_t *v, int v_stride, int sws_mode) {
  if (!gray || !y || !u || !v || w < 2 || h < 2 || (w & 1) || (h & 1)) return NES_EINVAL;
  if (g_stride < w || y_stride < w || u_stride < w / 2 || v_stride < w / 2) return NES_EINVAL;
  if (sws_mode != NES_SWS_X86 && sws_mode != NES_SWS_C) return NES_EINVAL;
  NES_TRY(ensure_device());
  std::lock_guard<std::mutex> lk(g_host_mu);
  DevVFilter flt;
  NES_TRY(get_dev_vfilter(h, &flt));
  const size_t px = (size_t)w * h, cw = (size_t)w / 2, chh = (size_t)h / 2;
  NES_TRY(g_in.reserve(px));
  NES_TRY(g_out.reserve(px + 2 * cw * chh));
  hipStream_t s = library_stream();
  uint8_t *d_g = g_in.as<uint8_t>();
  uint8_t *d_y = g_out.as<uint8_t>(), *d_u = d_y + px, *d_v = d_u + cw * chh;
  NES_HIP_TRY(hipMemcpy2DAsync(d_g, w, gray, g_stride, w, h, hipMemcpyHostToDevice, s));
  dim3 grid((unsigned)((w + 255) / 256), (unsigned)h);
  gray_kernel<false><<<grid, 256, 0, s>>>(d_g, w, w, h, d_y, w, d_u, (int)cw, d_v, (int)cw,
                                          flt.coef, flt.pos, flt.taps, sws_mode, nullptr, 0.0f,
                                          nullptr);
  NES_HIP_TRY(hipGetLastError());
  NES_HIP_TRY(hipMemcpy2DAsync(y, y_stride, d_y, w, w, h, hipMemcpyDeviceToHost, s));
  NES_HIP_TRY(hipMemcpy2DAsync(u, u_stride, d_u, cw, cw, chh, hipMemcpyDeviceToHost, s));
  NES_HIP_TRY(hipMemcpy2DAsync(v, v_stride, d_v, cw, cw, chh, hipMemcpyDeviceToHost, s));
  NES_HIP_TRY(hipStreamSynchronize(s));
  return NES_OK;
}

// Host-side view of the vertical chroma filter (for inspection and tests).
int nes_sws_vchroma_filter(int h, int16_t *coef, int32_t *pos, int max_taps,
                           int *taps) {
  if (h < 2 || !coef || !pos || !taps) return NES_EINVAL;
  VFilter f = build_vchroma_filter(h);
  if (f.taps > max_taps) return NES_ERANGE;
  std::memcpy(coef, f.coef.data(), f.coef.size() * sizeof(int16_t));
  std::memcpy(pos, f.pos.data(), f.pos.size() * sizeof(int32_t));
  *taps = f.taps;
  return NES_OK;
}

}  // extern "C"
