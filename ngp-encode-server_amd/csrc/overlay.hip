// overlay.hip -- text overlay of the frame-prep step, on the device.
//
// Replaces RenderTextContext (src/base/video/render_text.cc:10-111) as used
// by process_frame_thread (src/encode.cpp:57-97).  The reference rasterises
// each character with FreeType under a global mutex and writes white RGB24
// pixels wherever the glyph bitmap is non-zero, one string at a time, on the
// CPU, for every frame.  Here the caller's FreeType rasterises the character
// set once (nes_glyph table + bitmaps, uploaded by nes_font_create); per frame
// the host only walks the pens (a few dozen integer additions), and one
// launch stamps every glyph of every string of every view in place on the
// device RGB24 frame, right before the YUV conversion reads it.
//
// Layout rules of render_string_to_frame (render_text.cc:48-109):
//   pen start  LEFT_TOP (50, 50); LEFT_BOTTOM (50, H - 50);
//              RIGHT_TOP (W - 250, 50); RIGHT_BOTTOM (W - 250, H - 50);
//              CENTER (W/2 - 300, H/2 - 100)  [x_box 300, y_box 100, margin 50]
//   '\n'       pen x back to the start, pen y + 20, nothing drawn
//   glyph      bitmap row q, column p lands on (pen_x + left + p,
//              pen_y - top + q), clipped to the frame; pen x += advance.
#include <cstdio>
#include <cstring>
#include <ctime>
#include <chrono>
#include <memory>

#include "render.h"

struct nes_font {
  nes_glyph glyphs[NES_FONT_GLYPHS];
  nes::DevBuf d_glyphs, d_bitmaps, d_place;
  std::mutex mu;
};

namespace nes {

static void pen_start(int position, int w, int h, int *x, int *y) {
  constexpr int kBoxX = 300, kBoxY = 100, kMargin = 50;
  switch (position) {
    case NES_TEXT_LEFT_BOTTOM: *x = kMargin; *y = h - kBoxY + kMargin; break;
    case NES_TEXT_RIGHT_TOP: *x = w - kBoxX + kMargin; *y = kMargin; break;
    case NES_TEXT_RIGHT_BOTTOM: *x = w - kBoxX + kMargin; *y = h - kBoxY + kMargin; break;
    case NES_TEXT_CENTER: *x = w / 2 - kBoxX; *y = h / 2 - kBoxY; break;
    default: *x = kMargin; *y = kMargin; break;  // LEFT_TOP and unknown values
  }
}

// Walk the pen over one string and append a placement per drawn glyph.
void overlay_layout(const nes_font *font, int position, const char *text, int w,
                    int h, int stride, uint64_t base, std::vector<GlyphPlacement> &out) {
  int pen_x, pen_y;
  pen_start(position, w, h, &pen_x, &pen_y);
  const int x_start = pen_x;
  for (const char *c = text; *c; ++c) {
    if (*c == '\n') {
      pen_x = x_start;
      pen_y += 20;
      continue;
    }
    const unsigned gi = (unsigned char)*c;
    const nes_glyph &g = font->glyphs[gi];
    if (g.width && g.rows) {
      const int x0 = pen_x + g.left, y0 = pen_y - g.top;
      // skip glyphs wholly outside the frame
      if (x0 < w && y0 < h && x0 + (int)g.width > 0 && y0 + (int)g.rows > 0)
        out.push_back({base, x0, y0, w, h, stride, gi});
    }
    pen_x += g.advance;
  }
}

// One workgroup per placed glyph; threads sweep its bitmap.  Overlapping
// glyphs write the same value (255), so their order does not matter.
__global__ __launch_bounds__(256) void overlay_stamp(const GlyphPlacement *place,
                                                     const nes_glyph *glyphs,
                                                     const uint8_t *bitmaps, uint8_t *rgb) {
  const GlyphPlacement pl = place[blockIdx.x];
  const nes_glyph g = glyphs[pl.glyph];
  const int n = (int)(g.width * g.rows);
  const uint8_t *bm = bitmaps + g.offset;
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    const int q = t / (int)g.width, p = t - q * (int)g.width;
    const int i = pl.x0 + p, j = pl.y0 + q;
    if (i < 0 || j < 0 || i >= pl.w || j >= pl.h) continue;
    if (bm[t]) {
      uint8_t *d = rgb + pl.base + (size_t)j * pl.stride + 3 * (size_t)i;
      d[0] = 255;
      d[1] = 255;
      d[2] = 255;
    }
  }
}

// Upload the placements through `staging` (device) and stamp them.  `place`
// must stay alive and `staging` unused by other work until the stream has
// passed this launch.
int overlay_launch(const nes_font *font, DevBuf &staging,
                   const std::vector<GlyphPlacement> &place, uint8_t *rgb, hipStream_t stream) {
  if (place.empty()) return NES_OK;
  if (place.size() > 0x7fffffffu) return NES_ERANGE;
  const size_t bytes = place.size() * sizeof(GlyphPlacement);
  NES_TRY(staging.reserve(bytes));
  NES_HIP_TRY(hipMemcpyAsync(staging.ptr, place.data(), bytes, hipMemcpyHostToDevice, stream));
  overlay_stamp<<<(unsigned)place.size(), 256, 0, stream>>>(
      staging.as<GlyphPlacement>(), font->d_glyphs.as<nes_glyph>(),
      font->d_bitmaps.as<uint8_t>(), rgb);
  NES_HIP_TRY(hipGetLastError());
  return NES_OK;
}

// std::fixed << std::showpos << std::setw(7) << std::setprecision(5)
// << std::setfill('0') of one float (src/encode.cpp:63-64): sign always
// shown, five decimals, left-filled with '0' to seven characters (only
// "+inf"-like strings are that short).
static void put_matrix_value(std::string &s, float v) {
  char tmp[64];
  std::snprintf(tmp, sizeof tmp, "%+.5f", (double)v);
  const size_t n = std::strlen(tmp);
  if (n < 7) s.append(7 - n, '0');
  s += tmp;
  s += ' ';
}

std::string overlay_matrix_text(const float c2w[12]) {
  std::string s;
  for (int i = 0; i < 12; ++i) {
    put_matrix_value(s, c2w[i]);
    if ((i + 1) % 4 == 0) s += '\n';
  }
  put_matrix_value(s, 0.f);
  put_matrix_value(s, 0.f);
  put_matrix_value(s, 0.f);
  put_matrix_value(s, 1.f);
  return s;
}

// timestamp() of src/encode.cpp:23-38: local "%T" then ".mmm".
std::string overlay_now_text() {
  using namespace std::chrono;
  const auto now = system_clock::now();
  const std::time_t t = system_clock::to_time_t(now);
  std::tm lt;
  localtime_r(&t, &lt);
  const long long ms = duration_cast<milliseconds>(now.time_since_epoch()).count() % 1000;
  char buf[64];
  std::snprintf(buf, sizeof buf, "%02d:%02d:%02d.%03lld", lt.tm_hour, lt.tm_min, lt.tm_sec, ms);
  return buf;
}

// The (position, string) pairs process_frame_thread draws for one frame.
void overlay_frame_texts(const nes_camera &cam, const char *timestamp, uint32_t flags,
                         std::vector<std::pair<int, std::string>> &out) {
  if (flags & NES_OVERLAY_MATRIX) out.emplace_back(NES_TEXT_CENTER, overlay_matrix_text(cam.c2w));
  if (flags & NES_OVERLAY_INDEX)
    out.emplace_back(NES_TEXT_LEFT_BOTTOM, "index=" + std::to_string(cam.index));
  if (flags & NES_OVERLAY_TIMESTAMP)
    out.emplace_back(NES_TEXT_LEFT_TOP, timestamp ? std::string(timestamp) : overlay_now_text());
  if (flags & NES_OVERLAY_DIRECTION)
    out.emplace_back(NES_TEXT_RIGHT_TOP,
                     cam.is_left ? "direction=left" : "direction=right");
}

void overlay_frame_layout(const nes_font *font, const nes_camera &cam, const char *timestamp,
                          uint32_t flags, int stride, uint64_t base,
                          std::vector<GlyphPlacement> &out) {
  std::vector<std::pair<int, std::string>> texts;
  overlay_frame_texts(cam, timestamp, flags, texts);
  for (auto &t : texts)
    overlay_layout(font, t.first, t.second.c_str(), (int)cam.width, (int)cam.height, stride,
                   base, out);
}

namespace {
std::mutex g_ovl_host_mu;
DevBuf g_ovl_in, g_ovl_out;
}  // namespace

}  // namespace nes

using namespace nes;

extern "C" {

int nes_font_create(const nes_glyph *glyphs, const uint8_t *bitmaps, size_t bitmap_bytes,
                    nes_font **out) {
  if (!glyphs || !out || (bitmap_bytes && !bitmaps)) return NES_EINVAL;
  for (int i = 0; i < NES_FONT_GLYPHS; ++i) {
    const nes_glyph &g = glyphs[i];
    const uint64_t n = (uint64_t)g.width * g.rows;
    if (g.width > 4096 || g.rows > 4096 || (uint64_t)g.offset + n > bitmap_bytes) {
      set_last_error("glyph " + std::to_string(i) + " bitmap lies outside the atlas");
      return NES_EINVAL;
    }
    if (g.advance < -65536 || g.advance > 65536 || g.left < -65536 || g.left > 65536 ||
        g.top < -65536 || g.top > 65536) {
      set_last_error("glyph " + std::to_string(i) + " metrics out of range");
      return NES_ERANGE;
    }
  }
  NES_TRY(ensure_device());
  auto f = std::make_unique<nes_font>();
  std::memcpy(f->glyphs, glyphs, sizeof f->glyphs);
  NES_TRY(f->d_glyphs.reserve(sizeof f->glyphs));
  NES_TRY(f->d_bitmaps.reserve(bitmap_bytes + 16));
  NES_HIP_TRY(hipMemcpy(f->d_glyphs.ptr, glyphs, sizeof f->glyphs, hipMemcpyHostToDevice));
  if (bitmap_bytes)
    NES_HIP_TRY(hipMemcpy(f->d_bitmaps.ptr, bitmaps, bitmap_bytes, hipMemcpyHostToDevice));
  *out = f.release();
  return NES_OK;
}

int nes_font_destroy(nes_font *font) {
  if (!font) return NES_EINVAL;
  NES_TRY(ensure_device());
  hipStreamSynchronize(library_stream());
  font->d_glyphs.release();
  font->d_bitmaps.release();
  font->d_place.release();
  delete font;
  return NES_OK;
}

int nes_overlay_device(nes_font *font, const nes_text *texts, int n_texts, uint8_t *rgb_dev,
                       int w, int h, int rgb_stride) {
  if (!font || (n_texts > 0 && !texts) || n_texts < 0 || !rgb_dev || w < 1 || h < 1 ||
      rgb_stride < 3 * w)
    return NES_EINVAL;
  for (int i = 0; i < n_texts; ++i)
    if (!texts[i].text) return NES_EINVAL;
  NES_TRY(ensure_device());
  std::vector<GlyphPlacement> place;
  for (int i = 0; i < n_texts; ++i)
    overlay_layout(font, texts[i].position, texts[i].text, w, h, rgb_stride, 0, place);
  std::lock_guard<std::mutex> lk(font->mu);
  hipStream_t s = library_stream();
  NES_TRY(overlay_launch(font, font->d_place, place, rgb_dev, s));
  NES_HIP_TRY(hipStreamSynchronize(s));
  return NES_OK;
}

int nes_overlay_texts(const nes_camera *cam, const char *timestamp, uint32_t flags, char *buf,
                      size_t buf_len) {
  if (!cam || !buf) return NES_EINVAL;
  std::vector<std::pair<int, std::string>> texts;
  overlay_frame_texts(*cam, timestamp, flags, texts);
  size_t need = 0;
  for (auto &t : texts) need += t.second.size() + 1;
  if (need > buf_len || need > 0x7fffffff) return NES_ERANGE;
  size_t off = 0;
  for (auto &t : texts) {
    std::memcpy(buf + off, t.second.c_str(), t.second.size() + 1);
    off += t.second.size() + 1;
  }
  return (int)need;
}

int nes_frameprep_overlay(nes_font *font, const nes_camera *cam, const char *timestamp,
                          uint32_t flags, const uint8_t *rgb_a, const float *depth_a,
                          const uint8_t *rgb_b, const float *depth_b, int w, int h,
                          int rgb_stride, uint8_t *y, int y_stride, uint8_t *u, int u_stride,
                          uint8_t *v, int v_stride, int sws_mode) {
  if (!font || !cam || !rgb_a || !y || !u || !v || w < 2 || h < 2 || (w & 1) || (h & 1))
    return NES_EINVAL;
  if (rgb_stride < 3 * w || y_stride < w || u_stride < w / 2 || v_stride < w / 2)
    return NES_EINVAL;
  if ((rgb_b == nullptr) != (depth_b == nullptr) || (rgb_b && !depth_a)) return NES_EINVAL;
  if (sws_mode != NES_SWS_X86 && sws_mode != NES_SWS_C) return NES_EINVAL;
  NES_TRY(ensure_device());
  nes_camera c = *cam;
  c.width = (uint32_t)w;
  c.height = (uint32_t)h;
  std::vector<GlyphPlacement> place;
  const size_t px = (size_t)w * h;
  const size_t rgb_bytes = 3 * px, dep_bytes = 4 * px;
  const size_t off_b = (rgb_bytes + 15) & ~size_t(15);
  const size_t off_da = (off_b + rgb_bytes + 15) & ~size_t(15);
  const size_t off_db = off_da + dep_bytes;
  // With a composite, stamping both sources is the same as stamping the
  // composited frame: a stamped pixel is white in A and in B, so it is white
  // whichever the depth test picks, and depths are untouched.
  overlay_frame_layout(font, c, timestamp, flags, 3 * w, 0, place);
  if (rgb_b) overlay_frame_layout(font, c, timestamp, flags, 3 * w, off_b, place);

  std::lock_guard<std::mutex> lk(g_ovl_host_mu);
  std::lock_guard<std::mutex> lk2(font->mu);
  const size_t cw = (size_t)w / 2, chh = (size_t)h / 2;
  NES_TRY(g_ovl_in.reserve(rgb_b ? off_db + dep_bytes + 64 : rgb_bytes + 64));
  NES_TRY(g_ovl_out.reserve(px + 2 * cw * chh + 64));
  hipStream_t s = library_stream();
  uint8_t *d_a = g_ovl_in.as<uint8_t>();
  uint8_t *d_b = rgb_b ? d_a + off_b : nullptr;
  float *d_da = rgb_b ? (float *)(d_a + off_da) : nullptr;
  float *d_db = rgb_b ? (float *)(d_a + off_db) : nullptr;
  NES_HIP_TRY(hipMemcpy2DAsync(d_a, 3 * w, rgb_a, rgb_stride, 3 * w, h, hipMemcpyHostToDevice, s));
  if (rgb_b) {
    NES_HIP_TRY(
        hipMemcpy2DAsync(d_b, 3 * w, rgb_b, rgb_stride, 3 * w, h, hipMemcpyHostToDevice, s));
    NES_HIP_TRY(hipMemcpyAsync(d_da, depth_a, dep_bytes, hipMemcpyHostToDevice, s));
    NES_HIP_TRY(hipMemcpyAsync(d_db, depth_b, dep_bytes, hipMemcpyHostToDevice, s));
  }
  NES_TRY(overlay_launch(font, font->d_place, place, d_a, s));
  uint8_t *d_y = g_ovl_out.as<uint8_t>();
  uint8_t *d_u = d_y + px, *d_v = d_u + cw * chh;
  NES_TRY(launch_frameprep(d_a, d_da, d_b, d_db, w, h, 1, 0, 0, d_y, w, 0, d_u, (int)cw, 0, d_v,
                           (int)cw, 0, sws_mode, s));
  NES_HIP_TRY(hipMemcpy2DAsync(y, y_stride, d_y, w, w, h, hipMemcpyDeviceToHost, s));
  NES_HIP_TRY(hipMemcpy2DAsync(u, u_stride, d_u, cw, cw, chh, hipMemcpyDeviceToHost, s));
  NES_HIP_TRY(hipMemcpy2DAsync(v, v_stride, d_v, cw, cw, chh, hipMemcpyDeviceToHost, s));
  NES_HIP_TRY(hipStreamSynchronize(s));
  return NES_OK;
}

}  // extern "C"
