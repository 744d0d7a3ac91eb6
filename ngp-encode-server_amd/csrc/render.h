// render.h -- device data structures of the render pipeline and the context.
#pragma once

#include <memory>

#include "nes_internal.h"

namespace nes {

// One camera of a render batch.  Ray slots of a view are 8x8 pixel tiles laid
// out tile-major so that a wavefront's 64 rays are a compact screen block.
struct ViewDesc {
  float c2w[12];
  float focal;
  uint32_t width, height, tiles_x;
  uint32_t slot_base, slot_count;  // global ray-slot range of the view
  uint32_t px_base;                // offset of the view in the output arrays
  uint32_t pad;
};

struct RayData {  // 32 bytes
  float o[3];
  float d[3];
  float t_start;
  uint32_t n_steps;  // candidate steps inside the box (0 = miss / no pixel)
};

struct RayAcc {  // 32 bytes, compositing state carried across passes
  float T, r, g, b, depth;
  uint32_t n_samples;
  uint32_t done;
  uint32_t pad;
};

// Sample record: ray slot in the high 25 bits, step index in the low 7.
constexpr uint32_t kMaxSlots = 1u << 25;
// View descriptors the pinned upload staging holds before it grows.
constexpr uint32_t kPinnedViews = 64;

// Per ray slot, the march writes counts[slot] = samples emitted by the pass in
// the low 16 bits and the step the next pass resumes at in the high 16 bits.
constexpr uint32_t kCountMask = 0xffffu;

// Words of RenderWork::total (device) and RenderWork::total_host (pinned),
// the same layout on both sides:
//   [0] sample cursor of the first pass, [1] sample cursor of a later pass
//   (0 in a single-pass frame), [2] samples evaluated by all passes of the
//   frame (added once per pass), then from word kTotTermBase kTotTermSlots
//   copies of the termination counters, one 128-byte line each (a wavefront
//   adds to copy wave % slots, and only when one of its rays reached
//   T < 1e-4 in the pass): [+0] rays that reached T < 1e-4, [+1] samples
//   those rays composited, [+2] samples evaluated behind that point, [+3]
//   rays that stopped before the last pass with steps left.
// The host copy of a pass is one transfer: word 0 or 1 after a pass that is
// not the last, words [0 or 1, kTotWords) after the last.
constexpr int kTotCursor = 0, kTotCursorLater = 1, kTotEvaluated = 2, kTotTermBase = 32,
              kTotTermSlots = 32, kTotTermStride = 32;
constexpr int kTotWords = kTotTermBase + kTotTermSlots * kTotTermStride;
constexpr size_t kTotBytes = sizeof(uint32_t) * kTotWords;

struct RenderWork {
  DevBuf views, rays, acc, counts, offsets, samples, sample_out, sample_pos, feat, out, total;
  DevBuf ray_sh;                  // mixed path: per ray slot 16 FP16 SH values
  DevBuf tinfo;                   // single-pass renders: per ray slot t_start, pixel
  DevBuf gray;                    // GRAY8 depth staging of nes_render_fetch
  uint32_t *total_host = nullptr;  // pinned
  hipEvent_t ev[12] = {};
  uint32_t n_views = 0, n_slots = 0;
  uint32_t capacity = 0;          // samples the sample buffers hold
  int capacity_precision = -1;    // feature layout the buffers were sized for
  size_t out_px = 0;
  std::vector<ViewDesc> host_views;
  // Pinned staging of the view descriptors: an upload from pageable memory
  // can block the host until the stream reaches it.  ev_views marks the
  // last upload, so the staging is rewritten only after it has been read.
  ViewDesc *views_pinned = nullptr;
  uint32_t views_pinned_cap = 0;
  hipEvent_t ev_views = nullptr;
  uint64_t budget = 0;            // device bytes the buffers may hold, 0 = no limit
  // Early termination learned from the previous frame of this work area:
  // samples the first pass gives each ray (kMaxSteps = one single pass).
  uint32_t first_pass_cap = kMaxSteps;
  uint32_t frame_passes = 1;      // passes the last frame launched
};

// What one frame evaluated and composited (for early-termination stats).
struct FrameCounts {
  uint64_t evaluated = 0, composited = 0, terminated = 0, term_samples = 0, wasted = 0,
           saved_rays = 0;
  uint32_t passes = 1;
};

// Device bytes currently held by a RenderWork's buffers.
size_t work_device_bytes(const RenderWork &w);
// Reserve every buffer a render of up to max_slots ray slots, max_px pixels
// and max_views views needs (with the initial four samples per ray slot),
// so that such renders allocate nothing.
int render_work_reserve(RenderWork &w, uint64_t max_slots, uint64_t max_px, uint32_t max_views,
                        int precision);

struct SceneDev {
  float *table_f32 = nullptr;      // entries x 2
  _Float16 *table_f16 = nullptr;   // entries x 2
  float *weights_f32 = nullptr;    // 9344
  _Float16 *wfrag = nullptr;       // 20 MFMA A-fragments x 64 lanes x 8
  uint32_t *occ = nullptr;         // 65536
};

int scene_upload(SceneDev &s, const nes_scene_desc *desc);
void scene_free(SceneDev &s);
// An empty scene copy whose device memory is freed with the last reference.
std::shared_ptr<SceneDev> new_scene();
// Permute the MLP weights into the A-operand fragments of the 20
// v_mfma_f32_16x16x32_f16 instructions one 16-sample tile uses.
void build_mlp_fragments(const float *w, _Float16 *frag);

// Byte offsets of the per-pixel outputs inside RenderWork::out.
struct OutLayout {
  size_t rgb24, depth, rgbf, nsamp, total;
};
OutLayout out_layout(size_t px);

// Stages of one render:
//   render_setup            views -> ray-slot layout, ray buffers, view upload
//   render_prepare_capacity sample buffers for the expected sample count
//   render_launch_pass      march_emit, encode, MLP, composite for march steps
//                           [k0, k1), then an async copy of the pass's sample
//                           total into work.total_host[0].  A total above
//                           work.capacity means the pass was dropped: grow with
//                           render_reserve_samples and launch the pass again.
int render_setup(RenderWork &w, const nes_camera *views, int n_views, hipStream_t stream);
int render_prepare_capacity(RenderWork &w, int precision);
int render_reserve_samples(RenderWork &w, int precision, size_t capacity);
// One pass: march steps [k0, k1) (or, with resume, from where the ray's
// previous pass stopped), at most `cap` samples per ray, then encode, MLPs
// and composite.
int render_launch_pass(const SceneDev &scene, RenderWork &w, int precision, int k0, int k1,
                       bool first, bool last, hipStream_t stream, hipEvent_t ev_eval_begin,
                       hipEvent_t ev_eval_end, hipEvent_t ev_encode_end, uint32_t cap = kMaxSteps,
                       bool resume = false);
// A whole frame without a host round trip: one pass when the work area's
// previous frame composited (almost) every sample it evaluated, else two:
// the first gives each ray at most first_pass_cap samples, the second
// continues the rays still above T = 1e-4 with all their remaining samples.
// The pass totals and composite counters land in total_host (read them after
// the stream has passed the launches).  Results are bit-identical to a single
// pass.
// ev (optional): 8 events recorded per pass p at [4p]: pass start, [4p+1]
// encode start, [4p+2] encode end, [4p+3] MLP end.
int render_launch_frame(const SceneDev &scene, RenderWork &w, int precision, hipStream_t stream,
                        const hipEvent_t *ev = nullptr);
// After the frame's stream work is complete: the largest pass total (a value
// above w.capacity means the frame must be launched again after growth) and
// the counts; updates first_pass_cap for the next frame.
uint32_t render_frame_max_total(const RenderWork &w);
FrameCounts render_frame_finish(RenderWork &w);
// Device bytes the sample buffers of `capacity` samples take (all buffers
// that grow with the sample count).
size_t sample_buffer_bytes(size_t capacity, int precision);
// Whether growing the sample buffers to `capacity` keeps the work area
// within its budget.
bool render_growth_fits(const RenderWork &w, int precision, size_t capacity);
// A frame within the current sample buffers, whatever the scene: passes that
// give each ray at most capacity / rays samples, each resuming where the
// previous one stopped (bit-identical to one pass; no overflow possible).
int render_launch_bounded(const SceneDev &scene, RenderWork &w, int precision, hipStream_t stream);

// Workgroups per XCD part of the encode (and of the training scatter) for a
// sample buffer of `capacity`: about one sample per thread, since the
// sample count itself stays on the device.  Measured on the stereo 1080p
// frame: 512 workgroups per part (grid-stride) 2.59 ms, 32768 (about one
// sample per thread) 2.20 ms.
inline uint32_t part_blocks(size_t capacity) {
  const size_t b = (capacity + 255) / 256;
  return (uint32_t)(b < 512 ? 512 : b > 49152 ? 49152 : b);
}

// Single-stage launchers (training step): march_emit over prepared rays,
// FP32 encode_parts, composite (which always stores the final per-ray state
// in work.acc for the backward).
int launch_march_emit(const uint32_t *occ, RenderWork &w, int k0, int k1, hipStream_t stream);
int launch_encode_f32(const float *table, RenderWork &w, hipStream_t stream);
int launch_mlp_f32(const float *weights, RenderWork &w, const uint32_t *n_ptr, hipStream_t stream);
int launch_composite(RenderWork &w, bool last, hipStream_t stream);

// Enqueue a full render of views on stream; returns after the sample count
// read-backs (one host sync per pass).  Outputs land in work.out.
int render_enqueue(const SceneDev &scene, RenderWork &work, int precision,
                   const nes_camera *views, int n_views, int steps_per_pass,
                   hipStream_t stream, nes_render_stats *stats);
int render_fetch(RenderWork &work, int view, const nes_view_out *out,
                 hipStream_t stream);
void render_work_free(RenderWork &w);

// Text overlay: one drawn glyph of one string on one RGB24 frame.
struct GlyphPlacement {  // 32 bytes
  uint64_t base;         // byte offset of the frame in the RGB24 buffer
  int32_t x0, y0;        // top-left corner of the glyph bitmap
  int32_t w, h, stride;  // frame size and RGB24 row pitch in bytes
  uint32_t glyph;
};
// Placements of the strings process_frame_thread draws on the frame of cam
// (its width and height are the frame's); timestamp NULL = local time now.
void overlay_frame_layout(const nes_font *font, const nes_camera &cam, const char *timestamp,
                          uint32_t flags, int stride, uint64_t base,
                          std::vector<GlyphPlacement> &out);
// Upload the placements through `staging` and stamp them on `rgb`; `place`
// stays alive and `staging` untouched until the stream has passed the launch.
int overlay_launch(const nes_font *font, DevBuf &staging,
                   const std::vector<GlyphPlacement> &place, uint8_t *rgb, hipStream_t stream);
std::string overlay_now_text();

}  // namespace nes

struct nes_ctx {
  // The scene renders read.  nes_ctx_update_scene uploads into another copy
  // and swaps the pointer between renders; a render or server batch holds a
  // reference to the copy it started with until its device work is complete,
  // so no frame mixes two scenes.  `spare` is the previous copy, reused by
  // the next update once nothing holds it.
  std::shared_ptr<nes::SceneDev> scene, spare;
  std::mutex scene_mu;  // guards scene / spare
  nes::RenderWork work;
  hipStream_t stream = nullptr;
  uint64_t budget = 0;
  std::mutex mu;  // serialises the context's own renders (work, stream)
};

namespace nes {
// The context's current scene (a reference that keeps it alive).
std::shared_ptr<SceneDev> ctx_scene(nes_ctx *ctx);
}  // namespace nes
