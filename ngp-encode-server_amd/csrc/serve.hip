// serve.hip -- continuous-batching render server (config: several VR clients,
// mixed per-eye resolutions, YUV420P frames out).
//
// Replaces the reference's request loop: one socket_client_thread per
// renderer connection, each with a single blocking FrameRequest in flight
// (src/server.cpp:116-212), feeding process_frame_thread's RGB24 -> YUV420P
// conversion of the scene and the GRAY8 depth (src/encode.cpp:42-119,
// include/base/video/rendered_frame.h:24-33).  Here:
//   * any number of client threads submit stereo requests (nes_server_submit);
//     like CameraManager (src/base/camera_manager.cc:33-69), only a client's
//     latest pose matters: a newer request supersedes one still waiting for
//     admission;
//   * one scheduler thread forms batches of waiting requests as soon as a
//     batch slot is free and fewer than two batches are queued or running on
//     the render stream.  A slot owns the device buffers of one batch
//     (allocated for the largest batch when the server starts, so no render
//     allocates) and a stream for its output work; the number of slots is
//     what the HBM budget affords, and each slot renders within its share of
//     the budget (a batch whose samples would not fit is rendered again in
//     bounded-memory passes instead of growing), so admission is by memory.
//     Renders run back to back on one render stream while earlier batches'
//     frame prep and copies overlap them, and a request arriving while others
//     render joins the next batch without waiting for them (continuous
//     batching);
//   * each batch is one asynchronous launch sequence on the scene the context
//     held when the batch was formed: the render (march, encode, MLPs,
//     composite; one pass, or two when the slot's previous batch evaluated
//     samples behind terminated rays -- the sample count never leaves the
//     device), frame prep of every view straight from the device RGB24 (and,
//     for depth-stream servers, the depth plane straight from the device
//     FP32 depth), and async copies of the YUV420P planes into host memory:
//     the server's pinned planes or the caller's own (nes_server_submit_into).
//     The scheduler polls completion events, so it never blocks on one batch
//     while another could progress.
// All per-ray work is independent of batch composition, so a request's frame
// is bit-identical alone or in any mix.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>

#include "render.h"

namespace nes {

using Clock = std::chrono::steady_clock;

static double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

constexpr int kMaxViews = 2;
constexpr int NES_ECANCELED_ = NES_ECANCELED;
// Batches queued or running on the render stream at once (see scheduler_loop).
constexpr int kRenderDepth = 2;

struct PinnedBlock {
  uint8_t *ptr = nullptr;
  size_t bytes = 0;
};

// Pinned host buffers recycled by exact size (frame sizes repeat).
class PinnedPool {
 public:
  ~PinnedPool() {
    for (auto &kv : free_)
      for (uint8_t *p : kv.second) hipHostFree(p);
  }
  int get(size_t bytes, PinnedBlock *out) {
    if (!bytes) {
      *out = PinnedBlock();
      return NES_OK;
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      auto &v = free_[bytes];
      if (!v.empty()) {
        out->ptr = v.back();
        out->bytes = bytes;
        v.pop_back();
        return NES_OK;
      }
    }
    NES_HIP_TRY(hipHostMalloc((void **)&out->ptr, bytes));
    out->bytes = bytes;
    return NES_OK;
  }
  void put(PinnedBlock b) {
    if (!b.ptr) return;
    std::lock_guard<std::mutex> lk(mu_);
    free_[b.bytes].push_back(b.ptr);
  }

 private:
  std::mutex mu_;
  std::map<size_t, std::vector<uint8_t *>> free_;
};

enum class ReqState { kWaiting, kRunning, kDone, kSuperseded, kFailed };

struct Request {
  uint64_t ticket = 0;
  uint32_t client = 0;
  int n_views = 0;
  nes_camera views[kMaxViews];
  Clock::time_point arrival, admitted, done;
  // arrivals of the requests this one superseded (served by this frame)
  std::vector<Clock::time_point> replaced;
  ReqState state = ReqState::kWaiting;
  int status = NES_OK;
  // Output planes per view: scene (planes[v]) and depth (dplanes[v]), either
  // the caller's or laid out tightly in `yuv` (pinned).
  PinnedBlock yuv;
  bool caller = false;
  nes_frame_planes planes[kMaxViews] = {}, dplanes[kMaxViews] = {};
  uint64_t rays = 0, samples = 0, batch = 0;
  uint32_t batch_requests = 0, batch_passes = 0;
};

static size_t yuv_bytes(const nes_camera &c) {
  return (size_t)c.width * c.height + 2 * (size_t)(c.width / 2) * (c.height / 2);
}

// Tight YUV420P planes of a w x h frame starting at p.
static nes_frame_planes tight_planes(uint8_t *p, const nes_camera &c) {
  nes_frame_planes f;
  f.data[0] = p;
  f.data[1] = p + (size_t)c.width * c.height;
  f.data[2] = f.data[1] + (size_t)(c.width / 2) * (c.height / 2);
  f.linesize[0] = (int32_t)c.width;
  f.linesize[1] = f.linesize[2] = (int32_t)(c.width / 2);
  return f;
}

struct Slot {
  hipStream_t stream = nullptr;    // frame prep and device-to-host copies
  hipEvent_t ev_rendered = nullptr;  // the batch's render is complete
  RenderWork work;
  DevBuf yuv;                            // device YUV420P planes of the batch
  DevBuf ovl_place;                      // text overlay placements (device)
  std::vector<GlyphPlacement> ovl_host;  // ... and their host copy
  hipEvent_t ev_done = nullptr;
  std::vector<std::shared_ptr<Request>> reqs;
  std::vector<nes_camera> views;
  std::shared_ptr<SceneDev> scene;  // the scene this batch renders
  uint64_t batch_id = 0;
  bool bounded = false;             // rendered in bounded-memory passes
  int stage = 0;                    // 0 idle, 2 in flight
  Clock::time_point launched;
};

}  // namespace nes

using namespace nes;

struct nes_server {
  nes_ctx *ctx = nullptr;
  nes_server_config cfg{};
  std::vector<std::unique_ptr<Slot>> slots;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> waiting;
  std::map<uint64_t, std::shared_ptr<Request>> tickets;
  std::map<uint32_t, std::shared_ptr<Request>> client_waiting;
  uint64_t next_ticket = 1, next_batch = 1;
  bool stop = false, hold = false;
  std::thread scheduler;
  // Every batch's render runs on this one stream, so renders never overlap
  // each other (concurrent encodes share the XCDs' L2s and run slower than
  // back to back); a slot's frame prep and copies run on its own stream and
  // overlap the next batch's render.
  hipStream_t render_stream = nullptr;
  PinnedPool pinned;
  // statistics
  std::vector<double> latencies_ms, latencies_all_ms;
  uint64_t completed = 0, superseded = 0, failed = 0, rays = 0, samples = 0, batches = 0;
  uint64_t composited = 0, bounded_batches = 0;
  Clock::time_point first_arrival, last_done;
  bool any_arrival = false;
  int last_error = NES_OK;
};

namespace nes {

// Device bytes one batch slot holds for max_rays ray slots, as the render
// allocates them: per ray slot the ray state (RayData, RayAcc), counts and
// offsets, the FP16 SH of the ray, its single-pass record, per-pixel outputs
// (RGB24, depth, float RGB, sample count) and the YUV420P frame (twice with
// the depth stream); per sample (the initial buffer holds 4 per ray slot, at
// least 2^20) the record, position, MLP output and the 8 feature planes;
// plus the 1/8 growth slack of every device buffer.  A batch never grows its
// sample buffers past the slot's share of the budget (it renders in bounded
// passes instead).
static size_t slot_bytes(uint32_t max_rays, int precision, bool depth) {
  const size_t r = max_rays;
  const size_t cap = std::max<size_t>(r * 4, (size_t)1 << 20);
  const size_t feat = precision != NES_PRECISION_MIXED ? sizeof(float4) : sizeof(uint2);
  const size_t per_ray = sizeof(RayData) + sizeof(RayAcc) + 8 + 32 + 8 + (3 + 4 + 12 + 2) +
                         (depth ? 4 : 2);
  const size_t per_sample = 4 + sizeof(float4) + sizeof(float4) + feat * 8;  // 8 XCD part planes
  return (r * per_ray + cap * per_sample) / 8 * 9 + ((size_t)1 << 20);
}

// Device-to-host copy of one view's YUV420P planes (tight on the device)
// into the request's planes.
static int copy_planes(const nes_frame_planes &dst, const uint8_t *src, const nes_camera &c,
                       hipStream_t stream) {
  const nes_frame_planes s = tight_planes(const_cast<uint8_t *>(src), c);
  const size_t w[3] = {c.width, c.width / 2, c.width / 2};
  const size_t h[3] = {c.height, c.height / 2, c.height / 2};
  for (int p = 0; p < 3; ++p)
    NES_HIP_TRY(hipMemcpy2DAsync(dst.data[p], (size_t)dst.linesize[p], s.data[p],
                                 (size_t)s.linesize[p], w[p], h[p], hipMemcpyDeviceToHost,
                                 stream));
  return NES_OK;
}

// Launch frame prep of every view of a finished render and the D2H copies.
static int launch_output(nes_server *s, Slot &sl) {
  RenderWork &w = sl.work;
  const OutLayout L = out_layout(w.out_px);
  const bool depth = s->cfg.depth_stream != 0;
  size_t yuv_total = 0;
  for (const nes_camera &c : sl.views) yuv_total += yuv_bytes(c) * (depth ? 2 : 1);
  NES_TRY(sl.yuv.reserve(yuv_total + 256));
  if (s->cfg.font && s->cfg.overlay_flags) {
    // the strings of every view, stamped by one launch before frame prep
    sl.ovl_host.clear();
    const std::string now = (s->cfg.overlay_flags & NES_OVERLAY_TIMESTAMP)
                                ? overlay_now_text() : std::string();
    for (size_t i = 0; i < sl.views.size(); ++i)
      overlay_frame_layout(s->cfg.font, sl.views[i], now.c_str(), s->cfg.overlay_flags,
                           3 * (int)sl.views[i].width, L.rgb24 + 3 * (size_t)w.host_views[i].px_base,
                           sl.ovl_host);
    NES_TRY(overlay_launch(s->cfg.font, sl.ovl_place, sl.ovl_host, w.out.as<uint8_t>(),
                           sl.stream));
  }
  // Frame prep of every view first, then every copy: each change between a
  // kernel and a copy on the slot stream is a hand-over between the
  // dispatcher and the copy engine that costs tens of microseconds, so the
  // batch makes it once instead of once per plane.
  const float scale = depth_scale(s->cfg.depth_far);
  size_t off = 0;
  size_t vi = 0;
  for (auto &rq : sl.reqs) {
    for (int v = 0; v < rq->n_views; ++v, ++vi) {
      const nes_camera &c = sl.views[vi];
      const ViewDesc &vd = w.host_views[vi];
      const int cw = (int)c.width, ch = (int)c.height;
      const uint8_t *rgb = w.out.as<uint8_t>() + L.rgb24 + 3 * (size_t)vd.px_base;
      const nes_frame_planes d = tight_planes(sl.yuv.as<uint8_t>() + off, c);
      NES_TRY(launch_frameprep(rgb, nullptr, nullptr, nullptr, cw, ch, 1, 0, 0, d.data[0], cw, 0,
                               d.data[1], cw / 2, 0, d.data[2], cw / 2, 0, s->cfg.sws_mode,
                               sl.stream));
      off += yuv_bytes(c);
      if (depth) {
        const float *dep = reinterpret_cast<const float *>(w.out.as<uint8_t>() + L.depth) +
                           vd.px_base;
        const nes_frame_planes dd = tight_planes(sl.yuv.as<uint8_t>() + off, c);
        NES_TRY(launch_depth_yuv(dep, cw, ch, scale, nullptr, dd.data[0], cw, dd.data[1], cw / 2,
                                 dd.data[2], cw / 2, s->cfg.sws_mode, sl.stream));
        off += yuv_bytes(c);
      }
    }
  }
  // The server's own planes of a request are laid out exactly as its part
  // of the staging (scene, then depth, per view): one copy per request.
  off = 0;
  for (auto &rq : sl.reqs) {
    size_t bytes = 0;
    for (int v = 0; v < rq->n_views; ++v) bytes += yuv_bytes(rq->views[v]) * (depth ? 2 : 1);
    if (!rq->caller) {
      NES_HIP_TRY(hipMemcpyAsync(rq->yuv.ptr, sl.yuv.as<uint8_t>() + off, bytes,
                                 hipMemcpyDeviceToHost, sl.stream));
    } else {
      size_t o = off;
      for (int v = 0; v < rq->n_views; ++v) {
        const nes_camera &c = rq->views[v];
        NES_TRY(copy_planes(rq->planes[v], sl.yuv.as<uint8_t>() + o, c, sl.stream));
        o += yuv_bytes(c);
        if (depth) {
          NES_TRY(copy_planes(rq->dplanes[v], sl.yuv.as<uint8_t>() + o, c, sl.stream));
          o += yuv_bytes(c);
        }
      }
    }
    off += bytes;
  }
  return NES_OK;
}

// Take waiting requests (FIFO) that fit one batch.
static void form_batch(nes_server *s, Slot &sl) {
  sl.reqs.clear();
  sl.views.clear();
  uint64_t slots_used = 0;
  while (!s->waiting.empty()) {
    auto rq = s->waiting.front();
    uint64_t need = 0;
    for (int v = 0; v < rq->n_views; ++v)
      need += (uint64_t)((rq->views[v].width + 7) / 8) * ((rq->views[v].height + 7) / 8) * 64;
    if (!sl.reqs.empty() && slots_used + need > s->cfg.max_batch_rays) break;
    s->waiting.pop_front();
    auto it = s->client_waiting.find(rq->client);
    if (it != s->client_waiting.end() && it->second == rq) s->client_waiting.erase(it);
    rq->state = ReqState::kRunning;
    rq->admitted = Clock::now();
    slots_used += need;
    for (int v = 0; v < rq->n_views; ++v) sl.views.push_back(rq->views[v]);
    sl.reqs.push_back(rq);
  }
}

static void finish_batch(nes_server *s, Slot &sl, int status) {
  const Clock::time_point now = Clock::now();
  FrameCounts fc;
  if (status == NES_OK) fc = render_frame_finish(sl.work);
  uint64_t batch_rays = 0;
  for (auto &rq : sl.reqs)
    for (int v = 0; v < rq->n_views; ++v) batch_rays += (uint64_t)rq->views[v].width * rq->views[v].height;
  std::lock_guard<std::mutex> lk(s->mu);
  for (auto &rq : sl.reqs) {
    rq->done = now;
    rq->status = status;
    rq->state = status == NES_OK ? ReqState::kDone : ReqState::kFailed;
    uint64_t r = 0;
    for (int v = 0; v < rq->n_views; ++v) r += (uint64_t)rq->views[v].width * rq->views[v].height;
    rq->rays = r;
    rq->samples = batch_rays ? (uint64_t)((double)fc.evaluated * (double)r / (double)batch_rays) : 0;
    rq->batch = sl.batch_id;
    rq->batch_requests = (uint32_t)sl.reqs.size();
    rq->batch_passes = fc.passes;
    if (status == NES_OK) {
      s->completed++;
      s->rays += r;
      const double lat = ms_between(rq->arrival, now);
      s->latencies_ms.push_back(lat);
      s->latencies_all_ms.push_back(lat);
      for (const Clock::time_point &a : rq->replaced) s->latencies_all_ms.push_back(ms_between(a, now));
      s->last_done = now;
    } else {
      s->failed++;
    }
  }
  s->samples += fc.evaluated;
  s->composited += fc.composited;
  s->bounded_batches += sl.bounded ? 1 : 0;
  s->batches++;
  sl.reqs.clear();
  sl.scene.reset();
  sl.stage = 0;
  s->cv_done.notify_all();
}

// Launch the whole batch and its done event: the render on the server's
// render stream, then frame prep and D2H on the slot's stream once the render
// is complete.
static int launch_batch(nes_server *s, Slot &sl) {
  hipStream_t rs = s->render_stream;
  if (sl.bounded)
    NES_TRY(render_launch_bounded(*sl.scene, sl.work, s->cfg.precision, rs));
  else
    NES_TRY(render_launch_frame(*sl.scene, sl.work, s->cfg.precision, rs));
  NES_HIP_TRY(hipEventRecord(sl.ev_rendered, rs));
  NES_HIP_TRY(hipStreamWaitEvent(sl.stream, sl.ev_rendered, 0));
  NES_TRY(launch_output(s, sl));
  NES_HIP_TRY(hipEventRecord(sl.ev_done, sl.stream));
  return NES_OK;
}

static void scheduler_loop(nes_server *s) {
  hipSetDevice(0);
  while (true) {
    bool progressed = false;
    for (auto &slp : s->slots) {
      Slot &sl = *slp;
      if (sl.stage != 2 || hipEventQuery(sl.ev_done) != hipSuccess) continue;
      progressed = true;
      const uint32_t n = render_frame_max_total(sl.work);
      if (!sl.bounded && n > sl.work.capacity) {
        // the frame dropped samples beyond capacity: grow within the slot's
        // share of the budget and run it again, or run it in bounded passes
        const size_t want = (size_t)n + n / 8 + 1024;
        int rc = NES_OK;
        if (render_growth_fits(sl.work, s->cfg.precision, want))
          rc = render_reserve_samples(sl.work, s->cfg.precision, want);
        else
          sl.bounded = true;
        if (rc == NES_OK) rc = launch_batch(s, sl);
        if (rc == NES_OK) continue;
        s->last_error = rc;
        hipStreamSynchronize(s->render_stream);
        hipStreamSynchronize(sl.stream);
        finish_batch(s, sl, rc);
      } else {
        finish_batch(s, sl, NES_OK);
      }
    }
    // admit waiting requests into free slots, keeping at most kRenderDepth
    // batches queued or running on the render stream: one renders while the
    // next waits, so the stream never drains while a slot is busy with its
    // output, and requests that arrive meanwhile join one batch instead of
    // being spread over many small ones
    {
      int rendering = 0;
      for (auto &slp : s->slots)
        if (slp->stage == 2 && hipEventQuery(slp->ev_rendered) != hipSuccess) ++rendering;
      std::unique_lock<std::mutex> lk(s->mu);
      for (auto &slp : s->slots) {
        Slot &sl = *slp;
        if (rendering >= kRenderDepth) break;
        if (sl.stage != 0 || s->waiting.empty() || s->hold) continue;
        form_batch(s, sl);
        if (sl.reqs.empty()) continue;
        sl.batch_id = s->next_batch++;
        lk.unlock();
        sl.scene = ctx_scene(s->ctx);  // kept until the batch completes
        sl.bounded = false;
        // the view upload goes with the render (the slot's previous batch is
        // complete, so its buffers are free)
        int rc = render_setup(sl.work, sl.views.data(), (int)sl.views.size(), s->render_stream);
        if (rc == NES_OK) rc = render_prepare_capacity(sl.work, s->cfg.precision);
        if (rc == NES_OK) rc = launch_batch(s, sl);
        sl.launched = Clock::now();
        if (rc != NES_OK) {
          s->last_error = rc;
          hipStreamSynchronize(s->render_stream);
          hipStreamSynchronize(sl.stream);
          finish_batch(s, sl, rc);
        } else {
          sl.stage = 2;
          ++rendering;
        }
        lk.lock();
        progressed = true;
      }
      bool busy = false;
      for (auto &slp : s->slots) busy = busy || slp->stage != 0;
      if (s->stop && !busy && (s->waiting.empty() || s->hold)) break;
      if (!progressed) {
        if (busy)
          s->cv_work.wait_for(lk, std::chrono::microseconds(50));
        else
          s->cv_work.wait_for(lk, std::chrono::milliseconds(5));
      }
    }
  }
}

static bool planes_ok(const nes_frame_planes &p, const nes_camera &c) {
  return p.data[0] && p.data[1] && p.data[2] && p.linesize[0] >= (int32_t)c.width &&
         p.linesize[1] >= (int32_t)(c.width / 2) && p.linesize[2] >= (int32_t)(c.width / 2);
}

static int submit(nes_server *s, uint32_t client, const nes_camera *views, int n_views,
                  const nes_frame_planes *planes, const nes_frame_planes *dplanes,
                  uint64_t *ticket) {
  if (!s || !views || n_views < 1 || n_views > kMaxViews || !ticket) return NES_EINVAL;
  const bool depth = s->cfg.depth_stream != 0;
  auto rq = std::make_shared<Request>();
  rq->client = client;
  rq->n_views = n_views;
  rq->caller = planes != nullptr;
  size_t total = 0;
  for (int v = 0; v < n_views; ++v) {
    const nes_camera &c = views[v];
    if (c.width < 2 || c.height < 2 || (c.width & 1) || (c.height & 1) || c.width > 16384 ||
        c.height > 16384 || !(c.fov_y_deg > 0.0f && c.fov_y_deg < 180.0f))
      return NES_EINVAL;
    if (rq->caller && (!planes_ok(planes[v], c) || (depth && (!dplanes || !planes_ok(dplanes[v], c))))) {
      set_last_error("caller planes missing or narrower than the frame");
      return NES_EINVAL;
    }
    rq->views[v] = c;
    total += yuv_bytes(c) * (depth ? 2 : 1);
  }
  if (!rq->caller) {
    NES_TRY(s->pinned.get(total, &rq->yuv));
    size_t off = 0;
    for (int v = 0; v < n_views; ++v) {
      rq->planes[v] = tight_planes(rq->yuv.ptr + off, views[v]);
      off += yuv_bytes(views[v]);
      if (depth) {
        rq->dplanes[v] = tight_planes(rq->yuv.ptr + off, views[v]);
        off += yuv_bytes(views[v]);
      }
    }
  } else {
    for (int v = 0; v < n_views; ++v) {
      rq->planes[v] = planes[v];
      if (depth) rq->dplanes[v] = dplanes[v];
    }
  }
  rq->arrival = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) {
      s->pinned.put(rq->yuv);
      return NES_EINVAL;
    }
    if (s->waiting.size() >= s->cfg.queue_capacity) {
      s->pinned.put(rq->yuv);
      return NES_EBUSY;
    }
    // latest pose wins: a newer request of the same client replaces one that
    // has not been admitted yet, and is the frame that serves it
    auto it = s->client_waiting.find(client);
    if (it != s->client_waiting.end()) {
      auto old = it->second;
      old->state = ReqState::kSuperseded;
      old->status = NES_ECANCELED_;
      rq->replaced = old->replaced;
      rq->replaced.push_back(old->arrival);
      s->waiting.erase(std::remove(s->waiting.begin(), s->waiting.end(), old), s->waiting.end());
      s->superseded++;
    }
    rq->ticket = s->next_ticket++;
    s->tickets[rq->ticket] = rq;
    s->client_waiting[client] = rq;
    s->waiting.push_back(rq);
    if (!s->any_arrival) {
      s->any_arrival = true;
      s->first_arrival = rq->arrival;
    }
    *ticket = rq->ticket;
  }
  s->cv_work.notify_one();
  s->cv_done.notify_all();
  return NES_OK;
}

static double percentile(std::vector<double> l, double p) {
  if (l.empty()) return 0.0;
  std::sort(l.begin(), l.end());
  const double idx = p * (double)(l.size() - 1);
  const size_t lo = (size_t)std::floor(idx), hi = std::min(lo + 1, l.size() - 1);
  return l[lo] + (l[hi] - l[lo]) * (idx - (double)lo);
}

}  // namespace nes

namespace nes {

// The slots' device buffers, events and streams and the render stream.
static void server_free_device(nes_server *s) {
  if (s->render_stream) hipStreamSynchronize(s->render_stream);
  for (auto &sl : s->slots) {
    if (sl->stream) hipStreamSynchronize(sl->stream);
    render_work_free(sl->work);
    sl->yuv.release();
    sl->ovl_place.release();
    if (sl->ev_done) hipEventDestroy(sl->ev_done);
    if (sl->ev_rendered) hipEventDestroy(sl->ev_rendered);
    if (sl->stream) hipStreamDestroy(sl->stream);
  }
  s->slots.clear();
  if (s->render_stream) hipStreamDestroy(s->render_stream);
  s->render_stream = nullptr;
}

}  // namespace nes

extern "C" {

int nes_server_create(nes_ctx *ctx, const nes_server_config *cfg, nes_server **out) {
  if (!ctx || !out) return NES_EINVAL;
  NES_TRY(ensure_device());
  auto s = std::make_unique<nes_server>();
  s->ctx = ctx;
  if (cfg) s->cfg = *cfg;
  if (!s->cfg.hbm_budget_bytes) s->cfg.hbm_budget_bytes = 48ull << 30;
  if (!s->cfg.max_batch_rays) s->cfg.max_batch_rays = 1u << 24;
  if (!s->cfg.max_inflight) s->cfg.max_inflight = 3;
  if (!s->cfg.queue_capacity) s->cfg.queue_capacity = 256;
  if (!(s->cfg.depth_far > 0.0f)) s->cfg.depth_far = NES_DEPTH_FAR_DEFAULT;
  if (s->cfg.max_batch_rays > kMaxSlots) return NES_ERANGE;
  if (s->cfg.precision != NES_PRECISION_FP32 && s->cfg.precision != NES_PRECISION_MIXED &&
      s->cfg.precision != NES_PRECISION_FP32_VALU)
    return NES_EINVAL;
  if (s->cfg.sws_mode != NES_SWS_X86 && s->cfg.sws_mode != NES_SWS_C) return NES_EINVAL;
  if (s->cfg.overlay_flags & ~NES_OVERLAY_ALL) return NES_EINVAL;
  if (s->cfg.depth_stream > 1) return NES_EINVAL;
  // admission by memory: as many batch slots as the budget holds, each
  // rendering within an equal share of it
  const bool depth = s->cfg.depth_stream != 0;
  const size_t per_slot = slot_bytes(s->cfg.max_batch_rays, s->cfg.precision, depth);
  const uint64_t afford = s->cfg.hbm_budget_bytes / per_slot;
  if (afford < 1) {
    set_last_error("HBM budget below one batch slot");
    return NES_ENOMEM;
  }
  const uint32_t n_slots = (uint32_t)std::min<uint64_t>(s->cfg.max_inflight, afford);
  const uint64_t share = s->cfg.hbm_budget_bytes / n_slots;
  // the YUV staging and overlay buffers of a slot sit outside its render work
  const uint64_t staging = (uint64_t)s->cfg.max_batch_rays * (depth ? 4 : 2) + (2u << 20);
  auto build = [&]() -> int {
    NES_HIP_TRY(hipStreamCreateWithFlags(&s->render_stream, hipStreamNonBlocking));
    for (uint32_t i = 0; i < n_slots; ++i) {
      s->slots.push_back(std::make_unique<Slot>());
      Slot &sl = *s->slots.back();
      sl.work.budget = share - staging;
      NES_HIP_TRY(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
      NES_HIP_TRY(hipEventCreateWithFlags(&sl.ev_done, hipEventDisableTiming));
      NES_HIP_TRY(hipEventCreateWithFlags(&sl.ev_rendered, hipEventDisableTiming));
      // Allocate the slot's buffers for the largest batch now: growing one
      // later frees the old buffer, and hipFree waits for the whole device,
      // draining the render stream while the scheduler thread is blocked.
      if (per_slot <= sl.work.budget) {
        NES_TRY(render_work_reserve(sl.work, s->cfg.max_batch_rays, s->cfg.max_batch_rays,
                                    kPinnedViews, s->cfg.precision));
        // YUV420P is 1.5 bytes per pixel, twice with the depth stream
        NES_TRY(sl.yuv.reserve((size_t)s->cfg.max_batch_rays * 3 / 2 * (depth ? 2 : 1) + 256));
      }
    }
    return NES_OK;
  };
  if (const int rc = build(); rc != NES_OK) {
    server_free_device(s.get());  // what was created before the failure
    return rc;
  }
  nes_server *raw = s.release();
  raw->scheduler = std::thread(scheduler_loop, raw);
  *out = raw;
  return NES_OK;
}

int nes_server_destroy(nes_server *s) {
  if (!s) return NES_EINVAL;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    // requests never admitted are cancelled
    for (auto &rq : s->waiting) {
      rq->state = ReqState::kSuperseded;
      rq->status = NES_ECANCELED_;
    }
    s->waiting.clear();
  }
  s->cv_work.notify_all();
  if (s->scheduler.joinable()) s->scheduler.join();
  server_free_device(s);
  for (auto &kv : s->tickets) s->pinned.put(kv.second->yuv);
  delete s;
  return NES_OK;
}

int nes_server_submit(nes_server *s, uint32_t client, const nes_camera *views, int n_views,
                      uint64_t *ticket) {
  return submit(s, client, views, n_views, nullptr, nullptr, ticket);
}

int nes_server_submit_into(nes_server *s, uint32_t client, const nes_camera *views, int n_views,
                           const nes_frame_planes *planes, const nes_frame_planes *depth_planes,
                           uint64_t *ticket) {
  if (!planes) return NES_EINVAL;
  return submit(s, client, views, n_views, planes, depth_planes, ticket);
}

int nes_server_hold(nes_server *s, int hold) {
  if (!s) return NES_EINVAL;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->hold = hold != 0;
  }
  s->cv_work.notify_all();
  return NES_OK;
}

int nes_server_wait(nes_server *s, uint64_t ticket, int timeout_ms, nes_frame_result *out) {
  if (!s || !out) return NES_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->tickets.find(ticket);
  if (it == s->tickets.end()) return NES_EINVAL;
  auto rq = it->second;
  auto ready = [&] { return rq->state != ReqState::kWaiting && rq->state != ReqState::kRunning; };
  if (timeout_ms < 0)
    s->cv_done.wait(lk, ready);
  else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
    return NES_EAGAIN;
  std::memset(out, 0, sizeof(*out));
  out->ticket = rq->ticket;
  out->index = rq->views[0].index;
  out->n_views = (uint32_t)rq->n_views;
  out->status = rq->status;
  if (rq->state == ReqState::kDone) {
    const bool depth = s->cfg.depth_stream != 0;
    for (int v = 0; v < rq->n_views; ++v) {
      const nes_camera &c = rq->views[v];
      out->width[v] = c.width;
      out->height[v] = c.height;
      out->y[v] = rq->planes[v].data[0];
      out->u[v] = rq->planes[v].data[1];
      out->v[v] = rq->planes[v].data[2];
      for (int p = 0; p < 3; ++p) out->linesize[v][p] = rq->planes[v].linesize[p];
      if (depth) {
        out->depth_y[v] = rq->dplanes[v].data[0];
        out->depth_u[v] = rq->dplanes[v].data[1];
        out->depth_v[v] = rq->dplanes[v].data[2];
        for (int p = 0; p < 3; ++p) out->depth_linesize[v][p] = rq->dplanes[v].linesize[p];
      }
    }
    out->latency_ms = ms_between(rq->arrival, rq->done);
    out->queue_ms = ms_between(rq->arrival, rq->admitted);
    out->rays = rq->rays;
    out->samples = rq->samples;
    out->batch = rq->batch;
    out->batch_requests = rq->batch_requests;
    out->batch_passes = rq->batch_passes;
  }
  return rq->status;
}

int nes_server_release(nes_server *s, uint64_t ticket) {
  if (!s) return NES_EINVAL;
  std::shared_ptr<Request> rq;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return NES_EINVAL;
    rq = it->second;
    if (rq->state == ReqState::kWaiting || rq->state == ReqState::kRunning) return NES_EBUSY;
    s->tickets.erase(it);
  }
  s->pinned.put(rq->yuv);
  rq->yuv = PinnedBlock();
  return NES_OK;
}

int nes_server_stats_get(nes_server *s, nes_server_stats *st) {
  if (!s || !st) return NES_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(st, 0, sizeof(*st));
  st->completed = s->completed;
  st->superseded = s->superseded;
  st->failed = s->failed;
  st->rays = s->rays;
  st->samples = s->samples;
  st->batches = s->batches;
  st->composited = s->composited;
  st->bounded_batches = s->bounded_batches;
  st->slots = (uint32_t)s->slots.size();
  if (!s->latencies_ms.empty()) {
    st->p50_ms = percentile(s->latencies_ms, 0.50);
    st->p99_ms = percentile(s->latencies_ms, 0.99);
    st->max_ms = *std::max_element(s->latencies_ms.begin(), s->latencies_ms.end());
    double sum = 0;
    for (double x : s->latencies_ms) sum += x;
    st->mean_ms = sum / (double)s->latencies_ms.size();
    st->p50_all_ms = percentile(s->latencies_all_ms, 0.50);
    st->p99_all_ms = percentile(s->latencies_all_ms, 0.99);
  }
  if (s->any_arrival && s->completed) {
    st->wall_s = std::chrono::duration<double>(s->last_done - s->first_arrival).count();
    if (st->wall_s > 0) st->mrays_per_s = (double)s->rays / st->wall_s / 1e6;
  }
  return NES_OK;
}

int nes_server_reset_stats(nes_server *s) {
  if (!s) return NES_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  s->latencies_ms.clear();
  s->latencies_all_ms.clear();
  s->completed = s->superseded = s->failed = s->rays = s->samples = s->batches = 0;
  s->composited = s->bounded_batches = 0;
  s->any_arrival = false;
  return NES_OK;
}

// Poisson load generator: n_clients independent clients, each with
// exponential inter-arrival times at rate_hz, a seeded head pose on a
// radius-2 orbit, a seeded FOV in [fov_min, fov_max] and a per-eye
// resolution drawn from res_wh; stereo pairs with 64 mm IPD.  Every frame
// that completes is waited for and released; returns the server statistics
// of the run (latency = arrival -> YUV420P planes in host memory).
int nes_server_simulate(nes_server *s, const nes_sim_config *sc, nes_server_stats *st) {
  if (!s || !sc || !st || sc->n_clients < 1 || !(sc->rate_hz > 0) || !(sc->duration_s > 0) ||
      sc->n_res < 1 || !sc->res_wh)
    return NES_EINVAL;
  NES_TRY(nes_server_reset_stats(s));
  struct Arrival {
    double t;
    uint32_t client;
    uint64_t k;
  };
  std::vector<Arrival> arr;
  for (uint32_t c = 0; c < sc->n_clients; ++c) {
    double t = 0;
    for (uint64_t k = 0;; ++k) {
      const double u = rand_uniform(sc->seed, 5000 + c, 4 * k);
      t += -std::log(1.0 - u) / sc->rate_hz;
      if (t >= sc->duration_s) break;
      arr.push_back({t, c, k});
    }
  }
  std::sort(arr.begin(), arr.end(), [](const Arrival &a, const Arrival &b) { return a.t < b.t; });
  std::mutex tmu;
  std::condition_variable tcv;
  std::deque<uint64_t> pending;
  bool gen_done = false;
  int first_err = NES_OK;
  std::thread collector([&] {
    while (true) {
      uint64_t t;
      {
        std::unique_lock<std::mutex> lk(tmu);
        tcv.wait(lk, [&] { return !pending.empty() || gen_done; });
        if (pending.empty()) return;
        t = pending.front();
        pending.pop_front();
      }
      nes_frame_result r;
      nes_server_wait(s, t, -1, &r);
      nes_server_release(s, t);
    }
  });
  const Clock::time_point t0 = Clock::now();
  for (const Arrival &a : arr) {
    std::this_thread::sleep_until(t0 + std::chrono::duration_cast<Clock::duration>(
                                           std::chrono::duration<double>(a.t)));
    const uint64_t base = 4 * a.k;
    const double az = 360.0 * rand_uniform(sc->seed, 6000 + a.client, base);
    const double el = 5.0 + 30.0 * rand_uniform(sc->seed, 6000 + a.client, base + 1);
    const double fu = rand_uniform(sc->seed, 6000 + a.client, base + 2);
    const uint32_t ri = (uint32_t)(rand_u64(sc->seed, 6000 + a.client, base + 3) % sc->n_res);
    nes_camera v[2];
    for (int e = 0; e < 2; ++e) {
      std::memset(&v[e], 0, sizeof(nes_camera));
      v[e].index = a.k;
      v[e].is_left = e == 0;
      v[e].width = sc->res_wh[2 * ri];
      v[e].height = sc->res_wh[2 * ri + 1];
      v[e].fov_y_deg = (float)(sc->fov_min + (sc->fov_max - sc->fov_min) * fu);
      nes_orbit_camera(az, el, 2.0, e == 0 ? -0.032 : 0.032, v[e].c2w);
    }
    uint64_t ticket = 0;
    const int rc = nes_server_submit(s, a.client, v, sc->stereo ? 2 : 1, &ticket);
    if (rc == NES_OK) {
      std::lock_guard<std::mutex> lk(tmu);
      pending.push_back(ticket);
      tcv.notify_one();
    } else if (rc != NES_EBUSY && first_err == NES_OK) {
      first_err = rc;
    }
  }
  {
    std::lock_guard<std::mutex> lk(tmu);
    gen_done = true;
  }
  tcv.notify_all();
  collector.join();
  NES_TRY(first_err);
  return nes_server_stats_get(s, st);
}

}  // extern "C"
