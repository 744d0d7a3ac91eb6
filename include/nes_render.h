/*
 * nes_render.h -- C-ABI of the MI355X render-and-frame-prep library
 * (libnes_render.so).
 *
 * The reference server (neserver) crosses two boundaries on this path:
 *   - a process boundary to an external instant-ngp renderer: one
 *     length-prefixed nesproto::FrameRequest out, one nesproto::RenderedFrame
 *     back per connection (src/server.cpp:116-212, proto/nes.proto:5-25);
 *   - the libswscale C ABI for frame prep (sws_getContext / sws_scale /
 *     sws_freeContext, src/base/video/type_managers.cc:143-155, called from
 *     include/base/video/rendered_frame.h:24-33).
 * This header replaces both with in-process calls into hand-written gfx950
 * kernels.  Plain C: no exceptions cross it, every entry point returns 0 or a
 * negative errno-style code (mirroring socket_send_blocking's -errno,
 * src/server.cpp:37).  Callers own host buffers; the library owns device
 * memory and streams.
 */
#ifndef NES_RENDER_H_
#define NES_RENDER_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ codes */
#define NES_OK 0
#define NES_EIO (-5)      /* device / runtime failure */
#define NES_EAGAIN (-11)  /* not finished yet (poll) */
#define NES_ENOMEM (-12)  /* device memory or HBM budget exhausted */
#define NES_EBUSY (-16)   /* queue full */
#define NES_ENODEV (-19)  /* no usable gfx950 device */
#define NES_EINVAL (-22)  /* bad argument (odd frame size, null pointer...) */
#define NES_ERANGE (-34)  /* size beyond a documented limit */

const char *nes_strerror(int code);

/* ----------------------------------------------------------------- camera */
/* nesproto::Camera (proto/nes.proto:5-10) + FrameRequest.index
 * (proto/nes.proto:12-16) + the vertical field of view the wire format lacks.
 * c2w is row-major 3x4 camera-to-world; columns are camera right, image-down,
 * forward and position (include/base/camera_manager.h:18-19). */
typedef struct nes_camera {
  uint64_t index;
  uint32_t is_left;
  uint32_t width;
  uint32_t height;
  float c2w[12];
  float fov_y_deg;
} nes_camera;

/* ------------------------------------------------------------------ scene */
#define NES_HASH_LEVELS 16
#define NES_HASH_FEATURES 2
#define NES_TABLE_ENTRIES 6098925u  /* sum over levels (dense levels 0-4) */
#define NES_MLP_WEIGHTS 9344u       /* W1 64x32, W2 16x64, W3 64x31, W4 64x64, W5 3x64 */
#define NES_OCC_WORDS 65536u        /* 128^3 bits */
#define NES_MAX_STEPS 128

typedef struct nes_scene_desc {
  const float *table;         /* NES_TABLE_ENTRIES x 2 f32, entry-major */
  const float *weights;       /* NES_MLP_WEIGHTS f32, row-major [out][in] */
  const uint32_t *occupancy;  /* NES_OCC_WORDS, bit x+128y+16384z */
} nes_scene_desc;

/* Fill host buffers with the seeded synthetic scene (any pointer may be NULL).
 * table_init 0: N(0, 0.1); 1: U(-1e-4, 1e-4). */
int nes_scene_synthesize(uint64_t seed, int table_init, float *table,
                         float *weights, uint32_t *occupancy);
/* Camera on a radius-r orbit around the cube centre, looking at it; the eye
 * is displaced by eye_offset along the camera right axis. */
int nes_orbit_camera(double azimuth_deg, double elevation_deg, double radius,
                     double eye_offset, float c2w[12]);

/* ---------------------------------------------------------------- context */
#define NES_PRECISION_FP32 0       /* FP32 table, FP32 MLPs on FP32 MFMA */
#define NES_PRECISION_MIXED 1      /* FP16 table, FP16 MFMA MLPs (FP32 accum) */
#define NES_PRECISION_FP32_VALU 2  /* FP32 table, FP32 MLPs per lane on the VALU
                                      (the plain FP32 GPU path) */

typedef struct nes_ctx nes_ctx;

/* Uploads the scene to the device.  hbm_budget_bytes bounds the device
 * buffers of the context's own renders (nes_render*, 0 = 16 GiB): a render
 * whose ray and sample buffers would exceed it fails with NES_ENOMEM and
 * releases them.  Replaces connecting to a renderer (socket_manage_thread,
 * src/server.cpp:216-283).  Destroy servers and wire endpoints built on a
 * context before the context. */
int nes_ctx_create(const nes_scene_desc *scene, uint64_t hbm_budget_bytes,
                   nes_ctx **out);
int nes_ctx_destroy(nes_ctx *ctx);
/* Replace the scene parameters (e.g. after training).  The new values go to a
 * second device copy of the scene, which then replaces the current one
 * between renders: a render or server batch already launched finishes on the
 * scene it started with, every later one uses the new scene (no frame mixes
 * the two). */
int nes_ctx_update_scene(nes_ctx *ctx, const nes_scene_desc *scene);

/* --------------------------------------------------------------- render */
/* One view's host outputs; any pointer may be NULL.  rgb24 is tightly packed
 * W*H*3, top row first (as RenderedFrame.frame, src/base/video/
 * rendered_frame.cc:14-18); depth is W*H f32 expected depth; rgb_f32 is the
 * unquantised colour W*H*3; n_samples W*H evaluated samples per pixel.
 * depth_gray8 is the W*H GRAY8 depth the reference's RenderedFrame.depth
 * carries (proto/nes.proto:24, consumed as AV_PIX_FMT_GRAY8 at
 * src/server.cpp:193-194): g = clamp(floor(fma(d, 255 / depth_far, 0.5)), 0,
 * 255) in FP32, NaN -> 0, computed on the device; depth_far <= 0 selects
 * NES_DEPTH_FAR_DEFAULT. */
#define NES_DEPTH_FAR_DEFAULT 4.0f
typedef struct nes_view_out {
  uint8_t *rgb24;
  float *depth;
  float *rgb_f32;
  uint16_t *n_samples;
  uint8_t *depth_gray8;
  float depth_far;
} nes_view_out;

typedef struct nes_render_stats {
  uint64_t rays;
  uint64_t samples;         /* evaluated network samples */
  double ms_total;          /* device time of the whole render */
  double ms_eval;           /* device time of the network kernels (encode + MLPs) */
  double ms_march;
  double ms_composite;
  double ms_encode;         /* hash-grid encode kernel (part of ms_eval) */
  double ms_mlp;            /* MLP kernel (part of ms_eval) */
  uint64_t composited;      /* samples composited (those before T < 1e-4) */
  uint64_t terminated_rays; /* rays that reached T < 1e-4 */
  uint32_t passes;          /* march passes the render launched */
} nes_render_stats;

/* Synchronous render of n views (e.g. a stereo pair) in one launch sequence:
 * the in-process replacement of one FrameRequest -> RenderedFrame round trip
 * (src/server.cpp:140-194).
 * steps_per_pass 0 (default): the whole frame without a host round trip, in
 *   one pass, or in two when the previous render of this context evaluated
 *   more than 5 % of its samples behind rays that had already reached
 *   T < 1e-4: the first pass then gives each ray as many samples as such rays
 *   needed on average, the second continues the rays still above 1e-4.
 * steps_per_pass >= 128: exactly one pass.
 * 1..127: windows of that many march steps, one host round trip each.
 * Every mode gives bit-identical results. */
int nes_render(nes_ctx *ctx, int precision, const nes_camera *views, int n_views,
               const nes_view_out *outs, int steps_per_pass,
               nes_render_stats *stats);

/* Device-resident variant for benchmarks and chaining: outputs stay in the
 * context's device buffers; nes_render_fetch copies view i to the host. */
int nes_render_device(nes_ctx *ctx, int precision, const nes_camera *views,
                      int n_views, int steps_per_pass, nes_render_stats *stats);
int nes_render_fetch(nes_ctx *ctx, int view, const nes_view_out *out);

/* ----------------------------------------------------------- frame prep */
#define NES_SWS_X86 0  /* libswscale 5.5 on x86-64, flags 0 (deployed path) */
#define NES_SWS_C 1    /* libswscale 5.5 portable C output paths */

/* Nearer-depth composite of two renderer outputs then RGB24 -> YUV420P with
 * the arithmetic of sws_scale(flags 0) (src/base/video/type_managers.cc:
 * 143-155).  Host pointers.  rgb_b/depth_b may be NULL (plain conversion of
 * A, which is what RenderedFrame::convert_frame does,
 * include/base/video/rendered_frame.h:24-33).  w and h must be even
 * (src/base/camera_manager.cc:38-44).  A pixel takes B when depth_b is nearer
 * (NaN counts as +inf; ties keep A). */
int nes_frameprep(const uint8_t *rgb_a, const float *depth_a,
                  const uint8_t *rgb_b, const float *depth_b, int w, int h,
                  int rgb_stride, uint8_t *y, int y_stride, uint8_t *u,
                  int u_stride, uint8_t *v, int v_stride, int sws_mode);

/* Same on device pointers, n_frames frames laid out frame_bytes apart in
 * each buffer (rgb: w*h*3, depth: w*h*4, planes: y_stride*h etc.); runs on the
 * library stream and returns after completion.  elapsed_ms (optional)
 * receives the kernel time measured with device events. */
int nes_frameprep_device(const uint8_t *rgb_a, const float *depth_a,
                         const uint8_t *rgb_b, const float *depth_b, int w,
                         int h, int n_frames, uint8_t *y, int y_stride,
                         uint8_t *u, int u_stride, uint8_t *v, int v_stride,
                         int sws_mode, float *elapsed_ms);

/* nes_frameprep_device without device allocation.  This entry point has no
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Its only scratch buffer, the vertical chroma filter of
 * the frame height, lives in a fixed device array of 144 KiB per GPU, held
 * until the process ends, so h is limited to 8192 (NES_ERANGE above).  The
 * array is shared by every call on a GPU and keeps the last call's table:
 * calls that overlap on one GPU race.  The work is enqueued on `stream` (a
 * hipStream_t) and the call returns without waiting; stream NULL uses the
 * null stream and returns after completion.  Same layout and results as
 * nes_frameprep_device. */
int nes_frameprep_device_static(const uint8_t *rgb_a, const float *depth_a,
                                const uint8_t *rgb_b, const float *depth_b, int w,
                                int h, int n_frames, uint8_t *y, int y_stride,
                                uint8_t *u, int u_stride, uint8_t *v, int v_stride,
                                int sws_mode, void *stream);

/* GRAY8 (full range) -> YUV420P, the depth stream of
 * RenderedFrame::convert_frame (include/base/video/rendered_frame.h:30-31). */
int nes_gray8_to_yuv420p(const uint8_t *gray, int w, int h, int g_stride,
                         uint8_t *y, int y_stride, uint8_t *u, int u_stride,
                         uint8_t *v, int v_stride, int sws_mode);

/* ------------------------------------------------------------ text overlay */
/* The frame-prep driver stamps four strings into the scene frame before the
 * YUV conversion (process_frame_thread, src/encode.cpp:57-97) with
 * RenderTextContext::render_string_to_frame (src/base/video/
 * render_text.cc:35-111): every non-zero pixel of a glyph bitmap turns the
 * RGB24 pixel under it white.  Glyph rasterisation stays with the caller's
 * FreeType (one FT_Load_Char(face, c, FT_LOAD_RENDER) per character, done
 * once per font); this library takes the resulting glyph table and does the
 * layout and the stamping on the device. */
typedef struct nes_glyph {
  int32_t left;      /* slot->bitmap_left */
  int32_t top;       /* slot->bitmap_top */
  uint32_t width;    /* slot->bitmap.width (also the row pitch used) */
  uint32_t rows;     /* slot->bitmap.rows */
  int32_t advance;   /* slot->advance.x >> 6 */
  uint32_t offset;   /* byte offset of the rows x width bitmap in `bitmaps` */
} nes_glyph;

#define NES_FONT_GLYPHS 256 /* indexed by the unsigned byte of the character */

/* RenderTextContext::RenderPosition (include/base/video/render_text.h:17-23) */
#define NES_TEXT_LEFT_TOP 0
#define NES_TEXT_LEFT_BOTTOM 1
#define NES_TEXT_RIGHT_TOP 2
#define NES_TEXT_RIGHT_BOTTOM 3
#define NES_TEXT_CENTER 4

typedef struct nes_text {
  int32_t position;  /* NES_TEXT_* */
  const char *text;  /* NUL-terminated; '\n' starts a line 20 px lower */
} nes_text;

/* The four strings of process_frame_thread, as bits of overlay_flags. */
#define NES_OVERLAY_MATRIX 1u    /* camera matrix, centre */
#define NES_OVERLAY_INDEX 2u     /* "index=<n>", left bottom */
#define NES_OVERLAY_TIMESTAMP 4u /* local "%T.mmm", left top */
#define NES_OVERLAY_DIRECTION 8u /* "direction=left|right", right top */
#define NES_OVERLAY_ALL 15u

typedef struct nes_font nes_font;
/* Replaces RenderTextContext's constructor (render_text.cc:10-33): glyphs
 * holds NES_FONT_GLYPHS entries; each bitmap must lie inside bitmaps. */
int nes_font_create(const nes_glyph *glyphs, const uint8_t *bitmaps,
                    size_t bitmap_bytes, nes_font **out);
int nes_font_destroy(nes_font *font);
/* render_string_to_frame for n texts, in place on a device RGB24 frame. */
int nes_overlay_device(nes_font *font, const nes_text *texts, int n_texts,
                       uint8_t *rgb_dev, int w, int h, int rgb_stride);
/* The strings process_frame_thread draws for a frame (src/encode.cpp:57-97),
 * NUL-separated into buf in the order matrix, index, timestamp, direction
 * (only those selected by flags); timestamp may be NULL for the current
 * local time.  Returns the bytes used, or a negative code. */
int nes_overlay_texts(const nes_camera *cam, const char *timestamp,
                      uint32_t flags, char *buf, size_t buf_len);
/* process_frame_thread's per-frame work in one call: the optional
 * nearer-depth composite, the selected strings of `cam` stamped into the
 * scene, then RGB24 -> YUV420P (host buffers, as nes_frameprep). */
int nes_frameprep_overlay(nes_font *font, const nes_camera *cam,
                          const char *timestamp, uint32_t flags,
                          const uint8_t *rgb_a, const float *depth_a,
                          const uint8_t *rgb_b, const float *depth_b, int w,
                          int h, int rgb_stride, uint8_t *y, int y_stride,
                          uint8_t *u, int u_stride, uint8_t *v, int v_stride,
                          int sws_mode);

/* -------------------------------------------------------------- training */
/* Differentiable render step for the same scene representation: L2
 * photometric loss, hand-written backward through compositing, the MLPs
 * (FP32 MFMA, activations recomputed per 16-sample tile rather than stored)
 * and an FP32 atomic scatter-add into the hash-table gradient (summed per
 * table slice in LDS, one slice per workgroup), then Adam.
 * The reference ships no trainer (its renderer loads a trained snapshot);
 * this is the training leg of the render path above. */
typedef struct nes_trainer nes_trainer;

typedef struct nes_train_config {
  uint32_t rays_per_batch; /* 0 = 2^18 */
  uint32_t n_poses;        /* target views in pose_c2w / targets */
  uint32_t width, height;  /* target image size */
  float fov_y_deg;         /* 0 = 60 */
  uint64_t seed;           /* ray sampling: splitmix64 stream 1000 + step */
  float lr, beta1, beta2, eps; /* 0 = 1e-2, 0.9, 0.99, 1e-15 */
} nes_train_config;

typedef struct nes_train_stats {
  double ms;        /* device time of the step */
  uint64_t rays, samples, step;
} nes_train_stats;

/* init: starting parameters and the (fixed) occupancy grid.  pose_c2w:
 * n_poses row-major 3x4 matrices; targets: n_poses x H x W x 3 f32 colours.
 * Both may be NULL for a trainer driven only by explicit rays. */
int nes_trainer_create(const nes_scene_desc *init, const nes_train_config *cfg,
                       const float *pose_c2w, const float *targets,
                       nes_trainer **out);
int nes_trainer_destroy(nes_trainer *trainer);
/* One step: sample the batch of the current step, forward, loss, backward,
 * Adam.  loss receives the batch loss before the update. */
int nes_trainer_step(nes_trainer *trainer, float *loss, nes_train_stats *stats);
/* Loss and gradients of step `step`'s batch without updating (host outputs
 * may be NULL). */
int nes_trainer_batch_grad(nes_trainer *trainer, uint64_t step, float *loss,
                           float *grad_table, float *grad_weights);
int nes_trainer_get_params(nes_trainer *trainer, float *table, float *weights);
int nes_trainer_set_params(nes_trainer *trainer, const float *table,
                           const float *weights);
/* Raw differentiable render of explicit rays (n x 3 origins, unit
 * directions) with the current parameters: rgb is n x 3 ... */
int nes_trainer_forward(nes_trainer *trainer, const float *ray_o,
                        const float *ray_d, uint32_t n, float *rgb);
/* ... and the gradient of sum(d_rgb * rgb) (d_rgb n x 3) for those rays with
 * the current parameters.  The backward recomputes the rays' forward itself,
 * so it depends on no state of earlier calls. */
int nes_trainer_backward(nes_trainer *trainer, const float *ray_o,
                         const float *ray_d, uint32_t n, const float *d_rgb,
                         float *grad_table, float *grad_weights);
/* Checkpoint: parameters (table and weights), Adam moments and the step
 * count in one binary file; load restores them into a trainer of the same
 * scene representation (a mismatched or truncated file changes nothing and
 * returns NES_EINVAL). */
int nes_trainer_save(nes_trainer *trainer, const char *path);
int nes_trainer_load(nes_trainer *trainer, const char *path);

/* --------------------------------------------------------------- serving */
/* Continuous-batching server: replaces the reference's per-connection loop
 * of one blocking FrameRequest in flight (socket_client_thread,
 * src/server.cpp:116-212) plus the per-eye process_frame_thread conversion
 * (src/encode.cpp:42-119).  Clients submit stereo requests from any thread;
 * a scheduler batches waiting requests into free batch slots (one HIP stream
 * and one batch of device buffers each; the HBM budget decides how many
 * slots exist and each slot's share of it), renders, converts every view to
 * YUV420P on the device and copies the planes to host memory.  A batch never
 * grows past its slot's share: when its samples would not fit, it is
 * rendered again in passes that give each ray at most capacity / rays
 * samples each (bit-identical results, more passes).  As in CameraManager
 * (src/base/camera_manager.cc:33-69) only a client's latest pose matters: a
 * newer request supersedes that client's request still waiting for
 * admission (it completes with NES_ECANCELED). */
#define NES_ECANCELED (-125)

typedef struct nes_server nes_server;

typedef struct nes_server_config {
  uint64_t hbm_budget_bytes; /* 0 = 48 GiB */
  uint32_t max_batch_rays;   /* ray slots per batch, 0 = 2^24 */
  uint32_t max_inflight;     /* batch slots, 0 = 3 (renders run one at a time,
                                at most two queued; the other slots hold earlier
                                batches' frame prep and copies) */
  uint32_t queue_capacity;   /* waiting requests, 0 = 256 */
  int precision;             /* NES_PRECISION_* */
  int sws_mode;              /* NES_SWS_* */
  nes_font *font;            /* text overlay font, or NULL for none */
  uint32_t overlay_flags;    /* NES_OVERLAY_* drawn on every view */
  /* 1: also emit each view's depth stream, GRAY8 depth (quantised against
   * depth_far as nes_view_out.depth_gray8) -> YUV420P, the second encoder
   * input per eye of the reference (src/main.cpp:195-217,
   * include/base/video/rendered_frame.h:30-31) */
  uint32_t depth_stream;
  float depth_far;           /* <= 0: NES_DEPTH_FAR_DEFAULT */
} nes_server_config;

/* Caller-owned YUV420P planes of one frame: an encoder frame's data[0..2]
 * and linesize[0..2] (e.g. an AVFrame from av_image_alloc with align 32,
 * src/base/video/type_managers.cc:116-141).  The server copies the device
 * planes straight into them (no staging copy); memory from nes_host_alloc is
 * pinned, so those copies are DMA transfers. */
typedef struct nes_frame_planes {
  uint8_t *data[3];
  int32_t linesize[3];
} nes_frame_planes;

typedef struct nes_frame_result {
  uint64_t ticket;
  uint64_t index;
  uint32_t n_views;
  uint32_t width[2], height[2];
  /* YUV420P planes, valid until nes_server_release: the server's tightly
   * packed planes, or the caller's planes of nes_server_submit_into */
  const uint8_t *y[2];
  const uint8_t *u[2];
  const uint8_t *v[2];
  double latency_ms; /* submit -> planes in host memory */
  double queue_ms;   /* submit -> admitted into a batch */
  uint64_t rays, samples;
  int status;        /* NES_OK, NES_ECANCELED (superseded) or an error */
  int32_t linesize[2][3];      /* of y, u, v per view */
  /* depth stream planes (depth_stream servers; else NULL) */
  const uint8_t *depth_y[2];
  const uint8_t *depth_u[2];
  const uint8_t *depth_v[2];
  int32_t depth_linesize[2][3];
  uint64_t batch;              /* batch the request was rendered in (from 1) */
  uint32_t batch_requests;     /* requests that batch held */
  uint32_t batch_passes;       /* march passes of that batch */
} nes_frame_result;

typedef struct nes_server_stats {
  uint64_t completed, superseded, failed, rays, samples, batches;
  uint32_t slots;
  double p50_ms, p99_ms, mean_ms, max_ms;
  double wall_s;       /* first arrival -> last completion */
  double mrays_per_s;  /* completed rays / wall_s */
  /* Latency including superseded requests: a superseded request is served
   * by the frame that replaced it, so it counts with the time from its own
   * arrival to that frame's completion. */
  double p50_all_ms, p99_all_ms;
  uint64_t composited;      /* samples composited (before T < 1e-4) */
  uint64_t bounded_batches; /* batches rendered in bounded-memory passes */
} nes_server_stats;

int nes_server_create(nes_ctx *ctx, const nes_server_config *cfg, nes_server **out);
int nes_server_destroy(nes_server *server);
/* n_views: 1 or 2 (stereo); even sizes.  NES_EBUSY when the queue is full. */
int nes_server_submit(nes_server *server, uint32_t client, const nes_camera *views,
                      int n_views, uint64_t *ticket);
/* Same, with the frame written into caller-owned planes: planes[v] for view
 * v's scene stream, depth_planes[v] for its depth stream (depth_stream
 * servers; may be NULL otherwise).  Each plane must hold its rows at the
 * given linesize (>= w for Y, >= w/2 for U and V) and stay valid until the
 * request completes. */
int nes_server_submit_into(nes_server *server, uint32_t client, const nes_camera *views,
                           int n_views, const nes_frame_planes *planes,
                           const nes_frame_planes *depth_planes, uint64_t *ticket);
/* hold != 0 stops admission (waiting requests stay queued, batches in flight
 * finish); hold == 0 resumes it.  Lets a caller fill the queue first so the
 * next batch takes every request that fits. */
int nes_server_hold(nes_server *server, int hold);
/* timeout_ms < 0 waits forever; NES_EAGAIN on timeout. */
int nes_server_wait(nes_server *server, uint64_t ticket, int timeout_ms,
                    nes_frame_result *out);
int nes_server_release(nes_server *server, uint64_t ticket);
int nes_server_stats_get(nes_server *server, nes_server_stats *stats);
int nes_server_reset_stats(nes_server *server);

/* Built-in Poisson load generator (benchmarks): n_clients clients at rate_hz
 * each for duration_s, seeded head poses on a radius-2 orbit, FOV uniform in
 * [fov_min, fov_max], per-eye resolution drawn from res_wh (n_res pairs). */
typedef struct nes_sim_config {
  uint32_t n_clients;
  double rate_hz;
  double duration_s;
  uint64_t seed;
  uint32_t n_res;
  const uint32_t *res_wh;
  double fov_min, fov_max;
  int stereo;
} nes_sim_config;
int nes_server_simulate(nes_server *server, const nes_sim_config *sim,
                        nes_server_stats *stats);

/* ------------------------------------------------- neserver wire protocol */
/* A renderer endpoint for an unmodified neserver: it accepts the TCP
 * connections of socket_manage_thread (src/server.cpp:216-283) and answers
 * each length-prefixed FrameRequest (8-byte host-endian length + proto3
 * message, src/server.cpp:23-112) with a RenderedFrame carrying RGB24 and
 * GRAY8 depth (proto/nes.proto:18-25, consumed at src/server.cpp:193-194).
 * Camera may carry an additive float field 5 (vertical FOV); without it the
 * server's default applies.  GRAY8 depth = clamp(floor(255 d / depth_far + 0.5)). */
typedef struct nes_wire_server nes_wire_server;

int nes_wire_decode_request(const uint8_t *msg, size_t len, float default_fov,
                            nes_camera *cam);
int nes_wire_encode_frame(const nes_camera *cam, const uint8_t *rgb24,
                          const uint8_t *gray8, uint8_t *out, size_t cap,
                          size_t *len);
int nes_wire_server_start(nes_ctx *ctx, const char *addr, uint16_t port,
                          int precision, float default_fov, float depth_far,
                          nes_wire_server **out);
int nes_wire_server_port(nes_wire_server *server);
uint64_t nes_wire_server_served(nes_wire_server *server);
int nes_wire_server_stop(nes_wire_server *server);

/* ------------------------------------------------------------ diagnostics */
/* Detailed text of the last error (shared by all threads). */
const char *nes_last_error(void);
/* The vertical chroma filter of an RGB24 WxH -> YUV420P WxH conversion
 * (libswscale initFilter, bicubic, 12-bit taps): coef is h/2 x taps. */
int nes_sws_vchroma_filter(int h, int16_t *coef, int32_t *pos, int max_taps,
                           int *taps);

/* ------------------------------------------------------ device memory */
/* Minimal device-memory helpers so callers (tests, benchmarks) need no other
 * GPU runtime. */
int nes_device_alloc(size_t bytes, void **ptr);
int nes_device_free(void *ptr);
/* Pinned host memory (e.g. for nes_frame_planes the server writes by DMA). */
int nes_host_alloc(size_t bytes, void **ptr);
int nes_host_free(void *ptr);
int nes_memcpy_h2d(void *dst, const void *src, size_t bytes);
int nes_memcpy_d2h(void *dst, const void *src, size_t bytes);
/* Device-to-device copy (a plain copy of known size, used to calibrate
 * memory counters); elapsed_ms optional. */
int nes_device_copy(void *dst, const void *src, size_t bytes, float *elapsed_ms);
/* Memory-hierarchy ceiling probe: random independent 4-byte gathers, one
 * address per lane, each XCD within its own region_bytes (a power of two,
 * 4 KiB .. 64 MiB).  A region inside the vector L1 gives the chip's L1
 * lane-address rate, one inside an XCD's L2 (but not L1) its L2 request
 * rate: lane_loads / elapsed_ms of the timed launch. */
int nes_device_gather_rate(size_t region_bytes, float *elapsed_ms, uint64_t *lane_loads);
/* LDS atomic add probe: every lane of one 1024-thread workgroup per CU adds
 * at random addresses of a 128 KB LDS array; kind 0 FP32, 1 FP64, 2 32-bit
 * integer, 3 FP32 by compare-and-swap, 4 two FP32 by one 64-bit
 * compare-and-swap (one lane op per pair).  lane_ops / elapsed_ms of the
 * timed launch is the chip's rate. */
int nes_device_lds_atomic_rate(int kind, float *elapsed_ms, uint64_t *lane_ops);
int nes_device_info(char *name, size_t name_len, int *cu_count,
                    uint64_t *hbm_bytes);

#ifdef __cplusplus
}
#endif

#endif /* NES_RENDER_H_ */
