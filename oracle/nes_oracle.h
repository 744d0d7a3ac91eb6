/*
 * nes_oracle.h -- CPU oracle for the render-and-frame-prep path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code; only tests/, the smoke check and the bench's CPU-baseline
 * leg use it, and only as the checker.
 *
 * Two halves:
 *   1. Frame prep: nearer-depth composite of two renderer outputs followed by
 *      a restatement of FFmpeg 4.2 / libswscale 5.5 RGB24->YUV420P under
 *      sws_getContext(..., flags = 0) as called from
 *      src/base/video/type_managers.cc:143-155 (via
 *      include/base/video/rendered_frame.h:24-33).  libswscale is not part of
 *      the reference tree and is absent here; the restatement follows its
 *      published C source (utils.c initFilter, input.c rgb24ToY_c /
 *      rgb24ToUV_half_c, swscale.c hScale16To15_c, output.c yuv2plane1_8_c /
 *      yuv2planeX_8_c, x86/swscale.c yuv2yuvX_sse3).
 *   2. Render: scalar FP64 restatement of the multiresolution hash-grid NeRF
 *      formulation (Mueller et al. 2022) that the external instant-ngp
 *      renderer evaluates behind the request/response of src/server.cpp:116-212.
 *      Sample positions are defined in FP32 arithmetic (ray generation, box
 *      clip, step positions, occupancy cell); encoding, MLPs and compositing
 *      are FP64.
 */
#ifndef NES_ORACLE_H_
#define NES_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- counter-based RNG (splitmix64) ---------------- */
uint64_t orc_mix64(uint64_t x);
uint64_t orc_rand_u64(uint64_t seed, uint64_t stream, uint64_t i);
double orc_uniform(uint64_t seed, uint64_t stream, uint64_t i);
double orc_normal(uint64_t seed, uint64_t stream, uint64_t i);

/* ---------------- scene constants ---------------- */
#define ORC_LEVELS 16
#define ORC_FEAT 2
#define ORC_LOG2_T 19
#define ORC_N_PARAMS_TABLE 6098925  /* entries (x2 features) */
#define ORC_N_WEIGHTS 9344
#define ORC_OCC_RES 128
#define ORC_OCC_WORDS 65536
#define ORC_MAX_STEPS 128

int orc_level_res(int level);
uint32_t orc_level_offset(int level);     /* entry offset of the level */
uint32_t orc_level_size(int level);       /* entries in the level */
uint32_t orc_grid_index(int level, uint32_t x, uint32_t y, uint32_t z);

/* Synthetic scene: table (ORC_N_PARAMS_TABLE*2 floats), weights
 * (ORC_N_WEIGHTS floats), occupancy bitfield (ORC_OCC_WORDS words).
 * table_init: 0 = N(0, 0.1), 1 = U(-1e-4, 1e-4).  Any pointer may be NULL. */
void orc_scene_generate(uint64_t seed, int table_init, float *table,
                        float *weights, uint32_t *occ);
double orc_occupancy_fraction(const uint32_t *occ);

/* ---------------- camera ---------------- */
typedef struct {
  uint64_t index;
  uint32_t width, height;
  float c2w[12];      /* row-major 3x4 camera-to-world */
  float fov_y_deg;    /* vertical field of view */
  uint32_t is_left;
} orc_camera;

void orc_orbit_camera(double azimuth_deg, double elevation_deg, double radius,
                      double eye_offset, float c2w[12]);
float orc_focal(uint32_t height, float fov_y_deg);
void orc_camera_ray(const orc_camera *cam, uint32_t px, uint32_t py,
                    float o[3], float d[3]);
/* Returns number of candidate steps (0 if the ray misses the box). */
int orc_ray_box(const float o[3], const float d[3], float *t_start, float *t_end);
float orc_step_t(int k, float t_start);
void orc_step_pos(const float o[3], const float d[3], float t, float x[3]);
int orc_occupied(const uint32_t *occ, const float x[3]);

/* ---------------- network (FP64) ---------------- */
void orc_encode(const float *table, const double x[3], double feat[32]);
void orc_sh4(const double d[3], double sh[16]);
void orc_mlp(const float *weights, const double feat[32], const double sh[16],
             double *sigma, double rgb[3]);

/* ---------------- render (FP64) ---------------- */
typedef struct {
  double rgb[3];
  double depth;
  double transmittance;
  int n_samples;
} orc_ray_result;

void orc_render_ray(const float *table, const float *weights,
                    const uint32_t *occ, const float o[3], const float d[3],
                    orc_ray_result *res);
/* Render n pixels (flat indices py*W+px) of one camera, threads>=1. */
void orc_render_pixels(const float *table, const float *weights,
                       const uint32_t *occ, const orc_camera *cam,
                       const uint32_t *pix, int n, double *rgb, double *depth,
                       int *n_samples, int threads);
uint8_t orc_quantize(double c);

/* ---------------- training (FP64) ---------------- */
/* loss = mean over rays and channels of (C - target)^2.  grad_* may be NULL
 * (accumulated into, not cleared).  grad_table has ORC_N_PARAMS_TABLE*2
 * entries, grad_weights ORC_N_WEIGHTS. */
double orc_loss_grad(const float *table, const float *weights,
                     const uint32_t *occ, int n_rays, const float *ray_o,
                     const float *ray_d, const float *target,
                     double *grad_table, double *grad_weights);
/* Loss with one parameter moved by delta in FP64 (finite differences). */
double orc_loss_perturbed(const float *table, const float *weights,
                          const uint32_t *occ, int n_rays, const float *ray_o,
                          const float *ray_d, const float *target,
                          long table_idx, long weight_idx, double delta);
void orc_adam_step(int n, double *p, const double *g, double *m, double *v,
                   int step, double lr, double beta1, double beta2, double eps);
void orc_training_rays(uint64_t seed, uint64_t step, int n_rays, int n_poses,
                       uint32_t width, uint32_t height, float *ray_o,
                       float *ray_d, uint32_t *pose_idx, uint32_t *pix_idx);
void orc_training_pose(uint64_t seed, int pose, float c2w[12]);

/* ---------------- frame prep: libswscale restatement ---------------- */
#define ORC_SWS_MODE_X86 0  /* FFmpeg 4.2 on x86-64: MMX/SSE3 vertical filter */
#define ORC_SWS_MODE_C 1    /* portable C paths (non-x86 build) */

/* initFilter() restated.  Writes up to dst_w*max_size coefficients; returns
 * the filter size (or -1 when max_size is too small). */
int orc_sws_init_filter(int x_inc, int src_w, int dst_w, int filter_align,
                        int one, int src_pos, int dst_pos, int16_t *coef,
                        int32_t *pos, int max_size);
/* The vertical chroma filter used for an RGB24 WxH -> YUV420P WxH context. */
int orc_sws_vchroma_filter(int h, int16_t *coef, int32_t *pos, int max_size);

/* RGB24 (packed, stride rgb_stride) -> YUV420P planes. */
int orc_rgb24_to_yuv420p(const uint8_t *rgb, int rgb_stride, int w, int h,
                         uint8_t *y, int y_stride, uint8_t *u, int u_stride,
                         uint8_t *v, int v_stride, int mode);
/* GRAY8 depth plane of a RenderedFrame from FP32 expected depth:
 * g = clamp(floor(fma(d, 255 / far, 0.5)), 0, 255) in FP32, NaN -> 0.  The
 * reference's external renderer quantised depth itself (proto/nes.proto:24
 * carries the bytes); this is the rule the library defines. */
void orc_depth_gray8(const float *depth, int n, float far, uint8_t *out);
/* GRAY8 (full range) -> YUV420P (limited range), the depth stream. */
int orc_gray8_to_yuv420p(const uint8_t *g, int g_stride, int w, int h,
                         uint8_t *y, int y_stride, uint8_t *u, int u_stride,
                         uint8_t *v, int v_stride, int mode);
/* Nearer-depth composite: pick A where key(dA) <= key(dB), NaN = +inf.
 * rgb_b/depth_b may be NULL (single input). */
void orc_composite(const uint8_t *rgb_a, const float *depth_a,
                   const uint8_t *rgb_b, const float *depth_b, int w, int h,
                   int rgb_stride, uint8_t *rgb_out, float *depth_out);
/* Threads used by the frame-prep row loops (default 1); outputs do not
 * depend on it. */
void orc_set_frameprep_threads(int n);
int orc_frameprep(const uint8_t *rgb_a, const float *depth_a,
                  const uint8_t *rgb_b, const float *depth_b, int w, int h,
                  uint8_t *y, int y_stride, uint8_t *u, int u_stride,
                  uint8_t *v, int v_stride, int mode);

/* Smallest |pre-activation| of the ReLU layers over one ray's evaluated
 * samples (a kink margin: FP32 and FP64 gradients may differ by a whole
 * unit's term when it is below FP32 rounding). */
double orc_ray_relu_margin(const float *table, const float *weights,
                           const uint32_t *occ, const float o[3], const float d[3]);

/* Text overlay: RenderTextContext::render_string_to_frame
 * (src/base/video/render_text.cc:35-111) on a tightly packed RGB24 frame,
 * with each character's FreeType glyph given as a table entry (bitmap_left,
 * bitmap_top, bitmap.width, bitmap.rows, advance.x >> 6, atlas offset). */
typedef struct orc_glyph {
  int32_t left, top;
  uint32_t width, rows;
  int32_t advance;
  uint32_t offset;
} orc_glyph;
void orc_render_string(uint8_t *surface, uint32_t width, uint32_t height,
                       const orc_glyph *glyphs, const uint8_t *bitmaps,
                       int position, const char *content);

#ifdef __cplusplus
}
#endif

#endif
