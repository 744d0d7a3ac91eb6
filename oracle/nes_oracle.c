/*
 * nes_oracle.c -- CPU oracle (test infrastructure only; see nes_oracle.h).
 *
 * Parity anchors:
 *  - frame prep follows the call chain src/encode.cpp:98 ->
 *    include/base/video/rendered_frame.h:24-33 ->
 *    src/base/video/type_managers.cc:143-155 (sws_getContext flags 0 +
 *    sws_scale of the whole frame) with AV_PIX_FMT_RGB24 sources
 *    (src/server.cpp:193-194) and AV_PIX_FMT_YUV420P destinations
 *    (src/main.cpp:195-217); the libswscale internals restated below are those
 *    of FFmpeg 4.2 (libswscale 5.5), the release the reference's Ubuntu 20.04
 *    dependency list (README.md:23-37) resolves to.
 *  - render: the request/response boundary of src/server.cpp:161-194 and the
 *    camera matrix layout of proto/nes.proto:5-10 /
 *    include/base/camera_manager.h:18-19 (row-major 3x4 camera-to-world,
 *    columns right / down / forward / position).
 */
#include "nes_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ======================================================================== */
/* RNG                                                                       */
/* ======================================================================== */

uint64_t orc_mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

uint64_t orc_rand_u64(uint64_t seed, uint64_t stream, uint64_t i) {
  return orc_mix64(orc_mix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) + i);
}

double orc_uniform(uint64_t seed, uint64_t stream, uint64_t i) {
  return (double)(orc_rand_u64(seed, stream, i) >> 11) * 0x1.0p-53;
}

double orc_normal(uint64_t seed, uint64_t stream, uint64_t i) {
  double u1 = orc_uniform(seed, stream, 2 * i);
  double u2 = orc_uniform(seed, stream, 2 * i + 1);
  double r = sqrt(-2.0 * log(1.0 - u1));
  return r * cos(6.283185307179586 * u2);
}

/* ======================================================================== */
/* Hash grid layout                                                          */
/* ======================================================================== */

/* N_l = floor(16 * b^l), b = (2048/16)^(1/15) = 2^(7/15). */
static const int kRes[ORC_LEVELS] = {16,  22,  30,  42,  58,   80,   111,  153,
                                     212, 294, 406, 561, 776, 1072, 1482, 2048};

int orc_level_res(int level) { return kRes[level]; }

static int level_dense(int level) {
  uint64_t n1 = (uint64_t)kRes[level] + 1;
  return n1 * n1 * n1 <= (1ull << ORC_LOG2_T);
}

uint32_t orc_level_size(int level) {
  uint64_t n1 = (uint64_t)kRes[level] + 1;
  return level_dense(level) ? (uint32_t)(n1 * n1 * n1) : (1u << ORC_LOG2_T);
}

uint32_t orc_level_offset(int level) {
  uint32_t off = 0;
  for (int l = 0; l < level; ++l) off += orc_level_size(l);
  return off;
}

uint32_t orc_grid_index(int level, uint32_t x, uint32_t y, uint32_t z) {
  if (level_dense(level)) {
    uint32_t n1 = (uint32_t)kRes[level] + 1;
    return x + y * n1 + z * n1 * n1;
  }
  uint32_t h = (x * 1u) ^ (y * 2654435761u) ^ (z * 805459861u);
  return h & ((1u << ORC_LOG2_T) - 1u);
}

/* ======================================================================== */
/* Scene                                                                     */
/* ======================================================================== */

enum {
  STREAM_TABLE = 1,
  STREAM_OCC = 2,
  STREAM_BLOB = 3,
  STREAM_WEIGHT0 = 10, /* + layer (1..5) */
  STREAM_POSE = 20,
  STREAM_RAYS = 1000   /* + training step */
};

static const int kLayerIn[5] = {32, 64, 31, 64, 64};
static const int kLayerOut[5] = {64, 16, 64, 64, 3};
static const int kLayerOff[5] = {0, 2048, 3072, 5056, 9152};

#define BLOB_SIGMA 0.2
#define BLOB_TAU 0.115

void orc_scene_generate(uint64_t seed, int table_init, float *table,
                        float *weights, uint32_t *occ) {
  if (table) {
    const uint64_t n = (uint64_t)ORC_N_PARAMS_TABLE * ORC_FEAT;
    for (uint64_t j = 0; j < n; ++j) {
      if (table_init == 0)
        table[j] = (float)(0.1 * orc_normal(seed, STREAM_TABLE, j));
      else
        table[j] = (float)((2.0 * orc_uniform(seed, STREAM_TABLE, j) - 1.0) * 1e-4);
    }
  }
  if (weights) {
    for (int l = 0; l < 5; ++l) {
      double a = sqrt(6.0 / (double)(kLayerIn[l] + kLayerOut[l]));
      int n = kLayerIn[l] * kLayerOut[l];
      for (int j = 0; j < n; ++j)
        weights[kLayerOff[l] + j] =
            (float)((2.0 * orc_uniform(seed, STREAM_WEIGHT0 + l + 1, j) - 1.0) * a);
    }
  }
  if (occ) {
    double c[3];
    for (int a = 0; a < 3; ++a)
      c[a] = 0.5 + 0.1 * (orc_uniform(seed, STREAM_BLOB, a) - 0.5);
    memset(occ, 0, sizeof(uint32_t) * ORC_OCC_WORDS);
    for (int k = 0; k < ORC_OCC_RES; ++k)
      for (int j = 0; j < ORC_OCC_RES; ++j)
        for (int i = 0; i < ORC_OCC_RES; ++i) {
          uint32_t cell = (uint32_t)(i + ORC_OCC_RES * (j + ORC_OCC_RES * k));
          double x = (i + 0.5) / ORC_OCC_RES - c[0];
          double y = (j + 0.5) / ORC_OCC_RES - c[1];
          double z = (k + 0.5) / ORC_OCC_RES - c[2];
          double rho = exp(-(x * x + y * y + z * z) / (2.0 * BLOB_SIGMA * BLOB_SIGMA));
          double u = orc_uniform(seed, STREAM_OCC, cell);
          if (rho * (0.75 + 0.5 * u) > BLOB_TAU) occ[cell >> 5] |= 1u << (cell & 31);
        }
  }
}

double orc_occupancy_fraction(const uint32_t *occ) {
  uint64_t n = 0;
  for (int i = 0; i < ORC_OCC_WORDS; ++i) n += (uint64_t)__builtin_popcount(occ[i]);
  return (double)n / (double)(ORC_OCC_RES * ORC_OCC_RES * ORC_OCC_RES);
}

/* ======================================================================== */
/* Camera and march geometry (FP32 by definition)                            */
/* ======================================================================== */

void orc_orbit_camera(double az_deg, double el_deg, double radius,
                      double eye_offset, float c2w[12]) {
  const double pi = 3.14159265358979323846;
  double az = az_deg * pi / 180.0, el = el_deg * pi / 180.0;
  double p[3] = {0.5 + radius * cos(el) * sin(az), 0.5 + radius * sin(el),
                 0.5 + radius * cos(el) * cos(az)};
  double f[3] = {0.5 - p[0], 0.5 - p[1], 0.5 - p[2]};
  double fn = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  for (int i = 0; i < 3; ++i) f[i] /= fn;
  /* right = forward x up(0,1,0) */
  double r[3] = {-f[2], 0.0, f[0]};
  double rn = sqrt(r[0] * r[0] + r[2] * r[2]);
  for (int i = 0; i < 3; ++i) r[i] /= rn;
  /* down = forward x right */
  double dn[3] = {f[1] * r[2] - f[2] * r[1], f[2] * r[0] - f[0] * r[2],
                  f[0] * r[1] - f[1] * r[0]};
  for (int i = 0; i < 3; ++i) {
    c2w[4 * i + 0] = (float)r[i];
    c2w[4 * i + 1] = (float)dn[i];
    c2w[4 * i + 2] = (float)f[i];
    c2w[4 * i + 3] = (float)(p[i] + eye_offset * r[i]);
  }
}

float orc_focal(uint32_t height, float fov_y_deg) {
  return (float)(0.5 * (double)height /
                 tan(0.5 * (double)fov_y_deg * 3.14159265358979323846 / 180.0));
}

void orc_camera_ray(const orc_camera *cam, uint32_t px, uint32_t py, float o[3],
                    float d[3]) {
  const float f = orc_focal(cam->height, cam->fov_y_deg);
  const float u = (((float)px + 0.5f) - 0.5f * (float)cam->width) / f;
  const float v = (((float)py + 0.5f) - 0.5f * (float)cam->height) / f;
  const float *m = cam->c2w;
  float w[3];
  for (int i = 0; i < 3; ++i)
    w[i] = fmaf(m[4 * i + 0], u, fmaf(m[4 * i + 1], v, m[4 * i + 2]));
  const float n = sqrtf(fmaf(w[0], w[0], fmaf(w[1], w[1], w[2] * w[2])));
  for (int i = 0; i < 3; ++i) {
    d[i] = w[i] / n;
    o[i] = m[4 * i + 3];
  }
}

#define ORC_DELTA 0.0135316469f /* (float)(sqrt(3)/128) */

int orc_ray_box(const float o[3], const float d[3], float *t_start,
                float *t_end) {
  float tn = -INFINITY, tf = INFINITY;
  for (int a = 0; a < 3; ++a) {
    if (d[a] == 0.0f) {
      if (o[a] < 0.0f || o[a] > 1.0f) return 0;
      continue;
    }
    float t0 = (0.0f - o[a]) / d[a];
    float t1 = (1.0f - o[a]) / d[a];
    float lo = t0 < t1 ? t0 : t1;
    float hi = t0 < t1 ? t1 : t0;
    tn = lo > tn ? lo : tn;
    tf = hi < tf ? hi : tf;
  }
  float ts = tn > 0.0f ? tn : 0.0f;
  *t_start = ts;
  *t_end = tf;
  if (!(tf > ts)) return 0;
  int n = 0;
  while (n < ORC_MAX_STEPS && orc_step_t(n, ts) < tf) ++n;
  return n;
}

float orc_step_t(int k, float t_start) {
  return fmaf((float)k + 0.5f, ORC_DELTA, t_start);
}

void orc_step_pos(const float o[3], const float d[3], float t, float x[3]) {
  for (int a = 0; a < 3; ++a) x[a] = fmaf(t, d[a], o[a]);
}

int orc_occupied(const uint32_t *occ, const float x[3]) {
  int c[3];
  for (int a = 0; a < 3; ++a) {
    int ci = (int)floorf(x[a] * 128.0f);
    c[a] = ci < 0 ? 0 : (ci > ORC_OCC_RES - 1 ? ORC_OCC_RES - 1 : ci);
  }
  uint32_t cell = (uint32_t)(c[0] + ORC_OCC_RES * (c[1] + ORC_OCC_RES * c[2]));
  return (occ[cell >> 5] >> (cell & 31)) & 1u;
}

/* ======================================================================== */
/* Network (FP64) with optional single-parameter perturbation                */
/* ======================================================================== */

typedef struct {
  const float *table;
  const float *weights;
  long pert_table; /* flat table float index or -1 */
  long pert_weight;
  double pert_delta;
} params_t;

static inline double tab(const params_t *p, uint64_t i) {
  double v = (double)p->table[i];
  if ((long)i == p->pert_table) v += p->pert_delta;
  return v;
}

static inline double wgt(const params_t *p, int i) {
  double v = (double)p->weights[i];
  if (i == p->pert_weight) v += p->pert_delta;
  return v;
}

typedef struct {
  uint32_t entry[8]; /* global entry index */
  double w[8];
} level_corners;

static void corners(int level, const double x[3], level_corners *lc) {
  const int n = kRes[level];
  uint32_t i0[3];
  double f[3];
  for (int a = 0; a < 3; ++a) {
    double xc = x[a] < 0.0 ? 0.0 : (x[a] > 1.0 ? 1.0 : x[a]);
    double pos = xc * (double)n;
    double fl = floor(pos);
    if (fl > n - 1) fl = n - 1;
    i0[a] = (uint32_t)fl;
    f[a] = pos - fl;
  }
  const uint32_t off = orc_level_offset(level);
  for (int c = 0; c < 8; ++c) {
    uint32_t b[3] = {(uint32_t)(c & 1), (uint32_t)((c >> 1) & 1), (uint32_t)((c >> 2) & 1)};
    double w = 1.0;
    for (int a = 0; a < 3; ++a) w *= b[a] ? f[a] : 1.0 - f[a];
    lc->entry[c] = off + orc_grid_index(level, i0[0] + b[0], i0[1] + b[1], i0[2] + b[2]);
    lc->w[c] = w;
  }
}

static void encode_p(const params_t *p, const double x[3], double feat[32]) {
  for (int l = 0; l < ORC_LEVELS; ++l) {
    level_corners lc;
    corners(l, x, &lc);
    double f0 = 0.0, f1 = 0.0;
    for (int c = 0; c < 8; ++c) {
      f0 += lc.w[c] * tab(p, 2ull * lc.entry[c] + 0);
      f1 += lc.w[c] * tab(p, 2ull * lc.entry[c] + 1);
    }
    feat[2 * l + 0] = f0;
    feat[2 * l + 1] = f1;
  }
}

void orc_encode(const float *table, const double x[3], double feat[32]) {
  params_t p = {table, NULL, -1, -1, 0.0};
  encode_p(&p, x, feat);
}

void orc_sh4(const double d[3], double sh[16]) {
  const double x = d[0], y = d[1], z = d[2];
  const double xx = x * x, yy = y * y, zz = z * z;
  sh[0] = 0.28209479177387814;
  sh[1] = -0.48860251190291987 * y;
  sh[2] = 0.48860251190291987 * z;
  sh[3] = -0.48860251190291987 * x;
  sh[4] = 1.0925484305920792 * x * y;
  sh[5] = -1.0925484305920792 * y * z;
  sh[6] = 0.94617469575755997 * zz - 0.31539156525251999;
  sh[7] = -1.0925484305920792 * x * z;
  sh[8] = 0.54627421529603959 * xx - 0.54627421529603959 * yy;
  sh[9] = 0.59004358992664352 * y * (-3.0 * xx + yy);
  sh[10] = 2.8906114426405538 * x * y * z;
  sh[11] = 0.45704579946446572 * y * (1.0 - 5.0 * zz);
  sh[12] = 0.3731763325901154 * z * (5.0 * zz - 3.0);
  sh[13] = 0.45704579946446572 * x * (1.0 - 5.0 * zz);
  sh[14] = 1.4453057213202769 * z * (xx - yy);
  sh[15] = 0.59004358992664352 * x * (-xx + 3.0 * yy);
}

typedef struct {
  double feat[32], h1p[64], out2[16], in3[31], h3p[64], h4p[64], out5[3];
  double sigma, rgb[3];
} mlp_trace;

static void dense(const params_t *p, int layer, const double *in, double *out) {
  const int ni = kLayerIn[layer], no = kLayerOut[layer], off = kLayerOff[layer];
  for (int o = 0; o < no; ++o) {
    double acc = 0.0;
    for (int i = 0; i < ni; ++i) acc += wgt(p, off + o * ni + i) * in[i];
    out[o] = acc;
  }
}

static inline double relu(double v) { return v > 0.0 ? v : 0.0; }

static void mlp_p(const params_t *p, const double feat[32], const double sh[16],
                  mlp_trace *t) {
  double h[64];
  memcpy(t->feat, feat, sizeof(t->feat));
  dense(p, 0, feat, t->h1p);
  for (int i = 0; i < 64; ++i) h[i] = relu(t->h1p[i]);
  dense(p, 1, h, t->out2);
  t->sigma = exp(t->out2[0]);
  for (int i = 0; i < 16; ++i) t->in3[i] = sh[i];
  for (int i = 0; i < 15; ++i) t->in3[16 + i] = t->out2[1 + i];
  dense(p, 2, t->in3, t->h3p);
  for (int i = 0; i < 64; ++i) h[i] = relu(t->h3p[i]);
  dense(p, 3, h, t->h4p);
  for (int i = 0; i < 64; ++i) h[i] = relu(t->h4p[i]);
  dense(p, 4, h, t->out5);
  for (int i = 0; i < 3; ++i) t->rgb[i] = 1.0 / (1.0 + exp(-t->out5[i]));
}

void orc_mlp(const float *weights, const double feat[32], const double sh[16],
             double *sigma, double rgb[3]) {
  params_t p = {NULL, weights, -1, -1, 0.0};
  mlp_trace t;
  mlp_p(&p, feat, sh, &t);
  *sigma = t.sigma;
  for (int i = 0; i < 3; ++i) rgb[i] = t.rgb[i];
}

/* ======================================================================== */
/* Render                                                                    */
/* ======================================================================== */

#define ORC_T_STOP 1e-4

typedef struct {
  float t;
  float x[3];
} sample_pos;

/* Occupied sample positions of one ray, in march order (<= 128). */
static int ray_samples(const uint32_t *occ, const float o[3], const float d[3],
                       sample_pos *s) {
  float ts, te;
  int n_steps = orc_ray_box(o, d, &ts, &te);
  int n = 0;
  for (int k = 0; k < n_steps; ++k) {
    float t = orc_step_t(k, ts);
    float x[3];
    orc_step_pos(o, d, t, x);
    if (!orc_occupied(occ, x)) continue;
    s[n].t = t;
    memcpy(s[n].x, x, sizeof(x));
    ++n;
  }
  return n;
}

static void render_ray_p(const params_t *p, const uint32_t *occ,
                         const float o[3], const float d[3],
                         orc_ray_result *res) {
  sample_pos s[ORC_MAX_STEPS];
  int n = ray_samples(occ, o, d, s);
  double dd[3] = {d[0], d[1], d[2]}, sh[16];
  orc_sh4(dd, sh);
  double T = 1.0, C[3] = {0, 0, 0}, D = 0.0;
  int used = 0;
  for (int i = 0; i < n; ++i) {
    double x[3] = {s[i].x[0], s[i].x[1], s[i].x[2]}, feat[32];
    mlp_trace tr;
    encode_p(p, x, feat);
    mlp_p(p, feat, sh, &tr);
    double a = -expm1(-tr.sigma * (double)ORC_DELTA);
    double w = T * a;
    for (int c = 0; c < 3; ++c) C[c] += w * tr.rgb[c];
    D += w * (double)s[i].t;
    T *= exp(-tr.sigma * (double)ORC_DELTA);
    ++used;
    if (T < ORC_T_STOP) break;
  }
  for (int c = 0; c < 3; ++c) res->rgb[c] = C[c];
  res->depth = D;
  res->transmittance = T;
  res->n_samples = used;
}

void orc_render_ray(const float *table, const float *weights,
                    const uint32_t *occ, const float o[3], const float d[3],
                    orc_ray_result *res) {
  params_t p = {table, weights, -1, -1, 0.0};
  render_ray_p(&p, occ, o, d, res);
}

/* Smallest |pre-activation| of the three ReLU layers over the samples one
 * ray evaluates (up to its termination).  An FP32 evaluation can take the
 * other side of a ReLU kink when this margin is below its rounding error, and
 * then its gradient differs from the FP64 one by that unit's whole term. */
double orc_ray_relu_margin(const float *table, const float *weights,
                           const uint32_t *occ, const float o[3], const float d[3]) {
  params_t p = {table, weights, -1, -1, 0.0};
  sample_pos s[ORC_MAX_STEPS];
  int n = ray_samples(occ, o, d, s);
  double dd[3] = {d[0], d[1], d[2]}, sh[16];
  orc_sh4(dd, sh);
  double T = 1.0, margin = INFINITY;
  for (int i = 0; i < n; ++i) {
    double x[3] = {s[i].x[0], s[i].x[1], s[i].x[2]}, feat[32];
    mlp_trace tr;
    encode_p(&p, x, feat);
    mlp_p(&p, feat, sh, &tr);
    for (int u = 0; u < 64; ++u) {
      margin = fmin(margin, fabs(tr.h1p[u]));
      margin = fmin(margin, fabs(tr.h3p[u]));
      margin = fmin(margin, fabs(tr.h4p[u]));
    }
    T *= exp(-tr.sigma * (double)ORC_DELTA);
    if (T < ORC_T_STOP) break;
  }
  return margin;
}

void orc_render_pixels(const float *table, const float *weights,
                       const uint32_t *occ, const orc_camera *cam,
                       const uint32_t *pix, int n, double *rgb, double *depth,
                       int *n_samples, int threads) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1)
#endif
  for (int i = 0; i < n; ++i) {
    uint32_t px = pix[i] % cam->width, py = pix[i] / cam->width;
    float o[3], d[3];
    orc_camera_ray(cam, px, py, o, d);
    orc_ray_result r;
    orc_render_ray(table, weights, occ, o, d, &r);
    for (int c = 0; c < 3; ++c) rgb[3 * i + c] = r.rgb[c];
    if (depth) depth[i] = r.depth;
    if (n_samples) n_samples[i] = r.n_samples;
  }
  (void)threads;
}

void orc_depth_gray8(const float *depth, int n, float far, uint8_t *out) {
  const float scale = 255.0f / far;
  for (int i = 0; i < n; ++i) {
    float q = floorf(fmaf(depth[i], scale, 0.5f));
    if (!(q >= 0.0f)) q = 0.0f; /* negative or NaN */
    if (q > 255.0f) q = 255.0f;
    out[i] = (uint8_t)q;
  }
}

uint8_t orc_quantize(double c) {
  double v = floor(c * 255.0 + 0.5);
  if (v < 0.0) v = 0.0;
  if (v > 255.0) v = 255.0;
  return (uint8_t)v;
}

/* ======================================================================== */
/* Training: loss, analytic gradient, Adam                                   */
/* ======================================================================== */

void orc_training_pose(uint64_t seed, int pose, float c2w[12]) {
  double az = 360.0 * orc_uniform(seed, STREAM_POSE, 2 * (uint64_t)pose);
  double el = 5.0 + 30.0 * orc_uniform(seed, STREAM_POSE, 2 * (uint64_t)pose + 1);
  orc_orbit_camera(az, el, 2.0, 0.0, c2w);
}

void orc_training_rays(uint64_t seed, uint64_t step, int n_rays, int n_poses,
                       uint32_t width, uint32_t height, float *ray_o,
                       float *ray_d, uint32_t *pose_idx, uint32_t *pix_idx) {
  for (int i = 0; i < n_rays; ++i) {
    uint64_t u = orc_rand_u64(seed, STREAM_RAYS + step, (uint64_t)i);
    uint32_t pose = (uint32_t)((u & 0xffffffffull) % (uint64_t)n_poses);
    uint32_t pix = (uint32_t)((u >> 32) % ((uint64_t)width * height));
    orc_camera cam;
    memset(&cam, 0, sizeof(cam));
    cam.width = width;
    cam.height = height;
    cam.fov_y_deg = 60.0f;
    orc_training_pose(seed, (int)pose, cam.c2w);
    orc_camera_ray(&cam, pix % width, pix / width, ray_o + 3 * i, ray_d + 3 * i);
    if (pose_idx) pose_idx[i] = pose;
    if (pix_idx) pix_idx[i] = pix;
  }
}

static void accumulate_grad(const params_t *p, const mlp_trace *tr,
                            const double x[3], double dsigma,
                            const double drgb[3], double *gtab, double *gw) {
  /* sigmoid */
  double d5[3], dh4[64], dh3[64], din3[31], dout2[16], dh1[64], dfeat[32];
  for (int c = 0; c < 3; ++c) d5[c] = drgb[c] * tr->rgb[c] * (1.0 - tr->rgb[c]);
  double h4[64], h3[64], h1[64];
  for (int i = 0; i < 64; ++i) {
    h4[i] = relu(tr->h4p[i]);
    h3[i] = relu(tr->h3p[i]);
    h1[i] = relu(tr->h1p[i]);
  }
  /* layer 5: 64 -> 3 */
  for (int i = 0; i < 64; ++i) dh4[i] = 0.0;
  for (int o = 0; o < 3; ++o)
    for (int i = 0; i < 64; ++i) {
      if (gw) gw[kLayerOff[4] + o * 64 + i] += d5[o] * h4[i];
      dh4[i] += wgt(p, kLayerOff[4] + o * 64 + i) * d5[o];
    }
  for (int i = 0; i < 64; ++i) dh4[i] = tr->h4p[i] > 0.0 ? dh4[i] : 0.0;
  /* layer 4: 64 -> 64 */
  for (int i = 0; i < 64; ++i) dh3[i] = 0.0;
  for (int o = 0; o < 64; ++o)
    for (int i = 0; i < 64; ++i) {
      if (gw) gw[kLayerOff[3] + o * 64 + i] += dh4[o] * h3[i];
      dh3[i] += wgt(p, kLayerOff[3] + o * 64 + i) * dh4[o];
    }
  for (int i = 0; i < 64; ++i) dh3[i] = tr->h3p[i] > 0.0 ? dh3[i] : 0.0;
  /* layer 3: 31 -> 64 */
  for (int i = 0; i < 31; ++i) din3[i] = 0.0;
  for (int o = 0; o < 64; ++o)
    for (int i = 0; i < 31; ++i) {
      if (gw) gw[kLayerOff[2] + o * 31 + i] += dh3[o] * tr->in3[i];
      din3[i] += wgt(p, kLayerOff[2] + o * 31 + i) * dh3[o];
    }
  /* density output: out2[0] -> sigma = exp, out2[1..15] -> geo */
  dout2[0] = dsigma * tr->sigma;
  for (int i = 0; i < 15; ++i) dout2[1 + i] = din3[16 + i];
  /* layer 2: 64 -> 16 */
  for (int i = 0; i < 64; ++i) dh1[i] = 0.0;
  for (int o = 0; o < 16; ++o)
    for (int i = 0; i < 64; ++i) {
      if (gw) gw[kLayerOff[1] + o * 64 + i] += dout2[o] * h1[i];
      dh1[i] += wgt(p, kLayerOff[1] + o * 64 + i) * dout2[o];
    }
  for (int i = 0; i < 64; ++i) dh1[i] = tr->h1p[i] > 0.0 ? dh1[i] : 0.0;
  /* layer 1: 32 -> 64 */
  for (int i = 0; i < 32; ++i) dfeat[i] = 0.0;
  for (int o = 0; o < 64; ++o)
    for (int i = 0; i < 32; ++i) {
      if (gw) gw[kLayerOff[0] + o * 32 + i] += dh1[o] * tr->feat[i];
      dfeat[i] += wgt(p, kLayerOff[0] + o * 32 + i) * dh1[o];
    }
  /* hash grid scatter */
  if (gtab) {
    for (int l = 0; l < ORC_LEVELS; ++l) {
      level_corners lc;
      corners(l, x, &lc);
      for (int c = 0; c < 8; ++c) {
        gtab[2ull * lc.entry[c] + 0] += lc.w[c] * dfeat[2 * l + 0];
        gtab[2ull * lc.entry[c] + 1] += lc.w[c] * dfeat[2 * l + 1];
      }
    }
  }
}

static double loss_grad_p(const params_t *p, const uint32_t *occ, int n_rays,
                          const float *ray_o, const float *ray_d,
                          const float *target, double *gtab, double *gw) {
  double loss = 0.0;
  const double norm = 1.0 / (3.0 * (double)n_rays);
  for (int r = 0; r < n_rays; ++r) {
    const float *o = ray_o + 3 * r, *d = ray_d + 3 * r;
    sample_pos s[ORC_MAX_STEPS];
    mlp_trace tr[ORC_MAX_STEPS];
    double xs[ORC_MAX_STEPS][3];
    double Ts[ORC_MAX_STEPS + 1], ws[ORC_MAX_STEPS];
    int n = ray_samples(occ, o, d, s);
    double dd[3] = {d[0], d[1], d[2]}, sh[16];
    orc_sh4(dd, sh);
    double T = 1.0, C[3] = {0, 0, 0};
    int used = 0;
    for (int i = 0; i < n; ++i) {
      double feat[32];
      for (int a = 0; a < 3; ++a) xs[i][a] = s[i].x[a];
      encode_p(p, xs[i], feat);
      mlp_p(p, feat, sh, &tr[i]);
      double a = -expm1(-tr[i].sigma * (double)ORC_DELTA);
      Ts[i] = T;
      ws[i] = T * a;
      for (int c = 0; c < 3; ++c) C[c] += ws[i] * tr[i].rgb[c];
      T *= exp(-tr[i].sigma * (double)ORC_DELTA);
      Ts[i + 1] = T;
      ++used;
      if (T < ORC_T_STOP) break;
    }
    double g[3];
    for (int c = 0; c < 3; ++c) {
      double e = C[c] - (double)target[3 * r + c];
      loss += e * e * norm;
      g[c] = 2.0 * e * norm;
    }
    if (!gtab && !gw) continue;
    /* suffix sums S_i = sum_{j>i} w_j c_j, walking backwards */
    double S[3] = {0, 0, 0};
    for (int i = used - 1; i >= 0; --i) {
      double drgb[3], cg = 0.0, sg = 0.0;
      for (int c = 0; c < 3; ++c) {
        drgb[c] = ws[i] * g[c];
        cg += tr[i].rgb[c] * g[c];
        sg += S[c] * g[c];
      }
      double dsigma = (double)ORC_DELTA * (Ts[i + 1] * cg - sg);
      accumulate_grad(p, &tr[i], xs[i], dsigma, drgb, gtab, gw);
      for (int c = 0; c < 3; ++c) S[c] += ws[i] * tr[i].rgb[c];
    }
  }
  return loss;
}

double orc_loss_grad(const float *table, const float *weights,
                     const uint32_t *occ, int n_rays, const float *ray_o,
                     const float *ray_d, const float *target,
                     double *grad_table, double *grad_weights) {
  params_t p = {table, weights, -1, -1, 0.0};
  return loss_grad_p(&p, occ, n_rays, ray_o, ray_d, target, grad_table,
                     grad_weights);
}

double orc_loss_perturbed(const float *table, const float *weights,
                          const uint32_t *occ, int n_rays, const float *ray_o,
                          const float *ray_d, const float *target,
                          long table_idx, long weight_idx, double delta) {
  params_t p = {table, weights, table_idx, weight_idx, delta};
  return loss_grad_p(&p, occ, n_rays, ray_o, ray_d, target, NULL, NULL);
}

void orc_adam_step(int n, double *p, const double *g, double *m, double *v,
                   int step, double lr, double beta1, double beta2, double eps) {
  const double c1 = 1.0 - pow(beta1, step), c2 = 1.0 - pow(beta2, step);
  for (int i = 0; i < n; ++i) {
    m[i] = beta1 * m[i] + (1.0 - beta1) * g[i];
    v[i] = beta2 * v[i] + (1.0 - beta2) * g[i] * g[i];
    double mh = m[i] / c1, vh = v[i] / c2;
    p[i] -= lr * mh / (sqrt(vh) + eps);
  }
}

/* ======================================================================== */
/* Frame prep: libswscale restatement                                       */
/* ======================================================================== */

/* BT.601 limited-range coefficients of libswscale's default rgb2yuv table
 * (swscale_internal.h RY..BV with RGB2YUV_SHIFT 15, selected by
 * fill_rgb2yuv_table for the default colorspace). */
#define RGB2YUV_SHIFT 15
#define SWS_RY 8414
#define SWS_GY 16519
#define SWS_BY 3208
#define SWS_RU (-4865)
#define SWS_GU (-9528)
#define SWS_BU 14392
#define SWS_RV 14392
#define SWS_GV (-12061)
#define SWS_BV (-2332)

static int64_t rounded_div(int64_t a, int64_t b) {
  return (a >= 0 ? a + (b >> 1) : a - (b >> 1)) / b;
}

static int ilog2(unsigned v) {
  int n = 0;
  while (v >>= 1) ++n;
  return n;
}

/* utils.c initFilter() for the bicubic (B = 0, C = 0.6) default scaler. */
int orc_sws_init_filter(int x_inc, int src_w, int dst_w, int filter_align,
                        int one, int src_pos, int dst_pos, int16_t *coef,
                        int32_t *pos, int max_size) {
  int sdiv = src_w / dst_w;
  int lg = ilog2((unsigned)(sdiv > 0 ? sdiv : 1));
  const int64_t fone = 1ll << (54 - (lg < 8 ? lg : 8));
  int filter_size;
  int64_t *filter;
  int32_t *fpos = (int32_t *)calloc((size_t)dst_w + 3, sizeof(int32_t));

  if (llabs((long long)x_inc - 0x10000) < 10 && src_pos == dst_pos) {
    filter_size = 1;
    filter = (int64_t *)calloc((size_t)dst_w, sizeof(int64_t));
    for (int i = 0; i < dst_w; ++i) {
      filter[i] = fone;
      fpos[i] = i;
    }
  } else {
    const int size_factor = 4; /* bicubic */
    if (x_inc <= 1 << 16)
      filter_size = 1 + size_factor;
    else
      filter_size = 1 + (size_factor * src_w + dst_w - 1) / dst_w;
    if (filter_size > src_w - 2) filter_size = src_w - 2;
    if (filter_size < 1) filter_size = 1;
    filter = (int64_t *)calloc((size_t)dst_w * filter_size, sizeof(int64_t));
    int64_t x_dst_in_src = ((dst_pos * (int64_t)x_inc) >> 7) - ((src_pos * 0x10000ll) >> 7);
    const int64_t B = 0;
    const int64_t C = (int64_t)(0.6 * (1 << 24));
    for (int i = 0; i < dst_w; ++i) {
      int xx = (int)((x_dst_in_src - (filter_size - 2) * (1ll << 16)) / (1 << 17));
      fpos[i] = xx;
      for (int j = 0; j < filter_size; ++j) {
        int64_t d = llabs(((int64_t)xx * (1 << 17)) - x_dst_in_src) << 13;
        int64_t c;
        if (x_inc > 1 << 16) d = d * dst_w / src_w;
        if (d >= 1ll << 31) {
          c = 0;
        } else {
          int64_t dd = (d * d) >> 30;
          int64_t ddd = (dd * d) >> 30;
          if (d < 1ll << 30)
            c = (12 * (1 << 24) - 9 * B - 6 * C) * ddd +
                (-18 * (1 << 24) + 12 * B + 6 * C) * dd +
                (6 * (1 << 24) - 2 * B) * (1 << 30);
          else
            c = (-B - 6 * C) * ddd + (6 * B + 30 * C) * dd +
                (-12 * B - 48 * C) * d + (8 * B + 24 * C) * (1 << 30);
        }
        c /= (1ll << 54) / fone;
        filter[i * filter_size + j] = c;
        xx++;
      }
      x_dst_in_src += 2 * (int64_t)x_inc;
    }
  }

  /* no src/dst vectors: filter2 == filter */
  const int f2size = filter_size;
  int64_t *f2 = filter;

  /* reduce: shift out near-zero leading taps, count trailing ones */
  int min_filter_size = 0;
  for (int i = dst_w - 1; i >= 0; --i) {
    int min = f2size;
    int64_t cut = 0;
    for (int j = 0; j < f2size; ++j) {
      cut += llabs(f2[i * f2size]);
      if ((double)cut > 0.002 * (double)fone) break;
      if (i < dst_w - 1 && fpos[i] >= fpos[i + 1]) break;
      int k;
      for (k = 1; k < f2size; ++k) f2[i * f2size + k - 1] = f2[i * f2size + k];
      f2[i * f2size + k - 1] = 0;
      fpos[i]++;
    }
    cut = 0;
    for (int j = f2size - 1; j > 0; --j) {
      cut += llabs(f2[i * f2size + j]);
      if ((double)cut > 0.002 * (double)fone) break;
      min--;
    }
    if (min > min_filter_size) min_filter_size = min;
  }
  /* x86 MMX builds: unscaled vertical filtering special case */
  if (min_filter_size == 1 && filter_align == 2) filter_align = 1;
  int out_size = (min_filter_size + (filter_align - 1)) & ~(filter_align - 1);
  if (out_size > max_size) {
    free(filter);
    free(fpos);
    return -1;
  }
  int64_t *fr = (int64_t *)calloc((size_t)dst_w * out_size, sizeof(int64_t));
  for (int i = 0; i < dst_w; ++i)
    for (int j = 0; j < out_size; ++j)
      fr[i * out_size + j] = j >= f2size ? 0 : f2[i * f2size + j];

  /* fix borders (clamp to edge) */
  for (int i = 0; i < dst_w; ++i) {
    int64_t *f = fr + (size_t)i * out_size;
    if (fpos[i] < 0) {
      for (int j = 1; j < out_size; ++j) {
        int left = j + fpos[i] > 0 ? j + fpos[i] : 0;
        f[left] += f[j];
        f[j] = 0;
      }
      fpos[i] = 0;
    }
    if (fpos[i] + out_size > src_w) {
      int shift = fpos[i] + (out_size - src_w < 0 ? out_size - src_w : 0);
      int64_t acc = 0;
      for (int j = out_size - 1; j >= 0; --j)
        if (fpos[i] + j >= src_w) {
          acc += f[j];
          f[j] = 0;
        }
      for (int j = out_size - 1; j >= 0; --j) f[j] = j < shift ? 0 : f[j - shift];
      fpos[i] -= shift;
      f[src_w - 1 - fpos[i]] += acc;
    }
  }

  /* normalise to `one` with error diffusion */
  for (int i = 0; i < dst_w; ++i) {
    int64_t err = 0, sum = 0;
    for (int j = 0; j < out_size; ++j) sum += fr[i * out_size + j];
    sum = (sum + one / 2) / one;
    if (!sum) sum = 1;
    for (int j = 0; j < out_size; ++j) {
      int64_t v = fr[i * out_size + j] + err;
      int64_t iv = rounded_div(v, sum);
      coef[i * out_size + j] = (int16_t)iv;
      err = v - iv * sum;
    }
    pos[i] = fpos[i];
  }
  free(fr);
  free(filter);
  free(fpos);
  return out_size;
}

/* get_local_pos() with the default -513 chroma position. */
static int local_pos(int chr_subsample) {
  int p = (128 << chr_subsample) - 128;
  p += 128;
  return p >> chr_subsample;
}

int orc_sws_vchroma_filter(int h, int16_t *coef, int32_t *pos, int max_size) {
  const int src_h = h, dst_h = (h + 1) >> 1;
  const int inc = (int)((((int64_t)src_h << 16) + (dst_h >> 1)) / dst_h);
  return orc_sws_init_filter(inc, src_h, dst_h, 2, 1 << 12, local_pos(0),
                             local_pos(1), coef, pos, max_size);
}

static inline uint8_t clip_u8(int v) { return v < 0 ? 0 : (v > 255 ? 255 : (uint8_t)v); }

/* rgb24ToY_c (14-bit intermediate) */
static inline int y14(int r, int g, int b) {
  return (SWS_RY * r + SWS_GY * g + SWS_BY * b + (32 << (RGB2YUV_SHIFT - 1)) +
          (1 << (RGB2YUV_SHIFT - 7))) >> (RGB2YUV_SHIFT - 6);
}

/* rgb24ToUV_half_c: r,g,b are sums of two horizontally adjacent pixels */
static inline int u14(int r, int g, int b) {
  return (SWS_RU * r + SWS_GU * g + SWS_BU * b + (256 << RGB2YUV_SHIFT) +
          (1 << (RGB2YUV_SHIFT - 6))) >> (RGB2YUV_SHIFT - 5);
}
static inline int v14(int r, int g, int b) {
  return (SWS_RV * r + SWS_GV * g + SWS_BV * b + (256 << RGB2YUV_SHIFT) +
          (1 << (RGB2YUV_SHIFT - 6))) >> (RGB2YUV_SHIFT - 5);
}

/* hScale16To15_c with the unscaled 1-tap filter (coefficient 1<<14), sh=13 */
static inline int to15(int v14v) {
  int v = (v14v * (1 << 14)) >> 13;
  return v < (1 << 15) - 1 ? v : (1 << 15) - 1;
}

/* Vertical chroma output of one sample column.  x86 = yuv2yuvX_sse3
 * (pmulhw taps, +((64 + 8*(n-1))>>4), >>3, packuswb); C = yuv2planeX_8_c
 * ((64<<12 + sum) >> 19, clip). */
static uint8_t vchroma(const int16_t *coef, int n, const int *src, int x86) {
  if (x86) {
    int16_t acc = (int16_t)((64 + 8 * (n - 1)) >> 4);
    for (int j = 0; j < n; ++j) {
      int prod = (int)coef[j] * (int)(int16_t)src[j];
      acc = (int16_t)(acc + (int16_t)(prod >> 16));
    }
    int v = acc >> 3;
    return clip_u8(v);
  }
  int val = 64 << 12;
  for (int j = 0; j < n; ++j) val += src[j] * coef[j];
  return clip_u8(val >> 19);
}

#define ORC_MAX_VTAPS 16

/* Threads for the frame-prep row loops (rows are independent, so the output
 * does not depend on it); 1 = one thread per frame, as process_frame_thread
 * runs one per eye (src/main.cpp:274-282). */
static int g_fp_threads = 1;
void orc_set_frameprep_threads(int n) { g_fp_threads = n > 0 ? n : 1; }

static int yuv420p_from_planes15(int w, int h, const int16_t *u15,
                                 const int16_t *v15, uint8_t *u, int u_stride,
                                 uint8_t *v, int v_stride, int mode) {
  const int cw = w >> 1, ch = h >> 1;
  int16_t *coef = (int16_t *)malloc(sizeof(int16_t) * (size_t)ch * ORC_MAX_VTAPS);
  int32_t *pos = (int32_t *)malloc(sizeof(int32_t) * (size_t)ch);
  int n = orc_sws_vchroma_filter(h, coef, pos, ORC_MAX_VTAPS);
  if (n < 0) {
    free(coef);
    free(pos);
    return -1;
  }
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(g_fp_threads)
#endif
  for (int k = 0; k < ch; ++k) {
    int su[ORC_MAX_VTAPS], sv[ORC_MAX_VTAPS];
    /* swscale.c: output rows dstY >= dstH-2 use the C output functions */
    int x86 = (mode == ORC_SWS_MODE_X86) && (2 * k < h - 2);
    for (int i = 0; i < cw; ++i) {
      for (int j = 0; j < n; ++j) {
        int row = pos[k] + j;
        if (row > h - 1) row = h - 1; /* coefficient is zero there */
        su[j] = u15[(size_t)row * cw + i];
        sv[j] = v15[(size_t)row * cw + i];
      }
      u[(size_t)k * u_stride + i] = vchroma(coef + (size_t)k * n, n, su, x86);
      v[(size_t)k * v_stride + i] = vchroma(coef + (size_t)k * n, n, sv, x86);
    }
  }
  free(coef);
  free(pos);
  return 0;
}

int orc_rgb24_to_yuv420p(const uint8_t *rgb, int rgb_stride, int w, int h,
                         uint8_t *y, int y_stride, uint8_t *u, int u_stride,
                         uint8_t *v, int v_stride, int mode) {
  if (w < 2 || h < 2 || (w & 1) || (h & 1)) return -1;
  const int cw = w >> 1;
  int16_t *u15 = (int16_t *)malloc(sizeof(int16_t) * (size_t)cw * h);
  int16_t *v15 = (int16_t *)malloc(sizeof(int16_t) * (size_t)cw * h);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(g_fp_threads)
#endif
  for (int r = 0; r < h; ++r) {
    const uint8_t *p = rgb + (size_t)r * rgb_stride;
    for (int x = 0; x < w; ++x) {
      int yy = to15(y14(p[3 * x], p[3 * x + 1], p[3 * x + 2]));
      y[(size_t)r * y_stride + x] = clip_u8((yy + 64) >> 7); /* yuv2plane1_8_c */
    }
    for (int i = 0; i < cw; ++i) {
      int rs = p[6 * i + 0] + p[6 * i + 3];
      int gs = p[6 * i + 1] + p[6 * i + 4];
      int bs = p[6 * i + 2] + p[6 * i + 5];
      u15[(size_t)r * cw + i] = (int16_t)to15(u14(rs, gs, bs));
      v15[(size_t)r * cw + i] = (int16_t)to15(v14(rs, gs, bs));
    }
  }
  int rc = yuv420p_from_planes15(w, h, u15, v15, u, u_stride, v, v_stride, mode);
  free(u15);
  free(v15);
  return rc;
}

int orc_gray8_to_yuv420p(const uint8_t *g, int g_stride, int w, int h,
                         uint8_t *y, int y_stride, uint8_t *u, int u_stride,
                         uint8_t *v, int v_stride, int mode) {
  if (w < 2 || h < 2 || (w & 1) || (h & 1)) return -1;
  const int cw = w >> 1;
  for (int r = 0; r < h; ++r)
    for (int x = 0; x < w; ++x) {
      /* hScale8To15: 1-tap unscaled, (g << 14) >> 7 */
      int v15 = ((int)g[(size_t)r * g_stride + x] * (1 << 14)) >> 7;
      /* lumRangeFromJpeg_c: full -> limited range */
      v15 = (v15 * 14071 + 33561947) >> 14;
      y[(size_t)r * y_stride + x] = clip_u8((v15 + 64) >> 7);
    }
  /* chroma planes of a gray source are constant 1<<14 before the vertical filter */
  int16_t *c15 = (int16_t *)malloc(sizeof(int16_t) * (size_t)cw * h);
  for (size_t i = 0; i < (size_t)cw * h; ++i) c15[i] = 1 << 14;
  int rc = yuv420p_from_planes15(w, h, c15, c15, u, u_stride, v, v_stride, mode);
  free(c15);
  return rc;
}

void orc_composite(const uint8_t *rgb_a, const float *depth_a,
                   const uint8_t *rgb_b, const float *depth_b, int w, int h,
                   int rgb_stride, uint8_t *rgb_out, float *depth_out) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(g_fp_threads)
#endif
  for (int r = 0; r < h; ++r)
    for (int x = 0; x < w; ++x) {
      size_t pi = (size_t)r * w + x;
      const uint8_t *src = rgb_a + (size_t)r * rgb_stride + 3 * x;
      float d = depth_a ? depth_a[pi] : 0.0f;
      if (rgb_b) {
        float ka = isnan(depth_a[pi]) ? INFINITY : depth_a[pi];
        float kb = isnan(depth_b[pi]) ? INFINITY : depth_b[pi];
        if (!(ka <= kb)) {
          src = rgb_b + (size_t)r * rgb_stride + 3 * x;
          d = depth_b[pi];
        }
      }
      memcpy(rgb_out + (size_t)r * 3 * w + 3 * x, src, 3);
      if (depth_out) depth_out[pi] = d;
    }
}

int orc_frameprep(const uint8_t *rgb_a, const float *depth_a,
                  const uint8_t *rgb_b, const float *depth_b, int w, int h,
                  uint8_t *y, int y_stride, uint8_t *u, int u_stride,
                  uint8_t *v, int v_stride, int mode) {
  if (w < 2 || h < 2 || (w & 1) || (h & 1)) return -1;
  uint8_t *tmp = (uint8_t *)malloc((size_t)w * h * 3);
  orc_composite(rgb_a, depth_a, rgb_b, depth_b, w, h, 3 * w, tmp, NULL);
  int rc = orc_rgb24_to_yuv420p(tmp, 3 * w, w, h, y, y_stride, u, u_stride, v,
                                v_stride, mode);
  free(tmp);
  return rc;
}

/* ------------------------------------------------------------ text overlay
 * render_text.cc:35-111.  The pen starts per position (x_box 300, y_box 100,
 * margin 50, :48-75) computed, as there, from the unsigned frame size and
 * stored in a signed int; '\n' returns the pen to its starting x and moves
 * it 20 px down (:80-84); every other byte loads its glyph (:86, indexed here
 * by the unsigned byte) and each non-zero bitmap byte at row q, column p
 * writes (255, 255, 255) at (pen_x + left + p, pen_y - top + q) when that
 * lies inside the frame (:92-105); the pen then advances (:108). */
void orc_render_string(uint8_t *surface, uint32_t width, uint32_t height,
                       const orc_glyph *glyphs, const uint8_t *bitmaps,
                       int position, const char *content) {
  const int x_box = 300, y_box = 100, margin = 50;
  int pen_x, pen_y;
  switch (position) {
    case 0: pen_x = margin; pen_y = margin; break;
    case 1: pen_x = margin; pen_y = (int)(height - y_box + margin); break;
    case 2: pen_x = (int)(width - x_box + margin); pen_y = margin; break;
    case 3:
      pen_x = (int)(width - x_box + margin);
      pen_y = (int)(height - y_box + margin);
      break;
    case 4: pen_x = (int)(width / 2 - x_box); pen_y = (int)(height / 2 - y_box); break;
    default: pen_x = margin; pen_y = margin; break;
  }
  const int orig_pen_x = pen_x;
  for (const char *ch = content; *ch; ++ch) {
    if (*ch == '\n') {
      pen_x = orig_pen_x;
      pen_y += 20;
      continue;
    }
    const orc_glyph *g = &glyphs[(unsigned char)*ch];
    const uint8_t *bm = bitmaps + g->offset;
    const int x_first = pen_x + g->left, y_first = pen_y - g->top;
    const int x_max = x_first + (int)g->width, y_max = y_first + (int)g->rows;
    for (int j = y_first, q = 0; j < y_max; ++j, ++q) {
      for (int i = x_first, p = 0; i < x_max; ++i, ++p) {
        if (i < 0 || j < 0 || (uint32_t)i >= width || (uint32_t)j >= height) continue;
        if (bm[(size_t)q * g->width + p]) {
          uint8_t *px = surface + ((size_t)j * width + (size_t)i) * 3;
          px[0] = 255;
          px[1] = 255;
          px[2] = 255;
        }
      }
    }
    pen_x += g->advance;
  }
}
