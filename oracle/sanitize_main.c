/*
 * sanitize_main.c -- drives every oracle entry point once on small inputs,
 * for a build with AddressSanitizer and UndefinedBehaviorSanitizer
 * (`make -C oracle sanitize`).  TEST INFRASTRUCTURE ONLY: it checks the
 * oracle's own memory safety and arithmetic, not its results (the pytest
 * suite pins those).  Buffers are allocated at their exact sizes so that any
 * read or write past an end is reported.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "nes_oracle.h"

#define N_TABLE 6098925
#define N_WEIGHTS 9344
#define OCC_WORDS 65536

static void *xmalloc(size_t n) {
  void *p = malloc(n ? n : 1);
  if (!p) {
    fprintf(stderr, "out of memory\n");
    exit(2);
  }
  return p;
}

static void fill_bytes(uint8_t *p, size_t n, uint64_t seed) {
  for (size_t i = 0; i < n; ++i) p[i] = (uint8_t)(orc_rand_u64(seed, 7, i) >> 56);
}

static void frameprep_case(int w, int h, int mode, int composite, uint64_t seed) {
  const size_t px = (size_t)w * h;
  uint8_t *a = xmalloc(3 * px), *b = xmalloc(3 * px);
  float *da = xmalloc(4 * px), *db = xmalloc(4 * px);
  fill_bytes(a, 3 * px, seed);
  fill_bytes(b, 3 * px, seed + 1);
  for (size_t i = 0; i < px; ++i) {
    da[i] = (float)(0.1 + 9.9 * orc_uniform(seed, 8, i));
    db[i] = (i % 17 == 0) ? NAN : (float)(0.1 + 9.9 * orc_uniform(seed, 9, i));
  }
  const int cw = w / 2, ch = h / 2;
  uint8_t *y = xmalloc(px), *u = xmalloc((size_t)cw * ch), *v = xmalloc((size_t)cw * ch);
  if (orc_frameprep(a, da, composite ? b : NULL, composite ? db : NULL, w, h, y, w, u, cw, v, cw,
                    mode) != 0) {
    fprintf(stderr, "orc_frameprep failed %dx%d\n", w, h);
    exit(1);
  }
  uint8_t *g = xmalloc(px);
  orc_depth_gray8(da, (int)px, 4.0f, g);
  if (orc_gray8_to_yuv420p(g, w, w, h, y, w, u, cw, v, cw, mode) != 0) exit(1);
  free(a); free(b); free(da); free(db); free(y); free(u); free(v); free(g);
}

int main(void) {
  /* frame prep, every mode, small and ragged sizes, composite and plain */
  const int sizes[][2] = {{2, 2}, {4, 2}, {2, 8}, {6, 4}, {10, 12}, {130, 98}, {258, 66}};
  for (size_t s = 0; s < sizeof(sizes) / sizeof(sizes[0]); ++s)
    for (int mode = 0; mode < 2; ++mode)
      for (int comp = 0; comp < 2; ++comp)
        frameprep_case(sizes[s][0], sizes[s][1], mode, comp, 100 + s);
  orc_set_frameprep_threads(4);
  frameprep_case(640, 480, ORC_SWS_MODE_X86, 1, 7);
  orc_set_frameprep_threads(1);
  for (int h = 2; h <= 64; h += 2) {
    int16_t coef[32 * 16];
    int32_t pos[32];
    if (orc_sws_vchroma_filter(h, coef, pos, 16) < 1) exit(1);
  }

  /* scene, camera, render */
  float *table = xmalloc(sizeof(float) * 2 * N_TABLE);
  float *weights = xmalloc(sizeof(float) * N_WEIGHTS);
  uint32_t *occ = xmalloc(sizeof(uint32_t) * OCC_WORDS);
  orc_scene_generate(1, 0, table, weights, occ);
  orc_camera cam;
  memset(&cam, 0, sizeof(cam));
  cam.width = 64;
  cam.height = 48;
  cam.fov_y_deg = 60.0f;
  orc_orbit_camera(30.0, 20.0, 2.0, 0.0, cam.c2w);
  const int n = 64 * 48;
  uint32_t *pix = xmalloc(sizeof(uint32_t) * n);
  for (int i = 0; i < n; ++i) pix[i] = (uint32_t)i;
  double *rgb = xmalloc(sizeof(double) * 3 * n), *depth = xmalloc(sizeof(double) * n);
  int *ns = xmalloc(sizeof(int) * n);
  orc_render_pixels(table, weights, occ, &cam, pix, n, rgb, depth, ns, 2);
  for (int i = 0; i < n; ++i) (void)orc_quantize(rgb[3 * i]);

  /* training: loss, gradients, a perturbed loss, Adam */
  const int nr = 24;
  float *ro = xmalloc(sizeof(float) * 3 * nr), *rd = xmalloc(sizeof(float) * 3 * nr);
  float *tgt = xmalloc(sizeof(float) * 3 * nr);
  uint32_t *pose = xmalloc(sizeof(uint32_t) * nr), *pidx = xmalloc(sizeof(uint32_t) * nr);
  orc_training_rays(0, 0, nr, 16, 640, 480, ro, rd, pose, pidx);
  for (int i = 0; i < 3 * nr; ++i) tgt[i] = (float)orc_uniform(3, 3, (uint64_t)i);
  double *gt = calloc((size_t)2 * N_TABLE, sizeof(double));
  double *gw = calloc(N_WEIGHTS, sizeof(double));
  if (!gt || !gw) exit(2);
  (void)orc_loss_grad(table, weights, occ, nr, ro, rd, tgt, gt, gw);
  (void)orc_loss_perturbed(table, weights, occ, nr, ro, rd, tgt, 12345, -1, 1e-3);
  (void)orc_loss_perturbed(table, weights, occ, nr, ro, rd, tgt, -1, 77, 1e-3);
  for (int i = 0; i < nr; ++i) (void)orc_ray_relu_margin(table, weights, occ, ro + 3 * i, rd + 3 * i);
  double p[8] = {0}, g[8] = {1, -1, 0, 2, -2, 1e-9, 0, 3}, m[8] = {0}, v[8] = {0};
  orc_adam_step(8, p, g, m, v, 1, 1e-2, 0.9, 0.99, 1e-15);

  /* text overlay: glyphs that fall partly and wholly outside the frame */
  orc_glyph glyphs[256];
  uint8_t atlas[256 * 12];
  for (int c = 0; c < 256; ++c) {
    glyphs[c].left = (c % 5) - 2;
    glyphs[c].top = 8 + (c % 3);
    glyphs[c].width = 3;
    glyphs[c].rows = 4;
    glyphs[c].advance = 4;
    glyphs[c].offset = (uint32_t)(12 * c);
  }
  fill_bytes(atlas, sizeof(atlas), 11);
  uint8_t *surface = xmalloc(3 * 40 * 30);
  memset(surface, 0, 3 * 40 * 30);
  for (int position = 0; position < 5; ++position)
    orc_render_string(surface, 40, 30, glyphs, atlas, position,
                      "index=12345\n+0.50000 -1.00000\nlong line that runs off the right edge");

  free(table); free(weights); free(occ); free(pix); free(rgb); free(depth); free(ns);
  free(ro); free(rd); free(tgt); free(pose); free(pidx); free(gt); free(gw); free(surface);
  printf("oracle sanitizer run: ok\n");
  return 0;
}
