/* A C program linked against libnes_render.so through include/nes_render.h
 * only, as the reference's C++ (src/server.cpp, src/encode.cpp) would call
 * it.  It prints one line per output buffer, "name checksum", with
 * checksum = sum_i byte[i] * (i % 65521 + 1) mod 2^64, which the tests
 * recompute from the Python binding's results.
 *   c_abi_client cpu   calls that need no device
 *   c_abi_client gpu   scene upload, a stereo render (RGB24, f32 depth,
 *                      GRAY8 depth), frame prep of both outputs and a
 *                      server round trip of the same views */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "nes_render.h"

static uint64_t checksum(const void *p, size_t n) {
  const uint8_t *b = (const uint8_t *)p;
  uint64_t s = 0;
  for (size_t i = 0; i < n; ++i) s += (uint64_t)b[i] * (uint64_t)(i % 65521 + 1);
  return s;
}

static void print(const char *name, const void *p, size_t n) {
  printf("%s %llu\n", name, (unsigned long long)checksum(p, n));
}

static int fail(const char *what, int rc) {
  fprintf(stderr, "%s: %d (%s) %s\n", what, rc, nes_strerror(rc), nes_last_error());
  return 1;
}

#define CHECK(call)                          \
  do {                                       \
    int rc_ = (call);                        \
    if (rc_ != NES_OK) return fail(#call, rc_); \
  } while (0)

static int cpu_part(void) {
  printf("strerror %s\n", nes_strerror(NES_EINVAL));
  float c2w[12];
  CHECK(nes_orbit_camera(30.0, 20.0, 2.0, -0.032, c2w));
  print("orbit", c2w, sizeof(c2w));
  float *w = (float *)malloc(NES_MLP_WEIGHTS * sizeof(float));
  uint32_t *occ = (uint32_t *)malloc(NES_OCC_WORDS * sizeof(uint32_t));
  CHECK(nes_scene_synthesize(1, 0, NULL, w, occ));
  print("weights", w, NES_MLP_WEIGHTS * sizeof(float));
  print("occupancy", occ, NES_OCC_WORDS * sizeof(uint32_t));
  int16_t coef[16 * 360];
  int32_t pos[360];
  int taps = 0;
  CHECK(nes_sws_vchroma_filter(720, coef, pos, 16, &taps));
  printf("vfilter_taps %d\n", taps);
  print("vfilter_coef", coef, (size_t)taps * 360 * sizeof(int16_t));
  free(w);
  free(occ);
  return 0;
}

static int gpu_part(void) {
  const int W = 160, H = 120;
  float *table = (float *)malloc((size_t)NES_TABLE_ENTRIES * 2 * sizeof(float));
  float *weights = (float *)malloc(NES_MLP_WEIGHTS * sizeof(float));
  uint32_t *occ = (uint32_t *)malloc(NES_OCC_WORDS * sizeof(uint32_t));
  CHECK(nes_scene_synthesize(1, 0, table, weights, occ));
  nes_scene_desc scene = {table, weights, occ};
  nes_ctx *ctx = NULL;
  CHECK(nes_ctx_create(&scene, 0, &ctx));

  nes_camera cams[2];
  memset(cams, 0, sizeof(cams));
  for (int e = 0; e < 2; ++e) {
    cams[e].index = 7;
    cams[e].is_left = e == 0;
    cams[e].width = W;
    cams[e].height = H;
    cams[e].fov_y_deg = 60.0f;
    CHECK(nes_orbit_camera(30.0, 20.0, 2.0, e == 0 ? -0.032 : 0.032, cams[e].c2w));
  }
  const size_t px = (size_t)W * H;
  uint8_t *rgb[2], *gray[2];
  float *depth[2];
  nes_view_out outs[2];
  for (int e = 0; e < 2; ++e) {
    rgb[e] = (uint8_t *)malloc(3 * px);
    gray[e] = (uint8_t *)malloc(px);
    depth[e] = (float *)malloc(px * sizeof(float));
    nes_view_out o = {rgb[e], depth[e], NULL, NULL, gray[e], 0.0f};
    outs[e] = o;
  }
  CHECK(nes_render(ctx, NES_PRECISION_MIXED, cams, 2, outs, 0, NULL));
  for (int e = 0; e < 2; ++e) {
    char name[32];
    snprintf(name, sizeof(name), "rgb24_%d", e);
    print(name, rgb[e], 3 * px);
    snprintf(name, sizeof(name), "gray8_%d", e);
    print(name, gray[e], px);
  }
  /* frame prep of the left eye, alone and composited with the right one */
  uint8_t *y = (uint8_t *)malloc(px), *u = (uint8_t *)malloc(px / 4), *v = (uint8_t *)malloc(px / 4);
  CHECK(nes_frameprep(rgb[0], NULL, NULL, NULL, W, H, 3 * W, y, W, u, W / 2, v, W / 2, NES_SWS_X86));
  print("yuv_y", y, px);
  print("yuv_u", u, px / 4);
  print("yuv_v", v, px / 4);
  CHECK(nes_frameprep(rgb[0], depth[0], rgb[1], depth[1], W, H, 3 * W, y, W, u, W / 2, v, W / 2,
                      NES_SWS_X86));
  print("comp_y", y, px);
  print("comp_u", u, px / 4);
  CHECK(nes_gray8_to_yuv420p(gray[0], W, H, W, y, W, u, W / 2, v, W / 2, NES_SWS_X86));
  print("depth_y", y, px);
  print("depth_u", u, px / 4);

  /* the same stereo request through the server */
  nes_server_config cfg;
  memset(&cfg, 0, sizeof(cfg));
  cfg.precision = NES_PRECISION_MIXED;
  cfg.max_batch_rays = 1u << 20;
  nes_server *srv = NULL;
  CHECK(nes_server_create(ctx, &cfg, &srv));
  uint64_t ticket = 0;
  CHECK(nes_server_submit(srv, 1, cams, 2, &ticket));
  nes_frame_result r;
  CHECK(nes_server_wait(srv, ticket, -1, &r));
  CHECK(r.status);
  print("server_y0", r.y[0], px);
  print("server_u0", r.u[0], px / 4);
  print("server_v1", r.v[1], px / 4);
  CHECK(nes_server_release(srv, ticket));
  CHECK(nes_server_destroy(srv));
  CHECK(nes_ctx_destroy(ctx));
  for (int e = 0; e < 2; ++e) {
    free(rgb[e]);
    free(gray[e]);
    free(depth[e]);
  }
  free(y);
  free(u);
  free(v);
  free(table);
  free(weights);
  free(occ);
  return 0;
}

int main(int argc, char **argv) {
  if (argc != 2) {
    fprintf(stderr, "usage: %s cpu|gpu\n", argv[0]);
    return 2;
  }
  if (!strcmp(argv[1], "cpu")) return cpu_part();
  if (!strcmp(argv[1], "gpu")) return gpu_part();
  fprintf(stderr, "unknown mode %s\n", argv[1]);
  return 2;
}
