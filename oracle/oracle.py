"""ctypes view of the CPU oracle (test infrastructure only).

Only tests/, the smoke check and the bench's CPU-baseline leg import this
module, and only as the checker.  It loads oracle/lib/libnes_oracle.so, built
by oracle/Makefile from oracle/nes_oracle.c.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libnes_oracle.so")

N_TABLE = 6098925
N_WEIGHTS = 9344
OCC_WORDS = 65536
MODE_X86 = 0
MODE_C = 1
DELTA = np.float32(0.0135316469)


class Camera(C.Structure):
    _fields_ = [("index", C.c_uint64), ("width", C.c_uint32), ("height", C.c_uint32),
                ("c2w", C.c_float * 12), ("fov_y_deg", C.c_float), ("is_left", C.c_uint32)]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _p(t):
    return C.POINTER(t)


def _declare(L):
    u8p, f32p, f64p, u32p, i32p, i16p = (_p(C.c_uint8), _p(C.c_float), _p(C.c_double),
                                         _p(C.c_uint32), _p(C.c_int32), _p(C.c_int16))
    sig = {
        "orc_mix64": (C.c_uint64, [C.c_uint64]),
        "orc_set_frameprep_threads": (None, [C.c_int]),
        "orc_ray_relu_margin": (C.c_double, [f32p, f32p, u32p, f32p, f32p]),
        "orc_render_string": (None, [u8p, C.c_uint32, C.c_uint32, _p(Glyph), u8p, C.c_int,
                                     C.c_char_p]),
        "orc_rand_u64": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64]),
        "orc_uniform": (C.c_double, [C.c_uint64, C.c_uint64, C.c_uint64]),
        "orc_normal": (C.c_double, [C.c_uint64, C.c_uint64, C.c_uint64]),
        "orc_level_res": (C.c_int, [C.c_int]),
        "orc_level_offset": (C.c_uint32, [C.c_int]),
        "orc_level_size": (C.c_uint32, [C.c_int]),
        "orc_grid_index": (C.c_uint32, [C.c_int, C.c_uint32, C.c_uint32, C.c_uint32]),
        "orc_scene_generate": (None, [C.c_uint64, C.c_int, f32p, f32p, u32p]),
        "orc_occupancy_fraction": (C.c_double, [u32p]),
        "orc_orbit_camera": (None, [C.c_double, C.c_double, C.c_double, C.c_double, f32p]),
        "orc_focal": (C.c_float, [C.c_uint32, C.c_float]),
        "orc_camera_ray": (None, [_p(Camera), C.c_uint32, C.c_uint32, f32p, f32p]),
        "orc_ray_box": (C.c_int, [f32p, f32p, f32p, f32p]),
        "orc_step_t": (C.c_float, [C.c_int, C.c_float]),
        "orc_occupied": (C.c_int, [u32p, f32p]),
        "orc_encode": (None, [f32p, f64p, f64p]),
        "orc_sh4": (None, [f64p, f64p]),
        "orc_mlp": (None, [f32p, f64p, f64p, f64p, f64p]),
        "orc_render_ray": (None, [f32p, f32p, u32p, f32p, f32p, _p(RayResult)]),
        "orc_render_pixels": (None, [f32p, f32p, u32p, _p(Camera), u32p, C.c_int, f64p, f64p,
                                     i32p, C.c_int]),
        "orc_quantize": (C.c_uint8, [C.c_double]),
        "orc_depth_gray8": (None, [f32p, C.c_int, C.c_float, u8p]),
        "orc_loss_grad": (C.c_double, [f32p, f32p, u32p, C.c_int, f32p, f32p, f32p, f64p, f64p]),
        "orc_loss_perturbed": (C.c_double, [f32p, f32p, u32p, C.c_int, f32p, f32p, f32p,
                                            C.c_long, C.c_long, C.c_double]),
        "orc_adam_step": (None, [C.c_int, f64p, f64p, f64p, f64p, C.c_int, C.c_double,
                                 C.c_double, C.c_double, C.c_double]),
        "orc_training_rays": (None, [C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.c_uint32,
                                     C.c_uint32, f32p, f32p, u32p, u32p]),
        "orc_training_pose": (None, [C.c_uint64, C.c_int, f32p]),
        "orc_sws_init_filter": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                          C.c_int, i16p, i32p, C.c_int]),
        "orc_sws_vchroma_filter": (C.c_int, [C.c_int, i16p, i32p, C.c_int]),
        "orc_rgb24_to_yuv420p": (C.c_int, [u8p, C.c_int, C.c_int, C.c_int, u8p, C.c_int, u8p,
                                           C.c_int, u8p, C.c_int, C.c_int]),
        "orc_gray8_to_yuv420p": (C.c_int, [u8p, C.c_int, C.c_int, C.c_int, u8p, C.c_int, u8p,
                                           C.c_int, u8p, C.c_int, C.c_int]),
        "orc_composite": (None, [u8p, f32p, u8p, f32p, C.c_int, C.c_int, C.c_int, u8p, f32p]),
        "orc_frameprep": (C.c_int, [u8p, f32p, u8p, f32p, C.c_int, C.c_int, u8p, C.c_int, u8p,
                                    C.c_int, u8p, C.c_int, C.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def ptr(a, t):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(t))


# ----------------------------------------------------------------- scene / camera

def scene(seed, table_init=0, table=True, weights=True, occ=True):
    L = lib()
    t = np.empty(N_TABLE * 2, np.float32) if table else None
    w = np.empty(N_WEIGHTS, np.float32) if weights else None
    o = np.empty(OCC_WORDS, np.uint32) if occ else None
    L.orc_scene_generate(seed, table_init, ptr(t, C.c_float), ptr(w, C.c_float),
                         ptr(o, C.c_uint32))
    return t, w, o


def orbit_c2w(az, el, radius=2.0, eye_offset=0.0):
    m = np.empty(12, np.float32)
    lib().orc_orbit_camera(az, el, radius, eye_offset, ptr(m, C.c_float))
    return m


def camera(width, height, c2w, fov=60.0, index=0, is_left=1):
    cam = Camera()
    cam.index, cam.width, cam.height = index, width, height
    cam.c2w[:] = [float(x) for x in c2w]
    cam.fov_y_deg = fov
    cam.is_left = is_left
    return cam


def camera_rays(cam, pix):
    """FP32 ray origins and directions of the given flat pixel indices."""
    L = lib()
    o = np.empty((len(pix), 3), np.float32)
    d = np.empty((len(pix), 3), np.float32)
    for i, p in enumerate(pix):
        L.orc_camera_ray(C.byref(cam), int(p) % cam.width, int(p) // cam.width,
                         ptr(o[i], C.c_float), ptr(d[i], C.c_float))
    return o, d


class RayResult(C.Structure):
    _fields_ = [("rgb", C.c_double * 3), ("depth", C.c_double), ("transmittance", C.c_double),
                ("n_samples", C.c_int)]


def render_ray(table, weights, occ, o, d):
    """(rgb, depth, transmittance, n_samples) of one FP32 ray."""
    out = RayResult()
    o = np.ascontiguousarray(o, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    lib().orc_render_ray(ptr(table, C.c_float), ptr(weights, C.c_float), ptr(occ, C.c_uint32),
                         ptr(o, C.c_float), ptr(d, C.c_float), C.byref(out))
    return np.array(out.rgb), out.depth, out.transmittance, out.n_samples


def ray_relu_margin(table, weights, occ, o, d):
    """Smallest |pre-activation| of the ReLU layers over one ray's samples."""
    o = np.ascontiguousarray(o, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    return lib().orc_ray_relu_margin(ptr(table, C.c_float), ptr(weights, C.c_float),
                                     ptr(occ, C.c_uint32), ptr(o, C.c_float), ptr(d, C.c_float))


def render_pixels(table, weights, occ, cam, pix, threads=8):
    L = lib()
    pix = np.ascontiguousarray(pix, np.uint32)
    n = len(pix)
    rgb = np.empty((n, 3), np.float64)
    depth = np.empty(n, np.float64)
    ns = np.empty(n, np.int32)
    L.orc_render_pixels(ptr(table, C.c_float), ptr(weights, C.c_float), ptr(occ, C.c_uint32),
                        C.byref(cam), ptr(pix, C.c_uint32), n, ptr(rgb, C.c_double),
                        ptr(depth, C.c_double), ptr(ns, C.c_int32), threads)
    return rgb, depth, ns


def encode(table, x):
    out = np.empty(32, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    lib().orc_encode(ptr(table, C.c_float), ptr(x, C.c_double), ptr(out, C.c_double))
    return out


def sh4(d):
    out = np.empty(16, np.float64)
    d = np.ascontiguousarray(d, np.float64)
    lib().orc_sh4(ptr(d, C.c_double), ptr(out, C.c_double))
    return out


def mlp(weights, feat, sh):
    feat = np.ascontiguousarray(feat, np.float64)
    sh = np.ascontiguousarray(sh, np.float64)
    sigma = C.c_double()
    rgb = np.empty(3, np.float64)
    lib().orc_mlp(ptr(weights, C.c_float), ptr(feat, C.c_double), ptr(sh, C.c_double),
                  C.byref(sigma), ptr(rgb, C.c_double))
    return sigma.value, rgb


# ----------------------------------------------------------------- frame prep

def vchroma_filter(h, max_size=16):
    ch = (h + 1) // 2
    coef = np.zeros(ch * max_size, np.int16)
    pos = np.zeros(ch, np.int32)
    n = lib().orc_sws_vchroma_filter(h, ptr(coef, C.c_int16), ptr(pos, C.c_int32), max_size)
    assert n > 0
    return coef[: ch * n].reshape(ch, n), pos


def rgb24_to_yuv420p(rgb, mode=MODE_X86):
    h, w, _ = rgb.shape
    rgb = np.ascontiguousarray(rgb, np.uint8)
    y = np.zeros((h, w), np.uint8)
    u = np.zeros((h // 2, w // 2), np.uint8)
    v = np.zeros((h // 2, w // 2), np.uint8)
    rc = lib().orc_rgb24_to_yuv420p(ptr(rgb, C.c_uint8), 3 * w, w, h, ptr(y, C.c_uint8), w,
                                    ptr(u, C.c_uint8), w // 2, ptr(v, C.c_uint8), w // 2, mode)
    if rc != 0:
        raise ValueError("orc_rgb24_to_yuv420p failed")
    return y, u, v


def gray8_to_yuv420p(g, mode=MODE_X86):
    h, w = g.shape
    g = np.ascontiguousarray(g, np.uint8)
    y = np.zeros((h, w), np.uint8)
    u = np.zeros((h // 2, w // 2), np.uint8)
    v = np.zeros((h // 2, w // 2), np.uint8)
    rc = lib().orc_gray8_to_yuv420p(ptr(g, C.c_uint8), w, w, h, ptr(y, C.c_uint8), w,
                                    ptr(u, C.c_uint8), w // 2, ptr(v, C.c_uint8), w // 2, mode)
    if rc != 0:
        raise ValueError("orc_gray8_to_yuv420p failed")
    return y, u, v


def composite(rgb_a, depth_a, rgb_b=None, depth_b=None):
    h, w, _ = rgb_a.shape
    out = np.empty((h, w, 3), np.uint8)
    dout = np.empty((h, w), np.float32)
    rgb_a = np.ascontiguousarray(rgb_a, np.uint8)
    depth_a = np.ascontiguousarray(depth_a, np.float32)
    if rgb_b is not None:
        rgb_b = np.ascontiguousarray(rgb_b, np.uint8)
        depth_b = np.ascontiguousarray(depth_b, np.float32)
    lib().orc_composite(ptr(rgb_a, C.c_uint8), ptr(depth_a, C.c_float), ptr(rgb_b, C.c_uint8),
                        ptr(depth_b, C.c_float), w, h, 3 * w, ptr(out, C.c_uint8),
                        ptr(dout, C.c_float))
    return out, dout


class Glyph(C.Structure):
    _fields_ = [("left", C.c_int32), ("top", C.c_int32), ("width", C.c_uint32),
                ("rows", C.c_uint32), ("advance", C.c_int32), ("offset", C.c_uint32)]


def render_string(rgb, glyphs, bitmaps, position, text):
    """render_string_to_frame on an (H, W, 3) uint8 array, in place."""
    assert rgb.dtype == np.uint8 and rgb.flags.c_contiguous and rgb.shape[2] == 3
    h, w, _ = rgb.shape
    tab = (Glyph * 256)(*[Glyph(*g) for g in glyphs])
    bm = np.ascontiguousarray(bitmaps, np.uint8).reshape(-1)
    lib().orc_render_string(rgb.ctypes.data_as(C.POINTER(C.c_uint8)), w, h, tab,
                            bm.ctypes.data_as(C.POINTER(C.c_uint8)), position,
                            text.encode("latin-1"))
    return rgb


def depth_gray8(depth, far=4.0):
    """GRAY8 depth plane of FP32 depth (the library's quantisation rule)."""
    d = np.ascontiguousarray(depth, np.float32)
    out = np.empty(d.shape, np.uint8)
    lib().orc_depth_gray8(ptr(d, C.c_float), d.size, far, ptr(out, C.c_uint8))
    return out


def set_frameprep_threads(n):
    lib().orc_set_frameprep_threads(int(n))


def frameprep(rgb_a, depth_a, rgb_b=None, depth_b=None, mode=MODE_X86):
    comp, _ = composite(rgb_a, depth_a, rgb_b, depth_b)
    return rgb24_to_yuv420p(comp, mode)


# ----------------------------------------------------------------- training

def training_rays(seed, step, n_rays, n_poses=16, width=640, height=480):
    o = np.empty((n_rays, 3), np.float32)
    d = np.empty((n_rays, 3), np.float32)
    pose = np.empty(n_rays, np.uint32)
    pix = np.empty(n_rays, np.uint32)
    lib().orc_training_rays(seed, step, n_rays, n_poses, width, height, ptr(o, C.c_float),
                            ptr(d, C.c_float), ptr(pose, C.c_uint32), ptr(pix, C.c_uint32))
    return o, d, pose, pix


def training_pose(seed, pose):
    m = np.empty(12, np.float32)
    lib().orc_training_pose(seed, pose, ptr(m, C.c_float))
    return m


def loss_grad(table, weights, occ, ray_o, ray_d, target, want_grad=True):
    n = len(ray_o)
    gt = np.zeros(N_TABLE * 2, np.float64) if want_grad else None
    gw = np.zeros(N_WEIGHTS, np.float64) if want_grad else None
    ray_o = np.ascontiguousarray(ray_o, np.float32)
    ray_d = np.ascontiguousarray(ray_d, np.float32)
    target = np.ascontiguousarray(target, np.float32)
    loss = lib().orc_loss_grad(ptr(table, C.c_float), ptr(weights, C.c_float),
                               ptr(occ, C.c_uint32), n, ptr(ray_o, C.c_float),
                               ptr(ray_d, C.c_float), ptr(target, C.c_float),
                               ptr(gt, C.c_double), ptr(gw, C.c_double))
    return loss, gt, gw


def loss_perturbed(table, weights, occ, ray_o, ray_d, target, table_idx=-1, weight_idx=-1,
                   delta=0.0):
    ray_o = np.ascontiguousarray(ray_o, np.float32)
    ray_d = np.ascontiguousarray(ray_d, np.float32)
    target = np.ascontiguousarray(target, np.float32)
    return lib().orc_loss_perturbed(ptr(table, C.c_float), ptr(weights, C.c_float),
                                    ptr(occ, C.c_uint32), len(ray_o), ptr(ray_o, C.c_float),
                                    ptr(ray_d, C.c_float), ptr(target, C.c_float),
                                    table_idx, weight_idx, delta)


def adam_step(p, g, m, v, step, lr=1e-2, beta1=0.9, beta2=0.99, eps=1e-15):
    lib().orc_adam_step(len(p), ptr(p, C.c_double), ptr(g, C.c_double), ptr(m, C.c_double),
                        ptr(v, C.c_double), step, lr, beta1, beta2, eps)
