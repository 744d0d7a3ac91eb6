"""Host-side binding of libnes_render.so (ctypes; no GPU framework needed).

The shipped compute path is the HIP library; this module only marshals
arguments.  It mirrors the reference's operator boundary for the render +
frame-prep path:

* ``Context.render`` replaces one FrameRequest -> RenderedFrame round trip of
  ``socket_client_thread`` (src/server.cpp:140-194);
* ``frameprep`` replaces ``RenderedFrame::convert_frame`` /
  ``SwsContextManager`` (include/base/video/rendered_frame.h:24-33,
  src/base/video/type_managers.cc:143-155).

Errors surface as :class:`NesError` carrying the library's negative errno code;
there is no fallback path -- a missing or broken library raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "lib", "libnes_render.so")

NES_OK = 0
NES_EAGAIN = -11
PRECISION_FP32 = 0
PRECISION_MIXED = 1
PRECISION_FP32_VALU = 2
DEPTH_FAR_DEFAULT = 4.0
SWS_X86 = 0
SWS_C = 1
TABLE_ENTRIES = 6098925
MLP_WEIGHTS = 9344
OCC_WORDS = 65536


class NesError(RuntimeError):
    def __init__(self, code, where, detail=""):
        self.code = code
        super().__init__(f"{where} failed: {code} ({detail})")


class Camera(C.Structure):
    """nes_camera: nesproto::Camera + FrameRequest.index + vertical FOV."""
    _fields_ = [("index", C.c_uint64), ("is_left", C.c_uint32), ("width", C.c_uint32),
                ("height", C.c_uint32), ("c2w", C.c_float * 12), ("fov_y_deg", C.c_float)]


class SceneDesc(C.Structure):
    _fields_ = [("table", C.POINTER(C.c_float)), ("weights", C.POINTER(C.c_float)),
                ("occupancy", C.POINTER(C.c_uint32))]


class ViewOut(C.Structure):
    _fields_ = [("rgb24", C.POINTER(C.c_uint8)), ("depth", C.POINTER(C.c_float)),
                ("rgb_f32", C.POINTER(C.c_float)), ("n_samples", C.POINTER(C.c_uint16)),
                ("depth_gray8", C.POINTER(C.c_uint8)), ("depth_far", C.c_float)]


class RenderStats(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("samples", C.c_uint64), ("ms_total", C.c_double),
                ("ms_eval", C.c_double), ("ms_march", C.c_double), ("ms_composite", C.c_double),
                ("ms_encode", C.c_double), ("ms_mlp", C.c_double), ("composited", C.c_uint64),
                ("terminated_rays", C.c_uint64), ("passes", C.c_uint32)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_batch_rays", C.c_uint32),
                ("max_inflight", C.c_uint32), ("queue_capacity", C.c_uint32),
                ("precision", C.c_int), ("sws_mode", C.c_int), ("font", C.c_void_p),
                ("overlay_flags", C.c_uint32), ("depth_stream", C.c_uint32),
                ("depth_far", C.c_float)]


class FramePlanes(C.Structure):
    """nes_frame_planes: an encoder frame's data[0..2] / linesize[0..2]."""
    _fields_ = [("data", C.POINTER(C.c_uint8) * 3), ("linesize", C.c_int32 * 3)]


class Glyph(C.Structure):
    _fields_ = [("left", C.c_int32), ("top", C.c_int32), ("width", C.c_uint32),
                ("rows", C.c_uint32), ("advance", C.c_int32), ("offset", C.c_uint32)]


class Text(C.Structure):
    _fields_ = [("position", C.c_int32), ("text", C.c_char_p)]


class FrameResult(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("index", C.c_uint64), ("n_views", C.c_uint32),
                ("width", C.c_uint32 * 2), ("height", C.c_uint32 * 2),
                ("y", C.POINTER(C.c_uint8) * 2), ("u", C.POINTER(C.c_uint8) * 2),
                ("v", C.POINTER(C.c_uint8) * 2), ("latency_ms", C.c_double),
                ("queue_ms", C.c_double), ("rays", C.c_uint64), ("samples", C.c_uint64),
                ("status", C.c_int), ("linesize", (C.c_int32 * 3) * 2),
                ("depth_y", C.POINTER(C.c_uint8) * 2), ("depth_u", C.POINTER(C.c_uint8) * 2),
                ("depth_v", C.POINTER(C.c_uint8) * 2), ("depth_linesize", (C.c_int32 * 3) * 2),
                ("batch", C.c_uint64), ("batch_requests", C.c_uint32),
                ("batch_passes", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("completed", C.c_uint64), ("superseded", C.c_uint64), ("failed", C.c_uint64),
                ("rays", C.c_uint64), ("samples", C.c_uint64), ("batches", C.c_uint64),
                ("slots", C.c_uint32), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("mean_ms", C.c_double), ("max_ms", C.c_double), ("wall_s", C.c_double),
                ("mrays_per_s", C.c_double), ("p50_all_ms", C.c_double),
                ("p99_all_ms", C.c_double), ("composited", C.c_uint64),
                ("bounded_batches", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SimConfig(C.Structure):
    _fields_ = [("n_clients", C.c_uint32), ("rate_hz", C.c_double), ("duration_s", C.c_double),
                ("seed", C.c_uint64), ("n_res", C.c_uint32), ("res_wh", C.POINTER(C.c_uint32)),
                ("fov_min", C.c_double), ("fov_max", C.c_double), ("stereo", C.c_int)]


class TrainConfig(C.Structure):
    _fields_ = [("rays_per_batch", C.c_uint32), ("n_poses", C.c_uint32), ("width", C.c_uint32),
                ("height", C.c_uint32), ("fov_y_deg", C.c_float), ("seed", C.c_uint64),
                ("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float)]


class TrainStats(C.Structure):
    _fields_ = [("ms", C.c_double), ("rays", C.c_uint64), ("samples", C.c_uint64),
                ("step", C.c_uint64)]


# C-ABI symbols declared in include/nes_render.h (checked by the CPU tests).
_SIGS = {}


def _sig(name, res, *args):
    _SIGS[name] = (res, list(args))


_P = C.POINTER
_sig("nes_strerror", C.c_char_p, C.c_int)
_sig("nes_last_error", C.c_char_p)
_sig("nes_scene_synthesize", C.c_int, C.c_uint64, C.c_int, _P(C.c_float), _P(C.c_float),
     _P(C.c_uint32))
_sig("nes_orbit_camera", C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _P(C.c_float))
_sig("nes_ctx_create", C.c_int, _P(SceneDesc), C.c_uint64, _P(C.c_void_p))
_sig("nes_ctx_destroy", C.c_int, C.c_void_p)
_sig("nes_ctx_update_scene", C.c_int, C.c_void_p, _P(SceneDesc))
_sig("nes_render", C.c_int, C.c_void_p, C.c_int, _P(Camera), C.c_int, _P(ViewOut), C.c_int,
     _P(RenderStats))
_sig("nes_render_device", C.c_int, C.c_void_p, C.c_int, _P(Camera), C.c_int, C.c_int,
     _P(RenderStats))
_sig("nes_render_fetch", C.c_int, C.c_void_p, C.c_int, _P(ViewOut))
_sig("nes_frameprep", C.c_int, _P(C.c_uint8), _P(C.c_float), _P(C.c_uint8), _P(C.c_float),
     C.c_int, C.c_int, C.c_int, _P(C.c_uint8), C.c_int, _P(C.c_uint8), C.c_int, _P(C.c_uint8),
     C.c_int, C.c_int)
_sig("nes_frameprep_device", C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
     C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int,
     _P(C.c_float))
_sig("nes_frameprep_device_static", C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
     C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
     C.c_int, C.c_void_p)
_sig("nes_gray8_to_yuv420p", C.c_int, _P(C.c_uint8), C.c_int, C.c_int, C.c_int, _P(C.c_uint8),
     C.c_int, _P(C.c_uint8), C.c_int, _P(C.c_uint8), C.c_int, C.c_int)
_sig("nes_sws_vchroma_filter", C.c_int, C.c_int, _P(C.c_int16), _P(C.c_int32), C.c_int,
     _P(C.c_int))
_sig("nes_trainer_create", C.c_int, _P(SceneDesc), _P(TrainConfig), _P(C.c_float), _P(C.c_float),
     _P(C.c_void_p))
_sig("nes_trainer_destroy", C.c_int, C.c_void_p)
_sig("nes_trainer_step", C.c_int, C.c_void_p, _P(C.c_float), _P(TrainStats))
_sig("nes_trainer_batch_grad", C.c_int, C.c_void_p, C.c_uint64, _P(C.c_float), _P(C.c_float),
     _P(C.c_float))
_sig("nes_trainer_get_params", C.c_int, C.c_void_p, _P(C.c_float), _P(C.c_float))
_sig("nes_trainer_set_params", C.c_int, C.c_void_p, _P(C.c_float), _P(C.c_float))
_sig("nes_trainer_forward", C.c_int, C.c_void_p, _P(C.c_float), _P(C.c_float), C.c_uint32,
     _P(C.c_float))
_sig("nes_trainer_backward", C.c_int, C.c_void_p, _P(C.c_float), _P(C.c_float), C.c_uint32,
     _P(C.c_float), _P(C.c_float), _P(C.c_float))
_sig("nes_trainer_save", C.c_int, C.c_void_p, C.c_char_p)
_sig("nes_trainer_load", C.c_int, C.c_void_p, C.c_char_p)
_sig("nes_wire_decode_request", C.c_int, C.c_char_p, C.c_size_t, C.c_float, _P(Camera))
_sig("nes_wire_encode_frame", C.c_int, _P(Camera), _P(C.c_uint8), _P(C.c_uint8), C.c_char_p,
     C.c_size_t, _P(C.c_size_t))
_sig("nes_wire_server_start", C.c_int, C.c_void_p, C.c_char_p, C.c_uint16, C.c_int, C.c_float,
     C.c_float, _P(C.c_void_p))
_sig("nes_wire_server_port", C.c_int, C.c_void_p)
_sig("nes_wire_server_served", C.c_uint64, C.c_void_p)
_sig("nes_wire_server_stop", C.c_int, C.c_void_p)
_sig("nes_font_create", C.c_int, _P(Glyph), _P(C.c_uint8), C.c_size_t, _P(C.c_void_p))
_sig("nes_font_destroy", C.c_int, C.c_void_p)
_sig("nes_overlay_device", C.c_int, C.c_void_p, _P(Text), C.c_int, C.c_void_p, C.c_int, C.c_int,
     C.c_int)
_sig("nes_overlay_texts", C.c_int, _P(Camera), C.c_char_p, C.c_uint32, C.c_char_p, C.c_size_t)
_sig("nes_frameprep_overlay", C.c_int, C.c_void_p, _P(Camera), C.c_char_p, C.c_uint32,
     _P(C.c_uint8), _P(C.c_float), _P(C.c_uint8), _P(C.c_float), C.c_int, C.c_int, C.c_int,
     _P(C.c_uint8), C.c_int, _P(C.c_uint8), C.c_int, _P(C.c_uint8), C.c_int, C.c_int)
_sig("nes_server_create", C.c_int, C.c_void_p, _P(ServerConfig), _P(C.c_void_p))
_sig("nes_server_destroy", C.c_int, C.c_void_p)
_sig("nes_server_submit", C.c_int, C.c_void_p, C.c_uint32, _P(Camera), C.c_int, _P(C.c_uint64))
_sig("nes_server_submit_into", C.c_int, C.c_void_p, C.c_uint32, _P(Camera), C.c_int,
     _P(FramePlanes), _P(FramePlanes), _P(C.c_uint64))
_sig("nes_server_hold", C.c_int, C.c_void_p, C.c_int)
_sig("nes_server_wait", C.c_int, C.c_void_p, C.c_uint64, C.c_int, _P(FrameResult))
_sig("nes_server_release", C.c_int, C.c_void_p, C.c_uint64)
_sig("nes_server_stats_get", C.c_int, C.c_void_p, _P(ServerStats))
_sig("nes_server_reset_stats", C.c_int, C.c_void_p)
_sig("nes_server_simulate", C.c_int, C.c_void_p, _P(SimConfig), _P(ServerStats))
_sig("nes_device_alloc", C.c_int, C.c_size_t, _P(C.c_void_p))
_sig("nes_device_free", C.c_int, C.c_void_p)
_sig("nes_host_alloc", C.c_int, C.c_size_t, _P(C.c_void_p))
_sig("nes_host_free", C.c_int, C.c_void_p)
_sig("nes_memcpy_h2d", C.c_int, C.c_void_p, C.c_void_p, C.c_size_t)
_sig("nes_memcpy_d2h", C.c_int, C.c_void_p, C.c_void_p, C.c_size_t)
_sig("nes_device_copy", C.c_int, C.c_void_p, C.c_void_p, C.c_size_t, _P(C.c_float))
_sig("nes_device_info", C.c_int, C.c_char_p, C.c_size_t, _P(C.c_int), _P(C.c_uint64))
_sig("nes_device_gather_rate", C.c_int, C.c_size_t, _P(C.c_float), _P(C.c_uint64))
_sig("nes_device_lds_atomic_rate", C.c_int, C.c_int, _P(C.c_float), _P(C.c_uint64))

_lib = None


def lib():
    """Load libnes_render.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} not built; run ngp-encode-server_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            if hasattr(L, name):
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
        _lib = L
    return _lib


def check(rc, where):
    if rc != NES_OK:
        detail = lib().nes_last_error().decode(errors="replace")
        raise NesError(rc, where, f"{lib().nes_strerror(rc).decode()}; {detail}")


def _ptr(a, t):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(t))


# ----------------------------------------------------------------- scene

def scene_synthesize(seed, table_init=0):
    table = np.empty(TABLE_ENTRIES * 2, np.float32)
    weights = np.empty(MLP_WEIGHTS, np.float32)
    occ = np.empty(OCC_WORDS, np.uint32)
    check(lib().nes_scene_synthesize(seed, table_init, _ptr(table, C.c_float),
                                     _ptr(weights, C.c_float), _ptr(occ, C.c_uint32)),
          "nes_scene_synthesize")
    return table, weights, occ


def orbit_camera(azimuth, elevation, radius=2.0, eye_offset=0.0):
    m = np.empty(12, np.float32)
    check(lib().nes_orbit_camera(azimuth, elevation, radius, eye_offset, _ptr(m, C.c_float)),
          "nes_orbit_camera")
    return m


def camera(width, height, c2w, fov=60.0, index=0, is_left=1):
    cam = Camera()
    cam.index, cam.is_left, cam.width, cam.height = index, is_left, width, height
    cam.c2w[:] = [float(x) for x in c2w]
    cam.fov_y_deg = fov
    return cam


class Context:
    """A device-resident scene (nes_ctx)."""

    def __init__(self, table, weights, occ, hbm_budget=0):
        self._keep = (np.ascontiguousarray(table, np.float32),
                      np.ascontiguousarray(weights, np.float32),
                      np.ascontiguousarray(occ, np.uint32))
        desc = SceneDesc(_ptr(self._keep[0], C.c_float), _ptr(self._keep[1], C.c_float),
                         _ptr(self._keep[2], C.c_uint32))
        h = C.c_void_p()
        check(lib().nes_ctx_create(C.byref(desc), hbm_budget, C.byref(h)), "nes_ctx_create")
        self.handle = h

    def update_scene(self, table, weights, occ):
        keep = (np.ascontiguousarray(table, np.float32), np.ascontiguousarray(weights, np.float32),
                np.ascontiguousarray(occ, np.uint32))
        desc = SceneDesc(_ptr(keep[0], C.c_float), _ptr(keep[1], C.c_float),
                         _ptr(keep[2], C.c_uint32))
        check(lib().nes_ctx_update_scene(self.handle, C.byref(desc)), "nes_ctx_update_scene")
        self._keep = keep

    def render(self, cams, precision=PRECISION_MIXED, steps_per_pass=0, depth_far=0.0):
        """Render views; returns ([dict per view], RenderStats).  Each dict
        holds rgb24, depth, rgb (float), n_samples and depth_gray8 (the
        RenderedFrame depth plane, quantised against depth_far)."""
        if isinstance(cams, Camera):
            cams = [cams]
        n = len(cams)
        arr = (Camera * n)(*cams)
        outs = (ViewOut * n)()
        results = []
        for i, c in enumerate(cams):
            r = {"rgb24": np.empty((c.height, c.width, 3), np.uint8),
                 "depth": np.empty((c.height, c.width), np.float32),
                 "rgb": np.empty((c.height, c.width, 3), np.float32),
                 "n_samples": np.empty((c.height, c.width), np.uint16),
                 "depth_gray8": np.empty((c.height, c.width), np.uint8)}
            outs[i] = ViewOut(_ptr(r["rgb24"], C.c_uint8), _ptr(r["depth"], C.c_float),
                              _ptr(r["rgb"], C.c_float), _ptr(r["n_samples"], C.c_uint16),
                              _ptr(r["depth_gray8"], C.c_uint8), depth_far)
            results.append(r)
        st = RenderStats()
        check(lib().nes_render(self.handle, precision, arr, n, outs, steps_per_pass, C.byref(st)),
              "nes_render")
        return results, st

    def render_device(self, cams, precision=PRECISION_MIXED, steps_per_pass=0):
        if isinstance(cams, Camera):
            cams = [cams]
        arr = (Camera * len(cams))(*cams)
        st = RenderStats()
        check(lib().nes_render_device(self.handle, precision, arr, len(cams), steps_per_pass,
                                      C.byref(st)), "nes_render_device")
        return st

    def close(self):
        if self.handle:
            check(lib().nes_ctx_destroy(self.handle), "nes_ctx_destroy")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ frame prep

def frameprep_pitched(rgb_rows_a, w, depth_a=None, rgb_rows_b=None, depth_b=None, mode=SWS_X86,
                      y_stride=None, uv_stride=None):
    """Nearer-depth composite + RGB24 -> YUV420P with row pitches, like
    sws_scale on planes with linesizes.  rgb_rows_*: (h, rgb_stride) uint8,
    pixels in the first 3w bytes of each row.  Returns the whole padded
    (h, y_stride) and (h/2, uv_stride) output arrays, pre-filled with 0xA5."""
    h, rgb_stride = rgb_rows_a.shape
    y_stride = y_stride or w
    uv_stride = uv_stride or w // 2
    ra = np.ascontiguousarray(rgb_rows_a, np.uint8)
    rb = None if rgb_rows_b is None else np.ascontiguousarray(rgb_rows_b, np.uint8)
    da = None if depth_a is None else np.ascontiguousarray(depth_a, np.float32)
    db = None if depth_b is None else np.ascontiguousarray(depth_b, np.float32)
    y = np.full((h, y_stride), 0xA5, np.uint8)
    u = np.full((h // 2, uv_stride), 0xA5, np.uint8)
    v = np.full((h // 2, uv_stride), 0xA5, np.uint8)
    check(lib().nes_frameprep(_ptr(ra, C.c_uint8), _ptr(da, C.c_float), _ptr(rb, C.c_uint8),
                              _ptr(db, C.c_float), w, h, rgb_stride, _ptr(y, C.c_uint8), y_stride,
                              _ptr(u, C.c_uint8), uv_stride, _ptr(v, C.c_uint8), uv_stride, mode),
          "nes_frameprep")
    return y, u, v


def frameprep(rgb_a, depth_a=None, rgb_b=None, depth_b=None, mode=SWS_X86):
    """Nearer-depth composite + RGB24 -> YUV420P (host arrays); returns (y, u, v)."""
    h, w, _ = rgb_a.shape
    rgb_a = np.ascontiguousarray(rgb_a, np.uint8)
    if depth_a is not None:
        depth_a = np.ascontiguousarray(depth_a, np.float32)
    if rgb_b is not None:
        rgb_b = np.ascontiguousarray(rgb_b, np.uint8)
        depth_b = np.ascontiguousarray(depth_b, np.float32)
    y = np.empty((h, w), np.uint8)
    u = np.empty((h // 2, w // 2), np.uint8)
    v = np.empty((h // 2, w // 2), np.uint8)
    check(lib().nes_frameprep(_ptr(rgb_a, C.c_uint8), _ptr(depth_a, C.c_float),
                              _ptr(rgb_b, C.c_uint8), _ptr(depth_b, C.c_float), w, h, 3 * w,
                              _ptr(y, C.c_uint8), w, _ptr(u, C.c_uint8), w // 2,
                              _ptr(v, C.c_uint8), w // 2, mode), "nes_frameprep")
    return y, u, v


# ------------------------------------------------------------ text overlay

TEXT_LEFT_TOP, TEXT_LEFT_BOTTOM, TEXT_RIGHT_TOP, TEXT_RIGHT_BOTTOM, TEXT_CENTER = range(5)
OVERLAY_MATRIX, OVERLAY_INDEX, OVERLAY_TIMESTAMP, OVERLAY_DIRECTION = 1, 2, 4, 8
OVERLAY_ALL = 15
FONT_GLYPHS = 256


class Font:
    """Glyph table uploaded once (RenderTextContext's FreeType face, rasterised
    by the caller).  glyphs: FONT_GLYPHS (left, top, width, rows, advance,
    offset) tuples; bitmaps: the uint8 atlas the offsets point into."""

    def __init__(self, glyphs, bitmaps):
        if len(glyphs) != FONT_GLYPHS:
            raise ValueError("need %d glyphs" % FONT_GLYPHS)
        self.glyphs = (Glyph * FONT_GLYPHS)(*[Glyph(*g) for g in glyphs])
        self.bitmaps = np.ascontiguousarray(bitmaps, np.uint8).reshape(-1)
        h = C.c_void_p()
        check(lib().nes_font_create(self.glyphs, _ptr(self.bitmaps, C.c_uint8),
                                    self.bitmaps.size, C.byref(h)), "nes_font_create")
        self.handle = h

    def close(self):
        if self.handle:
            check(lib().nes_font_destroy(self.handle), "nes_font_destroy")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def overlay_device(font, texts, rgb_dev, w, h, stride=None):
    """render_string_to_frame for [(position, str)] in place on a device
    RGB24 frame (DeviceBuffer or raw pointer)."""
    arr = (Text * max(len(texts), 1))(*[Text(p, t.encode()) for p, t in texts])
    ptr = rgb_dev.ptr if isinstance(rgb_dev, DeviceBuffer) else rgb_dev
    check(lib().nes_overlay_device(font.handle, arr, len(texts), ptr, w, h,
                                   3 * w if stride is None else stride), "nes_overlay_device")


def overlay_texts(cam, timestamp=None, flags=OVERLAY_ALL):
    """The strings process_frame_thread draws for cam, in the order matrix,
    index, timestamp, direction (those selected by flags)."""
    buf = C.create_string_buffer(4096)
    n = lib().nes_overlay_texts(C.byref(cam), None if timestamp is None else timestamp.encode(),
                                flags, buf, len(buf))
    check(n if n < 0 else NES_OK, "nes_overlay_texts")
    return [t.decode() for t in buf.raw[:n].split(b"\0")[:-1]]


def frameprep_overlay(font, cam, rgb_a, depth_a=None, rgb_b=None, depth_b=None, flags=OVERLAY_ALL,
                      timestamp=None, mode=SWS_X86):
    """Composite, text overlay of cam's strings, RGB24 -> YUV420P; (y, u, v)."""
    h, w, _ = rgb_a.shape
    rgb_a = np.ascontiguousarray(rgb_a, np.uint8)
    if depth_a is not None:
        depth_a = np.ascontiguousarray(depth_a, np.float32)
    if rgb_b is not None:
        rgb_b = np.ascontiguousarray(rgb_b, np.uint8)
        depth_b = np.ascontiguousarray(depth_b, np.float32)
    y = np.empty((h, w), np.uint8)
    u = np.empty((h // 2, w // 2), np.uint8)
    v = np.empty((h // 2, w // 2), np.uint8)
    check(lib().nes_frameprep_overlay(
        font.handle, C.byref(cam), None if timestamp is None else timestamp.encode(), flags,
        _ptr(rgb_a, C.c_uint8), _ptr(depth_a, C.c_float), _ptr(rgb_b, C.c_uint8),
        _ptr(depth_b, C.c_float), w, h, 3 * w, _ptr(y, C.c_uint8), w, _ptr(u, C.c_uint8), w // 2,
        _ptr(v, C.c_uint8), w // 2, mode), "nes_frameprep_overlay")
    return y, u, v


def gray8_to_yuv420p(g, mode=SWS_X86):
    h, w = g.shape
    g = np.ascontiguousarray(g, np.uint8)
    y = np.empty((h, w), np.uint8)
    u = np.empty((h // 2, w // 2), np.uint8)
    v = np.empty((h // 2, w // 2), np.uint8)
    check(lib().nes_gray8_to_yuv420p(_ptr(g, C.c_uint8), w, h, w, _ptr(y, C.c_uint8), w,
                                     _ptr(u, C.c_uint8), w // 2, _ptr(v, C.c_uint8), w // 2,
                                     mode), "nes_gray8_to_yuv420p")
    return y, u, v


def vchroma_filter(h, max_taps=16):
    ch = h // 2
    coef = np.zeros(ch * max_taps, np.int16)
    pos = np.zeros(ch, np.int32)
    taps = C.c_int()
    check(lib().nes_sws_vchroma_filter(h, _ptr(coef, C.c_int16), _ptr(pos, C.c_int32), max_taps,
                                       C.byref(taps)), "nes_sws_vchroma_filter")
    return coef[: ch * taps.value].reshape(ch, taps.value), pos


class DeviceBuffer:
    """Raw device allocation through the library (no other GPU runtime)."""

    def __init__(self, nbytes):
        p = C.c_void_p()
        check(lib().nes_device_alloc(nbytes, C.byref(p)), "nes_device_alloc")
        self.ptr, self.nbytes = p, nbytes

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        check(lib().nes_memcpy_h2d(self.ptr, arr.ctypes.data_as(C.c_void_p), arr.nbytes),
              "nes_memcpy_h2d")

    def download(self, arr):
        assert arr.nbytes <= self.nbytes and arr.flags["C_CONTIGUOUS"]
        check(lib().nes_memcpy_d2h(arr.ctypes.data_as(C.c_void_p), self.ptr, arr.nbytes),
              "nes_memcpy_d2h")
        return arr

    def free(self):
        if self.ptr:
            check(lib().nes_device_free(self.ptr), "nes_device_free")
            self.ptr = None


def frameprep_device(rgb_a, depth_a, rgb_b, depth_b, w, h, n_frames, y, u, v, mode=SWS_X86,
                     y_stride=None, uv_stride=None):
    """Device-pointer frame prep over n_frames packed frames (output planes
    y_stride * h and uv_stride * h/2 bytes apart); returns kernel ms."""
    ms = C.c_float()
    vp = lambda b: b.ptr if b is not None else None
    ys, uvs = y_stride or w, uv_stride or w // 2
    check(lib().nes_frameprep_device(vp(rgb_a), vp(depth_a), vp(rgb_b), vp(depth_b), w, h,
                                     n_frames, vp(y), ys, vp(u), uvs, vp(v), uvs, mode,
                                     C.byref(ms)), "nes_frameprep_device")
    return ms.value


def frameprep_device_static(rgb_a, depth_a, rgb_b, depth_b, w, h, n_frames, y, u, v,
                            mode=SWS_X86, y_stride=None, uv_stride=None):
    """frameprep_device through nes_frameprep_device_static (no device
    allocation; null stream, returns after completion)."""
    vp = lambda b: b.ptr if b is not None else None
    ys, uvs = y_stride or w, uv_stride or w // 2
    check(lib().nes_frameprep_device_static(vp(rgb_a), vp(depth_a), vp(rgb_b), vp(depth_b), w, h,
                                            n_frames, vp(y), ys, vp(u), uvs, vp(v), uvs, mode,
                                            None), "nes_frameprep_device_static")


def device_copy(dst, src, nbytes):
    """dst/src: DeviceBuffer; returns the copy time in ms."""
    ms = C.c_float()
    check(lib().nes_device_copy(dst.ptr, src.ptr, nbytes, C.byref(ms)), "nes_device_copy")
    return ms.value


def gather_rate(region_bytes):
    """Random one-address-per-lane 4-byte gathers within a per-XCD region:
    (lane loads per second, kernel ms)."""
    ms = C.c_float()
    n = C.c_uint64()
    check(lib().nes_device_gather_rate(region_bytes, C.byref(ms), C.byref(n)),
          "nes_device_gather_rate")
    return n.value / (ms.value * 1e-3), ms.value


LDS_ATOMIC_KINDS = {"f32": 0, "f64": 1, "u32": 2, "f32_cas": 3, "f32x2_cas": 4}


def lds_atomic_rate(kind):
    """LDS atomic adds per second for the chip (kind "f32", "f64", "u32",
    "f32_cas", an FP32 add by compare-and-swap, or "f32x2_cas", two FP32
    adds by one 64-bit compare-and-swap, counted once),
    every lane of one workgroup per CU adding at random addresses."""
    ms = C.c_float()
    n = C.c_uint64()
    check(lib().nes_device_lds_atomic_rate(LDS_ATOMIC_KINDS[kind], C.byref(ms), C.byref(n)),
          "nes_device_lds_atomic_rate")
    return n.value / (ms.value * 1e-3)


def device_info():
    name = C.create_string_buffer(64)
    cus = C.c_int()
    hbm = C.c_uint64()
    check(lib().nes_device_info(name, 64, C.byref(cus), C.byref(hbm)), "nes_device_info")
    return name.value.decode(), cus.value, hbm.value


# --------------------------------------------------------------- serving

NES_ECANCELED = -125


class HostPlanes:
    """YUV420P planes of one w x h frame in pinned host memory
    (nes_host_alloc) with encoder linesizes: FFALIGN(width, align) for Y and
    FFALIGN(width / 2, align) for U and V, as av_image_alloc(..., align)
    lays them out (src/base/video/type_managers.cc:116-141)."""

    def __init__(self, w, h, align=32):
        a = lambda x: (x + align - 1) // align * align
        self.w, self.h = w, h
        self.linesize = (a(w), a(w // 2), a(w // 2))
        self.rows = (h, h // 2, h // 2)
        sizes = [ls * r for ls, r in zip(self.linesize, self.rows)]
        p = C.c_void_p()
        check(lib().nes_host_alloc(sum(sizes), C.byref(p)), "nes_host_alloc")
        self.ptr, self.nbytes = p, sum(sizes)
        base = C.cast(p, C.POINTER(C.c_uint8))
        raw = np.ctypeslib.as_array(base, (self.nbytes,))
        raw[:] = 0xA5  # padding bytes must survive the server's writes
        self.raw = raw
        offs = np.cumsum([0] + sizes[:-1])
        self.planes = FramePlanes()
        for i in range(3):
            self.planes.data[i] = C.cast(C.c_void_p(p.value + int(offs[i])), C.POINTER(C.c_uint8))
            self.planes.linesize[i] = self.linesize[i]
        self.views = [raw[o:o + s].reshape(r, ls) for o, s, r, ls in
                      zip(offs, sizes, self.rows, self.linesize)]

    def yuv(self):
        """(y, u, v) of the frame (copies, without the row padding)."""
        return (self.views[0][:, :self.w].copy(), self.views[1][:, :self.w // 2].copy(),
                self.views[2][:, :self.w // 2].copy())

    def padding_intact(self):
        return all((v[:, n:] == 0xA5).all() for v, n in
                   zip(self.views, (self.w, self.w // 2, self.w // 2)))

    def free(self):
        if self.ptr:
            check(lib().nes_host_free(self.ptr), "nes_host_free")
            self.ptr = None


def _copy_planes(ptrs, lines, w, h):
    out = []
    for p, ls, (pw, ph) in zip(ptrs, lines, ((w, h), (w // 2, h // 2), (w // 2, h // 2))):
        a = np.ctypeslib.as_array(p, (ph, ls))
        out.append(a[:, :pw].copy())
    return tuple(out)


class Server:
    """Continuous-batching render server over a Context (nes_server)."""

    def __init__(self, ctx, hbm_budget=0, max_batch_rays=0, max_inflight=0, queue_capacity=0,
                 precision=PRECISION_MIXED, sws_mode=SWS_X86, font=None, overlay_flags=0,
                 depth_stream=False, depth_far=0.0):
        self.ctx = ctx
        self.font = font  # must outlive the server
        self.depth_stream = bool(depth_stream)
        cfg = ServerConfig(hbm_budget, max_batch_rays, max_inflight, queue_capacity, precision,
                           sws_mode, font.handle if font is not None else None, overlay_flags,
                           int(depth_stream), depth_far)
        h = C.c_void_p()
        check(lib().nes_server_create(ctx.handle, C.byref(cfg), C.byref(h)), "nes_server_create")
        self.handle = h

    def submit(self, client, views):
        if isinstance(views, Camera):
            views = [views]
        arr = (Camera * len(views))(*views)
        t = C.c_uint64()
        check(lib().nes_server_submit(self.handle, client, arr, len(views), C.byref(t)),
              "nes_server_submit")
        return t.value

    def submit_into(self, client, views, planes, depth_planes=None):
        """Like submit, the planes written into caller-owned HostPlanes."""
        if isinstance(views, Camera):
            views = [views]
        arr = (Camera * len(views))(*views)
        pa = (FramePlanes * len(views))(*[p.planes for p in planes])
        da = None if depth_planes is None else (FramePlanes * len(views))(
            *[p.planes for p in depth_planes])
        t = C.c_uint64()
        check(lib().nes_server_submit_into(self.handle, client, arr, len(views), pa, da,
                                           C.byref(t)), "nes_server_submit_into")
        return t.value

    def hold(self, on=True):
        check(lib().nes_server_hold(self.handle, int(on)), "nes_server_hold")

    def wait(self, ticket, timeout_ms=-1, release=True, depth=False):
        """Returns (status, [(y, u, v) per view] copied out, FrameResult); with
        depth=True also the depth stream: (status, scene, depth, result)."""
        r = FrameResult()
        rc = lib().nes_server_wait(self.handle, ticket, timeout_ms, C.byref(r))
        planes, dplanes = [], []
        if rc == NES_OK:
            for i in range(r.n_views):
                w, h = r.width[i], r.height[i]
                planes.append(_copy_planes((r.y[i], r.u[i], r.v[i]), r.linesize[i], w, h))
                if self.depth_stream:
                    dplanes.append(_copy_planes((r.depth_y[i], r.depth_u[i], r.depth_v[i]),
                                                r.depth_linesize[i], w, h))
        elif rc not in (NES_ECANCELED, NES_EAGAIN):
            check(rc, "nes_server_wait")
        if release and rc != NES_EAGAIN:
            check(lib().nes_server_release(self.handle, ticket), "nes_server_release")
        if depth:
            return rc, planes, dplanes, r
        return rc, planes, r

    def stats(self):
        st = ServerStats()
        check(lib().nes_server_stats_get(self.handle, C.byref(st)), "nes_server_stats_get")
        return st

    def simulate(self, n_clients=8, rate_hz=72.0, duration_s=2.0, seed=4,
                 resolutions=((640, 480), (1440, 1600), (2064, 2208)), fov=(60.0, 90.0),
                 stereo=True):
        res = np.ascontiguousarray(np.array(resolutions, np.uint32).reshape(-1))
        sc = SimConfig(n_clients, rate_hz, duration_s, seed, len(resolutions),
                       res.ctypes.data_as(C.POINTER(C.c_uint32)), fov[0], fov[1], int(stereo))
        st = ServerStats()
        check(lib().nes_server_simulate(self.handle, C.byref(sc), C.byref(st)),
              "nes_server_simulate")
        return st

    def close(self):
        if self.handle:
            check(lib().nes_server_destroy(self.handle), "nes_server_destroy")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# --------------------------------------------------------------- training

class Trainer:
    """Differentiable render + Adam over the hash-grid scene (nes_trainer)."""

    def __init__(self, table, weights, occ, poses=None, targets=None, width=0, height=0,
                 rays_per_batch=1 << 18, fov=60.0, seed=0, lr=1e-2, beta1=0.9, beta2=0.99,
                 eps=1e-15):
        keep = (np.ascontiguousarray(table, np.float32), np.ascontiguousarray(weights, np.float32),
                np.ascontiguousarray(occ, np.uint32))
        desc = SceneDesc(_ptr(keep[0], C.c_float), _ptr(keep[1], C.c_float),
                         _ptr(keep[2], C.c_uint32))
        n_poses = 0 if poses is None else len(poses)
        cfg = TrainConfig(rays_per_batch, n_poses, width, height, fov, seed, lr, beta1, beta2, eps)
        p = None if poses is None else np.ascontiguousarray(poses, np.float32).reshape(-1)
        t = None if targets is None else np.ascontiguousarray(targets, np.float32).reshape(-1)
        h = C.c_void_p()
        check(lib().nes_trainer_create(C.byref(desc), C.byref(cfg), _ptr(p, C.c_float),
                                       _ptr(t, C.c_float), C.byref(h)), "nes_trainer_create")
        self.handle = h

    def step(self):
        loss = C.c_float()
        st = TrainStats()
        check(lib().nes_trainer_step(self.handle, C.byref(loss), C.byref(st)), "nes_trainer_step")
        return loss.value, st

    def batch_grad(self, step):
        loss = C.c_float()
        gt = np.empty(TABLE_ENTRIES * 2, np.float32)
        gw = np.empty(MLP_WEIGHTS, np.float32)
        check(lib().nes_trainer_batch_grad(self.handle, step, C.byref(loss), _ptr(gt, C.c_float),
                                           _ptr(gw, C.c_float)), "nes_trainer_batch_grad")
        return loss.value, gt, gw

    def get_params(self):
        t = np.empty(TABLE_ENTRIES * 2, np.float32)
        w = np.empty(MLP_WEIGHTS, np.float32)
        check(lib().nes_trainer_get_params(self.handle, _ptr(t, C.c_float), _ptr(w, C.c_float)),
              "nes_trainer_get_params")
        return t, w

    def set_params(self, table=None, weights=None):
        t = None if table is None else np.ascontiguousarray(table, np.float32)
        w = None if weights is None else np.ascontiguousarray(weights, np.float32)
        check(lib().nes_trainer_set_params(self.handle, _ptr(t, C.c_float), _ptr(w, C.c_float)),
              "nes_trainer_set_params")

    def forward(self, ray_o, ray_d):
        o = np.ascontiguousarray(ray_o, np.float32)
        d = np.ascontiguousarray(ray_d, np.float32)
        rgb = np.empty((len(o), 3), np.float32)
        check(lib().nes_trainer_forward(self.handle, _ptr(o, C.c_float), _ptr(d, C.c_float), len(o),
                                        _ptr(rgb, C.c_float)), "nes_trainer_forward")
        return rgb

    def backward(self, ray_o, ray_d, d_rgb):
        """Gradients of sum(d_rgb * rgb(ray_o, ray_d)) (the forward of these
        rays is recomputed inside)."""
        o = np.ascontiguousarray(ray_o, np.float32)
        d = np.ascontiguousarray(ray_d, np.float32)
        g = np.ascontiguousarray(d_rgb, np.float32)
        if g.shape != (len(o), 3) or d.shape != o.shape:
            raise ValueError("ray_o, ray_d and d_rgb must be n x 3")
        gt = np.empty(TABLE_ENTRIES * 2, np.float32)
        gw = np.empty(MLP_WEIGHTS, np.float32)
        check(lib().nes_trainer_backward(self.handle, _ptr(o, C.c_float), _ptr(d, C.c_float), len(o),
                                         _ptr(g, C.c_float), _ptr(gt, C.c_float),
                                         _ptr(gw, C.c_float)), "nes_trainer_backward")
        return gt, gw

    def save(self, path):
        check(lib().nes_trainer_save(self.handle, os.fsencode(path)), "nes_trainer_save")

    def load(self, path):
        check(lib().nes_trainer_load(self.handle, os.fsencode(path)), "nes_trainer_load")

    def close(self):
        if self.handle:
            check(lib().nes_trainer_destroy(self.handle), "nes_trainer_destroy")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# --------------------------------------------------------- wire protocol

def wire_decode_request(msg, default_fov=60.0):
    cam = Camera()
    check(lib().nes_wire_decode_request(msg, len(msg), default_fov, C.byref(cam)),
          "nes_wire_decode_request")
    return cam


def wire_encode_frame(cam, rgb24=None, gray8=None):
    r = None if rgb24 is None else np.ascontiguousarray(rgb24, np.uint8)
    g = None if gray8 is None else np.ascontiguousarray(gray8, np.uint8)
    n = C.c_size_t()
    check(lib().nes_wire_encode_frame(C.byref(cam), _ptr(r, C.c_uint8), _ptr(g, C.c_uint8), None, 0,
                                      C.byref(n)), "nes_wire_encode_frame")
    buf = C.create_string_buffer(n.value)
    check(lib().nes_wire_encode_frame(C.byref(cam), _ptr(r, C.c_uint8), _ptr(g, C.c_uint8), buf,
                                      n.value, C.byref(n)), "nes_wire_encode_frame")
    return buf.raw[: n.value]


class WireServer:
    """neserver-compatible renderer endpoint (length-prefixed nes.proto over TCP)."""

    def __init__(self, ctx, port=0, addr="127.0.0.1", precision=PRECISION_MIXED, default_fov=60.0,
                 depth_far=4.0):
        h = C.c_void_p()
        check(lib().nes_wire_server_start(ctx.handle, addr.encode(), port, precision, default_fov,
                                          depth_far, C.byref(h)), "nes_wire_server_start")
        self.handle = h
        self.port = lib().nes_wire_server_port(h)

    def served(self):
        return lib().nes_wire_server_served(self.handle)

    def close(self):
        if self.handle:
            check(lib().nes_wire_server_stop(self.handle), "nes_wire_server_stop")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
