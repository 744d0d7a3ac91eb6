"""The render boundary's two outputs as the reference consumes them:
RenderedFrame{RGB24, GRAY8 depth} (proto/nes.proto:18-25, src/server.cpp:
193-194), and YUV420P written with the encoder's frame layout (linesizes
FFALIGN(w, 32), src/base/video/type_managers.cc:116-141), bit-exact against
the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP32, MIXED, FP32_VALU = 0, 1, 2


def ffalign(x, a=32):
    return (x + a - 1) // a * a


def cams(nes, oracle, w, h, az=30.0, el=20.0, eye=0.0):
    c2w = nes.orbit_camera(az, el, 2.0, eye)
    return nes.camera(w, h, c2w), oracle.camera(w, h, c2w)


@pytest.mark.parametrize("prec", [FP32, MIXED])
@pytest.mark.parametrize("far", [0.0, 2.5])
def test_render_gray8_depth_bit_exact(nes, oracle, gpu_ctx, prec, far):
    """The GRAY8 plane returned by the render equals the oracle's quantiser
    applied to the render's own FP32 depth; far 0 selects the default 4."""
    cam, _ = cams(nes, oracle, 320, 240)
    (o,), _ = gpu_ctx.render(cam, precision=prec, depth_far=far)
    want = oracle.depth_gray8(o["depth"], far if far > 0 else nes.DEPTH_FAR_DEFAULT)
    assert np.array_equal(o["depth_gray8"], want)
    assert o["depth_gray8"].max() > 0  # the blob is in view


def test_render_gray8_depth_stream_matches_oracle_conversion(nes, oracle, gpu_ctx):
    """GRAY8 depth from the render, converted like RenderedFrame's depth
    stream (rendered_frame.h:30-31), equals the oracle's GRAY8 -> YUV420P."""
    cam, _ = cams(nes, oracle, 640, 480, az=70.0)
    (o,), _ = gpu_ctx.render(cam, precision=MIXED)
    for mode in (0, 1):
        got = nes.gray8_to_yuv420p(o["depth_gray8"], mode)
        ref = oracle.gray8_to_yuv420p(o["depth_gray8"], mode)
        for g, r in zip(got, ref):
            assert np.array_equal(g, r)


@pytest.mark.parametrize("w,h", [(1440, 1600), (2064, 2208), (1280, 720)])
@pytest.mark.parametrize("mode", [0, 1])
def test_encoder_linesizes_host_path(nes, oracle, w, h, mode):
    """Output planes at the encoder's linesizes (2064 -> 2080 / 1056, 1440 ->
    1440 / 736) and RGB rows with a pitch above 3w: pixels bit-exact, row
    padding untouched."""
    rng = np.random.default_rng(w + h + mode)
    pitch = 3 * w + 64
    rows_a = rng.integers(0, 256, (h, pitch), dtype=np.uint8)
    rows_b = rng.integers(0, 256, (h, pitch), dtype=np.uint8)
    da = rng.uniform(0.1, 10.0, (h, w)).astype(np.float32)
    db = rng.uniform(0.1, 10.0, (h, w)).astype(np.float32)
    ys, uvs = ffalign(w), ffalign(w // 2)
    y, u, v = nes.frameprep_pitched(rows_a, w, da, rows_b, db, mode=mode, y_stride=ys,
                                    uv_stride=uvs)
    a = np.ascontiguousarray(rows_a[:, :3 * w]).reshape(h, w, 3)
    b = np.ascontiguousarray(rows_b[:, :3 * w]).reshape(h, w, 3)
    ry, ru, rv = oracle.frameprep(a, da, b, db, mode=mode)
    assert np.array_equal(y[:, :w], ry)
    assert np.array_equal(u[:, :w // 2], ru) and np.array_equal(v[:, :w // 2], rv)
    assert (y[:, w:] == 0xA5).all() and (u[:, w // 2:] == 0xA5).all() and (v[:, w // 2:] == 0xA5).all()


@pytest.mark.parametrize("w,h", [(2064, 2208), (1440, 1600)])
def test_encoder_linesizes_device_batch(nes, oracle, w, h):
    """Device frame prep of 3 frames into planes with encoder linesizes."""
    n, ys, uvs = 3, ffalign(w), ffalign(w // 2)
    rng = np.random.default_rng(h)
    A = rng.integers(0, 256, (n, h, w, 3), dtype=np.uint8)
    B = rng.integers(0, 256, (n, h, w, 3), dtype=np.uint8)
    DA = rng.uniform(0.1, 10.0, (n, h, w)).astype(np.float32)
    DB = rng.uniform(0.1, 10.0, (n, h, w)).astype(np.float32)
    bufs = [nes.DeviceBuffer(x.nbytes) for x in (A, DA, B, DB)]
    for buf, x in zip(bufs, (A, DA, B, DB)):
        buf.upload(x)
    Y = nes.DeviceBuffer(n * ys * h)
    U = nes.DeviceBuffer(n * uvs * h // 2)
    V = nes.DeviceBuffer(n * uvs * h // 2)
    for buf in (Y, U, V):
        buf.upload(np.full(buf.nbytes, 0xA5, np.uint8))
    nes.frameprep_device(*bufs, w, h, n, Y, U, V, mode=0, y_stride=ys, uv_stride=uvs)
    y = Y.download(np.empty((n, h, ys), np.uint8))
    u = U.download(np.empty((n, h // 2, uvs), np.uint8))
    v = V.download(np.empty((n, h // 2, uvs), np.uint8))
    for i in range(n):
        ry, ru, rv = oracle.frameprep(A[i], DA[i], B[i], DB[i], mode=0)
        assert np.array_equal(y[i, :, :w], ry)
        assert np.array_equal(u[i, :, :w // 2], ru) and np.array_equal(v[i, :, :w // 2], rv)
    assert (y[:, :, w:] == 0xA5).all() and (u[:, :, w // 2:] == 0xA5).all()
    for buf in bufs + [Y, U, V]:
        buf.free()


def test_fp32_valu_path_matches_oracle(nes, oracle, gpu_ctx, scene1):
    """The plain FP32 path (FP32 MLPs on the VALU): the oracle's sample
    sets, colour within 2e-5 and depth within 1e-4, and the same samples as
    the FP32 MFMA path."""
    cam, ocam = cams(nes, oracle, 160, 120, az=50.0)
    (o,), st = gpu_ctx.render(cam, precision=FP32_VALU)
    (m,), _ = gpu_ctx.render(cam, precision=FP32)
    pix = np.arange(160 * 120, dtype=np.uint32)
    t, w, occ = scene1
    rgb, depth, ns = oracle.render_pixels(t, w, occ, ocam, pix, threads=16)
    assert np.array_equal(o["n_samples"].reshape(-1), ns)
    assert np.abs(o["rgb"].reshape(-1, 3) - rgb).max() < 2e-5
    assert np.abs(o["depth"].reshape(-1) - depth).max() < 1e-4
    assert np.abs(o["rgb"] - m["rgb"]).max() < 2e-5
