"""Device text overlay: bit-exact with the oracle's render_string_to_frame
restatement, alone, fused into frame prep (with the nearer-depth composite)
and inside the serving loop."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PRINTABLE = [chr(c) for c in range(32, 127)]


def random_font(seed):
    """256 glyphs with FreeType-like metrics and half-empty bitmaps."""
    rng = np.random.default_rng(seed)
    glyphs, chunks, off = [], [], 0
    for c in range(256):
        w = int(rng.integers(0, 21))
        r = int(rng.integers(0, 25))
        if c == 32:
            w = r = 0
        bm = rng.integers(0, 256, w * r).astype(np.uint8)
        bm[rng.random(w * r) < 0.5] = 0
        glyphs.append((int(rng.integers(-3, 6)), int(rng.integers(0, 23)), w, r,
                       int(rng.integers(0, 15)), off))
        chunks.append(bm)
        off += bm.size
    return glyphs, np.concatenate(chunks)


def random_texts(rng, n):
    out = []
    for _ in range(n):
        k = int(rng.integers(0, 60))
        s = "".join(rng.choice(PRINTABLE + ["\n"], k))
        out.append((int(rng.integers(0, 5)), s))
    return out


@pytest.mark.parametrize("w,h,seed", [(640, 480, 0), (200, 100, 1), (1920, 1080, 2),
                                      (64, 48, 3), (2064, 2208, 4), (1440, 1600, 5)])
def test_overlay_device_matches_oracle(nes, oracle, w, h, seed):
    glyphs, bitmaps = random_font(100 + seed)
    rng = np.random.default_rng(seed)
    texts = random_texts(rng, 6) + [(p, "".join(PRINTABLE)) for p in range(5)]
    rgb = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    expect = rgb.copy()
    for pos, s in texts:
        oracle.render_string(expect, glyphs, bitmaps, pos, s)
    print(f"overlay {w}x{h} seed {seed}", flush=True)
    buf = nes.DeviceBuffer(rgb.nbytes)
    try:
        buf.upload(rgb)
        with nes.Font(glyphs, bitmaps) as font:
            nes.overlay_device(font, texts, buf, w, h)
        got = buf.download(np.empty_like(rgb))
    finally:
        buf.free()
    assert np.array_equal(got, expect)
    assert not np.array_equal(got, rgb)  # something was drawn


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("w,h", [(1280, 720), (642, 360)])
def test_frameprep_overlay_matches_oracle(nes, oracle, w, h, mode):
    glyphs, bitmaps = random_font(7)
    rng = np.random.default_rng(w + mode)
    ra = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    rb = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    da = rng.uniform(0.1, 10.0, (h, w)).astype(np.float32)
    db = rng.uniform(0.1, 10.0, (h, w)).astype(np.float32)
    c2w = nes.orbit_camera(30.0, 20.0)
    cam = nes.camera(w, h, c2w, index=123456789, is_left=1)
    ts = "01:02:03.456"
    texts = nes.overlay_texts(cam, timestamp=ts)
    comp, _ = oracle.composite(ra, da, rb, db)
    positions = [4, 1, 0, 2]  # matrix centre, index, timestamp, direction
    for pos, s in zip(positions, texts):
        oracle.render_string(comp, glyphs, bitmaps, pos, s)
    ey, eu, ev = oracle.rgb24_to_yuv420p(comp, mode)
    print(f"frameprep_overlay {w}x{h} mode {mode}", flush=True)
    with nes.Font(glyphs, bitmaps) as font:
        y, u, v = nes.frameprep_overlay(font, cam, ra, da, rb, db, timestamp=ts, mode=mode)
        # single source, no composite
        y1, u1, v1 = nes.frameprep_overlay(font, cam, ra, timestamp=ts, mode=mode)
    assert np.array_equal(y, ey) and np.array_equal(u, eu) and np.array_equal(v, ev)
    single = ra.copy()
    for pos, s in zip(positions, texts):
        oracle.render_string(single, glyphs, bitmaps, pos, s)
    sy, su, sv = oracle.rgb24_to_yuv420p(single, mode)
    assert np.array_equal(y1, sy) and np.array_equal(u1, su) and np.array_equal(v1, sv)


def test_server_overlay_matches_direct(nes, oracle, gpu_ctx):
    glyphs, bitmaps = random_font(9)
    flags = nes.OVERLAY_MATRIX | nes.OVERLAY_INDEX | nes.OVERLAY_DIRECTION
    views = [nes.camera(640, 480, nes.orbit_camera(40.0, 15.0, 2.0, e), index=42,
                        is_left=int(e < 0)) for e in (-0.032, 0.032)]
    print("server overlay 640x480 stereo", flush=True)
    with nes.Font(glyphs, bitmaps) as font:
        with nes.Server(gpu_ctx, font=font, overlay_flags=flags) as srv:
            rc, frames, _ = srv.wait(srv.submit(0, views))
    assert rc == 0
    outs, _ = gpu_ctx.render(views, precision=nes.PRECISION_MIXED)
    for cam, o, (y, u, v) in zip(views, outs, frames):
        rgb = np.ascontiguousarray(o["rgb24"])
        for pos, s in zip([4, 1, 2], nes.overlay_texts(cam, flags=flags)):
            oracle.render_string(rgb, glyphs, bitmaps, pos, s)
        ey, eu, ev = oracle.rgb24_to_yuv420p(rgb, 0)
        assert np.array_equal(y, ey) and np.array_equal(u, eu) and np.array_equal(v, ev)
