"""Text overlay: the oracle's restatement of render_string_to_frame
(src/base/video/render_text.cc:35-111) pinned by hand-computed pixel
positions, and the overlay strings of process_frame_thread
(src/encode.cpp:57-97) produced by the C-ABI (host-only, no device)."""
import numpy as np
import pytest

LEFT_TOP, LEFT_BOTTOM, RIGHT_TOP, RIGHT_BOTTOM, CENTER = range(5)


def tiny_font():
    """'A': 3x2 bitmap [[1,0,2],[0,3,0]] at left 1, top 2, advance 4;
    'B': 1x1 bitmap [9] at left -2, top 0, advance 5; ' ': advance 7."""
    glyphs = [(0, 0, 0, 0, 0, 0)] * 256
    glyphs[ord("A")] = (1, 2, 3, 2, 4, 0)
    glyphs[ord("B")] = (-2, 0, 1, 1, 5, 6)
    glyphs[ord(" ")] = (0, 0, 0, 0, 7, 0)
    bitmaps = np.array([1, 0, 2, 0, 3, 0, 9], np.uint8)
    return glyphs, bitmaps


def stamped(oracle, w, h, position, text, seed=0):
    glyphs, bitmaps = tiny_font()
    rng = np.random.default_rng(seed)
    bg = rng.integers(0, 255, (h, w, 3), dtype=np.uint8)  # never 255
    out = oracle.render_string(bg.copy(), glyphs, bitmaps, position, text)
    changed = np.argwhere(np.any(out != bg, axis=2))
    assert np.all(out[changed[:, 0], changed[:, 1]] == 255)
    return sorted((int(x), int(y)) for y, x in changed)


@pytest.mark.parametrize("w,h,position,text,expect", [
    # pen (50, 50): glyph box starts at (51, 48)
    (400, 200, LEFT_TOP, "A", [(51, 48), (52, 49), (53, 48)]),
    # second glyph 4 px further right
    (400, 200, LEFT_TOP, "AA", [(51, 48), (52, 49), (53, 48), (55, 48), (56, 49), (57, 48)]),
    # newline: back to x 50, 20 px down
    (400, 200, LEFT_TOP, "A\nA", [(51, 48), (51, 68), (52, 49), (52, 69), (53, 48), (53, 68)]),
    # space advances 7 without drawing; B sits 2 px left of its pen
    (400, 200, LEFT_TOP, " B", [(55, 50)]),
    # centre of 640x480: pen (320 - 300, 240 - 100)
    (640, 480, CENTER, "A", [(21, 138), (22, 139), (23, 138)]),
    # right bottom of 640x480: pen (640 - 250, 480 - 50)
    (640, 480, RIGHT_BOTTOM, "A", [(391, 428), (392, 429), (393, 428)]),
    (640, 480, LEFT_BOTTOM, "A", [(51, 428), (52, 429), (53, 428)]),
    (640, 480, RIGHT_TOP, "A", [(391, 48), (392, 49), (393, 48)]),
    # right top of a 200-wide frame: pen x = -50, clipped on the left
    (200, 100, RIGHT_TOP, "A" * 14, [(0, 49), (1, 48), (3, 48), (4, 49), (5, 48)]),
    # centre pen (-250, -70) of a 100x60 frame: nothing visible
    (100, 60, CENTER, "AAAA", []),
    # bottom row clipped: frame height 49
    (100, 49, LEFT_TOP, "A", [(51, 48), (53, 48)]),
    # unknown position behaves as left top
    (400, 200, 7, "A", [(51, 48), (52, 49), (53, 48)]),
])
def test_render_string_known_answers(oracle, w, h, position, text, expect):
    assert stamped(oracle, w, h, position, text) == sorted(expect)


def test_render_string_negative_pen_from_unsigned_size(oracle):
    # height 40: pen y = 40 - 100 + 50 = -10 (unsigned arithmetic stored in
    # an int), so the glyph box starts at y = -12 and nothing is drawn
    assert stamped(oracle, 300, 40, LEFT_BOTTOM, "AB") == []


def test_render_string_idempotent_and_order_free(oracle):
    glyphs, bitmaps = tiny_font()
    rng = np.random.default_rng(3)
    bg = rng.integers(0, 255, (120, 400, 3), dtype=np.uint8)
    a = oracle.render_string(bg.copy(), glyphs, bitmaps, LEFT_TOP, "AB A\nBA")
    a = oracle.render_string(a, glyphs, bitmaps, CENTER, "ABAB")
    b = oracle.render_string(bg.copy(), glyphs, bitmaps, CENTER, "ABAB")
    b = oracle.render_string(b, glyphs, bitmaps, LEFT_TOP, "AB A\nBA")
    b = oracle.render_string(b, glyphs, bitmaps, LEFT_TOP, "AB A\nBA")
    assert np.array_equal(a, b)


def test_overlay_strings(nes):
    c2w = [1, 0, 0, 0.5, 0, -1, 0, 0.5, 0, 0, -1, 0.5]
    cam = nes.camera(640, 480, c2w, index=17, is_left=0)
    texts = nes.overlay_texts(cam, timestamp="12:34:56.789")
    assert texts == [
        "+1.00000 +0.00000 +0.00000 +0.50000 \n"
        "+0.00000 -1.00000 +0.00000 +0.50000 \n"
        "+0.00000 +0.00000 -1.00000 +0.50000 \n"
        "+0.00000 +0.00000 +0.00000 +1.00000 ",
        "index=17", "12:34:56.789", "direction=right"]
    cam.is_left = 1
    assert nes.overlay_texts(cam, flags=nes.OVERLAY_DIRECTION) == ["direction=left"]
    assert nes.overlay_texts(cam, flags=nes.OVERLAY_INDEX | nes.OVERLAY_DIRECTION) == [
        "index=17", "direction=left"]
    assert nes.overlay_texts(cam, flags=0) == []


def test_overlay_matrix_number_format(nes):
    # showpos + fixed + precision 5, width 7 left-filled with '0'
    vals = [float("inf"), -0.0, 12.345678, -3.5e-6, 1e7, float("-inf"), 0.123455, 2.0 / 3,
            -1.0, 0.5, 1e-30, 99.999996]
    cam = nes.camera(64, 48, vals)
    text = nes.overlay_texts(cam, flags=nes.OVERLAY_MATRIX)[0]
    rows = text.split("\n")
    assert len(rows) == 4
    tokens = " ".join(rows).split()
    assert tokens[:2] == ["000+inf", "-0.00000"]
    assert tokens[5] == "000-inf"
    for tok, v in zip(tokens[2:12], vals[2:12]):
        if np.isinf(v):
            continue
        assert tok == "%+.5f" % np.float32(v)
    assert tokens[12:] == ["+0.00000", "+0.00000", "+0.00000", "+1.00000"]
    assert all(r.endswith(" ") for r in rows)


def test_overlay_timestamp_shape(nes):
    cam = nes.camera(64, 48, [0] * 12)
    (ts,) = nes.overlay_texts(cam, flags=nes.OVERLAY_TIMESTAMP)
    assert len(ts) == 12 and ts[2] == ":" and ts[5] == ":" and ts[8] == "."
    assert all(ts[i].isdigit() for i in (0, 1, 3, 4, 6, 7, 9, 10, 11))


def test_font_rejects_glyph_outside_atlas(nes):
    glyphs, bitmaps = tiny_font()
    glyphs = list(glyphs)
    glyphs[ord("C")] = (0, 0, 4, 4, 4, 1)  # 16 bytes from offset 1, atlas has 7
    with pytest.raises(nes.NesError) as e:
        nes.Font(glyphs, bitmaps)
    assert e.value.code == -22
