"""Fuzzing the FrameRequest decoder (CPU; the bytes come from the network):
arbitrary, truncated and mutated messages never crash it; whatever it accepts
the protobuf runtime parses to the same request (it may refuse more, never
accept what the runtime rejects or read it differently)."""
import numpy as np
import pytest
from google.protobuf.message import DecodeError
from hypothesis import HealthCheck, given, settings, strategies as st

from test_wire import FrameRequest, request


def decode(nes, msg):
    try:
        return nes.wire_decode_request(msg, 60.0)
    except nes.NesError as e:
        assert e.code == -22
        return None


def agrees_with_runtime(cam, msg):
    try:
        r = FrameRequest.FromString(msg)
    except DecodeError:
        raise AssertionError("accepted bytes the protobuf runtime rejects")
    assert cam.index == r.index and cam.is_left == int(r.is_left)
    assert (cam.width, cam.height) == (r.camera.width, r.camera.height)
    assert len(r.camera.matrix) == 12
    got = np.array(cam.c2w, np.float32)
    want = np.array(list(r.camera.matrix), np.float32)
    # bit-exact, except that the Python runtime hands floats over as doubles,
    # which quiets a signalling NaN's payload: NaN must stay NaN
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint32), want[~nan].view(np.uint32))


FUZZ = settings(max_examples=400, deadline=None, suppress_health_check=[HealthCheck.too_slow])
floats = st.floats(width=32, allow_nan=False)
valid = st.builds(lambda i, l, w, h, m: request(i, l, w, h, m).SerializeToString(),
                  st.integers(0, 2 ** 64 - 1), st.booleans(), st.integers(0, 2 ** 32 - 1),
                  st.integers(0, 2 ** 32 - 1), st.lists(floats, min_size=12, max_size=12))


@FUZZ
@given(st.binary(max_size=300))
def test_random_bytes(nes, data):
    cam = decode(nes, data)
    if cam is not None:
        agrees_with_runtime(cam, data)


@FUZZ
@given(valid)
def test_valid_requests_round_trip(nes, msg):
    cam = decode(nes, msg)
    assert cam is not None
    agrees_with_runtime(cam, msg)


@FUZZ
@given(valid, st.data())
def test_truncated_and_mutated(nes, msg, data):
    cut = data.draw(st.integers(0, len(msg)))
    for m in (msg[:cut],
              msg[:cut] + bytes([data.draw(st.integers(0, 255))]) + msg[cut + 1:],
              msg[:cut] + data.draw(st.binary(max_size=16)) + msg[cut:]):
        cam = decode(nes, m)
        if cam is not None:
            agrees_with_runtime(cam, m)


@pytest.mark.parametrize("msg", [b"\x00\x00" + request(1, True, 4, 4, [0.0] * 12).SerializeToString(),
                                 request(1, True, 4, 4, [0.0] * 12).SerializeToString() * 2])
def test_refuses_invalid_tag_and_repeated_camera(nes, msg):
    assert decode(nes, msg) is None
