"""neserver's renderer protocol end to end: length-prefixed FrameRequest in,
RenderedFrame with the library's RGB24 (and quantised depth) out."""
import socket
import struct

import numpy as np
import pytest

from test_wire import RenderedFrame, request

pytestmark = pytest.mark.gpu


def lpf_send(sock, payload):
    sock.sendall(struct.pack("=Q", len(payload)) + payload)  # host-endian size_t


def lpf_recv(sock):
    def exact(n):
        buf = b""
        while len(buf) < n:
            chunk = sock.recv(n - len(buf))
            assert chunk, "renderer closed the connection"
            buf += chunk
        return buf
    (n,) = struct.unpack("=Q", exact(8))
    return exact(n)


def test_neserver_protocol_round_trip(nes, oracle, gpu_ctx):
    c2w = nes.orbit_camera(30.0, 20.0)
    with nes.WireServer(gpu_ctx, depth_far=4.0) as srv:
        with socket.create_connection(("127.0.0.1", srv.port), timeout=30) as s:
            for index, (w, h, left) in enumerate([(320, 240, True), (256, 128, False)]):
                lpf_send(s, request(index, left, w, h, c2w.tolist()).SerializeToString())
                f = RenderedFrame.FromString(lpf_recv(s))
                assert f.index == index and f.is_left == left
                assert (f.camera.width, f.camera.height) == (w, h)
                (ref,), _ = gpu_ctx.render(nes.camera(w, h, c2w, fov=60.0))
                assert f.frame == ref["rgb24"].tobytes()
                gray = np.frombuffer(f.depth, np.uint8).reshape(h, w)
                assert np.array_equal(gray, oracle.depth_gray8(ref["depth"], 4.0))
                assert np.array_equal(gray, ref["depth_gray8"])
        assert srv.served() == 2
