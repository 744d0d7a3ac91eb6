"""The hand-written proto3 codec of the neserver wire protocol against the
protobuf runtime (schema of proto/nes.proto:5-25 built as a descriptor)."""
import numpy as np
import pytest
from google.protobuf import descriptor_pb2, descriptor_pool, message_factory

F = descriptor_pb2.FieldDescriptorProto


def _schema():
    fdp = descriptor_pb2.FileDescriptorProto(name="nes_test.proto", package="nesproto",
                                             syntax="proto3")
    cam = fdp.message_type.add(name="Camera")
    cam.field.add(name="is_left", number=2, type=F.TYPE_BOOL, label=F.LABEL_OPTIONAL)
    cam.field.add(name="width", number=3, type=F.TYPE_UINT32, label=F.LABEL_OPTIONAL)
    cam.field.add(name="height", number=4, type=F.TYPE_UINT32, label=F.LABEL_OPTIONAL)
    cam.field.add(name="matrix", number=12, type=F.TYPE_FLOAT, label=F.LABEL_REPEATED)
    req = fdp.message_type.add(name="FrameRequest")
    req.field.add(name="index", number=1, type=F.TYPE_UINT64, label=F.LABEL_OPTIONAL)
    req.field.add(name="camera", number=2, type=F.TYPE_MESSAGE, type_name=".nesproto.Camera",
                  label=F.LABEL_OPTIONAL)
    req.field.add(name="is_left", number=3, type=F.TYPE_BOOL, label=F.LABEL_OPTIONAL)
    rf = fdp.message_type.add(name="RenderedFrame")
    rf.field.add(name="index", number=1, type=F.TYPE_UINT64, label=F.LABEL_OPTIONAL)
    rf.field.add(name="camera", number=2, type=F.TYPE_MESSAGE, type_name=".nesproto.Camera",
                 label=F.LABEL_OPTIONAL)
    rf.field.add(name="is_left", number=3, type=F.TYPE_BOOL, label=F.LABEL_OPTIONAL)
    rf.field.add(name="frame", number=6, type=F.TYPE_BYTES, label=F.LABEL_OPTIONAL)
    rf.field.add(name="depth", number=7, type=F.TYPE_BYTES, label=F.LABEL_OPTIONAL)
    pool = descriptor_pool.DescriptorPool()
    pool.Add(fdp)
    get = lambda n: message_factory.GetMessageClass(pool.FindMessageTypeByName("nesproto." + n))
    return get("Camera"), get("FrameRequest"), get("RenderedFrame")


Camera, FrameRequest, RenderedFrame = _schema()
INITIAL = [1.0, 0.0, 0.0, 0.5, 0.0, -1.0, 0.0, 0.5, 0.0, 0.0, -1.0, 0.5]  # camera_manager.h:18-19


def request(index, is_left, w, h, matrix):
    r = FrameRequest(index=index, is_left=is_left)
    r.camera.is_left = is_left
    r.camera.width, r.camera.height = w, h
    r.camera.matrix.extend(matrix)
    return r


def test_decode_request(nes):
    m = np.random.default_rng(0).normal(size=12).astype(np.float32)
    msg = request(123456789012, True, 1280, 720, m.tolist()).SerializeToString()
    cam = nes.wire_decode_request(msg, 60.0)
    assert cam.index == 123456789012 and cam.is_left == 1
    assert (cam.width, cam.height) == (1280, 720)
    assert np.array_equal(np.array(cam.c2w, np.float32), m) and cam.fov_y_deg == 60.0
    cam = nes.wire_decode_request(request(0, False, 2, 4, INITIAL).SerializeToString(), 75.0)
    assert cam.index == 0 and cam.is_left == 0 and cam.fov_y_deg == 75.0
    assert list(cam.c2w) == INITIAL


def test_decode_unpacked_matrix_and_unknown_fields(nes):
    # unpacked repeated float (wire type 5 per element) + an unknown varint and
    # length-delimited field, plus the additive FOV field 5
    body = b"".join([b"\x18\x80\x05", b"\x20\xd0\x02"] +
                    [b"\x65" + np.float32(v).tobytes() for v in INITIAL] +
                    [b"\x2d" + np.float32(90.0).tobytes(), b"\x48\x07", b"\x52\x02ab"])
    msg = b"\x08\x05" + b"\x12" + bytes([len(body)]) + body + b"\x18\x01"
    cam = nes.wire_decode_request(msg, 60.0)
    assert (cam.width, cam.height, cam.index, cam.is_left) == (640, 336, 5, 1)
    assert list(cam.c2w) == INITIAL and cam.fov_y_deg == 90.0


@pytest.mark.parametrize("bad", [b"", b"\x08", b"\x12\x05abc", b"\x12\x02\x18\x01",
                                 b"\x0f\x00"])
def test_decode_rejects_malformed(nes, bad):
    with pytest.raises(nes.NesError):
        nes.wire_decode_request(bad)


def test_encode_frame(nes):
    cam = nes.camera(8, 6, INITIAL, index=77, is_left=1)
    rgb = np.arange(8 * 6 * 3, dtype=np.uint8)
    gray = np.arange(8 * 6, dtype=np.uint8)
    msg = nes.wire_encode_frame(cam, rgb, gray)
    f = RenderedFrame.FromString(msg)
    assert f.index == 77 and f.is_left and f.camera.is_left
    assert (f.camera.width, f.camera.height) == (8, 6)
    assert list(f.camera.matrix) == INITIAL
    assert f.frame == rgb.tobytes() and f.depth == gray.tobytes()
