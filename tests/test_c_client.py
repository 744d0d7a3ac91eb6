"""The C-ABI from C: tests/c_client/c_abi_client.c is compiled with gcc
against include/nes_render.h and linked to libnes_render.so, as the
reference's C++ would be, and its outputs are compared with the Python
binding's results for the same calls (checksums of every buffer)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "ngp-encode-server_amd", "lib")


def checksum(a):
    b = np.ascontiguousarray(a).view(np.uint8).ravel().astype(np.uint64)
    w = (np.arange(b.size, dtype=np.uint64) % np.uint64(65521)) + np.uint64(1)
    return int((b * w).sum(dtype=np.uint64))


@pytest.fixture(scope="module")
def client(tmp_path_factory, nes):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    exe = str(tmp_path_factory.mktemp("c_client") / "c_abi_client")
    subprocess.run(["gcc", "-std=c99", "-O2", "-Wall", "-Werror",
                    "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c_client", "c_abi_client.c"),
                    "-L", LIBDIR, "-lnes_render", "-Wl,-rpath," + LIBDIR, "-o", exe],
                   check=True)
    return exe


def run(exe, mode):
    p = subprocess.run([exe, mode], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    out = {}
    for line in p.stdout.splitlines():
        k, v = line.split(" ", 1)
        out[k] = v
    return out


def test_c_client_host_calls(client, nes):
    got = run(client, "cpu")
    assert got["strerror"] == nes.lib().nes_strerror(-22).decode()  # NES_EINVAL
    assert int(got["orbit"]) == checksum(nes.orbit_camera(30.0, 20.0, 2.0, -0.032))
    _, w, o = nes.scene_synthesize(1)
    assert int(got["weights"]) == checksum(w) and int(got["occupancy"]) == checksum(o)
    coef, _ = nes.vchroma_filter(720)
    assert int(got["vfilter_taps"]) == coef.shape[1]
    assert int(got["vfilter_coef"]) == checksum(coef)


@pytest.mark.gpu
def test_c_client_render_frameprep_server(client, nes, gpu_ctx):
    got = run(client, "gpu")
    cams = [nes.camera(160, 120, nes.orbit_camera(30.0, 20.0, 2.0, e), fov=60.0, index=7,
                       is_left=int(e < 0)) for e in (-0.032, 0.032)]
    outs, _ = gpu_ctx.render(cams, precision=nes.PRECISION_MIXED)
    for e in range(2):
        assert int(got[f"rgb24_{e}"]) == checksum(outs[e]["rgb24"])
        assert int(got[f"gray8_{e}"]) == checksum(outs[e]["depth_gray8"])
    y, u, v = nes.frameprep(outs[0]["rgb24"])
    assert [int(got[k]) for k in ("yuv_y", "yuv_u", "yuv_v")] == [checksum(y), checksum(u), checksum(v)]
    y, u, _ = nes.frameprep(outs[0]["rgb24"], outs[0]["depth"], outs[1]["rgb24"], outs[1]["depth"])
    assert [int(got["comp_y"]), int(got["comp_u"])] == [checksum(y), checksum(u)]
    y, u, _ = nes.gray8_to_yuv420p(outs[0]["depth_gray8"])
    assert [int(got["depth_y"]), int(got["depth_u"])] == [checksum(y), checksum(u)]
    with nes.Server(gpu_ctx, max_batch_rays=1 << 20) as srv:
        rc, frames, _ = srv.wait(srv.submit(1, cams))
    assert rc == 0
    assert int(got["server_y0"]) == checksum(frames[0][0])
    assert int(got["server_u0"]) == checksum(frames[0][1])
    assert int(got["server_v1"]) == checksum(frames[1][2])
