"""Benchmark: mixed-precision stereo render throughput on one MI355X.

One step = one render of a stereo pair, 1920x1080 per eye with 64 mm IPD
(4,147,200 rays), seeded scene (16-level hash grid, 128^3 occupancy blob,
Xavier MLPs), through the C-ABI (nes_render_device: outputs stay in HBM).
Prints one JSON line.  --gpus above 1 is refused: the path runs on one GPU.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ngp-encode-server_amd"))

METRIC = "Mrays/s mixed-prec vs FP32 at ≥45 dB PSNR; p99 frame latency, 8 clients"
FLOP_PER_SAMPLE = 18688      # density 32-64-16 + colour 31-64-64-3 MLPs, 2 FLOP/MAC
GATHER_BYTES_PER_SAMPLE = 512  # 16 levels x 8 corners x 2 FP16 features
PEAK_FP16_TFLOPS = 2500.0    # MI355X dense FP16 MFMA
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stereo_views(N, w=1920, h=1080):
    return [N.camera(w, h, N.orbit_camera(30.0, 20.0, 2.0, e), fov=60.0, is_left=int(e < 0))
            for e in (-0.032, 0.032)]


def time_render(N, ctx, views, precision, steps, warmup):
    for _ in range(warmup):
        ctx.render_device(views, precision)
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        stats.append(ctx.render_device(views, precision))
    wall = time.perf_counter() - t0
    return wall, stats


def host_cpu():
    """nproc and the CPU model name of the host the CPU baselines run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return os.cpu_count() or 1, model


def oracle_compare(O, table, weights, occ, views, outs, stride, threads):
    """FP64 oracle on every stride-th pixel of each view: (rays, seconds,
    PSNR of outs' float RGB, max |depth error|)."""
    total_rays, total_t, ref, got, derr = 0, 0.0, [], [], 0.0
    for v, o in zip(views, outs):
        oc = O.camera(v.width, v.height, list(v.c2w), fov=v.fov_y_deg)
        pix = np.arange(0, v.width * v.height, stride, dtype=np.uint32)
        t0 = time.perf_counter()
        rgb, depth, _ = O.render_pixels(table, weights, occ, oc, pix, threads=threads)
        total_t += time.perf_counter() - t0
        total_rays += len(pix)
        ref.append(rgb)
        got.append(o["rgb"].reshape(-1, 3)[pix])
        derr = max(derr, float(np.abs(o["depth"].reshape(-1)[pix] - depth).max()))
    ref, got = np.concatenate(ref), np.concatenate(got)
    mse = float(np.mean((ref - got) ** 2))
    psnr = 10 * math.log10(1.0 / mse) if mse > 0 else float("inf")
    return total_rays, total_t, psnr, derr


def cpu_baseline(N, table, weights, occ, ctx, views, budget_s):
    """FP64 oracle on the same rays (whole frames, or an even stride of them
    when that would exceed about budget_s seconds); its pixels are also the
    PSNR and depth-error reference of the mixed and FP32 frames."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    nproc, model = host_cpu()
    threads = min(16, nproc)
    outs, _ = ctx.render(views, precision=N.PRECISION_MIXED)
    outs32, _ = ctx.render(views, precision=N.PRECISION_FP32)
    w, h = views[0].width, views[0].height
    probe = np.arange(0, w * h, 97, dtype=np.uint32)
    oc = O.camera(w, h, list(views[0].c2w), fov=views[0].fov_y_deg)
    t0 = time.perf_counter()
    O.render_pixels(table, weights, occ, oc, probe, threads=threads)
    per_ray = (time.perf_counter() - t0) / len(probe)
    stride = max(1, int(math.ceil(per_ray * w * h * len(views) / budget_s)))
    rays, secs, psnr, derr = oracle_compare(O, table, weights, occ, views, outs, stride, threads)
    _, _, psnr32, derr32 = oracle_compare(O, table, weights, occ, views, outs32, stride * 16,
                                          threads)
    what = "both full 1920x1080 eyes" if stride == 1 else f"every {stride}th pixel of both eyes"
    base = {"value": rays / secs / 1e6, "unit": "Mrays/s", "cores": threads, "nproc": nproc,
            "cpu_model": model, "kind": "port",
            "sample": f"FP64 oracle render of {rays} rays ({what}), {secs:.1f} s"}
    acc = {"mixed_psnr_db": round(psnr, 2), "mixed_max_depth_err": derr,
           "fp32_psnr_db": round(psnr32, 2), "fp32_max_depth_err": derr32,
           "rays_compared": rays}
    return base, acc


def config1_leg(N, table, weights, occ, ctx, steps=20):
    """640x480, 60 deg, radius-2 orbit: the plain FP32 path (VALU MLPs), the
    FP32 MFMA path and the mixed path against the FP64 oracle over the whole
    frame, plus the oracle's own rate on the host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    nproc, model = host_cpu()
    threads = min(16, nproc)
    view = [N.camera(640, 480, N.orbit_camera(30.0, 20.0), fov=60.0)]
    out = {}
    for name, prec in (("fp32_valu", N.PRECISION_FP32_VALU), ("fp32", N.PRECISION_FP32),
                       ("mixed", N.PRECISION_MIXED)):
        wall, st = time_render(N, ctx, view, prec, steps, 3)
        outs, _ = ctx.render(view, precision=prec)
        rays, secs, psnr, derr = oracle_compare(O, table, weights, occ, view, outs, 1, threads)
        out[name] = {"mrays_per_s": round(640 * 480 * steps / wall / 1e6, 2),
                     "ms_per_frame": round(wall / steps * 1e3, 4),
                     "psnr_db": round(psnr, 2), "max_depth_err": derr}
        out["oracle"] = {"mrays_per_s": round(rays / secs / 1e6, 3), "threads": threads,
                         "nproc": nproc, "cpu_model": model, "rays": rays}
    return out


def opaque_leg(N, views, steps=10):
    """Early termination in the timed path: a constant-density opaque scene
    (sigma = 1500, every cell occupied; each ray stops at T < 1e-4 after its
    first sample) at stereo 1080p, single pass against the adaptive default
    (two passes once a frame has shown the waste)."""
    table = np.full(N.TABLE_ENTRIES * 2, 1.0, np.float32)
    w = np.zeros(N.MLP_WEIGHTS, np.float32)
    w[0:2048] = 1.0 / 32.0
    w[2048:2048 + 64] = math.log(1500.0) / 64.0
    occ = np.full(N.OCC_WORDS, 0xFFFFFFFF, np.uint32)
    out = {}
    with N.Context(table, w, occ) as c:
        for name, spp in (("single_pass", 128), ("adaptive", 0)):
            for _ in range(2):
                c.render_device(views, N.PRECISION_MIXED, spp)
            t0 = time.perf_counter()
            for _ in range(steps):
                st = c.render_device(views, N.PRECISION_MIXED, spp)
            ms = (time.perf_counter() - t0) / steps * 1e3
            out[name] = {"ms_per_frame": round(ms, 4), "passes": st.passes,
                         "evaluated": st.samples, "composited": st.composited,
                         "waste": round(1 - st.composited / max(st.samples, 1), 4)}
    return out


def ceilings_leg(N, samples, ms_encode, l2_req_per_launch, l1_acc_per_launch=None):
    """Measured gather ceilings (nes_device_gather_rate): random one-lane
    4-byte loads inside 16 KiB per XCD (vector-L1 hits: lane-address rate) and
    inside 2 MiB per XCD (L1 misses, L2 hits: L2 request rate), and the
    encode's rates against them."""
    l1, _ = N.gather_rate(16 << 10)
    l2, _ = N.gather_rate(2 << 20)
    # corner fetches: 16 levels x 8 corners per sample.  The encode issues
    # fewer lane addresses than that (one per corner pair on dense levels,
    # one per 16 B chunk on hashed levels 5-13), so this can exceed the
    # lane-address ceiling.
    fetches = 128 * samples / (ms_encode * 1e-3)
    out = {"l1_lane_loads_per_s": l1, "l2_requests_per_s": l2,
           "encode_corner_fetches_per_s": fetches,
           "encode_fetches_per_l1_lane_slot": round(fetches / l1, 4),
           # LDS atomic adds (every lane, random addresses): why the training
           # scatter sums its slices in FP64
           "lds_atomic_adds_per_s": {k: N.lds_atomic_rate(k) for k in ("f32", "f64", "u32", "f32_cas", "f32x2_cas")}}
    if l1_acc_per_launch:
        out["encode_l1_accesses_frac"] = round(l1_acc_per_launch / (ms_encode * 1e-3) / l1, 4)
    if l2_req_per_launch:
        rate = l2_req_per_launch / (ms_encode * 1e-3)
        out.update({"encode_l2_requests_per_s": rate, "encode_l2_frac": round(rate / l2, 4),
                    # level 15 on XCD 0: no reuse between samples, so 4 L2
                    # requests (one per y/z corner pair) per sample
                    "xcd0_l2_requests_per_s": 4 * samples / (ms_encode * 1e-3),
                    "xcd0_l2_frac": round(4 * samples / (ms_encode * 1e-3) / (l2 / 8), 4)})
    return out


def cpu_frameprep_baseline(threads, w=1280, h=720, runs=100, warmup=10):
    """The repo's CPU frame prep (composite + RGB24 -> YUV420P restatement,
    filter built per frame as SwsContextManager does) on one 1280x720 pair:
    median ms per frame of `runs` after `warmup`, on 1 thread (one
    process_frame_thread per eye) and on `threads` threads (rows split)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    rng = np.random.default_rng(0)
    a = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    b = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    da = rng.uniform(0.1, 10.0, (h, w)).astype(np.float32)
    db = rng.uniform(0.1, 10.0, (h, w)).astype(np.float32)
    out = {}
    for key, n in (("ms_per_frame_1t", 1), ("ms_per_frame_mt", threads)):
        O.set_frameprep_threads(n)
        times = []
        for i in range(warmup + runs):
            t0 = time.perf_counter()
            O.frameprep(a, da, b, db)
            if i >= warmup:
                times.append((time.perf_counter() - t0) * 1e3)
        out[key] = float(np.median(times))
    O.set_frameprep_threads(1)
    out.update({"threads": threads, "kind": "port", "runs": runs, "warmup": warmup,
                "sample": f"2 x {w}x{h} RGB24 + f32 depth -> composite -> YUV420P"})
    return out


def serving_leg(N, ctx, seconds):
    """8 Poisson clients at 72 Hz, stereo, per-eye {640x480, 1440x1600, 2064x2208}."""
    with N.Server(ctx) as srv:
        srv.simulate(n_clients=8, rate_hz=72.0, duration_s=0.3, seed=99)  # warm-up
        st = srv.simulate(n_clients=8, rate_hz=72.0, duration_s=seconds, seed=4)
    d = st.as_dict()
    d["offered_requests_per_s"] = 8 * 72.0
    return d


def frameprep_leg(N, n_frames=256, w=1280, h=720):
    """Frame prep on device: two 1280x720 RGB24 + depth inputs composited and
    converted, 256 frames per launch (the bit-exact kernel)."""
    rng = np.random.default_rng(0)
    a = rng.integers(0, 256, (n_frames, h, w, 3), dtype=np.uint8)
    b = rng.integers(0, 256, (n_frames, h, w, 3), dtype=np.uint8)
    da = rng.uniform(0.1, 10.0, (n_frames, h, w)).astype(np.float32)
    db = rng.uniform(0.1, 10.0, (n_frames, h, w)).astype(np.float32)
    bufs = [N.DeviceBuffer(x.nbytes) for x in (a, da, b, db)]
    for buf, x in zip(bufs, (a, da, b, db)):
        buf.upload(x)
    Y = N.DeviceBuffer(n_frames * w * h)
    U = N.DeviceBuffer(n_frames * w * h // 4)
    V = N.DeviceBuffer(n_frames * w * h // 4)
    N.frameprep_device(*bufs, w, h, n_frames, Y, U, V)
    ms = sorted(N.frameprep_device(*bufs, w, h, n_frames, Y, U, V) for _ in range(10))[5]
    for buf in bufs + [Y, U, V]:
        buf.free()
    # the practical streaming ceiling: a plain 1 GiB device-to-device copy
    # (read + write), the same byte mix a pure streaming kernel moves
    nbytes = 1 << 30
    src, dst = N.DeviceBuffer(nbytes), N.DeviceBuffer(nbytes)
    N.device_copy(dst, src, nbytes)
    copy_ms = sorted(N.device_copy(dst, src, nbytes) for _ in range(5))[2]
    src.free()
    dst.free()
    copy_gbs = 2 * nbytes / copy_ms / 1e6
    bytes_per_frame = w * h * 15.5
    gbs = bytes_per_frame * n_frames / ms / 1e6
    return {"frames_per_launch": n_frames, "ms_per_launch": ms,
            "frames_per_s": n_frames / ms * 1e3,
            "achieved_GBs": gbs, "hbm_frac": gbs / PEAK_HBM_GBS,
            "copy_ceiling_GBs": round(copy_gbs, 1),
            "frac_of_copy_ceiling": round(gbs / copy_gbs, 4)}


def training_leg(N, ctx, occ, steps=12, warmup=3):
    """Differentiable training step: 2^18 rays per batch drawn from 16 seeded
    poses at 640x480, targets rendered by the seeded teacher scene, student
    table U(-1e-4, 1e-4); forward, L2 loss, backward (composite, MLPs on
    FP32 MFMA, table gradient summed per level slice in LDS) and Adam per
    step."""
    rng = np.random.default_rng(3)
    poses = np.stack([N.orbit_camera(360.0 * rng.random(), 5.0 + 30.0 * rng.random())
                      for _ in range(16)])
    outs, _ = ctx.render([N.camera(640, 480, p) for p in poses], precision=N.PRECISION_FP32)
    targets = np.stack([o["rgb"] for o in outs])
    ts, ws, _ = N.scene_synthesize(4, table_init=1)
    rays = 1 << 18
    with N.Trainer(ts, ws, occ, poses=poses, targets=targets, width=640, height=480,
                   rays_per_batch=rays, seed=0) as tr:
        for _ in range(warmup):
            tr.step()
        ms, samples, losses = [], [], []
        for _ in range(steps):
            loss, st = tr.step()
            ms.append(st.ms)
            samples.append(st.samples)
            losses.append(loss)
    m = float(np.median(ms))
    n = float(np.mean(samples))
    return {"rays_per_step": rays, "ms_per_step": round(m, 3),
            "mrays_per_s": round(rays / m / 1e3, 2), "samples_per_step": int(n),
            # forward + data gradient + weight gradient (recompute not credited)
            "mlp_tflops": round(3 * FLOP_PER_SAMPLE * n / (m * 1e-3) / 1e12, 2),
            "loss_first": losses[0], "loss_last": losses[-1]}


def load_profile(name, key):
    """A figure from a committed rocprofv3 summary under profiles/."""
    p = os.path.join(ROOT, "profiles", name)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get(key)
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-frameprep", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--skip-serve", action="store_true")
    ap.add_argument("--skip-train", action="store_true")
    ap.add_argument("--skip-extra", action="store_true",
                    help="skip the config-1, opaque-scene and ceiling legs")
    ap.add_argument("--train-only", action="store_true",
                    help="only the training leg (for profiling its kernels)")
    ap.add_argument("--serve-only", action="store_true",
                    help="only the serving leg (for profiling its kernels)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 device copies of 1 GiB (memory-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the render path runs on one GPU; --gpus {args.gpus} is not supported "
              "(requests and tiles shard across GPUs with no exchange; only training would "
              "all-reduce its gradients)", file=sys.stderr)
        sys.exit(2)

    import nes_amd as N
    if args.copy_calibration:
        nbytes = 1 << 30
        a, b = N.DeviceBuffer(nbytes), N.DeviceBuffer(nbytes)
        ms = [N.device_copy(b, a, nbytes) for _ in range(5)]
        print(json.dumps({"copy_bytes": nbytes, "ms": ms}))
        return
    name, cus, _ = N.device_info()
    log(f"device {name} ({cus} CUs)")
    table, weights, occ = N.scene_synthesize(1)
    views = stereo_views(N)
    rays = sum(v.width * v.height for v in views)
    with N.Context(table, weights, occ) as ctx:
        if args.train_only:
            print(json.dumps({"training": training_leg(N, ctx, occ, steps=args.steps,
                                                       warmup=args.warmup)}))
            return
        if args.serve_only:
            print(json.dumps({"serving": serving_leg(N, ctx, args.serve_seconds)}))
            return
        wall_m, st_m = time_render(N, ctx, views, N.PRECISION_MIXED, args.steps, args.warmup)
        log(f"mixed: {wall_m / args.steps * 1e3:.3f} ms/step")
        nf = max(3, args.steps // 4)
        wall_f, st_f = time_render(N, ctx, views, N.PRECISION_FP32, nf, args.warmup)
        log(f"fp32 (MFMA MLPs): {wall_f / nf * 1e3:.3f} ms/step")
        wall_v, st_v = time_render(N, ctx, views, N.PRECISION_FP32_VALU, nf, args.warmup)
        log(f"fp32 (VALU MLPs): {wall_v / nf * 1e3:.3f} ms/step")
        samples = st_m[-1].samples
        ms_enc = float(np.mean([s.ms_encode for s in st_m]))
        ms_mlp = float(np.mean([s.ms_mlp for s in st_m]))
        ms_march = float(np.mean([s.ms_march for s in st_m]))
        ms_comp = float(np.mean([s.ms_composite for s in st_m]))
        value = rays * args.steps / wall_m / 1e6
        fp32 = rays * nf / wall_f / 1e6
        fp32_valu = rays * nf / wall_v / 1e6
        gathered_gbs = GATHER_BYTES_PER_SAMPLE * samples / (ms_enc * 1e-3) / 1e9
        mlp_tflops = FLOP_PER_SAMPLE * samples / (ms_mlp * 1e-3) / 1e12
        traffic = load_profile("encode_parts_f16_stereo1080p_pmc.json", "hbm_bytes_per_launch")
        roofline = {
            "kernel": "encode_parts<f16>", "bound": "hbm",
            "achieved": round(gathered_gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(gathered_gbs / PEAK_HBM_GBS, 4), "traffic": traffic,
            # what the counters and per-XCD timings show: the gathers are L1/L2
            # hits, each XCD is bound by the vector L1's gather slots (one lane
            # address per lane and instruction), XCD 0 (level 15, no reuse)
            # also by its L2 request rate; HBM carries far less than 512 B/sample
            "limiter": "vector-L1 gather slots on every XCD, and the L2 request rate of XCD 0 "
                       "(level 15); HBM is not the limiter",
        }
        if traffic:
            hbm = traffic / (ms_enc * 1e-3) / 1e9
            roofline.update({"hbm_measured_GBs": round(hbm, 1),
                             "hbm_measured_frac": round(hbm / PEAK_HBM_GBS, 4)})
        result = {
            "metric": METRIC, "value": round(value, 3), "unit": "Mrays/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(wall_m / args.steps * 1e3, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "fp16",
            "data": "synthetic (seeded scene: N(0,0.1) hash table, Xavier MLPs, 30% occupancy blob)",
            "config": {"workload": "stereo render 1920x1080 x 2 eyes (64 mm IPD), mixed "
                                   "precision: FP16 table, FP16 MFMA MLPs, FP32 compositing",
                       "rays_per_step": rays, "samples_per_step": int(samples),
                       "samples_per_ray": round(samples / rays, 4)},
            "fp32_mrays": round(fp32, 3), "fp32_valu_mrays": round(fp32_valu, 3),
            "speedup_vs_fp32": round(value / fp32, 3),
            "speedup_vs_fp32_valu": round(value / fp32_valu, 3),
            "stage_ms": {"march": round(ms_march, 4), "encode": round(ms_enc, 4),
                         "mlp": round(ms_mlp, 4), "composite": round(ms_comp, 4),
                         "fp32_encode": round(float(np.mean([s.ms_encode for s in st_f])), 4),
                         "fp32_mlp": round(float(np.mean([s.ms_mlp for s in st_f])), 4),
                         "fp32_valu_mlp": round(float(np.mean([s.ms_mlp for s in st_v])), 4)},
            "roofline": roofline,
            "mlp_roofline": {"kernel": "mlp_mixed", "bound": "mfma",
                             "achieved": round(mlp_tflops, 2), "peak": PEAK_FP16_TFLOPS,
                             "unit": "TFLOP/s", "frac": round(mlp_tflops / PEAK_FP16_TFLOPS, 5)},
            "psnr_db": None, "max_depth_err": None, "p99_frame_latency_ms": None,
            "cpu_baseline": None,
        }
        if not args.skip_extra:
            result["ceilings"] = ceilings_leg(
                N, samples, ms_enc,
                load_profile("encode_parts_f16_stereo1080p_l2.json", "l2_requests_per_launch"),
                load_profile("encode_parts_f16_stereo1080p_l2.json", "l1_accesses_per_launch"))
            result["config1_640x480"] = config1_leg(N, table, weights, occ, ctx)
            result["opaque_stereo1080p"] = opaque_leg(N, views)
        if not args.skip_frameprep:
            result["frameprep"] = frameprep_leg(N)
        if not args.skip_train:
            result["training"] = training_leg(N, ctx, occ)
        if not args.skip_serve:
            srv = serving_leg(N, ctx, args.serve_seconds)
            prof = load_profile("serving_8clients_roofline.json", "summary")
            if prof:
                srv["rocprof"] = prof
            result["serving"] = srv
            result["p99_frame_latency_ms"] = round(srv["p99_ms"], 3)
        if not args.skip_cpu:
            base, acc = cpu_baseline(N, table, weights, occ, ctx, views, args.cpu_seconds)
            result["cpu_baseline"] = base
            result["psnr_db"] = acc["mixed_psnr_db"]
            result["max_depth_err"] = acc["mixed_max_depth_err"]
            result["accuracy_vs_fp64"] = acc
            if "frameprep" in result:
                cf = cpu_frameprep_baseline(base["cores"])
                cf["nproc"], cf["cpu_model"] = base["nproc"], base["cpu_model"]
                result["frameprep"]["cpu"] = cf
                result["frameprep"]["gpu_vs_cpu_1t"] = round(
                    cf["ms_per_frame_1t"] / (result["frameprep"]["ms_per_launch"] /
                                             result["frameprep"]["frames_per_launch"]), 1)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
