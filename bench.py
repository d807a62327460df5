"""Benchmark of the MI355X per-patch GoP decoder.

Workload (default): the config-1 stream -- synthetic humanoid, 10-bit grid,
~800k points/frame, ~860 patches of lognormal 200-3000 points, GoP 8,
quantisation step 4 -- four consecutive GoPs of one seeded stream, generated
on the host, packed and resident in HBM before timing.  One step = decode of
one whole GoP (8 frames); the steps cycle over the four GoPs.

Reported: `value` = decoded Mpts/s of the batched persistent decode (whole
GoP per launch), the K steps enqueued back to back between one synchronize
before and one after (`ms_per_step`; `ms_per_step_synced` repeats them with a
synchronize and the library's kernel timing events after each, which is also
where the kernel times of the roofline come from); beside it the per-frame
drop-in decode (launch_cudaProcess,
one launch per frame), frames/s of both, p50/p99 per-frame latency, the HBM
roofline fraction of the dominant kernel, and the CPU oracle decoding a
bounded sample of the same stream on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "patchvvc-plus_amd"))
sys.path.insert(0, str(REPO))

HBM_PEAK_GBS = 8000.0
ENGINE_CLOCK_HZ = 2.4e9   # MI355X peak engine clock
XCDS = 8                  # MI355X compute dies; GRBM_GUI_ACTIVE reads the sum of their cycles
B_OUT_NATIVE = 15   # xyz f32 x3 + RGB8
B_OUT_POINTS = 24   # reference Points AoS


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pct(a, q):
    return float(np.percentile(np.asarray(a, np.float64), q)) if len(a) else float("nan")


def _pmc_means(csv_path: Path, kernel: str) -> dict:
    """Mean value per counter (over the dispatches of `kernel`) of a
    rocprofv3 counter-collection CSV."""
    import csv
    vals: dict[str, list[float]] = {}
    with open(csv_path, newline="") as fh:
        for row in csv.DictReader(fh):
            if kernel in row["Kernel_Name"]:
                vals.setdefault(row.get("Counter_Name", "?"), []).append(float(row["Counter_Value"]))
    return {k: float(np.mean(v)) for k, v in vals.items()}


def measure_counters(args) -> dict | None:
    """Per persistent_kernel launch, from rocprofv3 --pmc passes of this
    script's batched leg (each a child process, before this process touches
    the GPU):
      * HBM bytes: FETCH_SIZE, then WRITE_SIZE, each scaled by what the same
        counter reads for a device copy of 256 MiB;
      * scalar issue: SQ_INSTS_SALU (instructions of the whole launch) with
        GRBM_GUI_ACTIVE (GPU cycles of the launch).
    None when rocprofv3 is absent or a pass fails."""
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if not prof:
        return None
    me = [sys.executable, "-u", str(Path(__file__).resolve())]
    leg = me + ["--only", "batched", "--steps", str(args.gops), "--warmup", "1", "--gops", str(args.gops),
                "--no-cpu", "--no-traffic", "--no-encode", "--no-quality", "--no-c2", "--preset", args.preset,
                "--seed", str(args.seed)]
    for k in ("waves", "lds", "wgs_per_cu", "queue_chunk", "priority_items", "lane_max_points"):
        v = getattr(args, k)
        if v:
            leg += ["--" + k.replace("_", "-"), str(v)]
    copy_mib = 256
    copy = me + ["--calibrate-copy", str(copy_mib)]
    res = {}
    with tempfile.TemporaryDirectory() as d:
        for tag, counters, cmd, kern in (("fetch", ["FETCH_SIZE"], leg, "persistent_kernel"),
                                          ("write", ["WRITE_SIZE"], leg, "persistent_kernel"),
                                          ("salu", ["SQ_INSTS_SALU", "GRBM_GUI_ACTIVE"], leg, "persistent_kernel"),
                                          ("copy_fetch", ["FETCH_SIZE"], copy, "copyBuffer"),
                                          ("copy_write", ["WRITE_SIZE"], copy, "copyBuffer")):
            out = Path(d) / tag
            full = [prof, "--pmc", *counters, "-f", "csv", "-d", str(out), "-o", "p", "--"] + cmd
            log("counter pass:", tag, " ".join(counters))
            try:
                r = subprocess.run(full, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=180)
            except subprocess.TimeoutExpired:
                log("counter pass timed out:", tag)
                return None
            csvs = list(out.rglob("*counter_collection.csv"))
            if r.returncode != 0 or not csvs:
                log("counter pass failed:", tag, r.returncode, r.stdout[-2000:].decode(errors="replace"))
                return None
            v = _pmc_means(csvs[0], kern)
            if not all(c in v for c in counters):
                log("counter pass without values:", tag, v)
                return None
            res[tag] = v
    copy_kib = copy_mib * 1024.0
    f_scale = copy_kib / res["copy_fetch"]["FETCH_SIZE"]
    w_scale = copy_kib / res["copy_write"]["WRITE_SIZE"]
    fetch_b = res["fetch"]["FETCH_SIZE"] * f_scale * 1024.0
    write_b = res["write"]["WRITE_SIZE"] * w_scale * 1024.0
    traffic = {"bytes": fetch_b + write_b, "fetch_bytes": fetch_b, "write_bytes": write_b,
               "fetch_kib_raw": res["fetch"]["FETCH_SIZE"], "write_kib_raw": res["write"]["WRITE_SIZE"],
               "fetch_scale": f_scale, "write_scale": w_scale,
               "calibration": f"device copy of {copy_mib} MiB reads FETCH_SIZE {res['copy_fetch']['FETCH_SIZE']:.1f} "
                              f"KiB, WRITE_SIZE {res['copy_write']['WRITE_SIZE']:.1f} KiB"}
    issue = {"salu_insts": res["salu"]["SQ_INSTS_SALU"], "gpu_cycles": res["salu"]["GRBM_GUI_ACTIVE"]}
    return {"traffic": traffic, "issue": issue}


def serving_leg(args, pvvc) -> dict:
    """Serving workload: streams of mixed density (one third each 100k pts /
    9-bit, 400k / 10-bit, 1M / 10-bit; GoP 8, qp 4), each stream generated
    from its own seed, Poisson arrivals at 30 fps per stream, continuous
    batching on the server, GoP state resident."""
    import concurrent.futures as cf

    from pvvc import serve
    n = args.serve_streams
    kinds = ["c3_100k", "c3_400k", "c3_1m"]
    counts = [n // 3, n // 3, n - 2 * (n // 3)]
    plan = [(kind, 1000 + args.seed * 997 + 31 * i + j) for i, (kind, c) in enumerate(zip(kinds, counts))
            for j in range(c)]
    t0 = time.time()
    workers = max(1, min(4, (os.cpu_count() or 8) // 4))

    def gen(item):
        kind, sd = item
        imgs = pvvc.synth_stream(kind, n_frames=8, seed=sd, threads=4)
        return kind, [pvvc.Frame.parse(b) for b in imgs]

    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        streams = list(ex.map(gen, plan))
    log(f"serving: generated {len(streams)} streams (one seed each) in {time.time() - t0:.1f}s")
    specs = [serve.StreamSpec(kind, frames) for kind, frames in streams]
    srv = pvvc.Server(lanes=args.serve_lanes)
    st = serve.run(srv, specs, fps=30.0, duration=args.serve_seconds, seed=args.seed)
    srv.close()
    st["pts_per_frame"] = {k: int(np.mean([f.points for kk, fr in streams if kk == k for f in fr])) for k in kinds}
    st["streams_per_kind"] = dict(zip(kinds, counts))
    st["seeds"] = f"{plan[0][1]}..{plan[-1][1]} (one per stream)"
    st["sustained"] = bool(st["frames_completed"] == st["frames_offered"] and st["frames_failed"] == 0
                           and st["p99_ms"] <= 33.0)
    for k in ("p50_ms", "p99_ms", "max_ms", "mpts_per_s", "offered_mpts_per_s", "submit_late_p99_ms",
              "queue_wait_p50_ms", "queue_wait_p99_ms", "service_p50_ms", "service_p99_ms", "frames_per_batch",
              "host_ms_per_batch"):
        st[k] = round(float(st[k]), 3)
    log("serving", st)
    return st


def c2_leg(args, pvvc) -> dict:
    """The batched configuration: one 16-frame GoP of the 11-bit stream
    (~3.3M points/frame, patches 50-20000) decoded by one persistent launch
    per step, beside the per-frame drop-in decode of the same frames."""
    gop = pvvc.SYNTH_PRESETS["c2"]["gop"]
    t0 = time.time()
    frames = [pvvc.Frame.parse(b) for b in pvvc.synth_stream("c2", n_frames=gop, seed=args.seed + 50,
                                                              threads=min(16, os.cpu_count() or 8))]
    log(f"c2: generated {gop} frames in {time.time() - t0:.1f}s")
    metas = [f.meta() for f in frames]
    n_patches = max(int(m["index"].max()) for m in metas) + 1
    max_pts = max(int(m["size"].max()) for m in metas)
    dec = pvvc.Decoder(n_patches, max_pts)
    b = dec.batch(frames)
    out = dec.outputs(b, pvvc.OUT_NATIVE)
    dec.decode(b, pvvc.MODE_PERSISTENT, out)
    pvvc.synchronize()
    ok = bool((b.status() == 0).all())
    # back to back between two synchronizes (ms_per_gop), then once more with
    # a synchronize per step for the kernel events
    t_c0 = time.perf_counter()
    for _ in range(args.c2_steps):
        dec.decode(b, pvvc.MODE_PERSISTENT, out)
    pvvc.synchronize()
    c2_ms = (time.perf_counter() - t_c0) * 1e3
    wall, pk = [], []
    prev = pvvc.set_timing(True)
    for _ in range(args.c2_steps):
        ts = time.perf_counter()
        dec.decode(b, pvvc.MODE_PERSISTENT, out)
        pvvc.synchronize()
        wall.append((time.perf_counter() - ts) * 1e3)
        pk.append(sum(pvvc.last_decode_kernel_times()))
    pvvc.set_timing(prev)
    state = pvvc.State(n_patches, max_pts)
    staged = [pvvc.StagedFrame(f) for f in frames]
    buf = pvvc.DeviceBuffer(24 * max(f.points for f in frames))
    for sf in staged:  # warm
        pvvc.launch_dropin(sf, state, buf)
    pvvc.synchronize()
    lat = []
    for sf in staged:
        ts = time.perf_counter()
        pvvc.launch_dropin(sf, state, buf)
        pvvc.synchronize()
        lat.append((time.perf_counter() - ts) * 1e3)
    gop_ms = c2_ms / max(1, args.c2_steps)
    pf_fps = len(lat) / (sum(lat) * 1e-3)
    res = {"frames": gop, "points_per_frame": int(np.mean([f.points for f in frames])),
           "patches_per_frame": int(np.mean([len(m["size"]) for m in metas])),
           "patch_points_min_max": [int(min(m["size"].min() for m in metas)), max_pts],
           "status_ok": ok, "steps": args.c2_steps,
           "batched_mpts_per_s": round(b.total_points / (gop_ms * 1e-3) / 1e6, 1),
           "batched_frames_per_s": round(gop / (gop_ms * 1e-3), 1), "ms_per_gop": round(gop_ms, 3),
           "ms_per_gop_synced": round(float(np.mean(wall)), 3),
           "persistent_kernel_ms": round(float(np.mean(pk)), 3),
           "per_frame_dropin": {"frames_per_s": round(pf_fps, 1), "p50_ms": round(pct(lat, 50), 3),
                                "p99_ms": round(pct(lat, 99), 3)},
           "batched_vs_per_frame": round(gop / (gop_ms * 1e-3) / pf_fps, 3)}
    log("c2", res)
    return res


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--preset", default="c1")
    ap.add_argument("--gops", type=int, default=4,
                    help="distinct GoPs (consecutive GoPs of one seeded stream), cycled over the steps")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="bound of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--waves", type=int, default=0)
    ap.add_argument("--lds", type=int, default=0)
    ap.add_argument("--wgs-per-cu", type=int, default=0)
    ap.add_argument("--queue-chunk", type=int, default=0)
    ap.add_argument("--priority-items", type=int, default=0)
    ap.add_argument("--lane-max-points", type=int, default=0,
                    help="decode the RLGR of slices up to this many points one slice per lane (default off)")
    ap.add_argument("--autotune", action="store_true", help="run the decoder autotuner on the first GoP before timing")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 FETCH/WRITE passes")
    ap.add_argument("--no-serving", action="store_true", help="skip the multi-stream serving leg")
    ap.add_argument("--no-encode", action="store_true", help="skip the GPU colour-encoder leg")
    ap.add_argument("--no-quality", action="store_true", help="skip the MSE / PSNR leg")
    ap.add_argument("--no-c2", action="store_true", help="skip the c2 (batched, 11-bit, GoP 16) leg")
    ap.add_argument("--c2-steps", type=int, default=5)
    ap.add_argument("--serve-streams", type=int, default=64)
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--serve-lanes", type=int, default=3)
    ap.add_argument("--only", choices=["all", "batched", "perframe", "serving"], default="all",
                    help="restrict the GPU legs (profiling runs)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; the multi-GPU path is design-only "
              "(patches shard by stream / patch-index range with no exchange).", file=sys.stderr)
        return 2

    counters = None
    if not args.no_traffic and not args.calibrate_copy and args.only == "all":
        counters = measure_counters(args)
        log("counters", counters)
    traffic = counters["traffic"] if counters else None

    import pvvc

    if args.calibrate_copy:
        import ctypes as C
        nbytes = args.calibrate_copy << 20
        src, dst = pvvc.DeviceBuffer(nbytes), pvvc.DeviceBuffer(nbytes)
        src.zero()
        for _ in range(3):
            pvvc.hip_lib().pvvc_memcpy_d2d(C.c_void_p(dst.ptr), C.c_void_p(src.ptr), nbytes)
        pvvc.synchronize()
        print(json.dumps({"calibrate_copy_bytes": nbytes, "copies": 3}))
        return 0

    info = pvvc.device_info()
    log("device", info)
    if args.only == "serving":
        print(json.dumps({"serving": serving_leg(args, pvvc)}))
        return 0
    gop = pvvc.SYNTH_PRESETS[args.preset]["gop"]
    t0 = time.time()
    images, sources = pvvc.synth_stream(args.preset, n_frames=gop * args.gops, seed=args.seed,
                                        threads=min(16, os.cpu_count() or 8), keep_source=True)
    log(f"generated {len(images)} frames in {time.time() - t0:.1f}s")
    frames = [pvvc.Frame.parse(b) for b in images]
    metas = [f.meta() for f in frames]
    n_patches = max(int(m["index"].max()) for m in metas) + 1
    max_pts = max(int(m["size"].max()) for m in metas)
    pts_per_frame = [f.points for f in frames]

    # ---------------- batched persistent decode (the headline) ----------------
    run_b = args.only in ("all", "batched")
    run_p = args.only in ("all", "perframe")
    dec = pvvc.Decoder(n_patches, max_pts)
    tun = {k: v for k, v in dict(waves_per_patch=args.waves, lds_bytes=args.lds, wgs_per_cu=args.wgs_per_cu,
                                 queue_chunk=args.queue_chunk, priority_items=args.priority_items,
                                 lane_max_points=args.lane_max_points).items() if v}
    if tun:
        dec.set_tuning(**tun)
    batches = [dec.batch(frames[g * gop:(g + 1) * gop]) for g in range(args.gops)]
    outs = [dec.outputs(b, pvvc.OUT_NATIVE) for b in batches]
    autotuned = None
    if args.autotune and run_b:
        t_at = time.time()
        best, best_ms = dec.autotune(batches[0], pvvc.MODE_PERSISTENT, outs[0], reps=3)
        autotuned = {"tuning": best, "ms": round(best_ms, 4), "seconds": round(time.time() - t_at, 1)}
        log("autotune", autotuned)
    for i in range(args.warmup if run_b else 0):
        dec.decode(batches[i % args.gops], pvvc.MODE_PERSISTENT, outs[i % args.gops])
    pvvc.synchronize()
    bad = [int((b.status() != 0).sum()) for b in batches]
    if any(bad):
        log("slice errors in warmup:", bad)
    # The timed run: K steps (one whole GoP each, the four GoPs in turn)
    # enqueued back to back on the decoder's stream, one synchronize before
    # and one after, so that the host's launch and wait per step are not in
    # the figure (value, ms_per_step).
    pvvc.synchronize()
    t_b0 = time.perf_counter()
    for i in range(args.steps if run_b else 0):
        dec.decode(batches[i % args.gops], pvvc.MODE_PERSISTENT, outs[i % args.gops])
    pvvc.synchronize()
    tot_ms = (time.perf_counter() - t_b0) * 1e3 if run_b and args.steps else float("nan")
    # The same K steps once more with a synchronize after each and the
    # library's timing events on (they are off in the timed run: each event
    # is a marker between kernels): the kernel time of every launch
    # (roofline) and the wall per step with the host's round trip and the
    # events in it (ms_per_step_synced).
    kernel_ms, wall, pk_ms, pk_launches = [], [], [], 0
    timing_prev = pvvc.set_timing(True)
    for i in range(args.steps if run_b else 0):
        b = batches[i % args.gops]
        ts = time.perf_counter()
        dec.decode(b, pvvc.MODE_PERSISTENT, outs[i % args.gops])
        pvvc.synchronize()
        wall.append((time.perf_counter() - ts) * 1e3)
        kernel_ms.append(pvvc.last_kernel_ms())
        kt = pvvc.last_decode_kernel_times()
        pk_ms.append(sum(kt))
        pk_launches += len(kt)
    pvvc.set_timing(timing_prev)
    step_pts = [b.total_points for b in batches]
    tot_pts = sum(step_pts[i % args.gops] for i in range(args.steps))
    kernel_ms = kernel_ms or [float("nan")]
    batched_mpts = tot_pts / (tot_ms * 1e-3) / 1e6
    batched_fps = args.steps * gop / (tot_ms * 1e-3)
    in_bytes = [b.input_bytes for b in batches]
    alg_bytes = [in_bytes[g] + B_OUT_NATIVE * step_pts[g] for g in range(args.gops)]
    avg_kernel_ms = float(np.mean(kernel_ms))
    # the dominant kernel: persistent_kernel launches alone, from their own events
    n_launch = max(1, pk_launches)
    pk_avg_ms = float(np.sum(pk_ms)) / n_launch if pk_ms else float("nan")
    alg_per_launch = float(np.sum([alg_bytes[i % args.gops] for i in range(args.steps)])) / n_launch
    achieved_gbs = alg_per_launch / (pk_avg_ms * 1e-3) / 1e9
    # resident GoP state read by predictive slices: intra geometry + reference YUV (12 B/pt)
    state_bytes = [b.state_bytes for b in batches]
    sym_per_step = [3 * p for p in step_pts]
    cu = int(info["cu"])
    tot_sym = sum(sym_per_step[i % args.gops] for i in range(args.steps))
    sym_per_s_cu = tot_sym / (float(np.sum(pk_ms)) * 1e-3) / cu if pk_ms else float("nan")
    # decode-issue bound: the RLGR loop runs on the scalar unit, one scalar
    # instruction per cycle per CU at most; SQ_INSTS_SALU per launch over that
    # capacity at the peak engine clock is the launch's floor
    issue = None
    if counters:
        salu = counters["issue"]["salu_insts"]
        cyc = counters["issue"]["gpu_cycles"]
        sym_launch = float(np.mean(sym_per_step))
        peak_rate = cu * ENGINE_CLOCK_HZ
        issue = {"bound": "salu", "unit": "G scalar instructions/s",
                 "salu_insts_per_launch": int(salu),
                 "salu_per_symbol": round(salu / sym_launch, 3),
                 "achieved": round(salu / (pk_avg_ms * 1e-3) / 1e9, 2),
                 "peak": round(peak_rate / 1e9, 2),
                 "frac": round(salu / (pk_avg_ms * 1e-3) / peak_rate, 4),
                 "busy_frac_profiled": round(salu * XCDS / (cu * cyc), 4) if cyc else None,
                 "floor_ms": round(salu / peak_rate * 1e3, 4),
                 "cycles_per_xcd_profiled": int(cyc / XCDS)}
    log(f"persistent: {avg_kernel_ms:.3f} ms/GoP all launches, {pk_avg_ms:.3f} ms per persistent_kernel, "
        f"{np.mean(wall):.3f} ms wall, {batched_mpts:.1f} Mpts/s")

    # ---------------- per-frame drop-in decode (launch_cudaProcess) ----------------
    state = pvvc.State(n_patches, max_pts)
    staged = [pvvc.StagedFrame(f) for f in frames]
    out_pts = pvvc.DeviceBuffer(24 * max(pts_per_frame))
    lat = []
    frame_kernel_ms = []
    for i in range(min(len(frames), gop) if run_p else 0):  # warm: one GoP
        pvvc.launch_dropin(staged[i], state, out_pts)
    pvvc.synchronize()
    n_pf = args.steps * gop if run_p else 0
    t_pf0 = time.perf_counter()
    for i in range(n_pf):
        k = i % len(frames)
        ts = time.perf_counter()
        pvvc.launch_dropin(staged[k], state, out_pts)
        pvvc.synchronize()
        lat.append((time.perf_counter() - ts) * 1e3)
    pf_ms = (time.perf_counter() - t_pf0) * 1e3
    # kernel times from the library's events, in a pass of their own
    timing_prev = pvvc.set_timing(True)
    for i in range(min(n_pf, 2 * gop)):
        pvvc.launch_dropin(staged[i % len(frames)], state, out_pts)
        pvvc.synchronize()
        frame_kernel_ms.append(pvvc.last_kernel_ms())
    pvvc.set_timing(timing_prev)
    pf_pts = sum(pts_per_frame[i % len(frames)] for i in range(n_pf))
    pf_fps = n_pf / (pf_ms * 1e-3) if n_pf else float("nan")
    pf_mpts = pf_pts / (pf_ms * 1e-3) / 1e6 if n_pf else float("nan")
    frame_kernel_ms = frame_kernel_ms or [float("nan")]
    log(f"per-frame drop-in: {np.mean(frame_kernel_ms):.3f} ms kernel, p50 {pct(lat, 50):.3f} p99 {pct(lat, 99):.3f} ms,"
        f" {pf_fps:.1f} fps")

    # ---------------- end to end per frame: container + H2D + decode ----------------
    # The reference harness times its per-slice copies plus the decode
    # (test_decode_only.cc:93-158, t0..t3); here one packed pinned copy per
    # frame.  "with_parse" also parses the .frame image (zstd unwrap) on the
    # host; "pipelined" parses frame k+1 while the GPU decodes frame k.
    e2e = None
    if run_p:
        cap = max(pvvc.Stager.bytes_for(f) for f in frames)
        stagers = [pvvc.Stager(cap), pvvc.Stager(cap)]
        state_e = pvvc.State(n_patches, max_pts)
        n_e = args.steps * gop
        seq = [i % len(images) for i in range(n_e)]
        for k in range(min(gop, len(images))):  # warm
            pvvc.launch_dropin(stagers[0].stage(frames[k]), state_e, out_pts)
            pvvc.synchronize()
        lat_ref, lat_all = [], []
        for i, k in enumerate(seq):
            t0 = time.perf_counter()
            f = pvvc.Frame.parse(images[k])
            t1 = time.perf_counter()
            pvvc.launch_dropin(stagers[0].stage(f), state_e, out_pts)
            pvvc.synchronize()
            t2 = time.perf_counter()
            lat_ref.append((t2 - t1) * 1e3)
            lat_all.append((t2 - t0) * 1e3)
        t0 = time.perf_counter()
        nxt = pvvc.Frame.parse(images[seq[0]])
        for i in range(n_e):
            f = nxt
            pvvc.launch_dropin(stagers[i & 1].stage(f), state_e, out_pts)
            if i + 1 < n_e:
                nxt = pvvc.Frame.parse(images[seq[i + 1]])
            pvvc.synchronize()
        pipe_s = time.perf_counter() - t0
        e2e = {"frames": n_e,
               "h2d_decode": {"frames_per_s": round(1e3 / float(np.mean(lat_ref)), 1),
                              "p50_ms": round(pct(lat_ref, 50), 4), "p99_ms": round(pct(lat_ref, 99), 4)},
               "with_parse": {"frames_per_s": round(1e3 / float(np.mean(lat_all)), 1),
                              "p50_ms": round(pct(lat_all, 50), 4), "p99_ms": round(pct(lat_all, 99), 4)},
               "pipelined_frames_per_s": round(n_e / pipe_s, 1),
               "pipelined_mpts_per_s": round(sum(pts_per_frame[k] for k in seq) / pipe_s / 1e6, 2),
               "staged_bytes_per_frame": int(np.mean([pvvc.Stager.bytes_for(f) for f in frames]))}
        for st_ in stagers:
            st_.close()
        log("end to end", e2e)

    # ---------------- forward path: GPU colour encoder ----------------
    # Re-encodes the GoP's colours: frame 0's octree per patch, leaf YUV of the
    # decoded frame (predictive frames: the residual against frame 0), the
    # stream's qp.  The CPU leg is the stream writer's encoder on the same
    # octrees (RAHT + RLGR, one thread) over frame 0, whose bytes the GPU
    # output must equal.
    encode = None
    if run_b and not args.no_encode:
        _, rgb_g = dec.decode_host(batches[0], pvvc.MODE_PERSISTENT, pvvc.OUT_NATIVE)
        rgbf = rgb_g.astype(np.float32)
        r_, g_, b_ = rgbf[:, 0], rgbf[:, 1], rgbf[:, 2]
        yuv = np.stack([r_ * np.float32(0.299) + g_ * np.float32(0.587) + b_ * np.float32(0.114),
                        r_ * np.float32(-0.168736) - g_ * np.float32(0.331264) + b_ * np.float32(0.5) + np.float32(128),
                        r_ * np.float32(0.5) - g_ * np.float32(0.418688) - b_ * np.float32(0.081312) + np.float32(128)],
                       1).astype(np.float32)
        m0 = metas[0]
        geo_of = {int(m0["index"][s_]): frames[0].geometry(s_) for s_ in range(frames[0].slice_cnt)}
        base = {}
        o = 0
        for s_ in range(frames[0].slice_cnt):
            base[int(m0["index"][s_])] = (o, int(m0["size"][s_]))
            o += int(m0["size"][s_])
        # packed job arrays (host-resident inputs, built before the timed calls)
        geos, npts_e, qps, base_idx = [], [], [], np.full(len(yuv), -1, np.int64)
        for f in range(gop):
            fo = batches[0].frame_offset[f]
            mf = metas[f]
            o = 0
            for s_ in range(frames[f].slice_cnt):
                p_, n_ = int(mf["index"][s_]), int(mf["size"][s_])
                if f:
                    b0, n0 = base[p_]
                    base_idx[fo + o: fo + o + n_] = np.arange(b0, b0 + n0)
                geos.append(geo_of[p_])
                npts_e.append(n_)
                qps.append(int(mf["qp"][s_]))
                o += n_
        yuv_jobs = yuv.copy()
        pm = base_idx >= 0
        yuv_jobs[pm] -= yuv[base_idx[pm]]
        geo_all = np.frombuffer(b"".join(geos), np.uint8)
        gsz_e = np.array([len(g) for g in geos], np.int64)
        npts_e = np.array(npts_e, np.int64)
        qps = np.array(qps, np.uint8)
        first_t = time.perf_counter()
        pvvc.encode_colors_packed(geo_all, gsz_e, yuv_jobs, npts_e, qps)  # sizes the kept buffers
        first_s = time.perf_counter() - first_t
        walls, kms = [], []
        for _ in range(5):
            t_e = time.perf_counter()
            out_all, off_e, size_e, st_e, ms_e = pvvc.encode_colors_packed(geo_all, gsz_e, yuv_jobs, npts_e, qps)
            walls.append(time.perf_counter() - t_e)
            kms.append(ms_e)
        wall_e, ms_e = float(np.median(walls)), float(np.median(kms))
        pts_e = int(npts_e.sum())
        ns0 = frames[0].slice_cnt
        poff_e = np.concatenate([[0], np.cumsum(npts_e)[:-1]])
        t_c = time.perf_counter()
        equal, cpu_pts = 0, 0
        for s_ in range(ns0):
            leaf_j = yuv_jobs[poff_e[s_]: poff_e[s_] + npts_e[s_]]
            col = pvvc.encode_colors_cpu_on_tree(geos[s_], leaf_j, int(qps[s_]))
            equal += col == out_all[off_e[s_]: off_e[s_] + size_e[s_]].tobytes()
            cpu_pts += len(leaf_j)
        cpu_s = time.perf_counter() - t_c
        encode = {"streams": len(geos), "points": pts_e, "kernel_ms": round(ms_e, 3),
                  "gpu_mpts_per_s_kernel": round(pts_e / (ms_e * 1e-3) / 1e6, 1),
                  "call_ms_median_of_5": round(wall_e * 1e3, 3),
                  "gpu_mpts_per_s_with_copies": round(pts_e / wall_e / 1e6, 1),
                  "first_call_ms": round(first_s * 1e3, 3),
                  "status_ok": int((st_e == 0).sum()),
                  "cpu_writer_mpts_per_s_1thread": round(cpu_pts / cpu_s / 1e6, 2),
                  "frame0_streams_equal_to_cpu": f"{equal}/{ns0}"}
        log("encode", encode)

    # ---------------- quality: MSE / PSNR of decoded frames vs their source ----------------
    # MSE::Compute + the PSNR PVVCDecompression prints (geometry peak 1024, Y
    # peak 256), on the GPU over the batched output of frames 0 and 1; the CPU
    # oracle's grid search on frame 0 gives the same neighbours and its time.
    quality = None
    if run_b and not args.no_quality:
        xyz_q, rgb_q = dec.decode_host(batches[0], pvvc.MODE_PERSISTENT, pvvc.OUT_NATIVE)
        per = []
        for f in range(min(2, gop)):
            o_, n_ = batches[0].frame_offset[f], batches[0].frame_points[f]
            dcl = (xyz_q[o_:o_ + n_], rgb_q[o_:o_ + n_])
            pvvc.quality.cloud_mse(sources[f], dcl)  # warm
            r_ = pvvc.quality.cloud_mse(sources[f], dcl, want_nn=not args.no_cpu and f == 0)
            row = {"frame": f, "source_points": len(sources[f][0]), "decoded_points": n_,
                   "geo_psnr_db": round(pvvc.quality.psnr(r_["geo"], pvvc.quality.GEO_PEAK), 3),
                   "y_psnr_db": round(pvvc.quality.psnr(r_["y"], pvvc.quality.Y_PEAK), 3),
                   "gpu_ms": round(r_["kernel_ms"], 3)}
            if not args.no_cpu and f == 0:
                from oracle import binding
                t_q = time.perf_counter()
                oq = binding.cloud_mse(*sources[f], *dcl, threads=min(16, os.cpu_count() or 8), want_nn=True)
                row["cpu_oracle_ms"] = round((time.perf_counter() - t_q) * 1e3, 1)
                row["neighbours_equal_to_oracle"] = bool(np.array_equal(oq["nn_p"], r_["nn_p"]) and
                                                         np.array_equal(oq["nn_q"], r_["nn_q"]))
            per.append(row)
        quality = {"frames": per}
        log("quality", quality)

    # ---------------- the batched configuration c2 ----------------
    c2 = None
    if run_b and not args.no_c2 and args.only == "all":
        c2 = c2_leg(args, pvvc)

    # ---------------- serving loop: many concurrent streams ----------------
    serving = None
    if not args.no_serving and args.only == "all":
        serving = serving_leg(args, pvvc)

    # ---------------- CPU baseline: oracle on a bounded sample ----------------
    cpu, accuracy = None, None
    if not args.no_cpu:
        import tempfile

        from oracle import binding

        threads = min(16, os.cpu_count() or 1)
        od = binding.StreamDecoder(fp64=False)
        with tempfile.TemporaryDirectory() as d:
            done_pts, done_ms, done_ms_ref, nfr = 0, 0.0, 0.0, 0
            frame_ms = []
            bufs = {"xyz": np.zeros((max(pts_per_frame), 3), np.float32),
                    "rgb": np.zeros((max(pts_per_frame), 3), np.uint8),
                    "status": np.zeros(max(f.slice_cnt for f in frames), np.int32)}
            paths = []
            for k, img in enumerate(images):
                p = Path(d) / f"cpu_{k}.frame"
                p.write_bytes(img)
                paths.append(p)
            t_start = time.time()
            while time.time() - t_start < args.cpu_seconds and nfr < 40 * len(frames):
                k = nfr % len(frames)
                of = binding.OracleFrame(str(paths[k]))
                ms_ref, ms_all = od.decode_timing_only(of, threads, bufs)
                done_ms += ms_all
                done_ms_ref += ms_ref
                frame_ms.append(ms_all)
                done_pts += pts_per_frame[k]
                nfr += 1
            cpu = {"value": round(done_pts / (done_ms * 1e-3) / 1e6, 3), "unit": "Mpts/s", "cores": threads,
                   "kind": "port",
                   "sample": f"{nfr} frames ({done_pts} points) of the same stream in decode order; "
                             f"SetSlice+GetPatch+motion wall time per frame, {threads} threads",
                   "value_reference_scope": round(done_pts / (done_ms_ref * 1e-3) / 1e6, 3),
                   "frames_per_s": round(nfr / (done_ms * 1e-3), 2),
                   "p50_frame_ms": round(pct(frame_ms, 50), 3), "p99_frame_ms": round(pct(frame_ms, 99), 3),
                   "cpu_model": cpu_model(), "nproc": os.cpu_count()}
            log("cpu baseline", cpu)

            # colour accuracy of the batched GPU output against the FP64 inverse-RAHT
            # oracle on the first two frames of the GoP (one intra, one predictive)
            if run_b:
                xyz_g, rgb_g = dec.decode_host(batches[0], pvvc.MODE_PERSISTENT, pvvc.OUT_NATIVE)
                o64 = binding.StreamDecoder(fp64=True)
                diffs, exact, geo_ok, cnt = [], 0, True, 0
                for k in range(min(2, gop)):
                    ref = o64.decode(binding.OracleFrame(str(paths[k])), threads)
                    o, n = batches[0].frame_offset[k], batches[0].frame_points[k]
                    geo_ok &= bool(np.array_equal(xyz_g[o:o + n], ref["xyz"]))
                    exact += int((rgb_g[o:o + n] == ref["rgb"]).all(1).sum())
                    cnt += n
                    diffs.append(rgb_g[o:o + n].astype(np.float64) - np.clip(ref["rgb64"], 0.0, 255.0))
                dd = np.concatenate(diffs)
                accuracy = {"frames": min(2, gop), "points": cnt, "geometry_bit_exact": geo_ok,
                            "rgb8_exact_vs_fp32_oracle": round(exact / max(1, cnt), 6),
                            "max_abs_err_vs_fp64": round(float(np.abs(dd).max()), 4),
                            "rms_err_vs_fp64": round(float(np.sqrt((dd ** 2).mean())), 4)}
                log("accuracy", accuracy)

    res = {
        "metric": "decoded Mpts/s (batched persistent GoP decode), 1 MI355X",
        "value": round(batched_mpts, 2),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(tot_ms / args.steps, 4),
        "ms_per_step_synced": round(float(np.mean(wall)), 4) if wall else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+int",
        "data": "synthetic (seeded humanoid stream, generated on the host)",
        "config": {"workload": f"{args.preset}: synthetic humanoid, GoP {gop}, "
                               f"{int(np.mean(pts_per_frame))} pts/frame, {len(metas[0]['size'])} patches/frame",
                   "frames_per_step": gop, "points_per_step": int(np.mean(step_pts)), "gops": args.gops,
                   "patches_per_frame": int(np.mean([len(m['size']) for m in metas])),
                   "tuning": dec.tuning(), "autotuned": autotuned},
        "frames_per_s": round(batched_fps, 1),
        "per_frame": {"frames_per_s": round(pf_fps, 1), "Mpts_per_s": round(pf_mpts, 2),
                      "kernel_ms_mean": round(float(np.mean(frame_kernel_ms)), 4),
                      "p50_ms": round(pct(lat, 50), 4), "p99_ms": round(pct(lat, 99), 4)},
        "batched_vs_per_frame": round(batched_fps / pf_fps, 3),
        "p99_frame_latency_ms": round(pct(lat, 99), 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved_gbs / HBM_PEAK_GBS, 5),
                     "traffic": round(traffic["bytes"]) if traffic else None,
                     "kernel": "persistent_kernel", "alg_bytes_per_launch": int(alg_per_launch),
                     "kernel_ms": round(pk_avg_ms, 4), "launches": pk_launches,
                     "traffic_detail": traffic, "issue": issue},
        "state_bytes_per_step": int(np.mean(state_bytes)),
        "c2": c2,
        "rlgr_symbols_per_s_per_cu": round(sym_per_s_cu, 1),
        "end_to_end_per_frame": e2e,
        "encode": encode,
        "quality": quality,
        "accuracy": accuracy,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(_finite(res), allow_nan=False))
    return 0


def _finite(x):
    """Legs that did not run leave NaN; JSON has no NaN, so they become null."""
    if isinstance(x, dict):
        return {k: _finite(v) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return [_finite(v) for v in x]
    if isinstance(x, float) and not np.isfinite(x):
        return None
    return x


if __name__ == "__main__":
    sys.exit(main())
