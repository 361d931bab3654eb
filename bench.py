"""Benchmark of the align_group_to_template hot path on one MI355X.

Workload (the headline configuration): 1000 frames of 512x512 uint16 in
groups of 50 frames, knots every 16 lines (33 knots), default alignment
settings (steady_scan.m:836-867), synthetic data (seeded generator), with
frames, template and geometry resident in HBM.  One step aligns all 20
groups and produces every output of the call, the masks included (written
in place into device buffers).

The headline streams the K steps through the continuous-batching server
(ss_server_*): every step has its own device copy of the frames and its own
output buffers, and step s is submitted once every group of step s-1 has
been admitted, so at most about two steps are resident and a step's slowest
groups share their ticks with the next step's groups.  Beside it:
barrier_per_step (all groups of a step in one batched submission, a
synchronisation per step), host_input_per_step (the block matrices in host
memory, uploaded inside the step: the mex boundary's PCIe-inclusive case),
end_to_end (a batched step followed by ss_resample of its 1000 frames,
device to device, with the bytes of the measurement plan), the
single-group path (p50/p99 group latency and its frames/s, and the ratio of
the headline to it), the batched uneven-groups workload and the serving
workload.

Precision: 'D' (the reference's default, FP64 pixel arithmetic) is the
headline; 'S' (FP32 pixel arithmetic, FP64 reductions and solve) and the
mixed-precision mode are timed beside it.

CPU baseline: the FP64 oracle (the C restatement) on the host cores of the
machine that holds the GPU, on group 0 with OpenMP and on the first frames of
group 0 with one thread; and the same source built with FP32 pixels, OpenMP.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "frames/s aligned at 512² + p50/p99 group latency, at ≤0.01 px RMS vs FP64"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "align_512_u16_traffic.md")


def measured_traffic(kernel, precision):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, copy-calibrated) on this workload, from the
    committed profile summary; None when absent."""
    key = f"hbm_bytes_per_launch[{kernel},{precision}]:"
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.strip().startswith(key):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def measured_limiter(kernel, precision):
    """Pipeline busy fractions of `kernel` (rocprofv3 --pmc) from the committed
    counter summary; None when absent."""
    out = {}
    try:
        with open(os.path.join(ROOT, "profiles", "align_512_u16_D_counters.md")) as f:
            for line in f:
                for key in ("td_busy", "valu_busy"):
                    tag = f"{key}[{kernel},{precision}]:"
                    if line.strip().startswith(tag):
                        out[key] = float(line.split(":", 1)[1])
    except OSError:
        pass
    if not out:
        return None
    out["source"] = "profiles/align_512_u16_D_counters.md (TD = texture-data return of the template gathers)"
    return out


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=1000)
    ap.add_argument("--group", type=int, default=50)
    ap.add_argument("--candidates", type=int, default=2)
    ap.add_argument("--precision", choices=["D", "S", "SB", "all"], default="all",
                    help="pixel precision of the timed runs (SB: 'S' with a BF16 template); "
                         "all = D (the headline), S, SB")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-latency", action="store_true", help="skip the single-group latency leg")
    ap.add_argument("--no-host-input", action="store_true",
                    help="skip the step with frames in host memory (uploaded inside the timed step)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving (continuous batching) leg")
    ap.add_argument("--no-uneven", action="store_true",
                    help="skip the batched uneven-groups workload (256 groups, 8-96 frames, 256^2/512^2)")
    ap.add_argument("--no-end-to-end", action="store_true",
                    help="skip the step followed by the resampling of its frames")
    ap.add_argument("--serve-seconds", type=float, default=4.0,
                    help="simulated acquisition time of the serving leg")
    ap.add_argument("--no-ibs", action="store_true",
                    help="skip the reconstruction (invert_bilinear_sampling) leg")
    ap.add_argument("--ibs-only", action="store_true",
                    help="run only the reconstruction leg (knots from one alignment step)")
    ap.add_argument("--stream-groups", type=int, default=-1,
                    help="time the steps as a stream through the continuous-batching server with at most "
                         "this many resident groups (-1: two steps' groups; 0: off, one barrier per step)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain device copy of known size (rocprof counter calibration)")
    return ap.parse_args()


def calibrate_copy():
    import torch
    n = 256 << 20
    a = torch.empty(n, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "copy", "bytes_read": n, "bytes_written": n, "launches": 5}))


def tuning_summary(entries):
    """The autotuner's choices per group shape: the chosen (samples, waves)
    variant of each pixel kernel and its trial-tick time against the default
    variant's."""
    out = []
    for e in entries:
        ks = {}
        gain = 0.0
        for name, k in e["kernels"].items():
            c = k["chosen"]
            ks[name] = {"chosen": {"samples": k["variants"][c][0], "waves": k["variants"][c][1]},
                        "trial_us": round(k["trial_us"][c], 1), "default_trial_us": round(k["trial_us"][0], 1)}
            gain += k["trial_us"][0] - k["trial_us"][c]
        out.append({"shape": f"{e['precision']}{'+bf16' if e['mixed'] else ''} frame_h {e['frame_h']} "
                             f"n_blocks {e['n_blocks']} frames {e['frames']}",
                    "kernels": ks, "trial_tick_us_saved": round(gain, 1)})
    return out


def serve_leg(args, torch, ssa, W):
    """Continuous batching of acquisition streams (16 x 512x512 at 30 Hz +
    2 x 1024x1024 at 15 Hz, int16, groups of 16-64 frames, jittered
    arrivals, 64 GiB budget): groups are submitted to a server as they
    arrive in real time; latency = completion - arrival."""
    t0 = time.time()
    streams, groups = W.config4(duration=args.serve_seconds)
    print(f"[bench] serving: {len(streams)} streams, {len(groups)} groups generated in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    h = ssa.Handle(0)
    geoms, calls = {}, []
    tids = []
    for st in streams:
        T = st["T"]
        xg, yg = W.gradients(T)
        tids.append((h.upload_template(T, xg, yg), xg, yg))
        size = st["size"]
        if size not in geoms:
            g = ssa.init_settings(size, size, T.shape, nknots=size // 16 + 1)
            geoms[size] = (g, h.upload_geometry(g["tfrac_bm"], g["basex_bm"], g["basey_bm"], g["base_mask_bm"]))
    dev = []
    for gr in groups:
        st = streams[gr["stream"]]
        geom, gid = geoms[st["size"]]
        nk = geom["nknots"]
        bm = ssa.image2blockmatrix(st["frames"][:, :, gr["first"]:gr["first"] + gr["count"]], geom)
        d = torch.from_numpy(bm.ravel(order="F")).to("cuda")
        dev.append(d)
        tid, xg, yg = tids[gr["stream"]]
        a = dict(flags="D", template_image=tid, xgrad=xg, ygrad=yg, tfrac=gid,
                 basex=geom["basex_bm"], basey=geom["basey_bm"], base_mask=geom["base_mask_bm"],
                 bmData=(d.data_ptr(), np.int16), est_dx_knots=np.zeros((nk, gr["count"])),
                 est_dy_knots=np.zeros((nk, gr["count"])), group_size=2 * (nk - 1), nsetframes=gr["count"],
                 min_points_per_block=geom["min_points_per_block"].astype(np.float64))
        a.update(W.default_args())
        calls.append(a)
    torch.cuda.synchronize()
    srv = ssa.Server(handle=h, hbm_budget_bytes=64 << 30, linesearch_candidates=args.candidates)
    lat, done, nxt, steps, resident_max = {}, 0, 0, 0, 0
    tickets = {}
    start = time.perf_counter()
    busy = 0.0
    while done < len(groups):
        now = time.perf_counter() - start
        while nxt < len(groups) and groups[nxt]["arrival"] <= now:
            tickets[srv.submit(calls[nxt])] = nxt
            nxt += 1
        if len(tickets) > done:
            t = time.perf_counter()
            nr, nq = srv.step()
            busy += time.perf_counter() - t
            steps += 1
            resident_max = max(resident_max, nr)
            end = time.perf_counter() - start
            for tk, r in srv.completed().items():
                if isinstance(r, Exception):
                    raise r
                gi = tickets[tk]
                lat[gi] = end - groups[gi]["arrival"]
                done += 1
        elif nxt < len(groups):
            time.sleep(max(0.0, min(0.002, groups[nxt]["arrival"] - now)))
    makespan = time.perf_counter() - start
    srv.close()
    ms = np.array([lat[i] for i in range(len(groups))]) * 1e3
    nfr = sum(g["count"] for g in groups)
    px = sum(g["count"] * streams[g["stream"]]["size"] ** 2 for g in groups)
    return {"streams": "16 x 512x512 @ 30 Hz + 2 x 1024x1024 @ 15 Hz, int16, precision D",
            "simulated_seconds": args.serve_seconds, "groups": len(groups), "frames": nfr,
            "p50_group_ms": float(np.percentile(ms, 50)), "p99_group_ms": float(np.percentile(ms, 99)),
            "max_group_ms": float(ms.max()), "frames_per_s": nfr / makespan,
            "megapixels_per_s": px / makespan / 1e6, "offered_megapixels_per_s": px / args.serve_seconds / 1e6,
            "steps": steps, "max_resident_groups": resident_max, "device_busy_fraction": busy / makespan}


def uneven_leg(args, torch, ssa, W):
    """256 groups of U{8..96} frames, half 256x256 and half 512x512, knots
    every 8 or 16 lines, each with its own template, all in one batched
    submission (precision D); group 0 is checked against the FP64 oracle."""
    t0 = time.time()
    groups = W.config2()
    print(f"[bench] uneven: {len(groups)} groups generated in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    h = ssa.Handle(0)
    geoms = {}
    calls, dev = [], []
    for gr in groups:
        T, size, lpk = gr["T"], gr["size"], gr["lines_per_knot"]
        key = (size, lpk)
        if key not in geoms:
            g = ssa.init_settings(size, size, T.shape, nknots=size // lpk + 1)
            geoms[key] = (g, h.upload_geometry(g["tfrac_bm"], g["basex_bm"], g["basey_bm"], g["base_mask_bm"]))
        geom, gid = geoms[key]
        xg, yg = W.gradients(T)
        tid = h.upload_template(T, xg, yg)
        n = gr["frames"].shape[2]
        nk = geom["nknots"]
        bm = ssa.image2blockmatrix(gr["frames"], geom)
        d = torch.from_numpy(bm.view(np.int16).ravel(order="F")).to("cuda")
        dev.append(d)
        a = dict(flags="D", template_image=tid, xgrad=xg, ygrad=yg, tfrac=gid,
                 basex=geom["basex_bm"], basey=geom["basey_bm"], base_mask=geom["base_mask_bm"],
                 bmData=(d.data_ptr(), np.uint16), est_dx_knots=np.zeros((nk, n)), est_dy_knots=np.zeros((nk, n)),
                 group_size=2 * (nk - 1), nsetframes=n,
                 min_points_per_block=geom["min_points_per_block"].astype(np.float64))
        a.update(W.default_args())
        calls.append(a)
    torch.cuda.synchronize()
    res = ssa.align_groups(calls, handle=h, linesearch_candidates=args.candidates, with_mask=False)
    t = time.perf_counter()
    res = ssa.align_groups(calls, handle=h, linesearch_candidates=args.candidates, with_mask=False)
    el = time.perf_counter() - t
    nfr = sum(gr["frames"].shape[2] for gr in groups)
    px = sum(gr["frames"].shape[2] * gr["size"] ** 2 for gr in groups)
    out = {"groups": len(groups), "frames": nfr, "gigapixels": px / 1e9, "seconds": el,
           "frames_per_s": nfr / el, "megapixels_per_s": px / el / 1e6,
           "gn_iterations_mean": float(np.mean([r["iterations"] for r in res]))}
    # the same batch with the launch variants tuned per group shape (same
    # handle: templates and geometries are resident per handle)
    h.tuning_mode(-1)
    ssa.align_groups(calls, handle=h, linesearch_candidates=args.candidates, with_mask=False)
    t = time.perf_counter()
    ssa.align_groups(calls, handle=h, linesearch_candidates=args.candidates, with_mask=False)
    out["frames_per_s_tuned_variants"] = nfr / (time.perf_counter() - t)
    out["tuned"] = tuning_summary(h.tuning())
    h.tuning_mode(-2)  # the default: variants by batch size
    if not args.no_cpu:
        from oracle import oracle as O
        gr = groups[0]
        geom = geoms[(gr["size"], gr["lines_per_knot"])][0]
        a = W.group_args(gr["T"], gr["frames"], geom, flags="D")
        ref = O.align_group_to_template(a, threads=args.cpu_threads or min(16, os.cpu_count() or 1), with_mask=False)
        n, nk = gr["frames"].shape[2], geom["nknots"]
        dx = (res[0]["knots_x"] - ref["knots_x"]).reshape(n, nk)[:-1]
        dy = (res[0]["knots_y"] - ref["knots_y"]).reshape(n, nk)[:-1]
        out["group0_knot_rms_vs_fp64_px"] = float(np.sqrt(np.mean(np.concatenate([dx.ravel(), dy.ravel()]) ** 2)))
    return out


def ibs_leg(args, torch, ssa, frames, geom, results, G, h):
    """invert_bilinear_sampling (steady_scan.m:290-404) of the workload's
    frames with the knots the alignment step estimated; frames and outputs
    resident in HBM.  The first frames are also reconstructed by the oracle
    (bit-exact check, and the CPU baseline of this leg)."""
    nk = geom["nknots"]
    n = min(frames.shape[2], G * len(results))
    kx = np.concatenate([r["knots_x"].reshape(G, nk).T for r in results], axis=1)[:, :n]
    ky = np.concatenate([r["knots_y"].reshape(G, nk).T for r in results], axis=1)[:, :n]
    th, tw = geom["desc"].template_h, geom["desc"].template_w
    d_in = torch.from_numpy(np.asfortranarray(frames[:, :, :n]).ravel(order="F").view(np.int16)).to("cuda")
    d_out = torch.empty(th * tw * n, dtype=torch.float32, device="cuda")
    d_mask = torch.empty(th * tw * n, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    info = {}

    def call():
        ssa.invert_bilinear_sampling((d_in.data_ptr(), np.uint16, n), geom, kx, ky,
                                     device_out=(d_out.data_ptr(), d_mask.data_ptr()), handle=h, info=info)

    call()
    torch.cuda.synchronize()
    t = time.perf_counter()
    call()
    torch.cuda.synchronize()
    el = time.perf_counter() - t
    it = info["iterations"]
    out = {"frames": n, "seconds": el, "frames_per_s": n / el,
           "gs_iterations_mean": float(it.mean()), "gs_iterations_max": int(it.max()),
           "mask_fraction": float(d_mask.float().mean().item()),
           "inputs": "uint16 frames and float32/uint8 outputs resident in HBM; knots of the alignment step"}
    print(f"[bench] reconstruction: {n} frames in {el:.3f}s ({n / el:.0f} frames/s), "
          f"Gauss-Seidel iterations mean {it.mean():.1f} max {it.max()}", file=sys.stderr, flush=True)
    # inpainting of the corrected frames (imgapfill_leastsquares, in place, single precision)
    from steadyscan_amd import _native as N
    missing = float((1 - d_mask.float().mean()).item())
    t = time.perf_counter()
    rc = N.lib().ss_fill_gaps(h.ptr, N.SS_F32, ctypes.c_void_p(d_out.data_ptr()), ctypes.c_void_p(d_out.data_ptr()),
                              th, tw, n, 1)
    torch.cuda.synchronize()
    el_fill = time.perf_counter() - t
    out["gap_fill"] = {"frames_per_s": n / el_fill, "seconds": el_fill, "missing_fraction": missing, "status": rc}
    # prefilter_foralignment of the raw frames (Gaussian sigma 1, class kept), in HBM
    d_pre = torch.empty_like(d_in)
    for rep in range(2):
        t = time.perf_counter()
        rc = N.lib().ss_prefilter(h.ptr, N.SS_U16, ctypes.c_void_p(d_in.data_ptr()), ctypes.c_void_p(d_pre.data_ptr()),
                                  frames.shape[0], frames.shape[1], n, 1.0, 0, 1)
        torch.cuda.synchronize()
        el_pre = time.perf_counter() - t
    out["prefilter"] = {"frames_per_s": n / el_pre, "seconds": el_pre, "pregauss": 1.0, "status": rc}
    print(f"[bench] gap fill: {n / el_fill:.0f} frames/s ({missing * 100:.1f}% missing); "
          f"prefilter (sigma 1): {n / el_pre:.0f} frames/s", file=sys.stderr, flush=True)
    # the driver's reconstruction loop (ss_reconstruct): host frames in, double
    # data_corrected out, sets of 500 frames restarting the check interval
    fr_host = np.asfortranarray(frames[:, :, :n])
    ssa.reconstruct(fr_host[:, :, :8], geom, kx[:, :8], ky[:, :8], frames_per_set=500, handle=h)
    t = time.perf_counter()
    ssa.reconstruct(fr_host, geom, kx, ky, frames_per_set=500, handle=h)
    el_drv = time.perf_counter() - t
    out["driver_loop"] = {"frames_per_s": n / el_drv, "seconds": el_drv, "frames_per_set": 500,
                          "inputs": "uint16 frames in host memory, float64 data_corrected and mask returned to host"}
    print(f"[bench] reconstruction loop (ss_reconstruct, host I/O, sets of 500): {n / el_drv:.0f} frames/s",
          file=sys.stderr, flush=True)
    if not args.no_cpu:
        from oracle import oracle as O
        m = 3
        gi = {}
        gout, gmask = ssa.invert_bilinear_sampling(frames[:, :, :m], geom, kx[:, :m], ky[:, :m], handle=h, info=gi)
        t = time.perf_counter()
        ref, rmask, rit = O.invert_bilinear_sampling(frames[:, :, :m].astype(np.float64), kx[:, :m], ky[:, :m],
                                                     geom, (th, tw))
        tcpu = time.perf_counter() - t
        same = bool(np.array_equal(gout, ref, equal_nan=True) and np.array_equal(gmask, rmask)
                    and np.array_equal(gi["iterations"], rit))
        out["oracle_bit_exact_first_frames"] = same
        out["cpu_baseline"] = {"value": m / tcpu, "unit": "frames/s", "cores": 1, "kind": "port",
                               "sample": f"oracle (C, one thread) on the first {m} frames, {tcpu:.1f} s"}
        from oracle import prefilter as P
        t = time.perf_counter()
        P.prefilter_foralignment(frames[:, :, :m], 1.0, False)
        out["prefilter"]["cpu_baseline"] = {"value": m / (time.perf_counter() - t), "unit": "frames/s", "cores": 1,
                                            "kind": "port", "sample": f"numpy FFT oracle, first {m} frames"}
        t = time.perf_counter()
        P.imgapfill_leastsquares(gout)
        out["gap_fill"]["cpu_baseline"] = {"value": m / (time.perf_counter() - t), "unit": "frames/s", "cores": 1,
                                           "kind": "port", "sample": f"scipy sparse direct oracle, first {m} frames"}
        print(f"[bench] reconstruction oracle: first {m} frames bit-exact {same}, {m / tcpu:.2f} frames/s on 1 core",
              file=sys.stderr, flush=True)
    return out


def host_cpu():
    """Host processor of this run: logical CPUs and the model name."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              "(groups are independent: on several GPUs each would align its own groups)", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    import torch

    import steadyscan_amd as ssa
    from steadyscan_amd import _native as N
    from steadyscan_amd import workloads as W

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    t0 = time.time()
    T, frames, geom, _ = W.config1(n_frames=args.frames)
    nk = geom["nknots"]
    nb = nk - 1
    G = args.group
    ngroups = args.frames // G
    fh = geom["basex_bm"].shape[0]
    print(f"[bench] generated {args.frames} frames in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    h = ssa.Handle(0)
    xg, yg = W.gradients(T)
    tid = h.upload_template(T, xg, yg)
    gid = h.upload_geometry(geom["tfrac_bm"], geom["basex_bm"], geom["basey_bm"], geom["base_mask_bm"])
    # frames resident in HBM, packed as block matrices (uint16 viewed as int16 bytes)
    dev = []
    for g in range(ngroups):
        bm = ssa.image2blockmatrix(frames[:, :, g * G:(g + 1) * G], geom)
        dev.append(torch.from_numpy(bm.view(np.int16).ravel(order="F")).to("cuda"))
    mask_bytes = fh * nb * G

    def new_masks():
        return [torch.empty(mask_bytes, dtype=torch.uint8, device="cuda") for _ in range(ngroups)]

    masks = new_masks()  # the step's masks, written in place by the kernels
    torch.cuda.synchronize()
    zeros = np.zeros((nk, G))

    def call(g, prec, bufs=None):
        d = (bufs or dev)[g]
        a = dict(flags=prec[0], template_bf16=(prec == "SB"), template_image=tid, xgrad=xg, ygrad=yg, tfrac=gid,
                 basex=geom["basex_bm"], basey=geom["basey_bm"], base_mask=geom["base_mask_bm"],
                 bmData=(d.data_ptr(), np.uint16), est_dx_knots=zeros, est_dy_knots=zeros,
                 group_size=2 * nb, nsetframes=G,
                 min_points_per_block=geom["min_points_per_block"].astype(np.float64))
        a.update(W.default_args())
        return a

    mask_ptrs = [m.data_ptr() for m in masks]

    def timed(prec):
        calls = [call(g, prec) for g in range(ngroups)]

        def step():
            return ssa.align_groups(calls, handle=h, linesearch_candidates=args.candidates, with_mask=mask_ptrs)

        res = None
        for _ in range(args.warmup):
            res = step()
        torch.cuda.synchronize()
        lat = []
        t_start = time.perf_counter()
        for _ in range(args.steps):
            t = time.perf_counter()
            res = step()
            lat.append(time.perf_counter() - t)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t_start
        print(f"[bench] precision {prec}: timed {args.steps} steps in {elapsed:.3f}s", file=sys.stderr, flush=True)
        # one more step with HIP events around every stage launch (plain
        # launches instead of the tick graph) for the per-kernel figures
        N.lib().ss_profile_enable(h.ptr, 1)
        step()
        launches = (ctypes.c_int64 * 6)()
        kms = (ctypes.c_double * 6)()
        kbytes = (ctypes.c_double * 6)()
        N.lib().ss_profile_read(h.ptr, launches, kms, kbytes)
        N.lib().ss_profile_enable(h.ptr, 0)
        per_kernel = {}
        for i in range(6):
            if launches[i]:
                per_kernel[N.KERNEL_CLASSES[i]] = {
                    "launches": int(launches[i]), "ms": kms[i], "avg_us": kms[i] / launches[i] * 1e3,
                    "frame_GBps": (kbytes[i] / (kms[i] / 1e3) / 1e9) if kms[i] > 0 else 0.0}
        dom = max(range(6), key=lambda i: kms[i])
        avg_s = kms[dom] / launches[dom] / 1e3 if launches[dom] else float("nan")
        bpl = kbytes[dom] / launches[dom] if launches[dom] else 0.0
        return dict(calls=calls, results=res, elapsed=elapsed, lat=lat, per_kernel=per_kernel, dom=dom,
                    avg_s=avg_s, bytes_per_launch=bpl, align_frame_bytes=float(sum(kbytes)),
                    achieved=bpl / avg_s / 1e9 if launches[dom] else 0.0,
                    fps=ngroups * G * args.steps / elapsed)

    def timed_stream(prec, max_groups):
        """The K steps as a stream through ss_server: every step has its own
        device copy of the frames and its own mask buffers; step s is
        submitted once every group of step s-1 has been admitted."""
        nsteps = max(args.steps, args.warmup)
        copies = [[d.clone() for d in dev] for _ in range(nsteps)]
        smasks = [new_masks() for _ in range(nsteps)]
        torch.cuda.synchronize()
        srv = ssa.Server(handle=h, max_groups=max_groups, linesearch_candidates=args.candidates)

        def run(k):
            tickets = []
            done = {}
            nxt, queued = 0, 0
            while len(done) < k * ngroups:
                if nxt < k and queued == 0:
                    for g in range(ngroups):
                        tickets.append(srv.submit(call(g, prec, copies[nxt]), mask=smasks[nxt][g].data_ptr()))
                    nxt += 1
                _, queued = srv.step()
                done.update(srv.completed())
            bad = [r for r in done.values() if isinstance(r, Exception)]
            if bad:
                raise bad[0]
            return [done[t] for t in tickets]

        if args.warmup:
            run(args.warmup)
        torch.cuda.synchronize()
        t = time.perf_counter()
        res = run(args.steps)
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        srv.close()
        same = all(np.array_equal(a["knots_x"], b["knots_x"]) and np.array_equal(a["knots_y"], b["knots_y"])
                   for a, b in zip(res[:ngroups], runs_ref.get(prec, res[:ngroups])))
        # the batched step's mask of group 0 in this precision
        ssa.align_groups([call(0, prec)], handle=h, linesearch_candidates=args.candidates,
                         with_mask=[masks[0].data_ptr()])
        same_mask = bool(torch.equal(smasks[0][0], masks[0]))
        del copies, smasks
        print(f"[bench] stream {prec} (<= {max_groups} resident groups): {args.steps} steps in {el:.3f}s "
              f"({ngroups * G * args.steps / el:.0f} frames/s), knots equal to the batched step: {same}",
              file=sys.stderr, flush=True)
        return {"frames_per_s": ngroups * G * args.steps / el, "ms_per_step": el / args.steps * 1e3,
                "max_resident_groups": max_groups, "knots_equal_to_batched": same,
                "mask_equal_to_batched": same_mask}

    runs_ref = {}
    precs = ["D", "S", "SB"] if args.precision == "all" else [args.precision]
    serving = None
    uneven = None
    if args.ibs_only:
        calls = [call(g, "D") for g in range(ngroups)]
        res = ssa.align_groups(calls, handle=h, linesearch_candidates=args.candidates, with_mask=False)
        print(json.dumps({"reconstruction": ibs_leg(args, torch, ssa, frames, geom, res, G, h)}))
        return
    runs = {p: timed(p) for p in precs}
    head = runs[precs[0]]
    results = head["results"]
    stream = None
    if args.stream_groups:
        runs_ref[precs[0]] = results
        stream = timed_stream(precs[0], 2 * ngroups if args.stream_groups < 0 else args.stream_groups)

    # the single-group path: one group per submission (latency, and its throughput)
    single = []
    if not args.no_latency:
        for g in range(ngroups):
            t = time.perf_counter()
            ssa.align_groups([head["calls"][g]], handle=h, linesearch_candidates=args.candidates,
                             with_mask=[mask_ptrs[g]])
            single.append(time.perf_counter() - t)
    single_ms = np.array(single) * 1e3
    single_fps = ngroups * G / float(np.sum(single)) if single else None

    # the mex boundary's case: block matrices in host memory, uploaded by the
    # library inside the timed step (PCIe-inclusive; never the headline value)
    host_in = None
    if not args.no_host_input:
        hcalls = []
        for g in range(ngroups):
            a = dict(head["calls"][g])
            a["bmData"] = ssa.image2blockmatrix(frames[:, :, g * G:(g + 1) * G], geom)
            hcalls.append(a)
        # warm-up step: the library's pinned staging and arena grow to the host-input size
        ssa.align_groups(hcalls, handle=h, linesearch_candidates=args.candidates, with_mask=mask_ptrs)
        torch.cuda.synchronize()
        t = time.perf_counter()
        hres = ssa.align_groups(hcalls, handle=h, linesearch_candidates=args.candidates, with_mask=mask_ptrs)
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        same = all(np.array_equal(a["knots_x"], b["knots_x"]) and np.array_equal(a["knots_y"], b["knots_y"])
                   for a, b in zip(hres, results))
        host_in = {"frames_per_s": ngroups * G / el, "ms_per_step": el * 1e3, "knots_equal_to_resident": same,
                   "mode": "one batched submission with the frames' block matrices in host (pageable) memory"}
        del hcalls, hres
        print(f"[bench] precision {precs[0]}, frames in host memory: one step in {el:.3f}s "
              f"({ngroups * G / el:.0f} frames/s), knots equal to the resident step: {same}",
              file=sys.stderr, flush=True)

    # end to end: a batched step, then the resampling (resample_frame, the
    # corrected frames) of all its frames, device to device; algorithmic
    # bytes of the measurement plan: the step's frame passes + the
    # resampling's frame read and float32 write
    e2e = None
    if not args.no_end_to_end:
        th, tw = geom["desc"].template_h, geom["desc"].template_w
        n = ngroups * G
        d_raw = torch.from_numpy(np.asfortranarray(frames[:, :, :n]).ravel(order="F").view(np.int16)).to("cuda")
        d_cor = torch.empty(th * tw * n, dtype=torch.float32, device="cuda")

        def e2e_step():
            res = ssa.align_groups(head["calls"], handle=h, linesearch_candidates=args.candidates,
                                   with_mask=mask_ptrs)
            kx = np.concatenate([r["knots_x"].reshape(G, nk).T for r in res], axis=1)
            ky = np.concatenate([r["knots_y"].reshape(G, nk).T for r in res], axis=1)
            ssa.resample_frames((d_raw.data_ptr(), np.uint16, n), geom, kx, ky, device_out=d_cor.data_ptr(),
                                handle=h)

        e2e_step()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            e2e_step()
        torch.cuda.synchronize()
        el = (time.perf_counter() - t) / args.steps
        s_pix = 2.0
        b_frame = head["align_frame_bytes"] / n + s_pix * frames.shape[0] * frames.shape[1] + 4.0 * th * tw
        e2e = {"frames_per_s": n / el, "ms_per_step": el * 1e3, "bytes_per_frame": b_frame,
               "achieved_GBps": b_frame * n / el / 1e9, "frac": b_frame * n / el / 1e9 / HBM_PEAK_GBS,
               "mode": "one batched step (masks in HBM) + ss_resample of its frames, device to device"}
        del d_raw, d_cor
        print(f"[bench] end to end (align + resample): {n / el:.0f} frames/s, {b_frame / 1e6:.1f} MB/frame",
              file=sys.stderr, flush=True)

    if not args.no_uneven:
        uneven = uneven_leg(args, torch, ssa, W)
        print(f"[bench] uneven: {uneven['frames_per_s']:.0f} frames/s, {uneven['megapixels_per_s']:.0f} Mpx/s",
              file=sys.stderr, flush=True)
    recon = None
    if not args.no_ibs:
        recon = ibs_leg(args, torch, ssa, frames, geom, results, G, h)
    if not args.no_serve:
        serving = serve_leg(args, torch, ssa, W)
        print(f"[bench] serving: p50 {serving['p50_group_ms']:.1f} ms p99 {serving['p99_group_ms']:.1f} ms "
              f"{serving['frames_per_s']:.0f} frames/s", file=sys.stderr, flush=True)

    # parity of group 0 against the FP64 oracle (same precision letter, which
    # only sets the mask-bound epsilon there), and the CPU baselines
    parity = {}
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        g = 0
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        fr = frames[:, :, g * G:g * G + G - 1]
        m = nk * (G - 1)
        refs = {}
        for prec in precs:
            a = W.group_args(T, frames, geom, first=g * G, count=G, flags=prec[0])
            t = time.perf_counter()
            if prec[0] not in refs:  # the oracle is FP64; the letter only sets the mask epsilon
                refs[prec[0]] = O.align_group_to_template(a, threads=threads, with_mask=False)
            ref = refs[prec[0]]
            tcpu = time.perf_counter() - t
            out = runs[prec]["results"][g]
            dx = (out["knots_x"] - ref["knots_x"]).reshape(G, nk)[:-1]
            dy = (out["knots_y"] - ref["knots_y"]).reshape(G, nk)[:-1]
            rms = float(np.sqrt(np.mean(np.concatenate([dx.ravel(), dy.ravel()]) ** 2)))
            # corrected frames from the GPU knots vs the oracle knots: the
            # oracle's own resampling (orc_resample_frame) of the oracle
            # knots against the GPU resampling of the GPU knots
            cg = ssa.resample_frames(fr, geom, out["knots_x"][:m], out["knots_y"][:m], handle=h)
            co = O.resample_frames(fr, geom, ref["knots_x"][:m], ref["knots_y"][:m])
            both = np.isfinite(cg) & np.isfinite(co)
            diff = np.abs(cg[both] - co[both])
            parity[prec] = dict(knot_rms=rms, corr_max=float(diff.max()) if diff.size else 0.0,
                                corr_rms=float(np.sqrt(np.mean(diff ** 2))) if diff.size else 0.0,
                                iters_gpu=int(out["iterations"]), iters_oracle=int(ref["iterations"]))
            print(f"[bench] oracle group 0 [{prec}]: iterations {ref['iterations']} gpu {out['iterations']} "
                  f"rms {rms:.2e} px ({tcpu:.1f} s)", file=sys.stderr, flush=True)
            if cpu is None:
                cpu = {"value": G / tcpu, "unit": "frames/s", "cores": threads, "kind": "port",
                       "sample": f"FP64 oracle (C restatement, OpenMP) on 1 group of {G} frames 512x512 "
                                 f"(group 0 of the workload), {ref['iterations']} GN iterations, {tcpu:.1f} s",
                       "host": host_cpu()}
        # the same oracle on one thread, on the first 40 frames of group 0
        n1 = 40
        a1 = W.group_args(T, frames, geom, first=0, count=n1, flags=precs[0][0])
        t = time.perf_counter()
        r1 = O.align_group_to_template(a1, threads=1, with_mask=False)
        t1 = time.perf_counter() - t
        cpu["single_thread"] = {"value": n1 / t1, "unit": "frames/s", "cores": 1,
                                "sample": f"FP64 oracle on one thread, the first {n1} frames of group 0 as "
                                          f"one group, {r1['iterations']} GN iterations, {t1:.1f} s"}
        print(f"[bench] oracle, one thread: {n1} frames in {t1:.1f} s", file=sys.stderr, flush=True)
        # FP32 pixel arithmetic: the oracle's source built with float pixels
        # (template, frames, warp, residuals, per-column sums), OpenMP, group 0
        af = W.group_args(T, frames, geom, first=0, count=G, flags="S")
        t = time.perf_counter()
        rf = O.align_group_to_template(af, threads=threads, with_mask=False, L=O.lib_f32())
        tf = time.perf_counter() - t
        cpu["fp32"] = {"value": G / tf, "unit": "frames/s", "cores": threads,
                       "sample": f"the oracle's source with FP32 pixels, OpenMP, on group 0 ({G} frames), "
                                 f"{rf['iterations']} GN iterations, {tf:.1f} s"}
        print(f"[bench] FP32-pixel oracle copy, {threads} threads: {G} frames in {tf:.1f} s", file=sys.stderr,
              flush=True)

    iters = [r["iterations"] for r in results]
    lat_ms = np.array(head["lat"]) * 1e3
    hp = precs[0]
    dom = head["dom"]
    other = {}
    for prec in precs[1:]:
        r = runs[prec]
        other[prec] = {"frames_per_s": r["fps"], "ms_per_step": r["elapsed"] / args.steps * 1e3,
                       "gn_iterations_mean": float(np.mean([x["iterations"] for x in r["results"]])),
                       "kernels_one_step": r["per_kernel"], **{k: v for k, v in parity.get(prec, {}).items()}}
    hpar = parity.get(hp, {})
    value = stream["frames_per_s"] if stream else head["fps"]
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": stream["ms_per_step"] if stream else head["elapsed"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if hp == "D" else "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.frames} frames 512x512 uint16 in groups of {G}, knots every 16 lines, "
                        "default settings (max_iter 120, smearing filter, exact solve), every output of the "
                        "call produced (masks in HBM); "
                        + ("the steps streamed through the continuous-batching server, each step on its own "
                           "copy of the frames, submitted once the previous step is admitted "
                           f"(at most {stream['max_resident_groups']} resident groups)" if stream else
                           "all groups of a step in one batched submission"),
            "barrier_per_step": {"frames_per_s": head["fps"], "ms_per_step": head["elapsed"] / args.steps * 1e3,
                                 "mode": "all groups of a step in one batched submission, synchronised per step"},
            "host_input_per_step": host_in,
            "end_to_end": e2e,
            "precision": hp,
            "frames_per_step": ngroups * G,
            "groups_per_step": ngroups,
            "p50_group_ms": float(np.percentile(single_ms, 50)) if single else None,
            "p99_group_ms": float(np.percentile(single_ms, 99)) if single else None,
            "group_latency_mode": "single-group submissions, inputs resident",
            "single_group_frames_per_s": single_fps,
            "ratio_vs_single_group": (value / single_fps) if single_fps else None,
            "batch_ms_p50": float(np.percentile(lat_ms, 50)),
            "gn_iterations_mean": float(np.mean(iters)),
            "gn_iterations_max": int(np.max(iters)),
            "linesearch_candidates_first_round": args.candidates,
            "knot_rms_vs_fp64_px": hpar.get("knot_rms"),
            "corrected_frame_max_abs_err": hpar.get("corr_max"),
            "corrected_frame_rms_err": hpar.get("corr_rms"),
            "kernels_one_step": head["per_kernel"],
            "storage": "u16 frames; " + ("f64 warp/residual/accumulation" if hp == "D" else
                                         "f32 warp/residual/accumulation") + "; f64 reductions and solve",
            "other_precisions": other,
            "serving": serving,
            "uneven_groups": uneven,
            "reconstruction": recon,
            "stream": stream,
        },
        "roofline": {
            "kernel": N.KERNEL_CLASSES[dom],
            "bound": "hbm",
            "achieved": head["achieved"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": head["achieved"] / HBM_PEAK_GBS,
            "traffic": measured_traffic(N.KERNEL_CLASSES[dom], hp),
            "bytes_per_launch": head["bytes_per_launch"],
            "avg_launch_us": head["avg_s"] * 1e6,
            "limiter": measured_limiter(N.KERNEL_CLASSES[dom], hp),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
