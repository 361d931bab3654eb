"""Benchmark of the align_group_to_template hot path on one MI355X.

Workload (the headline configuration): 1000 frames of 512x512 uint16 in
groups of 50 frames, knots every 16 lines (33 knots), default alignment
settings (steady_scan.m:836-867), synthetic data (seeded generator), with
frames, template and geometry resident in HBM.  One step aligns all 20
groups.  The headline times the K steps as one stream of groups through
the continuous-batching server (ss_server_*; every group of the K steps may
be resident, within the server's HBM budget), so a step's last, slowest
groups share their ticks with other steps' groups; the knots are checked
equal to the batched path's.  The
same steps with one batched submission (ss_align_batch) and a
synchronisation per step are reported beside it (barrier_per_step), and
one such step with the block matrices in host memory, as the mex boundary
hands them over, uploaded inside the timed step (host_input_per_step, the
PCIe-inclusive rate).

Precision: 'D' (the reference's default, FP64 pixel arithmetic) is the
headline; 'S' (FP32 pixel arithmetic, FP64 reductions and solve) is timed
beside it.

Printed JSON line: frames/s (whole job); group latency p50/p99 of
single-group submissions (the plain path) and the batch latency; knot RMS
and corrected-frame error of one group against the FP64 oracle; the
roofline of the dominant kernel (HIP events on the library's stream); the
oracle timed on the host as the CPU baseline.
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
    ap.add_argument("--uneven", action="store_true",
                    help="also time the batched uneven-groups workload (256 groups, 8-96 frames, 256^2/512^2)")
    ap.add_argument("--serve-seconds", type=float, default=4.0,
                    help="simulated acquisition time of the serving leg")
    ap.add_argument("--no-ibs", action="store_true",
                    help="skip the reconstruction (invert_bilinear_sampling) leg")
    ap.add_argument("--ibs-only", action="store_true",
                    help="run only the reconstruction leg (knots from one alignment step)")
    ap.add_argument("--stream-groups", type=int, default=-1,
                    help="time the steps as one stream through the continuous-batching server with at most "
                         "this many resident groups (-1: every group of the timed steps, within the server's "
                         "HBM budget; 0: off, one barrier per step)")
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


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              "(groups are independent; see DESIGN.md for how the path shards)", file=sys.stderr)
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
    torch.cuda.synchronize()
    zeros = np.zeros((nk, G))

    def call(g, prec):
        a = dict(flags=prec[0], template_bf16=(prec == "SB"), template_image=tid, xgrad=xg, ygrad=yg, tfrac=gid,
                 basex=geom["basex_bm"], basey=geom["basey_bm"], base_mask=geom["base_mask_bm"],
                 bmData=(dev[g].data_ptr(), np.uint16), est_dx_knots=zeros, est_dy_knots=zeros,
                 group_size=2 * nb, nsetframes=G,
                 min_points_per_block=geom["min_points_per_block"].astype(np.float64))
        a.update(W.default_args())
        return a

    def timed(prec):
        calls = [call(g, prec) for g in range(ngroups)]

        def step():
            return ssa.align_groups(calls, handle=h, linesearch_candidates=args.candidates, with_mask=False)

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
                    avg_s=avg_s, bytes_per_launch=bpl,
                    achieved=bpl / avg_s / 1e9 if launches[dom] else 0.0,
                    fps=ngroups * G * args.steps / elapsed)

    def timed_stream(prec, max_groups):
        """The same K steps as one stream of groups through ss_server: a
        step's groups are submitted together and a new group is admitted as
        soon as a resident one finishes, so the next step's groups fill the
        ticks where the last groups of a step run alone."""
        calls = [call(g, prec) for g in range(ngroups)]
        srv = ssa.Server(handle=h, max_groups=max_groups, linesearch_candidates=args.candidates)

        def run(nsteps):
            tickets = [srv.submit(c) for _ in range(nsteps) for c in calls]
            done = {}
            while len(done) < len(tickets):
                srv.step()
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
        print(f"[bench] stream {prec} (<= {max_groups} resident groups): {args.steps} steps in {el:.3f}s "
              f"({ngroups * G * args.steps / el:.0f} frames/s), knots equal to the batched step: {same}",
              file=sys.stderr, flush=True)
        return {"frames_per_s": ngroups * G * args.steps / el, "ms_per_step": el / args.steps * 1e3,
                "max_resident_groups": max_groups, "knots_equal_to_batched": same}

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
        stream = timed_stream(precs[0], ngroups * max(args.steps, args.warmup) if args.stream_groups < 0
                              else args.stream_groups)

    # latency of the plain path: one group per submission
    single = []
    if not args.no_latency:
        for g in range(ngroups):
            t = time.perf_counter()
            ssa.align_groups([head["calls"][g]], handle=h, linesearch_candidates=args.candidates, with_mask=False)
            single.append(time.perf_counter() - t)
    single_ms = np.array(single) * 1e3

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
        ssa.align_groups(hcalls, handle=h, linesearch_candidates=args.candidates, with_mask=False)
        torch.cuda.synchronize()
        t = time.perf_counter()
        hres = ssa.align_groups(hcalls, handle=h, linesearch_candidates=args.candidates, with_mask=False)
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

    if args.uneven:
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
    # only sets the mask-bound epsilon there), and the CPU baseline
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
            # corrected frames from the GPU knots vs the oracle knots (same splat)
            cg = ssa.resample_frames(fr, geom, out["knots_x"][:m], out["knots_y"][:m], handle=h)
            co = ssa.resample_frames(fr, geom, ref["knots_x"][:m], ref["knots_y"][:m], handle=h)
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
                                 f"(group 0 of the workload), {ref['iterations']} GN iterations, {tcpu:.1f} s"}

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
    line = {
        "metric": METRIC,
        "value": stream["frames_per_s"] if stream else head["fps"],
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
                        "default settings (max_iter 120, smearing filter, exact solve); "
                        + ("the steps' groups streamed through the continuous-batching server "
                           f"(at most {stream['max_resident_groups']} resident)" if stream else
                           "all groups of a step in one batched submission"),
            "barrier_per_step": {"frames_per_s": head["fps"], "ms_per_step": head["elapsed"] / args.steps * 1e3,
                                 "mode": "all groups of a step in one batched submission, synchronised per step"},
            "host_input_per_step": host_in,
            "precision": hp,
            "frames_per_step": ngroups * G,
            "groups_per_step": ngroups,
            "p50_group_ms": float(np.percentile(single_ms, 50)) if single else None,
            "p99_group_ms": float(np.percentile(single_ms, 99)) if single else None,
            "group_latency_mode": "single-group submissions, inputs resident",
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
