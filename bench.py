#!/usr/bin/env python3
"""Potential-solve benchmark on one MI355X.

Workload: the 256x128x256 spiral-arm torus (JMAX=256, KMAX=128, LMAX=256,
DELR=DELZ=1/256, seeded), plain FP64.  One solve is BDYGEN + POT3 on one
density field, device resident, COEF for the K=2 plane.  A step is one batch
of B independent fields (different seeds, each on its own prepared grid)
solved by one chy_solve_batch call, which runs them as batched launches of
up to 32 fields (default B=32: one launch per kernel); `value` is solves per
second over the timed steps (B solves per step) with every density already in
HBM.  B=1 is the single-field solve.

The line also carries:
  * p50/p99 latency of one solve (a separate pass of single-field chy_solve
    calls timed with HIP events on the solve's stream; --latency-steps 0 skips
    it) and of one batch step;
  * the error of field 0 against the fixture solutions (reference FP64 and the
    80-bit solve of the same discrete system);
  * the roofline of the dominant kernel of the timed region: HIP events
    around every stage on the worker stream the batched launches run on give
    its average launch duration (`batch_stages_ms_per_launch`), one launch
    processes `fields_per_launch` fields; `roofline.traffic` is the HBM bytes
    per launch that rocprofv3 FETCH_SIZE/WRITE_SIZE recorded for that kernel
    with launches of that many fields (profiles/potsolve_256x128x256_batch<B>_
    hbm_traffic.txt, or ..._hbm_traffic.txt for single fields), or null;
  * `stages_ms`: per-stage times of single-field solves launched kernel by
    kernel (--stage-steps 0 skips them);
  * the reference CPU solve timed on this host;
  * `source_grav`: the hydro step's momentum sources that read PHI
    (radhydro/source.f:41-79) on the same grid, per call, with the oracle's
    C loop timed beside it, and the star potential added to PHI before them
    (pot/ExternalPot.f, `source_grav.external_pot`) and the disk's pull on
    the star (find_force of pot/wiggle.F, `source_grav.find_force`)
    (--source-calls 0 skips them).

    python bench.py [--steps K] [--warmup W] [--batch B] [--workers W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "chymera-uo_amd"))

METRIC = "pot3 solves/sec @256x128x256 refined vs plain FP64 at 1e-10; p50/p99 latency"
HBM_PEAK_GBS = 8000.0        # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 (matrix and vector)


def stage_work(g):
    """Algorithmic bytes / flops of one launch of each stage at grid g."""
    J, K, L = g.jmax, g.kmax, g.lmax
    M, NZ = J - 1, K - 1
    lanes = J * K * L * 8              # every (J,K) lane x L samples, FP64
    modes = M * NZ * L * 8             # interior mode unknowns
    gemm = 2.0 * NZ * NZ * M * L       # one z transform
    return {
        "bdygen": dict(bytes=modes, flops=None),
        "fft_fwd": dict(bytes=2 * lanes, flops=None),
        "zfwd": dict(bytes=lanes + modes, flops=gemm),
        "rsolve": dict(bytes=2 * modes, flops=None),
        "zinv": dict(bytes=2 * modes, flops=gemm),
        "fft_inv": dict(bytes=2 * lanes, flops=None),
        "ghost": dict(bytes=None, flops=None),
        "fft_bdy": dict(bytes=None, flops=None),
    }


# stage -> kernel whose rocprofv3 counters stand for it
STAGE_KERNEL = {"zfwd": "k_zgemm_reg<true, 128, 64, 1, double>", "zinv": "k_zgemm_reg<false, 128, 64, 1, double>",
                "fft_fwd": "k_fft_forward<16, 8, 32, double, double>",
                "fft_inv": "k_fft_inverse<16, 8, 32, double, false>", "rsolve": "k_rsolve<4, double>"}


def hbm_traffic(stage, batch):
    """HBM bytes per launch of the stage's kernel as rocprofv3 FETCH_SIZE +
    WRITE_SIZE recorded them, each scaled by what the counter read for a
    device-to-device copy of known size: profiles/..._hbm_traffic.txt for
    single-field launches, profiles/..._batch<B>_hbm_traffic.txt for launches
    of B fields."""
    name_b = "" if batch == 1 else f"_batch{batch}"
    f = ROOT / "profiles" / f"potsolve_256x128x256{name_b}_hbm_traffic.txt"
    name = STAGE_KERNEL.get(stage)
    if name is None or not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("kernel="):
            kv = dict(p.split("=", 1) for p in line.split("\t"))
            if kv["kernel"] == name:
                return float(kv["traffic_bytes"])
    return None


def hetero_batch(ctx, n, reps=3):
    """One chy_solve_batch over n independent fields of mixed grids (the grid
    and torus of fields.random_field for seeds 1000.., generated on the GPU),
    one prepared grid per shape; fields per second over `reps` batches."""
    import torch

    import chypot
    from chypot import fields
    prepared, jobs, points = {}, [], 0
    for i in range(n):
        g, rho_d = fields.random_field_device(1000 + i)
        key = (g.jmax, g.kmax, g.lmax)
        if key not in prepared:
            prepared[key] = chypot.PreparedGrid(ctx, g)
        jobs.append(dict(grid=prepared[key], rho=rho_d, phi=torch.zeros_like(rho_d),
                         coef=torch.zeros((2, g.lmax // 2, g.kmax + 2, g.jmax + 2), dtype=torch.float64,
                                          device="cuda"),
                         coef_mode=chypot.COEF_K2))
        points += g.jmax * g.kmax * g.lmax
    ctx.solve_batch(jobs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.solve_batch(jobs)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    out = {"fields": n, "distinct_grids": len(prepared), "batch_ms": dt * 1e3,
           "fields_per_s": n / dt, "grid_points_per_s": points / dt}
    for pg in prepared.values():
        pg.close()
    return out


def serve_run(ctx, args):
    """The serving loop (chypot.serve) over Poisson arrivals: throughput and
    per-request latency from arrival to completion."""
    import torch

    from chypot import serve
    mix = [((128, 64, 128), 0.6), ((256, 128, 256), 0.3), ((512, 256, 512), 0.1)]
    dens = {shape: serve.density_pool(shape, 2, seed=17, on_device=True) for shape, _ in mix}
    torch.cuda.synchronize()
    reqs = serve.poisson_requests(args.serve_rate, args.serve, mix, dens, seed=5)
    srv = serve.Server(ctx, hbm_budget=int(args.serve_budget_gib * (1 << 30)), max_batch=16, workers=8)
    srv.prepare([shape for shape, _ in mix])
    t0 = time.perf_counter()
    done = srv.run(reqs)
    wall = time.perf_counter() - t0
    srv.close()
    out = serve.latency_summary(done)
    out.update({"offered_rate_per_s": args.serve_rate, "seconds": args.serve,
                "completed_per_s": len(done) / wall,
                "mix": "60% 128x64x128, 30% 256x128x256, 10% 512x256x512"})
    return out


def source_row(ctx, stream, g, n, with_cpu):
    """The momentum sources of radhydro/source.f:41-79 on the solve's grid
    (hydro grid = potential grid, JSTART=2): per-call time from HIP events on
    the solve stream around n calls; algorithmic bytes are 80 per cell of the
    T/A range (RHO, OMEGA, PHI, P read; S, T, A read and written); with_cpu
    also times the oracle's C loop on this host (1 thread) on the same data."""
    import torch

    import chypot
    rng = np.random.default_rng(7)
    shape = (g.lmax, g.kmax + 2, g.jmax + 2)
    host = [rng.standard_normal(shape) for _ in range(7)]
    rho, om, phi, p, s, t, a = (torch.from_numpy(x).to("cuda") for x in host)
    r = torch.from_numpy(g.r()).to("cuda")
    d = chypot.SourceDesc(g.jmax, g.kmax, g.lmax, g.jmax, g.kmax, 2, 1.0e-3, g.delr, g.delz, g.dtheta)
    for _ in range(3):
        ctx.source_grav(d, r, rho, om, phi, p, s, t, a)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(n):
        ctx.source_grav(d, r, rho, om, phi, p, s, t, a)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / n
    cells = (g.jmax + 1 - d.jstart + 1) * g.kmax * g.lmax
    nbytes = 80 * cells
    out = {"kernel": "k_source_grav", "grid": g.name, "calls": n, "us_per_call": ms * 1e3,
           "algorithmic_bytes": nbytes, "achieved_GBs": nbytes / (ms * 1e-3) / 1e9,
           "frac_hbm": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    # the star potential of ExternalPot on the same PHI: 24 bytes per cell
    # (PHI read and written, STARPHI written)
    e = chypot.ExtPotDesc(g.jmax, g.kmax, g.lmax, g.jmax, g.kmax, g.dtheta, 1.0, 0.0, 0.02, 0.3)
    rhf, zhf = torch.from_numpy(g.rhf()).to("cuda"), torch.from_numpy(g.zhf()).to("cuda")
    ctx.external_pot(e, rhf, zhf, phi, s)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(n):
        ctx.external_pot(e, rhf, zhf, phi, s)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms_e = ev0.elapsed_time(ev1) / n
    eb = 24 * (g.jmax + 2) * (g.kmax + 2) * g.lmax
    out["external_pot"] = {"kernel": "k_external_pot", "us_per_call": ms_e * 1e3, "algorithmic_bytes": eb,
                           "achieved_GBs": eb / (ms_e * 1e-3) / 1e9,
                           "frac_hbm": eb / (ms_e * 1e-3) / 1e9 / HBM_PEAK_GBS}
    # the disk's pull on the star (find_force of wiggle, JMIN=4): RHO read over
    # J=1..JMAX+1, K=2..KMAX+1; each call reads (fx, fy) back and waits, so the
    # time includes that round trip
    fd = chypot.ForceDesc(g.jmax, g.kmax, g.lmax, 4, g.delr, g.delz, g.dtheta, 0.01)
    ctx.find_force(fd, rhf, zhf, rho, 0.01, 0.0)
    t0 = time.perf_counter()
    for _ in range(n):
        ctx.find_force(fd, rhf, zhf, rho, 0.01, 0.0)
    ms_f = (time.perf_counter() - t0) / n * 1e3
    fb = 8 * (g.jmax + 1) * g.kmax * g.lmax
    out["find_force"] = {"kernel": "k_find_force + k_force_total", "us_per_call": ms_f * 1e3,
                         "algorithmic_bytes": fb, "achieved_GBs": fb / (ms_f * 1e-3) / 1e9,
                         "includes": "launch, readback of (fx, fy) and the wait"}
    del rho, om, phi, p, s, t, a
    if with_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import pyoracle
        f = [np.asfortranarray(x.transpose(2, 1, 0)) for x in host]
        secs = []
        for _ in range(3):
            t0 = time.perf_counter()
            pyoracle.source_grav(g.jmax, g.kmax, g.lmax, g.jmax, g.kmax, 2, d.delt, d.dr, d.dz, d.dtheta,
                                 g.r(), f[0], f[1], f[2], f[3], f[4], f[5], f[6])
            secs.append(time.perf_counter() - t0)
        out["cpu_oracle_ms"] = float(np.median(secs)) * 1e3
        out["cpu_cores"] = 1
    return out


def cpu_baseline(g, rho, nrep):
    """Reference BDYGEN+POT3 (flang -O3 real-8 build of pot/*.f, oracle/_ref) on
    this host, 1 thread; falls back to the oracle's C restatement."""
    exe = ROOT / "oracle" / "_ref" / f"r8_{g.name}" / "refpot"
    if exe.exists():
        with tempfile.TemporaryDirectory() as td:
            fin, fout = pathlib.Path(td) / "in.bin", pathlib.Path(td) / "out.bin"
            with open(fin, "wb") as f:
                f.write(np.array([g.jmax, g.kmax, g.lmax, 2, 10, 8, 1, nrep], dtype="<i4").tobytes())
                f.write(np.array([g.delr, g.delz, 0.0, 0.0], dtype="<f8").tobytes())
                f.write(np.asarray(rho, dtype="<f8").tobytes(order="F"))
            env = dict(os.environ, OMP_NUM_THREADS="1", OMP_STACKSIZE="512M")
            subprocess.run([str(exe), "solve", str(fin), str(fout)], check=True, env=env,
                           stdout=subprocess.DEVNULL)
            raw = fout.read_bytes()
        n = int(np.prod(g.shape))
        nc = (g.jmax + 2) * (g.kmax + 2) * (g.lmax // 2) * 2
        off = (2 * n + 1 + nc + 3) * 8
        secs = np.frombuffer(raw, dtype="<f8", count=nrep, offset=off + 4)
        p50 = float(np.median(secs))
        return {"value": 1.0 / p50, "unit": "solves/s", "cores": 1, "kind": "reference",
                "sample": f"{nrep} solves (after 1 warm-up) of the same 256x128x256 field, "
                          f"BDYGEN+POT3 of pot/*.f built with flang -O3 -fdefault-real-8, 1 thread; "
                          f"p50 {p50:.3f} s"}
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    pyoracle.solve(g, rho)
    secs = []
    for _ in range(nrep):
        t = time.perf_counter()
        pyoracle.solve(g, rho)
        secs.append(time.perf_counter() - t)
    p50 = float(np.median(secs))
    return {"value": 1.0 / p50, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"{nrep} solves of the same field by the C restatement, {threads} threads"}


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch", type=int, default=32, help="independent fields per step")
    ap.add_argument("--workers", type=int, default=0, help="worker streams (default min(batch, 8))")
    ap.add_argument("--refine-steps", type=int, default=2,
                    help="correction steps of the refined path (0: do not measure it)")
    ap.add_argument("--hetero", type=int, default=0,
                    help="also time one batch of this many heterogeneous fields (J in 64..256, "
                         "K in 32..128, L in 64..256; GPU-generated tori)")
    ap.add_argument("--serve", type=float, default=0.0,
                    help="also run the serving loop for this many seconds of Poisson arrivals "
                         "(60%% 128x64x128, 30%% 256x128x256, 10%% 512x256x512)")
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="arrivals per second")
    ap.add_argument("--serve-budget-gib", type=float, default=64.0, help="HBM budget of the server")
    ap.add_argument("--latency-steps", type=int, default=-1,
                    help="single-field chy_solve calls for p50/p99 (default: --steps; 0 skips)")
    ap.add_argument("--stage-steps", type=int, default=-1,
                    help="single-field stage-timed solves for stages_ms (default: --steps; 0 skips)")
    ap.add_argument("--source-calls", type=int, default=20,
                    help="also time this many calls of the SOURCE momentum sources (0: skip)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="before the solves, copy a buffer of this many MiB device to device "
                         "three times (a kernel of known traffic for scaling HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the potential solve of one field runs on one MI355X; "
              "splitting a field over several GPUs is not implemented", file=sys.stderr)
        sys.exit(2)

    import torch

    import chypot
    from chypot import fields

    g = fields.unit_grid(256, 128, 256)
    B = max(1, args.batch)
    W = args.workers or min(B, 8)
    rho = fields.density_spiral_torus(g, seed=2)
    ctx = chypot.Context(0)
    # one explicit (default-priority) stream for the solves and the timing events
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    ctx.set_stream(stream)
    ctx.workers(W)
    jobs = []
    for b in range(B):
        rho_b = rho if b == 0 else fields.density_spiral_torus(g, seed=2 + b)
        jobs.append(dict(grid=chypot.PreparedGrid(ctx, g), rho=chypot.to_device(rho_b),
                         phi=torch.zeros((g.lmax, g.kmax + 2, g.jmax + 2), dtype=torch.float64,
                                         device="cuda"),
                         coef=torch.zeros((2, g.lmax // 2, g.kmax + 2, g.jmax + 2), dtype=torch.float64,
                                          device="cuda"),
                         coef_mode=chypot.COEF_K2))
    j0 = jobs[0]

    def step():
        ctx.solve_batch(jobs)

    def single():
        j0["grid"].solve(j0["rho"], j0["phi"], coef=j0["coef"], coef_mode=chypot.COEF_K2)

    if args.copy_mib > 0:
        src = torch.ones(args.copy_mib << 17, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # single-field passes (latency, stage times) run only when asked for, so
    # that a profile of the batch alone sees only batched launches
    n_lat = args.steps if args.latency_steps < 0 else args.latency_steps
    n_st = args.steps if args.stage_steps < 0 else args.stage_steps
    for _ in range(args.warmup):
        step()
        if n_lat > 0 or n_st > 0:
            single()
    torch.cuda.synchronize()

    def timed(fn, n):
        starts = [torch.cuda.Event(enable_timing=True) for _ in range(n)]
        ends = [torch.cuda.Event(enable_timing=True) for _ in range(n)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(n):
            starts[i].record(stream)
            fn()
            ends[i].record(stream)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        return t1 - t0, np.array([a.elapsed_time(b) for a, b in zip(starts, ends)])

    # timed region: batch steps (batched launches on a worker stream), with
    # HIP events around every stage on the stream each launch runs on: the
    # per-launch durations of the roofline come from the timed region itself
    ctx.profile(True)
    elapsed, lat_batch = timed(step, args.steps)
    tstages = ctx.stage_times()
    ctx.profile(False)
    value = B * args.steps / elapsed
    phi = chypot.to_host(j0["phi"])
    # single-solve latency: one field per call on the context's stream
    lat = timed(single, n_lat)[1] if n_lat > 0 else None

    # the refined path (binary32 correction solves on a binary64 residual):
    # single-field solves of field 0, timed the same way, and its error
    # against the plain binary64 result of the timed region (skipped with
    # --refine-steps 0, e.g. to profile the plain path alone)
    ref_phi = torch.zeros_like(j0["phi"])
    ref_coef = torch.zeros_like(j0["coef"])

    def refined():
        j0["grid"].solve_refined(j0["rho"], ref_phi, steps=args.refine_steps, coef=ref_coef,
                                 coef_mode=chypot.COEF_K2)

    if args.refine_steps > 0:
        for _ in range(args.warmup):
            refined()
        el_ref, lat_ref = timed(refined, args.steps)
        _, rel_corr = j0["grid"].solve_refined(j0["rho"], ref_phi, steps=args.refine_steps, coef=ref_coef,
                                               coef_mode=chypot.COEF_K2, report=True)
        phi_ref = chypot.to_host(ref_phi)

    # single-field steps launched kernel by kernel with HIP events around every
    # stage on the solve's stream: the single-field stage times
    stages = None
    if n_st > 0:
        ctx.profile(True)
        for _ in range(n_st):
            single()
        stages = ctx.stage_times()
        ctx.profile(False)

    # accuracy of the timed result against the fixture solutions
    fx = np.load(ROOT / "tests" / "golden" / "solve_large_256x128x256.npz")
    scale = float(fx["phi_absmax"])

    def err(a, b):
        return float(np.abs(a - b).max() / scale)

    e_ref = max(err(phi[:, 1, :], fx["phi_k2"]), err(phi[:, :, 0], fx["phi_l1"]))
    refined_line = None if args.refine_steps <= 0 else {
        "solves_per_s": args.steps / el_ref,
        "p50_ms": float(np.percentile(lat_ref, 50)),
        "p99_ms": float(np.percentile(lat_ref, 99)),
        "steps": args.refine_steps,
        "rel_correction_per_step": rel_corr,
        "max_rel_err_vs_plain_fp64": float(np.abs(phi_ref - phi).max() / np.abs(phi).max()),
        "max_rel_err_vs_reference_fp64": max(err(phi_ref[:, 1, :], fx["phi_k2"]),
                                             err(phi_ref[:, :, 0], fx["phi_l1"])),
    }
    e_ext = max(err(phi[:, 1, :], fx["ext_phi_k2"]), err(phi[:, :, 0], fx["ext_phi_l1"]))

    # roofline of the dominant stage of the timed region: its average launch
    # duration and the fields one launch processes (B when the batch is one
    # group of batched launches)
    work = stage_work(g)
    t_avg = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in tstages.items()}
    dom = max(t_avg, key=lambda k: tstages[k][0])
    avg_ms, launches = t_avg[dom], tstages[dom][1]
    per_launch = B * args.steps / launches
    wk = work[dom]
    if wk["flops"] is not None and wk["flops"] / wk["bytes"] > FP64_PEAK_TFLOPS * 1e12 / (HBM_PEAK_GBS * 1e9):
        achieved = per_launch * wk["flops"] / (avg_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK_TFLOPS}
    else:
        achieved = per_launch * (wk["bytes"] or 0) / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS}
    roof.update({"traffic": hbm_traffic(dom, int(round(per_launch))), "kernel": dom, "avg_ms": avg_ms,
                 "launches": launches, "fields_per_launch": per_launch})
    hetero = hetero_batch(ctx, args.hetero) if args.hetero > 0 else None
    serving = serve_run(ctx, args) if args.serve > 0 else None
    source = (source_row(ctx, stream, g, args.source_calls, not args.no_cpu_baseline)
              if args.source_calls > 0 else None)
    cpu = None if args.no_cpu_baseline else cpu_baseline(g, rho, args.cpu_reps)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "plain FP64 potential solve (BDYGEN+POT3), spiral-arm torus",
                   "grid": g.name, "path": "plain_fp64", "batch": B, "workers": W,
                   "coef": "K=2 plane", "maxtrm": 10, "isym": 2, "npoint": 8},
        "p50_ms": None if lat is None else float(np.percentile(lat, 50)),
        "p99_ms": None if lat is None else float(np.percentile(lat, 99)),
        "single_solves_per_s": None if lat is None else float(1e3 / np.mean(lat)),
        "batch_step_p50_ms": float(np.percentile(lat_batch, 50)),
        "batch_step_p99_ms": float(np.percentile(lat_batch, 99)),
        "max_rel_err_vs_reference_fp64": e_ref,
        "max_rel_err_vs_exact": e_ext,
        "refined_path": refined_line,
        "hetero_batch": hetero,
        "serving": serving,
        "source_grav": source,
        "stages_ms": None if stages is None else {k: round(v[0] / v[1] if v[1] else 0.0, 5)
                                                   for k, v in stages.items()},
        "batch_stages_ms_per_launch": {k: round(v, 5) for k, v in t_avg.items() if tstages[k][1]},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
