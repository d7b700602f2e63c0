#!/usr/bin/env python3
"""Potential-solve benchmark on one MI355X.

Metric: pot3 solves/sec at 256x128x256, the refined solve against the plain
FP64 solve, at 1e-10 max relative PHI error; p50/p99 latency.

Workload: the 256x128x256 spiral-arm torus (JMAX=256, KMAX=128, LMAX=256,
DELR=DELZ=1/256, seeded), one field per prepared grid, every density already
resident in HBM, COEF for the K=2 plane.  One solve is BDYGEN + POT3 on one
field.  A step is one batched call over B fields (default 32, one launch per
kernel for all of them):

  * `value`: solves/s of the refined solve (chy_solve_batch_refined: iterative
    refinement of the mode systems in the azimuthal spectrum, binary32 mode
    solves on the FP32 matrix cores, binary64 residuals and binary64 r solves
    for the orders m <= 4, --refine-steps correction steps), timed over
    --steps steps after --warmup;
  * `plain_fp64`: the same for the plain FP64 solve (chy_solve_batch, BLKTRI's
    cyclic reduction with the reference's root tables -- the path that
    matches the reference to 1e-12), and `plain_fp64_eigen` for the opt-in z
    eigenbasis solver; `vs_plain` = value / plain_fp64 solves/s;
  * errors: the refined result against the plain FP64 one (the 1e-10
    criterion) and against the reference's fixture planes, and per
    refinement step (1..4) against the plain FP64 result;
  * p50/p99 latency of single-field solves (HIP events on the solve stream)
    for both paths;
  * `rooflines`: every stage of both paths in their timed regions (HIP events
    around each stage on the stream the batched launches run on): algorithmic
    bytes per field (each input read and each output written once: forward transform 134.2 MB,
    cyclic reduction 133.4 MB, inverse 135.8 MB at 256x128x256) over the
    stage time, against 8 TB/s and (`frac_of_copy`) against a 2 GiB device
    copy timed in the same run (`roofline.hbm_copy_GBs`, read + write, beyond
    the Infinity Cache); the z transforms also against the matrix-core
    peak, the cyclic reduction against the FP64 vector peak; where profiles/
    holds a rocprofv3 FETCH_SIZE/WRITE_SIZE summary of the workload, the HBM
    bytes counted per launch beside the algorithmic ones; `roofline` is the
    dominant stage of the headline (refined) timed region;
  * `hetero_batch` (64 fields of mixed grids, refined) and `serving` (Poisson
    arrivals over 60% 128x64x128, 30% 256x128x256, 10% 512x256x512 under an
    HBM budget) -- both on by default (--hetero 0 / --serve 0 skip them);
  * `cpu_baseline`: the reference's own BDYGEN+POT3 (pot/*.f built with flang
    -O3 -fdefault-real-8 under oracle/_ref) on this host's cores and on one
    core, 3 warm-ups and --cpu-reps timed solves each, p50/p99.

    python bench.py [--steps K] [--warmup W] [--batch B] [--gpus 1]
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
FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 (matrix; the vector FP64 peak is the same)
FP32_PEAK_TFLOPS = 157.3     # MI355X FP32 (matrix)
PEAKS = {"mfma_f64": FP64_PEAK_TFLOPS, "mfma_f32": FP32_PEAK_TFLOPS, "valu_f64": FP64_PEAK_TFLOPS}


def cr_solves_per_mode(n):
    """r solves of BLKTRI's cyclic reduction of one mode with n = 2^k - 1 z
    rows, in the partial-fraction form the GPU runs: each reduction level's
    odd rows apply a product with 2h - 1 divisor roots, each even row two
    chains of h - 1 solves; each back-substitution row its two chains and a
    product of 2h - 1 roots (pot/blktri.f:72-170)."""
    kx = int(round(np.log2(n + 1)))
    tot = 0
    for lv in range(kx - 1):
        h = 1 << lv
        tot += len(range(h, n + 1, 2 * h)) * (2 * h - 1) + len(range(2 * h, n - h + 1, 2 * h)) * 2 * (h - 1)
    for lv in range(kx - 1, -1, -1):
        h, nrow = 1 << lv, (1 << (kx - lv)) - 1
        for jj in range(1, nrow + 1, 2):
            tot += (h - 1) * ((jj > 1) + (jj < nrow)) + 2 * h - 1
    return tot


def stage_work(g, path, steps):
    """Algorithmic bytes (and flops, with the peak they are priced against)
    per field of one solve, per stage, summed over the stage's launches in
    that solve.  Transforms and the cyclic reduction, each input read and
    each output written once:
    forward 2 J K L 8 bytes, reduction J K L 8 + M NZ L 8, inverse J K L 8 +
    (J+2)(K+2) L 8.  The reduction's flops: 10 per r row per solve (8 for a
    tridiagonal elimination and back-substitution, 2 to combine), over the
    solves of every mode system."""
    J, K, L = g.jmax, g.kmax, g.lmax
    M, NZ = J - 1, K - 1
    PP = ((NZ + 63) // 64) * 64
    JP = 64 * (1 << max(0, int(np.ceil(np.log2(max(1, M / 64))))))
    lanes = J * K * L * 8
    modes = M * NZ * L * 8
    gemm = 2.0 * NZ * NZ * M * L
    none = dict(bytes=None, flops=None, peak=None)
    w = {"fft_fwd": dict(bytes=2 * lanes, flops=None, peak=None),
         "fft_inv": dict(bytes=lanes + (J + 2) * (K + 2) * L * 8, flops=None, peak=None),
         "bdygen": none, "ghost": none}
    if path == "cr":
        w["cr"] = dict(bytes=lanes + modes, flops=10.0 * M * L * cr_solves_per_mode(NZ), peak="valu_f64")
    elif path == "eigen":
        w["zfwd"] = dict(bytes=modes + PP * JP * L * 8, flops=gemm, peak="mfma_f64")
        w["rsolve"] = dict(bytes=2 * modes, flops=None, peak=None)
        w["zinv"] = dict(bytes=PP * JP * L * 8 + modes, flops=gemm, peak="mfma_f64")
    elif path == "refined_fused":  # one kernel per mode plane (1 or 2 steps)
        w["planes"] = (dict(bytes=modes * 2, flops=2 * gemm, peak="mfma_f32") if steps == 1 else
                       dict(bytes=2 * modes + modes + modes, flops=4 * gemm, peak="mfma_f32"))
    else:  # refined, binary32 mode solves; the first step reads the binary64 right-hand side
        m32 = modes // 2
        w["zfwd"] = dict(bytes=modes + m32 + (steps - 1) * 2 * m32, flops=steps * gemm, peak="mfma_f32")
        w["rsolve"] = dict(bytes=steps * 2 * m32, flops=None, peak=None)
        # X written, then read and written; two steps keep their one stored
        # iterate (a binary32 sum) in binary32
        xs = m32 if steps == 2 else modes
        xb = [m32 + (modes if steps == 1 else xs)] + [m32 + 2 * modes] * (steps - 2) + [m32 + xs + modes] * (steps > 1)
        w["zinv"] = dict(bytes=sum(xb), flops=steps * gemm, peak="mfma_f32")
        w["resid"] = dict(bytes=(steps - 1) * (modes + xs + m32), flops=None, peak=None)
    return w


# stage -> name prefixes of the kernels its timed launches run
STAGE_KERNELS = {"fft_fwd": ("k_fft_forward<",), "fft_inv": ("k_fft_inverse<",),
                 "zfwd": ("k_zgemm_reg<true,",), "zinv": ("k_zgemm_reg<false,",),
                 "rsolve": ("k_rsolve<",), "cr": ("k_cr_group<",), "resid": ("k_spec_residual",),
                 "planes": ("k_refine_plane<",),
                 "bdygen": ("k_bdy_",), "ghost": ("k_zaxphi<",)}


def hbm_traffic(stage, launches_per_step, batch, workload):
    """HBM bytes per timed launch of a stage, from rocprofv3 FETCH_SIZE and
    WRITE_SIZE per dispatch of each of its kernels (each counter scaled by
    what it read for a device-to-device copy of known size), weighted by the
    kernel's dispatches per batch step: profiles/<workload>_batch<B>_
    hbm_traffic.txt.  None without that summary."""
    f = ROOT / "profiles" / f"{workload}_batch{batch}_hbm_traffic.txt"
    pre = STAGE_KERNELS.get(stage)
    if pre is None or not f.exists() or launches_per_step <= 0:
        return None
    tot, seen = 0.0, False
    for line in f.read_text().splitlines():
        if line.startswith("kernel="):
            kv = dict(p.split("=", 1) for p in line.split("\t"))
            if kv["kernel"].startswith(pre):
                tot += float(kv["traffic_bytes"]) * float(kv["per_step"])
                seen = True
    return tot / launches_per_step if seen else None


def copy_rate(stream, mib=2048, reps=10):
    """GB/s (read + write) of a device-to-device copy of `mib` MiB, beyond the
    256 MB Infinity Cache: the practical rate of a kernel that streams as
    many bytes in as out, measured in the same run as the stages."""
    import torch
    src = torch.ones(mib << 17, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    with torch.cuda.stream(stream):
        for _ in range(2):
            dst.copy_(src)
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record(stream)
        for _ in range(reps):
            dst.copy_(src)
        t1.record(stream)
    t1.synchronize()
    gbs = 2.0 * src.numel() * 8 * reps / (t0.elapsed_time(t1) * 1e-3) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    return gbs


def rooflines(stage_times, work, nsolve, fields_per_launch, steps, workload, copy_gbs=None):
    """Per stage of a timed region of nsolve solves over `steps` batch steps:
    total ms, launches, the algorithmic rate against the HBM peak (and
    against a large device copy's rate measured in the same run, copy_gbs,
    and the compute peak the stage's flops are priced against), and the HBM
    bytes rocprofv3 counted per launch (profiles/) beside the algorithmic
    ones."""
    out = {}
    for st, (ms, n) in stage_times.items():
        if not n or st not in work:
            continue
        w = work[st]
        e = {"ms_total": round(ms, 4), "launches": n, "ms_per_launch": round(ms / n, 5),
             "fields_per_launch": fields_per_launch}
        if w["bytes"]:
            gbs = w["bytes"] * nsolve / (ms * 1e-3) / 1e9
            e.update(bytes_per_field=w["bytes"], achieved_GBs=round(gbs, 1), frac_hbm=round(gbs / HBM_PEAK_GBS, 4))
            if copy_gbs:
                e["frac_of_copy"] = round(gbs / copy_gbs, 4)
            tr = hbm_traffic(st, n / steps, fields_per_launch, workload)
            if tr is not None:
                alg = w["bytes"] * nsolve / n
                e.update(traffic_bytes_per_launch=round(tr), traffic_vs_algorithmic=round(tr / alg, 3))
        if w["flops"]:
            tf = w["flops"] * nsolve / (ms * 1e-3) / 1e12
            e.update(flops_per_field=w["flops"], achieved_TFLOPs=round(tf, 2), compute_peak=w["peak"],
                     frac_compute=round(tf / PEAKS[w["peak"]], 4))
        out[st] = e
    return out


def hetero_batch(ctx, n, refine_steps, reps=3):
    """The heterogeneous batch: one batched call over n independent fields of mixed grids
    (fields.random_field_device for seeds 1000.., generated on the GPU), each
    on its own prepared grid (fields on identical grids share launches);
    fields per second over `reps` calls, refined and plain."""
    import torch

    import chypot
    from chypot import fields
    from chypot import autotune
    jobs, points, shapes = [], 0, {}
    for i in range(n):
        g, rho_d = fields.random_field_device(1000 + i)
        pg = chypot.PreparedGrid(ctx, g)
        key = (g.jmax, g.kmax, g.lmax)
        shapes.setdefault(key, [g, 0])[1] += 1
        jobs.append(dict(grid=pg, rho=rho_d, phi=torch.zeros_like(rho_d),
                         coef=torch.zeros(pg.coef_shape, dtype=torch.float64, device="cuda"),
                         coef_mode=chypot.COEF_K2))
        points += g.jmax * g.kmax * g.lmax
    # launch shapes of every grid class in the set, by the autotuner, at the
    # batch sizes the set has (1 and its largest group)
    t_tune = time.perf_counter()
    most = max(c for _, c in shapes.values())
    table = autotune.tune(ctx, [g for g, _ in shapes.values()], batch_sizes=tuple(sorted({1, most})),
                          refine_steps=refine_steps, reps=5)
    for j in jobs:
        autotune.apply(j["grid"], table)
    tune_s = time.perf_counter() - t_tune
    # many small groups: more worker streams overlap more of them (16 measured
    # 5-6% faster than 8 on this set)
    workers = 16
    ctx.workers(workers)
    out = {"fields": n, "distinct_grids": len(shapes), "refine_steps": refine_steps, "workers": workers,
           "autotune": {"seconds": round(tune_s, 2), "batch_sizes": sorted({1, most}),
                        "one_kernel_per_plane_classes": sum(e["refine_fused"] > 0 for e in table.values())}}
    for label, steps in (("refined", refine_steps), ("plain_fp64", 0)):
        ctx.solve_batch(jobs, refine_steps=steps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.solve_batch(jobs, refine_steps=steps)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        out[label] = {"batch_ms": dt * 1e3, "fields_per_s": n / dt, "grid_points_per_s": points / dt}
    for j in jobs:
        j["grid"].close()
    return out


def serve_run(ctx, args, tune_table=None):
    """The served mix: the serving loop (chypot.serve) over Poisson arrivals,
    refined where the grid allows it: throughput and per-request latency
    from arrival to completion."""
    import torch

    from chypot import serve
    mix = [((128, 64, 128), 0.6), ((256, 128, 256), 0.3), ((512, 256, 512), 0.1)]
    dens = {shape: serve.density_pool(shape, 2, seed=17, on_device=True) for shape, _ in mix}
    torch.cuda.synchronize()
    reqs = serve.poisson_requests(args.serve_rate, args.serve, mix, dens, seed=5)
    srv = serve.Server(ctx, hbm_budget=int(args.serve_budget_gib * (1 << 30)), max_batch=16, workers=8,
                       refine_steps=args.refine_steps, tune_table=tune_table)
    srv.prepare([shape for shape, _ in mix], per_shape=4)
    t0 = time.perf_counter()
    done = srv.run(reqs)
    wall = time.perf_counter() - t0
    srv.close()
    out = serve.latency_summary(done)
    out.update({"offered_rate_per_s": args.serve_rate, "seconds": args.serve, "budget_gib": args.serve_budget_gib,
                "completed_per_s": len(done) / wall, "refine_steps": args.refine_steps,
                "mix": "60% 128x64x128, 30% 256x128x256, 10% 512x256x512 (refined where KMAX <= 256: all)"})
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


def _refpot_times(exe, g, rho, threads, nwarm, nrep):
    with tempfile.TemporaryDirectory() as td:
        fin, fout = pathlib.Path(td) / "in.bin", pathlib.Path(td) / "out.bin"
        with open(fin, "wb") as f:
            f.write(np.array([g.jmax, g.kmax, g.lmax, 2, 10, 8, nwarm, nrep], dtype="<i4").tobytes())
            f.write(np.array([g.delr, g.delz, 0.0, 0.0], dtype="<f8").tobytes())
            f.write(np.asarray(rho, dtype="<f8").tobytes(order="F"))
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_STACKSIZE="512M")
        subprocess.run([str(exe), "solve", str(fin), str(fout)], check=True, env=env, stdout=subprocess.DEVNULL)
        raw = fout.read_bytes()
    n = int(np.prod(g.shape))
    nc = (g.jmax + 2) * (g.kmax + 2) * (g.lmax // 2) * 2
    off = (2 * n + 1 + nc + 3) * 8
    used = int(np.frombuffer(raw, dtype="<i4", count=1, offset=off)[0])
    return used, np.frombuffer(raw, dtype="<f8", count=nrep, offset=off + 4).copy()


def host_cores():
    """Cores this process may use (the affinity mask, capped by OMP_NUM_THREADS when set)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    return min(n, int(env)) if env and env.isdigit() and int(env) > 0 else n


def cpu_baseline(g, rho, nrep):
    """The reference's BDYGEN+POT3 (pot/*.f, flang -O3 -fdefault-real-8,
    oracle/_ref) on this host: all cores and one core, 3 warm-ups and nrep
    timed solves each.  Without that build (it is made in a tree holding the
    reference) it says so on stderr and times the oracle's C restatement."""
    exe = ROOT / "oracle" / "_ref" / f"r8_{g.name}" / "refpot"
    cores = host_cores()
    if exe.exists():
        res = {}
        for label, th in (("all_cores", cores), ("one_core", 1)):
            used, secs = _refpot_times(exe, g, rho, th, 3, nrep)
            res[label] = {"threads": used, "p50_s": float(np.percentile(secs, 50)),
                          "p99_s": float(np.percentile(secs, 99)), "solves_per_s": float(1.0 / np.mean(secs))}
        a = res["all_cores"]
        return {"value": a["solves_per_s"], "unit": "solves/s", "cores": a["threads"], "kind": "reference",
                "sample": f"3 warm-ups then {nrep} timed BDYGEN+POT3 solves of the 256x128x256 field by the "
                          f"reference's pot/*.f (flang -O3 -fdefault-real-8, OpenMP), {a['threads']} threads; "
                          f"p50 {a['p50_s'] * 1e3:.1f} ms, p99 {a['p99_s'] * 1e3:.1f} ms",
                "one_core": res["one_core"], "all_cores": a}
    print(f"bench.py: the reference build {exe} is missing (build it with __graft_entry__.build() where the "
          "reference tree is present); cpu_baseline times the oracle's C restatement instead", file=sys.stderr)
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle
    pyoracle.use_cyclic_reduction(True)
    pyoracle.solve(g, rho)
    secs = []
    for _ in range(nrep):
        t = time.perf_counter()
        pyoracle.solve(g, rho)
        secs.append(time.perf_counter() - t)
    pyoracle.use_cyclic_reduction(False)
    return {"value": 1.0 / float(np.mean(secs)), "unit": "solves/s", "cores": cores, "kind": "port",
            "sample": f"{nrep} solves of the same field by the oracle's C restatement (BLKTRI's reduction), "
                      f"{cores} threads; the reference build was missing"}


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=32, help="independent fields per step")
    ap.add_argument("--workers", type=int, default=0, help="worker streams (default min(batch, 8))")
    ap.add_argument("--refine-steps", type=int, default=2, help="correction steps of the refined solve")
    ap.add_argument("--paths", default="refined,cr,eigen,refined_fused",
                    help="which paths to time (refined is the headline; a profile may time one)")
    ap.add_argument("--cpu-reps", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--hetero", type=int, default=64,
                    help="also time one batch of this many heterogeneous fields, refined and plain (0: skip)")
    ap.add_argument("--serve", type=float, default=0.5,
                    help="also run the serving loop for this many seconds of Poisson arrivals (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=200.0, help="arrivals per second")
    ap.add_argument("--serve-budget-gib", type=float, default=64.0, help="HBM budget of the server")
    ap.add_argument("--latency-steps", type=int, default=-1,
                    help="single-field solves per path for p50/p99 (default: 2 x --steps; 0 skips)")
    ap.add_argument("--source-calls", type=int, default=20,
                    help="also time this many calls of the SOURCE momentum sources (0: skip)")
    ap.add_argument("--error-steps", type=int, default=4,
                    help="refinement steps whose error against the plain FP64 result is reported "
                         "(single-field solves after the timed region; 0 skips)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the library's default launch shapes instead of timing the candidates first")
    ap.add_argument("--cr-lanes", type=int, default=0, help="set the cyclic reduction's lanes per r system (64, 32)")
    ap.add_argument("--refine-fused", type=int, default=-1,
                    help="set the largest fields-per-launch of the one-kernel-per-plane refined solve")
    ap.add_argument("--coef-full-steps", type=int, default=5,
                    help="batch steps of the refined solve writing COEF for every K (0: skip)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="before the solves, copy a buffer of this many MiB device to device "
                         "three times (a kernel of known traffic for scaling HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the potential solve of one field runs on one MI355X; "
              "splitting a field over several GPUs is not implemented", file=sys.stderr)
        sys.exit(2)
    paths = [p for p in args.paths.split(",") if p]

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
        pg = chypot.PreparedGrid(ctx, g)
        jobs.append(dict(grid=pg, rho=chypot.to_device(rho_b),
                         phi=torch.zeros(pg.field_shape, dtype=torch.float64, device="cuda"),
                         coef=torch.zeros(pg.coef_shape, dtype=torch.float64, device="cuda"),
                         coef_mode=chypot.COEF_K2))
    j0 = jobs[0]
    tune_table = None
    if args.autotune:
        # launch shapes of the bench grid for single fields and for the batch
        from chypot import autotune
        t_tune = time.perf_counter()
        tune_table = autotune.tune(ctx, [g], batch_sizes=(1, B), refine_steps=args.refine_steps, reps=5)
        for j in jobs:
            autotune.apply(j["grid"], tune_table)
        ctx.set_stream(stream)
        tune_table[autotune.grid_key(g)]["seconds"] = round(time.perf_counter() - t_tune, 2)
    for j in jobs:  # shapes given on the command line win
        if args.cr_lanes:
            j["grid"].cr_lanes = args.cr_lanes
        if args.refine_fused >= 0:
            j["grid"].refine_fused = args.refine_fused

    copy_gbs = copy_rate(stream)
    if args.copy_mib > 0:
        src = torch.ones(args.copy_mib << 17, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

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

    def set_solver(kind):
        for j in jobs:
            j["grid"].mode_solver = kind

    def batch_step(steps):
        return lambda: ctx.solve_batch(jobs, refine_steps=steps)

    def single(steps):
        if steps:
            return lambda: j0["grid"].solve_refined(j0["rho"], j0["phi"], steps=steps, coef=j0["coef"],
                                                    coef_mode=chypot.COEF_K2)
        return lambda: j0["grid"].solve(j0["rho"], j0["phi"], coef=j0["coef"], coef_mode=chypot.COEF_K2)

    n_lat = 2 * args.steps if args.latency_steps < 0 else args.latency_steps
    fx = np.load(ROOT / "tests" / "golden" / "solve_large_256x128x256.npz")
    scale = float(fx["phi_absmax"])

    def err_ref(phi):
        return max(float(np.abs(phi[:, 1, :] - fx["phi_k2"]).max() / scale),
                   float(np.abs(phi[:, :, 0] - fx["phi_l1"]).max() / scale))

    def run_path(label, solver, steps):
        """warm-up, then the timed region (stage events on), then latency."""
        set_solver(solver)
        fn = batch_step(steps)
        for _ in range(args.warmup):
            fn()
        ctx.profile(True)
        elapsed, lat_batch = timed(fn, args.steps)
        st = ctx.stage_times()
        ctx.profile(False)
        phi = chypot.to_host(j0["phi"])
        res = {"solves_per_s": B * args.steps / elapsed, "ms_per_step": elapsed / args.steps * 1e3,
               "batch_step_p50_ms": float(np.percentile(lat_batch, 50)),
               "batch_step_p99_ms": float(np.percentile(lat_batch, 99)),
               "max_rel_err_vs_reference_fp64": err_ref(phi)}
        kind = "refined" if steps else ("cr" if solver == chypot.MODES_CR else "eigen")
        if steps and st.get("planes", (0, 0))[1]:
            kind = "refined_fused"
        work = stage_work(g, kind, steps)
        res["rooflines"] = rooflines(st, work, B * args.steps, B, args.steps, f"potsolve_{g.name}_{kind}", copy_gbs)
        if n_lat > 0:
            one = single(steps)
            for _ in range(3):
                one()
            lat = timed(one, n_lat)[1]
            res.update(p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                       single_solves_per_s=float(1e3 / np.mean(lat)))
        return res, phi, elapsed

    out_paths = {}
    phis = {}
    elapsed_ref = None
    if "cr" in paths:
        out_paths["plain_fp64"], phis["cr"], _ = run_path("plain_fp64", chypot.MODES_CR, 0)
        out_paths["plain_fp64"]["mode_solver"] = "cyclic reduction (BLKTRI's root tables)"
    if "eigen" in paths:
        out_paths["plain_fp64_eigen"], phis["eigen"], _ = run_path("eigen", chypot.MODES_EIGEN, 0)
        out_paths["plain_fp64_eigen"]["mode_solver"] = "z eigenbasis (opt-in)"
    set_solver(chypot.MODES_CR)
    refined = None
    if "refined" in paths:
        refined, phis["refined"], elapsed_ref = run_path("refined", chypot.MODES_CR, args.refine_steps)
        refined["refine_steps"] = args.refine_steps
        refined["refine_m64"] = j0["grid"].refine_m64
        # the error against the plain FP64 result, after each refinement step (field 0)
        plain = phis.get("cr")
        if plain is None and args.error_steps > 0:
            tmp = torch.zeros_like(j0["phi"])
            j0["grid"].solve(j0["rho"], tmp)
            ctx.sync()
            plain = chypot.to_host(tmp)
        if plain is not None:
            pscale = float(np.abs(plain).max())
            refined["max_rel_err_vs_plain_fp64"] = float(np.abs(phis["refined"] - plain).max() / pscale)
            per_step = []
            tmp = torch.zeros_like(j0["phi"])
            for st in range(1, args.error_steps + 1):
                j0["grid"].solve_refined(j0["rho"], tmp, steps=st)
                ctx.sync()
                per_step.append(float(np.abs(chypot.to_host(tmp) - plain).max() / pscale))
            refined["error_vs_plain_fp64_per_step"] = per_step
    if refined is not None and args.coef_full_steps > 0:
        # the same batched refined solve writing COEF for every K, as the
        # reference's POT3 does (the timed region above writes the K=2 plane),
        # beside the K=2 form timed the same way (HIP graph replays)
        cf = {}
        for mode, key in ((chypot.COEF_K2, "k2_plane"), (chypot.COEF_FULL, "every_k")):
            for j in jobs:
                j["coef_mode"] = mode
            fn = batch_step(args.refine_steps)
            for _ in range(2):
                fn()
            el, lat = timed(fn, args.coef_full_steps)
            cf[key] = {"solves_per_s": B * args.coef_full_steps / el, "batch_step_p50_ms": float(np.percentile(lat, 50))}
        for j in jobs:
            j["coef_mode"] = chypot.COEF_K2
        cf["steps"] = args.coef_full_steps
        refined["coef"] = cf
    if "refined_fused" in paths:
        # the same refined solve with all mode work of a plane in one kernel
        # (binary32 plane in LDS, one HBM round trip): the comparison row
        tuned_fused = j0["grid"].refine_fused
        for j in jobs:
            j["grid"].refine_fused = True
        fused, phi_f, _ = run_path("refined_fused", chypot.MODES_CR, args.refine_steps)
        for j in jobs:
            j["grid"].refine_fused = tuned_fused
        ref_phi = phis.get("refined")
        if ref_phi is not None:
            fused["max_rel_diff_vs_refined"] = float(np.abs(phi_f - ref_phi).max() / np.abs(ref_phi).max())
        out_paths["refined_one_kernel_per_plane"] = fused

    hetero = hetero_batch(ctx, args.hetero, args.refine_steps) if args.hetero > 0 else None
    ctx.workers(W)
    serving = serve_run(ctx, args, tune_table) if args.serve > 0 else None
    ctx.set_stream(stream)
    source = (source_row(ctx, stream, g, args.source_calls, not args.no_cpu_baseline)
              if args.source_calls > 0 else None)
    cpu = None if args.no_cpu_baseline else cpu_baseline(g, rho, args.cpu_reps)

    head = (refined if refined is not None else out_paths.get("plain_fp64") or out_paths.get("plain_fp64_eigen")
            or out_paths.get("refined_one_kernel_per_plane"))
    value = head["solves_per_s"]
    plain_v = out_paths.get("plain_fp64", {}).get("solves_per_s")
    # roofline: the dominant stage of the headline timed region
    rl = head["rooflines"]
    dom = max(rl, key=lambda k: rl[k]["ms_total"])
    d = rl[dom]
    if d.get("frac_compute", 0) > d.get("frac_hbm", 0):
        roof = {"bound": "mfma" if d["compute_peak"].startswith("mfma") else "valu", "achieved": d["achieved_TFLOPs"],
                "peak": PEAKS[d["compute_peak"]], "unit": "TFLOP/s", "frac": d["frac_compute"]}
    else:
        roof = {"bound": "hbm", "achieved": d.get("achieved_GBs"), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": d.get("frac_hbm")}
    roof.update({"hbm_copy_GBs": round(copy_gbs, 1), "hbm_frac_of_copy": d.get("frac_of_copy"),
                 "traffic": d.get("traffic_bytes_per_launch"), "algorithmic_bytes_per_launch":
                 None if not d.get("bytes_per_field") else round(d["bytes_per_field"] * B * args.steps / d["launches"]),
                 "kernel": dom, "avg_ms": d["ms_per_launch"], "launches": d["launches"], "fields_per_launch": B})
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if refined is not None else "f64",
        "data": "synthetic",
        "config": {"workload": "refined potential solve (BDYGEN+POT3), batched, spiral-arm torus"
                   if refined is not None else "plain FP64 potential solve (BDYGEN+POT3), batched, spiral-arm torus",
                   "grid": g.name, "path": "refined" if refined is not None else "plain_fp64", "batch": B,
                   "workers": W, "refine_steps": args.refine_steps if refined is not None else 0,
                   "coef": "K=2 plane", "maxtrm": 10, "isym": 2, "npoint": 8},
        "vs_plain": None if (refined is None or not plain_v) else value / plain_v,
        "vs_plain_eigen": None if (refined is None or "plain_fp64_eigen" not in out_paths)
        else value / out_paths["plain_fp64_eigen"]["solves_per_s"],
        "p50_ms": head.get("p50_ms"),
        "p99_ms": head.get("p99_ms"),
        "refined": refined,
        **out_paths,
        "hetero_batch": hetero,
        "serving": serving,
        "source_grav": source,
        "roofline": roof,
        "autotune": tune_table,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
