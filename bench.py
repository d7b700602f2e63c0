#!/usr/bin/env python3
"""Potential-solve benchmark on one MI355X.

A step is one potential solve -- BDYGEN + POT3 on one density field, device
resident, COEF K=2 plane -- of the 256x128x256 spiral-arm torus (JMAX=256,
KMAX=128, LMAX=256, DELR=DELZ=1/256, seeded).  `value` is solves per second over
the timed steps with the density already in HBM.  The line also carries the
per-step p50/p99 latency, the error against the fixture solutions (reference
FP64 and the 80-bit solve of the same discrete system), per-stage times from
HIP events on the solve's stream, the roofline of the dominant stage and the
reference CPU solve timed on this host.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
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
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="before the solves, copy a buffer of this many MiB device to device "
                         "three times (a kernel of known traffic for scaling HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the potential solve of one field runs on one MI355X "
              "(multi-GPU sharding is described in DESIGN.md, not run)", file=sys.stderr)
        sys.exit(2)

    import torch

    import chypot
    from chypot import fields

    g = fields.unit_grid(256, 128, 256)
    rho = fields.density_spiral_torus(g, seed=2)
    ctx = chypot.Context(0)
    # one explicit (default-priority) stream for the solves and the timing events
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    ctx.set_stream(stream)
    pg = chypot.PreparedGrid(ctx, g)
    rho_d = chypot.to_device(rho)
    phi_d = torch.zeros_like(rho_d)
    coef_d = torch.zeros((2, g.lmax // 2, g.kmax + 2, g.jmax + 2), dtype=torch.float64, device="cuda")

    def step():
        pg.solve(rho_d, phi_d, coef=coef_d, coef_mode=chypot.COEF_K2)

    if args.copy_mib > 0:
        src = torch.ones(args.copy_mib << 17, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # timed region: graph replays of the solve (stage timing off)
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        step()
        ends[i].record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    lat = np.array([s.elapsed_time(e) for s, e in zip(starts, ends)])
    elapsed = t1 - t0
    value = args.steps / elapsed
    phi = chypot.to_host(phi_d)

    # the same steps launched kernel by kernel with HIP events around every
    # stage on the solve's stream: per-kernel durations for the roofline
    ctx.profile(True)
    for _ in range(args.steps):
        step()
    stages = ctx.stage_times()
    ctx.profile(False)

    # accuracy of the timed result against the fixture solutions
    fx = np.load(ROOT / "tests" / "golden" / "solve_large_256x128x256.npz")
    scale = float(fx["phi_absmax"])

    def err(a, b):
        return float(np.abs(a - b).max() / scale)

    e_ref = max(err(phi[:, 1, :], fx["phi_k2"]), err(phi[:, :, 0], fx["phi_l1"]))
    e_ext = max(err(phi[:, 1, :], fx["ext_phi_k2"]), err(phi[:, :, 0], fx["ext_phi_l1"]))

    work = stage_work(g)
    stage_avg = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in stages.items()}
    dom = max(stage_avg, key=stage_avg.get)
    avg_ms = stage_avg[dom]
    wk = work[dom]
    if wk["flops"] is not None and wk["flops"] / wk["bytes"] > FP64_PEAK_TFLOPS * 1e12 / (HBM_PEAK_GBS * 1e9):
        achieved = wk["flops"] / (avg_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None, "kernel": dom,
                "avg_ms": avg_ms, "launches": stages[dom][1]}
    else:
        achieved = (wk["bytes"] or 0) / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dom, "avg_ms": avg_ms,
                "launches": stages[dom][1]}

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
                   "grid": g.name, "path": "plain_fp64", "coef": "K=2 plane",
                   "maxtrm": 10, "isym": 2, "npoint": 8},
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "max_rel_err_vs_reference_fp64": e_ref,
        "max_rel_err_vs_exact": e_ext,
        "stages_ms": {k: round(v, 5) for k, v in stage_avg.items()},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
