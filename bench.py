#!/usr/bin/env python3
"""Headline benchmark of the closed-shell CC hot path on one MI355X.

Workload: synthetic density-fitted closed-shell system, o=64 occupied,
v=512 virtual, naux=2048 (B ~ N(0, 0.05^2) mirrored in p<=q, eps_occ
U(-1.5,-0.3), eps_vir U(0.1,3.0), MP2 start).  One step = one CCSD
iteration on device-resident data: T1-dressed DF residual (every
contraction on MFMA), Jacobi step, DIIS (restart every 8), energy.

Prints one JSON line.  value = seconds per CCSD iteration (lower is
better); the line also carries the algorithmic TFLOP/s of the iteration,
the roofline of the dominant kernel measured with HIP events on its own
stream, and a CPU baseline (this repository's numpy restatement of the
algorithm, timed on the host on a smaller instance and scaled by the
algorithmic flop ratio).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}  # MI355X dense matrix TFLOP/s


def f_ccsd(o: int, v: int, naux: int) -> float:
    """Algorithmic flops of one closed-shell DF-CCSD iteration (multiply-add = 2):
    +/- ladder, DF ladder-integral formation, ring (8 o^3 v^3), hole ladder,
    DF ring integrals and singles."""
    npv, npo = v * (v + 1) / 2, o * (o + 1) / 2
    return (4 * npv * npv * npo + naux * v**3 * (v + 1) + 8 * o**3 * v**3 + 4 * o**4 * v**2
            + 4 * naux * o**2 * v**2 + 2 * naux * o * v**3 + 4 * o**2 * v**3)


def sync():
    try:
        import torch

        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def cpu_baseline(sample=(16, 128, 512), seed=3):
    """The BLAS restatement of the residual in oracle/ (numpy + multithreaded
    OpenBLAS) on a smaller instance, scaled to o=64/v=512/naux=2048 by the
    algorithmic flop ratio."""
    from dalton_amd.synthetic import make_system
    from oracle import ccsd as occ
    from oracle import ccsd_blas

    e0, B0 = make_system(4, 12, 20, seed)  # warm the BLAS threads
    s0 = occ.DFSystem(e0, B0, 4)
    ccsd_blas.residual(s0, *occ.mp2_guess(s0))
    o, v, naux = sample
    eps, B = make_system(o, v, naux, seed)
    sysm = occ.DFSystem(eps, B, o)
    t1, t2 = occ.mp2_guess(sysm)
    t0 = time.perf_counter()
    ccsd_blas.residual(sysm, t1, t2)
    dt = time.perf_counter() - t0
    try:
        from threadpoolctl import threadpool_info

        info = threadpool_info()
        cores = max(int(i.get("num_threads", 1)) for i in info) if info else 1
    except Exception:
        cores = os.cpu_count() or 1
    scale = f_ccsd(64, 512, 2048) / f_ccsd(o, v, naux)
    return {
        "value": round(dt * scale, 2),
        "unit": "s/iter",
        "cores": cores,
        "kind": "port",
        "sample": f"one residual of the BLAS restatement (oracle/ccsd_blas.py) at o={o} v={v} naux={naux} "
                  f"({dt:.2f} s), scaled x{scale:.1f} by algorithmic flops to o=64 v=512 naux=2048",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--o", type=int, default=64)
    ap.add_argument("--v", type=int, default=512)
    ap.add_argument("--naux", type=int, default=2048)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--pt-triples", type=int, default=24,
                    help="(T) triples timed, spread evenly over the i>=j>=k list (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1 "
              "(how the path would shard is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import dalton_amd as d
    from dalton_amd.cc import profile_start, profile_stop
    from dalton_amd.synthetic import make_system

    o, v, naux = args.o, args.v, args.naux
    t0 = time.perf_counter()
    eps, B = make_system(o, v, naux, args.seed)
    print(f"[bench] system o={o} v={v} naux={naux} generated in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    ctx = d.CCContext(eps, B, o)
    del B
    e_mp2 = ctx.guess()
    print(f"[bench] E_MP2 = {e_mp2:.10f}", file=sys.stderr, flush=True)
    for w in range(args.warmup):
        e, n = ctx.iterate()
        print(f"[bench] warmup {w + 1}: E = {e:.10f} |Omega| = {n:.3e}", file=sys.stderr, flush=True)
    sync()
    profile_start()
    t0 = time.perf_counter()
    hist = []
    for s in range(args.steps):
        e, n = ctx.iterate()
        hist.append(e)
        print(f"[bench] step {s + 1}: E = {e:.10f} |Omega| = {n:.3e}", file=sys.stderr, flush=True)
    sync()
    dt = time.perf_counter() - t0
    prof = profile_stop()

    pt = None
    if args.pt_triples > 0:
        _, _, _, ntr = ctx.pt_partial(0, 0, 1)  # build (T) integral layouts (untimed setup)
        stride = max(1, ntr // args.pt_triples)
        sync()
        profile_start()
        t1 = time.perf_counter()
        e4, e5, fl, _ = ctx.pt_partial(0, args.pt_triples, stride)
        sync()
        dtp = time.perf_counter() - t1
        pprof = profile_stop()
        pk = pprof["kernels"]
        pdom = max(pk, key=lambda k: pk[k]["ms"])
        pt = {"triples_timed": args.pt_triples, "triples_total": ntr, "stride": stride,
              "seconds": round(dtp, 4), "tflops": round(fl / dtp / 1e12, 3),
              "est_full_seconds": round(dtp / args.pt_triples * ntr, 1),
              "flops_per_triple": fl / args.pt_triples, "e4_sample": e4, "e5_sample": e5,
              "dominant_kernel": pdom,
              "kernels": {k: {"count": v_["count"], "ms": round(v_["ms"], 3),
                              "tflops": round(v_["flops"] / max(v_["ms"], 1e-9) / 1e9, 2)} for k, v_ in pk.items()}}
        print(f"[bench] (T): {args.pt_triples} triples in {dtp:.3f} s, {fl / dtp / 1e12:.2f} TFLOP/s", file=sys.stderr,
              flush=True)
    ctx.close()

    spi = dt / args.steps
    F = f_ccsd(o, v, naux)
    kern = prof["kernels"]
    dom = max(kern, key=lambda k: kern[k]["ms"])
    kd = kern[dom]
    achieved = kd["flops"] / (kd["ms"] * 1e-3) / 1e12
    roof = {
        "bound": "mfma",
        "achieved": round(achieved, 3),
        "peak": PEAK["fp64"],
        "unit": "TFLOP/s",
        "frac": round(achieved / PEAK["fp64"], 4),
        "traffic": None,
        "kernel": dom,
        "launches": kd["count"],
        "avg_launch_ms": round(kd["ms"] / kd["count"], 4),
        "share_of_step": round(kd["ms"] * 1e-3 / dt, 4),
    }
    out = {
        "metric": "CCSD s/iter + (T) TFLOP/s at o64/v512: mixed vs FP64 MFMA, |dE|<=1uEh",
        "value": round(spi, 4),
        "unit": "s/iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(spi * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"synthetic DF-CCSD iteration, o={o} v={v} naux={naux}, seed {args.seed}, MP2 start",
                   "o": o, "v": v, "naux": naux, "precision": "fp64"},
        "ccsd_tflops": round(F / spi / 1e12, 3),
        "ccsd_flops_per_iter": F,
        "energies": hist,
        "pt": pt,
        "roofline": roof,
        "phases_ms_per_iter": {k: round(v_ / args.steps, 2) for k, v_ in prof["phases"].items()},
        "kernels": {k: {"count": v_["count"], "ms": round(v_["ms"], 3),
                        "tflops": round(v_["flops"] / max(v_["ms"], 1e-9) / 1e9, 2)} for k, v_ in kern.items()},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
