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

# HBM-side bytes per launch of the dominant kernels at o=64 v=512 naux=2048,
# from two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) scaled by the same
# counters read on a copy of known size; the summaries and commands are in
# profiles/gemm_bf16x3_o64_v512_traffic.md and profiles/dgemm_o64_v512_traffic.md.
TRAFFIC_BYTES = {"k_gemm_bf16x3": 59.0e9, "k_dgemm<0,1,false,128>": 1.4979e12}

# Attainable dense rates of the vendor BLAS on this card (torch.matmul 8192^3,
# profiles/vendor_blas_attainable_rates.md): bf16 1086, fp64 71.5 TFLOP/s.
ATTAINABLE = {"bf16x3": 1086.2 / 3, "fp64": 71.5}


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


def cpu_baseline(o=32, v=256, naux=1024, seed=3):
    """The BLAS restatement of the residual in oracle/ (numpy + multithreaded
    OpenBLAS on the host cores): one whole residual of the o=32 v=256
    naux=1024 synthetic system at its MP2 amplitudes, and that time scaled to
    o=64 v=512 naux=2048 by the ratio of the algorithmic flop counts."""
    from dalton_amd.synthetic import make_system
    from oracle import ccsd as occ
    from oracle import ccsd_blas

    e0, B0 = make_system(4, 12, 20, seed)  # warm the BLAS threads
    s0 = occ.DFSystem(e0, B0, 4)
    ccsd_blas.residual(s0, *occ.mp2_guess(s0))
    eps, B = make_system(o, v, naux, seed)
    sysm = occ.DFSystem(eps, B, o)
    t1, t2 = occ.mp2_guess(sysm)
    t0 = time.perf_counter()
    ccsd_blas.residual(sysm, t1, t2)
    dt = time.perf_counter() - t0
    cores = len(os.sched_getaffinity(0))
    try:
        from threadpoolctl import threadpool_info

        blas = max((int(i.get("num_threads", 1)) for i in threadpool_info() if i.get("user_api") == "blas"),
                   default=cores)
    except Exception:
        blas = cores
    scale = f_ccsd(64, 512, 2048) / f_ccsd(o, v, naux)
    return {
        "value": round(dt * scale, 2),
        "unit": "s/iter",
        "cores": min(cores, blas),
        "kind": "port",
        "sample": f"one whole residual of the BLAS restatement (oracle/ccsd_blas.py, numpy + OpenBLAS with "
                  f"{min(cores, blas)} threads) at o={o} v={v} naux={naux}: {dt:.2f} s, "
                  f"{f_ccsd(o, v, naux) / dt / 1e9:.1f} GFLOP/s; scaled x{scale:.1f} by algorithmic flops "
                  f"to o=64 v=512 naux=2048",
        "sample_seconds": round(dt, 3),
        "sample_gflops": round(f_ccsd(o, v, naux) / dt / 1e9, 2),
    }


def cpu_baseline_child():
    """Child-process side of the CPU baseline: the system is built and the
    BLAS threads warmed at once, the timed residual starts when the parent
    writes a line to stdin, and the result is printed as one JSON line."""
    from dalton_amd.synthetic import make_system

    make_system(4, 12, 20, 3)  # imports and first numpy calls
    sys.stdout.write("ready\n")
    sys.stdout.flush()
    if sys.stdin.readline().strip() != "go":  # the parent ended (or gave up) without starting it
        return
    print(json.dumps(cpu_baseline()), flush=True)


def start_cpu_baseline():
    """Starts the CPU baseline in a child process before this process touches
    the GPU; the parent lets it run (go()) once the headline timing is over,
    so it shares the host with the later GPU legs only.  Two host cores are
    left to the process that drives the GPU."""
    import subprocess

    threads = max(1, len(os.sched_getaffinity(0)) - 2)
    env = dict(os.environ)
    for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
        env[k] = str(threads)
    p = subprocess.Popen([sys.executable, os.path.abspath(__file__), "--workload", "cpu-baseline"],
                         stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True, env=env)
    import atexit

    atexit.register(lambda: p.poll() is None and p.kill())  # never outlives a failed run
    return p


def finish_cpu_baseline(p, log):
    out, _ = p.communicate(timeout=900)
    if p.returncode != 0:
        raise RuntimeError(f"CPU baseline process failed with status {p.returncode}")
    res = json.loads([ln for ln in out.splitlines() if ln.startswith("{")][-1])
    log(f"CPU baseline: {res['sample']}")
    return res


def convergent_leg(o, v, naux, seed, ntrip, log, threnr=1e-9, thrvec=1e-7, maxit=60, scale=0.5):
    """The metric's accuracy gate at the headline shape.  The same synthetic
    distribution with the integral scale scale * dalton_amd.synthetic.batch_sigma,
    at which CCSD converges (at scale 1 the FP64 solve at o=64 v=512 takes 48
    iterations, 7.8 minutes; at 0.5 it converges in about a third of that):
    solved in FP64 (every iteration timed, as the
    CCSD_ENERGY loop: |dE| < threnr and |Omega| < thrvec) and in mixed
    precision (the library's loop, ccgpu_ccsd_energy, whose last two
    iterations are FP64), each followed by (T) over ntrip triples spread
    evenly over the i>=j>=k list.  Reports the errors of E_CCSD, E_(T) and
    the amplitudes of the mixed solve against the FP64 one, the FP64 residual
    norm of the mixed amplitudes (one FP64 ccgpu_rhs) and both solve times."""
    import dalton_amd as d
    from dalton_amd.synthetic import batch_sigma, make_system

    sigma = scale * batch_sigma(v, naux)
    t0 = time.perf_counter()
    eps, B = make_system(o, v, naux, seed, sigma=sigma)
    log(f"convergent system o={o} v={v} naux={naux} sigma={sigma:.6f} generated in {time.perf_counter() - t0:.1f} s")
    res = {"workload": f"synthetic o={o} v={v} naux={naux} seed {seed}, B ~ N(0, sigma^2) with "
                       f"sigma = {sigma:.6f} ({scale} x batch_sigma), MP2 start",
           "threnr": threnr, "thrvec": thrvec}
    F = f_ccsd(o, v, naux)

    def pt_subset(c, prec):
        _, _, _, ntr = c.pt_partial(0, 1, 1, precision=prec)  # (T) integral layouts, one triple (untimed)
        stride = max(1, ntr // ntrip)
        sync()
        t0 = time.perf_counter()
        e4, e5, fl, _ = c.pt_partial(0, ntrip, stride, precision=prec)
        sync()
        dt = time.perf_counter() - t0
        return {"e4": e4, "e5": e5, "seconds": round(dt, 4), "tflops": round(fl / max(dt, 1e-9) / 1e12, 2),
                "triples": ntrip, "stride": stride, "triples_total": ntr}

    # FP64: the CCSD_ENERGY loop from the host, one timed iteration at a time
    with d.CCContext(eps, B, o) as c:
        e_prev = c.guess()
        times, hist, norms = [], [], []
        conv = False
        for it in range(1, maxit + 1):
            sync()
            t0 = time.perf_counter()
            e, n = c.iterate(d.FP64)
            times.append(time.perf_counter() - t0)
            hist.append(e)
            norms.append(n)
            log(f"convergent fp64 iteration {it}: E = {e:.12f} dE = {e - e_prev:.2e} |Omega| = {n:.3e} "
                f"({times[-1]:.2f} s)")
            conv = abs(e - e_prev) < threnr and n < thrvec
            e_prev = e
            if conv:
                break
        if not conv:
            raise RuntimeError("convergent leg: the FP64 solve did not converge")
        t1_d, t2_d = c.amplitudes()
        pt_d = pt_subset(c, d.FP64)
        log(f"convergent fp64 (T) subset: E4 = {pt_d['e4']:.14f} E5 = {pt_d['e5']:.14f}")
    res["fp64"] = {"iterations": len(times), "solve_seconds": round(sum(times), 3),
                   "s_per_iter": round(sum(times) / len(times), 4),
                   "s_per_iter_min": round(min(times), 4), "ccsd_tflops": round(F * len(times) / sum(times) / 1e12, 2),
                   "e_ccsd": hist[-1], "history": hist, "omega_norms": norms, "pt_subset": pt_d}
    # mixed: the library's loop (BF16x3 until |dE| < 1e-6 and |Omega| < 1e-3, then FP64)
    with d.CCContext(eps, B, o) as c:
        del B
        sync()
        t0 = time.perf_counter()
        r = c.ccsd_energy(maxit=maxit, threnr=threnr, thrvec=thrvec, precision=d.MIXED)
        sync()
        tm = time.perf_counter() - t0
        log(f"convergent mixed: {r['niter']} iterations in {tm:.2f} s, E = {r['e_ccsd']:.12f}")
        t1_m, t2_m = c.amplitudes()
        om1, om2 = c.rhs(t1_m, t2_m, d.FP64)  # FP64 residual of the mixed amplitudes
        rnorm = float(np.sqrt(np.dot(om1, om1) + np.dot(om2, om2)))
        pt_m = pt_subset(c, d.MIXED)
        log(f"convergent mixed (T) subset: E4 = {pt_m['e4']:.14f} E5 = {pt_m['e5']:.14f}")
    res["mixed"] = {"iterations": r["niter"], "solve_seconds": round(tm, 3), "e_ccsd": r["e_ccsd"],
                    "history": [float(x) for x in r["history"]], "pt_subset": pt_m,
                    "fp64_residual_norm": rnorm}
    dccsd = abs(r["e_ccsd"] - hist[-1])
    dpt = abs(pt_m["e4"] + pt_m["e5"] - pt_d["e4"] - pt_d["e5"])
    ntr, nsub = pt_d["triples_total"], pt_d["triples"]
    res.update({
        "abs_dE_ccsd": dccsd,
        "abs_dE_pt_subset": dpt,
        "abs_dE_pt_full_estimate": dpt * ntr / nsub,
        "abs_dE_ccsd_t_estimate": dccsd + dpt * ntr / nsub,
        "pt_subset_note": f"(T) over {nsub} of {ntr} triples (stride {pt_d['stride']}); the full-list error is "
                          f"estimated by scaling with {ntr}/{nsub}",
        "max_abs_dT1": float(np.abs(t1_m - t1_d).max()),
        "max_abs_dT2": float(np.abs(t2_m - t2_d).max()),
        "residual_norm_mixed_amplitudes_fp64": rnorm,
        "solve_speedup_mixed_vs_fp64": round(res["fp64"]["solve_seconds"] / max(tm, 1e-9), 3),
        "meets_1uEh": bool(dccsd + dpt * ntr / nsub <= 1e-6),
    })
    log(f"convergent: |dE_CCSD| = {dccsd:.2e}, |dE_(T)| subset {dpt:.2e} (full estimate "
        f"{dpt * ntr / nsub:.2e}), max|dT2| = {res['max_abs_dT2']:.2e}, |R| = {rnorm:.2e}")
    return res


def lhtr_leg(nvec, log, kern_table, o=32, v=256, naux=1024, seed=4):
    """Left transformation of nvec random unit trial vectors at the MP2
    amplitudes of the o=32 v=256 naux=1024 synthetic system, one call per
    precision (the forward intermediates are built in an untimed first call)."""
    import dalton_amd as d
    from dalton_amd.cc import profile_start, profile_stop
    from dalton_amd.synthetic import make_system

    eps, B = make_system(o, v, naux, seed)
    ctx = d.CCContext(eps, B, o)
    del B
    ctx.guess()
    rng = np.random.default_rng(seed)
    L1 = rng.standard_normal((nvec, ctx.nt1))
    L2 = rng.standard_normal((nvec, ctx.nt2))
    nrm = np.sqrt((L1 * L1).sum(1) + (L2 * L2).sum(1))
    L1 /= nrm[:, None]
    L2 /= nrm[:, None]
    res = {"workload": f"synthetic o={o} v={v} naux={naux} seed {seed}, MP2 amplitudes, {nvec} unit vectors",
           "flops_per_vector_nominal": 2.0 * f_ccsd(o, v, naux)}
    out = {}
    for label, prec in (("mixed", d.MIXED), ("fp64", d.FP64)):
        ctx.lhtr(L1[0], L2[0], prec)  # forward intermediates at this precision (untimed)
        sync()
        profile_start()
        t0 = time.perf_counter()
        R1, R2 = ctx.lhtr(L1, L2, prec)
        sync()
        dt = time.perf_counter() - t0
        prof = profile_stop()
        gem = sum(k_["flops"] for k_ in prof.get("kernels", {}).values())
        gms = sum(k_["ms"] for k_ in prof.get("kernels", {}).values())
        res[label] = {"seconds": round(dt, 4), "s_per_vector": round(dt / nvec, 4),
                      "gemm_tflops": round(gem / max(gms, 1e-9) / 1e9, 2),
                      "gemm_flops_per_vector": gem / nvec,
                      "phases_ms_per_vector": {k_: round(v_ / nvec, 2) for k_, v_ in prof.get("phases", {}).items()},
                      "kernels": kern_table(prof)}
        out[label] = (R1, R2)
        log(f"lhtr {label}: {nvec} vectors in {dt:.3f} s")
    a1, a2 = out["fp64"]
    b1, b2 = out["mixed"]
    res["max_rel_dev_mixed_vs_fp64"] = float(max(np.abs(a1 - b1).max(), np.abs(a2 - b2).max())
                                             / max(np.abs(a1).max(), np.abs(a2).max()))
    res["speedup_mixed_vs_fp64"] = round(res["fp64"]["seconds"] / res["mixed"]["seconds"], 3)
    ctx.close()
    return res


def kern_table_of(prof):
    return {k: {"count": v_["count"], "ms": round(v_["ms"], 3),
                "tflops": round(v_["flops"] / max(v_["ms"], 1e-9) / 1e9, 2)} for k, v_ in prof.get("kernels", {}).items()}


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def pt_leg(ctx, ntrip, o, v, log):
    """(T) over ntrip triples spread evenly over the list, mixed then FP64,
    from the context's current amplitudes; per-kernel times from the
    library's own event profiler."""
    import dalton_amd as d
    from dalton_amd.cc import profile_start, profile_stop

    pt = {}
    for label, prec in (("mixed", d.MIXED), ("fp64", d.FP64)):
        # integral layouts of this precision and one triple, whose launches
        # settle the block tiles of the mixed products (untimed)
        _, _, _, ntr = ctx.pt_partial(0, 1, 1, precision=prec)
        stride = max(1, ntr // ntrip)
        sync()
        profile_start()
        t0 = time.perf_counter()
        e4, e5, fl, _ = ctx.pt_partial(0, ntrip, stride, precision=prec)
        sync()
        dtp = time.perf_counter() - t0
        pp = profile_stop()
        pt[label] = {"seconds": round(dtp, 4), "tflops": round(fl / dtp / 1e12, 3),
                     "est_full_seconds": round(dtp / ntrip * ntr, 1), "e4": e4, "e5": e5,
                     "kernels": kern_table_of(pp)}
        log(f"(T) {label}: {ntrip} triples in {dtp:.3f} s, {fl / dtp / 1e12:.2f} TFLOP/s")
    pt["triples_timed"] = ntrip
    pt["triples_total"] = ntr
    pt["stride"] = stride
    pt["flops_per_triple"] = 12.0 * v**4 + 12.0 * o * v**3
    for e in ("e4", "e5"):  # relative, or absolute where the FP64 value is zero (E5 at t1 = 0)
        dev = abs(pt["mixed"][e] - pt["fp64"][e])
        pt[f"rel_err_{e}_mixed_vs_fp64"] = dev / abs(pt["fp64"][e]) if pt["fp64"][e] != 0.0 else dev
    pt["speedup_mixed_vs_fp64"] = round(pt["fp64"]["seconds"] / max(pt["mixed"]["seconds"], 1e-9), 3)
    return pt


def pt_workload(o, v, naux, seed, ntrip):
    """The (T) leg alone (MP2 amplitudes), for profiling the (T) kernels."""
    import dalton_amd as d
    from dalton_amd.synthetic import make_system

    eps, B = make_system(o, v, naux, seed)
    with d.CCContext(eps, B, o) as c:
        c.guess()
        pt = pt_leg(c, ntrip, o, v, _log)
    return {"metric": f"(T) TFLOP/s at o{o}/v{v}, {ntrip} triples, MP2 amplitudes", "unit": "TFLOP/s",
            "workload": f"synthetic o={o} v={v} naux={naux} seed {seed}", "pt": pt}


def ao_workload(nbas=120, rank=1500, seed=6):
    """AO -> density-fitting front end (ccgpu_df_factor_ao) on a synthetic
    positive semidefinite pair supermatrix of known rank (V = X^T X over the
    nbas(nbas+1)/2 AO pairs) and random orthonormal MO coefficients, against
    the oracle's single-column pivoted Cholesky on the host cores."""
    import dalton_amd as d
    from oracle import cholesky as ch

    rng = np.random.default_rng(seed)
    M = nbas * (nbas + 1) // 2
    X = rng.standard_normal((rank, M)) / np.sqrt(rank)
    V = X.T @ X
    il = np.tril_indices(M)
    packed = V[il]
    C = np.linalg.qr(rng.standard_normal((nbas, nbas)))[0]
    cmo = np.ascontiguousarray(C.T).reshape(-1)
    res = {"metric": "AO pivoted Cholesky + MO transform", "unit": "s",
           "workload": f"synthetic PSD pair supermatrix, nbas={nbas} ({M} pairs), rank {rank}, seed {seed}",
           "threshold": 1e-10}
    sync()  # the first call imports torch: outside the timed region
    d.df_factor_ao(np.eye(4).reshape(-1), np.eye(10)[np.tril_indices(10)], 4, 1e-10)  # code objects, allocator
    t0 = time.perf_counter()
    B = d.df_factor_ao(cmo, packed, nbas, 1e-10)  # synchronous: returns with B on the host
    dt = time.perf_counter() - t0
    K = B.shape[0]
    res["gpu"] = {"seconds": round(dt, 4), "naux": K, "ms_per_vector": round(dt / max(K, 1) * 1e3, 4),
                  "note": "host arrays in and out: includes the 0.2 GB upload of the "
                  "packed supermatrix and the download of B"}
    _log(f"ao: GPU {K} vectors in {dt:.3f} s")
    t0 = time.perf_counter()
    L, _, conv = ch.pivoted_cholesky(V, 1e-10)
    dc = time.perf_counter() - t0
    res["cpu_baseline"] = {"value": round(dc, 3), "unit": "s", "kind": "port", "cores": len(os.sched_getaffinity(0)),
                           "sample": "oracle pivoted_cholesky on the same supermatrix (decomposition only)"}
    _log(f"ao: oracle {L.shape[0]} vectors in {dc:.3f} s")
    Bo = np.einsum("ap,kab,bq->kpq", C, _pairs_square(L, nbas), C, optimize=True)
    g = np.einsum("kpq,krs->pqrs", B[:, :24, :24], B[:, :24, :24], optimize=True)
    go = np.einsum("kpq,krs->pqrs", Bo[:, :24, :24], Bo[:, :24, :24], optimize=True)
    res["naux_oracle"] = int(L.shape[0])
    res["max_abs_dev_integrals_24_mos"] = float(np.abs(g - go).max())
    return res


def _pairs_square(L, nbas):
    il = np.tril_indices(nbas)
    S = np.zeros((L.shape[0], nbas, nbas))
    S[:, il[0], il[1]] = L
    S[:, il[1], il[0]] = L
    return S


def pt48_workload(o=48, v=384, naux=1536, seed=5):
    """(T) over every i>=j>=k triple at o=48 v=384 naux=1536, FP64 and mixed,
    after converging CCSD in the same precision.  The integral scale is the
    convergent one of dalton_amd.synthetic.batch_sigma, so the energies are
    finite and |dE_CCSD(T)| between the precisions is meaningful."""
    import dalton_amd as d
    from dalton_amd.synthetic import batch_sigma, make_system

    sigma = batch_sigma(v, naux)
    eps, B = make_system(o, v, naux, seed, sigma=sigma)
    res = {"metric": "(T) over all i>=j>=k at o=48 v=384, FP64 vs mixed", "unit": "s",
           "workload": f"synthetic o={o} v={v} naux={naux} seed {seed} sigma {sigma:.5f}",
           "flops_pt": 2.0 * o**3 * v**4 + 2.0 * o**4 * v**3}
    for label, prec in (("fp64", d.FP64), ("mixed", d.MIXED)):
        with d.CCContext(eps, B, o) as c:
            t0 = time.perf_counter()
            r = c.ccsd_energy(maxit=80, threnr=1e-9, precision=prec)
            sync()
            tc = time.perf_counter() - t0
            _log(f"{label}: CCSD {r['niter']} it in {tc:.1f} s, E = {r['e_ccsd']:.12f}")
            _, _, _, ntr = c.pt_partial(0, 1, 1, precision=prec)  # (T) integral layouts, one triple (untimed)
            sync()
            t0 = time.perf_counter()
            e4 = e5 = 0.0
            nchunk = 8  # the full triple list in contiguous ranges, one progress line each
            for q in range(nchunk):
                a, b = q * ntr // nchunk, (q + 1) * ntr // nchunk
                p4, p5, _, _ = c.pt_partial(a, b - a, 1, precision=prec)
                e4 += p4
                e5 += p5
                _log(f"{label}: (T) triples {a}..{b - 1} of {ntr} done at {time.perf_counter() - t0:.1f} s")
            sync()
            tp = time.perf_counter() - t0
            _log(f"{label}: (T) {tp:.1f} s, E4 = {e4:.12f} E5 = {e5:.12f}")
        res[label] = {"ccsd_seconds": round(tc, 2), "ccsd_iterations": r["niter"], "e_ccsd": r["e_ccsd"],
                      "pt_seconds": round(tp, 3), "pt_tflops": round(res["flops_pt"] / tp / 1e12, 2),
                      "e4": e4, "e5": e5}
    m, f = res["mixed"], res["fp64"]
    res["abs_dE_ccsd"] = abs(m["e_ccsd"] - f["e_ccsd"])
    res["abs_dE_pt"] = abs(m["e4"] + m["e5"] - f["e4"] - f["e5"])
    res["abs_dE_ccsd_t"] = abs(m["e_ccsd"] + m["e4"] + m["e5"] - f["e_ccsd"] - f["e4"] - f["e5"])
    res["pt_speedup_mixed_vs_fp64"] = round(f["pt_seconds"] / m["pt_seconds"], 3)
    return res


def lambda_workload(o=32, v=256, naux=1024, seed=4, thr=1e-8):
    """Ground-state multipliers at o=32 v=256 naux=1024: CCSD converged in
    FP64 (untimed), then the L0 equations lambda A + eta = 0 solved to
    |r| < thr from the FP64 context and from the mixed one (BF16x3 phase,
    then FP64), and the unrelaxed density of each; the integral scale is the
    convergent one of dalton_amd.synthetic.batch_sigma."""
    import dalton_amd as d
    from dalton_amd.cc import profile_start, profile_stop
    from dalton_amd.synthetic import batch_sigma, make_system

    sigma = batch_sigma(v, naux)
    eps, B = make_system(o, v, naux, seed, sigma=sigma)
    res = {"metric": "L0 multiplier solve at o=32 v=256 naux=1024, FP64 vs mixed", "unit": "s",
           "workload": f"synthetic o={o} v={v} naux={naux} seed {seed} sigma {sigma:.5f}, thr {thr:g}",
           "flops_per_transform_nominal": 2.0 * f_ccsd(o, v, naux)}
    rng = np.random.default_rng(seed)
    V = rng.standard_normal((o + v, o + v))
    V = 0.5 * (V + V.T)
    with d.CCContext(eps, B, o) as c:
        del B
        t0 = time.perf_counter()
        r = c.ccsd_energy(maxit=80, threnr=1e-10, precision=d.FP64)
        sync()
        _log(f"CCSD {r['niter']} iterations in {time.perf_counter() - t0:.1f} s")
        res["ccsd_iterations"] = r["niter"]
        c.lhtr(np.zeros(c.nt1), np.zeros(c.nt2), d.FP64)  # library warm-up (untimed)
        out = {}
        for label, prec in (("fp64", d.FP64), ("mixed", d.MIXED)):
            sync()
            profile_start()
            t0 = time.perf_counter()
            l1, l2, it, rn = c.lambda_solve(thr=thr, maxit=100, maxred=20, precision=prec)
            sync()
            dt = time.perf_counter() - t0
            prof = profile_stop()
            n64, nmx = c.lambda_stats()
            D = c.onedens(l1, l2)
            gem = sum(k_["flops"] for k_ in prof.get("kernels", {}).values())
            gms = sum(k_["ms"] for k_ in prof.get("kernels", {}).values())
            res[label] = {"seconds": round(dt, 3), "transforms": it, "transforms_fp64": n64,
                          "transforms_bf16x3": nmx, "s_per_transform": round(dt / it, 4),
                          "resnorm": rn, "gemm_tflops": round(gem / max(gms, 1e-9) / 1e9, 2),
                          "property_V": float(np.sum(D * V)), "kernels": kern_table_of(prof)}
            out[label] = (l1, l2)
            _log(f"lambda {label}: {it} transforms ({n64} FP64, {nmx} BF16x3) in {dt:.2f} s, |r| = {rn:.2e}")
    a1, a2 = out["fp64"]
    b1, b2 = out["mixed"]
    res["max_abs_dev_lambda_mixed_vs_fp64"] = float(max(np.abs(a1 - b1).max(), np.abs(a2 - b2).max()))
    res["abs_dev_property_mixed_vs_fp64"] = abs(res["mixed"]["property_V"] - res["fp64"]["property_V"])
    res["speedup_mixed_vs_fp64"] = round(res["fp64"]["seconds"] / res["mixed"]["seconds"], 3)
    return res


def restart_workload(o=64, v=512, naux=2048, seed=3):
    """Dalton 'R0' amplitude file at o=64 v=512: MP2 amplitudes of the
    synthetic system written with ccgpu_save_amplitudes (device -> host ->
    Fortran records; the 4.3 GB doubles record is split into gfortran
    subrecords) and read back with ccgpu_restart_amplitudes, on local disk."""
    import os
    import shutil
    import struct
    import tempfile

    import dalton_amd as d
    from dalton_amd.synthetic import make_system

    eps, B = make_system(o, v, naux, seed)
    res = {"metric": "R0 file save / restart at o=64 v=512", "unit": "s",
           "workload": f"synthetic o={o} v={v} naux={naux} seed {seed}, MP2 amplitudes"}
    tmp = tempfile.mkdtemp(prefix="ccgpu_rsp_")
    try:
        with d.CCContext(eps, B, o) as c:
            del B
            e_mp2 = c.guess()
            path = os.path.join(tmp, d.rsp_filename("R0"))
            sync()
            t0 = time.perf_counter()
            c.save_amplitudes(path)
            ts = time.perf_counter() - t0
            size = os.path.getsize(path)
            _log(f"save: {size / 1e9:.2f} GB in {ts:.2f} s")
            with open(path, "rb") as fh:  # count the subrecords of the doubles record (the fourth)
                recs = []
                while len(recs) < 4:
                    nsub = 0
                    while True:
                        (lead,) = struct.unpack("<i", fh.read(4))
                        fh.seek(abs(lead), 1)
                        fh.read(4)
                        nsub += 1
                        if lead >= 0:
                            break
                    recs.append(nsub)
            t0 = time.perf_counter()
            e = c.restart_amplitudes(path)
            tr = time.perf_counter() - t0
            _log(f"restart: {tr:.2f} s, E = {e:.12f} (saved {e_mp2:.12f})")
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    res.update({"file_bytes": size, "doubles_subrecords": recs[3], "save_seconds": round(ts, 3),
                "save_GBps": round(size / ts / 1e9, 3), "restart_seconds": round(tr, 3),
                "restart_GBps": round(size / tr / 1e9, 3), "energy_saved": e_mp2, "energy_restarted": e,
                "abs_dE": abs(e - e_mp2)})
    return res


def batch_workload(nsys=32, seed=5):
    """32 uneven systems (o ~ U{5..30}, v = 8..12 o, naux = 4v, convergent
    integral scale) converged together: FP64, one HIP graph replay per
    iteration after an autotuning pass; then the same work issued eagerly
    (no graph) for comparison.  Energies of both runs must agree."""
    import dalton_amd as d
    from dalton_amd.synthetic import batch_shapes, make_batch

    shapes = batch_shapes(nsys, seed)
    t0 = time.perf_counter()
    systems = make_batch(shapes, seed)
    _log(f"{nsys} systems generated in {time.perf_counter() - t0:.1f} s")
    out = {"metric": "32 uneven CCSD systems converged together (FP64)", "unit": "s",
           "workload": f"batch_shapes({nsys}, seed {seed}): " + " ".join(f"{o}/{v}/{n}" for o, v, n in shapes),
           "flops_per_round_all_systems": sum(f_ccsd(o, v, n) for o, v, n in shapes)}
    with d.CCBatch(systems) as b:
        runs = {}

        def solve(label, graph):
            sync()
            t0 = time.perf_counter()
            r = b.energy(maxit=100, threnr=1e-8, graph=graph, tune=False)
            sync()
            dt = time.perf_counter() - t0
            runs[label] = r
            out[label] = {"seconds": round(dt, 3), "rounds": int(r["niter"].max()), **b.stats()}
            _log(f"batch {label}: {dt:.2f} s, rounds {int(r['niter'].max())}")

        solve("graph_default_tiles", True)
        sync()
        t0 = time.perf_counter()
        b.tune()
        out["tune_seconds"] = round(time.perf_counter() - t0, 3)
        _log(f"batch autotuning pass: {out['tune_seconds']:.2f} s")
        solve("graph", True)
        out["speedup_tuned_vs_default"] = round(out["graph_default_tiles"]["seconds"] / out["graph"]["seconds"], 3)
        # algorithmic work actually done: each system's iterations
        done = sum(f_ccsd(o, v, n) * int(k) for (o, v, n), k in zip(shapes, runs["graph"]["niter"]))
        out["tflops_graph"] = round(done / out["graph"]["seconds"] / 1e12, 2)
    out["niter"] = [int(x) for x in runs["graph"]["niter"]]
    out["e_ccsd"] = [float(x) for x in runs["graph"]["e_ccsd"]]
    out["max_abs_dE_tuned_vs_default"] = float(np.abs(runs["graph"]["e_ccsd"]
                                                      - runs["graph_default_tiles"]["e_ccsd"]).max())
    # the same systems solved one after another, each on its own (FP64)
    sync()
    t0 = time.perf_counter()
    ind_e, ind_it = [], []
    for eps, B, o in systems:
        with d.CCContext(eps, B, o) as c:
            r = c.ccsd_energy(maxit=100, threnr=1e-8)
        ind_e.append(r["e_ccsd"])
        ind_it.append(r["niter"])
    sync()
    out["individual"] = {"seconds": round(time.perf_counter() - t0, 3)}
    _log(f"individual solves: {out['individual']['seconds']:.2f} s")
    out["max_abs_dE_batch_vs_individual"] = float(np.abs(np.array(ind_e) - runs["graph"]["e_ccsd"]).max())
    out["same_iterations_as_individual"] = bool(list(ind_it) == [int(x) for x in runs["graph"]["niter"]])
    out["speedup_batch_vs_individual"] = round(out["individual"]["seconds"] / out["graph"]["seconds"], 3)
    return out


def copy_workload(nbytes=2 << 30, reps=3):
    """A plain device copy of known size (torch's copy kernel): the
    calibration run for the FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    n = nbytes // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    return {"workload": "device copy", "bytes_read": nbytes, "bytes_written": nbytes, "reps": reps,
            "GBps_read_plus_write": round(2 * nbytes / dt / 1e9, 1)}


def blas_workload(n=8192, reps=5):
    """The vendor BLAS (torch.matmul -> hipBLASLt / rocBLAS) on square
    products, as the attainable-rate calibration for this card: bf16 and fp64
    dense GEMM TFLOP/s."""
    import torch

    out = {"workload": f"torch.matmul {n}^3", "reps": reps}
    for name, dt in (("bf16", torch.bfloat16), ("fp32", torch.float32), ("fp64", torch.float64)):
        a = torch.randn(n, n, device="cuda", dtype=torch.float32).to(dt)
        b = torch.randn(n, n, device="cuda", dtype=torch.float32).to(dt)
        c = a @ b
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            c = a @ b
        torch.cuda.synchronize()
        dt_ = (time.perf_counter() - t0) / reps
        out[name + "_tflops"] = round(2.0 * n**3 / dt_ / 1e12, 1)
        del a, b, c
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed mixed-precision iterations")
    ap.add_argument("--fp64-steps", type=int, default=2,
                    help="timed FP64 iterations of the headline system (the convergent leg times every FP64 "
                         "iteration of a whole solve at the same shape)")
    ap.add_argument("--no-accuracy", dest="accuracy", action="store_false",
                    help="skip the convergent-scale leg (FP64 and mixed solves at the same shape)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--o", type=int, default=64)
    ap.add_argument("--v", type=int, default=512)
    ap.add_argument("--naux", type=int, default=2048)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--pt-triples", type=int, default=256,
                    help="(T) triples timed, spread evenly over the i>=j>=k list (0: skip)")
    ap.add_argument("--lhtr-vectors", type=int, default=16,
                    help="left-transformation trial vectors timed at o=32 v=256 naux=1024 (0: skip)")
    ap.add_argument("--workload", choices=["ccsd", "pt", "pt48", "batch", "ao", "lambda", "restart", "copy", "blas",
                                           "accuracy", "cpu-baseline"], default="ccsd",
                    help="ccsd: the headline line (default); pt: the (T) leg alone at --o/--v; "
                         "ao: the AO Cholesky front end; lambda: the L0 multiplier solve at "
                         "o=32 v=256 in both precisions; restart: the R0 amplitude file at o=64 v=512; pt48: full (T) at o=48 v=384 in both "
                         "precisions on a convergent system; batch: 32 uneven systems converged in "
                         "one graph replay per iteration; accuracy: the convergent-scale FP64 and mixed "
                         "solves at --o/--v/--naux alone")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: runs on one GPU only; --gpus must be 1 (the CC path is not sharded "
              "across GPUs)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "cpu-baseline":
        cpu_baseline_child()
        return
    if args.workload == "pt":
        print(json.dumps(pt_workload(args.o, args.v, args.naux, args.seed, args.pt_triples)), flush=True)
        return
    if args.workload == "ao":
        print(json.dumps(ao_workload()), flush=True)
        return
    if args.workload == "lambda":
        print(json.dumps(lambda_workload()), flush=True)
        return
    if args.workload == "restart":
        print(json.dumps(restart_workload()), flush=True)
        return
    if args.workload == "pt48":
        print(json.dumps(pt48_workload()), flush=True)
        return
    if args.workload == "batch":
        print(json.dumps(batch_workload()), flush=True)
        return
    if args.workload == "accuracy":
        print(json.dumps(convergent_leg(args.o, args.v, args.naux, args.seed, args.pt_triples, _log)), flush=True)
        return
    if args.workload == "copy":
        print(json.dumps(copy_workload()), flush=True)
        return
    if args.workload == "blas":
        print(json.dumps(blas_workload()), flush=True)
        return

    # the CPU baseline's process is started before this one touches the GPU
    cpu = start_cpu_baseline() if not args.no_cpu_baseline else None

    import dalton_amd as d
    from dalton_amd.cc import profile_start, profile_stop
    from dalton_amd.synthetic import make_system

    log = _log
    o, v, naux = args.o, args.v, args.naux
    t0 = time.perf_counter()
    eps, B = make_system(o, v, naux, args.seed)
    log(f"system o={o} v={v} naux={naux} generated in {time.perf_counter() - t0:.1f} s")
    ctx = d.CCContext(eps, B, o)
    del B
    e_mp2 = ctx.guess()
    log(f"E_MP2 = {e_mp2:.10f}")
    period = ctx.mxdiis

    def steps(prec, nsteps, label, timed):
        """nsteps iterations from the MP2 start.  The integrals of this
        distribution are far outside the convergent regime, and the iteration
        overflows after its first DIIS restart, so the MP2 start is re-seeded
        (outside the timer) at the start of every DIIS cycle of MXDIIS steps;
        every step is checked to stay finite.  Returns s/step, energies,
        residual norms and the library's event profile."""
        hist, norms, dt = [], [], 0.0
        if timed:
            profile_start()
        for s_ in range(nsteps):
            if s_ % period == 0:
                ctx.guess()  # MP2 amplitudes, empty DIIS history (untimed)
            sync()
            t0 = time.perf_counter()
            e, n = ctx.iterate(prec)  # returns after the step (its energy is read back)
            dt += time.perf_counter() - t0
            if not (np.isfinite(e) and np.isfinite(n)):
                raise RuntimeError(f"{label} step {s_ + 1}: non-finite energy or residual norm ({e}, {n})")
            hist.append(e)
            norms.append(n)
            log(f"{label} step {s_ + 1}: E = {e:.10f} |Omega| = {n:.3e}")
        return dt / max(nsteps, 1), hist, norms, (profile_stop() if timed else None)

    for w in range(args.warmup):  # allocates the iteration workspace
        e, n = ctx.iterate(d.MIXED)
        log(f"warmup {w + 1} (mixed): E = {e:.10f} |Omega| = {n:.3e}")
    if args.steps < 1:  # profiling runs of the FP64 leg alone
        spi_d, hist_d, _, prof_d = steps(d.FP64, max(1, args.fp64_steps), "fp64", True)
        print(json.dumps({"fp64_s_per_iter": spi_d, "kernels": kern_table_of(prof_d)}), flush=True)
        ctx.close()
        if cpu is not None:
            cpu.kill()
        return
    spi_m, hist_m, norm_m, prof_m = steps(d.MIXED, args.steps, "mixed", True)
    fp64_leg = args.fp64_steps > 0
    if fp64_leg:
        spi_d, hist_d, norm_d, prof_d = steps(d.FP64, args.fp64_steps, "fp64", True)

    # the matrix cores' own sustained rate on random operands (no memory
    # traffic), measured now on this card: the ceiling beside the data-sheet peak
    mfma_only = {"bf16": d.mfma_ceiling(d.MIXED), "fp64": d.mfma_ceiling(d.FP64)}
    log(f"MFMA-only rates on random operands: bf16 {mfma_only['bf16']:.1f}, fp64 {mfma_only['fp64']:.1f} TFLOP/s")

    def roofline(prof, dt_total, peak, peak_note, attainable=None, ceiling=None):
        kern = prof.get("kernels", {})
        if not kern:  # dry run: nothing was launched
            return None
        dom = max(kern, key=lambda k: kern[k]["ms"])
        kd = kern[dom]
        ach = kd["flops"] / (kd["ms"] * 1e-3) / 1e12
        tb = TRAFFIC_BYTES.get(dom) if (o, v, naux) == (64, 512, 2048) else None
        return {"bound": "mfma", "achieved": round(ach, 3), "peak": peak, "unit": "TFLOP/s",
                "frac": round(ach / peak, 4), "traffic": tb, "traffic_unit": "bytes per launch",
                "kernel": dom, "launches": kd["count"],
                "avg_launch_ms": round(kd["ms"] / kd["count"], 4),
                "share_of_step": round(kd["ms"] * 1e-3 / dt_total, 4), "peak_note": peak_note,
                "attainable": attainable,
                "frac_attainable": round(ach / attainable, 4) if attainable else None,
                "mfma_only": round(ceiling, 1) if ceiling else None,
                "frac_mfma_only": round(ach / ceiling, 4) if ceiling else None}

    F = f_ccsd(o, v, naux)
    if cpu is not None:  # headline timing done: the CPU baseline may run beside the remaining GPU legs
        cpu.stdin.write("go\n")
        cpu.stdin.flush()
    pt = pt_leg(ctx, args.pt_triples, o, v, log) if args.pt_triples > 0 else {}
    ctx.close()
    del ctx

    conv = convergent_leg(o, v, naux, args.seed, args.pt_triples, log) if args.accuracy else None
    lh = lhtr_leg(args.lhtr_vectors, log, kern_table_of) if args.lhtr_vectors > 0 else None

    fp64 = None
    if fp64_leg:
        fp64 = {"s_per_iter": round(spi_d, 4), "steps": args.fp64_steps, "ccsd_tflops": round(F / spi_d / 1e12, 3),
                "energies": hist_d, "omega_norms": norm_d,
                "roofline": roofline(prof_d, spi_d * args.fp64_steps, PEAK["fp64"], "FP64 matrix",
                                     ATTAINABLE["fp64"], mfma_only["fp64"]),
                "phases_ms_per_iter": {k: round(v_ / args.fp64_steps, 2) for k, v_ in prof_d.get("phases", {}).items()},
                "kernels": kern_table_of(prof_d)}
    out = {
        "metric": "CCSD s/iter + (T) TFLOP/s at o64/v512: mixed vs FP64 MFMA, |dE|<=1uEh",
        "value": round(spi_m, 4),
        "unit": "s/iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(spi_m * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3/f32-accumulate (bulk contractions); f64 update, DIIS, energy",
        "data": "synthetic",
        "config": {"workload": f"synthetic DF-CCSD iterations from the MP2 start (re-seeded every "
                               f"{period} steps), o={o} v={v} naux={naux}, B ~ N(0, 0.05^2), seed {args.seed}",
                   "o": o, "v": v, "naux": naux, "precision": "mixed"},
        "ccsd_tflops": round(F / spi_m / 1e12, 3),
        "ccsd_flops_per_iter": F,
        "fp64": fp64,
        "speedup_mixed_vs_fp64": round(spi_d / spi_m, 3) if fp64_leg else None,
        "energies": hist_m,
        "omega_norms": norm_m,
        "rel_dE_per_step_mixed_vs_fp64": [abs(a - b) / max(abs(b), 1e-300) for a, b in zip(hist_m, hist_d)]
        if fp64_leg else None,
        "roofline": roofline(prof_m, spi_m * args.steps, round(PEAK["bf16"] / 3, 1),
                             "BF16 dense matrix peak / 3 (three bf16 products per FP32-equivalent product)",
                             round(ATTAINABLE["bf16x3"], 1), mfma_only["bf16"] / 3),
        "phases_ms_per_iter": {k: round(v_ / args.steps, 2) for k, v_ in prof_m.get("phases", {}).items()},
        "kernels": kern_table_of(prof_m),
        "pt": pt or None,
        "convergent": conv,
        "lhtr": lh,
    }
    if cpu is not None:
        out["cpu_baseline"] = finish_cpu_baseline(cpu, log)
        out["cpu_baseline"]["note"] = ("timed in a child process on the host cores minus two, while this "
                                       "process drove the later GPU legs")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
