"""Batched uneven systems (one HIP graph replay per iteration) against their
individual solves."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _systems(n=6, seed=21, omin=3, omax=8):
    from dalton_amd.synthetic import batch_shapes, make_batch

    shapes = batch_shapes(n, seed, omin=omin, omax=omax)
    return shapes, make_batch(shapes, seed)


def _individual(systems, precision=0, threnr=1e-10):
    import dalton_amd as d

    out = []
    for eps, B, o in systems:
        with d.CCContext(eps, B, o) as c:
            r = c.ccsd_energy(maxit=80, threnr=threnr, precision=precision)
            t1, t2 = c.amplitudes()
        out.append((r, t1, t2))
    return out


@pytest.mark.parametrize("graph", [True, False])
def test_batch_fp64_matches_individual_solves(graph):
    import dalton_amd as d

    shapes, systems = _systems()
    print(f"batch fp64 graph={graph}: shapes {shapes}", flush=True)
    ref = _individual(systems)
    with d.CCBatch(systems, ladder_budget_bytes=2e6) as b:
        res = b.energy(maxit=80, threnr=1e-10, graph=graph)
        amps = [b.amplitudes(k) for k in range(len(systems))]
        st = b.stats()
    print(f"  stats {st}, niter {list(res['niter'])}", flush=True)
    for k, (r, t1, t2) in enumerate(ref):
        assert abs(res["e_mp2"][k] - r["e_mp2"]) < 1e-12
        assert res["niter"][k] == r["niter"], (k, res["niter"][k], r["niter"])
        assert abs(res["e_ccsd"][k] - r["e_ccsd"]) < 1e-10, (k, res["e_ccsd"][k], r["e_ccsd"])
        assert np.abs(amps[k][0] - t1).max() < 1e-9 and np.abs(amps[k][1] - t2).max() < 1e-9
    if graph:  # the first round runs eagerly (it sizes the scratch), the rest are replays
        assert st["replays"] == int(max(res["niter"])) - 1 and st["captures"] >= 1
    else:
        assert st["replays"] == 0


def test_batch_mixed_within_1uEh_of_fp64():
    import dalton_amd as d

    shapes, systems = _systems(n=5, seed=33)
    print(f"batch mixed: shapes {shapes}", flush=True)
    ref = _individual(systems)
    with d.CCBatch(systems) as b:
        res = b.energy(maxit=80, threnr=1e-10, precision=d.MIXED)
    for k, (r, _, _) in enumerate(ref):
        assert abs(res["e_ccsd"][k] - r["e_ccsd"]) < 1e-6, (k, res["e_ccsd"][k], r["e_ccsd"])


def test_single_context_unchanged_after_autotuning():
    """The autotuner's cached choices are used by every later GEMM of the
    same shape; a single-system solve afterwards still matches its earlier
    result (tile choice only changes summation order)."""
    import dalton_amd as d

    _, systems = _systems(n=2, seed=44)
    ref = _individual(systems[:1])
    with d.CCBatch(systems) as b:
        b.energy(maxit=80, threnr=1e-10)
    again = _individual(systems[:1])
    assert abs(ref[0][0]["e_ccsd"] - again[0][0]["e_ccsd"]) < 1e-11


def test_small_batch_matches_oracle_solves():
    """Six uneven systems (o 3..5, v = 8..10 o, naux = 4v) converged together
    in one graph replay per round, against the oracle's CCSD_ENERGY loop
    (oracle.ccsd.solve: same iteration count, energies to 1e-10 Eh), so the
    batch path is pinned to the oracle and not only to single HIP solves."""
    import dalton_amd as d
    from dalton_amd.synthetic import batch_shapes, make_batch
    from oracle import ccsd as occ

    shapes = batch_shapes(6, 77, omin=3, omax=5, kmin=8, kmax=10)
    systems = make_batch(shapes, 77)
    print(f"small batch vs oracle: shapes {shapes}", flush=True)
    with d.CCBatch(systems) as b:
        res = b.energy(maxit=80, threnr=1e-10, graph=True)
    for k, (eps, B, o) in enumerate(systems):
        ref = occ.solve(occ.DFSystem(eps, B, o), maxiter=80, threnr=1e-10)
        print(f"  system {k}: oracle {len(ref['history'])} it E = {ref['e_ccsd']:.12f}, "
              f"batch {res['niter'][k]} it E = {res['e_ccsd'][k]:.12f}", flush=True)
        assert res["niter"][k] == len(ref["history"])
        assert abs(res["e_mp2"][k] - ref["e_mp2"]) < 1e-12
        assert abs(res["e_ccsd"][k] - ref["e_ccsd"]) < 1e-10


def test_c5b_batch_as_configured_matches_individual_fp64_solves():
    """C5b at its configured shape: 32 systems with o ~ U{5..30},
    v = 8..12 o and naux = 4v (batch_shapes(32, seed 5), the convergent
    integral scale batch_sigma), converged together in FP64 with one HIP
    graph replay per round, against each system solved on its own:
    the same iteration count and the energy to 1e-10 Eh, per system."""
    import time

    import dalton_amd as d
    from dalton_amd.synthetic import batch_shapes, make_batch

    shapes = batch_shapes(32, 5)
    t0 = time.perf_counter()
    systems = make_batch(shapes, 5)
    print(f"C5b: 32 systems generated in {time.perf_counter() - t0:.1f} s: {shapes}", flush=True)
    ind = []
    for k, (eps, B, o) in enumerate(systems):
        with d.CCContext(eps, B, o) as c:
            r = c.ccsd_energy(maxit=100, threnr=1e-8)
        ind.append((r["e_ccsd"], r["niter"]))
        print(f"  individual {k} o={o} v={eps.size - o}: {r['niter']} it, E = {r['e_ccsd']:.12f}", flush=True)
    t0 = time.perf_counter()
    with d.CCBatch(systems) as b:
        res = b.energy(maxit=100, threnr=1e-8, graph=True, tune=False)
        st = b.stats()
    print(f"  batch: {time.perf_counter() - t0:.1f} s, {st}", flush=True)
    for k, (e, n) in enumerate(ind):
        assert res["niter"][k] == n, (k, res["niter"][k], n)
        assert abs(res["e_ccsd"][k] - e) < 1e-10, (k, res["e_ccsd"][k], e)
    assert st["replays"] >= 1
