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
