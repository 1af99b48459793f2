"""HIP left transformation (CC_LHTR) against the autograd oracle and against
central finite differences of the CC Lagrangian lambda . Omega(t) formed with
the HIP vector function (ccgpu_rhs)."""
import numpy as np
import pytest

from conftest import load_molecule
from oracle import ccsd as occ
from oracle import lhtr as olh

pytestmark = pytest.mark.gpu


def _system(name):
    if name.startswith("synthetic"):
        from dalton_amd.synthetic import make_system

        o, v, naux = (int(x[1:]) for x in name.split("_")[1:])
        eps, B = make_system(o, v, naux, seed=11)
        return eps, B, o
    eps, B, o, _ = load_molecule(name)
    return eps, B, o


def _vectors(nt1, nt2, rng, scale_t=0.03):
    t1 = scale_t * rng.standard_normal(nt1)
    t2 = scale_t * rng.standard_normal(nt2)
    l1 = rng.standard_normal(nt1)
    l2 = rng.standard_normal(nt2)
    return t1, t2, l1, l2


def _lagrangian(c, t1, t2, l1, l2, prec=0):
    om1, om2 = c.rhs(t1, t2, prec)
    return float(np.dot(om1, l1) + np.dot(om2, l2))


@pytest.mark.parametrize("name", ["h2o_ccpvdz_test", "synthetic_o4_v23_n60"])
def test_lhtr_matches_oracle(name):
    import dalton_amd as d

    eps, B, o = _system(name)
    print(f"lhtr vs oracle {name}", flush=True)
    v = eps.size - o
    sysm = occ.DFSystem(eps, B, o)
    rng = np.random.default_rng(5)
    with d.CCContext(eps, B, o) as c:
        t1, t2, l1, l2 = _vectors(c.nt1, c.nt2, rng)
        c.set_amplitudes(t1, t2)
        r1, r2 = c.lhtr(l1, l2, convention="natural")
        q1, q2 = olh.left_transform(sysm, t1, t2, l1, l2)
    s = max(1.0, np.abs(q1).max(), np.abs(q2).max())
    assert np.abs(r1 - q1).max() < 1e-12 * s, np.abs(r1 - q1).max()
    assert np.abs(r2 - q2).max() < 1e-12 * s, np.abs(r2 - q2).max()
    assert v > 0


def test_lhtr_at_converged_ccsd_matches_oracle():
    """At Dalton's converged H2O/cc-pVDZ ground state (the state CC_LHTR is
    called at), several trial vectors in one call."""
    import dalton_amd as d

    eps, B, o, _ = load_molecule("h2o_ccpvdz_test")
    print("lhtr at converged H2O", flush=True)
    sysm = occ.DFSystem(eps, B, o)
    rng = np.random.default_rng(6)
    with d.CCContext(eps, B, o) as c:
        c.ccsd_energy(maxit=60, threnr=1e-12)
        t1, t2 = c.amplitudes()
        L1 = rng.standard_normal((3, c.nt1))
        L2 = rng.standard_normal((3, c.nt2))
        R1, R2 = c.lhtr(L1, L2, convention="natural")
    for k in range(3):
        q1, q2 = olh.left_transform(sysm, t1, t2, L1[k], L2[k])
        s = max(1.0, np.abs(q2).max())
        assert np.abs(R1[k] - q1).max() < 1e-12 * s
        assert np.abs(R2[k] - q2).max() < 1e-12 * s


@pytest.mark.parametrize("name", ["h2o_ccpvdz_test", "synthetic_o6_v40_n90"])
def test_lhtr_matches_finite_differences(name):
    """32 random directions, central differences with h = 1e-5 (HIP rhs)."""
    import dalton_amd as d

    eps, B, o = _system(name)
    print(f"lhtr vs finite differences {name}", flush=True)
    rng = np.random.default_rng(7)
    with d.CCContext(eps, B, o) as c:
        t1, t2, l1, l2 = _vectors(c.nt1, c.nt2, rng)
        c.set_amplitudes(t1, t2)
        r1, r2 = c.lhtr(l1, l2, convention="natural")
        h = 1e-5
        worst = 0.0
        for _ in range(32):
            e1 = rng.standard_normal(c.nt1)
            e2 = rng.standard_normal(c.nt2)
            nrm = np.sqrt(np.dot(e1, e1) + np.dot(e2, e2))
            e1 /= nrm
            e2 /= nrm
            fd = (_lagrangian(c, t1 + h * e1, t2 + h * e2, l1, l2)
                  - _lagrangian(c, t1 - h * e1, t2 - h * e2, l1, l2)) / (2 * h)
            an = float(np.dot(r1, e1) + np.dot(r2, e2))
            worst = max(worst, abs(fd - an))
    print(f"  max |FD - rho.d| = {worst:.3e}", flush=True)
    assert worst < 1e-7


def _c4_check(sigma, label, strict, ndir=16):
    """o=32, v=256, naux=1024 at the MP2 amplitudes (plus small singles): 16
    unit trial vectors in one call, checked along ndir random unit directions
    (direction d with vector d mod 16) by central differences of
    lambda . Omega with h = 1e-5; the batched call equals one-vector calls.
    The particle ladder runs in several a-batches at this size.  The FD
    round-off floor is estimated from the spread of the differences at h, 2h
    and h/2: an adjoint error would show in all of them alike.  The FP32-MFMA
    transformation of the same vectors is compared with the FP64 one."""
    import dalton_amd as d
    from dalton_amd.synthetic import make_system

    o, v, naux = 32, 256, 1024
    eps, B = make_system(o, v, naux, seed=2, sigma=sigma)
    print(f"lhtr o32 v256 naux1024, 16 vectors, {label}", flush=True)
    rng = np.random.default_rng(8)
    with d.CCContext(eps, B, o) as c:
        c.guess()
        t1, t2 = c.amplitudes()
        t1 = 0.01 * rng.standard_normal(c.nt1)
        nv = 16
        L1 = rng.standard_normal((nv, c.nt1))
        L2 = rng.standard_normal((nv, c.nt2))
        nrm = np.sqrt((L1 * L1).sum(1) + (L2 * L2).sum(1))
        L1 /= nrm[:, None]
        L2 /= nrm[:, None]
        c.set_amplitudes(t1, t2)
        R1, R2 = c.lhtr(L1, L2, convention="natural")
        s1, s2 = c.lhtr(L1[5], L2[5], convention="natural")
        assert np.abs(s1 - R1[5]).max() == 0.0 and np.abs(s2 - R2[5]).max() == 0.0
        F1, F2 = c.lhtr(L1, L2, d.FP32, convention="natural")
        dev32 = max(np.abs(F1 - R1).max(), np.abs(F2 - R2).max()) / max(np.abs(R1).max(), np.abs(R2).max())
        print(f"  FP32-MFMA vs FP64 left transformation: max rel deviation {dev32:.2e}", flush=True)
        assert dev32 < 1e-4
        h = 1e-5
        worst = 0.0
        for dd in range(ndir):
            k = dd % nv
            e1 = rng.standard_normal(c.nt1)
            e2 = rng.standard_normal(c.nt2)
            en = np.sqrt(np.dot(e1, e1) + np.dot(e2, e2))
            e1 /= en
            e2 /= en

            def fd(hh):
                return (_lagrangian(c, t1 + hh * e1, t2 + hh * e2, L1[k], L2[k])
                        - _lagrangian(c, t1 - hh * e1, t2 - hh * e2, L1[k], L2[k])) / (2 * hh)

            f1 = fd(h)
            an = float(np.dot(R1[k], e1) + np.dot(R2[k], e2))
            scale = max(1.0, abs(an))
            rel = abs(f1 - an) / scale
            if strict:
                bound = 1e-7
            else:  # FD noise floor from the spread of three step sizes
                f2, f3 = fd(2 * h), fd(0.5 * h)
                bound = max(1e-7, 4.0 * max(abs(f1 - f2), abs(f1 - f3)) / scale)
            print(f"  direction {dd} (vector {k}): rho.d = {an:.12e} FD = {f1:.12e} rel = {rel:.2e} "
                  f"bound = {bound:.1e}", flush=True)
            assert rel < bound, (k, rel, bound)
            worst = max(worst, rel)
        print(f"  max rel |FD - rho.d| = {worst:.2e}", flush=True)


def test_lhtr_c4_shape_sixteen_vectors():
    """Integral distribution of the o32/v256 workload (entries N(0, 0.05^2)):
    within 1e-7 or within the measured FD round-off floor."""
    _c4_check(0.05, "sigma 0.05", strict=False)


def test_lhtr_c4_shape_convergent_scale_strict():
    """The same shape with physically sized integrals (batch_sigma): the FD
    round-off floor is low there, so 1e-7 relative is required outright."""
    from dalton_amd.synthetic import batch_sigma

    _c4_check(batch_sigma(256, 1024), "convergent scale, 32 directions", strict=True, ndir=32)


@pytest.mark.parametrize("budget_mb", ["2", "0.05"])
def test_lhtr_and_rhs_with_batched_ladder(budget_mb, monkeypatch):
    """The particle ladder split into several a-batches (small memory budget):
    rhs and the left transformation against the oracle."""
    import dalton_amd as d

    monkeypatch.setenv("CCGPU_LADDER_BUDGET_MB", budget_mb)
    eps, B, o = _system("synthetic_o6_v40_n90")
    print(f"batched ladder, budget {budget_mb} MB", flush=True)
    sysm = occ.DFSystem(eps, B, o)
    rng = np.random.default_rng(12)
    with d.CCContext(eps, B, o) as c:
        t1, t2, l1, l2 = _vectors(c.nt1, c.nt2, rng)
        om1, om2 = c.rhs(t1, t2)
        c.set_amplitudes(t1, t2)
        r1, r2 = c.lhtr(l1, l2, convention="natural")
    v = eps.size - o
    p1, p2 = occ.residual(sysm, occ.unpack_t1(t1, o, v), occ.unpack_t2(t2, o, v))
    p1, p2 = occ.pack_t1(p1), occ.pack_t2(p2)
    s = max(1.0, np.abs(p2).max())
    assert np.abs(om1 - p1).max() < 1e-12 * s and np.abs(om2 - p2).max() < 1e-12 * s
    q1, q2 = olh.left_transform(sysm, t1, t2, l1, l2)
    s = max(1.0, np.abs(q2).max())
    assert np.abs(r1 - q1).max() < 1e-12 * s and np.abs(r2 - q2).max() < 1e-12 * s


def test_lhtr_mixed_close_to_fp64():
    import dalton_amd as d

    eps, B, o = _system("synthetic_o6_v40_n90")
    print("lhtr mixed vs fp64", flush=True)
    rng = np.random.default_rng(9)
    with d.CCContext(eps, B, o) as c:
        t1, t2, l1, l2 = _vectors(c.nt1, c.nt2, rng)
        c.set_amplitudes(t1, t2)
        a1, a2 = c.lhtr(l1, l2, d.FP64)
        b1, b2 = c.lhtr(l1, l2, d.MIXED)
    s = max(np.abs(a1).max(), np.abs(a2).max())
    err = max(np.abs(a1 - b1).max(), np.abs(a2 - b2).max()) / s
    print(f"  max rel deviation mixed vs fp64 = {err:.3e}", flush=True)
    assert err < 1e-4


def test_autograd_function_backward_is_lhtr():
    import torch

    import dalton_amd as d
    from dalton_amd.autograd import ccsd_residual

    eps, B, o = _system("synthetic_o4_v23_n60")
    print("autograd backward", flush=True)
    sysm = occ.DFSystem(eps, B, o)
    rng = np.random.default_rng(10)
    with d.CCContext(eps, B, o) as c:
        t1, t2, l1, l2 = _vectors(c.nt1, c.nt2, rng)
        x1 = torch.tensor(t1, requires_grad=True)
        x2 = torch.tensor(t2, requires_grad=True)
        om1, om2 = ccsd_residual(c, x1, x2)
        (om1 @ torch.tensor(l1) + om2 @ torch.tensor(l2)).backward()
    q1, q2 = olh.left_transform(sysm, t1, t2, l1, l2)
    s = max(1.0, np.abs(q2).max())
    assert np.abs(x1.grad.numpy() - q1).max() < 1e-12 * s
    assert np.abs(x2.grad.numpy() - q2).max() < 1e-12 * s


def test_autograd_function_on_device_tensors():
    """The autograd Function with amplitudes on the GPU (torch device
    tensors, whose addresses go through the C ABI): forward and backward stay
    on the device and give the same bits as the host-array path, and the
    backward equals the oracle's left transformation."""
    import torch

    import dalton_amd as d
    from dalton_amd.autograd import ccsd_residual

    eps, B, o = _system("synthetic_o4_v23_n60")
    print("autograd backward on device tensors", flush=True)
    sysm = occ.DFSystem(eps, B, o)
    rng = np.random.default_rng(12)
    with d.CCContext(eps, B, o) as c:
        t1, t2, l1, l2 = _vectors(c.nt1, c.nt2, rng)
        h1 = torch.tensor(t1, requires_grad=True)
        h2 = torch.tensor(t2, requires_grad=True)
        hom1, hom2 = ccsd_residual(c, h1, h2)
        (hom1 @ torch.tensor(l1) + hom2 @ torch.tensor(l2)).backward()
        x1 = torch.tensor(t1, device="cuda", requires_grad=True)
        x2 = torch.tensor(t2, device="cuda", requires_grad=True)
        om1, om2 = ccsd_residual(c, x1, x2)
        assert om1.is_cuda and om2.is_cuda
        (om1 @ torch.tensor(l1, device="cuda") + om2 @ torch.tensor(l2, device="cuda")).backward()
        assert x1.grad.is_cuda and x2.grad.is_cuda
        g1, g2 = x1.grad.cpu().numpy(), x2.grad.cpu().numpy()
        assert np.array_equal(om1.detach().cpu().numpy(), hom1.detach().numpy())
        assert np.array_equal(om2.detach().cpu().numpy(), hom2.detach().numpy())
        assert np.array_equal(g1, h1.grad.numpy()) and np.array_equal(g2, h2.grad.numpy())
    q1, q2 = olh.left_transform(sysm, t1, t2, l1, l2)
    s = max(1.0, np.abs(q2).max())
    assert np.abs(g1 - q1).max() < 1e-12 * s
    assert np.abs(g2 - q2).max() < 1e-12 * s


def test_lhtr_hf_molecule_matches_oracle():
    """HF/6-31G** (the reference's cc_energy_triples molecule) at its converged
    ground state: the left transformation of two vectors against the oracle."""
    import dalton_amd as d

    eps, B, o = _system("hf_631gss")
    print("lhtr HF at the converged ground state", flush=True)
    sysm = occ.DFSystem(eps, B, o)
    rng = np.random.default_rng(13)
    with d.CCContext(eps, B, o) as c:
        c.ccsd_energy(maxit=60, threnr=1e-10)
        t1, t2 = c.amplitudes()
        L1 = rng.standard_normal((2, c.nt1))
        L2 = rng.standard_normal((2, c.nt2))
        R1, R2 = c.lhtr(L1, L2, convention="natural")
    for k in range(2):
        q1, q2 = olh.left_transform(sysm, t1, t2, L1[k], L2[k])
        s = max(1.0, np.abs(q2).max())
        assert np.abs(R1[k] - q1).max() < 1e-12 * s and np.abs(R2[k] - q2).max() < 1e-12 * s


def test_lhtr_is_daltons_left_product_of_its_fd_jacobian():
    """ccgpu_lhtr in Dalton's left-vector convention against the reference's
    own definition of its Jacobian: the records of CCLR_FDJAC
    (cc/cclr_util.F:946-1400: central differences of the vector function,
    here ccgpu_rhs, with the doubles diagonal rows halved and the diagonal
    doubles records doubled) applied from the left as CCLR_DUMTRR does
    (:1624, NSIDE = -1), at the converged ground state of water/STO-3G and at
    random amplitudes of a synthetic system."""
    import dalton_amd as d
    from dalton_amd.cc import scale_pdiag

    for label in ("h2o_sto3g_exgr", "synthetic_o3_v5_n12"):
        if label.startswith("synthetic"):
            from dalton_amd.synthetic import make_system

            eps, B = make_system(3, 5, 12, 4)
            o = 3
        else:
            eps, B, o, _ = load_molecule(label)
        print(f"lhtr vs FDJAC protocol {label}", flush=True)
        rng = np.random.default_rng(17)
        with d.CCContext(eps, B, o) as c:
            if label.startswith("synthetic"):
                t1 = 0.03 * rng.standard_normal(c.nt1)
                t2 = 0.03 * rng.standard_normal(c.nt2)
                c.set_amplitudes(t1, t2)
            else:
                c.ccsd_energy(maxit=60, threnr=1e-12)
                t1, t2 = c.amplitudes()
            nt1, nt2 = c.nt1, c.nt2
            diag = {ai * (ai + 3) // 2 for ai in range(nt1)}
            delta = 1e-4
            recs = np.empty((nt1 + nt2, nt1 + nt2))
            for nu in range(nt1 + nt2):
                a1, a2 = t1.copy(), t2.copy()
                x, k = (a1, nu) if nu < nt1 else (a2, nu - nt1)
                x[k] -= 0.5 * delta
                m1, m2 = c.rhs(a1, a2)
                x[k] += delta
                p1, p2 = c.rhs(a1, a2)
                col = np.concatenate([p1 - m1, scale_pdiag(p2 - m2, 0.5)]) / delta
                recs[nu] = 2.0 * col if (nu >= nt1 and k in diag) else col
            L1 = rng.standard_normal((2, nt1))
            L2 = rng.standard_normal((2, nt2))
            R1, R2 = c.lhtr(L1, L2)
        for q in range(2):
            dum = recs @ np.concatenate([L1[q], L2[q]])
            got = np.concatenate([R1[q], R2[q]])
            err = np.abs(dum - got).max() / np.abs(got).max()
            print(f"  vector {q}: max rel |DUMTRR - lhtr| = {err:.2e}", flush=True)
            assert err < 1e-7
