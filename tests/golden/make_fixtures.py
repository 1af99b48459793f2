"""Generate the molecular fixtures under tests/golden/ (run in a container
that has the reference's basis library; the outputs are committed).

For each molecule: RHF in the Dalton basis set, canonical MOs, MO
integrals, and an exact pivoted-Cholesky factor B with
(pq|rs) = sum_P B[P,p,q] B[P,r,s] to ~1e-13.  Stored: eps, B packed over
p<=q (np.triu_indices order), the SCF and nuclear-repulsion energies.

Molecules (geometries from the reference's test inputs):
* h2o_ccpvdz_test   DALTON/test/newccsdpt_h2o/newccsdpt_h2o.mol (Angstrom)
* hf_631gss         DALTON/test/cc_energy_triples:34-40 (bohr); the
                    degenerate pi pairs are rotated into eigenfunctions of
                    the y -> -y reflection, as Dalton's symmetry-blocked SCF
                    produces them (Dalton's DIIS inner product runs over
                    packed ai<=bj amplitudes, which is not invariant under
                    rotations inside a degenerate set, so the printed
                    iteration energies depend on this choice)
* h2o_ccpvdz_base   the water geometry O (0,0,0), H (0,+-0.757,0.586) A
                    used for the timing/parity configuration (no Dalton
                    output exists for it)
* h2o_sto3g_exgr    DALTON/test/cc_exgr_prop (STO-3G, bohr), whose output
                    holds the unrelaxed CCSD dipole moment; this fixture also
                    stores the MO dipole-length and second-moment integrals
                    (origin 0, as printed there), dip[d][p][q] and
                    smom[d][e][p][q]

Usage:  python tests/golden/make_fixtures.py [/path/to/reference]
        python tests/golden/make_fixtures.py /path/to/reference --ao [--only NAME]   (AO fixture; default h2o_ccpvdz_test)
        python tests/golden/make_fixtures.py /path/to/reference --only NAME   (one molecule)
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import gto  # noqa: E402

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
A = gto.ANG2BOHR_DALTON

MOLECULES = {
    "h2o_ccpvdz_test": dict(
        basis="cc-pVDZ",
        atoms=[(8, np.array([0.0, -0.07579, 5.0]) * A),
               (1, np.array([0.86681, 0.60144, 5.0]) * A),
               (1, np.array([-0.86681, 0.60144, 5.0]) * A)],
        nocc=5, adapt=False),
    "hf_631gss": dict(
        basis="6-31G**",
        atoms=[(9, np.array([0.0, 0.0, 1.7632])), (1, np.array([0.0, 0.0, 0.0]))],
        nocc=5, adapt=True),
    "h2o_ccpvdz_base": dict(
        basis="cc-pVDZ",
        atoms=[(8, np.array([0.0, 0.0, 0.0]) * A),
               (1, np.array([0.0, 0.757, 0.586]) * A),
               (1, np.array([0.0, -0.757, 0.586]) * A)],
        nocc=5, adapt=False),
    "h2o_sto3g_exgr": dict(
        basis="STO-3G",
        atoms=[(8, np.array([0.0, 0.0, -0.13594222])),
               (1, np.array([0.0, 1.4476145, 1.07875057])),
               (1, np.array([0.0, -1.4476145, 1.07875057]))],
        nocc=5, adapt=False, dipole=True),
}


def build(name: str, spec: dict) -> dict:
    bas = gto.Basis(spec["atoms"], os.path.join(REF, "basis", spec["basis"]))
    S, T, V = gto.one_electron(bas)
    G = gto.two_electron(bas)
    enuc = bas.nuclear_repulsion()
    E, eps, C = gto.rhf(S, T + V, G, spec["nocc"], enuc)
    if spec["adapt"]:
        C = gto.adapt_degenerate(C, eps, S, gto.reflection_y(bas))
    Gmo = np.einsum("pqrs,pa,qb,rc,sd->abcd", G, C, C, C, C, optimize=True)
    B = gto.mo_cholesky(Gmo)
    err = np.abs(np.einsum("Ppq,Prs->pqrs", B, B, optimize=True) - Gmo).max()
    iu = np.triu_indices(bas.nbf)
    print(f"{name}: nbf={bas.nbf} naux={B.shape[0]} E_nuc={enuc:.12f} E_SCF={E:.12f} chol_err={err:.1e}")
    out = dict(eps=eps, Bpk=B[:, iu[0], iu[1]], escf=E, enuc=enuc, nocc=spec["nocc"])
    if spec.get("dipole"):
        M = gto.dipole(bas)
        out["dip"] = np.einsum("dab,ap,bq->dpq", M, C, C, optimize=True)
        out["smom"] = np.einsum("deab,ap,bq->depq", gto.second_moment(bas), C, C, optimize=True)
        mu = -2.0 * np.einsum("dii->d", out["dip"][:, : spec["nocc"], : spec["nocc"]])
        print(f"{name}: electronic SCF dipole {mu}")
    return out


def build_ao(name: str, spec: dict) -> dict:
    """AO-side fixture of one molecule for the AO -> density-fitting front end:
    the AO two-electron integrals as the packed lower triangle of the pair
    supermatrix (pairs (a,b), a >= b, at a(a+1)/2 + b; element (P,Q), P >= Q,
    at P(P+1)/2 + Q), the canonical MO coefficients column by column
    (cmo[p*nbas + a] = C[a,p]) and the orbital energies."""
    bas = gto.Basis(spec["atoms"], os.path.join(REF, "basis", spec["basis"]))
    S, T, V = gto.one_electron(bas)
    G = gto.two_electron(bas)
    enuc = bas.nuclear_repulsion()
    E, eps, C = gto.rhf(S, T + V, G, spec["nocc"], enuc)
    n = bas.nbf
    il = np.tril_indices(n)
    Vs = G[il[0], il[1]][:, il[0], il[1]]
    ip = np.tril_indices(Vs.shape[0])
    print(f"{name} (AO): nbf={n} pairs={Vs.shape[0]} E_SCF={E:.12f}")
    out = dict(eri_pairs=Vs[ip[0], ip[1]], cmo=np.ascontiguousarray(C.T).reshape(-1), eps=eps, escf=E,
               enuc=enuc, nocc=spec["nocc"], nbas=n)
    if spec.get("dipole"):
        out["dip_ao"] = gto.dipole(bas)  # [d][mu][nu], origin 0
    return out


if __name__ == "__main__":
    if "--ao" in sys.argv:
        name = sys.argv[sys.argv.index("--only") + 1] if "--only" in sys.argv else "h2o_ccpvdz_test"
        d = build_ao(name, MOLECULES[name])
        np.savez_compressed(os.path.join(HERE, name + "_ao.npz"), **d)
        sys.exit(0)
    only = sys.argv[sys.argv.index("--only") + 1] if "--only" in sys.argv else None
    for name, spec in MOLECULES.items():
        if only is not None and name != only:
            continue
        d = build(name, spec)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)
