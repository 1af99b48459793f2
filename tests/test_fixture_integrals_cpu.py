"""The fixture generator's one-electron moment integrals (tests/golden/gto.py)
against closed forms: a normalised s Gaussian has <x> = A_x and
<x^2> = A_x^2 + 1/(4 alpha); for any basis on one centre a shift of the
centre by delta moves <x> by delta S and <x^2> by 2 delta <x> + delta^2 S."""
import os
import sys

import numpy as np

from conftest import ROOT

sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import gto  # noqa: E402

BASIS = """a 8
$ S-TYPE FUNCTIONS
    2    1    0
      5.0  0.4
      0.9  0.7
$ P-TYPE FUNCTIONS
    1    1    0
      1.1  1.0
$ D-TYPE FUNCTIONS
    1    1    0
      0.7  1.0
a 1
$ S-TYPE FUNCTIONS
    1    1    0
      0.8  1.0
"""


def _basis(tmp_path, atoms):
    p = tmp_path / "basis"
    p.write_text(BASIS)
    return gto.Basis(atoms, str(p))


def test_single_s_gaussian_moments(tmp_path):
    A = np.array([0.3, -1.2, 0.7])
    b = _basis(tmp_path, [(1, A)])
    S, _, _ = gto.one_electron(b)
    M1 = gto.dipole(b)
    M2 = gto.second_moment(b)
    assert abs(S[0, 0] - 1.0) < 1e-14
    assert np.abs(M1[:, 0, 0] - A).max() < 1e-14
    for d in range(3):
        assert abs(M2[d, d, 0, 0] - (A[d] ** 2 + 1 / (4 * 0.8))) < 1e-13
        for e in range(3):
            if e != d:
                assert abs(M2[d, e, 0, 0] - A[d] * A[e]) < 1e-13


def test_moments_follow_a_shift_of_the_centre(tmp_path):
    """s, p and pure-d shells on one centre: exact translation identities."""
    A = np.array([0.1, 0.4, -0.3])
    delta = np.array([0.5, -0.25, 1.5])
    b0 = _basis(tmp_path, [(8, A)])
    b1 = _basis(tmp_path, [(8, A + delta)])
    S, _, _ = gto.one_electron(b0)
    m1a, m1b = gto.dipole(b0), gto.dipole(b1)
    m2a, m2b = gto.second_moment(b0), gto.second_moment(b1)
    for d in range(3):
        assert np.abs(m1b[d] - (m1a[d] + delta[d] * S)).max() < 1e-12
        for e in range(3):
            ref = m2a[d, e] + delta[d] * m1a[e] + delta[e] * m1a[d] + delta[d] * delta[e] * S
            assert np.abs(m2b[d, e] - ref).max() < 1e-12
    # hermitian, and the second moments symmetric in (d, e)
    assert np.abs(m1a - m1a.transpose(0, 2, 1)).max() < 1e-13
    assert np.abs(m2a - m2a.transpose(1, 0, 2, 3)).max() < 1e-13
    assert np.abs(m2a - m2a.transpose(0, 1, 3, 2)).max() < 1e-13


def test_two_centre_dipole_equals_overlap_weighted_position(tmp_path):
    """<a| x - O |b> summed with the density of a normalised function
    reproduces the centroid: for two s functions on different centres the
    dipole of the pair equals the Gaussian product centre times the overlap."""
    A = np.array([0.0, 0.0, 0.0])
    B = np.array([0.0, 0.0, 1.4])
    b = _basis(tmp_path, [(1, A), (1, B)])
    S, _, _ = gto.one_electron(b)
    M1 = gto.dipole(b)
    P = (A + B) / 2  # equal exponents: the product centre is the midpoint
    assert np.abs(M1[:, 0, 1] - P * S[0, 1]).max() < 1e-14
