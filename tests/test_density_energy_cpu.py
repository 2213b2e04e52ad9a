"""Pins of the oracle's density restatement (oracle.density, compute_density
of src/node/dyn/util/compute_density.C) and the extension table entries of
the energy / density fast paths.  CPU only."""
import numpy as np
import pytest

import oracle


def _lattice(a=0.1, half=3):
    g = np.arange(-half, half + 1) * a
    return np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T.copy()


def test_density_simple_cubic_lattice_closed_form():
    """Centre of a simple cubic lattice of spacing a: 6 neighbours at a, 12
    at a sqrt 2, so the 12th nearest is at a sqrt 2 and the 11 nearest hold
    11 m: density = 11 m / (4/3 pi (2 a^2)^1.5).  k = 6 stops at the first
    shell: 5 m / (4/3 pi a^3)."""
    a = 0.1
    x = _lattice(a)
    m = np.full(len(x), 0.5)
    c = len(x) // 2
    rho, dk2 = oracle.density(m, x, 12, c, 1)
    want = 11 * 0.5 / ((4.0 / 3.0) * np.pi * (2 * a * a) ** 1.5)
    assert abs(rho[0] - want) <= 1e-13 * want
    assert abs(dk2[0] - 2 * a * a) <= 1e-16
    rho6, dk6 = oracle.density(m, x, 6, c, 1)
    assert abs(rho6[0] - 5 * 0.5 / ((4.0 / 3.0) * np.pi * a ** 3)) <= 1e-13 * rho6[0]
    assert abs(dk6[0] - a * a) <= 1e-17


def test_density_brute_force_sort():
    """Against a sort of all distances (unequal masses, no ties)."""
    rng = np.random.default_rng(2)
    n = 400
    x = rng.normal(size=(n, 3))
    m = rng.uniform(0.1, 2.0, n)
    for k in (2, 12, 31):
        rho, dk2 = oracle.density(m, x, k)
        for i in range(0, n, 37):
            d2 = ((x - x[i]) ** 2).sum(1)
            d2[i] = np.inf
            order = np.argsort(d2, kind="stable")
            want_dk2 = d2[order[k - 1]]
            want = m[order[:k - 1]].sum() / ((4.0 / 3.0) * np.pi * want_dk2 ** 1.5)
            assert abs(dk2[i] - want_dk2) <= 1e-15 * want_dk2
            assert abs(rho[i] - want) <= 1e-12 * want


def test_density_coincident_kth_neighbour_gives_zero():
    x = np.zeros((5, 3))
    x[4] = [1.0, 0.0, 0.0]
    m = np.ones(5)
    rho, dk2 = oracle.density(m, x, 2, 0, 1)
    assert dk2[0] == 0.0 and rho[0] == 0.0


def test_extension_entries_declared():
    from starlab_amd import _abi
    assert "g6x_potential" in _abi.G6X and "g6x_densities" in _abi.G6X
