"""Pinning of the oracle's density (daub_to_isf + partial_density).

The density is tied to the local Hamiltonian's potential energy, which is
pinned independently (test_oracle_operators.py): for any potential V,
    sum_p V(p) rho_s(p) = sum_{orbitals of spin s} hfac * epot(V),
because epot = sum V psir^2 with the same psir.  A random V makes this a
pointwise check of rho.  Further: partition of unity (a constant coarse
orbital gives a constant psir on a periodic box), positivity, accumulation.
"""
import numpy as np
import pytest

import oracle
from helpers import dense_region, mixed_sphere_region, okeys, random_pot, random_psi


@pytest.mark.parametrize("reg", [
    dense_region("P", (9, 8, 10), (0.4, 0.45, 0.5)),
    mixed_sphere_region((14, 12, 13), 0.4, 3),
    dense_region("F", (7, 9, 6), (0.5, 0.4, 0.45)),
], ids=["P", "F-sphere", "F-dense"])
def test_density_matches_potential_energy(reg):
    psi = random_psi(reg, 3, 5)
    hf = np.array([0.7, 1.3, 2.0])
    sp = np.array([0, 1, 0], dtype=np.int32)
    rho = oracle.density(reg.geocode, reg.n, reg.nfl, reg.nfu, okeys(reg), psi,
                         hf, sp, nspin=2)
    for seed in (6, 7):
        pot = random_pot(reg, seed)
        _, _, ep = oracle.locham(reg.geocode, reg.n, reg.nfl, reg.nfu, reg.hgrid,
                                 okeys(reg), pot, psi)
        for s in (0, 1):
            lhs = np.dot(rho[s], pot)
            rhs = np.dot(hf[sp == s], ep[sp == s])
            assert abs(lhs - rhs) < 1e-12 * abs(rhs)
    assert np.all(rho >= 0.0)


def test_density_constant_orbital_is_uniform():
    """sum_k phi(x - k) = 1 and the magic filter sums to 1: a constant coarse
    orbital (no wavelets) has a constant psir on a periodic box"""
    from bigdft_hip.descriptors import region_from_grids
    n = (7, 8, 9)
    full = np.ones(tuple(x + 1 for x in n), dtype=bool)
    reg = region_from_grids("P", n, (0.4,) * 3, full, np.zeros_like(full))
    psi = np.full(reg.ncoef, 0.3)
    psir = oracle.daub_to_isf("P", n, reg.nfl, reg.nfu, okeys(reg), psi)
    assert np.ptp(psir) < 1e-13 * abs(psir).max()
    rho = oracle.density("P", n, reg.nfl, reg.nfu, okeys(reg), psi[None, :],
                         np.array([2.0]))
    assert np.allclose(rho[0], 2.0 * psir[0] ** 2, rtol=1e-13, atol=0)


def test_density_accumulates_and_skips_zero_weights():
    reg = dense_region("P", (5, 6, 5), (0.45,) * 3)
    psi = random_psi(reg, 2, 9)
    r1 = oracle.density("P", reg.n, reg.nfl, reg.nfu, okeys(reg), psi,
                        np.array([1.0, 0.0]))
    r0 = np.full_like(r1, 1e-20)
    r2 = oracle.density("P", reg.n, reg.nfl, reg.nfu, okeys(reg), psi,
                        np.array([1.0, 0.0]), rho=r0)
    assert np.array_equal(r2, r0 + r1)
    ps = oracle.daub_to_isf("P", reg.n, reg.nfl, reg.nfu, okeys(reg), psi[0])
    assert np.array_equal(r1[0], 1.0 * (ps * ps))
