"""Pinning of the oracle's confining potential (apply_potential_lr_conf*,
bigdft/src/orbitals/potential.f90:203-560; confpot_data,
locregs/locreg_operations.f90:23-30).

The confinement adds prefac * r2^(potorder/2) to the local potential, r2 the
squared distance of the ISF point hh*(i + ioffset) from rxyzConf.  Checks:
  * with ioffset placing the box's ISF points at their physical positions
    (periodic: -1, fine index i-1 at (i-1) h/2; free: -15, fine index i-15),
    the Hamiltonian with confinement equals the Hamiltonian of the potential
    V + prefac |r - rc|^order tabulated on the physical grid (the geometry
    used by every other potential in these tests);
  * potorder/2 is an integer division: order 1 is the constant shift
    prefac, order 3 equals order 2;
  * potorder 0 is no confinement;
  * the confinement acts only where the potential box covers the
    wavefunction box (a window through ishift), psir is zero elsewhere.
"""
import numpy as np
import pytest

import oracle
from helpers import dense_region, field, mixed_sphere_region, okeys, random_pot, random_psi


def conf_field(geocode, n, h, prefac, order, rc):
    half = order // 2
    return field(geocode, n, h, lambda X, Y, Z: prefac * (
        (X - rc[0]) ** 2 + (Y - rc[1]) ** 2 + (Z - rc[2]) ** 2) ** half)


def run(reg, pot, psi, conf=None):
    return oracle.locham(reg.geocode, reg.n, reg.nfl, reg.nfu, reg.hgrid, okeys(reg), pot,
                         psi, conf=conf)


CASES = [("P", dense_region("P", (9, 7, 11), (0.45, 0.4, 0.5)), -1),
         ("F", mixed_sphere_region((12, 14, 11), 0.45, 5), -15)]


@pytest.mark.parametrize("name,reg,off", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("order", [2, 4, 6])
def test_confinement_is_physical_potential(name, reg, off, order):
    h = reg.hgrid
    rc = [0.37 * reg.n[d] * h[d] for d in range(3)]
    prefac = 3e-3
    pot = random_pot(reg, 4)
    psi = random_psi(reg, 2, 9)
    conf = np.concatenate([oracle.confpot(order, prefac, [x / 2 for x in h], rc, [off] * 3)] * 2)
    h1, k1, p1 = run(reg, pot, psi, conf)
    h2, k2, p2 = run(reg, pot + conf_field(reg.geocode, reg.n, h, prefac, order, rc), psi)
    assert np.abs(h1 - h2).max() < 1e-12 * np.abs(h2).max()
    assert np.allclose(p1, p2, rtol=1e-12, atol=0) and np.allclose(k1, k2, rtol=1e-13, atol=0)
    # the confinement is not a no-op
    h0, _, p0 = run(reg, pot, psi)
    assert np.abs(p1 - p0).min() > 1e-6 * np.abs(p0).max()


def test_order_is_halved_by_integer_division():
    reg = CASES[1][1]
    h = reg.hgrid
    rc, hh = (2.0, 2.5, 1.5), [x / 2 for x in h]
    pot = random_pot(reg, 6)
    psi = random_psi(reg, 1, 3)
    h1, _, p1 = run(reg, pot, psi, oracle.confpot(1, 0.25, hh, rc, (-15,) * 3))
    h2, _, p2 = run(reg, pot + 0.25, psi)
    assert np.abs(h1 - h2).max() < 1e-12 * np.abs(h2).max() and abs(p1[0] - p2[0]) < 1e-12 * abs(p2[0])
    h3, _, p3 = run(reg, pot, psi, oracle.confpot(3, 0.01, hh, rc, (-15,) * 3))
    h4, _, p4 = run(reg, pot, psi, oracle.confpot(2, 0.01, hh, rc, (-15,) * 3))
    assert np.abs(h3 - h4).max() < 1e-14 * np.abs(h4).max() and abs(p3[0] - p4[0]) < 1e-13 * abs(p4[0])
    h5, _, p5 = run(reg, pot, psi, oracle.confpot(0, 5.0, hh, rc, (-15,) * 3))
    h6, _, p6 = run(reg, pot, psi)
    assert np.abs(h5 - h6).max() < 1e-14 * np.abs(h6).max() and abs(p5[0] - p6[0]) < 1e-13 * abs(p6[0])


def test_confinement_only_inside_the_potential_window():
    dims = (20, 18, 22)
    pdims = (14, 15, 12)
    ishift = (3, -2, 6)
    rng = np.random.default_rng(2)
    psir = rng.standard_normal(dims[0] * dims[1] * dims[2])
    pot = rng.uniform(-1, 0, pdims[0] * pdims[1] * pdims[2])
    conf = oracle.confpot(4, 0.02, (0.2, 0.22, 0.25), (1.0, 2.0, 3.0), (5, -3, 2))
    out, ep = oracle.apply_potential_conf(psir, pot, dims, conf, potdims=pdims, ishift=ishift)
    P = psir.reshape(dims[::-1]).transpose(2, 1, 0)           # [i1, i2, i3]
    V = pot.reshape(pdims[::-1]).transpose(2, 1, 0)
    exp = np.zeros_like(P)
    e = 0.0
    for i3 in range(dims[2]):
        for i2 in range(dims[1]):
            for i1 in range(dims[0]):
                q = (i1 - ishift[0], i2 - ishift[1], i3 - ishift[2])
                if all(0 <= q[d] < pdims[d] for d in range(3)):
                    x = [0.2 * (i1 + 1 + 5) - 1.0, 0.22 * (i2 + 1 - 3) - 2.0,
                         0.25 * (i3 + 1 + 2) - 3.0]
                    v = V[q] + 0.02 * (x[0] ** 2 + x[1] ** 2 + x[2] ** 2) ** 2
                    exp[i1, i2, i3] = v * P[i1, i2, i3]
                    e += v * P[i1, i2, i3] ** 2
    got = out.reshape(dims[::-1]).transpose(2, 1, 0)
    assert np.abs(got - exp).max() < 1e-13 * np.abs(exp).max()
    assert abs(ep - e) < 1e-12 * abs(e)
