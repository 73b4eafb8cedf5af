"""Keys (keyg/keyv) construction: product (numpy) vs oracle (C restatement of
segkeys, bigdft/src/locregs/initlocregs.f90:1573-1686) bit-exactly, plus
brute-force decoding, and the compression known-answer case of the
reference's conv_check (bigdft/src/OpenCL/conv_check.f90:1622-1660: every
(i2,i3) row of a 77^3 grid with i1 in [5,71]; the reference's
tests/libs/OpenCL/OpenCL.ref.yaml reports 3,177,944 coefficients for it)."""
import numpy as np
import pytest

import oracle
from bigdft_hip.descriptors import segkeys, sphere_logrid, sphere_region


def decode(n, nseg, keyg, keyv):
    """set of (linear index -> vector index) by brute force"""
    m1, m2 = n[0] + 1, n[1] + 1
    out = {}
    for s in range(nseg):
        j0, j1, jj = int(keyg[2 * s]), int(keyg[2 * s + 1]), int(keyv[s])
        assert (j0 - 1) // m1 == (j1 - 1) // m1, "segment crosses a row"
        for j in range(j0, j1 + 1):
            out[j - 1] = jj - 1 + (j - j0)
    return out


@pytest.mark.parametrize("seed", range(6))
def test_segkeys_matches_oracle_and_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = tuple(int(x) for x in rng.integers(0, 9, size=3))
    lg = rng.random(tuple(x + 1 for x in n)) < rng.uniform(0.1, 0.9)
    nseg, nvctr, kg, kv = segkeys(lg)
    onseg, onvctr, okg, okv = oracle.segkeys(lg)
    assert (nseg, nvctr) == (onseg, onvctr)
    assert np.array_equal(kg, okg) and np.array_equal(kv, okv)
    m = decode(n, nseg, kg, kv)
    idx = np.nonzero(lg.ravel(order="F"))[0]
    assert sorted(m.keys()) == list(idx)
    assert sorted(m.values()) == list(range(nvctr))
    # keys are ordered like the grid: vector index increases with position
    assert [m[i] for i in idx] == list(range(nvctr))


def test_empty_and_full_grids():
    nseg, nvctr, kg, kv = segkeys(np.zeros((3, 4, 5), dtype=bool))
    assert (nseg, nvctr, kg.size, kv.size) == (0, 0, 0, 0)
    nseg, nvctr, kg, kv = segkeys(np.ones((3, 4, 5), dtype=bool))
    assert nseg == 20 and nvctr == 60
    assert kg[0] == 1 and kg[-1] == 60 and kv[-1] == 58


def test_conv_check_compression_pattern():
    n1 = 76
    lg = np.zeros((n1 + 1,) * 3, dtype=bool)
    lg[5:n1 - 5 + 1] = True
    nseg, nvctr, kg, kv = segkeys(lg)
    assert nseg == (n1 + 1) ** 2
    assert 8 * nvctr == 3177944
    i3, i2 = 3, 11
    iseg = i2 + (n1 + 1) * i3
    assert kg[2 * iseg] == i3 * (n1 + 1) ** 2 + i2 * (n1 + 1) + 5 + 1
    assert kg[2 * iseg + 1] == i3 * (n1 + 1) ** 2 + i2 * (n1 + 1) + (n1 - 5) + 1
    assert kv[iseg] == 1 + iseg * (n1 - 10 + 1)


def test_sphere_criterion_inclusive():
    h = 0.5
    lg = sphere_logrid((8, 8, 8), (h, h, h), (2.0, 2.0, 2.0), 1.0)
    # points exactly at distance 1.0 (i = 2, 6 along one axis) are inside
    assert lg[2, 4, 4] and lg[6, 4, 4] and lg[4, 2, 4]
    assert not lg[1, 4, 4]
    brute = np.array([[[((i * h - 2) ** 2 + (j * h - 2) ** 2 + (k * h - 2) ** 2) <= 1.0
                        for k in range(9)] for j in range(9)] for i in range(9)])
    assert np.array_equal(lg, brute)


def test_sphere_region_local_box():
    reg, nl = sphere_region((40, 40, 40), (0.4,) * 3, (8.1, 7.3, 9.2), 3.0, 1.5)
    assert reg.geocode == "F"
    assert all(reg.nfl[d] >= 0 and reg.nfu[d] <= reg.n[d] for d in range(3))
    assert reg.nvctr_f < reg.nvctr_c
    # every point of the region is within radius of the centre
    m = decode(reg.n, reg.nseg_c, reg.keyg_c, reg.keyv_c)
    m1, m2 = reg.n[0] + 1, reg.n[1] + 1
    for lin in list(m)[::37]:
        i1, i2, i3 = lin % m1, (lin // m1) % m2, lin // (m1 * m2)
        r2 = (((i1 + nl[0]) * 0.4 - 8.1) ** 2 + ((i2 + nl[1]) * 0.4 - 7.3) ** 2 +
              ((i3 + nl[2]) * 0.4 - 9.2) ** 2)
        assert r2 <= 9.0 + 1e-12


def test_sphere_region_matches_full_scan():
    """the bounding-range evaluation selects exactly the points of a scan of
    the whole global box"""
    rng = np.random.default_rng(5)
    n_glob, h = (30, 27, 33), (0.4, 0.37, 0.41)
    for _ in range(6):
        c = [rng.uniform(-1.0, n_glob[d] * h[d] + 1.0) for d in range(3)]
        rc, rf = rng.uniform(1.0, 4.0), rng.uniform(0.5, 2.0)
        full = sphere_logrid(n_glob, h, c, rc)
        if not full.any():
            continue
        reg, nl = sphere_region(n_glob, h, c, rc, rf)
        idx = np.nonzero(full)
        assert nl == tuple(int(idx[d].min()) for d in range(3))
        nu = tuple(int(idx[d].max()) for d in range(3))
        assert reg.n == tuple(nu[d] - nl[d] for d in range(3))
        assert reg.nvctr_c == int(full.sum())
        box = full[nl[0]:nu[0] + 1, nl[1]:nu[1] + 1, nl[2]:nu[2] + 1]
        fine = sphere_logrid(reg.n, h, c, rf, origin=nl) & box
        assert reg.nvctr_f == int(fine.sum())
