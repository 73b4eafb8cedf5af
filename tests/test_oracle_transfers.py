"""Pinning of the oracle's region transfers (psi_to_locreg2, Lpsi_to_global2).

Brute force: every compressed point of the local region is mapped by its
grid coordinates (+ the box offset ns) to the global region's compressed
index, read off the keys independently of the segment-intersection loops of
the reference; both directions must agree with it bit for bit, and a
global -> local -> global round trip restores the global values at the
local points.
"""
import numpy as np
import pytest

import oracle
from bigdft_hip.descriptors import region_from_grids, sphere_logrid
from helpers import okeys


def lin_points(keyg, nseg):
    """0-based linear grid index of each compressed point, in order"""
    out = []
    for s in range(nseg):
        out.append(np.arange(int(keyg[2 * s]) - 1, int(keyg[2 * s + 1])))
    return np.concatenate(out) if out else np.zeros(0, dtype=np.int64)


def setup(seed):
    rng = np.random.default_rng(seed)
    ng = (30, 28, 32)
    h = (0.4, 0.4, 0.4)
    cg = [ng[d] * h[d] / 2 for d in range(3)]
    coarse_g = sphere_logrid(ng, h, cg, 6.0)
    fine_g = sphere_logrid(ng, h, cg, 4.0) & coarse_g
    glob = region_from_grids("F", ng, h, coarse_g, fine_g)
    # local box inside the global one, local spheres inside the global ones
    nl = tuple(int(x) for x in rng.integers(10, 16, size=3))
    ns = tuple(int(rng.integers(3, ng[d] - nl[d] - 3)) for d in range(3))
    cl = [cg[d] + rng.uniform(-0.8, 0.8) for d in range(3)]
    coarse_l = sphere_logrid(nl, h, cl, 2.5, origin=ns)
    fine_l = sphere_logrid(nl, h, cl, 1.5, origin=ns) & coarse_l
    sub = coarse_g[ns[0]:ns[0] + nl[0] + 1, ns[1]:ns[1] + nl[1] + 1, ns[2]:ns[2] + nl[2] + 1]
    subf = fine_g[ns[0]:ns[0] + nl[0] + 1, ns[1]:ns[1] + nl[1] + 1, ns[2]:ns[2] + nl[2] + 1]
    coarse_l &= sub
    fine_l &= subf
    loc = region_from_grids("F", nl, h, coarse_l, fine_l)
    return glob, loc, ng, nl, ns


def brute_map(glob, loc, ng, nl, ns):
    """(coarse, fine) global compressed index of each local compressed point"""
    def to_glob(lin):
        m1, m2 = nl[0] + 1, nl[1] + 1
        i1, i2, i3 = lin % m1, (lin // m1) % m2, lin // (m1 * m2)
        return (i1 + ns[0]) + (ng[0] + 1) * ((i2 + ns[1]) + (ng[1] + 1) * (i3 + ns[2]))
    gc = {int(p): k for k, p in enumerate(lin_points(glob.keyg_c, glob.nseg_c))}
    gf = {int(p): k for k, p in enumerate(lin_points(glob.keyg_f, glob.nseg_f))}
    mc = np.array([gc.get(int(p), -1) for p in to_glob(lin_points(loc.keyg_c, loc.nseg_c))])
    mf = np.array([gf.get(int(p), -1) for p in to_glob(lin_points(loc.keyg_f, loc.nseg_f))])
    return mc, mf


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_transfers_match_brute_force(seed):
    glob, loc, ng, nl, ns = setup(seed)
    assert loc.nvctr_c > 0 and loc.nvctr_f > 0
    mc, mf = brute_map(glob, loc, ng, nl, ns)
    assert np.all(mc >= 0) and np.all(mf >= 0)
    rng = np.random.default_rng(seed + 10)
    gpsi = rng.standard_normal(glob.ncoef)
    lpsi = oracle.psi_to_locreg2(nl, ng, ns, okeys(loc), okeys(glob), gpsi)
    exp = np.concatenate([gpsi[mc], gpsi[glob.nvctr_c:].reshape(-1, 7)[mf].ravel()])
    assert np.array_equal(lpsi, exp)
    lsrc = rng.standard_normal(loc.ncoef)
    out, n = oracle.lpsi_to_global2(nl, ng, ns, okeys(loc), okeys(glob), lsrc,
                                    np.zeros(glob.ncoef))
    assert n == loc.nvctr_c + loc.nvctr_f
    ref = np.zeros(glob.ncoef)
    ref[mc] = lsrc[:loc.nvctr_c]
    ref[glob.nvctr_c:].reshape(-1, 7)[mf] = lsrc[loc.nvctr_c:].reshape(-1, 7)
    assert np.array_equal(out, ref)
    # round trip: global -> local -> global restores the local points
    back, _ = oracle.lpsi_to_global2(nl, ng, ns, okeys(loc), okeys(glob), lpsi,
                                     np.zeros(glob.ncoef))
    touched = ref != 0.0
    assert np.array_equal(back[touched], gpsi[touched])


def test_psi_to_locreg_zero_outside_global():
    """local points absent from the global region stay zero"""
    glob, loc, ng, nl, ns = setup(4)
    full = np.ones(tuple(x + 1 for x in nl), dtype=bool)
    loc2 = region_from_grids("F", nl, loc.hgrid, full, np.zeros_like(full))
    mc, _ = brute_map(glob, loc2, ng, nl, ns)
    gpsi = np.random.default_rng(5).standard_normal(glob.ncoef)
    lpsi = oracle.psi_to_locreg2(nl, ng, ns, okeys(loc2), okeys(glob), gpsi)
    # the reference's copy assumes every local point is in the global
    # region; where rows are fully inside or fully outside it the result is
    # the per-point map
    rows_ok = True
    for s in range(loc2.nseg_c):
        a, b = int(loc2.keyv_c[s]) - 1, int(loc2.keyv_c[s]) - 1 + \
            int(loc2.keyg_c[2 * s + 1]) - int(loc2.keyg_c[2 * s])
        seg = mc[a:b + 1]
        rows_ok &= bool(np.all(seg >= 0) or np.all(seg < 0))
    if rows_ok:
        exp = np.where(mc >= 0, gpsi[np.maximum(mc, 0)], 0.0)
        assert np.array_equal(lpsi, exp)
    assert np.all(lpsi[mc < 0] == 0.0) if rows_ok else True
