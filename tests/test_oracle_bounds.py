"""Pinning of the oracle's convolution bounds (locreg_bounds and what it
calls, bigdft/src/locregs/bounds.f90:555-1123).

The oracle follows the reference's mechanism (one flat logical work grid,
filled from intervals and read back with rotated axes).  Three independent
checks pin it:
  * every stage against a direct per-axis formulation: for an output row
    (named output axes), scan the named input axis for the first/last index
    whose interval, spread by the free-BC filter footprint [2i-14, 2i+16],
    covers the row; squares / squares_1d as 2x2 / 2x1 block min-max;
  * the meaning of the arrays: every coarse (fine) grid point's footprint
    lies inside the grow and shrink bounds of the rows it touches, and every
    row bound by a kinetic array holds the row's first and last point;
  * the real-space bound ibyyzz_r against the actual support of psir from
    the oracle's daub_to_isf with random coefficients: contained everywhere,
    and equal (up to the 2x2 block union) on convex sphere regions.
"""
import numpy as np
import pytest

import oracle
from bigdft_hip.descriptors import region_from_grids
from helpers import mixed_sphere_region, okeys

BIG, SMALL = 1000, -1000


def irregular_region(n, seed, pc=0.35, pf=0.4, empty_fine=False):
    rng = np.random.default_rng(seed)
    shape = tuple(x + 1 for x in n)
    coarse = rng.random(shape) < pc
    fine = coarse & (rng.random(shape) < pf)
    if empty_fine:
        fine[:] = False
    return region_from_grids("F", n, (0.4, 0.4, 0.4), coarse, fine)


REGIONS = {
    "sphere_a": lambda: mixed_sphere_region((14, 17, 12), 0.45, 3),
    "sphere_b": lambda: mixed_sphere_region((19, 16, 21), 0.4, 8),
    "irregular": lambda: irregular_region((11, 14, 9), 1),
    "irregular_odd": lambda: irregular_region((13, 9, 16), 2, pc=0.15),
    "no_fine": lambda: irregular_region((8, 11, 10), 3, empty_fine=True),
    "flat": lambda: irregular_region((7, 0, 5), 4, pc=0.6),
}


def grids(reg):
    shape = tuple(x + 1 for x in reg.n)
    out = []
    for kg, ns in ((reg.keyg_c, reg.nseg_c), (reg.keyg_f, reg.nseg_f)):
        g = np.zeros(int(np.prod(shape)), dtype=bool)
        for s in range(ns):
            g[int(kg[2 * s]) - 1:int(kg[2 * s + 1])] = True
        out.append(g.reshape(shape, order="F"))
    return out


def bounds_of(reg):
    buf = oracle.locreg_bounds(reg.n, reg.nfl, reg.nfu, okeys(reg))
    return oracle.unpack_bounds(buf, reg.n, reg.nfl, reg.nfu)


def first_last(mask, axis, lo):
    """first/last index (offset lo) along axis where mask holds, else the
    1000/-1000 sentinels"""
    has = mask.any(axis)
    f = np.argmax(mask, axis) + lo
    l = mask.shape[axis] - 1 - np.argmax(np.flip(mask, axis), axis) + lo
    return np.stack([np.where(has, f, BIG), np.where(has, l, SMALL)])


def covers(ib, C):
    """cov[C, a, b]: C lies in the footprint of interval ib[:, a, b]"""
    C = C[:, None, None]
    return (C >= 2 * ib[0][None] - 14) & (C <= 2 * ib[1][None] + 16)


def stage_inv(ib, a0, b0, C0, C1):
    """out(C, a) = first/last b with C in footprint(ib(a, b))"""
    return first_last(covers(ib, np.arange(C0, C1 + 1)), 2, b0)


def stage_rot(ib, a0, b0, C0, C1):
    """out(b, C) = first/last a with C in footprint(ib(a, b))"""
    m = covers(ib, np.arange(C0, C1 + 1)).transpose(2, 0, 1)   # [b, C, a]
    return first_last(m, 2, a0)


def squares(ib):
    ib = ib.copy()
    d1, d2 = ib.shape[1] - 1, ib.shape[2] - 1
    if d1 == 0 or d2 == 0:
        return ib
    m1, m2 = 2 * ((d1 - 1) // 2 + 1), 2 * ((d2 - 1) // 2 + 1)
    blk = ib[:, :m1, :m2].reshape(2, m1 // 2, 2, m2 // 2, 2)
    lo = blk[0].min(axis=(1, 3))
    hi = blk[1].max(axis=(1, 3))
    ib[0, :m1, :m2] = np.repeat(np.repeat(lo, 2, 0), 2, 1)
    ib[1, :m1, :m2] = np.repeat(np.repeat(hi, 2, 0), 2, 1)
    return ib


def squares_1d(ib, lo1):
    """pairs (2j, 2j+1) of the first axis, j from lo1/2 (lo1 even)"""
    assert lo1 % 2 == 0
    ib = ib.copy()
    m = 2 * ((ib.shape[1] - 1) // 2)
    p = ib[:, :m].reshape(2, m // 2, 2, -1)
    ib[0, :m] = np.repeat(p[0].min(axis=1), 2, 0)
    ib[1, :m] = np.repeat(p[1].max(axis=1), 2, 0)
    return ib


def direct_bounds(reg):
    """every array from its per-axis definition"""
    lc, lf = grids(reg)
    n1, n2, n3 = reg.n
    (l1, l2, l3), (u1, u2, u3) = reg.nfl, reg.nfu
    out = {}
    for tag, g in (("c", lc), ("f", lf)):
        out["ibyz_" + tag] = first_last(g, 0, 0)          # [i2, i3]
        out["ibxz_" + tag] = first_last(g, 1, 0)          # [i1, i3]
        out["ibxy_" + tag] = first_last(g, 2, 0)          # [i1, i2]
    # shrink
    out["ibzzx_c"] = stage_inv(out["ibxy_c"], 0, 0, -14, 2 * n3 + 16)
    out["ibyyzz_c"] = stage_inv(out["ibzzx_c"], -14, 0, -14, 2 * n2 + 16)
    out["ibxy_ff"] = out["ibxy_f"][:, l1:u1 + 1, l2:u2 + 1]
    out["ibzzx_f"] = stage_inv(out["ibxy_ff"], l1, l2, 2 * l3 - 14, 2 * u3 + 16)
    out["ibyyzz_f"] = stage_inv(out["ibzzx_f"], 2 * l3 - 14, l1, 2 * l2 - 14,
                                2 * u2 + 16)
    r = out["ibyyzz_c"]
    has = r[0] != BIG
    out["ibyyzz_r"] = squares(np.stack([np.where(has, 2 * r[0], BIG),
                                        np.where(has, 2 * r[1] + 30, SMALL)]))
    # grow (the coarse kinetic ibyz is squared in place first)
    out["ibyz_ff"] = out["ibyz_f"][:, l2:u2 + 1, l3:u3 + 1]
    out["ibyz_c"] = squares(out["ibyz_c"])
    out["ibzxx_c"] = squares(stage_rot(out["ibyz_c"], 0, 0, -14, 2 * n1 + 16))
    out["ibxxyy_c"] = squares(stage_rot(out["ibzxx_c"], 0, -14, -14, 2 * n2 + 16))
    out["ibzxx_f"] = stage_rot(out["ibyz_ff"], l2, l3, 2 * l1 - 14, 2 * u1 + 16)
    out["ibxxyy_f"] = squares_1d(
        stage_rot(out["ibzxx_f"], l3, 2 * l1 - 14, 2 * l2 - 14, 2 * u2 + 16),
        2 * l1 - 14)
    return out


@pytest.mark.parametrize("name", list(REGIONS))
def test_layout(name):
    reg = REGIONS[name]()
    off, dims = oracle.bounds_layout(reg.n, reg.nfl, reg.nfu)
    assert off[0] == 0 and np.all(np.diff(off) > 0)
    for k in range(17):
        lo1, hi1, lo2, hi2 = dims[k]
        assert off[k + 1] - off[k] == 2 * (hi1 - lo1 + 1) * (hi2 - lo2 + 1)
    n1, n2, n3 = reg.n
    assert tuple(dims[16]) == (-14, 2 * n2 + 16, -14, 2 * n3 + 16)
    assert tuple(dims[10]) == (2 * reg.nfl[0] - 14, 2 * reg.nfu[0] + 16,
                               2 * reg.nfl[1] - 14, 2 * reg.nfu[1] + 16)


@pytest.mark.parametrize("name", list(REGIONS))
def test_every_stage_matches_direct_formulation(name):
    reg = REGIONS[name]()
    got = bounds_of(reg)
    ref = direct_bounds(reg)
    for k in oracle.BOUNDS_NAMES:
        a = got[k][0]
        assert a.shape == ref[k].shape, k
        assert np.array_equal(a, ref[k]), k


def footprint(i):
    return 2 * i - 14, 2 * i + 16


@pytest.mark.parametrize("name", ["sphere_a", "irregular", "irregular_odd"])
def test_bounds_hold_every_footprint(name):
    """each grid point lies inside the bound of every row its filter
    footprint reaches (the arrays limit the convolution loops)"""
    reg = REGIONS[name]()
    b = bounds_of(reg)
    lc, lf = grids(reg)

    def at(key, i, j):
        a, lo1, lo2 = b[key]
        return a[:, i - lo1, j - lo2]

    for tag, g in (("c", lc), ("f", lf)):
        pts = np.argwhere(g)
        rng = np.random.default_rng(0)
        for i1, i2, i3 in pts[rng.choice(len(pts), min(len(pts), 60), replace=False)]:
            for key, i, j, v in (("ibyz_" + tag, i2, i3, i1),
                                 ("ibxz_" + tag, i1, i3, i2),
                                 ("ibxy_" + tag, i1, i2, i3)):
                lo, hi = at(key, i, j)
                assert lo <= v <= hi, key
            s = "c" if tag == "c" else "f"
            zzx, yyzz = ("ibzzx_" + s, "ibyyzz_" + s)
            zxx, xxyy = ("ibzxx_" + s, "ibxxyy_" + s)
            for J3 in range(*footprint(i3)):
                lo, hi = at(zzx, J3, i1)
                assert lo <= i2 <= hi
                for J2 in range(*footprint(i2)):
                    lo, hi = at(yyzz, J2, J3)
                    assert lo <= i1 <= hi
            for J1 in range(*footprint(i1)):
                lo, hi = at(zxx, i3, J1)
                assert lo <= i2 <= hi
                for J2 in range(*footprint(i2)):
                    lo, hi = at(xxyy, J1, J2)
                    assert lo <= i3 <= hi
        if tag == "c":
            # the real-space bound holds the coarse point's fine footprint
            for i1, i2, i3 in pts[:40]:
                lo, hi = at("ibyyzz_r", 2 * i2, 2 * i3)
                assert lo <= 2 * i1 and 2 * i1 + 30 <= hi


@pytest.mark.parametrize("name", ["sphere_a", "sphere_b", "irregular"])
def test_real_space_bound_is_psir_support(name):
    reg = REGIONS[name]()
    a, lo1, lo2 = bounds_of(reg)["ibyyzz_r"]
    assert lo1 == -14 and lo2 == -14
    psi = np.random.default_rng(11).standard_normal(reg.ncoef)
    nf = tuple(2 * x + 31 for x in reg.n)
    psir = oracle.daub_to_isf("F", reg.n, reg.nfl, reg.nfu, okeys(reg), psi)
    nz = psir.reshape(nf, order="F") != 0.0
    sup = first_last(nz, 0, 0)            # [k, P2, P3], 0-based fine indices
    has = sup[0] != BIG
    # contained: psir vanishes outside the bound
    assert np.all(a[0][has] <= sup[0][has]) and np.all(sup[1][has] <= a[1][has])
    if name.startswith("sphere"):
        # tight on convex regions, up to the 2x2 block union of squares
        assert np.array_equal(squares(sup), a)
    assert has.sum() > 0.2 * has.size


def test_sentinels_and_squares_blocks():
    reg = REGIONS["irregular_odd"]()
    b = bounds_of(reg)
    for k in oracle.BOUNDS_NAMES:
        a = b[k][0]
        empty = a[0] == BIG
        assert np.all(a[1][empty] == SMALL), k
        assert np.all(a[0][~empty] <= a[1][~empty]), k
        assert np.all((a[0][~empty] >= -14) & (a[1][~empty] < 999)), k
    for k in ("ibyz_c", "ibzxx_c", "ibxxyy_c", "ibyyzz_r"):
        a = b[k][0]
        m1, m2 = 2 * ((a.shape[1] - 1) // 2), 2 * ((a.shape[2] - 1) // 2)
        blk = a[:, :m1, :m2]
        assert np.array_equal(blk[:, 0::2, 0::2], blk[:, 1::2, 1::2]), k
        assert np.array_equal(blk[:, 0::2, 0::2], blk[:, 1::2, 0::2]), k
