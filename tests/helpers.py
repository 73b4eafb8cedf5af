"""Shared helpers of the test-suite (inputs, regions, metrics)."""
import numpy as np

from bigdft_hip.descriptors import dense_region, region_from_grids, sphere_region

import oracle


def okeys(region):
    """oracle.Keys of a descriptors.Region"""
    return oracle.Keys(region.nseg_c, region.nvctr_c, region.keyg_c,
                       region.keyv_c, region.nseg_f, region.nvctr_f,
                       region.keyg_f, region.keyv_f)


def fine_axis(geocode, n, h):
    """fine coordinates (bohr) of one axis"""
    if geocode == "P":
        return np.arange(2 * n + 2) * (0.5 * h)
    return (np.arange(2 * n + 31) - 14) * (0.5 * h)


def field(geocode, n, hgrid, fun):
    """fun(x,y,z) on the fine grid, flattened x fastest"""
    x, y, z = (fine_axis(geocode, n[d], hgrid[d]) for d in range(3))
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    return fun(X, Y, Z).ravel(order="F")


def random_pot(region, seed, lo=-2.0, hi=0.0):
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, size=region.n_isf)


def random_psi(region, norb, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((norb, region.ncoef))


def rel_err(a, b):
    """max |a-b| / max(|b|, 1e-3 max|b|) per element (the per-element
    relative error, guarded near the zeros of b)."""
    a, b = np.asarray(a), np.asarray(b)
    scale = np.maximum(np.abs(b), 1e-3 * np.abs(b).max())
    return float((np.abs(a - b) / scale).max())


def mixed_sphere_region(n, h, seed, rc_frac=0.45, rf_frac=0.25):
    """free-BC region on an n-box with a spherical coarse region and a
    smaller fine one around a random centre"""
    rng = np.random.default_rng(seed)
    ctr = [(0.3 + 0.4 * rng.random()) * n[d] * h for d in range(3)]
    R = min(n) * h
    from bigdft_hip.descriptors import sphere_logrid
    coarse = sphere_logrid(n, (h, h, h), ctr, rc_frac * R)
    fine = sphere_logrid(n, (h, h, h), ctr, rf_frac * R)
    return region_from_grids("F", n, (h, h, h), coarse, fine)


__all__ = ["dense_region", "region_from_grids", "sphere_region", "okeys",
           "field", "random_pot", "random_psi", "rel_err",
           "mixed_sphere_region"]
