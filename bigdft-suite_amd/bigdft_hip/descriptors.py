"""Compression descriptors (keyg/keyv) and localisation regions.

Restates the reference's key construction:
  - num_segkeys / segkeys, bigdft/src/locregs/initlocregs.f90:1573-1686:
    one segment per maximal run of present points along x in each (i2,i3)
    row, keyg(1:2) = 1-based linear start/end i3*(n1+1)*(n2+1)+i2*(n1+1)+i1+1,
    keyv = 1-based running offset; coarse and fine keys are separate arrays,
    each with its own offsets starting at 1 (bigdft/src/init.f90:149-166);
  - spherical regions, num_segkeys_sphere / segkeys_Sphere,
    initlocregs.f90:140-286: a point belongs when
    (i1*hx-cx)^2 + (i2*hy-cy)^2 + (i3*hz-cz)^2 <= locrad^2 (inclusive).
Vectorised with numpy; the result is bit-identical to the loops above.
"""
from dataclasses import dataclass

import numpy as np


def segkeys(logrid):
    """Keys of a boolean grid of shape (n1+1, n2+1, n3+1) (x fastest).

    Returns (nseg, nvctr, keyg int32 (2*nseg,), keyv int32 (nseg,))."""
    lg = np.asarray(logrid, dtype=bool)
    m1, m2, m3 = lg.shape
    # rows in Fortran order: i2 fastest, then i3
    rows = lg.transpose(2, 1, 0).reshape(m2 * m3, m1)  # [i3*m2+i2, i1]
    pad = np.zeros((rows.shape[0], 1), dtype=bool)
    d = np.diff(np.concatenate([pad, rows, pad], axis=1).astype(np.int8),
                axis=1)
    r_s, c_s = np.nonzero(d == 1)     # segment starts (row, i1)
    r_e, c_e = np.nonzero(d == -1)    # segment ends, exclusive
    nseg = r_s.size
    nvctr = int(lg.sum())
    if nseg == 0:
        return 0, 0, np.zeros(0, np.int32), np.zeros(0, np.int32)
    i3 = r_s // m2
    i2 = r_s % m2
    base = i3 * (m1 * m2) + i2 * m1
    keyg = np.empty((nseg, 2), dtype=np.int64)
    keyg[:, 0] = base + c_s + 1
    keyg[:, 1] = base + (c_e - 1) + 1
    lens = c_e - c_s
    keyv = np.concatenate([[1], 1 + np.cumsum(lens)[:-1]])
    return (nseg, nvctr, keyg.astype(np.int32).ravel(),
            keyv.astype(np.int32))


def sphere_logrid(n, hgrid, center, radius, origin=(0, 0, 0)):
    """Points (i1,i2,i3) of the box (0:n1,0:n2,0:n3) whose global position
    ((i+origin)*h) lies within radius of center (inclusive, as
    num_segkeys_sphere, initlocregs.f90:233-239)."""
    ax = [((np.arange(n[d] + 1) + origin[d]) * hgrid[d] - center[d]) ** 2
          for d in range(3)]
    dist2 = ax[0][:, None, None] + ax[1][None, :, None] + ax[2][None, None, :]
    return dist2 <= radius * radius


@dataclass
class Region:
    """A localisation region: box, fine box and keys (host side)."""
    geocode: str
    n: tuple
    nfl: tuple
    nfu: tuple
    hgrid: tuple
    nseg_c: int
    nvctr_c: int
    keyg_c: np.ndarray
    keyv_c: np.ndarray
    nseg_f: int
    nvctr_f: int
    keyg_f: np.ndarray
    keyv_f: np.ndarray

    @property
    def ncoef(self):
        return self.nvctr_c + 7 * self.nvctr_f

    @property
    def nfine(self):
        add = 2 if self.geocode == "P" else 31
        return tuple(2 * self.n[d] + add for d in range(3))

    @property
    def n_isf(self):
        a, b, c = self.nfine
        return a * b * c


def region_from_grids(geocode, n, hgrid, coarse, fine):
    """Region from boolean coarse/fine grids of shape (n1+1,n2+1,n3+1)."""
    nseg_c, nvctr_c, kg_c, kv_c = segkeys(coarse)
    nseg_f, nvctr_f, kg_f, kv_f = segkeys(fine)
    fine = np.asarray(fine, dtype=bool)
    if nvctr_f > 0:
        idx = np.nonzero(fine)
        nfl = tuple(int(idx[d].min()) for d in range(3))
        nfu = tuple(int(idx[d].max()) for d in range(3))
    else:
        nfl, nfu = (0, 0, 0), tuple(int(x) for x in n)
    return Region(geocode, tuple(int(x) for x in n), nfl, nfu,
                  tuple(float(h) for h in hgrid), nseg_c, nvctr_c, kg_c, kv_c,
                  nseg_f, nvctr_f, kg_f, kv_f)


def dense_region(geocode, n, hgrid):
    """Every coarse point carries 1 scaling + 7 wavelet coefficients."""
    full = np.ones(tuple(int(x) + 1 for x in n), dtype=bool)
    return region_from_grids(geocode, n, hgrid, full, full)


def sphere_region(n_glob, hgrid, center, rcoarse, rfine):
    """Free-BC region of the points of a global box (0:n_glob) within rcoarse
    (coarse) and rfine (fine) of center; the local box is the bounding box of
    the coarse points.  Returns (region, nl) with nl the box origin in the
    global grid."""
    # the criterion is evaluated on the index range that can hold points
    # (one point of margin per side against rounding), which selects the same
    # points as a scan of the whole global box
    lo = [max(0, int(np.floor((center[d] - rcoarse) / hgrid[d])) - 1) for d in range(3)]
    hi = [min(int(n_glob[d]), int(np.ceil((center[d] + rcoarse) / hgrid[d])) + 1)
          for d in range(3)]
    sub = sphere_logrid(tuple(hi[d] - lo[d] for d in range(3)), hgrid, center,
                        rcoarse, origin=lo)
    idx = np.nonzero(sub)
    if idx[0].size == 0:
        raise ValueError("no coarse point within rcoarse of the centre")
    nl = tuple(int(idx[d].min()) + lo[d] for d in range(3))
    nu = tuple(int(idx[d].max()) + lo[d] for d in range(3))
    n = tuple(nu[d] - nl[d] for d in range(3))
    coarse = sub[nl[0] - lo[0]:nu[0] - lo[0] + 1, nl[1] - lo[1]:nu[1] - lo[1] + 1,
                 nl[2] - lo[2]:nu[2] - lo[2] + 1]
    fine = sphere_logrid(n, hgrid, center, rfine, origin=nl)
    fine &= coarse
    return region_from_grids("F", n, hgrid, coarse, fine), nl
