"""Seeded synthetic inputs for the local-Hamiltonian workloads.

All generators take a torch.Generator-compatible seed and a device, so the
same call gives the same numbers on the same device type.  Positions: the
fine (ISF) point p of a periodic box sits at p*h/2, of a free box at
(p-14)*h/2 (fine index -14..2n+16, bigdft/src/orbitals/growshrink.f90:12-37),
both relative to the box origin.
"""
import math

import numpy as np
import torch


def gen(seed, device="cpu"):
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    return g


def fine_coords(geocode, n, hgrid, origin=(0.0, 0.0, 0.0), device="cpu"):
    """1D coordinate arrays (bohr) of the fine points along each axis."""
    out = []
    for d in range(3):
        if geocode == "P":
            p = torch.arange(2 * n[d] + 2, dtype=torch.float64, device=device)
        else:
            p = torch.arange(2 * n[d] + 31, dtype=torch.float64,
                             device=device) - 14.0
        out.append(origin[d] + p * (0.5 * hgrid[d]))
    return out


def gaussian_params(seed, cell, ngauss=8):
    """8-Gaussian potential: centres uniform in the cell, depth U[-5,-1] Ha,
    width U[0.5,2] bohr (numpy, host)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, 1.0, size=(ngauss, 3)) * np.asarray(cell)
    depth = rng.uniform(-5.0, -1.0, size=ngauss)
    width = rng.uniform(0.5, 2.0, size=ngauss)
    return c, depth, width


def gaussian_field(params, coords, cell=None):
    """V(x,y,z) = sum_i depth_i exp(-|r-c_i|^2/(2 w_i^2)) on the tensor grid
    of coords (Fortran order flattened, x fastest).  With cell given, the
    minimum-image distance is used (periodic)."""
    c, depth, width = params
    x, y, z = coords
    dev = x.device
    V = torch.zeros((z.numel(), y.numel(), x.numel()), dtype=torch.float64,
                    device=dev)
    for i in range(len(depth)):
        d2 = []
        for ax, a in enumerate((x, y, z)):
            t = a - float(c[i][ax])
            if cell is not None:
                L = float(cell[ax])
                t = t - L * torch.round(t / L)
            d2.append(t * t)
        e = torch.exp(-d2[0] / (2 * width[i] ** 2))
        f = torch.exp(-d2[1] / (2 * width[i] ** 2))
        g = torch.exp(-d2[2] / (2 * width[i] ** 2))
        V += float(depth[i]) * g[:, None, None] * f[None, :, None] * e[None, None, :]
    return V.reshape(-1)


def uniform_potential(nfine, lo, hi, seed, device="cpu"):
    """V ~ U[lo,hi] on the fine box (x fastest)."""
    g = gen(seed, device)
    n = nfine[0] * nfine[1] * nfine[2]
    return lo + (hi - lo) * torch.rand(n, dtype=torch.float64, generator=g,
                                       device=device)


def coefficients(norb, ncoef, seed, device="cpu", normalise=False):
    """norb x ncoef coefficients drawn N(0,1); optionally unit 2-norm."""
    g = gen(seed, device)
    psi = torch.randn((norb, ncoef), dtype=torch.float64, generator=g,
                      device=device)
    if normalise:
        psi /= torch.linalg.vector_norm(psi, dim=1, keepdim=True)
    return psi


def envelope(region, sigma, device="cpu"):
    """exp(-r^2/(2 sigma^2)) per coefficient, r = distance (bohr) of the
    coefficient's coarse point from the box centre."""
    from .descriptors import segkeys  # noqa: F401  (keys already in region)
    n, h = region.n, region.hgrid
    ctr = [0.5 * n[d] * h[d] for d in range(3)]

    def pts(keyg, nseg):
        lin = []
        for s in range(nseg):
            j0, j1 = int(keyg[2 * s]), int(keyg[2 * s + 1])
            lin.append(np.arange(j0 - 1, j1))
        lin = np.concatenate(lin) if lin else np.zeros(0, dtype=np.int64)
        m1, m2 = n[0] + 1, n[1] + 1
        i1 = lin % m1
        i2 = (lin // m1) % m2
        i3 = lin // (m1 * m2)
        r2 = ((i1 * h[0] - ctr[0]) ** 2 + (i2 * h[1] - ctr[1]) ** 2 +
              (i3 * h[2] - ctr[2]) ** 2)
        return np.exp(-r2 / (2 * sigma * sigma))
    ec = pts(region.keyg_c, region.nseg_c)
    ef = np.repeat(pts(region.keyg_f, region.nseg_f), 7)
    return torch.from_numpy(np.concatenate([ec, ef])).to(device)
