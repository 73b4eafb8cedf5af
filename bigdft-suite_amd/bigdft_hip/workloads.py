"""Synthetic batches of the measured workloads, built on a device.

* ``uneven_free``: free-BC boxes with N_d ~ U{nmin..nmax} coarse points per
  axis (seeded), dense keys, coefficients N(0,1) * exp(-r^2/32) (r = distance
  of the coefficient's coarse point from the box centre, bohr), and a
  potential per box sampled on the box's fine grid (n_d' = 2n_d+31, zero-padded
  halos included) from one 8-Gaussian field placed in a global cell.  Every
  box has its own potential array (``pot_off``, ``potdim``).
* ``sphere_batch``: orbitals of a global free box of (n_glob+1)^3 coarse points,
  each a sphere of coarse points (radius rc) with a finer sphere (radius rf)
  around a centre drawn uniformly in the box (inclusive criterion of
  num_segkeys_sphere, bigdft/src/locregs/initlocregs.f90:140-286); the local
  box is the bounding box of the coarse points.  All orbitals read one global
  potential on the global fine grid through ``ishift`` = -2*nl (the local box
  origin nl in coarse points), as apply_potential_lr places a region's
  potential (bigdft/src/orbitals/potential.f90:850-879).

* ``dense_periodic``: norb orbitals sharing one periodic box of (n+1)^3
  coarse points with dense keys, coefficients N(0,1) normalised per orbital,
  and one potential, a sum of 8 Gaussians (centres uniform in the cell,
  depth U[-5,-1] Ha, width U[0.5,2] bohr, minimum-image distance) on the
  box's fine grid.

Each returns a dict with the host regions and the device arrays the batched
entry points take, plus ``units`` = sum of the boxes' fine grid sizes.
"""
import numpy as np
import torch

from . import inputs
from .descriptors import dense_region, sphere_region


def _envelopes(regions, sigma, device):
    cache = {}
    out = []
    for r in regions:
        key = id(r)
        if key not in cache:
            cache[key] = inputs.envelope(r, sigma, device=device)
        out.append(cache[key])
    return out


def uneven_free(nbox, seed, hgrid=0.4, nmin=16, nmax=40, cell=40.0, device="cuda"):
    rng = np.random.default_rng(seed)
    Ns = rng.integers(nmin, nmax + 1, size=(nbox, 3))
    origins = rng.uniform(0.0, cell, size=(nbox, 3))
    h = (hgrid, hgrid, hgrid)
    shapes = {}
    regions = []
    for b in range(nbox):
        key = tuple(int(x) for x in Ns[b])
        if key not in shapes:
            shapes[key] = dense_region("F", tuple(x - 1 for x in key), h)
        regions.append(shapes[key])
    ncoef = np.array([r.ncoef for r in regions], dtype=np.int64)
    psi_off = np.concatenate([[0], np.cumsum(ncoef)[:-1]]).astype(np.int64)
    psi = inputs.coefficients(1, int(ncoef.sum()), seed + 2, device=device)[0]
    env = _envelopes(regions, 4.0, device)
    for b in range(nbox):
        psi[psi_off[b]:psi_off[b] + ncoef[b]] *= env[b]
    nisf = np.array([r.n_isf for r in regions], dtype=np.int64)
    pot_off = np.concatenate([[0], np.cumsum(nisf)[:-1]]).astype(np.int64)
    pot = torch.empty(int(nisf.sum()), dtype=torch.float64, device=device)
    gp = inputs.gaussian_params(seed + 1, [cell] * 3)
    for b in range(nbox):
        r = regions[b]
        xyz = inputs.fine_coords("F", r.n, r.hgrid, origin=tuple(origins[b]),
                                 device=device)
        pot[pot_off[b]:pot_off[b] + nisf[b]] = inputs.gaussian_field(gp, xyz)
    potdim = np.array([r.nfine for r in regions], dtype=np.int32).ravel()
    return dict(regions=regions, psi=psi, psi_off=psi_off, ncoef=ncoef, pot=pot,
                pot_off=pot_off, potdim=potdim, ishift=None,
                units=int(nisf.sum()), origins=origins)


def sphere_batch(norb, seed, n_glob=95, hgrid=0.4, rc=7.0, rf=3.5, device="cuda"):
    rng = np.random.default_rng(seed)
    h = (hgrid, hgrid, hgrid)
    ng = (n_glob, n_glob, n_glob)
    regions, nls = [], []
    for _ in range(norb):
        c = rng.uniform(0.0, n_glob * hgrid, size=3)
        reg, nl = sphere_region(ng, h, tuple(c), rc, rf)
        regions.append(reg)
        nls.append(nl)
    ncoef = np.array([r.ncoef for r in regions], dtype=np.int64)
    psi_off = np.concatenate([[0], np.cumsum(ncoef)[:-1]]).astype(np.int64)
    psi = inputs.coefficients(1, int(ncoef.sum()), seed + 2, device=device)[0]
    for b in range(norb):
        seg = psi[psi_off[b]:psi_off[b] + ncoef[b]]
        seg /= torch.linalg.vector_norm(seg)
    cell = [(n_glob + 1) * hgrid] * 3
    gp = inputs.gaussian_params(seed + 1, cell)
    xyz = inputs.fine_coords("F", ng, h, device=device)
    pot = inputs.gaussian_field(gp, xyz)
    gdim = tuple(2 * n_glob + 31 for _ in range(3))
    potdim = np.array([gdim] * norb, dtype=np.int32).ravel()
    ishift = np.array([[-2 * x for x in nl] for nl in nls], dtype=np.int32).ravel()
    units = int(sum(r.n_isf for r in regions))
    return dict(regions=regions, psi=psi, psi_off=psi_off, ncoef=ncoef, pot=pot,
                pot_off=np.zeros(norb, dtype=np.int64), potdim=potdim,
                ishift=ishift, units=units, nl=nls, global_nfine=gdim)


def local_potential(batch, b):
    """The potential box of orbital b as the oracle takes it (host, the
    region's own fine grid; zero where the shifted window leaves the array)."""
    r = batch["regions"][b]
    pd = batch["potdim"][3 * b:3 * b + 3]
    p0 = int(batch["pot_off"][b])
    pot = batch["pot"][p0:p0 + int(np.prod(pd))].cpu().numpy().reshape(pd[::-1])
    sh = (0, 0, 0) if batch["ishift"] is None else batch["ishift"][3 * b:3 * b + 3]
    nf = r.nfine
    out = np.zeros(nf[::-1])
    src = []
    dst = []
    for d in range(3):
        q0 = -int(sh[d])             # potential index of local point 0
        a = max(0, -q0)
        e = min(nf[d], int(pd[d]) - q0)
        dst.append(slice(a, max(a, e)))
        src.append(slice(a + q0, max(a, e) + q0))
    out[dst[2], dst[1], dst[0]] = pot[src[2], src[1], src[0]]
    return out.ravel()


def dense_periodic(n, norb, seed, hgrid=0.35, device="cuda"):
    h = (hgrid, hgrid, hgrid)
    reg = dense_region("P", (n, n, n), h)
    psi = inputs.coefficients(norb, reg.ncoef, seed, device=device,
                              normalise=True).reshape(-1)
    cell = [(n + 1) * hgrid] * 3
    gp = inputs.gaussian_params(seed + 1, cell)
    pot = inputs.gaussian_field(gp, inputs.fine_coords("P", (n, n, n), h, device=device),
                                cell=cell).contiguous()
    ncoef = np.full(norb, reg.ncoef, dtype=np.int64)
    return dict(regions=[reg] * norb, psi=psi, ncoef=ncoef,
                psi_off=np.arange(norb, dtype=np.int64) * reg.ncoef, pot=pot,
                pot_off=np.zeros(norb, dtype=np.int64),
                potdim=np.array(list(reg.nfine) * norb, dtype=np.int32), ishift=None,
                units=reg.n_isf * norb)
