"""Density through the C ABI vs the oracle (daub_to_isf + partial_density).

Bars: FP64 within 1e-12 of max(rho) per point and 1e-12 relative in the
integral; FP32 within 1e-6 of max(rho) per point.  Spin components,
accumulation onto an existing rho (the OpenCL layer's 1e-20 prefill) and
the kwgt*occup/volume_element weighting of the high-level call are checked.
"""
import numpy as np
import pytest

import oracle
from helpers import dense_region, mixed_sphere_region, okeys, random_psi

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import bigdft_hip
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    c = bigdft_hip.Context(0)
    yield c
    c.close()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


CASES = [("P", dense_region("P", (23, 19, 21), (0.45, 0.4, 0.5))),
         ("F", dense_region("F", (15, 13, 17), (0.4, 0.45, 0.5))),
         ("sphere", mixed_sphere_region((24, 22, 26), 0.4, 31)),
         ("P47", dense_region("P", (47, 47, 47), (0.35,) * 3)),
         # z lines of 211 and 260 fine points: the z kernel works in rounds
         ("F_large", dense_region("F", (30, 24, 90), (0.4, 0.45, 0.5))),
         ("P_large", dense_region("P", (37, 45, 129), (0.4, 0.45, 0.5)))]


@pytest.mark.parametrize("name,reg", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32"])
def test_density_vs_oracle(ctx, name, reg, precision):
    norb = 5
    print(f"density {name} n={reg.n} precision={precision}", flush=True)
    psi = random_psi(reg, norb, 77)
    hf = np.array([2.0, 1.0, 0.5, 2.0, 0.0]) * 3.7
    sp = np.array([0, 1, 0, 1, 1], dtype=np.int32)
    ro = oracle.density(reg.geocode, reg.n, reg.nfl, reg.nfu, okeys(reg), psi, hf,
                        sp, nspin=2)
    lr = ctx.region(reg)
    rho = dev(np.full(ro.shape, 1e-20))
    ctx.density(lr, dev(psi), hf, rho, spin=sp, nspin=2, precision=precision)
    rg = rho.cpu().numpy() - 1e-20
    lr.close()
    tol = 1e-12 if precision == 0 else 1e-6
    for s in (0, 1):
        err = np.abs(rg[s] - ro[s]).max() / np.abs(ro[s]).max()
        ie = abs(rg[s].sum() - ro[s].sum()) / abs(ro[s].sum())
        print(f"  spin {s}: max err {err:.2e}, integral {ie:.2e}", flush=True)
        assert err < tol and ie < tol


def test_density_accumulates_and_weights(ctx):
    import bigdft_hip
    reg = dense_region("P", (19, 19, 19), (0.4,) * 3)
    psi = random_psi(reg, 3, 88)
    lr = ctx.region(reg)
    nf = reg.n_isf
    occ = np.array([2.0, 2.0, 1.0])
    rho = torch.zeros((1, nf), dtype=torch.float64, device="cuda")
    bigdft_hip.local_partial_density(ctx, lr, occ, dev(psi), rho, kwgt=0.5)
    r1 = rho.cpu().numpy().copy()
    vol = 0.2 ** 3
    ro = oracle.density("P", reg.n, reg.nfl, reg.nfu, okeys(reg), psi, 0.5 * occ / vol)
    assert np.abs(r1 - ro).max() < 1e-12 * np.abs(ro).max()
    # a second call adds the same density again
    bigdft_hip.local_partial_density(ctx, lr, occ, dev(psi), rho, kwgt=0.5)
    assert np.abs(rho.cpu().numpy() - 2 * ro).max() < 1e-12 * np.abs(ro).max()
    # spin-down orbitals need two components
    with pytest.raises(ValueError):
        bigdft_hip.local_partial_density(ctx, lr, occ, dev(psi), rho,
                                         spinsgn=[1, -1, 1], nspin=1)
    lr.close()


def test_density_argument_errors(ctx):
    import bigdft_hip
    reg = dense_region("P", (7, 7, 7), (0.4,) * 3)
    lr = ctx.region(reg)
    psi = dev(random_psi(reg, 2, 1))
    rho = torch.zeros((2, reg.n_isf), dtype=torch.float64, device="cuda")
    with pytest.raises(bigdft_hip.BigDFTError, match="ARG"):
        ctx.density(lr, psi, [1.0, 1.0], rho, spin=[0, 2], nspin=2)
    with pytest.raises(bigdft_hip.BigDFTError, match="UNSUPPORTED"):
        ctx.density(lr, psi, [1.0, 1.0], rho, precision=bigdft_hip.PREC_BF16X4)
    rho3 = torch.zeros((3, reg.n_isf), dtype=torch.float64, device="cuda")
    with pytest.raises(bigdft_hip.BigDFTError, match="ARG"):
        ctx.density(lr, psi, [1.0, 1.0], rho3, nspin=3)
    with pytest.raises(ValueError, match="rho"):   # too short for nspin = 3
        ctx.density(lr, psi, [1.0, 1.0], rho, nspin=3)
    lr.close()


def test_density_reuses_its_plan_across_calls(ctx):
    """repeated calls on one region keep the plan: new orbital arrays,
    another orbital count and another precision in between give the
    oracle's densities each time"""
    reg = dense_region("F", (13, 15, 11), (0.45, 0.4, 0.5))
    lr = ctx.region(reg)
    ok = okeys(reg)
    for k, (norb, prec) in enumerate([(4, 0), (4, 0), (2, 0), (4, 1), (4, 0)]):
        psi = random_psi(reg, norb, 300 + k)
        hf = np.linspace(0.5, 2.0, norb)
        ro = oracle.density(reg.geocode, reg.n, reg.nfl, reg.nfu, ok, psi, hf)
        rho = torch.zeros((1, reg.n_isf), dtype=torch.float64, device="cuda")
        ctx.density(lr, dev(psi), hf, rho, precision=prec)
        err = np.abs(rho.cpu().numpy() - ro).max() / np.abs(ro).max()
        print(f"  call {k}: norb {norb} precision {prec}: {err:.2e}", flush=True)
        assert err < (1e-12 if prec == 0 else 1e-6)
    lr.close()


@pytest.fixture(scope="module")
def ctx_lines():
    import os

    import bigdft_hip
    os.environ["BDFT_LINE_PASSES"] = "1"
    try:
        c = bigdft_hip.Context(0)
    finally:
        del os.environ["BDFT_LINE_PASSES"]
    yield c
    c.close()


@pytest.mark.parametrize("name,reg", CASES[:3], ids=[c[0] for c in CASES[:3]])
def test_density_line_and_plane_grow_agree(ctx, ctx_lines, name, reg):
    """the plane grow (default) and the x/y line passes give the oracle's
    density on the same orbitals"""
    psi = random_psi(reg, 3, 91)
    hf = np.array([2.0, 1.5, 0.5])
    ro = oracle.density(reg.geocode, reg.n, reg.nfl, reg.nfu, okeys(reg), psi, hf)
    for c in (ctx, ctx_lines):
        lr = c.region(reg)
        rho = torch.zeros((1, reg.n_isf), dtype=torch.float64, device="cuda")
        c.density(lr, dev(psi), hf, rho)
        lr.close()
        err = np.abs(rho.cpu().numpy() - ro).max() / np.abs(ro).max()
        print(f"  {name}: {err:.2e}", flush=True)
        assert err < 1e-12
