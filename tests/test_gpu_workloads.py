"""The full-size batched workloads through the C ABI.

Uneven free boxes (2048 boxes, N_d ~ U{16..40}, own potential per box) and
compressed spheres (1024 orbitals of a 96^3 global box reading one global
potential through ishift).  At these sizes the oracle checks a seeded sample
of orbitals (FP64 within 1e-12; FP32 within 2e-6 max|hpsi| per element and
1e-6 in the energies; split-BF16 within 1e-4 max|hpsi| per element); every
orbital is checked through size-independent properties:
  * <psi | hpsi> = ekin + epot (hpsi starts at zero; FP64, 1e-10 relative),
  * FP32 energies within 1e-6 relative of the FP64 ones,
  * split-FP16 forward and split-BF16 (four products) energies within 1e-6
    relative of max(|E|, 1e-12 |ekin|) per orbital.  On these workloads
    some boxes lie far from every Gaussian of the potential: their epot
    (1e-25..1e-41 Ha) comes from the far halo of the free box, where psir is
    made of the smallest filter tails (fil2 down to 1.6e-9).  Two 11-bit
    fp16 pieces cannot represent both ends of that range (5-bit exponent)
    and two bf16 pieces carry 2^-17 per value, so the plain relative error
    of such an epot is not meaningful for them; FP32 meets it everywhere.
    The number of orbitals outside the plain relative bar is printed.
"""
import numpy as np
import pytest

import oracle
from helpers import okeys, rel_err

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import bigdft_hip
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    c = bigdft_hip.Context(0)
    yield c
    c.close()


def apply(ctx, lrs, batch, precision):
    n = len(lrs)
    h = torch.zeros_like(batch["psi"])
    ek = torch.zeros(n, dtype=torch.float64, device="cuda")
    ep = torch.zeros_like(ek)
    ctx.locham_batch(lrs, batch["pot"], batch["psi"], h, ek, ep,
                     pot_off=batch["pot_off"], potdim=batch["potdim"],
                     ishift=batch["ishift"], psi_off=batch["psi_off"],
                     precision=precision)
    ctx.sync()
    return h, ek, ep


def run_workload(ctx, batch, sample, label):
    import bigdft_hip
    from bigdft_hip import workloads
    regs = batch["regions"]
    cache = {}
    lrs = []
    for r in regs:
        if id(r) not in cache:
            cache[id(r)] = ctx.region(r)
        lrs.append(cache[id(r)])
    print(f"{label}: {len(regs)} orbitals, {batch['units']} grid points", flush=True)
    h64, ek64, ep64 = apply(ctx, lrs, batch, bigdft_hip.PREC_FP64)
    h32, ek32, ep32 = apply(ctx, lrs, batch, bigdft_hip.PREC_FP32)
    hbf, ekbf, epbf = apply(ctx, lrs, batch, bigdft_hip.PREC_BF16X4)
    h16, ek16, ep16 = apply(ctx, lrs, batch, bigdft_hip.PREC_F16X3)
    # <psi|hpsi> per orbital
    seg = torch.repeat_interleave(
        torch.arange(len(regs), device="cuda"),
        torch.from_numpy(batch["ncoef"]).cuda())
    dots = torch.zeros(len(regs), dtype=torch.float64, device="cuda")
    dots.index_add_(0, seg, batch["psi"] * h64)
    tot = ek64 + ep64
    ident = float(((dots - tot).abs() / tot.abs()).max())
    def rel(a, b):
        return float(((a - b).abs() / b.abs()).max())

    def rel_scaled(a, b):
        floor = 1e-12 * ek64.abs()
        return float(((a - b).abs() / torch.maximum(b.abs(), floor)).max())

    def n_plain(a, b):
        return int((((a - b).abs() / b.abs()) > 1e-6).sum())
    f_k, f_p = rel(ek32, ek64), rel(ep32, ep64)
    print(f"{label}: identity {ident:.2e}, fp32 ekin {f_k:.2e} epot {f_p:.2e}", flush=True)
    for name, ek, ep in (("bf16x4", ekbf, epbf), ("f16x3", ek16, ep16)):
        print(f"{label}: {name} ekin {rel_scaled(ek, ek64):.2e} epot "
              f"{rel_scaled(ep, ep64):.2e} (plain relative: max {rel(ep, ep64):.2e}, "
              f"{n_plain(ep, ep64)} of {len(regs)} orbitals above 1e-6)", flush=True)
    assert ident < 1e-10
    assert f_k < 1e-6 and f_p < 1e-6
    assert rel_scaled(ek16, ek64) < 1e-6 and rel_scaled(ep16, ep64) < 1e-6
    assert rel_scaled(ekbf, ek64) < 1e-6 and rel_scaled(epbf, ep64) < 5e-5
    for b in sample:
        r = regs[b]
        o0, nc = int(batch["psi_off"][b]), int(batch["ncoef"][b])
        psi = batch["psi"][o0:o0 + nc].cpu().numpy()[None, :]
        pot = workloads.local_potential(batch, b)
        ho, eo, po = oracle.locham("F", r.n, r.nfl, r.nfu, r.hgrid, okeys(r), pot, psi)
        g = h64[o0:o0 + nc].cpu().numpy()
        assert rel_err(g, ho[0]) < 1e-12
        assert abs(float(ek64[b]) - eo[0]) < 1e-12 * abs(eo[0])
        assert abs(float(ep64[b]) - po[0]) < 1e-12 * abs(po[0])
        g32 = h32[o0:o0 + nc].cpu().numpy()
        assert np.abs(g32 - ho[0]).max() < 2e-6 * np.abs(ho[0]).max()
        gb = hbf[o0:o0 + nc].cpu().numpy()
        assert np.abs(gb - ho[0]).max() < 1e-4 * np.abs(ho[0]).max()
        g16 = h16[o0:o0 + nc].cpu().numpy()
        assert np.abs(g16 - ho[0]).max() < 1e-4 * np.abs(ho[0]).max()
    for l in cache.values():
        l.close()


def test_uneven_free_boxes_2048(ctx):
    from bigdft_hip import workloads
    batch = workloads.uneven_free(2048, 404, device="cuda")
    rng = np.random.default_rng(1)
    run_workload(ctx, batch, rng.choice(2048, 6, replace=False), "uneven free")


def test_compressed_spheres_1024(ctx):
    from bigdft_hip import workloads
    batch = workloads.sphere_batch(1024, 505, device="cuda")
    rng = np.random.default_rng(2)
    run_workload(ctx, batch, rng.choice(1024, 6, replace=False), "spheres")
