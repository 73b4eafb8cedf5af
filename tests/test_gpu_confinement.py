"""Local Hamiltonian with a confining potential per orbital through the C
ABI (bdft_locham_conf_batch_d) vs the oracle's apply_potential_lr_conf*.

Bars as for the plain Hamiltonian: FP64 within 1e-12 of max|hpsi| per
element and in the energies, FP32 within 2e-6 of max|hpsi| and 1e-6 in the
energies, split-BF16 / split-FP16 within 1e-4 of max|hpsi|.  The batched
case places compressed spheres in a 64^3 global box, reads the global
potential through ishift and puts the confinement centre at the orbital's
own centre in global coordinates (ioffset = -ishift - 15).
"""
import numpy as np
import pytest

import oracle
from helpers import dense_region, mixed_sphere_region, okeys, random_pot, random_psi

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


BARS = {0: 1e-12, 1: 2e-6, 3: 1e-4, 4: 1e-4}


def apply(ctx, lrs, pot, psi, precision, conf, **kw):
    n = len(lrs)
    h = torch.zeros_like(psi)
    ek = torch.zeros(n, dtype=torch.float64, device="cuda")
    ep = torch.zeros_like(ek)
    ctx.locham_batch(lrs, pot, psi, h, ek, ep, precision=precision, conf=conf, **kw)
    ctx.sync()
    return h.cpu().numpy(), ek.cpu().numpy(), ep.cpu().numpy()


CASES = [("P", dense_region("P", (15, 13, 17), (0.45, 0.4, 0.5)), -1),
         ("F", mixed_sphere_region((22, 19, 24), 0.42, 13), -15)]


@pytest.mark.parametrize("precision,zpot", [(0, 0), (1, 0), (3, 0), (4, 0), (3, 1), (4, 1)],
                         ids=["fp64", "fp32", "bf16x4", "f16x3", "bf16x4-mzp", "f16x3-mzp"])
@pytest.mark.parametrize("name,reg,off", CASES, ids=[c[0] for c in CASES])
def test_confinement_vs_oracle(ctx, name, reg, off, precision, zpot):
    """zpot 1: the matrix-core z potential and backward passes (k_mzp,
    k_mbwd), 0: the automatic choice of units"""
    import bigdft_hip
    print(f"confinement {name} n={reg.n} precision={precision} zpot={zpot}", flush=True)
    ctx.set_option("zpot", zpot)
    ctx.set_option("bwd", zpot)
    norb = 3
    h = reg.hgrid
    pot = random_pot(reg, 3, lo=-1.0, hi=0.0)
    psi = random_psi(reg, norb, 17)
    rc = [[(0.3 + 0.2 * k) * reg.n[d] * h[d] for d in range(3)] for k in range(norb)]
    orders = [4, 2, 0]          # the last orbital has no confinement
    conf = np.concatenate([bigdft_hip.confpot(orders[k], 2e-3 * (k + 1), [x / 2 for x in h],
                                              rc[k], [off] * 3) for k in range(norb)])
    oconf = np.concatenate([oracle.confpot(orders[k], 2e-3 * (k + 1), [x / 2 for x in h],
                                           rc[k], [off] * 3) for k in range(norb)])
    ho, eko, epo = oracle.locham(reg.geocode, reg.n, reg.nfl, reg.nfu, h, okeys(reg), pot, psi,
                                 conf=oconf)
    lr = ctx.region(reg)
    hg, ekg, epg = apply(ctx, [lr] * norb, dev(pot), dev(psi), precision, conf)
    lr.close()
    ctx.set_option("zpot", 0)
    ctx.set_option("bwd", 0)
    err = np.abs(hg - ho).max() / np.abs(ho).max()
    ee = max(np.abs(ekg - eko).max() / np.abs(eko).max(), np.abs(epg - epo).max() / np.abs(epo).max())
    print(f"  hpsi {err:.2e} energies {ee:.2e}", flush=True)
    assert err < BARS[precision]
    if precision <= 1:
        assert ee < (1e-12 if precision == 0 else 1e-6)


def test_confinement_batch_global_potential(ctx):
    import bigdft_hip
    from bigdft_hip import workloads
    batch = workloads.sphere_batch(48, 77, n_glob=63, device="cuda")
    regs = batch["regions"]
    cache = {}
    lrs = [cache.setdefault(id(r), ctx.region(r)) for r in regs]
    ish = batch["ishift"].reshape(-1, 3)
    rng = np.random.default_rng(4)
    conf = np.concatenate([bigdft_hip.confpot(4, rng.uniform(1e-3, 1e-2), (0.2,) * 3,
                                              rng.uniform(5, 20, 3), -ish[b] - 15)
                           for b in range(len(regs))])
    kw = dict(pot_off=batch["pot_off"], potdim=batch["potdim"], ishift=batch["ishift"],
              psi_off=batch["psi_off"])
    h64, ek64, ep64 = apply(ctx, lrs, batch["pot"], batch["psi"], 0, conf, **kw)
    _, _, ep0 = apply(ctx, lrs, batch["pot"], batch["psi"], 0, None, **kw)
    assert np.all(ep64 > ep0)      # a positive confinement raises every epot
    # <psi|hpsi> = ekin + epot per orbital
    psi = batch["psi"].cpu().numpy()
    for b in range(len(regs)):
        o, nc = int(batch["psi_off"][b]), int(batch["ncoef"][b])
        d = float(psi[o:o + nc] @ h64[o:o + nc])
        assert abs(d - ek64[b] - ep64[b]) < 1e-10 * abs(ek64[b] + ep64[b])
    for b in rng.choice(len(regs), 4, replace=False):
        r = regs[b]
        o, nc = int(batch["psi_off"][b]), int(batch["ncoef"][b])
        c = oracle.confpot(4, conf["prefac"][b], (0.2,) * 3, conf["rxyzConf"][b],
                           conf["ioffset"][b])
        ho, eo, po = oracle.locham("F", r.n, r.nfl, r.nfu, r.hgrid, okeys(r),
                                   workloads.local_potential(batch, b), psi[None, o:o + nc],
                                   conf=c)
        assert np.abs(h64[o:o + nc] - ho[0]).max() < 1e-12 * np.abs(ho[0]).max()
        assert abs(ep64[b] - po[0]) < 1e-12 * abs(po[0]) and abs(ek64[b] - eo[0]) < 1e-12 * abs(eo[0])
    # the F16X3 path on the same batch
    h16, ek16, ep16 = apply(ctx, lrs, batch["pot"], batch["psi"], 4, conf, **kw)
    assert np.abs(h16 - h64).max() < 1e-4 * np.abs(h64).max()
    assert np.abs(ep16 - ep64).max() < 1e-5 * np.abs(ep64).max()
    for l in cache.values():
        l.close()


def test_confinement_changes_invalidate_the_cached_plan(ctx):
    import bigdft_hip
    reg = dense_region("P", (11, 11, 11), (0.4,) * 3)
    lr = ctx.region(reg)
    pot, psi = dev(random_pot(reg, 1)), dev(random_psi(reg, 1, 2))
    c1 = bigdft_hip.confpot(2, 1e-2, (0.2,) * 3, (1.0, 1.0, 1.0), (-1,) * 3)
    c2 = bigdft_hip.confpot(2, 2e-2, (0.2,) * 3, (1.0, 1.0, 1.0), (-1,) * 3)
    _, _, e0 = apply(ctx, [lr], pot, psi, 0, None)
    _, _, e1 = apply(ctx, [lr], pot, psi, 0, c1)
    _, _, e2 = apply(ctx, [lr], pot, psi, 0, c2)
    _, _, e3 = apply(ctx, [lr], pot, psi, 0, None)
    # epot is linear in the confinement prefactor
    assert abs((e2[0] - e0[0]) - 2 * (e1[0] - e0[0])) < 1e-10 * abs(e2[0])
    assert e3[0] == e0[0] and e1[0] != e0[0]
    with pytest.raises(ValueError):
        apply(ctx, [lr], pot, psi, 0, np.concatenate([c1, c2]))
    lr.close()
