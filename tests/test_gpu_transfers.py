"""Region transfers through the C ABI vs the oracle: bit-exact."""
import numpy as np
import pytest

import oracle
from helpers import okeys
from test_oracle_transfers import setup

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


def test_transfers_batch_bit_exact(ctx):
    cases = [setup(s) for s in (1, 2, 3)]
    glob = cases[0][0]
    ng = cases[0][2]
    g = ctx.region(glob)
    locs = [ctx.region(c[1]) for c in cases]
    ns = np.array([c[4] for c in cases], dtype=np.int32)
    rng = np.random.default_rng(7)
    gpsi = rng.standard_normal((3, glob.ncoef))
    lnc = [c[1].ncoef for c in cases]
    loff = np.concatenate([[0], np.cumsum(lnc)[:-1]])
    lpsi = torch.full((sum(lnc),), np.nan, dtype=torch.float64, device="cuda")
    ctx.psi_to_locreg(g, locs, ns, dev(gpsi), lpsi)
    got = lpsi.cpu().numpy()
    for i, (gl, lo, _, nl, nsi) in enumerate(cases):
        exp = oracle.psi_to_locreg2(nl, ng, nsi, okeys(lo), okeys(gl), gpsi[i])
        assert np.array_equal(got[loff[i]:loff[i] + lnc[i]], exp)
    # local -> global onto a global array holding other values
    lsrc = rng.standard_normal(sum(lnc))
    base = rng.standard_normal((3, glob.ncoef))
    gout = dev(base)
    ctx.locreg_to_psi(g, locs, ns, dev(lsrc), gout)
    got = gout.cpu().numpy()
    for i, (gl, lo, _, nl, nsi) in enumerate(cases):
        exp, n = oracle.lpsi_to_global2(nl, ng, nsi, okeys(lo), okeys(gl),
                                        lsrc[loff[i]:loff[i] + lnc[i]], base[i])
        assert n == lo.nvctr_c + lo.nvctr_f
        assert np.array_equal(got[i], exp)
    for l in locs:
        l.close()
    g.close()


def test_locreg_to_psi_rejects_points_outside(ctx):
    import bigdft_hip
    from bigdft_hip.descriptors import region_from_grids
    glob, loc, ng, nl, ns = setup(4)
    full = np.ones(tuple(x + 1 for x in nl), dtype=bool)
    loc2 = region_from_grids("F", nl, loc.hgrid, full, np.zeros_like(full))
    g, l = ctx.region(glob), ctx.region(loc2)
    gpsi = torch.zeros(glob.ncoef, dtype=torch.float64, device="cuda")
    lp = dev(np.ones(loc2.ncoef))
    with pytest.raises(bigdft_hip.BigDFTError, match="DESC"):
        ctx.locreg_to_psi(g, [l], np.array(ns, dtype=np.int32), lp, gpsi)
    # the other direction fills the missing points with zeros
    out = torch.full((loc2.ncoef,), np.nan, dtype=torch.float64, device="cuda")
    ctx.psi_to_locreg(g, [l], np.array(ns, dtype=np.int32),
                      dev(np.arange(glob.ncoef, dtype=np.float64)), out)
    assert not torch.isnan(out).any()
    l.close()
    g.close()


def test_transfers_sphere_batch_round_trip(ctx):
    """48 compressed spheres of a 64^3 free global box (tiles of several
    thousand coefficients, coarse and fine parts of many sizes): local
    orbitals equal the oracle's psi_to_locreg2, and writing them back into
    a zero global orbital restores exactly the points each sphere holds"""
    from bigdft_hip import workloads
    from bigdft_hip.descriptors import dense_region
    sb = workloads.sphere_batch(48, 5, n_glob=63, device="cuda")
    regs = sb["regions"]
    ng = (63, 63, 63)
    glob = dense_region("F", ng, (0.4,) * 3)
    g = ctx.region(glob)
    cache = {}
    lrs = [cache.setdefault(id(r), ctx.region(r)) for r in regs]
    ns = np.array(sb["nl"], dtype=np.int32).reshape(-1, 3)
    rng = np.random.default_rng(8)
    gh = rng.standard_normal((len(regs), glob.ncoef))
    goff = np.arange(len(regs), dtype=np.int64) * glob.ncoef
    loff = sb["psi_off"]
    lp = torch.full_like(sb["psi"], np.nan)
    ctx.psi_to_locreg(g, lrs, ns, dev(gh.ravel()), lp, goff=goff, loff=loff)
    lh = lp.cpu().numpy()
    gk = okeys(glob)
    for i in rng.choice(len(regs), 5, replace=False):
        r = regs[i]
        exp = oracle.psi_to_locreg2(r.n, ng, sb["nl"][i], okeys(r), gk, gh[i])
        assert np.array_equal(lh[loff[i]:loff[i] + r.ncoef], exp)
    back = torch.zeros(len(regs) * glob.ncoef, dtype=torch.float64, device="cuda")
    ctx.locreg_to_psi(g, lrs, ns, lp, back, loff=loff, goff=goff)
    bh = back.cpu().numpy().reshape(len(regs), -1)
    for i in range(len(regs)):
        nz = bh[i] != 0.0
        assert np.array_equal(bh[i][nz], gh[i][nz])
        assert nz.sum() == regs[i].ncoef
    # the same batch again (descriptors reused) on other values
    lp2 = torch.full_like(sb["psi"], np.nan)
    ctx.psi_to_locreg(g, lrs, ns, dev(2.0 * gh.ravel()), lp2, goff=goff, loff=loff)
    assert np.array_equal(lp2.cpu().numpy(), 2.0 * lh)
    for l in cache.values():
        l.close()
    g.close()
