"""bdft_locham_batch_fixed_d (no device allocation, one fixed device array
per GPU) against bdft_locham_batch_d on the same inputs, through the C ABI:
two calls in a row, the second on a smaller batch, bit for bit (the same
kernels see the same boxes), in FP64 and in the split-FP16 precision; a
batch split into several chunks of the fixed array; an orbital too large
for the array refused before anything runs."""
import numpy as np
import pytest

from helpers import dense_region, random_psi

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import bigdft_hip
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    c = bigdft_hip.Context(0)
    yield c
    c.close()


def make_batch(regs, seed):
    """psi of every orbital one after the other, each box's own potential,
    hpsi starting from non-zero values (the entry points accumulate)"""
    rng = np.random.default_rng(seed)
    psi = np.concatenate([random_psi(r, 1, seed + i)[0] for i, r in enumerate(regs)])
    pots = [rng.uniform(-2.0, 0.0, size=r.n_isf) for r in regs]
    pot_off = np.cumsum([0] + [p.size for p in pots[:-1]]).astype(np.int64)
    h0 = rng.standard_normal(psi.size)
    return (torch.from_numpy(np.concatenate(pots)).cuda(), torch.from_numpy(psi).cuda(),
            torch.from_numpy(h0).cuda(), pot_off)


def both(ctx, lrs, batch, precision):
    pot, psi, h0, pot_off = batch
    n = len(lrs)
    out = []
    for fixed in (False, True):
        h = h0.clone()
        ek = torch.zeros(n, dtype=torch.float64, device="cuda")
        ep = torch.zeros_like(ek)
        f = ctx.locham_batch_fixed if fixed else ctx.locham_batch
        print(f"  {'fixed' if fixed else 'allocating'} entry point, {n} orbitals, "
              f"precision {precision}", flush=True)
        f(lrs, pot, psi, h, ek, ep, pot_off=pot_off, precision=precision)
        ctx.sync()
        torch.cuda.synchronize()
        out.append((h.cpu().numpy(), ek.cpu().numpy(), ep.cpu().numpy()))
    return out


@pytest.mark.parametrize("precision", [0, 4])
def test_fixed_matches_allocating_twice(ctx, precision):
    shapes = [("F", (17, 22, 19)), ("F", (25, 16, 30)), ("P", (23, 23, 23)),
              ("F", (31, 27, 20)), ("F", (15, 15, 15)), ("P", (19, 21, 25))]
    regs = [dense_region(g, n, (0.4, 0.4, 0.4)) for g, n in shapes]
    lrs_all = [ctx.region(r) for r in regs]
    for n in (6, 3):   # the second call on a smaller batch
        lrs = lrs_all[:n]
        batch = make_batch(regs[:n], 100 + n)
        (ha, eka, epa), (hf, ekf, epf) = both(ctx, lrs, batch, precision)
        assert np.array_equal(ha, hf)
        assert np.array_equal(eka, ekf) and np.array_equal(epa, epf)
        assert np.abs(ha - batch[2].cpu().numpy()).max() > 0   # something was added
    for l in lrs_all:
        l.close()


@pytest.mark.parametrize("precision", [0, 4])
def test_fixed_in_several_chunks(ctx, precision):
    """periodic 48^3 boxes take 21 MB of working grids each in FP64 (half in
    F16X3), so 14 orbitals of them and of free boxes run in several chunks
    of the fixed array; every chunk runs the kernels the whole batch would
    (the choices that depend on the largest box are the batch's)"""
    import bigdft_hip
    regs = [dense_region("P", (47, 47, 47), (0.35, 0.35, 0.35)),
            dense_region("F", (30, 25, 28), (0.4, 0.4, 0.4))]
    lrs_all = [ctx.region(r) for r in regs]
    for norb in (14, 5):
        rs = [regs[i % 2] for i in range(norb)]
        lrs = [lrs_all[i % 2] for i in range(norb)]
        batch = make_batch(rs, 7 + norb)
        (ha, eka, epa), (hf, ekf, epf) = both(ctx, lrs, batch, precision)
        assert np.array_equal(ha, hf)
        assert np.array_equal(eka, ekf) and np.array_equal(epa, epf)
    for l in lrs_all:
        l.close()


def test_fixed_refuses_an_orbital_larger_than_the_array(ctx):
    import bigdft_hip
    reg = dense_region("P", (95, 95, 95), (0.35, 0.35, 0.35))
    lr = ctx.region(reg)
    pot = torch.zeros(reg.n_isf, dtype=torch.float64, device="cuda")
    psi = torch.zeros(reg.ncoef, dtype=torch.float64, device="cuda")
    h = torch.zeros_like(psi)
    ek = torch.zeros(1, dtype=torch.float64, device="cuda")
    with pytest.raises(bigdft_hip.BigDFTError) as e:
        ctx.locham_batch_fixed([lr], pot, psi, h, ek, ek.clone(), precision=0)
    assert e.value.code == 5 and "fixed device array" in str(e.value)
    assert float(h.abs().max()) == 0.0
    lr.close()
