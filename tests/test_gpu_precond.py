"""Preconditioner (ocl_preconditioner_generic_k) through the C ABI vs the
oracle: FP64, within 1e-10 of max|x| per element after up to 12 CG steps
(the CG scalars are FP64 reductions in a different order than the
oracle's serial sums); ncong = 0 is bit-exact (x = S x)."""
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


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


CASES = [("P", dense_region("P", (19, 17, 21), (0.45, 0.4, 0.5))),
         ("F", mixed_sphere_region((24, 21, 26), 0.42, 13)),
         # more than 64 coarse points on x: the x line pass in its BIG form
         ("F_large", dense_region("F", (79, 20, 24), (0.4, 0.45, 0.5)))]


@pytest.mark.parametrize("seq", ["plane", "lines"])
@pytest.mark.parametrize("ncong", [0, 1, 5, 12])
@pytest.mark.parametrize("name,reg", CASES, ids=[c[0] for c in CASES])
def test_precond_vs_oracle(ctx, ctx_lines, name, reg, ncong, seq):
    """plane: x and y operator in one plane pass (k_plane_kin_pc); lines:
    the x and y line passes"""
    print(f"precond {name} n={reg.n} ncong={ncong} {seq}", flush=True)
    ctx = ctx if seq == "plane" else ctx_lines
    norb = 3
    rng = np.random.default_rng(ncong + 3)
    b = random_psi(reg, norb, 5)
    x0 = 0.05 * random_psi(reg, norb, 6)
    cpr = np.array([0.2, 0.55, 1.3])
    lr = ctx.region(reg)
    xg = dev(x0.ravel())
    ctx.precond_cg([lr] * norb, cpr, ncong, xg, dev(b.ravel()))
    got = xg.cpu().numpy().reshape(norb, -1)
    lr.close()
    for i in range(norb):
        exp = oracle.precond_cg(reg.geocode, reg.n, reg.hgrid, cpr[i], ncong, okeys(reg), x0[i],
                                b[i])
        if ncong == 0:
            assert np.array_equal(got[i], exp)
        else:
            err = np.abs(got[i] - exp).max() / np.abs(exp).max()
            print(f"  orbital {i}: {err:.2e}", flush=True)
            assert err < 1e-10
    del rng


def test_precond_batch_of_spheres(ctx):
    from bigdft_hip import workloads
    batch = workloads.sphere_batch(24, 31, n_glob=47, device="cuda")
    regs = batch["regions"]
    cache = {}
    lrs = [cache.setdefault(id(r), ctx.region(r)) for r in regs]
    rng = np.random.default_rng(2)
    cpr = rng.uniform(0.2, 1.0, len(regs))
    b = batch["psi"].clone()
    x = 0.1 * batch["psi"]
    x0 = x.cpu().numpy()
    ctx.precond_cg(lrs, cpr, 6, x, b, psi_off=batch["psi_off"])
    got = x.cpu().numpy()
    bh = b.cpu().numpy()
    assert np.array_equal(bh, batch["psi"].cpu().numpy())     # b is not modified
    for i in rng.choice(len(regs), 4, replace=False):
        r = regs[i]
        o, nc = int(batch["psi_off"][i]), int(batch["ncoef"][i])
        exp = oracle.precond_cg("F", r.n, r.hgrid, cpr[i], 6, okeys(r), x0[o:o + nc], bh[o:o + nc])
        assert np.abs(got[o:o + nc] - exp).max() < 1e-10 * np.abs(exp).max()
    for l in cache.values():
        l.close()


def test_precond_host_entry(ctx):
    reg = CASES[1][1]
    b = random_psi(reg, 2, 9)
    x = 0.02 * random_psi(reg, 2, 10)
    exp = [oracle.precond_cg("F", reg.n, reg.hgrid, c, 4, okeys(reg), x[i], b[i])
           for i, c in enumerate((0.3, 0.7))]
    lr = ctx.region(reg)
    ctx.precondition_host(lr, [0.3, 0.7], 4, x, b)
    lr.close()
    for i in range(2):
        assert np.abs(x[i] - exp[i]).max() < 1e-10 * np.abs(exp[i]).max()


def test_precond_batch_mixing_periodic_and_free(ctx):
    """one preconditioner call over orbitals of a periodic and a free region
    (per-orbital cprecr and offsets) gives each orbital the oracle's result"""
    regs = [CASES[0][1], CASES[1][1], CASES[0][1]]
    lrs_u = {id(r): ctx.region(r) for r in regs}
    lrs = [lrs_u[id(r)] for r in regs]
    bs = [random_psi(r, 1, 40 + i)[0] for i, r in enumerate(regs)]
    xs = [0.05 * random_psi(r, 1, 50 + i)[0] for i, r in enumerate(regs)]
    off = np.cumsum([0] + [b.size for b in bs])[:-1]
    cpr = np.array([0.3, 0.8, 1.1])
    x = dev(np.concatenate(xs))
    ctx.precond_cg(lrs, cpr, 5, x, dev(np.concatenate(bs)), psi_off=off)
    got = x.cpu().numpy()
    for i, r in enumerate(regs):
        exp = oracle.precond_cg(r.geocode, r.n, r.hgrid, cpr[i], 5, okeys(r), xs[i], bs[i])
        g = got[off[i]:off[i] + bs[i].size]
        assert np.abs(g - exp).max() < 1e-10 * np.abs(exp).max()
    for l in lrs_u.values():
        l.close()
