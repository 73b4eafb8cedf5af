"""Launch options and the autotuner on the GPU: every option set gives the
oracle's Hamiltonian (FP64 within 1e-12 of max|hpsi| and in the energies,
F16X3 within 2e-6 of max|hpsi| and 1e-6 in the energies), and a tuned
context keeps its choice in the cache file and produces the same result."""
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


CASES = [("P", dense_region("P", (23, 19, 21), (0.45, 0.4, 0.5))),
         ("F", mixed_sphere_region((24, 21, 26), 0.42, 13))]


def run(ctx, lr, norb, pot, psi, precision):
    h = torch.zeros_like(psi)
    ek = torch.zeros(norb, dtype=torch.float64, device="cuda")
    ep = torch.zeros_like(ek)
    ctx.locham_batch([lr] * norb, pot, psi, h, ek, ep, precision=precision)
    return h.cpu().numpy(), ek.cpu().numpy(), ep.cpu().numpy()


@pytest.mark.parametrize("name,reg", CASES, ids=[c[0] for c in CASES])
def test_every_option_set_matches_the_oracle(ctx, name, reg):
    import bigdft_hip
    from bigdft_hip import autotune
    norb = 3
    pot = random_pot(reg, 21)
    psi = random_psi(reg, norb, 22)
    ho, eko, epo = oracle.locham(reg.geocode, reg.n, reg.nfl, reg.nfu, reg.hgrid, okeys(reg),
                                 pot, psi)
    lr = ctx.region(reg)
    dp, ds = dev(pot), dev(psi.ravel())
    for prec, bar, ebar in ((bigdft_hip.PREC_FP64, 1e-12, 1e-12),
                            (bigdft_hip.PREC_F16X3, 2e-6, 1e-6)):
        for opts in autotune.candidates(prec):
            autotune.apply(ctx, opts)
            assert ctx.get_option("line_passes") == opts["line_passes"]
            h, ek, ep = run(ctx, lr, norb, dp, ds, prec)
            err = np.abs(h.reshape(norb, -1) - ho).max() / np.abs(ho).max()
            ee = max(np.abs(ek - eko).max() / np.abs(eko).max(),
                     np.abs(ep - epo).max() / np.abs(epo).max())
            print(f"  {name} precision {prec} {opts}: hpsi {err:.2e} energies {ee:.2e}",
                  flush=True)
            assert err < bar and ee < ebar
    autotune.apply(ctx, {"line_passes": 0, "zgroup": 0, "zpot": 0, "bwd": 0, "chunk_mb": 0, "tile_order": 0, "persist": 0})
    lr.close()


def test_tune_chooses_caches_and_keeps_results(ctx, tmp_path):
    import bigdft_hip
    from bigdft_hip import autotune
    reg = dense_region("P", (31, 31, 31), (0.4,) * 3)
    norb = 6
    pot, psi = dev(random_pot(reg, 5)), dev(random_psi(reg, norb, 6).ravel())
    lr = ctx.region(reg)
    prec = bigdft_hip.PREC_F16X3
    h0, ek0, ep0 = run(ctx, lr, norb, pot, psi, prec)
    cache = autotune.TuneCache(str(tmp_path / "tune.json"))
    key, opts, times = autotune.tune(ctx, [lr] * norb, pot, psi, prec, cache=cache)
    print(f"  {key}: {opts} {times}", flush=True)
    assert key.endswith("|P:31x31x31:f16x3")
    assert len(times) == len(autotune.candidates(prec)) + len(autotune.CHUNK_MB) + 2
    assert dict(opts, chunk_mb=0, tile_order=0, persist=0) in autotune.candidates(prec)
    assert (tmp_path / "tune.json").exists()
    h1, ek1, ep1 = run(ctx, lr, norb, pot, psi, prec)
    assert np.abs(h1 - h0).max() < 2e-6 * np.abs(h0).max()
    assert np.abs(ep1 - ep0).max() < 1e-6 * np.abs(ep0).max()
    # a new tuner on the same file answers from it
    _, opts2, times2 = autotune.tune(ctx, [lr] * norb, pot, psi, prec,
                                     cache=autotune.TuneCache(str(tmp_path / "tune.json")))
    assert opts2 == opts and times2 is None
    autotune.apply(ctx, {"line_passes": 0, "zgroup": 0, "zpot": 0, "bwd": 0, "chunk_mb": 0, "tile_order": 0, "persist": 0})
    with pytest.raises(bigdft_hip.BigDFTError):
        ctx.set_option("zgroup", 5)
    lr.close()


def test_tune_on_a_box_that_only_fits_some_options(ctx, tmp_path):
    """a 60^3 free box in F16X3: the matrix-core backward pass and the
    per-pass sequence do not fit in LDS; the tuner skips them and its choice
    gives the oracle's energies"""
    import bigdft_hip
    from bigdft_hip import autotune
    reg = dense_region("F", (59, 59, 59), (0.45, 0.4, 0.5))
    pot, psi = random_pot(reg, 31), random_psi(reg, 1, 32)
    lr = ctx.region(reg)
    dp, ds = dev(pot), dev(psi.ravel())
    prec = bigdft_hip.PREC_F16X3
    key, opts, times = autotune.tune(ctx, [lr], dp, ds, prec,
                                     cache=autotune.TuneCache(str(tmp_path / "t.json")))
    print(f"  {key}: {opts}; skipped {[k for k, v in times.items() if v is None]}", flush=True)
    assert opts["bwd"] != 1 and opts["line_passes"] == 0
    assert any(v is None for v in times.values())
    h, ek, ep = run(ctx, lr, 1, dp, ds, prec)
    ho, eko, epo = oracle.locham("F", reg.n, reg.nfl, reg.nfu, reg.hgrid, okeys(reg), pot, psi)
    assert np.abs(h - ho.ravel()).max() < 1e-4 * np.abs(ho).max()
    assert abs(ek[0] - eko[0]) < 1e-6 * abs(eko[0]) and abs(ep[0] - epo[0]) < 1e-6 * abs(epo[0])
    autotune.apply(ctx, {"line_passes": 0, "zgroup": 0, "zpot": 0, "bwd": 0, "chunk_mb": 0, "tile_order": 0, "persist": 0})
    lr.close()


@pytest.mark.parametrize("prec", [0, 1, 4], ids=["fp64", "fp32", "f16x3"])
def test_tile_order_leaves_results_bit_equal(ctx, prec):
    """BDFT_OPT_TILE_ORDER and BDFT_OPT_PERSIST only change which
    work-group takes which tile: hpsi and the energies are bit-equal in every
    order, on a batch mixing periodic, free and sphere boxes (plane and line
    sequences, the z potential pass per tile and as a persistent work queue)"""
    from bigdft_hip import autotune
    regs = [CASES[0][1], CASES[1][1], dense_region("F", (17, 21, 15), (0.4,) * 3)]
    rs = [regs[i % 3] for i in range(7)]
    lrs_of = {id(r): ctx.region(r) for r in regs}
    lrs = [lrs_of[id(r)] for r in rs]
    pots = [random_pot(r, 60 + i) for i, r in enumerate(rs)]
    psi = dev(np.concatenate([random_psi(r, 1, 70 + i)[0] for i, r in enumerate(rs)]))
    pot = dev(np.concatenate(pots))
    kw = dict(pot_off=np.concatenate([[0], np.cumsum([p.size for p in pots])[:-1]]),
              potdim=np.array([r.nfine for r in rs]).ravel(), precision=prec)
    for lines in (0, 1):
        outs = []
        for order, persist in ((0, 0), (1, 0), (0, 1), (1, 1)):
            print(f"  precision {prec} line_passes {lines} tile_order {order} persist {persist}",
                  flush=True)
            autotune.apply(ctx, {"line_passes": lines, "zgroup": 0, "zpot": 0, "bwd": 0,
                                 "chunk_mb": 0, "tile_order": order, "persist": persist})
            h = torch.zeros_like(psi)
            ek = torch.zeros(len(rs), dtype=torch.float64, device="cuda")
            ep = torch.zeros_like(ek)
            ctx.locham_batch(lrs, pot, psi, h, ek, ep, **kw)
            ctx.sync()
            outs.append((h, ek, ep))
        for o in outs[1:]:
            for a, b in zip(outs[0], o):
                assert torch.equal(a, b)
    autotune.apply(ctx, {"line_passes": 0, "zgroup": 0, "zpot": 0, "bwd": 0, "chunk_mb": 0,
                         "tile_order": 0})
    for l in lrs_of.values():
        l.close()
