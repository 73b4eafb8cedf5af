"""HIP path vs the oracle, through the C ABI, on seeded inputs.

Bars: compression transfers bit-exact; FP64 local Hamiltonian within 1e-12
per element (relative error guarded near zeros: |a-b| / max(|b|,
1e-3 max|b|)) and 1e-12 relative in ekin and epot.
"""
import numpy as np
import pytest

import oracle
from helpers import (dense_region, mixed_sphere_region, okeys, random_pot,
                     random_psi, rel_err, region_from_grids, sphere_region)

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def ctx():
    import bigdft_hip
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    c = bigdft_hip.Context(0)
    print("device LDS per work-group, plane sequence:", c.info(), flush=True)
    torch.cuda.synchronize()
    yield c
    c.close()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def run_gpu(ctx, region, pot, psi, hpsi0=None, precision=0):
    lr = ctx.region(region)
    norb = psi.shape[0]
    d_psi = dev(psi)
    d_h = dev(np.zeros_like(psi) if hpsi0 is None else hpsi0)
    d_pot = dev(pot)
    ek = torch.zeros(norb, dtype=torch.float64, device="cuda")
    ep = torch.zeros(norb, dtype=torch.float64, device="cuda")
    ctx.locham(lr, d_pot, d_psi, d_h, ek, ep, precision)
    ctx.sync()
    out = d_h.cpu().numpy(), ek.cpu().numpy(), ep.cpu().numpy()
    lr.close()
    return out


def check(region, norb, seed, ctx):
    print(f"case {region.geocode} n={region.n} nvctr=({region.nvctr_c},"
          f"{region.nvctr_f}) norb={norb} seed={seed}", flush=True)
    psi = random_psi(region, norb, seed)
    pot = random_pot(region, seed + 1)
    h0 = random_psi(region, norb, seed + 2)
    hg, eg, pg = run_gpu(ctx, region, pot, psi, h0)
    ho, eo, po = oracle.locham(region.geocode, region.n, region.nfl,
                               region.nfu, region.hgrid, okeys(region), pot,
                               psi, h0)
    assert rel_err(hg - h0, ho - h0) < TOL
    assert rel_err(eg, eo) < TOL
    assert rel_err(pg, po) < TOL


@pytest.mark.parametrize("n", [(0, 0, 0), (3, 4, 2), (7, 6, 8), (9, 13, 5)])
def test_periodic_small(ctx, n):
    check(dense_region("P", n, (0.45, 0.4, 0.5)), 2, 10 + sum(n), ctx)


def test_periodic_config1(ctx):
    """24^3 coarse, hgrid 0.45, V ~ U[-2,0] (the single-orbital case)"""
    check(dense_region("P", (23, 23, 23), (0.45,) * 3), 1, 1, ctx)


@pytest.mark.parametrize("n", [(4, 5, 6), (11, 9, 10), (15, 15, 15)])
def test_free_dense(ctx, n):
    check(dense_region("F", n, (0.4, 0.45, 0.5)), 2, 20 + sum(n), ctx)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_free_sphere(ctx, seed):
    check(mixed_sphere_region((14, 12, 16), 0.4, seed), 2, 30 + seed, ctx)


@pytest.mark.parametrize("geo,host", [("P", False), ("F", False), ("P", True)])
def test_two_spin_potentials(ctx, geo, host):
    """pot(n1i,n2i,n3i,2): orbitals whose spinsgn is not positive read the
    second potential (rhopot_down, interface_ocl.f90:487-491)"""
    reg = dense_region(geo, (9, 10, 8), (0.45, 0.4, 0.5))
    norb = 5
    print(f"case nspin=2 {geo} host={host}", flush=True)
    psi = random_psi(reg, norb, 61)
    up, dn = random_pot(reg, 62), random_pot(reg, 63, -4.0, 1.0)
    sg = np.array([1.0, -1.0, -1.0, 1.0, -0.5])
    h0 = random_psi(reg, norb, 64)
    lr = ctx.region(reg)
    if host:
        hg = h0.copy()
        eg, pg = ctx.locham_host(lr, np.concatenate([up, dn]), psi, hg, nspin=2, spinsgn=sg)
    else:
        d_h = dev(h0)
        ek = torch.zeros(norb, dtype=torch.float64, device="cuda")
        ep = torch.zeros_like(ek)
        ctx.locham(lr, dev(np.concatenate([up, dn])), dev(psi), d_h, ek, ep, nspin=2,
                   spinsgn=sg)
        ctx.sync()
        hg, eg, pg = d_h.cpu().numpy(), ek.cpu().numpy(), ep.cpu().numpy()
    for o in range(norb):
        v = up if sg[o] > 0 else dn
        ho, eo, po = oracle.locham(reg.geocode, reg.n, reg.nfl, reg.nfu, reg.hgrid,
                                   okeys(reg), v, psi[o:o + 1], h0[o:o + 1])
        assert rel_err(hg[o] - h0[o], ho[0] - h0[o]) < TOL
        assert rel_err(eg[o:o + 1], eo) < TOL
        assert rel_err(pg[o:o + 1], po) < TOL
    # a spin-down orbital without a second potential is refused
    import bigdft_hip
    with pytest.raises(bigdft_hip.BigDFTError, match="nspin"):
        ctx.locham_host(lr, up, psi, h0.copy(), nspin=1, spinsgn=sg)
    lr.close()


def test_stream_switch_orders_shared_buffers(ctx):
    """bdft_ctx_set_stream: a call queued on the new stream waits for the
    old stream's work, which shares the context's one-shot workspace; the
    same batch queued on two streams through the raw C ABI (no ordering by
    the Python layer) gives the same result twice"""
    import ctypes

    import bigdft_hip
    L = bigdft_hip.lib()
    reg = dense_region("P", (31, 31, 31), (0.45,) * 3)
    norb = 8
    psi, pot = random_psi(reg, norb, 71), random_pot(reg, 72)
    c = bigdft_hip.Context(0)
    lr = c.region(reg)
    d_psi, d_pot = dev(psi), dev(pot)
    outs = [(torch.zeros_like(d_psi), torch.zeros(norb, dtype=torch.float64, device="cuda"),
             torch.zeros(norb, dtype=torch.float64, device="cuda")) for _ in range(2)]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    torch.cuda.synchronize()
    P = ctypes.c_void_p
    for (h, ek, ep), st in zip(outs, streams):
        c._check(L.bdft_ctx_set_stream(c.handle, P(st.cuda_stream)))
        c._check(L.bdft_locham_d(c.handle, lr.handle, norb, 1, P(d_pot.data_ptr()), None,
                                 P(d_psi.data_ptr()), P(h.data_ptr()), P(ek.data_ptr()),
                                 P(ep.data_ptr()), 0))
    torch.cuda.synchronize()
    for a, b in zip(outs[0], outs[1]):
        assert torch.equal(a, b)
    c._stream = streams[1]
    lr.close()
    c.close()


def test_device_arrays_are_checked(ctx):
    """CPU, mistyped, strided or short tensors are refused before any launch"""
    reg = dense_region("P", (7, 7, 7), (0.45,) * 3)
    lr = ctx.region(reg)
    psi = torch.zeros((2, reg.ncoef), dtype=torch.float64, device="cuda")
    pot = torch.zeros(reg.n_isf, dtype=torch.float64, device="cuda")
    ek = torch.zeros(2, dtype=torch.float64, device="cuda")
    bad = [(psi.cpu(), psi), (psi.float(), psi), (psi[:, ::2], psi), (psi[:1], psi)]
    for p, h in bad:
        with pytest.raises(ValueError):
            ctx.locham(lr, pot, p, h, ek, ek.clone())
    with pytest.raises(ValueError):
        ctx.locham(lr, pot[:10], psi, psi.clone(), ek, ek.clone())
    lr.close()


def test_free_no_fine_points(ctx):
    lg = np.ones((9, 8, 7), dtype=bool)
    reg = region_from_grids("F", (8, 7, 6), (0.4,) * 3, lg, np.zeros_like(lg))
    check(reg, 1, 40, ctx)


@pytest.mark.parametrize("geo", ["P", "F"])
def test_uncompress_compress_bit_exact(ctx, geo):
    if geo == "P":
        reg = dense_region("P", (9, 7, 8), (0.4,) * 3)
        lg = np.zeros((10, 8, 9), dtype=bool)
        lg[2:7] = True   # partial rows
        reg = region_from_grids("P", (9, 7, 8), (0.4,) * 3, lg, lg)
    else:
        reg = mixed_sphere_region((14, 12, 16), 0.4, 4)
    norb = 3
    psi = random_psi(reg, norb, 50)
    lr = ctx.region(reg)
    vol = 8 * np.prod([x + 1 for x in reg.n])
    psig = torch.full((norb * vol,), 7.0, dtype=torch.float64, device="cuda")
    ctx.uncompress(lr, norb, dev(psi), psig)
    h0 = random_psi(reg, norb, 51)
    d_h = dev(h0)
    ctx.compress_acc(lr, norb, psig, d_h)
    ctx.sync()
    g = psig.cpu().numpy().reshape(norb, vol)
    for i in range(norb):
        o = oracle.uncompress_psig(reg.n, okeys(reg), psi[i])
        assert np.array_equal(g[i], o)
        oc = oracle.compress_acc_mixed(reg.n, okeys(reg), o, h0[i])
        assert np.array_equal(d_h.cpu().numpy()[i], oc)
    # compress(uncompress(psi)) onto zeros gives psi back exactly
    z = torch.zeros((norb, reg.ncoef), dtype=torch.float64, device="cuda")
    ctx.compress_acc(lr, norb, psig, z)
    ctx.sync()
    assert np.array_equal(z.cpu().numpy(), psi)
    lr.close()


def test_conv_check_pattern_bit_exact(ctx):
    """keys and values of bigdft/src/OpenCL/conv_check.f90:1622-1660"""
    n1 = 76
    lg = np.zeros((n1 + 1,) * 3, dtype=bool)
    lg[5:n1 - 4] = True
    reg = region_from_grids("P", (n1,) * 3, (0.4,) * 3, lg, lg)
    nv = reg.nvctr_c
    psi = np.concatenate([np.arange(1, nv + 1, dtype=np.float64),
                          (np.arange(1, nv + 1)[:, None] +
                           0.1 * np.arange(1, 8)[None, :]).ravel()])[None, :]
    lr = ctx.region(reg)
    psig = torch.empty(8 * (n1 + 1) ** 3, dtype=torch.float64, device="cuda")
    ctx.uncompress(lr, 1, dev(psi), psig)
    z = torch.zeros_like(dev(psi))
    ctx.compress_acc(lr, 1, psig, z)
    ctx.sync()
    assert np.array_equal(z.cpu().numpy(), psi)
    assert np.array_equal(psig.cpu().numpy(),
                          oracle.uncompress_psig(reg.n, okeys(reg), psi[0]))
    lr.close()


def test_uneven_batch_with_potential_offsets(ctx):
    """orbitals in different free-BC regions, each with its own potential box
    inside one concatenated array (linear-scaling layout)"""
    import bigdft_hip
    rng = np.random.default_rng(60)
    regs = [mixed_sphere_region(tuple(int(x) for x in rng.integers(6, 15, 3)),
                                0.4, 61 + i) for i in range(5)]
    regs.append(dense_region("F", (5, 9, 7), (0.4,) * 3))
    lrs = [ctx.region(r) for r in regs]
    psis = [random_psi(r, 1, 70 + i)[0] for i, r in enumerate(regs)]
    pots = [random_pot(r, 80 + i) for i, r in enumerate(regs)]
    psi_off = np.cumsum([0] + [p.size for p in psis])[:-1]
    pot_off = np.cumsum([0] + [p.size for p in pots])[:-1]
    d_psi = dev(np.concatenate(psis))
    d_h = torch.zeros_like(d_psi)
    d_pot = dev(np.concatenate(pots))
    potdim = np.array([r.nfine for r in regs], dtype=np.int32).ravel()
    ek = torch.zeros(len(regs), dtype=torch.float64, device="cuda")
    ep = torch.zeros_like(ek)
    ctx.locham_batch(lrs, d_pot, d_psi, d_h, ek, ep, pot_off=pot_off,
                     potdim=potdim, psi_off=psi_off)
    ctx.sync()
    h = d_h.cpu().numpy()
    for i, r in enumerate(regs):
        ho, eo, po = oracle.locham("F", r.n, r.nfl, r.nfu, r.hgrid, okeys(r),
                                   pots[i], psis[i])
        assert rel_err(h[psi_off[i]:psi_off[i] + psis[i].size], ho[0]) < TOL
        assert rel_err(ek.cpu().numpy()[i], eo[0]) < TOL
        assert rel_err(ep.cpu().numpy()[i], po[0]) < TOL
    for l in lrs:
        l.close()


@pytest.mark.parametrize("precision", [0, 1, 4], ids=["fp64", "fp32", "f16x3"])
def test_batch_mixing_periodic_and_free_boxes(ctx, precision):
    """one batch of periodic and free boxes of different sizes (the kernel
    sequence is chosen for the batch, the boundary conditions per box):
    FP64 within 1e-12 per element and in the energies; FP32 and F16X3
    within 1e-4 of max|hpsi| per orbital and 1e-6 in the energies"""
    regs = [dense_region("P", (23, 19, 21), (0.45, 0.4, 0.5)),
            dense_region("F", (11, 9, 13), (0.4, 0.45, 0.5)),
            mixed_sphere_region((20, 18, 22), 0.42, 7),
            dense_region("P", (17, 17, 17), (0.4,) * 3),
            dense_region("F", (29, 21, 17), (0.45,) * 3)]
    lrs = [ctx.region(r) for r in regs]
    psis = [random_psi(r, 1, 90 + i)[0] for i, r in enumerate(regs)]
    pots = [random_pot(r, 95 + i) for i, r in enumerate(regs)]
    psi_off = np.cumsum([0] + [p.size for p in psis])[:-1]
    pot_off = np.cumsum([0] + [p.size for p in pots])[:-1]
    d_psi = dev(np.concatenate(psis))
    d_h = torch.zeros_like(d_psi)
    potdim = np.array([r.nfine for r in regs], dtype=np.int32).ravel()
    ek = torch.zeros(len(regs), dtype=torch.float64, device="cuda")
    ep = torch.zeros_like(ek)
    ctx.locham_batch(lrs, dev(np.concatenate(pots)), d_psi, d_h, ek, ep, pot_off=pot_off,
                     potdim=potdim, psi_off=psi_off, precision=precision)
    ctx.sync()
    h, ekh, eph = d_h.cpu().numpy(), ek.cpu().numpy(), ep.cpu().numpy()
    for i, r in enumerate(regs):
        ho, eo, po = oracle.locham(r.geocode, r.n, r.nfl, r.nfu, r.hgrid, okeys(r), pots[i],
                                   psis[i])
        hi = h[psi_off[i]:psi_off[i] + psis[i].size]
        ee = max(abs(ekh[i] - eo[0]) / abs(eo[0]), abs(eph[i] - po[0]) / abs(po[0]))
        if precision == 0:
            assert rel_err(hi, ho[0]) < TOL and ee < TOL
        else:
            assert np.abs(hi - ho[0]).max() < 1e-4 * np.abs(ho[0]).max() and ee < 1e-6
    for l in lrs:
        l.close()


def test_potential_box_shift(ctx):
    """global potential, region placed by ishift (apply_potential_lr_nobounds
    with ishift), compared with the oracle on the extracted box"""
    reg, nl = sphere_region((30, 30, 30), (0.4,) * 3, (6.1, 5.7, 6.6), 3.0, 1.5)
    gn = tuple(2 * 30 + 31 for _ in range(3))
    rng = np.random.default_rng(90)
    gpot = rng.uniform(-2, 0, size=np.prod(gn))
    ishift = [-2 * nl[d] for d in range(3)]
    nf = reg.nfine
    G = gpot.reshape(gn, order="F")
    box = G[2 * nl[0]:2 * nl[0] + nf[0], 2 * nl[1]:2 * nl[1] + nf[1],
            2 * nl[2]:2 * nl[2] + nf[2]].ravel(order="F")
    psi = random_psi(reg, 1, 91)
    lr = ctx.region(reg)
    d_psi = dev(psi)
    d_h = torch.zeros_like(d_psi)
    ek = torch.zeros(1, dtype=torch.float64, device="cuda")
    ep = torch.zeros_like(ek)
    ctx.locham_batch([lr], dev(gpot), d_psi, d_h, ek, ep, pot_off=[0],
                     potdim=np.array(gn, dtype=np.int32), ishift=np.array(ishift))
    ctx.sync()
    ho, eo, po = oracle.locham("F", reg.n, reg.nfl, reg.nfu, reg.hgrid,
                               okeys(reg), box, psi)
    assert rel_err(d_h.cpu().numpy(), ho) < TOL
    assert rel_err(ep.cpu().numpy(), po) < TOL
    lr.close()


def test_graph_plan_replays(ctx):
    """captured plan replayed several times with eager work in between; every
    replay adds the same H psi"""
    reg = dense_region("P", (11, 11, 11), (0.4,) * 3)
    psi = random_psi(reg, 4, 100)
    pot = random_pot(reg, 101)
    lr = ctx.region(reg)
    d_psi, d_pot = dev(psi), dev(pot)
    d_h = torch.zeros_like(d_psi)
    ek = torch.zeros(4, dtype=torch.float64, device="cuda")
    ep = torch.zeros_like(ek)
    plan = ctx.plan([lr] * 4, d_pot, d_psi, d_h, ek, ep)
    ho, eo, po = oracle.locham("P", reg.n, reg.nfl, reg.nfu, reg.hgrid,
                               okeys(reg), pot, psi)
    prev = np.zeros_like(psi)
    for rep in range(4):
        plan.execute()
        ctx.sync()
        cur = d_h.cpu().numpy()
        assert rel_err(cur - prev, ho) < 1e-11
        assert rel_err(ek.cpu().numpy(), eo) < TOL
        prev = cur
        _ = (torch.randn(1000, device="cuda") * 2).sum().item()  # eager work
    plan.close()
    lr.close()


@pytest.mark.parametrize("geo,n,precision", [("P", (23, 23, 23), 4), ("F", (19, 17, 21), 4),
                                              ("P", (23, 23, 23), 3)])
def test_graph_branches_match_the_direct_launches(ctx, geo, n, precision):
    """matrix-core plans run the z kinetic and z potential passes as parallel
    branches of their graph; every replay equals the direct (one-stream)
    launch bit for bit, with eager work between the replays"""
    reg = dense_region(geo, n, (0.4, 0.45, 0.5))
    norb = 3
    psi = random_psi(reg, norb, 120)
    pot = random_pot(reg, 121)
    lr = ctx.region(reg)
    d_psi, d_pot = dev(psi), dev(pot)
    h1 = torch.zeros_like(d_psi)
    e1 = torch.zeros(norb, dtype=torch.float64, device="cuda")
    p1 = torch.zeros_like(e1)
    ctx.locham_batch([lr] * norb, d_pot, d_psi, h1, e1, p1, precision=precision)
    ctx.sync()
    h2 = torch.zeros_like(d_psi)
    e2 = torch.zeros_like(e1)
    p2 = torch.zeros_like(e1)
    plan = ctx.plan([lr] * norb, d_pot, d_psi, h2, e2, p2, precision=precision)
    for rep in range(3):
        h2.zero_()
        plan.execute()
        ctx.sync()
        assert torch.equal(h2, h1)
        assert torch.equal(e2, e1) and torch.equal(p2, p1)
        _ = (torch.randn(1000, device="cuda") * 2).sum().item()  # eager work
    plan.close()
    lr.close()


def test_host_entry_point(ctx):
    reg = dense_region("F", (6, 7, 5), (0.45,) * 3)
    psi = random_psi(reg, 2, 110)
    pot = random_pot(reg, 111)
    lr = ctx.region(reg)
    h = np.zeros_like(psi)
    ek, ep = ctx.locham_host(lr, pot, psi, h)
    ho, eo, po = oracle.locham("F", reg.n, reg.nfl, reg.nfu, reg.hgrid,
                               okeys(reg), pot, psi)
    assert rel_err(h, ho) < TOL and rel_err(ek, eo) < TOL and rel_err(ep, po) < TOL
    lr.close()


def test_descriptor_errors_are_reported(ctx):
    import bigdft_hip
    reg = dense_region("P", (3, 3, 3), (0.4,) * 3)
    bad = reg.keyv_c.copy()
    bad[2] += 1
    reg.keyv_c = bad
    with pytest.raises(bigdft_hip.BigDFTError) as e:
        ctx.region(reg)
    assert e.value.code == 3
    # 257 coarse points along x: transfers accept it, the Hamiltonian
    # (at most 256 per axis) does not
    big = dense_region("P", (256, 3, 3), (0.4,) * 3)
    lr = ctx.region(big)
    with pytest.raises(bigdft_hip.BigDFTError) as e:
        run = torch.zeros(big.ncoef, dtype=torch.float64, device="cuda")
        ek = torch.zeros(1, dtype=torch.float64, device="cuda")
        ctx.locham(lr, torch.zeros(big.n_isf, dtype=torch.float64, device="cuda"),
                   run, run.clone(), ek, ek.clone())
    assert e.value.code == 5
    lr.close()


def test_empty_batch_is_an_error(ctx):
    import bigdft_hip
    z = torch.zeros(1, dtype=torch.float64, device="cuda")
    with pytest.raises(bigdft_hip.BigDFTError) as e:
        ctx.locham_batch([], z, z, z, z, z)
    assert e.value.code == 1


@pytest.mark.parametrize("precision,n", [(0, 63), (4, 59)], ids=["fp64", "f16x3"])
def test_largest_free_box(ctx, precision, n):
    """the largest supported free box, one orbital: 64 coarse points per
    axis (157 fine) in FP64, 60 in the split precisions, which run the
    plane-fused sequence with the FP32 backward plane pass there (the
    matrix-core backward kernel would not fit in LDS); FP64 within 1e-12,
    F16X3 within 1e-4 of max|hpsi| and 1e-6 in the energies"""
    reg = dense_region("F", (n, n, n), (0.45, 0.4, 0.5))
    print(f"largest free box {reg.n} precision {precision}", flush=True)
    psi = random_psi(reg, 1, 4242)
    pot = random_pot(reg, 4243)
    hg, eg, pg = run_gpu(ctx, reg, pot, psi, precision=precision)
    ho, eo, po = oracle.locham("F", reg.n, reg.nfl, reg.nfu, reg.hgrid, okeys(reg), pot, psi)
    eh = np.abs(hg - ho).max() / np.abs(ho).max()
    ee = max(abs(eg[0] - eo[0]) / abs(eo[0]), abs(pg[0] - po[0]) / abs(po[0]))
    print(f"  hpsi {eh:.2e} energies {ee:.2e}", flush=True)
    if precision == 0:
        assert rel_err(hg, ho) < TOL and ee < TOL
    else:
        assert eh < 1e-4 and ee < 1e-6


def test_split_precision_refuses_a_forced_unit_beyond_its_lds(ctx):
    """a free box whose matrix-core backward pass does not fit in LDS while
    the option asks for the matrix cores: an error with the LDS figure, not
    a launch (with the automatic units such a batch runs the FP32 passes,
    test_large_boxes)"""
    import bigdft_hip
    reg = dense_region("F", (59, 59, 59), (0.45, 0.4, 0.5))
    ctx.set_option("bwd", 1)
    try:
        with pytest.raises(bigdft_hip.BigDFTError, match="LDS") as e:
            run_gpu(ctx, reg, random_pot(reg, 1), random_psi(reg, 1, 2), precision=4)
        assert e.value.code == 5
    finally:
        ctx.set_option("bwd", 0)


# Boxes beyond the line kernels' register-held outputs (N1 > 64, or more
# than 192 fine points on x or z) and beyond the matrix-core kernels (more
# than 64 coarse points: the split precisions run the FP32 passes there).
# (37, 45, 64) is the 76 x 92 x 130 fine grid of the reference's kernel
# check (bigdft/tests/libs/OpenCL/Makefile.am:28-31).
LARGE = [("P", (37, 45, 64)), ("F", (63, 63, 63)), ("F", (70, 40, 90)),
         ("P", (127, 127, 127))]


@pytest.mark.parametrize("geo,n", LARGE, ids=[f"{g}{n[0]}x{n[1]}x{n[2]}" for g, n in LARGE])
@pytest.mark.parametrize("precision", [0, 4], ids=["fp64", "f16x3"])
def test_large_boxes(ctx, geo, n, precision):
    reg = dense_region(geo, n, (0.45, 0.4, 0.5))
    print(f"large box {geo} {n} fine {reg.nfine} precision {precision}", flush=True)
    psi = random_psi(reg, 1, 4300 + sum(n))
    pot = random_pot(reg, 4301 + sum(n))
    hg, eg, pg = run_gpu(ctx, reg, pot, psi, precision=precision)
    ho, eo, po = oracle.locham(geo, reg.n, reg.nfl, reg.nfu, reg.hgrid, okeys(reg), pot, psi)
    eh = np.abs(hg - ho).max() / np.abs(ho).max()
    ee = max(abs(eg[0] - eo[0]) / abs(eo[0]), abs(pg[0] - po[0]) / abs(po[0]))
    print(f"  hpsi {eh:.2e} energies {ee:.2e}", flush=True)
    if precision == 0:
        assert rel_err(hg, ho) < TOL and ee < TOL
    else:
        assert eh < 1e-4 and ee < 1e-6


# ---------------------------------------------------------------------------
# mixed precision (float working type; FP64 input, potential, energies and
# hpsi accumulation).  Bars: ekin, epot within 1e-6 relative of the FP64
# oracle; hpsi within 2e-6 of max|hpsi| per element.
# ---------------------------------------------------------------------------
MIXED_E = 1e-6
MIXED_H = 2e-6


@pytest.mark.parametrize("geo,n,seed", [("P", (7, 6, 8), 1), ("P", (23, 23, 23), 2),
                                        ("F", (11, 9, 10), 3), ("F", (15, 15, 15), 4)])
def test_mixed_fp32(ctx, geo, n, seed):
    reg = dense_region(geo, n, (0.45, 0.4, 0.5))
    psi = random_psi(reg, 2, 200 + seed)
    pot = random_pot(reg, 300 + seed)
    hg, eg, pg = run_gpu(ctx, reg, pot, psi, precision=1)
    ho, eo, po = oracle.locham(geo, reg.n, reg.nfl, reg.nfu, reg.hgrid,
                               okeys(reg), pot, psi)
    print(f"mixed {geo} {n}: hpsi {np.abs(hg - ho).max() / np.abs(ho).max():.2e} "
          f"ekin {np.abs(eg - eo).max() / np.abs(eo).max():.2e} "
          f"epot {np.abs(pg - po).max() / np.abs(po).max():.2e}")
    assert np.abs(hg - ho).max() < MIXED_H * np.abs(ho).max()
    assert np.all(np.abs(eg - eo) < MIXED_E * np.abs(eo))
    assert np.all(np.abs(pg - po) < MIXED_E * np.abs(po))


def test_mixed_sphere_batch(ctx):
    regs = [mixed_sphere_region((14, 12, 16), 0.4, s) for s in (5, 6)]
    for i, reg in enumerate(regs):
        psi = random_psi(reg, 1, 400 + i)
        pot = random_pot(reg, 500 + i)
        hg, eg, pg = run_gpu(ctx, reg, pot, psi, precision=1)
        ho, eo, po = oracle.locham("F", reg.n, reg.nfl, reg.nfu, reg.hgrid,
                                   okeys(reg), pot, psi)
        assert np.abs(hg - ho).max() < MIXED_H * np.abs(ho).max()
        assert abs(eg[0] - eo[0]) < MIXED_E * abs(eo[0])
        assert abs(pg[0] - po[0]) < MIXED_E * abs(po[0])


# ---------------------------------------------------------------------------
# both kernel sequences: plane-fused (default where its planes fit in LDS)
# and the line passes; the small periodic boxes wrap each stencil several
# times (the plane kernels' line maps take k modulo N)
# ---------------------------------------------------------------------------
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


@pytest.mark.parametrize("geo,n", [("P", (23, 19, 18)), ("F", (15, 13, 17)),
                                   ("P", (31, 31, 31)), ("P", (7, 11, 5)),
                                   ("P", (2, 0, 4))])
def test_line_and_plane_sequences_agree_with_oracle(ctx, ctx_lines, geo, n):
    reg = dense_region(geo, n, (0.41, 0.43, 0.39))
    for c in (ctx, ctx_lines):
        check(reg, 2, 700 + n[0], c)


@pytest.mark.parametrize("seed", [11, 12])
def test_plane_sparse_keys(ctx, seed):
    """sphere regions (inverse maps) on boxes large enough for the plane path"""
    check(mixed_sphere_region((24, 22, 26), 0.4, seed), 2, 800 + seed, ctx)


def test_mixed_line_sequence(ctx_lines):
    reg = dense_region("P", (23, 23, 23), (0.45,) * 3)
    psi = random_psi(reg, 2, 900)
    pot = random_pot(reg, 901)
    hg, eg, pg = run_gpu(ctx_lines, reg, pot, psi, precision=1)
    ho, eo, po = oracle.locham("P", reg.n, reg.nfl, reg.nfu, reg.hgrid,
                               okeys(reg), pot, psi)
    assert np.abs(hg - ho).max() < MIXED_H * np.abs(ho).max()
    assert np.all(np.abs(eg - eo) < MIXED_E * np.abs(eo))
    assert np.all(np.abs(pg - po) < MIXED_E * np.abs(po))


# ---------------------------------------------------------------------------
# split-BF16 matrix-core path (FP32 accumulation, FP64 energies and
# accumulation).  Bars for the four-product form: ekin, epot within 1e-6
# relative of the FP64 oracle; hpsi within 1e-4 of max|hpsi| per element.
# The three-product form is held to 1e-4 per element and 2e-5 in the
# energies (its filter carries a fixed 2^-16 representation error).
# ---------------------------------------------------------------------------
BF_E = 1e-6
BF_H = 1e-4


def set_passes(c, seq):
    """units of the z potential and backward passes of the plane-fused
    matrix-core sequence (launch options)"""
    v = {"plane-mfma": 1, "plane-valu": 2}.get(seq, 0)
    c.set_option("zpot", v)
    c.set_option("bwd", v)


@pytest.mark.parametrize("seq", ["plane", "plane-mfma", "plane-valu", "lines"])
@pytest.mark.parametrize("geo,n,seed", [("P", (23, 23, 23), 1), ("P", (7, 6, 8), 2),
                                        ("F", (15, 13, 17), 3), ("P", (31, 19, 24), 4),
                                        ("F", (39, 39, 39), 5), ("P", (47, 20, 9), 6)])
def test_bf16x4(ctx, ctx_lines, seq, geo, n, seed):
    """plane: the plane-fused matrix-core sequence with the automatic choice
    of units for the z potential and backward passes; plane-mfma: both on
    the matrix cores; plane-valu: both on the FP32 vector units (where they
    apply); lines: the per-pass sequence"""
    reg = dense_region(geo, n, (0.45, 0.4, 0.5))
    print(f"bf16x4 case {seq} {geo} {n}", flush=True)
    psi = random_psi(reg, 2, 1000 + seed)
    pot = random_pot(reg, 1100 + seed)
    c = ctx_lines if seq == "lines" else ctx
    set_passes(c, seq)
    try:
        hg, eg, pg = run_gpu(c, reg, pot, psi, precision=3)
    finally:
        set_passes(c, "plane")
    ho, eo, po = oracle.locham(geo, reg.n, reg.nfl, reg.nfu, reg.hgrid,
                               okeys(reg), pot, psi)
    eh = np.abs(hg - ho).max() / np.abs(ho).max()
    ee = np.max(np.abs(eg - eo) / np.abs(eo))
    ep_ = np.max(np.abs(pg - po) / np.abs(po))
    print(f"bf16x4 {seq} {geo} {n}: hpsi {eh:.2e} ekin {ee:.2e} epot {ep_:.2e}",
          flush=True)
    assert eh < BF_H
    assert ee < BF_E and ep_ < BF_E


def test_bf16x4_sphere_batch(ctx):
    regs = [mixed_sphere_region((24, 22, 26), 0.4, s) for s in (21, 22, 23)]
    lrs = [ctx.region(r) for r in regs]
    psis = [random_psi(r, 1, 1200 + i)[0] for i, r in enumerate(regs)]
    pots = [random_pot(r, 1300 + i) for i, r in enumerate(regs)]
    psi_off = np.cumsum([0] + [p.size for p in psis])[:-1]
    pot_off = np.cumsum([0] + [p.size for p in pots])[:-1]
    d_psi = dev(np.concatenate(psis))
    d_h = torch.zeros_like(d_psi)
    potdim = np.array([r.nfine for r in regs], dtype=np.int32).ravel()
    ek = torch.zeros(len(regs), dtype=torch.float64, device="cuda")
    ep = torch.zeros_like(ek)
    ctx.locham_batch(lrs, dev(np.concatenate(pots)), d_psi, d_h, ek, ep,
                     pot_off=pot_off, potdim=potdim, psi_off=psi_off, precision=3)
    ctx.sync()
    h = d_h.cpu().numpy()
    for i, r in enumerate(regs):
        ho, eo, po = oracle.locham("F", r.n, r.nfl, r.nfu, r.hgrid, okeys(r),
                                   pots[i], psis[i])
        hs = h[psi_off[i]:psi_off[i] + psis[i].size]
        assert np.abs(hs - ho[0]).max() < BF_H * np.abs(ho).max()
        assert abs(ek.cpu().numpy()[i] - eo[0]) < BF_E * abs(eo[0])
        assert abs(ep.cpu().numpy()[i] - po[0]) < BF_E * abs(po[0])
    for l in lrs:
        l.close()


def test_bf16x3_three_products(ctx):
    reg = dense_region("P", (23, 23, 23), (0.45,) * 3)
    psi = random_psi(reg, 2, 1400)
    pot = random_pot(reg, 1401)
    hg, eg, pg = run_gpu(ctx, reg, pot, psi, precision=2)
    ho, eo, po = oracle.locham("P", reg.n, reg.nfl, reg.nfu, reg.hgrid,
                               okeys(reg), pot, psi)
    print(f"bf16x3: hpsi {np.abs(hg - ho).max() / np.abs(ho).max():.2e} "
          f"ekin {np.max(np.abs(eg - eo) / np.abs(eo)):.2e} "
          f"epot {np.max(np.abs(pg - po) / np.abs(po)):.2e}")
    assert np.abs(hg - ho).max() < BF_H * np.abs(ho).max()
    assert np.all(np.abs(eg - eo) < 2e-5 * np.abs(eo))
    assert np.all(np.abs(pg - po) < 2e-5 * np.abs(po))


# split-FP16 forward transforms and kinetic (power-of-two scaling per box),
# split-BF16 transposed transforms: same bars as the four-product BF16 form
@pytest.mark.parametrize("seq", ["plane", "plane-mfma", "plane-valu", "lines"])
@pytest.mark.parametrize("geo,n,seed", [("P", (23, 23, 23), 1), ("P", (7, 6, 8), 2),
                                        ("F", (15, 13, 17), 3), ("F", (39, 39, 39), 5),
                                        ("P", (47, 20, 9), 6)])
def test_f16x3(ctx, ctx_lines, seq, geo, n, seed):
    """sequences as in test_bf16x4"""
    reg = dense_region(geo, n, (0.45, 0.4, 0.5))
    print(f"f16x3 case {seq} {geo} {n}", flush=True)
    psi = random_psi(reg, 2, 1500 + seed) * 1e-3
    pot = random_pot(reg, 1600 + seed)
    c = ctx_lines if seq == "lines" else ctx
    set_passes(c, seq)
    try:
        hg, eg, pg = run_gpu(c, reg, pot, psi, precision=4)
    finally:
        set_passes(c, "plane")
    ho, eo, po = oracle.locham(geo, reg.n, reg.nfl, reg.nfu, reg.hgrid,
                               okeys(reg), pot, psi)
    eh = np.abs(hg - ho).max() / np.abs(ho).max()
    ee = np.max(np.abs(eg - eo) / np.abs(eo))
    ep_ = np.max(np.abs(pg - po) / np.abs(po))
    print(f"f16x3 {seq} {geo} {n}: hpsi {eh:.2e} ekin {ee:.2e} epot {ep_:.2e}",
          flush=True)
    assert eh < BF_H
    assert ee < BF_E and ep_ < BF_E


@pytest.mark.parametrize("prec", [0, 4], ids=["fp64", "f16x3"])
def test_chunked_batches_match_one_chunk(ctx, prec):
    """BDFT_OPT_CHUNK_MB: the batch in chunks of orbitals on one workspace
    (one-shot calls and captured plans) gives the one-chunk result bit for
    bit: every orbital runs the same kernels on the same data"""
    import bigdft_hip
    regs = [dense_region("F", (19, 23, 17), (0.4,) * 3), dense_region("P", (23, 19, 21),
                                                                        (0.45, 0.4, 0.5)),
            mixed_sphere_region((20, 18, 22), 0.4, 7)]
    norb = 9
    rs = [regs[i % 3] for i in range(norb)]
    cache = {id(r): ctx.region(r) for r in regs}
    lrs = [cache[id(r)] for r in rs]
    psi = dev(np.concatenate([random_psi(r, 1, 300 + i)[0] for i, r in enumerate(rs)]))
    pots = [random_pot(r, 400 + i) for i, r in enumerate(rs)]
    pot = dev(np.concatenate(pots))
    kw = dict(pot_off=np.concatenate([[0], np.cumsum([p.size for p in pots])[:-1]]),
              potdim=np.array([r.nfine for r in rs]).ravel(), precision=prec)
    outs = []
    for mb, use_plan in ((0, False), (6, False), (6, True), (1, True)):
        print(f"case chunk_mb={mb} plan={use_plan} precision={prec}", flush=True)
        ctx.set_option("chunk_mb", mb)
        h = torch.zeros_like(psi)
        ek = torch.zeros(norb, dtype=torch.float64, device="cuda")
        ep = torch.zeros_like(ek)
        if use_plan:
            plan = ctx.plan(lrs, pot, psi, h, ek, ep, **kw)
            plan.execute()
            plan.close()
        else:
            ctx.locham_batch(lrs, pot, psi, h, ek, ep, **kw)
        ctx.sync()
        outs.append((h, ek, ep))
    ctx.set_option("chunk_mb", 0)
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert torch.equal(a, b)
    for l in cache.values():
        l.close()
