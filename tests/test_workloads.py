"""Host-side checks of the synthetic batch generators (small sizes)."""
import numpy as np
import torch

from bigdft_hip import inputs, workloads


def test_uneven_free_layout():
    b = workloads.uneven_free(12, 3, device="cpu")
    regs = b["regions"]
    assert len(regs) == 12
    for r in regs:
        assert r.geocode == "F" and all(15 <= n <= 39 for n in r.n)
        assert r.nvctr_c == r.nvctr_f == int(np.prod([n + 1 for n in r.n]))
    assert b["psi"].numel() == int(b["ncoef"].sum())
    assert np.array_equal(b["psi_off"][1:], np.cumsum(b["ncoef"])[:-1])
    assert b["pot"].numel() == b["units"] == sum(r.n_isf for r in regs)
    # the window of box 2 is its own potential array
    p0 = int(b["pot_off"][2])
    lp = workloads.local_potential(b, 2)
    assert np.array_equal(lp, b["pot"][p0:p0 + regs[2].n_isf].numpy())
    # envelope: coarse coefficient at the box centre carries weight ~1, the
    # corner ones exp(-r^2/32)
    env = inputs.envelope(regs[0], 4.0).numpy()
    n, h = regs[0].n, regs[0].hgrid
    r2 = sum((0.5 * n[d] * h[d]) ** 2 for d in range(3))
    assert abs(env[0] - np.exp(-r2 / 32.0)) < 1e-14


def test_sphere_batch_potential_window():
    b = workloads.sphere_batch(6, 11, n_glob=40, device="cpu")
    gd = b["global_nfine"]
    assert b["pot"].numel() == int(np.prod(gd))
    gp = inputs.gaussian_params(12, [41 * 0.4] * 3)
    for i, (r, nl) in enumerate(zip(b["regions"], b["nl"])):
        lp = workloads.local_potential(b, i).reshape(r.nfine[::-1])
        # inside the global fine grid: the field at the local fine points;
        # outside: zero (the potential box ends there)
        xyz = inputs.fine_coords("F", r.n, r.hgrid,
                                 origin=tuple(nl[d] * r.hgrid[d] for d in range(3)))
        ref = inputs.gaussian_field(gp, xyz).numpy().reshape(r.nfine[::-1])
        inside = np.ones(r.nfine[::-1], dtype=bool)
        for d in range(3):
            q = np.arange(r.nfine[d]) + 2 * nl[d]
            ok = (q >= 0) & (q < gd[d])
            sl = [None, None, None]
            sl[2 - d] = slice(None)
            inside &= ok[tuple(sl)]
        assert np.allclose(lp[inside], ref[inside], rtol=0, atol=1e-12)
        assert np.all(lp[~inside] == 0.0)
        seg = b["psi"][b["psi_off"][i]:b["psi_off"][i] + b["ncoef"][i]]
        assert abs(float(torch.linalg.vector_norm(seg)) - 1.0) < 1e-12
