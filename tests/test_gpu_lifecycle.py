"""Handle lifetimes through the C ABI: a context destroyed before its
regions and captured plans releases their device memory and detaches them;
destroying them afterwards frees only the host objects, and executing a
detached plan is BDFT_ERR_ARG instead of a use of freed memory."""
import ctypes

import numpy as np
import pytest

from helpers import dense_region, random_psi

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def test_context_destroyed_before_its_regions_and_plans():
    import bigdft_hip
    lib = bigdft_hip.lib()
    ctx = bigdft_hip.Context(0)
    reg = dense_region("P", (11, 11, 11), (0.45,) * 3)
    lr = ctx.region(reg)
    psi = torch.from_numpy(random_psi(reg, 2, 3).ravel()).cuda()
    pot = torch.full((reg.n_isf,), -1.0, dtype=torch.float64, device="cuda")
    h = torch.zeros_like(psi)
    ek = torch.zeros(2, dtype=torch.float64, device="cuda")
    ep = torch.zeros_like(ek)
    plan = ctx.plan([lr, lr], pot, psi, h, ek, ep)
    plan.execute()
    ctx.sync()
    assert float(ek.abs().min()) > 0
    print("  context destroyed first", flush=True)
    ctx.close()
    assert lib.bdft_plan_execute(plan.handle) == 1
    nf = (ctypes.c_int * 3)()
    nc = ctypes.c_int64()
    assert lib.bdft_lr_info(lr.handle, nf, ctypes.byref(nc)) == 0
    assert tuple(nf) == reg.nfine and nc.value == reg.ncoef
    plan.close()
    lr.close()
    # a new context on the device works as before
    ctx2 = bigdft_hip.Context(0)
    lr2 = ctx2.region(reg)
    h2 = torch.zeros_like(psi)
    ek2 = torch.zeros_like(ek)
    ctx2.locham_batch([lr2, lr2], pot, psi, h2, ek2, torch.zeros_like(ek))
    ctx2.sync()
    assert torch.equal(h2, h) and torch.equal(ek2, ek)
    lr2.close()
    ctx2.close()
