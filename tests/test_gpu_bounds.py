"""Convolution bounds (locreg_bounds) through the C ABI vs the oracle:
bit-exact integers, device and host entry points."""
import numpy as np
import pytest

import oracle
from helpers import mixed_sphere_region, okeys
from test_oracle_bounds import REGIONS, irregular_region

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import bigdft_hip
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    c = bigdft_hip.Context(0)
    yield c
    c.close()


CASES = dict(REGIONS)
CASES["sphere_40"] = lambda: mixed_sphere_region((40, 37, 33), 0.4, 21)
CASES["irregular_big"] = lambda: irregular_region((45, 30, 52), 9, pc=0.2)


@pytest.mark.parametrize("name", list(CASES))
def test_bounds_bit_exact(ctx, name):
    import bigdft_hip
    reg = CASES[name]()
    print(f"bounds {name} n={reg.n} nfl={reg.nfl} nfu={reg.nfu}", flush=True)
    ref = oracle.locreg_bounds(reg.n, reg.nfl, reg.nfu, okeys(reg))
    lr = ctx.region(reg)
    host = ctx.locreg_bounds(lr)
    assert host.shape == ref.shape
    bad = [k for k, (a, b) in enumerate(zip(oracle.unpack_bounds(host, reg.n, reg.nfl, reg.nfu).values(),
                                            oracle.unpack_bounds(ref, reg.n, reg.nfl, reg.nfu).values()))
           if not np.array_equal(a[0], b[0])]
    assert not bad, [bigdft_hip.BOUNDS_NAMES[k] for k in bad]
    dev = torch.full((ref.size,), 7, dtype=torch.int32, device="cuda")
    ctx.locreg_bounds(lr, out=dev)
    ctx.sync()
    assert np.array_equal(dev.cpu().numpy(), ref)
    lr.close()


def test_bounds_errors(ctx):
    import bigdft_hip
    from bigdft_hip.descriptors import dense_region
    lr = ctx.region(dense_region("P", (7, 7, 7), (0.4,) * 3))
    with pytest.raises(bigdft_hip.BigDFTError, match="UNSUPPORTED"):
        ctx.locreg_bounds(lr)
    lr.close()
