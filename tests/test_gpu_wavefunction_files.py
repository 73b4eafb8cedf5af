"""Orbitals read from BigDFT wavefunction files (binary) through the
Hamiltonian on the GPU: the region rebuilt from the files' points and the
coefficients must give the oracle's hpsi and energies (FP64, 1e-12)."""
import numpy as np
import pytest

import oracle
from bigdft_hip import wavefunction_io as wio
from helpers import mixed_sphere_region, okeys, random_pot, random_psi

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def test_hamiltonian_on_loaded_orbitals(tmp_path):
    import bigdft_hip
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    reg0 = mixed_sphere_region((26, 23, 25), 0.4, 19)
    psi0 = random_psi(reg0, 4, 21)
    paths = []
    for i in range(4):
        p = tmp_path / wio.wfn_filename("wavefunction", i + 1, binary=True)
        wio.write_one_wave(str(p), reg0, psi0[i], iorb=i + 1, eval=-0.2 - 0.1 * i)
        paths.append(str(p))
    reg, psi, _ = wio.load_wavefunctions(paths, "F")
    pot = random_pot(reg, 8)
    ctx = bigdft_hip.Context(0)
    lr = ctx.region(reg)
    h = torch.zeros(psi.size, dtype=torch.float64, device="cuda")
    ek = torch.zeros(4, dtype=torch.float64, device="cuda")
    ep = torch.zeros_like(ek)
    ctx.locham_batch([lr] * 4, torch.from_numpy(pot).cuda(),
                     torch.from_numpy(psi.ravel()).cuda(), h, ek, ep)
    ctx.sync()
    ho, eo, po = oracle.locham("F", reg.n, reg.nfl, reg.nfu, reg.hgrid, okeys(reg), pot, psi)
    hg = h.cpu().numpy().reshape(4, -1)
    assert np.abs(hg - ho).max() < 1e-12 * np.abs(ho).max()
    assert np.allclose(ek.cpu().numpy(), eo, rtol=1e-12, atol=0)
    assert np.allclose(ep.cpu().numpy(), po, rtol=1e-12, atol=0)
    lr.close()
    ctx.close()
