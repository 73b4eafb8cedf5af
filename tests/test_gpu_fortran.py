"""The Fortran iso_c_binding module (bigdft-suite_amd/fortran) drives the
library like local_hamiltonian_OCL + finish_hamiltonian_OCL would; its
weighted energies must match the oracle on the same inputs."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from helpers import dense_region, okeys

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "bigdft-suite_amd", "fortran", "locham_hip_driver")


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built (no amdflang)")
@pytest.mark.parametrize("n,norb,lines,nspin", [(11, 3, 0, 1), (23, 2, 0, 1), (23, 2, 1, 1),
                                                (11, 4, 0, 2), (23, 3, 1, 2)])
def test_fortran_driver_matches_oracle(n, norb, lines, nspin):
    r = subprocess.run([DRIVER, str(n), str(norb), "0", str(lines), str(nspin)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    vals = {l.split()[0]: float(l.split()[1]) for l in r.stdout.splitlines() if l.strip()}
    reg = dense_region("P", (n, n, n), (0.45,) * 3)
    i = np.arange(1, reg.ncoef + 1, dtype=np.float64)
    psi = np.stack([np.sin(0.37 * i + 1.3 * o) * np.cos(0.011 * i)
                    for o in range(1, norb + 1)])
    occup = 2.0 / np.arange(1, norb + 1)
    ip = np.arange(1, reg.n_isf + 1, dtype=np.float64)
    pot = -1.0 - np.cos(0.05 * ip)
    # nspin 2: orbitals 1,3,.. read pot(:,1), orbitals 2,4,.. pot(:,2)
    # (rhopot_up / rhopot_down by spinsgn, interface_ocl.f90:487-491)
    pot_dn = -0.5 - np.sin(0.07 * ip)
    ek, ep = np.zeros(norb), np.zeros(norb)
    for o in range(norb):
        v = pot_dn if (nspin == 2 and o % 2 == 1) else pot
        _, ek[o:o + 1], ep[o:o + 1] = oracle.locham("P", reg.n, reg.nfl, reg.nfu, reg.hgrid,
                                                    okeys(reg), v, psi[o:o + 1])
    assert abs(vals["ekin_sum"] - (occup * ek).sum()) < 1e-12 * abs((occup * ek).sum())
    assert abs(vals["epot_sum"] - (occup * ep).sum()) < 1e-12 * abs((occup * ep).sum())
    assert abs(vals["psi_hpsi"] - (ek + ep).sum()) < 1e-11 * abs((ek + ep).sum())
    # density: orbitals 1,3,.. spin up, 2,4,.. down; hfac = occup/0.225^3
    hf = occup / 0.225 ** 3
    sp = np.arange(norb, dtype=np.int32) % 2
    rho = oracle.density("P", reg.n, reg.nfl, reg.nfu, okeys(reg), psi, hf, sp, nspin=2)
    up = np.dot(rho[0], pot)
    assert abs(vals["rho_up_pot"] - up) < 1e-12 * abs(up)
    dn = rho[1].sum() + 1e-20 * reg.n_isf
    assert abs(vals["rho_dn_sum"] - dn) < 1e-12 * abs(dn)
