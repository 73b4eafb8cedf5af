"""Pinning of the oracle's preconditioner (the GPU part of
preconditionall_OCL: ocl_preconditioner_generic_k, bigdft/src/OpenCL/
Preconditioner.c:198-358, real case).

The operator A = S Ana (K + c) Syn S is restated literally (fine-grid
kinetic between synthesis and analysis; free BC on the grid grown by 7
coarse points per side).  Checks:
  * A x = S (T S x + c S x) with T the kinetic of the oracle's local
    Hamiltonian at zero potential (periodic: the same fine-grid kinetic
    between the periodic transforms; free: ConvolkineticT, the wavelet-basis
    filters), so the grown/shrunk free transforms and the kinetic between
    them are checked against an independent formulation;
  * A is symmetric and its diagonal is 1: A2 and B2 of the scaling are the
    diagonal elements of the scaling- and wavelet-function kinetic, so the
    scaling is a Jacobi preconditioner (periodic boxes of at least 16 coarse
    points, so that the 29-tap fine kinetic does not wrap onto itself);
  * the conjugate-gradient loop solves A x = b (residual 1e-9 after 80
    iterations, shift 3), ncong = 0 returns S x, and each iteration lowers the error
    in the A-norm.
"""
import numpy as np
import pytest

import oracle
from helpers import dense_region, mixed_sphere_region, okeys, random_psi

CASES = [("P", dense_region("P", (15, 16, 15), (0.45, 0.4, 0.5))),
         ("F", mixed_sphere_region((10, 12, 9), 0.45, 5))]


def scale_vec(reg, scal, x):
    y = np.array(x, copy=True)
    y[:reg.nvctr_c] *= scal[0]
    y[reg.nvctr_c:] = (y[reg.nvctr_c:].reshape(-1, 7) * scal[1:]).ravel()
    return y


@pytest.mark.parametrize("name,reg", CASES, ids=[c[0] for c in CASES])
def test_operator_is_scaled_hamiltonian_kinetic(name, reg):
    c = 0.37
    x = random_psi(reg, 1, 3)[0]
    scal = oracle.precond_scal(reg.hgrid, c)
    y = oracle.precond_apply(reg.geocode, reg.n, reg.hgrid, c, okeys(reg), x)
    sx = scale_vec(reg, scal, x)
    zero = np.zeros(reg.n_isf)
    t, _, _ = oracle.locham(reg.geocode, reg.n, reg.nfl, reg.nfu, reg.hgrid, okeys(reg), zero,
                            sx)
    exp = scale_vec(reg, scal, t[0] + c * sx)
    assert np.abs(y - exp).max() < 1e-12 * np.abs(exp).max()


@pytest.mark.parametrize("name,reg", CASES, ids=[c[0] for c in CASES])
def test_operator_symmetric_unit_diagonal(name, reg):
    c = 0.2
    u, v = random_psi(reg, 2, 8)
    k = okeys(reg)
    au = oracle.precond_apply(reg.geocode, reg.n, reg.hgrid, c, k, u)
    av = oracle.precond_apply(reg.geocode, reg.n, reg.hgrid, c, k, v)
    assert abs(u @ av - v @ au) < 1e-12 * abs(u @ av)
    rng = np.random.default_rng(1)
    idx = list(rng.choice(reg.nvctr_c, 3, replace=False)) + \
        list(reg.nvctr_c + rng.choice(7 * reg.nvctr_f, 7, replace=False))
    for i in idx:
        e = np.zeros(reg.ncoef)
        e[i] = 1.0
        ae = oracle.precond_apply(reg.geocode, reg.n, reg.hgrid, c, k, e)
        assert abs(ae[i] - 1.0) < 1e-12
        assert np.abs(ae).max() <= 1.0 + 1e-12


@pytest.mark.parametrize("name,reg", CASES, ids=[c[0] for c in CASES])
def test_conjugate_gradient_solves_the_system(name, reg):
    c = 3.0     # condition number ~ 20: 80 iterations reach 1e-9
    k = okeys(reg)
    rng = np.random.default_rng(4)
    b = rng.standard_normal(reg.ncoef)
    x0 = rng.standard_normal(reg.ncoef) * 0.1
    scal = oracle.precond_scal(reg.hgrid, c)
    assert np.array_equal(oracle.precond_cg(reg.geocode, reg.n, reg.hgrid, c, 0, k, x0, b),
                          scale_vec(reg, scal, x0))
    xs = oracle.precond_cg(reg.geocode, reg.n, reg.hgrid, c, 80, k, x0, b)
    x = scale_vec(reg, 1.0 / scal, xs)
    res = oracle.precond_apply(reg.geocode, reg.n, reg.hgrid, c, k, x) - b
    assert np.linalg.norm(res) < 1e-9 * np.linalg.norm(b)
    # the A-norm error decreases monotonically with the iteration count
    errs = []
    for it in (1, 2, 4, 8):
        xi = scale_vec(reg, 1.0 / scal,
                       oracle.precond_cg(reg.geocode, reg.n, reg.hgrid, c, it, k, x0, b)) - x
        errs.append(xi @ oracle.precond_apply(reg.geocode, reg.n, reg.hgrid, c, k, xi))
    assert all(errs[i + 1] < errs[i] for i in range(len(errs) - 1))
