"""Pins the oracle's operators by mathematics (no reference vectors exist).

Each check is chosen so that a plausible restatement mistake fails it:
  - wavelet synthesis/analysis of convolut_optim_per.f90:10-304 reconstruct
    perfectly and are each other's transpose (orthonormal D16 basis), and the
    synthesis matches the dense refinement-relation matrix;
  - the magic filter pair convrot_n_per / convrot_t_per (:519-766) are
    adjoint;
  - the periodic kinetic (:1169-1629) has the exact plane-wave symbol and
    gives 1/(4 sigma^2) for a Gaussian;
  - the periodic chain (synthesis, magic, potential) reproduces sampled point
    values exactly: epot of a Gaussian equals sum V psi(x)^2, and ekin equals
    the analytic 3/(4 sigma^2) of a 3D Gaussian;
  - the full local Hamiltonian (P and F) is symmetric, its energies satisfy
    <psi|H psi> = ekin + epot and epot = <psi|H(V) psi - H(0) psi>;
  - the free-BC restatement (dense form of comb_grow_all / comb_shrink /
    ConvolkineticT) agrees with the periodic one on a localised orbital.
"""
import numpy as np
import pytest

import oracle
from helpers import dense_region, field, mixed_sphere_region, okeys, random_psi


def test_synthesis_analysis_perfect_reconstruction():
    rng = np.random.default_rng(1)
    for n in (0, 3, 7, 12):
        m = 2 * n + 2
        x = rng.standard_normal((m, 5))
        y = oracle.syn_rot_per(n, x)           # (5, m)
        back = oracle.ana_rot_per(n, y.T)      # (5, m)
        np.testing.assert_allclose(back.T, x, rtol=0, atol=2e-15)
        z = rng.standard_normal((m, 4))
        assert abs(np.sum(oracle.syn_rot_per(n, x)[:4] * z.T)
                   - np.sum(x[:, :4] * oracle.ana_rot_per(n, z).T)) < 1e-13


def test_synthesis_matches_refinement_matrix():
    F = oracle.filters()
    n = 9
    N, m = n + 1, 2 * n + 2
    S = np.zeros((m, m))
    for k in range(N):
        for mm in range(-7, 9):
            S[(2 * k + mm) % m, k] += F["ch"][mm + 7]
            S[(2 * k + mm) % m, N + k] += F["cg"][mm + 7]
    x = np.random.default_rng(2).standard_normal((m, 3))
    np.testing.assert_allclose(oracle.syn_rot_per(n, x).T, S @ x, atol=2e-15)


def test_magic_filter_adjoint():
    rng = np.random.default_rng(3)
    for n1 in (5, 17, 47):
        x = rng.standard_normal((n1 + 1, 6))
        y = rng.standard_normal((n1 + 1, 6))
        lhs = np.sum(oracle.convrot_n_per(n1, x).T * y)
        rhs = np.sum(x * oracle.convrot_t_per(n1, y).T)
        assert abs(lhs - rhs) < 1e-12 * max(1.0, abs(lhs))


def test_kinetic_plane_wave_symbol():
    F = oracle.filters()
    a = np.concatenate([F["kin_a"][1:][::-1], F["kin_a"]])
    M, h = 64, 0.3
    i = np.arange(M)
    for k in (1, 2, 3, 5):
        x = np.cos(2 * np.pi * k * i / M)
        y, _ = oracle.kinetic_per(x.reshape(M, 1, 1), (h, h, h))
        lam = -0.5 / h ** 2 * np.sum(a * np.cos(2 * np.pi * k * np.arange(-14, 15) / M))
        np.testing.assert_allclose(y.ravel(), lam * x, atol=1e-12 * abs(lam))
        q = 2 * np.pi * k / (M * h)
        assert abs(lam - 0.5 * q * q) < 1e-8 * 0.5 * q * q


def _magic_symbol(M):
    F = oracle.filters()
    w = np.zeros(M)
    for l in range(-8, 8):
        w[l % M] += F["magic"][l + 8]
    # (M s)(i) = sum_l fil(l) s(i+l): eigenvalue sum_l fil(l) e^{+i q l}
    return np.conj(np.fft.fft(w))


def test_gaussian_kinetic_1d():
    M, h, sig = 96, 0.35, 1.4
    x = (np.arange(M) - M / 2) * h
    f = np.exp(-x ** 2 / (2 * sig ** 2))
    s = np.real(np.fft.ifft(np.fft.fft(f) / _magic_symbol(M)))
    y, _ = oracle.kinetic_per(s.reshape(M, 1, 1), (h, h, h))
    ratio = np.dot(s, y.ravel()) / np.dot(s, s)
    assert abs(ratio - 1.0 / (4 * sig ** 2)) < 1e-7 / (4 * sig ** 2)


def _gaussian_wavelets(n, h, sig):
    """wavelet coefficients (dense periodic keys order) of a centred Gaussian,
    and its point values on the fine grid"""
    M = 2 * n + 2
    x = (np.arange(M) - M / 2) * (0.5 * h)
    g = np.exp(-x ** 2 / (2 * sig ** 2))
    f3 = g[:, None, None] * g[None, :, None] * g[None, None, :]
    w = _magic_symbol(M)
    s3 = np.real(np.fft.ifftn(np.fft.fftn(f3) / (w[:, None, None] * w[None, :, None]
                                                 * w[None, None, :])))
    psig = np.zeros(M ** 3)
    buf = np.ascontiguousarray(s3.ravel(order="F"))
    oracle.lib().oref_analyse_per_self(n, n, n, buf, psig)
    # psig (0:n,2,0:n,2,0:n,2) -> dense compressed order [c ; f(7,...)]
    P = psig.reshape((n + 1, 2, n + 1, 2, n + 1, 2), order="F")
    c = P[:, 0, :, 0, :, 0].ravel(order="F")
    fine = np.stack([P[:, t & 1, :, (t >> 1) & 1, :, (t >> 2) & 1].ravel(order="F")
                     for t in range(1, 8)], axis=1).ravel()
    return np.concatenate([c, fine]), f3.ravel(order="F")


def test_periodic_chain_reproduces_gaussian():
    n, h, sig = 23, 0.45, 1.2
    reg = dense_region("P", (n, n, n), (h, h, h))
    psi, f = _gaussian_wavelets(n, h, sig)
    rng = np.random.default_rng(4)
    V = rng.uniform(-2, 0, size=f.size)
    _, ek, ep = oracle.locham("P", reg.n, reg.nfl, reg.nfu, reg.hgrid,
                              okeys(reg), V, psi)
    np.testing.assert_allclose(ep[0], np.sum(V * f * f), rtol=1e-13)
    norm = np.dot(psi, psi)
    assert abs(ek[0] / norm - 3.0 / (4 * sig ** 2)) < 1e-7


@pytest.mark.parametrize("geo", ["P", "F"])
def test_energy_identities_and_symmetry(geo):
    n = (7, 6, 8)
    h = (0.45, 0.4, 0.5)
    reg = dense_region(geo, n, h) if geo == "P" else mixed_sphere_region(n, 0.45, 5)
    if geo == "F":
        h = reg.hgrid
    psi = random_psi(reg, 2, 6)
    rng = np.random.default_rng(7)
    V = rng.uniform(-2, 0, size=reg.n_isf)
    run = lambda pot, p: oracle.locham(geo, reg.n, reg.nfl, reg.nfu, h, okeys(reg), pot, p)
    hp, ek, ep = run(V, psi)
    h0, ek0, ep0 = run(np.zeros_like(V), psi)
    for i in range(2):
        e = np.dot(psi[i], hp[i])
        assert abs(e - (ek[i] + ep[i])) < 1e-12 * abs(e)
        assert ep0[i] == 0.0
        assert abs(np.dot(psi[i], h0[i]) - ek0[i]) < 1e-12 * abs(ek0[i])
        assert abs(ek[i] - ek0[i]) < 1e-13 * abs(ek[i])
        assert abs(np.dot(psi[i], hp[i] - h0[i]) - ep[i]) < 1e-12 * abs(ep[i])
    # symmetric operator: <psi0|H psi1> = <H psi0|psi1>
    a, b = np.dot(psi[0], hp[1]), np.dot(hp[0], psi[1])
    assert abs(a - b) < 1e-12 * max(abs(a), 1.0)


def test_free_matches_periodic_on_localised_orbital():
    n, h = (39, 39, 39), (0.4, 0.4, 0.4)
    rP = dense_region("P", n, h)
    rF = dense_region("F", n, h)
    rng = np.random.default_rng(8)
    m = np.zeros((40, 40, 40), dtype=bool)
    m[15:25, 15:25, 15:25] = True
    mask = np.concatenate([m.ravel(order="F"), np.repeat(m.ravel(order="F"), 7)])
    psi = rng.standard_normal(rP.ncoef) * mask
    fun = lambda X, Y, Z: -np.exp(-((X - 8) ** 2 + (Y - 8.5) ** 2 + (Z - 7.5) ** 2) / 9.0)
    VP = field("P", n, h, fun)
    VF = field("F", n, h, fun)
    hP, eP, pP = oracle.locham("P", n, rP.nfl, rP.nfu, h, okeys(rP), VP, psi)
    hF, eF, pF = oracle.locham("F", n, rF.nfl, rF.nfu, h, okeys(rF), VF, psi)
    np.testing.assert_allclose(hF, hP, rtol=0, atol=1e-12 * np.abs(hP).max())
    assert abs(eF[0] - eP[0]) < 1e-12 * abs(eP[0])
    assert abs(pF[0] - pP[0]) < 1e-12 * abs(pP[0])


def test_free_fine_subbox_matches_full_fine_region():
    """free BC with the fine region in a sub-box (nfl != 0): the oracle
    places x_f(7, nfl1:nfu1, ...) by the region's fine bounds
    (uncompress_forstandard, un-compress.f90:286-360; ConvolkineticT and
    comb_grow/shrink bounds, convolut_ib_optim.f90:921-1063).  The same
    orbital stored in a region whose fine grid is the whole box (nfl = 0)
    is the same function, so H psi at the sphere's points is identical."""
    n, h = (14, 12, 16), 0.4
    sph = mixed_sphere_region(n, h, 9)
    assert sph.nfl != (0, 0, 0) and sph.nfu != n
    full = dense_region("F", n, (h, h, h))
    rng = np.random.default_rng(10)
    psi_s = rng.standard_normal(sph.ncoef)
    # the same coefficients in the full region's compressed layout
    psig = oracle.uncompress_psig(n, okeys(sph), psi_s)
    psi_f = oracle.compress_acc_mixed(n, okeys(full), psig, np.zeros(full.ncoef))
    rng2 = np.random.default_rng(11)
    V = rng2.uniform(-2, 0, size=sph.n_isf)
    hs, es, ps = oracle.locham("F", n, sph.nfl, sph.nfu, sph.hgrid, okeys(sph), V,
                               psi_s[None, :])
    hf, ef, pf = oracle.locham("F", n, full.nfl, full.nfu, full.hgrid, okeys(full), V,
                               psi_f[None, :])
    # hpsi of the full region at the sphere's points
    hg = oracle.uncompress_psig(n, okeys(full), hf[0])
    h_at = oracle.compress_acc_mixed(n, okeys(sph), hg, np.zeros(sph.ncoef))
    np.testing.assert_allclose(hs[0], h_at, rtol=0, atol=1e-13 * np.abs(h_at).max())
    assert abs(es[0] - ef[0]) < 1e-13 * abs(ef[0])
    assert abs(ps[0] - pf[0]) < 1e-13 * abs(pf[0])
