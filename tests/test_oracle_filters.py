"""Pins the oracle's filter tables by the identities that define them.

The reference holds no kernel-level golden vectors for this path, so each
table copied into the oracle is checked against mathematics that an error in
one digit, a sign, an index shift or a swapped table would break:
  - Daubechies-16 orthonormality (sum ch = sqrt2, double-shift orthogonality)
    and the quadrature-mirror relation cg(l) = (-1)^l ch(1-l)
    (bigdft/src/orbitals/convolut_optim_per.f90:25-42);
  - the combined synthesis+magic table fil2 (bigdft/src/orbitals/v.inc) equals
    the composition of ch/cg with the magic filter
    (convolut_optim_per.f90:537-553) -- an independent table of the reference;
  - the wavelet-basis kinetic tables a, c, e (convolut_ib_optim.f90:962-1038)
    satisfy the two-scale relations a = 4 A_phi K S_phi (a is a fixed point),
    c = 4 A_psi K S_phi, e = 4 A_psi K S_psi, and a has sum 0 and second
    moment 2 (exact second derivative of x^2).
"""
import numpy as np
import pytest

import oracle


@pytest.fixture(scope="module")
def F():
    return oracle.filters()


def test_daubechies_orthonormal(F):
    ch = F["ch"]
    assert abs(ch.sum() - np.sqrt(2.0)) < 1e-15
    assert abs((ch ** 2).sum() - 1.0) < 1e-15
    for m in range(1, 8):
        s = np.dot(ch[:16 - 2 * m], ch[2 * m:])
        assert abs(s) < 1e-15, m


def test_quadrature_mirror(F):
    ch, cg = F["ch"], F["cg"]
    for l in range(-7, 9):
        assert cg[l + 7] == (-1) ** l * ch[(1 - l) + 7]


def test_magic_filter_partition_of_unity(F):
    assert abs(F["magic"].sum() - 1.0) < 1e-15


def _compose_fil2(filt, mag):
    def m(j):
        return mag[j + 8] if -8 <= j <= 7 else 0.0
    return np.array([sum(filt[k + 7] * m(k - j) for k in range(-7, 9))
                     for j in range(-14, 17)])


def test_fil2_is_synthesis_times_magic(F):
    for s, filt in ((0, F["ch"]), (1, F["cg"])):
        comp = _compose_fil2(filt, F["magic"])
        np.testing.assert_allclose(comp, F["fil2"][s], rtol=0, atol=3e-16)


def _sym(half):
    full = np.zeros(29)
    full[14:] = half
    full[:14] = half[1:][::-1]
    return full


def _two_scale(f_out, f_in, a):
    def af(t):
        return a[t + 14] if -14 <= t <= 14 else 0.0
    return np.array([4.0 * sum(f_out[m + 7] * f_in[mp + 7] * af(2 * l + mp - m)
                               for m in range(-7, 9) for mp in range(-7, 9))
                     for l in range(-14, 15)])


def test_kinetic_tables_two_scale(F):
    a = _sym(F["kin_a"])
    e = _sym(F["kin_e"])
    ch, cg = F["ch"], F["cg"]
    np.testing.assert_allclose(_two_scale(ch, ch, a), a, rtol=0, atol=5e-15)
    np.testing.assert_allclose(_two_scale(cg, ch, a), F["kin_c"], rtol=0,
                               atol=5e-15)
    np.testing.assert_allclose(_two_scale(cg, cg, a), e, rtol=0, atol=1e-14)


def test_kinetic_moments(F):
    a = _sym(F["kin_a"])
    l = np.arange(-14, 15)
    assert abs(a.sum()) < 1e-14
    assert abs((a * l ** 2).sum() - 2.0) < 1e-14
    assert abs((a * l).sum()) < 1e-15
