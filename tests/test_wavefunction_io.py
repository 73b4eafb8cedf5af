"""Wavefunction files (writeonewave / read_psi_compress layout,
bigdft/src/modules/io.f90:1616-1780, 2639-2732).

No wavefunction file ships with the reference, so the layout is pinned by
an independent producer: tests/golden/write_wave.f90, compiled with the
image's Fortran compiler, writes one orbital with Fortran's own unformatted
records and edit descriptors.  Checks: our reader recovers its values bit
for bit (binary) and to the printed digits (plain), our binary writer is
byte-identical to it and our plain writer's formatted lines equal its
lines; round trips through our writer on sphere regions return the same
keys (segkeys of the file's points) and the same coefficients.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from bigdft_hip import wavefunction_io as wio
from bigdft_hip.descriptors import region_from_grids
from helpers import mixed_sphere_region, random_psi

HERE = os.path.dirname(os.path.abspath(__file__))
FC = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")


def golden_values():
    n = (5, 3, 4)
    i1, i2, i3 = np.meshgrid(*[np.arange(x + 1) for x in n], indexing="ij")
    coarse = (i1 + 2 * i2 + 3 * i3) % 4 != 0
    fine = coarse & ((i1 * i2 + i3) % 3 == 0)
    reg = region_from_grids("F", n, (0.45, 0.4, 0.5), coarse, fine)
    k = np.arange(1, reg.nvctr_c + 1)
    vc = ((k * 37) % 101 - 50) / 64.0 * 2.0 ** ((k % 5) - 2)
    kf = np.arange(1, reg.nvctr_f + 1)[:, None]
    j = np.arange(1, 8)[None, :]
    vf = (((7 * kf + j) * 53) % 97 - 48) / 32.0 * 2.0 ** (((kf + j) % 7) - 3)
    return reg, np.concatenate([vc, vf.ravel()])


@pytest.fixture(scope="module")
def fortran_files(tmp_path_factory):
    if not os.path.exists(FC):
        pytest.skip("no Fortran compiler")
    d = tmp_path_factory.mktemp("wave")
    exe = d / "write_wave"
    r = subprocess.run([FC, "-O0", os.path.join(HERE, "golden", "write_wave.f90"), "-o",
                        str(exe)], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("Fortran build failed: " + r.stderr[-500:])
    fb = d / wio.wfn_filename("wavefunction", 7, binary=True)
    ft = d / wio.wfn_filename("wavefunction", 7, binary=False)
    subprocess.run([str(exe), str(fb), str(ft)], check=True)
    return fb, ft


def test_reads_fortran_binary(fortran_files, tmp_path):
    fb, _ = fortran_files
    reg, vals = golden_values()
    w = wio.read_one_wave(str(fb))
    assert w["iorb"] == 7 and w["eval"] == -0.3125
    assert w["hgrid"] == (0.45, 0.4, 0.5) and w["n"] == (5, 3, 4)
    assert np.array_equal(w["rxyz"], [[1, 2, 3], [-0.5, 0.25, 1.125]])
    assert np.array_equal(w["psi"], vals)
    r2 = wio.region_of_wave(w, "F")
    for k in ("keyg_c", "keyv_c", "keyg_f", "keyv_f"):
        assert np.array_equal(getattr(r2, k), getattr(reg, k)), k
    # our binary writer reproduces the compiler's records byte for byte
    out = tmp_path / "ours.bin.b0007"
    wio.write_one_wave(str(out), reg, vals, iorb=7, eval=-0.3125, rxyz=w["rxyz"], binary=True)
    assert out.read_bytes() == fb.read_bytes()


def test_reads_fortran_text(fortran_files, tmp_path):
    _, ft = fortran_files
    _, vals = golden_values()
    w = wio.read_one_wave(str(ft))
    assert w["iorb"] == 7 and w["n"] == (5, 3, 4) and w["rxyz"].shape == (2, 3)
    nc = len(w["coords_c"])
    assert np.allclose(w["psi"][:nc], vals[:nc], rtol=1e-12, atol=0)
    assert np.allclose(w["psi"][nc:], vals[nc:], rtol=1e-10, atol=0)
    # the formatted lines (rxyz and points) equal the compiler's
    reg, _ = golden_values()
    out = tmp_path / "ours.b0007"
    wio.write_one_wave(str(out), reg, vals, iorb=7, eval=-0.3125, rxyz=w["rxyz"], binary=False)
    theirs = ft.read_text().splitlines()
    ours = out.read_text().splitlines()
    assert len(theirs) == len(ours)
    fixed = list(range(4, 6)) + list(range(7, len(ours)))
    for i in fixed:
        assert ours[i] == theirs[i], (i, ours[i], theirs[i])


@pytest.mark.parametrize("binary", [True, False])
def test_round_trip_sphere(tmp_path, binary):
    reg = mixed_sphere_region((18, 16, 20), 0.4, 7)
    psi = random_psi(reg, 3, 4)
    paths = []
    for i in range(3):
        p = tmp_path / wio.wfn_filename("wavefunction", i + 1, binary=binary)
        wio.write_one_wave(str(p), reg, psi[i], iorb=i + 1, eval=-0.1 * i, binary=binary)
        paths.append(str(p))
    r2, got, ev = wio.load_wavefunctions(paths, "F")
    for k in ("keyg_c", "keyv_c", "keyg_f", "keyv_f", "nfl", "nfu"):
        assert np.array_equal(getattr(r2, k), getattr(reg, k)), k
    if binary:
        assert np.array_equal(got, psi)
    else:
        nc = reg.nvctr_c
        assert np.allclose(got[:, :nc], psi[:, :nc], rtol=1e-11, atol=0)
        assert np.allclose(got[:, nc:], psi[:, nc:], rtol=1e-9, atol=0)
    assert np.allclose(ev, [0.0, -0.1, -0.2])


def test_names_and_errors(tmp_path):
    assert wio.wfn_filename("wavefunction", 12, ikpt=3, spin="D", binary=True) == \
        "wavefunction-k003-DR.bin.b0012"
    assert wio.parse_wfn_filename("data/wavefunction-k001-NI.b0005") == (5, 1, "N", True, False)
    with pytest.raises(ValueError):
        wio.parse_wfn_filename("psi.dat")
    assert wio.fortran_e(0.0, 19, 12) == " 0.000000000000E+00"
    assert wio.fortran_e(-1.5e-120, 17, 10) == "-0.1500000000-119"
    reg = mixed_sphere_region((10, 9, 11), 0.45, 2)
    p = tmp_path / "wavefunction-k001-NR.bin.b0001"
    wio.write_one_wave(str(p), reg, random_psi(reg, 1, 1)[0])
    data = bytearray(p.read_bytes())
    data[-3] ^= 0x40                       # break the last record's end marker
    p.write_bytes(bytes(data))
    with pytest.raises(ValueError):
        wio.read_one_wave(str(p))
