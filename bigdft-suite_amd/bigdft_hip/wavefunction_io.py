"""BigDFT wavefunction files: one compressed orbital per file.

Format of writeonewave (bigdft/src/modules/io.f90:2639-2732), read back by
io_read_descr / read_psi_compress (io.f90:1616-1780) and readonewave
(bigdft/src/wavelib/i-o.f90:298-410):

    iorb, eval                  (int, real)
    hx, hy, hz                  (coarse grid spacings)
    n1, n2, n3
    nat
    rxyz(1:3, iat)              nat lines, format (3(1x,e24.17))
    nvctr_c, nvctr_f
    i1, i2, i3, psi_c           nvctr_c lines, format (3(i4),1x,e19.12)
    i1, i2, i3, psi_f(1:7)      nvctr_f lines, format (3(i4),7(1x,e17.10))

The points come in key order (segment after segment, i1 fastest).  The
binary variant writes the same items as Fortran unformatted sequential
records: each record is framed by its byte length as a 4-byte integer,
integers are 4 bytes and reals 8 bytes, little endian.  File names follow
wfn_filename (bigdft/src/restart.f90:598-618):
<base>-k<ikpt:03>-<N|U|D|A|B><R|I>.b<iorb:04> (plain) or .bin.b<iorb:04>.

Reading needs the same grid as the writer (readonewave's direct copy,
i-o.f90:347-354); a different grid needs reformatonewave's interpolation,
which is not part of this module.
"""
import os
import re
import struct

import numpy as np

from .descriptors import region_from_grids

_I4 = struct.Struct("<i")


def wfn_filename(base, iorb, ikpt=1, spin="N", imag=False, binary=True):
    """file name of one orbital (wfn_filename, restart.f90:598-618)"""
    kind = "I" if imag else "R"
    ext = ".bin.b" if binary else ".b"
    return f"{base}-k{ikpt:03d}-{spin}{kind}{ext}{iorb:04d}"


_NAME = re.compile(r"-k(\d{3})-([NUDAB])([RI])\.(bin\.)?b(\d+)$")


def parse_wfn_filename(path):
    """(iorb, ikpt, spin code, imaginary, binary) from a file name, as
    readwavedescr reads them (bigdft/src/wavelib/i-o.f90:413-470)"""
    m = _NAME.search(os.path.basename(path))
    if not m:
        raise ValueError(f"not a wavefunction file name: {path}")
    return (int(m.group(5)), int(m.group(1)), m.group(2), m.group(3) == "I",
            m.group(4) is not None)


def fortran_e(x, w, d):
    """Fortran Ew.d edit of x: sign, '0.', d digits, exponent E+xx (two
    exponent digits, or three without the E beyond 99)"""
    x = float(x)
    if x == 0.0 or not np.isfinite(x):
        m, e = ("0." + "0" * d, 0) if x == 0.0 else (None, None)
        if m is None:
            return f"{x}".rjust(w)
        s = f"{m}E+00"
        return s.rjust(w)
    s = f"{abs(x):.{d - 1}E}"           # d.ddddE+xx: mantissa in [1, 10)
    mant, ex = s.split("E")
    digits = mant.replace(".", "")
    e = int(ex) + 1
    if len(digits) > d:                 # rounding carried into a new digit
        digits = digits[:d]
    body = "0." + digits
    if abs(e) <= 99:
        es = f"E{'+' if e >= 0 else '-'}{abs(e):02d}"
    else:
        es = f"{'+' if e >= 0 else '-'}{abs(e):03d}"
    out = ("-" if x < 0 else "") + body + es
    if len(out) > w:
        return "*" * w
    return out.rjust(w)


def _list_line(*vals):
    """a list-directed output line (the layout of one compiler; readers
    split on blanks)"""
    parts = []
    for v in vals:
        if isinstance(v, (int, np.integer)):
            parts.append(f"{int(v):12d}")
        else:
            parts.append(f"{float(v):25.16E}")
    return " " + " ".join(parts) + "\n"


def _points(region):
    """(coords_c (nvctr_c, 3), coords_f (nvctr_f, 3)) in key order"""
    out = []
    n1, n2 = region.n[0] + 1, region.n[1] + 1
    for keyg, nseg in ((region.keyg_c, region.nseg_c), (region.keyg_f, region.nseg_f)):
        pts = []
        for s in range(nseg):
            j0, j1 = int(keyg[2 * s]), int(keyg[2 * s + 1])
            lin = np.arange(j0 - 1, j1)
            i1 = lin % n1
            i2 = (lin // n1) % n2
            i3 = lin // (n1 * n2)
            pts.append(np.stack([i1, i2, i3], axis=1))
        out.append(np.concatenate(pts).astype(np.int32) if pts
                   else np.zeros((0, 3), dtype=np.int32))
    return out[0], out[1]


def write_one_wave(path, region, psi, iorb=1, eval=0.0, rxyz=None, binary=True):
    """write one orbital (coefficients [psi_c ; psi_f(7, .)]) in the layout
    of writeonewave"""
    psi = np.asarray(psi, dtype=np.float64).ravel()
    nc, nf = region.nvctr_c, region.nvctr_f
    if psi.size != nc + 7 * nf:
        raise ValueError("coefficient count does not match the region")
    rxyz = np.zeros((0, 3)) if rxyz is None else np.asarray(rxyz, dtype=np.float64)
    cc, cf = _points(region)
    pc, pf = psi[:nc], psi[nc:].reshape(nf, 7)
    if binary:
        with open(path, "wb") as f:
            def rec(fmt, *vals):
                body = struct.pack("<" + fmt, *vals)
                f.write(_I4.pack(len(body)) + body + _I4.pack(len(body)))
            rec("id", iorb, eval)
            rec("ddd", *region.hgrid)
            rec("iii", *region.n)
            rec("i", len(rxyz))
            for r in rxyz:
                rec("ddd", *r)
            rec("ii", nc, nf)
            for k in range(nc):
                rec("iiid", *cc[k], pc[k])
            for k in range(nf):
                rec("iii7d", *cf[k], *pf[k])
        return
    with open(path, "w") as f:
        f.write(_list_line(iorb, eval))
        f.write(_list_line(*region.hgrid))
        f.write(_list_line(*region.n))
        f.write(_list_line(len(rxyz)))
        for r in rxyz:
            f.write("".join(" " + fortran_e(v, 24, 17) for v in r) + "\n")
        f.write(_list_line(nc, nf))
        for k in range(nc):
            f.write("".join(f"{int(v):4d}" for v in cc[k]) + " " + fortran_e(pc[k], 19, 12)
                    + "\n")
        for k in range(nf):
            f.write("".join(f"{int(v):4d}" for v in cf[k])
                    + "".join(" " + fortran_e(v, 17, 10) for v in pf[k]) + "\n")


def _fnum(tok):
    return float(tok.replace("D", "E").replace("d", "e"))


def read_one_wave(path, binary=None):
    """read one orbital file; returns a dict with iorb, eval, hgrid, n,
    rxyz, coords_c, coords_f (point coordinates in file order) and psi
    ([psi_c ; psi_f(7, .)])"""
    if binary is None:
        binary = ".bin." in os.path.basename(path)
    if binary:
        data = open(path, "rb").read()
        pos = 0

        def rec(fmt):
            nonlocal pos
            (ln,) = _I4.unpack_from(data, pos)
            body = data[pos + 4:pos + 4 + ln]
            (ln2,) = _I4.unpack_from(data, pos + 4 + ln)
            if ln2 != ln or struct.calcsize("<" + fmt) != ln:
                raise ValueError(f"{path}: bad record at byte {pos}")
            pos += 8 + ln
            return struct.unpack("<" + fmt, body)
        iorb, ev = rec("id")
        hgrid = rec("ddd")
        n = rec("iii")
        (nat,) = rec("i")
        rxyz = np.array([rec("ddd") for _ in range(nat)]).reshape(nat, 3)
        nc, nf = rec("ii")
        # the point records have a fixed size: read them as arrays
        dt_c = np.dtype([("l0", "<i4"), ("i", "<i4", 3), ("v", "<f8"), ("l1", "<i4")])
        dt_f = np.dtype([("l0", "<i4"), ("i", "<i4", 3), ("v", "<f8", 7), ("l1", "<i4")])
        ac = np.frombuffer(data, dtype=dt_c, count=nc, offset=pos)
        pos += nc * dt_c.itemsize
        af = np.frombuffer(data, dtype=dt_f, count=nf, offset=pos)
        pos += nf * dt_f.itemsize
        for a, size in ((ac, 20), (af, 68)):
            if a.size and (np.any(a["l0"] != size) or np.any(a["l1"] != size)):
                raise ValueError(f"{path}: bad point record")
        coords_c, coords_f = ac["i"].astype(np.int32), af["i"].astype(np.int32)
        psi = np.concatenate([ac["v"].astype(np.float64), af["v"].ravel().astype(np.float64)])
    else:
        lines = open(path).read().splitlines()
        it = iter(lines)

        def toks():
            return next(it).split()
        t = toks()
        iorb, ev = int(t[0]), _fnum(t[1])
        hgrid = tuple(_fnum(x) for x in toks()[:3])
        n = tuple(int(x) for x in toks()[:3])
        nat = int(toks()[0])
        rxyz = np.array([[_fnum(x) for x in toks()[:3]] for _ in range(nat)]).reshape(nat, 3)
        t = toks()
        nc, nf = int(t[0]), int(t[1])
        coords_c = np.zeros((nc, 3), dtype=np.int32)
        coords_f = np.zeros((nf, 3), dtype=np.int32)
        psi = np.zeros(nc + 7 * nf)
        for k in range(nc):
            line = next(it)
            coords_c[k] = [int(line[0:4]), int(line[4:8]), int(line[8:12])]
            psi[k] = _fnum(line[12:].split()[0])
        for k in range(nf):
            line = next(it)
            coords_f[k] = [int(line[0:4]), int(line[4:8]), int(line[8:12])]
            psi[nc + 7 * k:nc + 7 * k + 7] = [_fnum(x) for x in line[12:].split()[:7]]
    return dict(iorb=int(iorb), eval=float(ev), hgrid=tuple(hgrid), n=tuple(n), rxyz=rxyz,
                coords_c=coords_c, coords_f=coords_f, psi=psi)


def region_of_wave(wave, geocode):
    """the region whose keys list the file's points in the file's order
    (the keys that wrote it): coarse and fine grids from the coordinates,
    then segkeys.  Raises if the points are not in key order."""
    n = wave["n"]
    shape = tuple(x + 1 for x in n)
    grids = []
    for cc in (wave["coords_c"], wave["coords_f"]):
        if cc.size and (cc.min(axis=0) < 0).any() or cc.size and (cc.max(axis=0) >= shape).any():
            raise ValueError("point outside the box")
        lin = cc[:, 0] + shape[0] * (cc[:, 1] + shape[1] * cc[:, 2]).astype(np.int64)
        if lin.size > 1 and np.any(np.diff(lin) <= 0):
            raise ValueError("points are not in key order")
        g = np.zeros(shape, dtype=bool)
        g[cc[:, 0], cc[:, 1], cc[:, 2]] = True
        grids.append(g)
    return region_from_grids(geocode, n, wave["hgrid"], grids[0], grids[1])


def load_wavefunctions(paths, geocode):
    """(region, psi (norb, ncoef), evals) of orbital files sharing one grid
    and one set of points (the restart case of readmywaves,
    bigdft/src/restart.f90:325-420, with no reformatting)"""
    waves = [read_one_wave(p) for p in paths]
    if not waves:
        raise ValueError("no files")
    w0 = waves[0]
    for w in waves[1:]:
        if (w["n"] != w0["n"] or w["hgrid"] != w0["hgrid"]
                or not np.array_equal(w["coords_c"], w0["coords_c"])
                or not np.array_equal(w["coords_f"], w0["coords_f"])):
            raise ValueError("the files do not share one grid and one set of points "
                             "(reformatting is not supported)")
    region = region_of_wave(w0, geocode)
    psi = np.stack([w["psi"] for w in waves])
    return region, psi, np.array([w["eval"] for w in waves])
