"""ctypes front-end of the CPU oracle (test infrastructure only).

The oracle restates the reference's FP64 Fortran routines in C
(``bdft_oracle.c``).  Only the tests, the smoke entry and the CPU-baseline leg
of the benchmark import this module, as the checker; the product never does.
Arrays are numpy, Fortran (column-major) semantics are kept by passing flat
buffers whose first index varies fastest.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_c_int3 = ctypes.c_int * 3
_c_dbl3 = ctypes.c_double * 3


def build():
    """Compile liboracle.so with gcc (always recompiles)."""
    subprocess.run(["make", "-B", "-C", _HERE], check=True,
                   stdout=subprocess.DEVNULL)


def _load():
    if not os.path.exists(_LIB):
        build()
    lib = ctypes.CDLL(_LIB)
    c_int, c_char = ctypes.c_int, ctypes.c_char
    lib.oref_filters.argtypes = [_dp] * 7
    for name in ("oref_ana_rot_per", "oref_syn_rot_per",
                 "oref_convrot_n_per", "oref_convrot_t_per"):
        getattr(lib, name).argtypes = [c_int, c_int, _dp, _dp]
    lib.oref_convolut_kinetic_per_T.argtypes = [c_int, c_int, c_int, _dp,
                                                _dp, _dp, _dp]
    lib.oref_synthese_per_self.argtypes = [c_int, c_int, c_int, _dp, _dp]
    lib.oref_analyse_per_self.argtypes = [c_int, c_int, c_int, _dp, _dp]
    lib.oref_uncompress_psig.argtypes = [c_int, c_int, c_int, c_int, _ip, _ip,
                                         c_int, _ip, _ip, _dp, _dp, _dp]
    lib.oref_compress_and_accumulate_mixed.argtypes = [
        c_int, c_int, c_int, c_int, _ip, _ip, c_int, _ip, _ip, _dp, _dp, _dp]
    lib.oref_apply_potential_lr.argtypes = [c_int] * 6 + [_c_int3, _dp, _dp,
                                                          _dp]
    lib.oref_comb_grow_all.argtypes = [c_int, c_int, c_int, _c_int3, _c_int3,
                                       _dp, _dp, _dp]
    lib.oref_comb_shrink.argtypes = [c_int, c_int, c_int, _c_int3, _c_int3,
                                     _dp, _dp, _dp]
    lib.oref_convolkinetic_T.argtypes = [c_int, c_int, c_int, _c_int3,
                                         _c_int3, _dp, _dp, _dp, _dp, _dp,
                                         _dp]
    lib.oref_locham_many.argtypes = [
        c_char, _c_int3, _c_int3, _c_int3, _dp, c_int, c_int, _ip, _ip,
        c_int, c_int, _ip, _ip, _dp, c_int, _dp, _dp, _dp, _dp]
    lib.oref_daub_to_isf.argtypes = [
        c_char, _c_int3, _c_int3, _c_int3, c_int, c_int, _ip, _ip, c_int,
        c_int, _ip, _ip, _dp, _dp]
    lib.oref_density.argtypes = [
        c_char, _c_int3, _c_int3, _c_int3, c_int, c_int, _ip, _ip, c_int,
        c_int, _ip, _ip, c_int, _dp, _dp, _ip, _dp]
    for name in ("oref_psi_to_locreg2", "oref_lpsi_to_global2"):
        getattr(lib, name).argtypes = [_c_int3, _c_int3, _c_int3] + [c_int] * 4 + \
            [_ip, _ip] + [c_int] * 3 + [_ip, _ip, _dp, _dp]
    lib.oref_lpsi_to_global2.restype = c_int
    lib.oref_num_segkeys.argtypes = [c_int] * 9 + [_u8, ctypes.POINTER(c_int),
                                                   ctypes.POINTER(c_int)]
    lib.oref_num_segkeys.restype = c_int
    lib.oref_segkeys.argtypes = [c_int] * 9 + [_u8, c_int, _ip, _ip]
    lib.oref_segkeys.restype = c_int
    lib.oref_num_threads.restype = c_int
    lib.oref_precond_scal.argtypes = [_dp, ctypes.c_double, _dp]
    lib.oref_precond_apply.argtypes = [c_char, _c_int3, _dp, ctypes.c_double, c_int,
                                       c_int, _ip, _ip, c_int, c_int, _ip, _ip, _dp, _dp]
    lib.oref_precond_cg.argtypes = [c_char, _c_int3, _dp, ctypes.c_double, c_int, c_int,
                                    c_int, _ip, _ip, c_int, c_int, _ip, _ip, _dp, _dp]
    lib.oref_locham_many_conf.argtypes = lib.oref_locham_many.argtypes + [
        ctypes.c_void_p]
    lib.oref_apply_potential_lr_conf.argtypes = [c_int] * 6 + [
        _c_int3, _dp, _dp, ctypes.c_void_p, _dp]
    _i64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
    lib.oref_bounds_layout.argtypes = [_c_int3, _c_int3, _c_int3, _i64, _ip]
    lib.oref_bounds_layout.restype = c_int
    lib.oref_locreg_bounds.argtypes = [_c_int3, _c_int3, _c_int3, c_int, _ip,
                                       c_int, _ip, _ip]
    lib.oref_locreg_bounds.restype = c_int
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def filters():
    """Filter tables as a dict of numpy arrays (index offsets in the keys)."""
    ch, cg, mag = np.zeros(16), np.zeros(16), np.zeros(16)
    ka, kc, ke = np.zeros(15), np.zeros(29), np.zeros(15)
    fil2 = np.zeros(62)
    lib().oref_filters(ch, cg, mag, ka, kc, ke, fil2)
    return {"ch": ch, "cg": cg, "magic": mag, "kin_a": ka, "kin_c": kc,
            "kin_e": ke, "fil2": fil2.reshape(2, 31)}


def num_threads():
    return lib().oref_num_threads()


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64).ravel(order="F") \
        if a.ndim > 1 else np.ascontiguousarray(a, dtype=np.float64)


# ---- 1D rotating kernels: x is (m, ndat) Fortran, returns (ndat, m) Fortran

def _rot(name, n_param, m, x):
    x = np.asarray(x, dtype=np.float64)
    ndat = x.size // m
    xf = np.ascontiguousarray(x.reshape(m, ndat, order="F").ravel(order="F"))
    y = np.zeros(ndat * m)
    getattr(lib(), name)(n_param, ndat, xf, y)
    return y.reshape(ndat, m, order="F")


def ana_rot_per(n, x):
    return _rot("oref_ana_rot_per", n, 2 * n + 2, x)


def syn_rot_per(n, x):
    return _rot("oref_syn_rot_per", n, 2 * n + 2, x)


def convrot_n_per(n1, x):
    return _rot("oref_convrot_n_per", n1, n1 + 1, x)


def convrot_t_per(n1, x):
    return _rot("oref_convrot_t_per", n1, n1 + 1, x)


def kinetic_per(x, hgrid):
    """y = K x on a periodic fine-scaling box x (Fortran 3D array), h per axis.
    Returns (y, kstrten[3])."""
    n1, n2, n3 = (s - 1 for s in x.shape)
    xf = np.ascontiguousarray(x.ravel(order="F"))
    y = np.zeros_like(xf)
    k = np.zeros(3)
    lib().oref_convolut_kinetic_per_T(n1, n2, n3,
                                      np.asarray(hgrid, dtype=np.float64),
                                      xf, y, k)
    return y.reshape(x.shape, order="F"), k


class Keys:
    """Compressed-form descriptor (wavefunctions_descriptors,
    locregs/compression.f90:29-39): int32 1-based keyg (2, nseg) and keyv."""

    def __init__(self, nseg_c, nvctr_c, keyg_c, keyv_c, nseg_f, nvctr_f,
                 keyg_f, keyv_f):
        self.nseg_c, self.nvctr_c = int(nseg_c), int(nvctr_c)
        self.nseg_f, self.nvctr_f = int(nseg_f), int(nvctr_f)
        self.keyg_c = np.ascontiguousarray(keyg_c, dtype=np.int32).ravel()
        self.keyv_c = np.ascontiguousarray(keyv_c, dtype=np.int32).ravel()
        self.keyg_f = np.ascontiguousarray(keyg_f, dtype=np.int32).ravel()
        self.keyv_f = np.ascontiguousarray(keyv_f, dtype=np.int32).ravel()
        if self.keyg_f.size == 0:
            self.keyg_f = np.zeros(2, dtype=np.int32)
            self.keyv_f = np.zeros(1, dtype=np.int32)

    @property
    def ncoef(self):
        return self.nvctr_c + 7 * self.nvctr_f


def segkeys(logrid, nl=None, nu=None):
    """num_segkeys + segkeys on a boolean grid (Fortran order, shape
    (n1+1, n2+1, n3+1)).  Returns (nseg, nvctr, keyg(2*nseg), keyv)."""
    lg = np.ascontiguousarray(np.asarray(logrid, dtype=np.uint8).ravel(order="F"))
    n1, n2, n3 = (s - 1 for s in logrid.shape)
    nl = nl or (0, 0, 0)
    nu = nu or (n1, n2, n3)
    mseg, mvctr = ctypes.c_int(0), ctypes.c_int(0)
    rc = lib().oref_num_segkeys(n1, n2, n3, nl[0], nu[0], nl[1], nu[1], nl[2],
                                nu[2], lg, ctypes.byref(mseg),
                                ctypes.byref(mvctr))
    assert rc == 0
    keyg = np.zeros(max(1, 2 * mseg.value), dtype=np.int32)
    keyv = np.zeros(max(1, mseg.value), dtype=np.int32)
    rc = lib().oref_segkeys(n1, n2, n3, nl[0], nu[0], nl[1], nu[1], nl[2],
                            nu[2], lg, mseg.value, keyg, keyv)
    assert rc == 0
    return mseg.value, mvctr.value, keyg[:2 * mseg.value], keyv[:mseg.value]


def uncompress_psig(n, keys, psi):
    n1, n2, n3 = n
    psig = np.zeros(8 * (n1 + 1) * (n2 + 1) * (n3 + 1))
    psi = np.ascontiguousarray(psi, dtype=np.float64)
    psi_f = psi[keys.nvctr_c:] if keys.nvctr_f > 0 else psi
    lib().oref_uncompress_psig(n1, n2, n3, keys.nseg_c, keys.keyg_c,
                               keys.keyv_c, keys.nseg_f, keys.keyg_f,
                               keys.keyv_f, psi[:max(1, keys.nvctr_c)].copy(),
                               np.ascontiguousarray(psi_f), psig)
    return psig


def compress_acc_mixed(n, keys, psig, hpsi):
    """hpsi += compress(psig); returns the updated copy."""
    n1, n2, n3 = n
    h = np.array(hpsi, dtype=np.float64, copy=True)
    hc = np.ascontiguousarray(h[:keys.nvctr_c]) if keys.nvctr_c else np.zeros(1)
    hf = np.ascontiguousarray(h[keys.nvctr_c:]) if keys.nvctr_f else np.zeros(7)
    lib().oref_compress_and_accumulate_mixed(
        n1, n2, n3, keys.nseg_c, keys.keyg_c, keys.keyv_c, keys.nseg_f,
        keys.keyg_f, keys.keyv_f, np.ascontiguousarray(psig, dtype=np.float64),
        hc, hf)
    if keys.nvctr_c:
        h[:keys.nvctr_c] = hc
    if keys.nvctr_f:
        h[keys.nvctr_c:] = hf
    return h


def comb_grow_all(n, nfl, nfu, x_c, x_f):
    n1, n2, n3 = n
    psir = np.zeros((2 * n1 + 31) * (2 * n2 + 31) * (2 * n3 + 31))
    lib().oref_comb_grow_all(n1, n2, n3, _c_int3(*nfl), _c_int3(*nfu),
                             _f(np.asarray(x_c)), _f(np.asarray(x_f)), psir)
    return psir


def comb_shrink(n, nfl, nfu, psir):
    n1, n2, n3 = n
    nc = (n1 + 1) * (n2 + 1) * (n3 + 1)
    nf = 7 * np.prod([nfu[i] - nfl[i] + 1 for i in range(3)])
    y_c, y_f = np.zeros(nc), np.zeros(max(1, int(nf)))
    lib().oref_comb_shrink(n1, n2, n3, _c_int3(*nfl), _c_int3(*nfu),
                           _f(np.asarray(psir)), y_c, y_f)
    return y_c, y_f


def convolkinetic_T(n, nfl, nfu, hgrid, x_c, x_f):
    n1, n2, n3 = n
    y_c = np.zeros_like(_f(np.asarray(x_c)))
    y_f = np.zeros_like(_f(np.asarray(x_f)))
    ek = np.zeros(1)
    lib().oref_convolkinetic_T(n1, n2, n3, _c_int3(*nfl), _c_int3(*nfu),
                               np.asarray(hgrid, dtype=np.float64),
                               _f(np.asarray(x_c)), _f(np.asarray(x_f)), y_c,
                               y_f, ek)
    return y_c, y_f, ek[0]


def apply_potential(psir, pot, dims, potdims=None, ishift=(0, 0, 0)):
    potdims = potdims or dims
    p = np.array(psir, dtype=np.float64, copy=True).ravel()
    ep = np.zeros(1)
    lib().oref_apply_potential_lr(dims[0], dims[1], dims[2], potdims[0],
                                  potdims[1], potdims[2], _c_int3(*ishift), p,
                                  np.ascontiguousarray(pot, dtype=np.float64).ravel(),
                                  ep)
    return p, ep[0]


# confpot_data (locregs/locreg_operations.f90:23-30), C layout of oref_confpot
CONFPOT_DTYPE = np.dtype([("potorder", np.int32), ("ioffset", np.int32, 3),
                          ("prefac", np.float64), ("hh", np.float64, 3),
                          ("rxyzConf", np.float64, 3), ("damping", np.float64)],
                         align=True)


def confpot(potorder, prefac, hh, center, ioffset, damping=1.0):
    c = np.zeros(1, dtype=CONFPOT_DTYPE)
    c["potorder"], c["prefac"], c["damping"] = potorder, prefac, damping
    c["hh"], c["rxyzConf"], c["ioffset"] = hh, center, ioffset
    return c


def apply_potential_conf(psir, pot, dims, conf, potdims=None, ishift=(0, 0, 0)):
    """apply_potential_lr_conf_nobounds on one psir (returns psir', epot)"""
    potdims = potdims or dims
    p = np.array(psir, dtype=np.float64, copy=True).ravel()
    ep = np.zeros(1)
    c = np.ascontiguousarray(conf, dtype=CONFPOT_DTYPE)
    lib().oref_apply_potential_lr_conf(dims[0], dims[1], dims[2], potdims[0],
                                       potdims[1], potdims[2], _c_int3(*ishift), p,
                                       np.ascontiguousarray(pot, dtype=np.float64).ravel(),
                                       c.ctypes.data, ep)
    return p, ep[0]


def locham(geocode, n, nfl, nfu, hgrid, keys, pot, psi, hpsi=None, conf=None):
    """Local Hamiltonian of norb orbitals sharing one locreg.

    psi: (norb, ncoef) C-contiguous; returns (hpsi, ekin[norb], epot[norb]).
    hpsi is accumulated onto the given array (zeros if None).  conf: None or
    norb CONFPOT_DTYPE records (the orbitals' confinement)."""
    psi = np.ascontiguousarray(psi, dtype=np.float64)
    if psi.ndim == 1:
        psi = psi[None, :]
    norb = psi.shape[0]
    h = np.zeros_like(psi) if hpsi is None else \
        np.array(hpsi, dtype=np.float64, copy=True).reshape(psi.shape)
    ek, ep = np.zeros(norb), np.zeros(norb)
    fn, extra = lib().oref_locham_many, []
    if conf is not None:
        conf = np.ascontiguousarray(conf, dtype=CONFPOT_DTYPE).ravel()
        assert conf.size == norb
        fn, extra = lib().oref_locham_many_conf, [conf.ctypes.data]
    fn(
        geocode.encode() if isinstance(geocode, str) else geocode,
        _c_int3(*n), _c_int3(*nfl), _c_int3(*nfu),
        np.asarray(hgrid, dtype=np.float64), keys.nseg_c, keys.nvctr_c,
        keys.keyg_c, keys.keyv_c, keys.nseg_f, keys.nvctr_f, keys.keyg_f,
        keys.keyv_f, np.ascontiguousarray(pot, dtype=np.float64).ravel(),
        norb, psi, h, ek, ep, *extra)
    return h, ek, ep


def _nfine(geocode, n):
    add = 2 if geocode == "P" else 31
    return tuple(2 * x + add for x in n)


def daub_to_isf(geocode, n, nfl, nfu, keys, psi):
    """psir on the fine ISF grid (x fastest) of one orbital."""
    psi = np.ascontiguousarray(psi, dtype=np.float64).ravel()
    nf = _nfine(geocode, n)
    out = np.zeros(nf[0] * nf[1] * nf[2])
    lib().oref_daub_to_isf(
        geocode.encode(), _c_int3(*n), _c_int3(*nfl), _c_int3(*nfu),
        keys.nseg_c, keys.nvctr_c, keys.keyg_c, keys.keyv_c, keys.nseg_f,
        keys.nvctr_f, keys.keyg_f, keys.keyv_f, psi, out)
    return out


def density(geocode, n, nfl, nfu, keys, psi, hfac, spin=None, nspin=1,
            rho=None):
    """rho(:, spin) += hfac * psir^2 over the orbitals (rows of psi);
    returns rho of shape (nspin, n1i*n2i*n3i)."""
    psi = np.ascontiguousarray(psi, dtype=np.float64)
    if psi.ndim == 1:
        psi = psi[None, :]
    norb = psi.shape[0]
    nf = _nfine(geocode, n)
    tot = nf[0] * nf[1] * nf[2]
    r = np.zeros((nspin, tot)) if rho is None else \
        np.array(rho, dtype=np.float64, copy=True).reshape(nspin, tot)
    sp = np.zeros(norb, dtype=np.int32) if spin is None else \
        np.ascontiguousarray(spin, dtype=np.int32)
    lib().oref_density(
        geocode.encode(), _c_int3(*n), _c_int3(*nfl), _c_int3(*nfu),
        keys.nseg_c, keys.nvctr_c, keys.keyg_c, keys.keyv_c, keys.nseg_f,
        keys.nvctr_f, keys.keyg_f, keys.keyv_f, norb, psi,
        np.ascontiguousarray(hfac, dtype=np.float64), sp, r)
    return r


def _cat_keys(keys):
    kg = np.concatenate([keys.keyg_c, keys.keyg_f]).astype(np.int32)
    kv = np.concatenate([keys.keyv_c, keys.keyv_f]).astype(np.int32)
    return np.ascontiguousarray(kg), np.ascontiguousarray(kv)


def psi_to_locreg2(nl, ng, ns, lkeys, gkeys, gpsi):
    """global compressed orbital -> local compressed orbital (free BC)."""
    kgl, kvl = _cat_keys(lkeys)
    kgg, kvg = _cat_keys(gkeys)
    lpsi = np.zeros(lkeys.nvctr_c + 7 * lkeys.nvctr_f)
    lib().oref_psi_to_locreg2(
        _c_int3(*nl), _c_int3(*ng), _c_int3(*ns), lkeys.nseg_c, lkeys.nseg_f,
        lkeys.nvctr_c, lkeys.nvctr_f, kgl, kvl, gkeys.nseg_c, gkeys.nseg_f,
        gkeys.nvctr_c, kgg, kvg, np.ascontiguousarray(gpsi, dtype=np.float64), lpsi)
    return lpsi


def lpsi_to_global2(nl, ng, ns, lkeys, gkeys, lpsi, psi):
    """local compressed orbital written into a copy of the global one;
    returns (psi, number of local points placed)."""
    kgl, kvl = _cat_keys(lkeys)
    kgg, kvg = _cat_keys(gkeys)
    out = np.array(psi, dtype=np.float64, copy=True)
    n = lib().oref_lpsi_to_global2(
        _c_int3(*nl), _c_int3(*ng), _c_int3(*ns), lkeys.nseg_c, lkeys.nseg_f,
        lkeys.nvctr_c, lkeys.nvctr_f, kgl, kvl, gkeys.nseg_c, gkeys.nseg_f,
        gkeys.nvctr_c, kgg, kvg, np.ascontiguousarray(lpsi, dtype=np.float64), out)
    return out, n


BOUNDS_NAMES = ("ibyz_c", "ibxz_c", "ibxy_c", "ibyz_f", "ibxz_f", "ibxy_f",
                "ibzxx_c", "ibxxyy_c", "ibyz_ff", "ibzxx_f", "ibxxyy_f",
                "ibzzx_c", "ibyyzz_c", "ibxy_ff", "ibzzx_f", "ibyyzz_f",
                "ibyyzz_r")


def bounds_layout(n, nfl, nfu):
    """(off[18], dims[17, 4]) of the convolution-bounds buffer: array k
    starts at off[k] and is ib(2, lo1:hi1, lo2:hi2) with dims[k] =
    (lo1, hi1, lo2, hi2)."""
    off = np.zeros(18, dtype=np.int64)
    dims = np.zeros(68, dtype=np.int32)
    if lib().oref_bounds_layout(_c_int3(*n), _c_int3(*nfl), _c_int3(*nfu), off, dims):
        raise ValueError("nfl > nfu")
    return off, dims.reshape(17, 4)


def unpack_bounds(buf, n, nfl, nfu):
    """dict name -> (ib, lo1, lo2); ib[k, i - lo1, j - lo2] is the
    reference's ib(k+1, i, j)"""
    off, dims = bounds_layout(n, nfl, nfu)
    out = {}
    for k, name in enumerate(BOUNDS_NAMES):
        lo1, hi1, lo2, hi2 = (int(x) for x in dims[k])
        a = np.asarray(buf[off[k]:off[k + 1]])
        a = a.reshape(hi2 - lo2 + 1, hi1 - lo1 + 1, 2).transpose(2, 1, 0)
        out[name] = (a, lo1, lo2)
    return out


def locreg_bounds(n, nfl, nfu, keys):
    """convolution bounds of a free-BC region (flat int32 buffer)"""
    off, _ = bounds_layout(n, nfl, nfu)
    buf = np.zeros(int(off[17]), dtype=np.int32)
    kc = np.ascontiguousarray(keys.keyg_c, dtype=np.int32)
    kf = np.ascontiguousarray(keys.keyg_f, dtype=np.int32)
    if lib().oref_locreg_bounds(_c_int3(*n), _c_int3(*nfl), _c_int3(*nfu),
                                keys.nseg_c, kc, keys.nseg_f, kf, buf):
        raise ValueError("nfl > nfu")
    return buf


def precond_scal(hgrid, c):
    """the 8 diagonal scalings (type t = s1 + 2 s2 + 4 s3)"""
    out = np.zeros(8)
    lib().oref_precond_scal(np.asarray(hgrid, dtype=np.float64), float(c), out)
    return out


def precond_apply(geocode, n, hgrid, c, keys, x):
    """y = S Ana (K + c) Syn S x (apply_hp of the OpenCL preconditioner)"""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    lib().oref_precond_apply(geocode.encode(), _c_int3(*n), np.asarray(hgrid, dtype=np.float64),
                             float(c), keys.nseg_c, keys.nvctr_c, keys.keyg_c, keys.keyv_c,
                             keys.nseg_f, keys.nvctr_f, keys.keyg_f, keys.keyv_f, x, y)
    return y


def precond_cg(geocode, n, hgrid, c, ncong, keys, x, b):
    """ocl_preconditioner_generic_k: returns the scaled CG solution"""
    x = np.array(x, dtype=np.float64, copy=True)
    b = np.array(b, dtype=np.float64, copy=True)
    lib().oref_precond_cg(geocode.encode(), _c_int3(*n), np.asarray(hgrid, dtype=np.float64),
                          float(c), int(ncong), keys.nseg_c, keys.nvctr_c, keys.keyg_c,
                          keys.keyv_c, keys.nseg_f, keys.nvctr_f, keys.keyg_f, keys.keyv_f,
                          x, b)
    return x
