"""Host side of the MI355X local-Hamiltonian library (ctypes over the C ABI).

Mirrors the accelerator interface the reference drives from
LocalHamiltonianApplication (bigdft/src/hpsiortho.f90:892-903):
  init_acceleration_OCL  -> Context(device)
  allocate_data_OCL      -> Context.region(Region)   (keys validated+uploaded)
  local_hamiltonian_OCL  -> local_hamiltonian(...)    (orbital batch on device)
  finish_hamiltonian_OCL -> the weighted energy sums returned by the above
Device memory is handled with torch tensors (plumbing only); every compute
call goes through libbigdft_hip.so, and importing this package fails loudly
when that library is missing.
"""
import ctypes
import os

import numpy as np

from . import descriptors  # noqa: F401
from .descriptors import Region, dense_region, sphere_region  # noqa: F401

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "lib", "libbigdft_hip.so")

PREC_FP64, PREC_FP32, PREC_BF16X3, PREC_BF16X4, PREC_F16X3 = 0, 1, 2, 3, 4
# launch options (BDFT_OPT_* of bigdft_hip.h)
OPTIONS = {"line_passes": 1, "zgroup": 2, "zpot": 3, "bwd": 4, "chunk_mb": 5,
           "tile_order": 6, "persist": 7}
# bytes of the fixed device array of bdft_locham_batch_fixed_d (BDFT_FIXED_BYTES)
FIXED_BYTES = 128 << 20
PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16x3": PREC_BF16X3,
              "bf16x4": PREC_BF16X4, "f16x3": PREC_F16X3}

ERRORS = {1: "BDFT_ERR_ARG", 2: "BDFT_ERR_HIP", 3: "BDFT_ERR_DESC",
          4: "BDFT_ERR_NOMEM", 5: "BDFT_ERR_UNSUPPORTED"}

# entry points declared in include/bigdft_hip.h
EXPORTS = ("bdft_ctx_create", "bdft_ctx_destroy", "bdft_ctx_set_stream",
           "bdft_last_error", "bdft_lr_create", "bdft_lr_destroy",
           "bdft_lr_info", "bdft_locham_d", "bdft_locham_batch_d",
           "bdft_plan_create", "bdft_plan_execute", "bdft_plan_destroy",
           "bdft_local_hamiltonian", "bdft_uncompress_d",
           "bdft_compress_acc_d", "bdft_profile_enable", "bdft_profile_read",
           "bdft_stage_name", "bdft_ctx_info", "bdft_sync", "bdft_density_d",
           "bdft_density", "bdft_psi_to_locreg_d", "bdft_locreg_to_psi_d",
           "bdft_bounds_layout", "bdft_locreg_bounds_d", "bdft_locreg_bounds",
           "bdft_locham_conf_batch_d", "bdft_precond_cg_d", "bdft_precondition",
           "bdft_ctx_set_option", "bdft_ctx_get_option",
           "bdft_locham_batch_fixed_d")


class BigDFTError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def load_library(path=LIB_PATH):
    """Load libbigdft_hip.so and declare the C signatures.  Raises if the
    library has not been built (there is no fallback path)."""
    if not os.path.exists(path):
        raise ImportError(f"{path} not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    I3 = ctypes.c_int * 3
    D3 = ctypes.c_double * 3
    sig = {
        "bdft_ctx_create": [I, ctypes.POINTER(P)],
        "bdft_ctx_destroy": [P],
        "bdft_ctx_set_stream": [P, P],
        "bdft_lr_create": [P, ctypes.c_char, I3, I3, I3, D3, I, I, P, P, I, I,
                           P, P, ctypes.POINTER(P)],
        "bdft_lr_destroy": [P],
        "bdft_lr_info": [P, I3, ctypes.POINTER(ctypes.c_int64)],
        "bdft_locham_d": [P, P, I, I, P, P, P, P, P, P, I],
        "bdft_locham_batch_d": [P, I, P, P, P, P, P, P, P, P, P, P, I],
        "bdft_locham_batch_fixed_d": [P, I, P, P, P, P, P, P, P, P, P, P, I],
        "bdft_locham_conf_batch_d": [P, I, P, P, P, P, P, P, P, P, P, P, P, I],
        "bdft_density_d": [P, P, I, P, P, P, I, P, I],
        "bdft_density": [P, P, I, P, P, P, I, P, I],
        "bdft_psi_to_locreg_d": [P, P, I, P, P, P, P, P, P],
        "bdft_locreg_to_psi_d": [P, P, I, P, P, P, P, P, P],
        "bdft_plan_create": [P, I, P, P, P, P, P, P, P, P, P, P, I,
                             ctypes.POINTER(P)],
        "bdft_plan_execute": [P],
        "bdft_plan_destroy": [P],
        "bdft_local_hamiltonian": [P, P, I, I, P, P, P, P, P, P, I],
        "bdft_uncompress_d": [P, P, I, P, P],
        "bdft_compress_acc_d": [P, P, I, P, P],
        "bdft_sync": [P],
        "bdft_profile_enable": [P, I],
        "bdft_profile_read": [P, I, P, P, I],
        "bdft_ctx_info": [P, P, P],
        "bdft_ctx_set_option": [P, I, I],
        "bdft_ctx_get_option": [P, I, P],
        "bdft_bounds_layout": [I3, I3, I3, P, P],
        "bdft_precond_cg_d": [P, I, P, P, P, I, P, P],
        "bdft_precondition": [P, P, I, P, I, P, P],
        "bdft_locreg_bounds_d": [P, P, P],
        "bdft_locreg_bounds": [P, P, P],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = I
    lib.bdft_last_error.argtypes = [P]
    lib.bdft_last_error.restype = ctypes.c_char_p
    lib.bdft_stage_name.argtypes = [I]
    lib.bdft_stage_name.restype = ctypes.c_char_p
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = load_library()
    return _lib


def _ptr(t):
    """Device (or host) address of a torch tensor / numpy array / None."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return ctypes.c_void_p(t.data_ptr())
    if isinstance(t, np.ndarray):
        return t.ctypes.data_as(ctypes.c_void_p)
    return ctypes.c_void_p(int(t))


def _dev(t, name, numel=0, dtype="float64"):
    """A device array argument: a contiguous CUDA torch tensor of the given
    dtype holding at least numel elements (ValueError otherwise), so that a
    CPU, mistyped, strided or short tensor never reaches a kernel."""
    import torch
    want = getattr(torch, dtype)
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name}: expected a torch tensor on the GPU, got {type(t).__name__}")
    if not t.is_cuda:
        raise ValueError(f"{name}: tensor is on {t.device}, expected a CUDA device")
    if t.dtype != want:
        raise ValueError(f"{name}: dtype {t.dtype}, expected {want}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: tensor is not contiguous")
    if t.numel() < numel:
        raise ValueError(f"{name}: {t.numel()} elements, at least {int(numel)} needed")
    return ctypes.c_void_p(t.data_ptr())


def _ncoef(lr):
    r = lr.region
    return r.nvctr_c + 7 * r.nvctr_f


def _batch_sizes(lrs, pot_off, potdim, psi_off):
    """elements the psi/hpsi and potential arrays of a batch must hold"""
    nc = np.array([_ncoef(l) for l in lrs], dtype=np.int64)
    # no psi_off: orbitals one after the other (as the library places them)
    so = np.cumsum(nc) - nc if psi_off is None else np.asarray(psi_off, dtype=np.int64)
    npsi = int((so + nc).max()) if len(lrs) else 0
    if potdim is None:
        pd = np.array([np.prod(l.region.nfine) for l in lrs], dtype=np.int64)
    else:
        pd = np.asarray(potdim, dtype=np.int64).reshape(-1, 3).prod(axis=1)
    po = np.zeros(len(lrs), dtype=np.int64) if pot_off is None else \
        np.asarray(pot_off, dtype=np.int64)
    npot = int((po + pd).max()) if len(lrs) else 0
    return npsi, npot


class LocReg:
    """A region registered on a context (bdft_lr)."""

    def __init__(self, ctx, region):
        self.ctx, self.region = ctx, region
        h = ctypes.c_void_p()
        I3, D3 = ctypes.c_int * 3, ctypes.c_double * 3
        kg_c = np.ascontiguousarray(region.keyg_c, dtype=np.int32)
        kv_c = np.ascontiguousarray(region.keyv_c, dtype=np.int32)
        kg_f = np.ascontiguousarray(region.keyg_f, dtype=np.int32)
        kv_f = np.ascontiguousarray(region.keyv_f, dtype=np.int32)
        ctx._check(lib().bdft_lr_create(
            ctx.handle, region.geocode.encode(), I3(*region.n),
            I3(*region.nfl), I3(*region.nfu), D3(*region.hgrid),
            region.nseg_c, region.nvctr_c, _ptr(kg_c), _ptr(kv_c),
            region.nseg_f, region.nvctr_f, _ptr(kg_f), _ptr(kv_f),
            ctypes.byref(h)))
        self.handle = h

    def info(self):
        nf = (ctypes.c_int * 3)()
        nc = ctypes.c_int64()
        lib().bdft_lr_info(self.handle, nf, ctypes.byref(nc))
        return tuple(nf), nc.value

    def close(self):
        if self.handle:
            lib().bdft_lr_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _stream_ordered(method):
    """Device-array entry points run on the context's stream: it first waits
    for the work already queued on torch's current stream (tensors the
    caller just allocated or filled), and torch's current stream then waits
    for the call, so results read through torch see it complete."""
    def wrapped(self, *args, **kwargs):
        ctx = self if isinstance(self, Context) else self.ctx
        ctx._enter()
        try:
            return method(self, *args, **kwargs)
        finally:
            ctx._leave()
    wrapped.__name__, wrapped.__doc__ = method.__name__, method.__doc__
    return wrapped


class Plan:
    """A captured batch (bdft_plan); execute() replays its hipGraph."""

    def __init__(self, ctx, handle, keep):
        self.ctx, self.handle, self._keep = ctx, handle, keep

    @_stream_ordered
    def execute(self):
        self.ctx._check(lib().bdft_plan_execute(self.handle))

    def close(self):
        if self.handle:
            lib().bdft_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """bdft_ctx on one device (init_acceleration_OCL)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        rc = lib().bdft_ctx_create(device, ctypes.byref(h))
        if rc:
            raise BigDFTError(rc, f"cannot create a context on device {device}")
        self.handle = h
        self.device = device
        # the torch stream the context runs on (default priority), so that
        # its work can be ordered against torch's
        import torch
        self._stream = torch.cuda.Stream(device=device)
        self._check(lib().bdft_ctx_set_stream(self.handle,
                                              ctypes.c_void_p(self._stream.cuda_stream)))

    def _events(self):
        import torch
        if getattr(self, "_ev", None) is None:   # two events, reused by every call
            with torch.cuda.device(self.device):
                self._ev = (torch.cuda.Event(), torch.cuda.Event())
        return torch, self._ev

    def _enter(self):
        torch, ev = self._events()
        cur = torch.cuda.current_stream(self.device)
        if cur != self._stream:
            ev[0].record(cur)
            self._stream.wait_event(ev[0])

    def _leave(self):
        torch, ev = self._events()
        cur = torch.cuda.current_stream(self.device)
        if cur != self._stream:
            ev[1].record(self._stream)
            cur.wait_event(ev[1])

    def _check(self, rc):
        if rc:
            msg = lib().bdft_last_error(self.handle)
            raise BigDFTError(rc, msg.decode() if msg else "")

    def region(self, region):
        return LocReg(self, region)

    def set_stream(self, stream):
        """Run on a torch stream (None: a new stream of the context's own)."""
        import torch
        if stream is None:
            stream = torch.cuda.Stream(device=self.device)
        self._check(lib().bdft_ctx_set_stream(self.handle,
                                              ctypes.c_void_p(stream.cuda_stream)))
        self._stream = stream

    def sync(self):
        self._check(lib().bdft_sync(self.handle))

    def info(self):
        """(LDS bytes per work-group, plane sequence enabled)"""
        lds, pl = ctypes.c_int64(), ctypes.c_int()
        self._check(lib().bdft_ctx_info(self.handle, ctypes.byref(lds),
                                        ctypes.byref(pl)))
        return lds.value, bool(pl.value)

    def set_option(self, name, value):
        """launch option (OPTIONS: "line_passes" 0/1, "zgroup" 0/2/3, "zpot"
        0/1/2, "bwd" 0/1/2, "chunk_mb" MiB, "tile_order" 0/1, "persist" 0/1); cached plans of the context are dropped when a
        value changes"""
        self._check(lib().bdft_ctx_set_option(self.handle, OPTIONS[name], int(value)))

    def get_option(self, name):
        v = ctypes.c_int()
        self._check(lib().bdft_ctx_get_option(self.handle, OPTIONS[name], ctypes.byref(v)))
        return v.value

    def profile(self, on=True):
        self._check(lib().bdft_profile_enable(self.handle, int(on)))

    def profile_read(self, reset=True):
        """{stage name: accumulated ms}, number of profiled launches"""
        n = 16
        ms = np.zeros(n)
        cnt = ctypes.c_int64()
        self._check(lib().bdft_profile_read(self.handle, n, _ptr(ms),
                                            ctypes.byref(cnt), int(reset)))
        out = {}
        for i in range(n):
            name = lib().bdft_stage_name(i).decode()
            if name:
                out[name] = float(ms[i])
        return out, cnt.value

    # ---- device-pointer entry points (torch tensors) ----
    @_stream_ordered
    def locham(self, lr, pot, psi, hpsi, ekin, epot, precision=PREC_FP64,
               nspin=1, spinsgn=None):
        """bdft_locham_d: norb = ekin.numel() consecutive orbitals of one
        region; pot holds nspin potentials (pot(n1i,n2i,n3i,nspin)), orbital
        i reads the spin-down one where spinsgn[i] <= 0 (host sequence)"""
        norb = ekin.numel()
        nc = _ncoef(lr) * norb
        npot = int(np.prod(lr.region.nfine)) * nspin
        sg = None if spinsgn is None else np.ascontiguousarray(spinsgn, dtype=np.float64)
        if sg is not None and sg.size != norb:
            raise ValueError("one spinsgn per orbital")
        self._check(lib().bdft_locham_d(self.handle, lr.handle, norb, int(nspin),
                                        _dev(pot, "pot", npot), _ptr(sg),
                                        _dev(psi, "psi", nc), _dev(hpsi, "hpsi", nc),
                                        _dev(ekin, "ekin", norb), _dev(epot, "epot", norb),
                                        precision))

    @staticmethod
    def _batch_args(lrs, pot_off, potdim, ishift, psi_off):
        arr = (ctypes.c_void_p * len(lrs))(*[l.handle.value for l in lrs])

        def a(x, dt):
            return None if x is None else np.ascontiguousarray(x, dtype=dt)
        return (arr, a(pot_off, np.int64), a(potdim, np.int32),
                a(ishift, np.int32), a(psi_off, np.int64))

    @_stream_ordered
    def locham_batch(self, lrs, pot, psi, hpsi, ekin, epot, pot_off=None,
                     potdim=None, ishift=None, psi_off=None,
                     precision=PREC_FP64, conf=None):
        """conf: None or one CONFPOT_DTYPE record per orbital (host)"""
        arr, po, pd, sh, so = self._batch_args(lrs, pot_off, potdim, ishift,
                                               psi_off)
        npsi, npot = _batch_sizes(lrs, po, pd, so)
        n = len(lrs)
        dp = (_dev(pot, "pot", npot), _dev(psi, "psi", npsi), _dev(hpsi, "hpsi", npsi),
              _dev(ekin, "ekin", n), _dev(epot, "epot", n))
        if conf is None:
            self._check(lib().bdft_locham_batch_d(
                self.handle, n, arr, dp[0], _ptr(po), _ptr(pd),
                _ptr(sh), dp[1], _ptr(so), dp[2], dp[3], dp[4], precision))
            return
        c = np.ascontiguousarray(conf, dtype=CONFPOT_DTYPE).ravel()
        if c.size != n:
            raise ValueError("one confinement record per orbital")
        self._check(lib().bdft_locham_conf_batch_d(
            self.handle, n, arr, dp[0], _ptr(po), _ptr(pd),
            _ptr(sh), dp[1], _ptr(so), dp[2], dp[3], dp[4], _ptr(c), precision))

    @_stream_ordered
    def locham_batch_fixed(self, lrs, pot, psi, hpsi, ekin, epot, pot_off=None,
                           potdim=None, ishift=None, psi_off=None,
                           precision=PREC_FP64):
        """bdft_locham_batch_fixed_d: locham_batch without device
        allocation (the library's fixed device array, FIXED_BYTES per
        device); returns when the batch is complete"""
        arr, po, pd, sh, so = self._batch_args(lrs, pot_off, potdim, ishift,
                                               psi_off)
        npsi, npot = _batch_sizes(lrs, po, pd, so)
        n = len(lrs)
        self._check(lib().bdft_locham_batch_fixed_d(
            self.handle, n, arr, _dev(pot, "pot", npot), _ptr(po), _ptr(pd), _ptr(sh),
            _dev(psi, "psi", npsi), _ptr(so), _dev(hpsi, "hpsi", npsi),
            _dev(ekin, "ekin", n), _dev(epot, "epot", n), precision))

    @_stream_ordered
    def density(self, lr, psi, hfac, rho, spin=None, nspin=1, precision=PREC_FP64):
        """rho[s] += hfac[i] psir_i^2 over the orbitals (rows of the device
        array psi) of spin s; hfac, spin: host sequences; rho: device,
        nspin x n1i*n2i*n3i, accumulated.  Synchronous."""
        norb = len(hfac)
        hf = np.ascontiguousarray(hfac, dtype=np.float64)
        sp = None if spin is None else np.ascontiguousarray(spin, dtype=np.int32)
        self._check(lib().bdft_density_d(
            self.handle, lr.handle, norb, _dev(psi, "psi", _ncoef(lr) * norb),
            hf.ctypes.data_as(ctypes.c_void_p),
            None if sp is None else sp.ctypes.data_as(ctypes.c_void_p),
            nspin, _dev(rho, "rho", int(np.prod(lr.region.nfine)) * nspin), precision))

    @_stream_ordered
    def _xfer(self, fn, glr, llrs, ns, src, soff, dst, doff):
        arr = (ctypes.c_void_p * len(llrs))(*[l.handle.value for l in llrs])
        nsa = np.ascontiguousarray(ns, dtype=np.int32).ravel()
        so = None if soff is None else np.ascontiguousarray(soff, dtype=np.int64)
        do = None if doff is None else np.ascontiguousarray(doff, dtype=np.int64)
        self._check(fn(self.handle, glr.handle, len(llrs), arr, _ptr(nsa), _ptr(src),
                       _ptr(so), _ptr(dst), _ptr(do)))

    def psi_to_locreg(self, glr, llrs, ns, gpsi, lpsi, goff=None, loff=None):
        """psi_to_locreg2: global orbitals -> local regions (device arrays)"""
        self._xfer(lib().bdft_psi_to_locreg_d, glr, llrs, ns, gpsi, goff, lpsi, loff)

    def locreg_to_psi(self, glr, llrs, ns, lpsi, gpsi, loff=None, goff=None):
        """Lpsi_to_global2: local regions -> their points of global orbitals"""
        self._xfer(lib().bdft_locreg_to_psi_d, glr, llrs, ns, lpsi, loff, gpsi, goff)

    @_stream_ordered
    def precond_cg(self, lrs, cprecr, ncong, x, b, psi_off=None):
        """ocl_preconditioner_generic_k over a batch: x (device, in/out) <-
        S y after ncong CG steps on S (T + c) S y = b from the guess x"""
        arr = (ctypes.c_void_p * len(lrs))(*[l.handle.value for l in lrs])
        c = np.ascontiguousarray(cprecr, dtype=np.float64)
        if c.size != len(lrs):
            raise ValueError("one cprecr per orbital")
        so = None if psi_off is None else np.ascontiguousarray(psi_off, dtype=np.int64)
        npsi, _ = _batch_sizes(lrs, None, None, so)
        self._check(lib().bdft_precond_cg_d(self.handle, len(lrs), arr, _ptr(so), _ptr(c),
                                            int(ncong), _dev(x, "x", npsi),
                                            _dev(b, "b", npsi)))

    def precondition_host(self, lr, cprecr, ncong, x, b):
        """host arrays (norb, ncoef); x is overwritten"""
        assert x.flags["C_CONTIGUOUS"] and x.dtype == np.float64
        bb = np.ascontiguousarray(b, dtype=np.float64)
        c = np.ascontiguousarray(cprecr, dtype=np.float64)
        self._check(lib().bdft_precondition(self.handle, lr.handle, c.size, _ptr(c),
                                            int(ncong), _ptr(x), _ptr(bb)))

    def locreg_bounds(self, lr, out=None):
        """locreg_bounds: the 17 convolution-bounds arrays of a free-BC
        region in one int32 buffer (layout: bounds_layout).  out: a device
        int32 tensor of bounds_layout(...)[0][17] elements (filled in stream
        order), else a numpy array is returned."""
        r = lr.region
        off, _ = bounds_layout(r.n, r.nfl, r.nfu)
        if out is not None:
            assert out.dtype == torch_int32() and out.numel() >= off[17]
            self._enter()
            try:
                self._check(lib().bdft_locreg_bounds_d(self.handle, lr.handle, _ptr(out)))
            finally:
                self._leave()
            return out
        buf = np.zeros(int(off[17]), dtype=np.int32)
        self._check(lib().bdft_locreg_bounds(self.handle, lr.handle, _ptr(buf)))
        return buf

    @_stream_ordered
    def plan(self, lrs, pot, psi, hpsi, ekin, epot, pot_off=None, potdim=None,
             ishift=None, psi_off=None, precision=PREC_FP64):
        arr, po, pd, sh, so = self._batch_args(lrs, pot_off, potdim, ishift,
                                               psi_off)
        npsi, npot = _batch_sizes(lrs, po, pd, so)
        n = len(lrs)
        h = ctypes.c_void_p()
        self._check(lib().bdft_plan_create(
            self.handle, n, arr, _dev(pot, "pot", npot), _ptr(po), _ptr(pd),
            _ptr(sh), _dev(psi, "psi", npsi), _ptr(so), _dev(hpsi, "hpsi", npsi),
            _dev(ekin, "ekin", n), _dev(epot, "epot", n), precision, ctypes.byref(h)))
        return Plan(self, h, (lrs, pot, psi, hpsi, ekin, epot))

    @_stream_ordered
    def uncompress(self, lr, norb, psi, psig):
        self._check(lib().bdft_uncompress_d(self.handle, lr.handle, norb,
                                            _ptr(psi), _ptr(psig)))

    @_stream_ordered
    def compress_acc(self, lr, norb, psig, hpsi):
        self._check(lib().bdft_compress_acc_d(self.handle, lr.handle, norb,
                                              _ptr(psig), _ptr(hpsi)))

    # ---- host-array entry point (numpy) ----
    def locham_host(self, lr, pot, psi, hpsi, precision=PREC_FP64, nspin=1,
                    spinsgn=None):
        """numpy in/out (bdft_local_hamiltonian); pot holds nspin potentials;
        hpsi is accumulated in place; returns (ekin, epot)."""
        psi = np.ascontiguousarray(psi, dtype=np.float64)
        norb = psi.shape[0] if psi.ndim > 1 else 1
        if not (hpsi.flags["C_CONTIGUOUS"] and hpsi.dtype == np.float64 and
                hpsi.size >= psi.size and psi.size >= norb * _ncoef(lr)):
            raise ValueError("psi and hpsi: contiguous float64, norb x ncoef")
        ek, ep = np.zeros(norb), np.zeros(norb)
        pot = np.ascontiguousarray(pot, dtype=np.float64)
        if pot.size < int(np.prod(lr.region.nfine)) * nspin:
            raise ValueError("pot: n1i*n2i*n3i*nspin values needed")
        sg = None if spinsgn is None else np.ascontiguousarray(spinsgn, dtype=np.float64)
        if sg is not None and sg.size != norb:
            raise ValueError("one spinsgn per orbital")
        self._check(lib().bdft_local_hamiltonian(
            self.handle, lr.handle, norb, int(nspin), _ptr(pot), _ptr(sg), _ptr(psi),
            _ptr(hpsi), _ptr(ek), _ptr(ep), precision))
        return ek, ep

    def close(self):
        if self.handle:
            lib().bdft_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def local_hamiltonian(ctx, lr, occup, pot, psi, hpsi, kwgt=1.0,
                      precision=PREC_FP64, nspin=1, spinsgn=None):
    """local_hamiltonian_OCL + finish_hamiltonian_OCL semantics
    (bigdft/src/interface_ocl.f90:391-607) on device tensors: hpsi += H_loc psi
    for every orbital, with pot = pot(n1i,n2i,n3i,nspin) and the spin-down
    potential for the orbitals whose spinsgn is not positive (:487-491), and
    the returned energies are
    ekin_sum = sum_i kwgt*occup_i*ekin_i, epot_sum = sum_i kwgt*occup_i*epot_i
    (the per-orbital weighting of hpsiortho.f90:946-954)."""
    import torch
    norb = len(occup)
    ekin = torch.zeros(norb, dtype=torch.float64, device=psi.device)
    epot = torch.zeros(norb, dtype=torch.float64, device=psi.device)
    ctx.locham(lr, pot, psi, hpsi, ekin, epot, precision, nspin=nspin, spinsgn=spinsgn)
    ctx.sync()
    w = np.asarray(occup, dtype=np.float64) * kwgt
    ek, ep = ekin.cpu().numpy(), epot.cpu().numpy()
    return float((w * ek).sum()), float((w * ep).sum()), ek, ep


def local_partial_density(ctx, lr, occup, psi, rho, kwgt=1.0, spinsgn=None,
                          nspin=1, precision=PREC_FP64):
    """local_partial_density_OCL semantics (bigdft/src/interface_ocl.f90:
    908-1040, CPU twin sumrho.f90:377-469) on device tensors:
    rho[s] += kwgt*occup_i/volume_element * psir_i^2, volume_element the
    fine-grid cell (hx/2)(hy/2)(hz/2); spin s = 0 where spinsgn_i > 0, else 1
    (sumrho.f90:515-519).  rho: device tensor (nspin, n1i*n2i*n3i)."""
    h = lr.region.hgrid
    vol = (0.5 * h[0]) * (0.5 * h[1]) * (0.5 * h[2])
    hfac = kwgt * (np.asarray(occup, dtype=np.float64) / vol)
    spin = None
    if spinsgn is not None:
        spin = np.where(np.asarray(spinsgn) > 0, 0, 1).astype(np.int32)
        if nspin == 1 and spin.any():
            raise ValueError("spin-down orbitals need nspin = 2")
    ctx.density(lr, psi, hfac, rho, spin=spin, nspin=nspin, precision=precision)


# bdft_confpot (confpot_data, bigdft/src/locregs/locreg_operations.f90:23-30)
CONFPOT_DTYPE = np.dtype([("potorder", np.int32), ("ioffset", np.int32, 3),
                          ("prefac", np.float64), ("hh", np.float64, 3),
                          ("rxyzConf", np.float64, 3), ("damping", np.float64)],
                         align=True)


def confpot(potorder, prefac, hh, center, ioffset, damping=1.0):
    """one bdft_confpot record"""
    c = np.zeros(1, dtype=CONFPOT_DTYPE)
    c["potorder"], c["prefac"], c["damping"] = potorder, prefac, damping
    c["hh"], c["rxyzConf"], c["ioffset"] = hh, center, ioffset
    return c


BOUNDS_NAMES = ("ibyz_c", "ibxz_c", "ibxy_c", "ibyz_f", "ibxz_f", "ibxy_f",
                "ibzxx_c", "ibxxyy_c", "ibyz_ff", "ibzxx_f", "ibxxyy_f",
                "ibzzx_c", "ibyyzz_c", "ibxy_ff", "ibzzx_f", "ibyyzz_f",
                "ibyyzz_r")


def bounds_layout(n, nfl, nfu):
    """(offset[18], dims[17, 4]) of the convolution-bounds buffer
    (bdft_bounds_layout): array k = BOUNDS_NAMES[k] starts at offset[k] and
    is the Fortran array ib(2, lo1:hi1, lo2:hi2), dims[k] = (lo1, hi1, lo2,
    hi2)."""
    I3 = ctypes.c_int * 3
    off = np.zeros(18, dtype=np.int64)
    dims = np.zeros(68, dtype=np.int32)
    rc = lib().bdft_bounds_layout(I3(*n), I3(*nfl), I3(*nfu), _ptr(off), _ptr(dims))
    if rc:
        raise BigDFTError(rc, "nfl > nfu or n < 0")
    return off, dims.reshape(17, 4)


def torch_int32():
    import torch
    return torch.int32
