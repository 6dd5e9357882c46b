"""ctypes binding of libopensps_hip.so (include/opensps_hip.h) with device buffers from PyTorch.

PyTorch provides device memory and streams only; every computation is a kernel of libopensps_hip.so.
Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import build as _build
from .params import NPAR

FP64 = 0
MIXED = 1

_lib = None


class BatchArgs(C.Structure):
    _fields_ = [("n", C.c_int), ("precision", C.c_int), ("split_parts", C.c_int), ("use_ops", C.c_int),
                ("prm", C.c_void_p),
                ("ssp_weights", C.c_void_p), ("mags", C.c_void_p), ("spec", C.c_void_p), ("mass", C.c_void_p),
                ("lbol", C.c_void_p), ("fband", C.c_void_p), ("young", C.c_void_p), ("old", C.c_void_p),
                ("idx", C.c_void_p), ("weights", C.c_void_p), ("op_count", C.c_void_p),
                ("ssp_weights_z", C.c_void_p), ("zred_mags", C.c_void_p), ("mdust", C.c_void_p),
                ("emlines", C.c_void_p)]


class GradArgs(C.Structure):
    _fields_ = [("fwd", BatchArgs), ("mags_bar", C.c_void_p), ("spec_bar", C.c_void_p), ("mass_bar", C.c_void_p),
                ("lbol_bar", C.c_void_p), ("params_bar", C.c_void_p), ("weights_bar", C.c_void_p),
                ("weights_z_bar", C.c_void_p)]


NBGRAD = 44


class SspOpsArgs(C.Structure):
    _fields_ = [("nz", C.c_int), ("nt", C.c_int), ("npts", C.c_int), ("ntpl", C.c_int), ("nent", C.c_int),
                ("age_ptr", C.c_void_p), ("mini", C.c_void_p), ("m1", C.c_void_p), ("m2", C.c_void_p),
                ("first", C.c_void_p), ("cm", C.c_void_p), ("cl", C.c_void_p), ("q", C.c_void_p),
                ("age_mmax", C.c_void_p), ("ent_tpl", C.c_void_p), ("ent_val", C.c_void_p), ("etpl", C.c_void_p)]


class OpenSPSHipError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"status {status}: {msg}")
        self.status = status


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_build.HIP_SO):
            raise OpenSPSHipError(1, f"{_build.HIP_SO} is not built (run __graft_entry__.build())")
        L = C.CDLL(_build.HIP_SO)
        vp, i, d = C.c_void_p, C.c_int, C.c_double
        L.opensps_hip_create.argtypes = [i, C.POINTER(vp)]
        L.opensps_hip_destroy.argtypes = [vp]
        L.opensps_hip_set_stream.argtypes = [vp, vp]
        L.opensps_hip_set_grid.argtypes = [vp, i, vp, i, vp, i, vp, vp, vp, vp]
        L.opensps_hip_set_bands.argtypes = [vp, i, vp]
        L.opensps_hip_set_mag_mask.argtypes = [vp, i, vp]
        L.opensps_hip_set_cosmology.argtypes = [vp, i, vp, vp]
        L.opensps_hip_set_sfh_tab.argtypes = [vp, i, vp, vp]
        L.opensps_hip_set_options.argtypes = [vp, i, i]
        L.opensps_hip_csp_batch.argtypes = [vp, C.POINTER(BatchArgs)]
        L.opensps_hip_csp_batch_noalloc.argtypes = [vp, C.POINTER(BatchArgs)]
        L.opensps_hip_sync.argtypes = [vp]
        L.opensps_hip_band_fluxes.argtypes = [i, i, i, i, vp, vp, vp, vp, vp, vp, vp, vp]
        L.opensps_hip_observed_frame.argtypes = [i, i, i, vp, vp, vp, vp, vp]
        L.opensps_hip_selftest_mfma.argtypes = [i, vp, vp, vp]
        L.opensps_hip_last_error.argtypes = [vp, C.c_char_p, i]
        L.opensps_hip_device_count.argtypes = [C.POINTER(i)]
        L.opensps_hip_set_ssp_ops.argtypes = [vp, C.POINTER(SspOpsArgs)]
        L.opensps_hip_free_ssp_ops.argtypes = [vp]
        L.opensps_hip_set_imf.argtypes = [vp, i, i, d, d, d, d]
        L.opensps_hip_csp_batch_grad.argtypes = [vp, C.POINTER(GradArgs)]
        L.opensps_hip_set_emission.argtypes = [vp, i, i]
        L.opensps_hip_set_dust_emission.argtypes = [vp, i, vp, i, vp, vp]
        L.opensps_hip_set_agn.argtypes = [vp, i, vp, vp]
        L.opensps_hip_set_nebular.argtypes = [vp, i, vp, i, vp, i, vp, i, vp, vp, i, vp, vp]
        L.opensps_hip_set_neb_options.argtypes = [vp, i, i, i]
        _lib = L
    return _lib


def last_error(ctx=None) -> str:
    buf = C.create_string_buffer(512)
    lib().opensps_hip_last_error(ctx, buf, 512)
    return buf.value.decode()


def check(st, ctx=None):
    if st != 0:
        raise OpenSPSHipError(st, last_error(ctx))


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _host(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


class DeviceContext:
    """One libopensps_hip context: resident bases, filters and cosmology on the GPU."""

    def __init__(self, device: int = 0):
        import torch  # noqa: F401  (device memory for the batch buffers)
        self._ctx = C.c_void_p()
        check(lib().opensps_hip_create(device, C.byref(self._ctx)))
        self.grid = None
        self.nbands = 0

    def close(self):
        if self._ctx:
            lib().opensps_hip_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        check(st, self._ctx)

    def set_stream(self, stream_handle):
        self._check(lib().opensps_hip_set_stream(self._ctx, C.c_void_p(stream_handle)))

    def set_grid(self, grid):
        lam, age, logz = _host(grid.lam), _host(grid.log_age), _host(grid.logz)
        spec, mass, lbol = _host(grid.spec), _host(grid.mass), _host(grid.lbol)
        self._check(lib().opensps_hip_set_grid(self._ctx, grid.nspec, lam.ctypes.data, grid.nt, age.ctypes.data,
                                               grid.nz, logz.ctypes.data, spec.ctypes.data, mass.ctypes.data,
                                               lbol.ctypes.data))
        self.grid = grid

    def set_bands(self, bands):
        b = _host(bands)
        self._check(lib().opensps_hip_set_bands(self._ctx, b.shape[0], b.ctypes.data))
        self.nbands = b.shape[0]

    def set_mag_mask(self, mask):
        m = None if mask is None else _host(mask, np.int32)
        self._check(lib().opensps_hip_set_mag_mask(self._ctx, self.nbands, None if m is None else m.ctypes.data))

    def set_sfh_tab(self, age_gyr, sfr):
        """The SFH table of sfh 2, 3: forward ages (Gyr) and SFRs (M_sun / yr)."""
        a = np.ascontiguousarray(np.asarray(age_gyr, dtype=np.float64) * 1.0e9)
        s = np.ascontiguousarray(sfr, dtype=np.float64)
        self._check(lib().opensps_hip_set_sfh_tab(self._ctx, a.size, a.ctypes.data, s.ctypes.data))

    def set_cosmology(self, z, dl):
        z, dl = _host(z), _host(dl)
        self._check(lib().opensps_hip_set_cosmology(self._ctx, z.size, z.ctypes.data, dl.ctypes.data))

    def set_options(self, add_igm=1, smooth_fast=1):
        self._check(lib().opensps_hip_set_options(self._ctx, int(add_igm), int(smooth_fast)))

    def set_ssp_ops(self, op):
        """Installs a synth.SSPOperator (the IMF-per-model operator set) on the context."""
        a = SspOpsArgs()
        a.nz, a.nt, a.npts, a.ntpl, a.nent = op.nz, op.nt, op.npts, op.ntpl, op.nent
        keep = []
        for k in ("age_ptr", "first", "ent_tpl"):
            v = _host(getattr(op, k), np.int32)
            keep.append(v)
            setattr(a, k, v.ctypes.data)
        for k in ("mini", "m1", "m2", "cm", "cl", "q", "age_mmax", "ent_val", "etpl"):
            v = _host(getattr(op, k))
            keep.append(v)
            setattr(a, k, v.ctypes.data)
        self._check(lib().opensps_hip_set_ssp_ops(self._ctx, C.byref(a)))

    def set_emission(self, add_dust_emission=1, add_agn_dust=1):
        self._check(lib().opensps_hip_set_emission(self._ctx, int(add_dust_emission), int(add_agn_dust)))

    def set_dust_emission(self, qpah, umin, dustem):
        """The Draine & Li templates on the grid: dustem (nqpah, 2 numin, nspec) (C order; for each (q_PAH, U_min) the
        U_max = U_min spectrum, then the power-law one), qpah, umin ascending; None removes them."""
        if dustem is None:
            self._check(lib().opensps_hip_set_dust_emission(self._ctx, 0, None, 0, None, None))
            return
        q, u = _host(qpah), _host(umin)
        t = np.ascontiguousarray(np.transpose(_host(dustem), (1, 0, 2)))   # (2 numin, nqpah, nspec) C order
        assert t.shape == (2 * u.size, q.size, self.grid.nspec)
        self._check(lib().opensps_hip_set_dust_emission(self._ctx, q.size, q.ctypes.data, u.size, u.ctypes.data,
                                                        t.ctypes.data))

    def set_agn(self, tau, agn):
        """The AGN torus templates on the grid: agn (ntau, nspec), tau ascending; None removes them."""
        if agn is None:
            self._check(lib().opensps_hip_set_agn(self._ctx, 0, None, None))
            return
        t, a = _host(tau), _host(agn)
        assert a.shape == (t.size, self.grid.nspec)
        self._check(lib().opensps_hip_set_agn(self._ctx, t.size, t.ctypes.data, a.ctypes.data))

    def set_nebular(self, library):
        """The CLOUDY library (logz, age (yr), logu, wavelength (nlam), continuum (nnz, nna, nnu, nlam) L_sun/Hz per unit
        Q, line wavelengths (nline), lines (nnz, nna, nnu, nline) L_sun per unit Q: datalib.nebular's tuple, or
        synth.nebular_library's), placed on the grid by the library (opensps_hip_set_nebular); None removes it."""
        if library is None:
            self._check(lib().opensps_hip_set_nebular(self._ctx, 0, None, 0, None, 0, None, 0, None, None, 0, None,
                                                      None))
            self.nline = 0
            return
        lz, age, lu, lam, cont, llam, lines = (_host(x) for x in library)
        assert cont.shape == (lz.size, age.size, lu.size, lam.size)
        assert lines.shape == (lz.size, age.size, lu.size, llam.size)
        self._check(lib().opensps_hip_set_nebular(self._ctx, lz.size, lz.ctypes.data, age.size, age.ctypes.data,
                                                  lu.size, lu.ctypes.data, lam.size, lam.ctypes.data,
                                                  cont.ctypes.data, llam.size, llam.ctypes.data, lines.ctypes.data))
        self.nline = llam.size

    def set_neb_options(self, add_neb_emission=1, add_neb_continuum=1, nebemlineinspec=1):
        self._check(lib().opensps_hip_set_neb_options(self._ctx, int(add_neb_emission), int(add_neb_continuum),
                                                      int(nebemlineinspec)))

    def free_ssp_ops(self):
        self._check(lib().opensps_hip_free_ssp_ops(self._ctx))

    def set_imf(self, imf_type=2, remnants=1, salp_ind=2.35, mdave=0.5, mlim_bh=40.0, mlim_ns=8.5):
        self._check(lib().opensps_hip_set_imf(self._ctx, int(imf_type), int(remnants), float(salp_ind), float(mdave),
                                              float(mlim_bh), float(mlim_ns)))

    def alloc_outputs(self, n, precision, want=("mags", "spec", "mass", "lbol")):
        import torch
        dev = "cuda"
        sdt = torch.float32 if precision == MIXED else torch.float64
        g = self.grid
        out = {"mags": torch.empty((n, self.nbands), dtype=torch.float64, device=dev)}
        shapes = {"spec": ((n, g.nspec), sdt), "mass": ((n,), torch.float64), "lbol": ((n,), torch.float64),
                  "fband": ((n, self.nbands), torch.float64), "young": ((n, g.nspec), sdt),
                  "old": ((n, g.nspec), sdt), "idx": ((n, 3), torch.int32), "weights": ((n, g.nt), torch.float64),
                  "op_count": ((n, 2), torch.int32), "mdust": ((n,), torch.float64),
                  "emlines": ((n, max(1, getattr(self, "nline", 0))), torch.float64)}
        for k in want:
            if k in shapes:
                out[k] = torch.empty(shapes[k][0], dtype=shapes[k][1], device=dev)
        return out

    def _batch_args(self, prm, out, precision, split_parts, ssp_weights, use_ops, ssp_weights_z=None, zred_mags=None):
        n = prm.shape[1]
        assert prm.shape[0] == NPAR and prm.is_contiguous() and prm.dtype.itemsize == 8
        g = self.grid
        isz = 4 if precision == MIXED else 8
        sizes = {"mags": (n * self.nbands, 8), "spec": (n * g.nspec, isz), "young": (n * g.nspec, isz),
                 "old": (n * g.nspec, isz), "mass": (n, 8), "lbol": (n, 8), "fband": (n * self.nbands, 8),
                 "idx": (3 * n, 4), "weights": (n * g.nt, 8), "op_count": (2 * n, 4), "mdust": (n, 8),
                 "emlines": (n * max(1, getattr(self, "nline", 0)), 8)}
        for k, (numel, itemsize) in sizes.items():
            t = out.get(k)
            if t is not None and (not t.is_contiguous() or t.element_size() != itemsize or t.numel() < numel):
                raise OpenSPSHipError(3, f"{k}: {numel} contiguous {itemsize}-byte elements required")
        if ssp_weights is not None and (ssp_weights.element_size() != 8 or ssp_weights.numel() < n * g.nt
                                        or not ssp_weights.is_contiguous()):
            raise OpenSPSHipError(3, "ssp_weights: nt x n float64 required")
        if ssp_weights_z is not None and (ssp_weights_z.element_size() != 8 or not ssp_weights_z.is_contiguous()
                                          or ssp_weights_z.numel() < n * g.nt * g.nz):
            raise OpenSPSHipError(3, "ssp_weights_z: nt x nz x n float64 required")
        if zred_mags is not None and (zred_mags.element_size() != 8 or not zred_mags.is_contiguous()
                                      or zred_mags.numel() < n):
            raise OpenSPSHipError(3, "zred_mags: n float64 required")
        a = BatchArgs()
        a.n, a.precision, a.split_parts, a.use_ops = n, precision, int(split_parts), int(use_ops)
        a.prm = prm.data_ptr()
        a.ssp_weights = ssp_weights.data_ptr() if ssp_weights is not None else None
        a.ssp_weights_z = ssp_weights_z.data_ptr() if ssp_weights_z is not None else None
        a.zred_mags = zred_mags.data_ptr() if zred_mags is not None else None
        for k in ("mags", "spec", "mass", "lbol", "fband", "young", "old", "idx", "weights", "op_count", "mdust",
                  "emlines"):
            setattr(a, k, out[k].data_ptr() if k in out else None)
        return a

    def launch_grad(self, prm, out, cot, grads, precision=MIXED, split_parts=0, ssp_weights=None, ssp_weights_z=None,
                    use_ops=0, zred_mags=None):
        """Enqueues a batch and its reverse pass. cot: device cotangents {"mags", "spec", "mass", "lbol"} (any
        absent); grads: device outputs {"params_bar" (n, NBGRAD), "weights_bar" (n, nt), "weights_z_bar"
        (n, nz, nt)} (any absent), float64. use_ops: the IMF per model (the operator), with its rows in params_bar;
        zred_mags: device (n) float64, each model's magnitude redshift (zred's row then holds the IGM's part)."""
        n = prm.shape[1]
        g = self.grid
        want = {"mags": n * self.nbands, "spec": n * g.nspec, "mass": n, "lbol": n}
        for k, numel in want.items():
            t = cot.get(k)
            if t is not None and (t.element_size() != 8 or not t.is_contiguous() or t.numel() < numel):
                raise OpenSPSHipError(3, f"{k}_bar: {numel} contiguous float64 required")
        gsz = {"params_bar": n * NBGRAD, "weights_bar": n * g.nt, "weights_z_bar": n * g.nt * g.nz}
        for k, numel in gsz.items():
            t = grads.get(k)
            if t is not None and (t.element_size() != 8 or not t.is_contiguous() or t.numel() < numel):
                raise OpenSPSHipError(3, f"{k}: {numel} contiguous float64 required")
        a = GradArgs()
        a.fwd = self._batch_args(prm, out, precision, split_parts, ssp_weights, use_ops, ssp_weights_z, zred_mags)
        for k in ("mags", "spec", "mass", "lbol"):
            setattr(a, k + "_bar", cot[k].data_ptr() if cot.get(k) is not None else None)
        for k in ("params_bar", "weights_bar", "weights_z_bar"):
            setattr(a, k, grads[k].data_ptr() if grads.get(k) is not None else None)
        self._check(lib().opensps_hip_csp_batch_grad(self._ctx, C.byref(a)))

    def launch(self, prm, out, precision=MIXED, split_parts=0, ssp_weights=None, use_ops=0, ssp_weights_z=None,
               noalloc=False, zred_mags=None):
        """Enqueues one batch on device-resident prm (NPAR, n) float64 and the buffers of out; noalloc: through
        opensps_hip_csp_batch_noalloc (the fixed scratch, no device allocation)."""
        a = self._batch_args(prm, out, precision, split_parts, ssp_weights, use_ops, ssp_weights_z, zred_mags)
        fn = lib().opensps_hip_csp_batch_noalloc if noalloc else lib().opensps_hip_csp_batch
        self._check(fn(self._ctx, C.byref(a)))

    def sync(self):
        self._check(lib().opensps_hip_sync(self._ctx))

    def batch(self, prm, precision=MIXED, split_parts=0, ssp_weights=None, want=("mags", "spec", "mass", "lbol"),
              use_ops=0, ssp_weights_z=None, noalloc=False, zred_mags=None):
        import torch
        n = prm.shape[1]
        out = self.alloc_outputs(n, precision, want)
        torch.cuda.synchronize()
        self.launch(prm, out, precision, split_parts, ssp_weights, use_ops, ssp_weights_z, noalloc, zred_mags)
        self.sync()
        return out

    def batch_grad(self, prm, cot, precision=MIXED, split_parts=0, ssp_weights=None, ssp_weights_z=None,
                   want=("mags", "spec", "mass", "lbol"), grads=("params_bar", "weights_bar", "weights_z_bar"),
                   use_ops=0, zred_mags=None):
        """The forward outputs and the reverse pass's outputs (device tensors)."""
        import torch
        n = prm.shape[1]
        g = self.grid
        out = self.alloc_outputs(n, precision, want)
        shapes = {"params_bar": (n, NBGRAD), "weights_bar": (n, g.nt), "weights_z_bar": (n, g.nz, g.nt)}
        gr = {k: torch.full(shapes[k], float("nan"), dtype=torch.float64, device="cuda") for k in grads}
        torch.cuda.synchronize()
        self.launch_grad(prm, out, cot, gr, precision, split_parts, ssp_weights, ssp_weights_z, use_ops, zred_mags)
        self.sync()
        out.update(gr)
        return out


def needs_split(prm_host) -> bool:
    """True if some model has a birth-cloud or an old-star screen (the young and old sums are then apart)."""
    from .params import SLOT
    tiny = np.finfo(np.float64).tiny
    d1, fob, d3 = prm_host[SLOT["dust1"]], prm_host[SLOT["frac_obrun"]], prm_host[SLOT["dust3"]]
    return bool(np.any(((d1 > tiny) & (1.0 - fob > tiny)) | (np.abs(d3) > tiny)))


def selftest_mfma(which, a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    d = np.zeros((64, 4), dtype=np.float64 if which == 0 else np.float32)
    check(lib().opensps_hip_selftest_mfma(which, a.ctypes.data, b.ctypes.data, d.ctypes.data))
    return d


def band_fluxes(precision, bands_t, jlo, jhi, do_band, wt, fobs_t):
    """Device tensors: bands_t (nspec, nbands) C order [= Fortran (nbands, nspec)], fobs_t (nspec, n) C order
    [= Fortran (n, nspec)]. Returns fband (n, nbands)."""
    import torch
    nspec, nbands = bands_t.shape
    n = fobs_t.shape[1]
    out = torch.zeros((n, nbands), dtype=torch.float64, device=bands_t.device)
    torch.cuda.synchronize()
    check(lib().opensps_hip_band_fluxes(precision, n, nspec, nbands, _ptr(bands_t), _ptr(jlo), _ptr(jhi),
                                        _ptr(do_band), _ptr(wt), _ptr(fobs_t), _ptr(out), None))
    return out


def observed_frame(precision, zc, lam, total):
    """Device tensors: zc (n), lam (nspec), total (n, nspec) C order. Returns fobs_t (nspec, n) C order."""
    import torch
    n, nspec = total.shape
    out = torch.zeros((nspec, n), dtype=total.dtype, device=total.device)
    torch.cuda.synchronize()
    check(lib().opensps_hip_observed_frame(precision, n, nspec, _ptr(zc), _ptr(lam), _ptr(total), _ptr(out), None))
    return out
