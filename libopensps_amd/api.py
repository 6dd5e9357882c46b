"""ctypes binding of the public C API (libopensps.so, include/opensps.h): the Python side of the drop-in.

The calls mirror the reference's C API for the batched path (include/opensps.h:132-455), with the same names,
argument meaning and status codes; the Fortran host behind them runs the batch on the GPU through
libopensps_hip.so. Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import build as _build
from .params import NPAR, PARAM_NAMES

NBGRAD = 44   # params_bar rows (opensps_get_nbgrad)

_lib = None

STATUS_MESSAGES = {0: "ok", 1: "bad handle or not set up", 2: "unsupported", 3: "unknown parameter or bad shape",
                   5: "no device", 10: "device error", 301: "not set up"}


class OpenSPSError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"status {status}: {msg}")
        self.status = status


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_build.API_SO):
            raise OpenSPSError(1, f"{_build.API_SO} is not built (run __graft_entry__.build())")
        L = C.CDLL(_build.API_SO)
        i, d, vp, ip, cp = C.c_int, C.c_double, C.c_void_p, C.POINTER(C.c_int), C.c_char_p
        L.opensps_context_create.argtypes = [ip, ip]
        L.opensps_context_destroy.argtypes = [i, ip]
        L.opensps_context_setup.argtypes = [i, cp, cp, cp, i, ip]
        L.opensps_context_set_ssp_basis.argtypes = [i, i, vp, i, vp, i, vp, d, vp, vp, vp, ip]
        L.opensps_context_set_int.argtypes = [i, cp, i, ip]
        L.opensps_context_set_float.argtypes = [i, cp, d, ip]
        L.opensps_context_set_str.argtypes = [i, cp, cp, ip]
        L.opensps_context_get_dims.argtypes = [i, ip, ip, ip]
        L.opensps_context_get_nbands.argtypes = [i, ip, ip]
        L.opensps_context_get_nz.argtypes = [i, ip, ip]
        L.opensps_context_get_nemline.argtypes = [i, ip, ip]
        L.opensps_context_get_lambda.argtypes = [i, vp, ip]
        L.opensps_context_get_bands.argtypes = [i, vp, ip]
        L.opensps_context_get_cosmology.argtypes = [i, ip, vp, vp, ip]
        L.opensps_context_get_params.argtypes = [i, vp, ip]
        L.opensps_context_get_filter_data.argtypes = [i, vp, vp, vp, ip]
        L.opensps_context_set_reference_spectra.argtypes = [i, i, vp, vp, i, i, vp, vp, i, ip]
        L.opensps_context_set_sfh_tab.argtypes = [i, i, vp, vp, vp, ip]
        L.opensps_filterset_create.argtypes = [ip, ip]
        L.opensps_filterset_add.argtypes = [i, cp, i, vp, vp, i, d, d, ip]
        L.opensps_filterset_count.argtypes = [i, ip, ip]
        L.opensps_filterset_free.argtypes = [i, ip]
        L.opensps_context_set_filters.argtypes = [i, i, i, i, i, ip]
        L.opensps_context_set_mag_compute.argtypes = [i, i, vp, ip]
        L.opensps_context_compute_csp_batch.argtypes = [i, i, i, vp, vp, vp] + [vp] * 8 + [ip]
        L.opensps_context_compute_csp_batch_parts.argtypes = [i, i, i, vp, vp, vp] + [vp] * 14 + [ip]
        L.opensps_context_compute_csp_batch_device.argtypes = [i, i, vp, vp, i, i, vp, vp, vp, vp, ip]
        L.opensps_context_compute_csp_batch_grad.argtypes = [i, i, i, vp, vp] + [vp] * 18 + [ip]
        L.opensps_context_compute_csp_batch_grad_rows.argtypes = [i, i, i, vp, vp] + [vp] * 16 + [i, vp, vp, ip]
        L.opensps_context_batch_spec_device.argtypes = [i, C.POINTER(vp), ip, ip, ip, ip]
        L.opensps_context_get_batch_maggies.argtypes = [i, vp, ip, ip, ip]
        L.opensps_get_last_error.argtypes = [ip, C.c_char_p, i]
        L.opensps_get_nbgrad.argtypes = [ip]
        L.opensps_get_grad_param_name.argtypes = [i, C.c_char_p, i, ip]
        L.opensps_get_version_string.restype = C.c_char_p
        L.opensps_context_build_ssp_ops.argtypes = [i, i, ip]
        L.opensps_context_build_ssp_ops_free_mods.argtypes = [i, i, ip]
        L.opensps_context_set_ssp_ops.argtypes = [i, i, i, i, i, i] + [vp] * 12 + [ip]
        L.opensps_context_set_dust_emission_templates.argtypes = [i, i, vp, i, vp, i, vp, vp, ip]
        L.opensps_context_set_agn_templates.argtypes = [i, i, vp, i, vp, vp, ip]
        L.opensps_context_set_nebular_tables.argtypes = [i, i, vp, i, vp, i, vp, i, vp, vp, i, vp, vp, ip]
        L.opensps_context_free_ssp_ops.argtypes = [i, ip]
        L.opensps_context_has_ssp_ops.argtypes = [i, ip, ip]
        L.opensps_context_evaluate_ssp_op.argtypes = [i, i, vp, vp, vp, vp, ip]
        _lib = L
    return _lib


def last_error() -> str:
    st = C.c_int()
    buf = C.create_string_buffer(512)
    lib().opensps_get_last_error(C.byref(st), buf, 512)
    return buf.value.decode()


def _check(st):
    v = st.value if isinstance(st, C.c_int) else int(st)
    if v != 0:
        raise OpenSPSError(v, last_error())


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _need(t, name, numel, itemsize, exact_shape=None):
    """Rejects a device tensor that the batch would overrun: wrong element size, too few elements, not contiguous,
    not on a GPU."""
    if t is None:
        return
    if not t.is_cuda or not t.is_contiguous():
        raise OpenSPSError(3, f"{name}: a contiguous GPU tensor is required")
    if t.element_size() != itemsize:
        raise OpenSPSError(3, f"{name}: {itemsize}-byte elements required for this precision, got {t.dtype}")
    if exact_shape is not None and tuple(t.shape) != tuple(exact_shape):
        raise OpenSPSError(3, f"{name}: shape {tuple(exact_shape)} required, got {tuple(t.shape)}")
    if t.numel() < numel:
        raise OpenSPSError(3, f"{name}: {numel} elements required, got {t.numel()}")


class Context:
    """One OpenSPS context (an int handle of the C API)."""

    def __init__(self):
        h, st = C.c_int(), C.c_int()
        lib().opensps_context_create(C.byref(h), C.byref(st))
        _check(st)
        self.handle = h.value
        self.nspec = self.nt = self.nz = self.nbands = 0
        self.mixed = False      # the context's kernel_precision ("mixed" or "double")

    def close(self):
        if self.handle:
            st = C.c_int()
            lib().opensps_context_destroy(self.handle, C.byref(st))
            self.handle = 0

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- settings ----
    def set_int(self, key, value):
        st = C.c_int()
        lib().opensps_context_set_int(self.handle, key.encode(), int(value), C.byref(st))
        _check(st)
        if key.strip().lower() == "mixed_precision":
            self.mixed = int(value) != 0

    def set_float(self, key, value):
        st = C.c_int()
        lib().opensps_context_set_float(self.handle, key.encode(), float(value), C.byref(st))
        _check(st)

    def set_str(self, key, value):
        st = C.c_int()
        lib().opensps_context_set_str(self.handle, key.encode(), value.encode(), C.byref(st))
        _check(st)
        if key.strip().lower() == "kernel_precision":
            self.mixed = value.strip().lower() == "mixed"

    def set_ssp_basis(self, grid, zsol=0.0142):
        st = C.c_int()
        lam, age, zl = _d(grid.lam), _d(grid.log_age), _d(zsol * 10.0 ** np.asarray(grid.logz))
        spec, mass, lbol = _d(grid.spec), _d(grid.mass), _d(grid.lbol)
        lib().opensps_context_set_ssp_basis(self.handle, grid.nspec, lam.ctypes.data, grid.nt, age.ctypes.data,
                                            grid.nz, zl.ctypes.data, float(zsol), spec.ctypes.data, mass.ctypes.data,
                                            lbol.ctypes.data, C.byref(st))
        _check(st)
        self.nspec, self.nt, self.nz = grid.nspec, grid.nt, grid.nz

    def set_reference_spectra(self, vega, sun):
        """vega, sun: (wavelength in Angstrom, flux, is_f_lambda) tables, e.g. datalib.reference_spectrum(root,
        "vega"); they give the filters' Vega and solar zero points and the Vega system (compute_vega_mags)."""
        (lv, fv, flv), (ls, fs, fls) = vega, sun
        lv, fv, ls, fs = _d(lv), _d(fv), _d(ls), _d(fs)
        st = C.c_int()
        lib().opensps_context_set_reference_spectra(self.handle, lv.size, lv.ctypes.data, fv.ctypes.data, int(bool(flv)),
                                                    ls.size, ls.ctypes.data, fs.ctypes.data, int(bool(fls)),
                                                    C.byref(st))
        _check(st)

    def set_sfh_tab(self, age_gyr, sfr, met=None):
        """The SFH table of sfh 2, 3 (opensps_context_set_sfh_tab): forward ages (Gyr), SFRs (M_sun / yr), Z."""
        a, s = _d(age_gyr), _d(sfr)
        m = None if met is None else _d(met)
        st = C.c_int()
        lib().opensps_context_set_sfh_tab(self.handle, a.size, a.ctypes.data, s.ctypes.data,
                                          None if m is None else m.ctypes.data, C.byref(st))
        _check(st)

    def set_filters(self, fset, resample=0, grid_policy=2, vega_ref_band=0):
        fs, st = C.c_int(), C.c_int()
        lib().opensps_filterset_create(C.byref(fs), C.byref(st))
        _check(st)
        nf = len(fset.names)
        det = fset.detector if isinstance(fset.detector, (list, tuple, np.ndarray)) else [fset.detector] * nf
        cal = getattr(fset, "cal", None) or [(0.0, 0.0)] * nf   # (lambda_ref, exponent) per filter; 0: none
        try:
            for k, (name, lam, resp) in enumerate(zip(fset.names, fset.lam, fset.resp)):
                l, r = _d(lam), _d(resp)
                lib().opensps_filterset_add(fs.value, name.encode(), l.size, l.ctypes.data, r.ctypes.data,
                                            int(det[k]), float(cal[k][0]), float(cal[k][1]), C.byref(st))
                _check(st)
            lib().opensps_context_set_filters(self.handle, fs.value, int(resample), int(grid_policy),
                                              int(vega_ref_band), C.byref(st))
            _check(st)
        finally:
            lib().opensps_filterset_free(fs.value, C.byref(C.c_int()))
        nb = C.c_int()
        lib().opensps_context_get_nbands(self.handle, C.byref(nb), C.byref(st))
        self.nbands = nb.value

    def set_mag_compute(self, mask):
        st = C.c_int()
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.int32)
        lib().opensps_context_set_mag_compute(self.handle, self.nbands if m is None else m.size,
                                              None if m is None else m.ctypes.data, C.byref(st))
        _check(st)

    def bands(self):
        """(nbands, nspec) bands on the grid."""
        out = np.zeros((self.nbands, self.nspec))
        st = C.c_int()
        lib().opensps_context_get_bands(self.handle, out.ctypes.data, C.byref(st))
        _check(st)
        return out

    def cosmology(self):
        """(z, dl): the context's cosmology table (redshift, luminosity distance in pc)."""
        nr, st = C.c_int(), C.c_int()
        lib().opensps_context_get_cosmology(self.handle, C.byref(nr), None, None, C.byref(st))
        _check(st)
        z, dl = np.zeros(nr.value), np.zeros(nr.value)
        lib().opensps_context_get_cosmology(self.handle, C.byref(nr), z.ctypes.data, dl.ctypes.data, C.byref(st))
        _check(st)
        return z, dl

    def params(self):
        out = np.zeros(NPAR)
        st = C.c_int()
        lib().opensps_context_get_params(self.handle, out.ctypes.data, C.byref(st))
        _check(st)
        return out

    def wave_eff(self):
        out = np.zeros(self.nbands)
        st = C.c_int()
        lib().opensps_context_get_filter_data(self.handle, out.ctypes.data, None, None, C.byref(st))
        _check(st)
        return out

    def filter_data(self):
        """(wave_eff, mag_vega, mag_sun) per band (opensps_context_get_filter_data)."""
        out = np.zeros((3, self.nbands))
        st = C.c_int()
        lib().opensps_context_get_filter_data(self.handle, out[0].ctypes.data, out[1].ctypes.data, out[2].ctypes.data,
                                              C.byref(st))
        _check(st)
        return out[0], out[1], out[2]

    # ---- the SSP operator (the IMF per model) ----
    def build_ssp_ops(self, all_metallicities=1):
        """opensps_context_build_ssp_ops: building from a data release is not part of this build (status 301)."""
        st = C.c_int()
        lib().opensps_context_build_ssp_ops(self.handle, int(all_metallicities), C.byref(st))
        _check(st)

    def set_dust_emission_templates(self, qpah, umin, lam, spectra):
        """opensps_context_set_dust_emission_templates: spectra (nqpah, numin, 2, nlam) as datalib.dust_emission
        returns them (per (q_PAH, U_min) the U_max = U_min spectrum, then the power-law one), lam in Angstrom."""
        q, u, lm = _d(qpah), _d(umin), _d(lam)
        sp = _d(spectra)
        assert sp.shape == (q.size, u.size, 2, lm.size)
        st = C.c_int()
        lib().opensps_context_set_dust_emission_templates(self.handle, q.size, q.ctypes.data, u.size, u.ctypes.data,
                                                          lm.size, lm.ctypes.data, sp.ctypes.data, C.byref(st))
        _check(st)

    def set_agn_templates(self, tau, lam, spectra):
        """opensps_context_set_agn_templates: spectra (ntau, nlam) as datalib.agn_torus returns them."""
        t, lm, sp = _d(tau), _d(lam), _d(spectra)
        assert sp.shape == (t.size, lm.size)
        st = C.c_int()
        lib().opensps_context_set_agn_templates(self.handle, t.size, t.ctypes.data, lm.size, lm.ctypes.data,
                                                sp.ctypes.data, C.byref(st))
        _check(st)

    def set_nebular_tables(self, logz, age, logu, lam, cont, line_lam, lines):
        """opensps_context_set_nebular_tables: the CLOUDY library as datalib.nebular returns it: axes logz, age (yr),
        logu, the continuum's wavelengths lam, cont (nnz, nna, nnu, nlam), the lines' wavelengths line_lam and lines
        (nnz, nna, nnu, nline)."""
        z, a, u, lm, co, ll, li = (_d(x) for x in (logz, age, logu, lam, cont, line_lam, lines))
        assert co.shape == (z.size, a.size, u.size, lm.size) and li.shape == (z.size, a.size, u.size, ll.size)
        st = C.c_int()
        lib().opensps_context_set_nebular_tables(self.handle, z.size, z.ctypes.data, a.size, a.ctypes.data, u.size,
                                                 u.ctypes.data, lm.size, lm.ctypes.data, co.ctypes.data, ll.size,
                                                 ll.ctypes.data, li.ctypes.data, C.byref(st))
        _check(st)

    @property
    def nemline(self):
        """opensps_context_get_nemline: the rows of emlines (the nebular library's lines, else 166)."""
        n, st = C.c_int(), C.c_int()
        lib().opensps_context_get_nemline(self.handle, C.byref(n), C.byref(st))
        _check(st)
        return n.value

    def set_ssp_ops(self, op):
        """Installs a caller-built operator (a synth.SSPOperator or anything with its fields) on the GPU."""
        ints = {k: np.ascontiguousarray(getattr(op, k), dtype=np.int32) for k in ("age_ptr", "first", "ent_tpl")}
        dbl = {k: _d(getattr(op, k)) for k in ("mini", "m1", "m2", "cm", "cl", "q", "age_mmax", "ent_val", "etpl")}
        P = lambda a: a.ctypes.data
        st = C.c_int()
        lib().opensps_context_set_ssp_ops(self.handle, op.nz, op.nt, op.npts, op.ntpl, op.nent, P(ints["age_ptr"]),
                                          P(dbl["mini"]), P(dbl["m1"]), P(dbl["m2"]), P(ints["first"]), P(dbl["cm"]),
                                          P(dbl["cl"]), P(dbl["q"]), P(dbl["age_mmax"]), P(ints["ent_tpl"]),
                                          P(dbl["ent_val"]), P(dbl["etpl"]), C.byref(st))
        _check(st)

    def free_ssp_ops(self):
        st = C.c_int()
        lib().opensps_context_free_ssp_ops(self.handle, C.byref(st))
        _check(st)

    def has_ssp_ops(self) -> bool:
        h, st = C.c_int(), C.c_int()
        lib().opensps_context_has_ssp_ops(self.handle, C.byref(h), C.byref(st))
        _check(st)
        return h.value == 1

    def evaluate_ssp_op(self, imf_type=-1, imf_params=None):
        """opensps_context_evaluate_ssp_op: (mass (nt,), log10 lbol (nt,), spec (nt, nspec)) of every age for one
        IMF; imf_params = (imf1, imf2, imf3, vdmc, mdave) or None for the context's."""
        mass, lbol, spec = np.zeros(self.nt), np.zeros(self.nt), np.zeros((self.nt, self.nspec))
        ip = None if imf_params is None else _d(imf_params)
        st = C.c_int()
        lib().opensps_context_evaluate_ssp_op(self.handle, int(imf_type), None if ip is None else ip.ctypes.data,
                                              mass.ctypes.data, lbol.ctypes.data, spec.ctypes.data, C.byref(st))
        _check(st)
        return mass, lbol, spec

    # ---- batches ----
    def compute_csp_batch(self, values: dict, n=None, ssp_weights=None, want=("mags", "spec", "mass", "lbol"),
                          parts=False, raw_status=False):
        """values: {name: array of n}. Returns a dict of (n, ...) C-order arrays (quantity fastest in memory)."""
        names = list(values.keys())
        if n is None:
            n = len(next(iter(values.values()))) if values else (ssp_weights.shape[0] if ssp_weights is not None else 0)
        keep = [_d(values[k]) for k in names]
        cnames = (C.c_char_p * max(len(names), 1))(*[k.encode() for k in names])
        cvals = (C.c_void_p * max(len(names), 1))(*[a.ctypes.data for a in keep])
        out = {"mags": np.zeros((n, self.nbands))}
        shapes = {"spec": (n, self.nspec), "mass": (n,), "lbol": (n,), "mdust": (n,), "sfr": (n,),
                  "emlines": (n, self.nemline), "young": (n, self.nspec), "old": (n, self.nspec), "mass_young": (n,),
                  "mass_old": (n,), "lbol_young": (n,), "lbol_old": (n,)}
        for k in want:
            if k in shapes:
                out[k] = np.zeros(shapes[k])
        if parts:
            for k in ("young", "old"):
                out.setdefault(k, np.zeros(shapes[k]))
        w = _d(ssp_weights) if ssp_weights is not None else None
        P = lambda k: out[k].ctypes.data if k in out else None
        st = C.c_int()
        if parts:
            lib().opensps_context_compute_csp_batch_parts(
                self.handle, n, len(names), C.cast(cnames, C.c_void_p), C.cast(cvals, C.c_void_p),
                None if w is None else w.ctypes.data, P("mags"), P("spec"), P("mass"), P("lbol"), P("mdust"),
                P("sfr"), P("emlines"), None, P("young"), P("old"), P("mass_young"), P("mass_old"),
                P("lbol_young"), P("lbol_old"), C.byref(st))
        else:
            lib().opensps_context_compute_csp_batch(
                self.handle, n, len(names), C.cast(cnames, C.c_void_p), C.cast(cvals, C.c_void_p),
                None if w is None else w.ctypes.data, P("mags"), P("spec"), P("mass"), P("lbol"), P("mdust"),
                P("sfr"), P("emlines"), None, C.byref(st))
        if raw_status:
            out["status"] = st.value
            return out
        _check(st)
        return out

    def compute_csp_batch_grad(self, values: dict, n=None, ssp_weights=None, ssp_weights_z=None, mags_bar=None,
                               spec_bar=None, mass_bar=None, lbol_bar=None, want=("mags",),
                               grads=("params_bar", "weights_bar", "weights_z_bar"), params_bar_rows=None,
                               raw_status=False, zred_mags=None):
        """The batch with gradients (opensps_context_compute_csp_batch_grad / _rows). values: {name: array of n};
        ssp_weights (n, ntfull); ssp_weights_z (n, nz, ntfull); zred_mags (n): each model's magnitude redshift;
        cotangents mags_bar (n, nbands), spec_bar (n, nspec), mass_bar, lbol_bar (n). Returns the forward outputs
        and params_bar (n, rows), weights_bar (n, ntfull), weights_z_bar (n, nz, ntfull)."""
        names = list(values.keys())
        if n is None:
            n = len(next(iter(values.values())))
        keep = [_d(values[k]) for k in names]
        cnames = (C.c_char_p * max(len(names), 1))(*[k.encode() for k in names])
        cvals = (C.c_void_p * max(len(names), 1))(*[a.ctypes.data for a in keep])
        rows = NBGRAD if params_bar_rows is None else int(params_bar_rows)
        out = {"mags": np.zeros((n, self.nbands))}
        shapes = {"spec": (n, self.nspec), "mass": (n,), "lbol": (n,), "mdust": (n,), "sfr": (n,),
                  "params_bar": (n, max(rows, 1)), "weights_bar": (n, self.nt),
                  "weights_z_bar": (n, self.nz, self.nt)}
        for k in tuple(want) + tuple(grads):
            if k in shapes:
                out[k] = np.zeros(shapes[k])
        opt = lambda a: None if a is None else _d(a)
        ins = {k: opt(v) for k, v in (("w", ssp_weights), ("wz", ssp_weights_z), ("mb", mags_bar), ("sb", spec_bar),
                                      ("mab", mass_bar), ("lb", lbol_bar), ("zm", zred_mags))}
        I = lambda k: None if ins[k] is None else ins[k].ctypes.data
        P = lambda k: out[k].ctypes.data if k in out else None
        st = C.c_int()
        args = [self.handle, n, len(names), C.cast(cnames, C.c_void_p), C.cast(cvals, C.c_void_p), I("w"), I("wz"), I("zm"),
                I("mb"), I("sb"), I("mab"), I("lb"), P("mags"), P("spec"), P("mass"), P("lbol"), P("mdust"), P("sfr"),
                None, None, P("params_bar")]
        if params_bar_rows is None:
            lib().opensps_context_compute_csp_batch_grad(*args, P("weights_bar"), P("weights_z_bar"), C.byref(st))
        else:
            lib().opensps_context_compute_csp_batch_grad_rows(*args, rows, P("weights_bar"), P("weights_z_bar"),
                                                              C.byref(st))
        if raw_status:
            out["status"] = st.value
            return out
        _check(st)
        return out

    def compute_csp_batch_params(self, prm, names, **kw):
        """The batch with the named rows of a (NPAR, n) parameter array as names[] / values[]."""
        from .params import SLOT
        return self.compute_csp_batch({k: prm[SLOT[k]] for k in names}, n=prm.shape[1], **kw)

    def compute_csp_batch_device(self, d_prm, d_mags, d_spec=None, d_mass=None, d_lbol=None, d_weights=None,
                                 precision=-1, split=0):
        """The batch on device-resident torch tensors (no host transfer). The tensors are checked against the
        batch before the launch: the kernels write n x nspec spectra of 4 bytes (mixed) or 8 bytes (FP64)."""
        n = d_prm.shape[1] if d_prm.dim() == 2 else -1
        mixed = self.mixed if precision < 0 else precision == 1
        _need(d_prm, "d_prm", NPAR * n, 8, exact_shape=(NPAR, n))
        _need(d_mags, "d_mags", n * self.nbands, 8)
        _need(d_spec, "d_spec", n * self.nspec, 4 if mixed else 8)
        _need(d_mass, "d_mass", n, 8)
        _need(d_lbol, "d_lbol", n, 8)
        _need(d_weights, "d_weights", n * self.nt, 8)
        st = C.c_int()
        ptr = lambda t: None if t is None else t.data_ptr()
        lib().opensps_context_compute_csp_batch_device(self.handle, d_prm.shape[1], ptr(d_prm), ptr(d_weights),
                                                       int(precision), int(split), ptr(d_mags), ptr(d_spec),
                                                       ptr(d_mass), ptr(d_lbol), C.byref(st))
        _check(st)

    def batch_spec_device(self):
        p, ns, n, b, st = C.c_void_p(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        lib().opensps_context_batch_spec_device(self.handle, C.byref(p), C.byref(ns), C.byref(n), C.byref(b),
                                                C.byref(st))
        _check(st)
        return p.value, ns.value, n.value, b.value

    def batch_maggies(self):
        nb, n, st = C.c_int(), C.c_int(), C.c_int()
        lib().opensps_context_get_batch_maggies(self.handle, None, C.byref(nb), C.byref(n), C.byref(st))
        _check(st)
        out = np.zeros((n.value, nb.value))
        lib().opensps_context_get_batch_maggies(self.handle, out.ctypes.data, C.byref(nb), C.byref(n), C.byref(st))
        _check(st)
        return out


def nbgrad() -> int:
    v = C.c_int()
    lib().opensps_get_nbgrad(C.byref(v))
    return v.value


def grad_param_name(i) -> str:
    buf = C.create_string_buffer(32)
    st = C.c_int()
    lib().opensps_get_grad_param_name(int(i), buf, 32, C.byref(st))
    _check(st)
    return buf.value.decode()


def version() -> str:
    return lib().opensps_get_version_string().decode()
