"""Reader of the reference's standard data-library format (format "opensps-data", version 1): a root index
opensps-data.json listing libraries/<kind>/<name>/, each a manifest.json with raw little-endian arrays (one array a
file, C order, no header), their dtype, shape and SHA-256 in the manifest. The reference reads it in Fortran
(src/io/opensps_datalib.f90, src/io/opensps_json.f90; the format: doc/DATA_FORMAT.md sections 2-3).

The kinds the batched path uses: reference_spectra (Vega and the Sun, for the filters' zero points and the Vega
system; the reference's load_standard_sed, src/io/opensps_io.f90:1516-1582), filters (the curves of
opensps_context_set_filters_library, src/spectra/opensps_context_filters.f90:148-170), dust_emission and agn_torus
(the templates of the dust emission and the AGN torus, load_dust_emission_table and load_agn_dust_models,
src/io/opensps_io.f90:959-1155). Host-side I/O only: the tables go to the context through its C ABI
(Context.set_reference_spectra, Context.set_filters, Context.set_dust_emission_templates,
Context.set_agn_templates)."""
from __future__ import annotations

import hashlib
import json
import os
from dataclasses import dataclass, field

import numpy as np

FORMAT = "opensps-data"
FORMAT_VERSION = 1
_DTYPES = {"<f4", "<f8", "<i4", "<i8"}


class DataLibraryError(RuntimeError):
    pass


def open_root(root: str) -> dict:
    """The root index (opensps-data.json) of a data release, checked for the format and its version."""
    path = os.path.join(root, "opensps-data.json")
    try:
        with open(path, encoding="utf-8") as f:
            idx = json.load(f)
    except OSError as e:
        raise DataLibraryError(f"{path}: {e}") from e
    if idx.get("format") != FORMAT or idx.get("format_version") != FORMAT_VERSION:
        raise DataLibraryError(f"{path}: not an {FORMAT} v{FORMAT_VERSION} index")
    return idx


def library_dir(root: str, kind: str, name: str) -> str:
    """The directory of library kind/name, as the root index lists it."""
    for lib in open_root(root).get("libraries", []):
        if lib.get("kind") == kind and lib.get("name") == name:
            return os.path.join(root, lib["path"])
    raise DataLibraryError(f"{root}: no library {kind}/{name}")


def read_manifest(libdir: str, kind: str | None = None) -> dict:
    path = os.path.join(libdir, "manifest.json")
    with open(path, encoding="utf-8") as f:
        m = json.load(f)
    if m.get("format") != FORMAT or m.get("format_version") != FORMAT_VERSION:
        raise DataLibraryError(f"{path}: not an {FORMAT} v{FORMAT_VERSION} manifest")
    if kind is not None and m.get("kind") != kind:
        raise DataLibraryError(f"{path}: kind {m.get('kind')!r}, expected {kind!r}")
    return m


def read_array(libdir: str, spec: dict, verify: bool = True) -> np.ndarray:
    """One array (or file-backed axis) of a manifest: exactly prod(shape) items of its dtype, its SHA-256 checked."""
    dt = spec.get("dtype")
    if dt not in _DTYPES:
        raise DataLibraryError(f"{libdir}: dtype {dt!r} is not one of {sorted(_DTYPES)}")
    shape = tuple(int(s) for s in spec["shape"])
    path = os.path.join(libdir, spec["file"])
    with open(path, "rb") as f:
        raw = f.read()
    want = int(np.prod(shape)) * np.dtype(dt).itemsize
    if len(raw) != want:
        raise DataLibraryError(f"{path}: {len(raw)} bytes, the manifest's shape {shape} of {dt} is {want}")
    if verify and "sha256" in spec and hashlib.sha256(raw).hexdigest() != spec["sha256"]:
        raise DataLibraryError(f"{path}: SHA-256 differs from the manifest's")
    return np.frombuffer(raw, dtype=dt).reshape(shape)


def _axis(libdir: str, spec: dict, verify: bool) -> np.ndarray:
    if "values" in spec:
        return np.asarray(spec["values"], dtype=np.float64)
    return read_array(libdir, spec, verify)


def reference_spectrum(root: str, name: str, verify: bool = True):
    """reference_spectra/<name> (vega, sun): (wavelength in Angstrom, flux as FP64, is_f_lambda) as
    read_reference_sed returns them (the flux promoted from its stored precision; metadata flux_density f_lambda or
    f_nu)."""
    d = library_dir(root, "reference_spectra", name)
    m = read_manifest(d, "reference_spectra")
    lam = np.asarray(_axis(d, m["axes"]["wavelength"], verify), dtype=np.float64)
    flux = np.asarray(read_array(d, m["arrays"]["flux"], verify), dtype=np.float64)
    fd = m.get("metadata", {}).get("flux_density")
    if fd not in ("f_lambda", "f_nu"):
        raise DataLibraryError(f"{d}: flux_density {fd!r} is neither f_lambda nor f_nu")
    if lam.shape != flux.shape:
        raise DataLibraryError(f"{d}: wavelength {lam.shape} and flux {flux.shape} differ")
    return lam, flux, fd == "f_lambda"


@dataclass
class FilterLibrary:
    """The curves of a filters/<name> library as tabulated (concatenated curves split by npoints), with each
    filter's detector (0 photon, 1 energy) and calibration (lambda_ref, exponent; lambda_ref 0: none)."""
    names: list = field(default_factory=list)
    lam: list = field(default_factory=list)
    resp: list = field(default_factory=list)
    detector: list = field(default_factory=list)
    cal: list = field(default_factory=list)


def filter_library(root: str, name: str = "allfilters", verify: bool = True) -> FilterLibrary:
    d = library_dir(root, "filters", name)
    m = read_manifest(d, "filters")
    npts = np.asarray(read_array(d, m["arrays"]["npoints"], verify), dtype=np.int64)
    wl = np.asarray(read_array(d, m["arrays"]["wavelength"], verify), dtype=np.float64)
    tr = np.asarray(read_array(d, m["arrays"]["transmission"], verify), dtype=np.float64)
    meta = m.get("metadata", {}).get("filters", [])
    if len(meta) != npts.size or int(npts.sum()) != wl.size or wl.size != tr.size or np.any(npts < 1):
        raise DataLibraryError(f"{d}: npoints, curves and per-filter metadata disagree")
    out = FilterLibrary()
    off = 0
    for k, n in enumerate(npts):
        e = meta[k]
        det = e.get("detector", "photon")
        if det not in ("photon", "energy"):
            raise DataLibraryError(f"{d}: filter {e.get('name')}: detector {det!r}")
        cal = e.get("calibration")
        out.names.append(e["name"])
        out.lam.append(wl[off:off + n].copy())
        out.resp.append(tr[off:off + n].copy())
        out.detector.append(0 if det == "photon" else 1)
        out.cal.append((float(cal["lambda_ref"]), float(cal["exponent"])) if cal else (0.0, 0.0))
        off += int(n)
    return out


def dust_emission(root: str, name: str = "DL07", verify: bool = True):
    """dust_emission/<name>: (qpah (nqpah), umin (numin), wavelength (nlam, A), spectra (nqpah, numin, 2, nlam) FP64),
    the arrays load_dust_emission_table reads (src/io/opensps_io.f90:986-1011): per (q_PAH, U_min) the spectrum of
    U_max = U_min, then of the power-law U distribution (components umin, umin_to_umax)."""
    d = library_dir(root, "dust_emission", name)
    m = read_manifest(d, "dust_emission")
    q = _axis(d, m["axes"]["qpah"], verify)
    u = _axis(d, m["axes"]["umin"], verify)
    lam = np.asarray(_axis(d, m["axes"]["wavelength"], verify), dtype=np.float64)
    sp = m["arrays"]["spectra"]
    if sp.get("axes") != ["qpah", "umin", "component", "wavelength"]:
        raise DataLibraryError(f"{d}: spectra axes {sp.get('axes')} are not [qpah, umin, component, wavelength]")
    s = np.asarray(read_array(d, sp, verify), dtype=np.float64)
    if s.shape != (q.size, u.size, 2, lam.size):
        raise DataLibraryError(f"{d}: spectra {s.shape} against the axes {(q.size, u.size, 2, lam.size)}")
    return q, u, lam, s


def agn_torus(root: str, name: str = "clumpy", verify: bool = True):
    """agn_torus/<name>: (tau (ntau), wavelength (nlam, A), spectra (ntau, nlam) FP64), the arrays
    load_agn_dust_models reads (src/io/opensps_io.f90:1110-1120)."""
    d = library_dir(root, "agn_torus", name)
    m = read_manifest(d, "agn_torus")
    tau = _axis(d, m["axes"]["tau"], verify)
    lam = np.asarray(_axis(d, m["axes"]["wavelength"], verify), dtype=np.float64)
    sp = m["arrays"]["spectra"]
    if sp.get("axes") != ["tau", "wavelength"]:
        raise DataLibraryError(f"{d}: spectra axes {sp.get('axes')} are not [tau, wavelength]")
    s = np.asarray(read_array(d, sp, verify), dtype=np.float64)
    if s.shape != (tau.size, lam.size):
        raise DataLibraryError(f"{d}: spectra {s.shape} against the axes {(tau.size, lam.size)}")
    return tau, lam, s


def nebular(root: str, name: str = "mist_nodust", verify: bool = True):
    """nebular/<name>: (logz (nnz), age (nna, yr), logu (nnu), wavelength (nlam, A), continuum (nnz, nna, nnu, nlam)
    L_sun/Hz per unit ionizing photon rate, line wavelengths (nline, A), lines (nnz, nna, nnu, nline) L_sun per unit
    ionizing photon rate), FP64: the arrays read_nebular_standard reads (src/io/opensps_io.f90:914-937)."""
    d = library_dir(root, "nebular", name)
    m = read_manifest(d, "nebular")
    ax = m["axes"]
    logz = np.asarray(ax["logZ"]["values"], dtype=np.float64)
    age = np.asarray(ax["age"]["values"], dtype=np.float64)
    logu = np.asarray(ax["logU"]["values"], dtype=np.float64)
    lam = np.asarray(_axis(d, ax["wavelength"], verify), dtype=np.float64)
    llam = np.asarray(_axis(d, ax["line_wavelength"], verify), dtype=np.float64)
    c, l = m["arrays"]["continuum"], m["arrays"]["lines"]
    if c.get("axes") != ["logZ", "age", "logU", "wavelength"]:
        raise DataLibraryError(f"{d}: continuum axes {c.get('axes')} are not [logZ, age, logU, wavelength]")
    if l.get("axes") != ["logZ", "age", "logU", "line_wavelength"]:
        raise DataLibraryError(f"{d}: lines axes {l.get('axes')} are not [logZ, age, logU, line_wavelength]")
    cont = np.asarray(read_array(d, c, verify), dtype=np.float64)
    lines = np.asarray(read_array(d, l, verify), dtype=np.float64)
    if cont.shape != (logz.size, age.size, logu.size, lam.size):
        raise DataLibraryError(f"{d}: continuum {cont.shape} against the axes")
    if lines.shape != (logz.size, age.size, logu.size, llam.size):
        raise DataLibraryError(f"{d}: lines {lines.shape} against the axes")
    return logz, age, logu, lam, cont, llam, lines
