"""The C-ABI libraries without a GPU: every declared symbol is exported, and the host logic of the public API
(handles, setters, the basis, filters and the statuses a batch returns before it reaches the device) behaves as
the reference's C API does (tests/test_c_batch.c:137-146, tests/test_c_ssp_operator.c:45-56)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from helpers import grid
from libopensps_amd import api, build, synth
from libopensps_amd.params import NPAR, PARAM_DEFAULTS, PARAM_NAMES

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = re.sub(r"//.*", "", txt)
    inline = set(re.findall(r"static\s+inline\s+\w+\s+(opensps_\w+)\s*\(", txt))   # header-only, not exported
    return sorted(set(re.findall(r"\b(opensps_\w+)\s*\(", txt)) - inline)


def _exported(so):
    out = subprocess.run(["nm", "-D", "--defined-only", so], stdout=subprocess.PIPE, text=True, check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if " T " in l}


@pytest.mark.parametrize("header,so", [("opensps.h", build.API_SO), ("opensps_hip.h", build.HIP_SO),
                                       ("opensps_emu.h", build.HIP_SO)])
def test_every_declared_symbol_is_exported(header, so):
    assert os.path.exists(so), f"{so} not built"
    C.CDLL(so)   # loads without a GPU
    missing = [s for s in _declared(header) if s not in _exported(so)]
    assert not missing, missing


def test_param_slots_match_header():
    """The Fortran host's slot table equals include/opensps_params.h (names and defaults, via get_params)."""
    with api.Context() as ctx:
        np.testing.assert_array_equal(ctx.params(), np.asarray(PARAM_DEFAULTS))
    src = open(os.path.join(ROOT, "libopensps_amd", "fortran", "opensps_host_types.f90")).read()
    names = re.search(r"PARAM_NAMES\(NPAR\) = \[character\(len=16\) :: &(.*?)\]", src, re.S).group(1)
    assert re.findall(r"'(\w+)'", names) == PARAM_NAMES
    assert NPAR == 35


def test_handles_and_setters():
    with api.Context() as ctx:
        ctx.set_int("add_igm_absorption", 1)
        ctx.set_float("dust2", 0.3)
        ctx.set_str("kernel_precision", "mixed")
        with pytest.raises(api.OpenSPSError) as e:
            ctx.set_int("no_such_key", 1)
        assert e.value.status == 101
        with pytest.raises(api.OpenSPSError) as e:
            ctx.set_float("no_such_key", 1.0)
        assert e.value.status == 102
        with pytest.raises(api.OpenSPSError) as e:
            ctx.set_str("kernel_precision", "quad")
        assert e.value.status == 103
        assert ctx.params()[PARAM_NAMES.index("dust2")] == 0.3
    st = C.c_int()
    api.lib().opensps_context_destroy(98765, C.byref(st))
    assert st.value == 1
    assert api.nbgrad() == 44
    assert api.grad_param_name(0) == "fagn" and api.grad_param_name(22) == "zred"


def test_setup_from_data_release_is_refused():
    with api.Context() as ctx:
        st = C.c_int()
        api.lib().opensps_context_setup(0, b"mist", b"miles", b"dl07", ctx.handle, C.byref(st))
        assert st.value == 301


def test_basis_and_filters_match_oracle():
    """The Fortran host's filters_build equals the oracle's (point and hat resampling), and dims / grids are kept."""
    g = grid(22)
    fs = synth.gaussian_filters(30, seed=7)
    for resample in (0, 1):
        with api.Context() as ctx:
            ctx.set_ssp_basis(g)
            ctx.set_filters(fs, resample=resample)
            b = ctx.bands()
            ref = O.build_bands(fs, g.lam, resample)
            # 1e-13 per pixel, and 1e-13 of the band's peak where a curve's end cuts a grid cell's last ~1 %: the hat
            # weights there differ two nearly equal wavelengths, so any form of the projection has ~1e-11 relative
            # rounding on values ~1e-9 of the peak
            peak = ref.max(axis=1, keepdims=True)
            assert np.all(np.abs(b - ref) <= 1e-13 * np.maximum(np.abs(ref), peak)), resample
            if resample == 0:
                np.testing.assert_allclose(b, ref, rtol=1e-13, atol=0)
            assert ctx.nbands == 30
            we = ctx.wave_eff()
            cen = np.sort(np.array([0.5 * (l[0] + l[-1]) for l in fs.lam]))
            assert np.all(np.abs(we / cen - 1.0) < 0.01)


def test_bad_basis_and_filters():
    g = grid(1)
    with api.Context() as ctx:
        bad = synth.SSPGrid(lam=g.lam[::-1].copy(), log_age=g.log_age, logz=g.logz, spec=g.spec, mass=g.mass,
                            lbol=g.lbol)
        with pytest.raises(api.OpenSPSError) as e:
            ctx.set_ssp_basis(bad)
        assert e.value.status == 3
        fs = synth.gaussian_filters(2)
        with pytest.raises(api.OpenSPSError) as e:
            ctx.set_filters(fs)
        assert e.value.status == 301
        ctx.set_ssp_basis(g)
        far = synth.FilterSet(lam=[np.linspace(2e8, 3e8, 50)], resp=[np.ones(50)], names=["far"])
        with pytest.raises(api.OpenSPSError) as e:
            ctx.set_filters(far, grid_policy=0)
        assert e.value.status in (321, 322)
        nanf = synth.FilterSet(lam=[np.array([1000.0, np.nan, 2000.0])], resp=[np.ones(3)], names=["nan"])
        with pytest.raises(api.OpenSPSError) as e:
            ctx.set_filters(nanf)
        assert e.value.status == 320


def test_batch_statuses_before_the_device():
    """Statuses the host returns before any device work: unknown name 3, an SSP parameter 2, no basis 1,
    spectral indices 2, dust emission with dust 2, n = 0 is a no-op."""
    g = grid(1)
    fs = synth.gaussian_filters(4)
    with api.Context() as ctx:
        r = ctx.compute_csp_batch({"tage": np.ones(3)}, raw_status=True)
        assert r["status"] == 1
        ctx.set_ssp_basis(g)
        ctx.set_filters(fs)
        r = ctx.compute_csp_batch({"no_such_param": np.ones(3)}, raw_status=True)
        assert r["status"] == 3
        r = ctx.compute_csp_batch({"imf_type": np.ones(3)}, raw_status=True)
        assert r["status"] == 2
        r = ctx.compute_csp_batch({"tage": np.ones(3), "dust2": np.ones(3)}, raw_status=True)
        assert r["status"] == 2      # add_dust_emission = 1 by default (reference src/core/opensps_context.f90:78)
        assert "dust emission" in api.last_error() and "templates" in api.last_error()
        ctx.set_int("add_dust_emission", 0)
        r = ctx.compute_csp_batch({}, n=0, raw_status=True)
        assert r["status"] == 0
        st = C.c_int()
        mags = np.zeros((2, 4))
        indx = np.zeros(10)
        api.lib().opensps_context_compute_csp_batch(ctx.handle, 2, 0, None, None, None, mags.ctypes.data, None, None,
                                                    None, None, None, None, indx.ctypes.data, C.byref(st))
        assert st.value == 2


def test_sfh_tab_statuses():
    """opensps_context_set_sfh_tab: ntab out of 1 .. 20000 is status 3."""
    with api.Context() as ctx:
        ctx.set_sfh_tab([0.0, 1.0], [1.0, 2.0], [0.01, 0.02])
        ctx.set_sfh_tab([0.0, 1.0], [1.0, 2.0])
        for bad in (np.zeros(0), np.zeros(20001)):
            with pytest.raises(api.OpenSPSError) as e:
                ctx.set_sfh_tab(bad, bad)
            assert e.value.status == 3


def test_batch_without_device_is_status_5():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    g = grid(1)
    with api.Context() as ctx:
        ctx.set_ssp_basis(g)
        ctx.set_filters(synth.gaussian_filters(4))
        ctx.set_int("add_dust_emission", 0)
        r = ctx.compute_csp_batch({"tage": np.full(2, 5.0), "sfh": np.full(2, 4.0)}, raw_status=True)
        assert r["status"] == 5


def test_batch_grad_statuses_before_the_device():
    """compute_csp_batch_grad's host checks (opensps_c_api.f90:729-735 for the row count): fewer params_bar rows
    than NBGRAD is status 3 with nothing written; spectral indices are status 2; both weight forms at once is
    status 3; zred_mags passes the host's checks (redshifted colours need it) and reaches the device."""
    g = grid(1, 600, 20)
    fs = synth.gaussian_filters(4, seed=7)
    with api.Context() as ctx:
        ctx.set_ssp_basis(g)
        ctx.set_filters(fs)
        n = 3
        vals = {"dust2": np.full(n, 0.3)}
        w = np.full((n, g.nt), 1.0 / g.nt)
        r = ctx.compute_csp_batch_grad(vals, ssp_weights=w, grads=("params_bar",), params_bar_rows=10,
                                       raw_status=True)
        assert r["status"] == 3 and np.all(r["params_bar"] == 0.0)
        r = ctx.compute_csp_batch_grad(vals, ssp_weights=w, ssp_weights_z=np.ones((n, 1, g.nt)), raw_status=True)
        assert r["status"] == 3
        zm = np.zeros(n)
        st = C.c_int()
        keep = [np.ascontiguousarray(vals["dust2"])]
        names = (C.c_char_p * 1)(b"dust2")
        vp = (C.c_void_p * 1)(keep[0].ctypes.data)
        mags = np.zeros((n, 4))
        indx = np.zeros((n, 4))
        api.lib().opensps_context_compute_csp_batch_grad(
            ctx.handle, n, 1, C.cast(names, C.c_void_p), C.cast(vp, C.c_void_p), w.ctypes.data, None, None,
            None, None, None, None, mags.ctypes.data, None, None, None, None, None, None, indx.ctypes.data,
            None, None, None, C.byref(st))
        assert st.value == 2
        import torch
        if not torch.cuda.is_available():
            ctx.set_int("add_dust_emission", 0)
            ctx.set_int("redshift_colors", 1)
            r = ctx.compute_csp_batch_grad(vals, ssp_weights=w, grads=(), raw_status=True)
            assert r["status"] == 2                       # redshifted colours without zred_mags
            r = ctx.compute_csp_batch_grad(vals, ssp_weights=w, grads=(), zred_mags=zm, raw_status=True)
            assert r["status"] == 5                       # past the host's checks: no device here


def test_dust_type_statuses():
    """dust_type 0, 2 (Calzetti) and 6 (Reddy) reach the device (status 5 here without one); the others are 2."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    g = grid(1)
    with api.Context() as ctx:
        ctx.set_ssp_basis(g)
        ctx.set_filters(synth.gaussian_filters(4))
        ctx.set_int("add_dust_emission", 0)
        for dt, want in ((1, 2), (2, 5), (3, 2), (4, 2), (5, 2), (6, 5)):
            ctx.set_int("dust_type", dt)
            r = ctx.compute_csp_batch({"tage": np.full(2, 5.0), "sfh": np.full(2, 4.0)}, raw_status=True)
            assert r["status"] == want, (dt, r["status"])



def test_imf_and_ssp_parameters_without_operator():
    """Without an SSP operator on the context the IMF's parameters change the SSPs: a batch naming them is status 2,
    as the reference's C test checks for imf1 (tests/test_c_batch.c:136-146); so is every parameter of
    is_ssp_sensitive_float_param (src/core/opensps_context.f90:958-969) that cannot act in the CSP."""
    g = grid(1)
    with api.Context() as ctx:
        ctx.set_ssp_basis(g)
        ctx.set_filters(synth.gaussian_filters(4))
        ctx.set_int("add_dust_emission", 0)
        assert not ctx.has_ssp_ops()
        for nm in ("imf1", "imf2", "imf3", "imf_upper_limit", "imf_lower_limit", "vdmc", "mdave", "tiny_logt",
                   "logt_wmb_hot", "frac_xrb", "agb_dust", "nebular_smooth_init", "pagb", "masscut", "afe"):
            r = ctx.compute_csp_batch({"tage": np.full(3, 5.0), nm: np.ones(3)}, raw_status=True)
            assert r["status"] == 2, (nm, r["status"])
        r = ctx.compute_csp_batch({"not_a_parameter": np.ones(3)}, raw_status=True)
        assert r["status"] == 3
        # fagn > 0 asks for the AGN stage (add_agn_dust = 1 by default): its templates are needed
        r = ctx.compute_csp_batch({"tage": np.full(3, 5.0), "fagn": np.full(3, 0.1)}, raw_status=True)
        assert r["status"] == 2 and "AGN" in api.last_error()


def test_emission_templates_api():
    """opensps_context_set_dust_emission_templates / _set_agn_templates: the test release's DL07 and clumpy libraries
    are taken; sizes below two or wavelengths not ascending are status 3; with the templates, dust emission and the
    AGN pass the host's checks (status 5 here: no device); the batch with gradients refuses them (status 2)."""
    import torch
    from libopensps_amd import datalib
    root = os.path.join(ROOT, "tests", "golden", "opensps-test-data")
    q, u, lam, sp = datalib.dust_emission(root)
    tau, alam, asp = datalib.agn_torus(root)
    g = grid(1)
    with api.Context() as ctx:
        ctx.set_ssp_basis(g)
        ctx.set_filters(synth.gaussian_filters(4))
        with pytest.raises(api.OpenSPSError) as e:
            ctx.set_dust_emission_templates(q, u, lam[::-1].copy(), sp)
        assert e.value.status == 3
        with pytest.raises(api.OpenSPSError) as e:
            ctx.set_agn_templates(tau[:1], alam, asp[:1])
        assert e.value.status == 3
        ctx.set_dust_emission_templates(q, u, lam, sp)
        ctx.set_agn_templates(tau, alam, asp)
        vals = {"tage": np.full(3, 5.0), "dust2": np.full(3, 0.5), "fagn": np.full(3, 0.1)}
        r = ctx.compute_csp_batch_grad(vals, mags_bar=np.ones((3, 4)), grads=("params_bar",), raw_status=True)
        assert r["status"] == 2 and "reverse pass" in api.last_error()
        if not torch.cuda.is_available():
            r = ctx.compute_csp_batch(vals, raw_status=True)
            assert r["status"] == 5


def test_nebular_tables_api():
    """opensps_context_set_nebular_tables: the test release's CLOUDY library is taken and sets get_nemline to its
    lines; axes not ascending or below two points are status 3; nebular emission without the tables, in the SSPs
    (add_neb_emission = 2) or with gradients is status 2; with them the batch passes the host's checks (status 5
    here: no device)."""
    import torch
    from libopensps_amd import datalib
    logz, age, logu, lam, cont, llam, lines = datalib.nebular(os.path.join(ROOT, "tests", "golden",
                                                                         "opensps-test-data"))
    g = grid(1)
    with api.Context() as ctx:
        ctx.set_ssp_basis(g)
        ctx.set_filters(synth.gaussian_filters(4))
        ctx.set_int("add_dust_emission", 0)
        assert ctx.nemline == 166
        vals = {"tage": np.full(3, 5.0), "gas_logu": np.full(3, -2.5), "gas_logz": np.zeros(3)}
        ctx.set_int("add_neb_emission", 1)
        r = ctx.compute_csp_batch(vals, raw_status=True)
        assert r["status"] == 2 and "set_nebular_tables" in api.last_error()
        with pytest.raises(api.OpenSPSError) as e:
            ctx.set_nebular_tables(logz[::-1].copy(), age, logu, lam, cont, llam, lines)
        assert e.value.status == 3
        with pytest.raises(api.OpenSPSError) as e:
            ctx.set_nebular_tables(logz[:1], age, logu, lam, cont[:1], llam, lines[:1])
        assert e.value.status == 3
        ctx.set_nebular_tables(logz, age, logu, lam, cont, llam, lines)
        assert ctx.nemline == llam.size
        r = ctx.compute_csp_batch_grad(vals, mags_bar=np.ones((3, 4)), grads=("params_bar",), raw_status=True)
        assert r["status"] == 2 and "nebular" in api.last_error()
        ctx.set_int("add_neb_emission", 2)
        r = ctx.compute_csp_batch(vals, raw_status=True)
        assert r["status"] == 2 and "add_neb_emission = 2" in api.last_error()
        ctx.set_int("add_neb_emission", 1)
        ctx.set_int("add_neb_continuum", 0)
        ctx.set_int("nebemlineinspec", 0)
        if not torch.cuda.is_available():
            r = ctx.compute_csp_batch(vals, want=("mags", "emlines"), raw_status=True)
            assert r["status"] == 5 and r["emlines"].shape == (3, llam.size)


def test_parameters_no_stage_reads_are_accepted():
    """Members of the parameter set that no stage of this build reads (dust curves of other types, nebular, dust
    emission without dust, fagn = 0) pass the host's checks and reach the device (status 5 here without one)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    g = grid(1)
    with api.Context() as ctx:
        ctx.set_ssp_basis(g)
        ctx.set_filters(synth.gaussian_filters(4))
        ctx.set_int("add_dust_emission", 0)
        vals = {k: np.ones(2) for k in ("pmetals", "dust_clumps", "uvb", "mwr", "gas_logu", "gas_logz", "gas_logn",
                                        "gas_logno", "gas_logco", "agn_tau", "duste_gamma", "duste_umin",
                                        "duste_qpah")}
        vals["fagn"] = np.zeros(2)
        vals["tage"] = np.full(2, 5.0)
        r = ctx.compute_csp_batch(vals, raw_status=True)
        assert r["status"] == 5
        ctx.set_float("mwr", 4.0)
        ctx.set_float("gas_logu", -3.0)


def test_ssp_operator_api_statuses():
    """The SSP-operator calls' statuses (the reference's tests/test_c_ssp_operator.c:14-59, 149-160): building from a
    data release is 301 here (this build takes the operator from the caller), bad handles 1, evaluate before an
    operator 303, has_ssp_ops 0, free is a no-op."""
    st = C.c_int()
    h = C.c_int()
    L = api.lib()
    for fn in ("opensps_context_build_ssp_ops", "opensps_context_build_ssp_ops_free_mods"):
        getattr(L, fn)(999, 0, C.byref(st))
        assert st.value == 1
    L.opensps_context_free_ssp_ops(999, C.byref(st))
    assert st.value == 1
    L.opensps_context_has_ssp_ops(999, C.byref(h), C.byref(st))
    assert st.value == 1
    L.opensps_context_evaluate_ssp_op(999, 2, None, None, None, None, C.byref(st))
    assert st.value == 1
    g = grid(3, 600, 20)
    with api.Context() as ctx:
        with pytest.raises(api.OpenSPSError) as e:
            ctx.build_ssp_ops(0)
        assert e.value.status == 301
        L.opensps_context_build_ssp_ops_free_mods(ctx.handle, 0, C.byref(st))
        assert st.value == 301
        op = synth.ssp_operator(g.lam, g.log_age, 3, npts=8, ntpl=20)
        with pytest.raises(api.OpenSPSError) as e:   # no basis yet
            ctx.set_ssp_ops(op)
        assert e.value.status == 1
        ctx.set_ssp_basis(g)
        assert not ctx.has_ssp_ops()
        with pytest.raises(api.OpenSPSError) as e:
            ctx.evaluate_ssp_op(2)
        assert e.value.status == 303
        ctx.free_ssp_ops()
        bad = synth.ssp_operator(g.lam, g.log_age[:10], 3, npts=8, ntpl=20)
        with pytest.raises(api.OpenSPSError) as e:   # ages other than the basis's
            ctx.set_ssp_ops(bad)
        assert e.value.status == 3


def test_batch_ext_null_arguments():
    """The header-only struct entry point (include/opensps.h, the reference's :366-421) returns status 1 for NULL
    options, compiled from C through the tests' shim."""
    shim = C.CDLL(build.SHIM_SO)
    st = C.c_int(-7)
    shim.shim_batch_ext(1, 0, 0, None, None, None, None, None, None, None, None, None, None, None, None, None, 0,
                        None, None, 1, C.byref(st))
    assert st.value == 1


def test_cosmology_getter_matches_oracle():
    """opensps_context_get_cosmology: the host's cosmology table (the one the distance term interpolates) equals the
    oracle's, which is pinned by quadrature of 1/E(z); it follows set_float H0."""
    with api.Context() as ctx:
        z, dl = ctx.cosmology()
        oz, odl = O.cosmology_table(72.0, 0.27, 0.73)
        np.testing.assert_array_equal(z, oz)
        np.testing.assert_allclose(dl, odl, rtol=1e-13, atol=0)
        ctx.set_float("H0", 70.0)
        z2, dl2 = ctx.cosmology()
        np.testing.assert_allclose(dl2[1:], dl[1:] * 72.0 / 70.0, rtol=1e-12)
