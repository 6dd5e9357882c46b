#!/usr/bin/env python3
"""Benchmark of the batched composite-population path on one MI355X.

Workload (the parametric CSP batch): 16 384 models on a synthetic SSP library of 107 ages x 22 metallicities x
5994 wavelengths; per model a delayed-tau SFH (tage U(0.1, 13.7) Gyr, tau U(0.1, 10) Gyr), logzsol U(-2, 0.3)
blended between the bracketing metallicities, power-law dust (dust2 U(0, 2), slope U(-1.2, -0.4)), redshift U(0, 4)
with Madau IGM, a Gaussian velocity dispersion U(50, 300) km/s, spectra plus 100 Gaussian-filter AB magnitudes.

A step is one batch of the 16 384 models through the public C API's device entry
(opensps_context_compute_csp_batch_device: the inputs resident in HBM, the spectra and magnitudes left there).
`value` is the mixed-precision path's models/s (split-BF16 age sums on the matrix cores, FP32 chain, FP64 SFH and
zero points); the FP64 GPU path is timed beside it, and the two are compared on the whole batch. The CPU baseline is
the FP64 oracle (the restatement of the reference's CPU path, OpenMP) on a bounded sample of the same models.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "CSP models/s, 16k batch, mixed vs FP64 GPU at ≤1e-4 rel err; emulator models/s"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
FP64_PEAK_TFLOPS = 78.6
# FP64 operations per (model, common isochrone point) of k_op_weights: the 12-FMA series of expm1(y) / y (24), y = e dl
# (1), the chain's next power phi = plo + plo y sx (3), w = c dl plo sx (3), the mass and L_bol sums (4)
OP_WEIGHT_FLOPS = 35
# FP64 operations per (model, common point) of the operator's reverse (k_op_adj_points): the forward weight's 35, h's
# series alongside (24), dw/de (5), and F, S, D with the cotangent ge (6 FMAs: 12)
OP_ADJ_FLOPS = 76
KERNELS = ["prep", "group", "age_sums", "chain"]
MAG_RANGE = 40.0
CHAIN_KERNEL = "k_chain<float, true, false>"
TRAFFIC_REPORT = "k_chain_f32_parametric16k_traffic.md"


def cpu_model() -> str:
    """The host CPU's model name (the first 'model name' of /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def product_tables(grid, fset):
    """The bands and the cosmology table as the product's host forms them (libopensps.so: filters_build and the
    cosmology of a context with the reference's default H0, Om0, OL0), for the legs that drive libopensps_hip."""
    from libopensps_amd import api
    with api.Context() as c:
        c.set_ssp_basis(grid)
        c.set_filters(fset)
        return c.bands(), c.cosmology()


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="CPU work of the baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-operator", action="store_true", help="skip the IMF-per-model workload")
    ap.add_argument("--no-weights", action="store_true", help="skip the 256-model weight-matrix workload")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the batch-adjoint workload")
    ap.add_argument("--no-emulator", action="store_true", help="skip the emulator workload")
    ap.add_argument("--adjoint-sfh", action="store_true",
                    help="also the batch adjoint with SFH parameters (params_bar's SFH rows: the SFH tangents)")
    ap.add_argument("--emu-train", type=int, default=4_000_000, help="emulator training spectra")
    ap.add_argument("--emu-epochs", type=int, default=200)
    ap.add_argument("--copy-probe", action="store_true",
                    help="end with a 256 MiB device-to-device copy (the calibration of the HBM counter runs)")
    return ap.parse_args()


def operator_workload(args, n=4096):
    """The IMF-per-model workload: 4096 models on the SSP operator (107 ages x 22 Z x 1500 isochrone points,
    10 000 blackbody templates on the 5994-pixel grid), per-model Kroupa slopes U(0.3, 3.3) and upper limits
    U(60, 120), with the parametric SFH, dust, IGM, redshift and smoothing of the headline workload and its 100
    filters. Timed through libopensps_hip (the device layer) in both precisions; the oracle on a bounded sample."""
    import numpy as np
    import torch
    from libopensps_amd import hip as H, synth

    t0 = time.perf_counter()
    from libopensps_amd import api
    grid = synth.ssp_grid(seed=1, nspec=5994, nt=107, nz=22)
    op = synth.ssp_operator(grid.lam, grid.log_age, 22, npts=1500, ntpl=10000, seed=11)
    fset = synth.gaussian_filters(100, seed=7)
    prm = synth.imf_params(n, seed=13, p=synth.parametric_params(n, seed=5))
    # through the public API: the operator installed with opensps_context_set_ssp_ops, the batch on device-resident
    # arrays (opensps_context_compute_csp_batch_device), each model's IMF from its slots
    ctx = api.Context()
    ctx.set_ssp_basis(grid)
    ctx.set_filters(fset)
    ctx.set_int("add_dust_emission", 0)
    ctx.set_int("add_agn_dust", 0)
    ctx.set_int("add_igm_absorption", 1)
    ctx.set_int("profile_kernels", 1)
    ctx.set_int("imf_type", 2)
    ctx.set_ssp_ops(op)
    bands = ctx.bands()
    cz, cdl = ctx.cosmology()
    d_prm = torch.as_tensor(prm, device="cuda")
    setup = time.perf_counter() - t0
    res = {}
    outs = {}
    for prec, name in ((1, "mixed"), (0, "fp64")):
        dt_ = torch.float32 if prec == 1 else torch.float64
        out = {"mags": torch.empty((n, ctx.nbands), dtype=torch.float64, device="cuda"),
               "spec": torch.empty((n, grid.nspec), dtype=dt_, device="cuda"),
               "mass": torch.empty(n, dtype=torch.float64, device="cuda"),
               "lbol": torch.empty(n, dtype=torch.float64, device="cuda")}
        run = lambda: ctx.compute_csp_batch_device(d_prm, out["mags"], d_spec=out["spec"], d_mass=out["mass"],
                                                   d_lbol=out["lbol"], precision=prec)
        run()
        kernel_times(ctx, reset=1)
        steps = max(2, args.steps // 4)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            run()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t1
        kms, _ = kernel_times(ctx, reset=1)
        oms = op_kernel_times(ctx, reset=1)
        res[name] = {"models_per_s": n * steps / dt, "ms_per_step": 1e3 * dt / steps,
                     "kernel_ms": dict(zip(["prep", "imf_weights_fold_sums", "template_product", "chain",
                                            "op_imf_weights", "op_fold_sums"], [round(x, 4) for x in kms + oms]))}
        outs[name] = {k: v.clone() for k, v in out.items()}
    ref = outs["fp64"]["spec"]
    peak = ref.abs().amax(dim=1, keepdim=True)
    mask = ref.abs() >= 1e-6 * peak
    rel = torch.where(mask, (outs["mixed"]["spec"].double() - ref).abs() / torch.where(mask, ref.abs(), torch.ones_like(ref)),
                      torch.zeros_like(ref))
    m64, mmx = outs["fp64"]["mags"], outs["mixed"]["mags"]
    fin = torch.isfinite(m64) & torch.isfinite(mmx)
    bright = torch.where(fin, m64, torch.full_like(m64, float("inf"))).amin(dim=1, keepdim=True)
    inr = fin & (m64 - bright <= MAG_RANGE)
    dmag = torch.where(inr, (mmx - m64).abs(), torch.zeros_like(m64))
    # template product: useful flops 2 nspec ntpl per model column
    kmix = res["mixed"]["kernel_ms"]["template_product"]
    flops = 2.0 * grid.nspec * op.ntpl * n
    roofs = {"template_product": {
        "bound": "mfma", "kernel": "k_mm_bf16x3 (template product)", "achieved": flops / (kmix * 1e-3) / 1e12,
        "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": flops / (kmix * 1e-3) / 1e12 / BF16_PEAK_TFLOPS,
        "algorithmic_per_launch": flops, "ms": kmix}}
    # the IMF weights (k_op_imf, k_op_slow, k_op_weights): every model's points in its two metallicities (each
    # metallicity of the operator holds npts / nz points), OP_WEIGHT_FLOPS FP64 operations a point, and the weights
    # written in the path's precision (FP32 mixed, FP64) for the segments
    pts = 2.0 * op.npts / op.nz * n
    for name, wsize in (("mixed", 4), ("fp64", 8)):
        kw = res[name]["kernel_ms"]["op_imf_weights"]
        if kw <= 0:
            continue
        fl, by = OP_WEIGHT_FLOPS * pts, wsize * pts
        t_fl, t_by = fl / (FP64_PEAK_TFLOPS * 1e12), by / (HBM_PEAK_GBS * 1e9)
        bound = "fp64" if t_fl >= t_by else "hbm"
        ach = fl / (kw * 1e-3) / 1e12 if bound == "fp64" else by / (kw * 1e-3) / 1e9
        peak = FP64_PEAK_TFLOPS if bound == "fp64" else HBM_PEAK_GBS
        roofs["imf_weights_" + name] = {
            "bound": bound, "kernel": "k_op_imf + k_op_slow + k_op_weights", "achieved": ach, "peak": peak,
            "unit": "TFLOP/s" if bound == "fp64" else "GB/s", "frac": ach / peak, "ms": kw,
            "fp64_tflops": fl / (kw * 1e-3) / 1e12, "weights_write_gbs": by / (kw * 1e-3) / 1e9,
            "algorithmic_per_launch": {"fp64_flops": fl, "weight_bytes": by, "model_points": pts}}
    dom = max(("template_product", "imf_weights_mixed"), key=lambda k: roofs.get(k, {}).get("ms", 0.0))
    roof = dict(roofs[dom], phase=dom)
    result = {"workload": f"IMF per model: {n} models, 107 ages x 22 Z x 1500 points, 10000 templates, 5994 px, "
                          "Kroupa slopes U(0.3,3.3), upper U(60,120), parametric SFH, dust, IGM, z, LSF, 100 filters",
              "value": res["mixed"]["models_per_s"], "unit": "models/s", "mixed": res["mixed"], "fp64": res["fp64"],
              "speedup_vs_fp64": res["mixed"]["models_per_s"] / res["fp64"]["models_per_s"],
              "accuracy_vs_fp64_gpu": {"max_spec_rel_err": float(rel.max()), "max_mag_err": float(dmag.max()),
                                       "mass_rel_err": float(((outs["mixed"]["mass"] / outs["fp64"]["mass"]) - 1).abs().max())},
              "roofline": roof, "rooflines": roofs, "setup_s": setup}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        m1 = 16
        sub = np.ascontiguousarray(prm[:, :m1])
        t1 = time.perf_counter()
        r = O.csp_batch(grid, bands, sub, add_igm=1, cosmo=(cz, cdl), want=("mags", "spec"), ops=op, imf_type=2,
                        remnants=1)
        dt = time.perf_counter() - t1
        hm = outs["mixed"]["mags"][:m1].cpu().numpy()
        ok = np.isfinite(r["mags"]) & np.isfinite(hm)
        br = np.min(np.where(ok, r["mags"], np.inf), axis=1, keepdims=True)
        okr = ok & (r["mags"] - br <= MAG_RANGE)
        result["cpu_baseline"] = {"value": m1 / dt, "unit": "models/s", "cores": O.num_threads(), "kind": "port",
                                  "sample": f"the first {m1} models, FP64 oracle (OpenMP, {O.num_threads()} threads), "
                                            f"{dt:.1f} s",
                                  "mixed_vs_oracle_max_mag_err": float(np.max(np.abs(hm[okr] - r["mags"][okr])))}
    ctx.close()
    result["adjoint"] = operator_adjoint(args, grid, op, fset, prm, (cz, cdl), n)
    return result


def operator_adjoint(args, grid, op, fset, prm, cosmo, n):
    """The IMF-per-model workload's reverse pass: the same 4096 models and operator, the chi^2 of the 100-band
    magnitudes against a held-out seed; params_bar with the IMF rows (imf1..imf3) through the operator's reverse
    (the transposed template product, then the points' weights' derivatives). Timed through libopensps_hip in both
    precisions."""
    import numpy as np
    import torch
    from libopensps_amd import hip as H, synth

    bands, _ = product_tables(grid, fset)
    ctx = H.DeviceContext(0)
    ctx.set_grid(grid)
    ctx.set_bands(bands)
    ctx.set_cosmology(*cosmo)
    ctx.set_options(1, 1)
    ctx.set_ssp_ops(op)
    ctx.set_imf(2, 1)
    d_prm = torch.as_tensor(prm, device="cuda")
    f64 = ctx.batch(d_prm, precision=H.FP64, use_ops=1, want=("mags",))
    loss, dmag = synth.chi2_cotangent(f64["mags"].cpu().numpy(), seed=31, detect_range=30.0)
    d_mb = torch.as_tensor(dmag, device="cuda")
    lib = H.lib()
    lib.opensps_hip_set_profiling.argtypes = [H.C.c_void_p, H.C.c_int]
    lib.opensps_hip_grad_kernel_times.argtypes = [H.C.c_void_p, H.C.c_void_p, H.C.c_void_p, H.C.c_int]
    res, outs = {}, {}
    for prec, name in ((H.MIXED, "mixed"), (H.FP64, "fp64")):
        out = ctx.alloc_outputs(n, prec, want=("mags",))
        grads = {"params_bar": torch.zeros((n, H.NBGRAD), dtype=torch.float64, device="cuda")}
        lib.opensps_hip_set_profiling(ctx._ctx, 1)
        ctx.launch_grad(d_prm, out, {"mags": d_mb}, grads, prec, use_ops=1)
        ctx.sync()
        lib.opensps_hip_grad_kernel_times(ctx._ctx, None, None, 1)
        steps = max(2, args.steps // 4)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            ctx.launch_grad(d_prm, out, {"mags": d_mb}, grads, prec, use_ops=1)
            ctx.sync()
        dt = time.perf_counter() - t1
        gms, gcnt = (H.C.c_double * 3)(), (H.C.c_int * 3)()
        lib.opensps_hip_grad_kernel_times(ctx._ctx, gms, gcnt, 1)
        gk = [gms[i] / max(gcnt[i], 1) for i in range(3)]
        res[name] = {"models_per_s": n * steps / dt, "ms_per_step": 1e3 * dt / steps,
                     "kernel_ms": dict(zip(["rev_chain", "rev_template_product", "rev_operator_out"],
                                           [round(x, 4) for x in gk]))}
        outs[name] = grads["params_bar"].cpu().numpy()
    lib.opensps_hip_set_profiling(ctx._ctx, 0)
    from libopensps_amd import api
    row_of = {api.grad_param_name(i): i for i in range(api.nbgrad())}
    acc = {}
    for k in ("imf1", "imf2", "imf3", "logzsol", "dust2", "zred"):
        a, b = outs["mixed"][:, row_of[k]], outs["fp64"][:, row_of[k]]
        acc[k] = float(np.max(np.abs(a - b)) / (np.max(np.abs(b)) + 1e-300))
    # the transposed template product: 2 ntpl nspec flops per (model, part) column, both parts
    flops = 2.0 * op.ntpl * grid.nspec * 2 * n
    tk = res["mixed"]["kernel_ms"]["rev_template_product"]
    roofs = {"rev_template_product": {
        "bound": "mfma", "kernel": "k_mm_bf16x3 (w_bar = templates^T y_bar)", "achieved": flops / (tk * 1e-3) / 1e12,
        "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": flops / (tk * 1e-3) / 1e12 / BF16_PEAK_TFLOPS,
        "algorithmic_per_launch": flops, "ms": tk}}
    # the operator's reverse (k_op_adj_t, k_op_imf_grad, k_op_adj_points, k_op_adj_slow, k_op_adj_out), dominated by
    # k_op_adj_points: per (model, point of its two metallicities) the forward weight's OP_WEIGHT_FLOPS, h's series
    # (24), dw/de (5) and the sums F, S, D with the cotangent ge (12): OP_ADJ_FLOPS FP64 operations; and the
    # template cotangents gathered, two per point (nent) in the path's precision
    to = res["mixed"]["kernel_ms"]["rev_operator_out"]
    pts = 2.0 * op.npts / op.nz * n
    fl, by = OP_ADJ_FLOPS * pts, 4.0 * op.nent * pts
    roofs["rev_operator_out"] = {
        "bound": "fp64", "kernel": "k_op_adj_points (+ k_op_adj_t, k_op_imf_grad, k_op_adj_slow, k_op_adj_out)",
        "achieved": fl / (to * 1e-3) / 1e12, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
        "frac": fl / (to * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, "gather_gbs": by / (to * 1e-3) / 1e9,
        "algorithmic_per_launch": {"fp64_flops": fl, "gather_bytes": by, "model_points": pts}, "ms": to}
    dom = max(roofs, key=lambda k: roofs[k]["ms"])
    roof = dict(roofs[dom], phase=dom)
    ctx.close()
    return {"workload": f"IMF per model, reverse: {n} models, the operator's 10000 templates, chi^2 of 100 filters; "
                        "forward + reverse, params_bar with imf1..imf3",
            "value": res["mixed"]["models_per_s"], "unit": "models/s (forward + gradient)", "mixed": res["mixed"],
            "fp64": res["fp64"], "speedup_vs_fp64": res["mixed"]["models_per_s"] / res["fp64"]["models_per_s"],
            "accuracy_vs_fp64_gpu": acc, "loss": loss, "roofline": roof, "rooflines": roofs}


def weight_matrix_workload(args, n=256):
    """The plain weight-matrix workload: 256 models, Dirichlet(0.5) weights over 107 ages (1 M_sun formed), one
    metallicity, 5994 px, no dust, z = 0, spectra and 20 Gaussian-filter magnitudes. Timed through the public C API's
    host-array entry (what a reference caller calls: opensps_context_compute_csp_batch with ssp_weights, spectra and
    magnitudes copied to the host) and its device-resident entry, in FP64 (the plain GPU path) and mixed precision;
    the FP64 oracle on the whole batch. At 256 models the batch is bound by launch latency and transfers."""
    import numpy as np
    import torch
    from libopensps_amd import api, synth

    grid = synth.ssp_grid(seed=1, nspec=5994, nt=107, nz=1)
    fset = synth.gaussian_filters(20, seed=7)
    w = synth.dirichlet_weights(n, grid.nt, seed=21)
    ctx = api.Context()
    ctx.set_ssp_basis(grid)
    ctx.set_filters(fset)
    ctx.set_int("add_dust_emission", 0)
    ctx.set_int("add_agn_dust", 0)
    ctx.set_int("add_igm_absorption", 0)
    ctx.set_int("profile_kernels", 1)
    d_prm = torch.as_tensor(synth.default_params(n), device="cuda")
    d_w = torch.as_tensor(w, device="cuda")
    d_mags = torch.empty((n, ctx.nbands), dtype=torch.float64, device="cuda")
    res = {}
    steps = max(10, args.steps)
    for prec, name in ((0, "fp64"), (1, "mixed")):
        ctx.set_str("kernel_precision", "double" if prec == 0 else "mixed")
        d_spec = torch.empty((n, grid.nspec), dtype=torch.float64 if prec == 0 else torch.float32, device="cuda")
        host = lambda: ctx.compute_csp_batch({}, n=n, ssp_weights=w, want=("mags", "spec"))
        dev = lambda: ctx.compute_csp_batch_device(d_prm, d_mags, d_spec=d_spec, d_weights=d_w, precision=prec)
        r = host()
        dev()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            host()
        dt_h = (time.perf_counter() - t1) / steps
        kernel_times(ctx, reset=1)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            dev()
        torch.cuda.synchronize()
        dt_d = (time.perf_counter() - t1) / steps
        kms, _ = kernel_times(ctx, reset=1)
        res[name] = {"host_api_models_per_s": n / dt_h, "host_api_ms": 1e3 * dt_h,
                     "device_models_per_s": n / dt_d, "device_ms": 1e3 * dt_d,
                     "kernel_ms": dict(zip(KERNELS, [round(x, 4) for x in kms]))}
        res[name + "_mags"] = r["mags"]
    out = {"workload": f"weight matrix: {n} models, Dirichlet(0.5) weights over 107 ages, 1 Z, 5994 px, no dust, "
                       "z = 0, 20 Gaussian filters",
           "value": res["fp64"]["host_api_models_per_s"], "unit": "models/s (FP64, public host-array call)",
           "fp64": res["fp64"], "mixed": res["mixed"],
           "mixed_vs_fp64_max_mag_err": float(np.nanmax(np.abs(res["mixed_mags"] - res["fp64_mags"])))}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        bands, cosmo = ctx.bands(), ctx.cosmology()
        reps = 5
        t1 = time.perf_counter()
        for _ in range(reps):
            ref = O.csp_batch(grid, bands, synth.default_params(n), ssp_weights=w, add_igm=0, cosmo=cosmo,
                              want=("mags", "spec"))
        dt = (time.perf_counter() - t1) / reps
        out["cpu_baseline"] = {"value": n / dt, "unit": "models/s", "cores": O.num_threads(), "kind": "port",
                               "cpu": cpu_model(),
                               "sample": f"the whole batch ({n} models) x {reps}, FP64 oracle (OpenMP, "
                                         f"{O.num_threads()} threads), {reps * dt:.2f} s",
                               "fp64_vs_oracle_max_mag_err": float(np.nanmax(np.abs(res["fp64_mags"] - ref["mags"])))}
    for k in ("fp64_mags", "mixed_mags"):
        res.pop(k)
    ctx.close()
    return out


def op_kernel_times(ctx, reset=1):
    """Per-batch device milliseconds of the SSP operator's two phases [IMF weights, ages' sums + segments + fold +
    mass sums] (opensps_context_op_kernel_times)."""
    import ctypes as C
    from libopensps_amd import api
    ms = (C.c_double * 2)()
    cnt = (C.c_int * 2)()
    st = C.c_int()
    api.lib().opensps_context_op_kernel_times.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                                          C.POINTER(C.c_int)]
    api.lib().opensps_context_op_kernel_times(ctx.handle, ms, cnt, reset, C.byref(st))
    return [ms[i] / max(cnt[i], 1) for i in range(2)]


def kernel_times(ctx, reset=1):
    """Per-launch device milliseconds of the batch's four kernel groups (opensps_context_kernel_times)."""
    import ctypes as C
    from libopensps_amd import api
    ms = (C.c_double * 4)()
    cnt = (C.c_int * 4)()
    st = C.c_int()
    api.lib().opensps_context_kernel_times.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int)]
    api.lib().opensps_context_kernel_times(ctx.handle, ms, cnt, reset, C.byref(st))
    return [ms[i] / max(cnt[i], 1) for i in range(4)], [cnt[i] for i in range(4)]


def adjoint_workload(args, n=8192, sfh=False):
    """The batch-adjoint workload: 8192 models on the 107 ages x 22 Z x 5994-pixel grid with weights over ages and
    metallicities (a log-normal SFH of 1 dex times a 22-bin enrichment history), dust, redshift U(0, 4) with IGM and
    smoothing U(50, 300) km/s, and the chi^2 loss of 100-band photometry against a held-out seed with 5 % errors.
    A step is the forward batch and its reverse pass (params_bar and weights_z_bar), timed through libopensps_hip
    in both precisions; the oracle's reverse pass on a bounded sample. sfh: the SFH parameters instead of the
    weights (half delayed tau, half Simha with truncation and a slope; the per-model metallicity blend), params_bar
    with the SFH rows."""
    import numpy as np
    import torch
    from libopensps_amd import hip as H, synth

    t0 = time.perf_counter()
    grid = synth.ssp_grid(seed=1, nspec=5994, nt=107, nz=22)
    fset = synth.gaussian_filters(100, seed=7)
    prm = synth.adjoint_params(n, seed=19)
    wz = synth.lognormal_weights_z(n, grid.log_age, grid.logz, seed=17)
    if sfh:
        from libopensps_amd.params import SLOT
        pp = synth.parametric_params(n, seed=19)
        for k in ("sfh", "tage", "tau", "logzsol"):
            prm[SLOT[k]] = pp[SLOT[k]]
        rng = np.random.default_rng(29)
        simha = np.arange(n) % 2 == 1
        prm[SLOT["sfh"]] = np.where(simha, 5.0, 4.0)
        prm[SLOT["sf_trunc"]] = np.where(simha, prm[SLOT["tage"]] * rng.uniform(0.3, 0.9, n), 0.0)
        prm[SLOT["sf_slope"]] = np.where(simha, rng.uniform(-1.0, 1.0, n), 0.0)
        wz = None
    bands, (cz, cdl) = product_tables(grid, fset)
    ctx = H.DeviceContext(0)
    ctx.set_grid(grid)
    ctx.set_bands(bands)
    ctx.set_cosmology(cz, cdl)
    ctx.set_options(1, 1)
    d_prm = torch.as_tensor(prm, device="cuda")
    d_wz = torch.as_tensor(wz, device="cuda") if wz is not None else None
    f64 = ctx.batch(d_prm, precision=H.FP64, ssp_weights_z=d_wz, want=("mags",))
    loss, dmag = synth.chi2_cotangent(f64["mags"].cpu().numpy(), seed=23)
    d_mb = torch.as_tensor(dmag, device="cuda")
    setup = time.perf_counter() - t0
    lib = H.lib()
    lib.opensps_hip_set_profiling.argtypes = [H.C.c_void_p, H.C.c_int]
    lib.opensps_hip_kernel_times.argtypes = [H.C.c_void_p, H.C.c_void_p, H.C.c_void_p, H.C.c_int]
    lib.opensps_hip_grad_kernel_times.argtypes = [H.C.c_void_p, H.C.c_void_p, H.C.c_void_p, H.C.c_int]
    res, outs = {}, {}
    for prec, name in ((H.MIXED, "mixed"), (H.FP64, "fp64")):
        out = ctx.alloc_outputs(n, prec, want=("mags",))
        grads = {"params_bar": torch.zeros((n, H.NBGRAD), dtype=torch.float64, device="cuda")}
        if wz is not None:
            grads["weights_z_bar"] = torch.zeros((n, grid.nz, grid.nt), dtype=torch.float64, device="cuda")
        lib.opensps_hip_set_profiling(ctx._ctx, 1)
        ctx.launch_grad(d_prm, out, {"mags": d_mb}, grads, prec, ssp_weights_z=d_wz)
        ctx.sync()
        for f in (lib.opensps_hip_kernel_times, lib.opensps_hip_grad_kernel_times):
            f(ctx._ctx, None, None, 1)
        steps = max(2, args.steps // 4)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            ctx.launch_grad(d_prm, out, {"mags": d_mb}, grads, prec, ssp_weights_z=d_wz)
            ctx.sync()
        dt = time.perf_counter() - t1
        ms, cnt = (H.C.c_double * 4)(), (H.C.c_int * 4)()
        gms, gcnt = (H.C.c_double * 3)(), (H.C.c_int * 3)()
        lib.opensps_hip_kernel_times(ctx._ctx, ms, cnt, 1)
        lib.opensps_hip_grad_kernel_times(ctx._ctx, gms, gcnt, 1)
        fk = [ms[i] / max(cnt[i], 1) for i in range(4)]
        gk = [gms[i] / max(gcnt[i], 1) for i in range(3)]
        res[name] = {"models_per_s": n * steps / dt, "ms_per_step": 1e3 * dt / steps,
                     "kernel_ms": dict(zip(["fwd_prep", "fwd_group", "fwd_product", "fwd_chain", "rev_chain",
                                            "rev_product", "rev_out"], [round(x, 4) for x in fk + gk]))}
        outs[name] = {k: v.detach().cpu().numpy() for k, v in grads.items()}
    lib.opensps_hip_set_profiling(ctx._ctx, 0)
    names = ("dust2", "dust_index", "zred", "sigma_smooth", "igm_factor")
    if sfh:
        names = names + ("logzsol", "tage", "tau", "sf_trunc", "sf_slope")
    from libopensps_amd import api
    row_of = {api.grad_param_name(i): i for i in range(api.nbgrad())}
    rows = {k: row_of[k] for k in names}
    acc = {}
    for k, r in rows.items():
        a, b = outs["mixed"]["params_bar"][:, r], outs["fp64"]["params_bar"][:, r]
        acc[k] = float(np.max(np.abs(a - b)) / (np.max(np.abs(b)) + 1e-300))
    if wz is not None:
        wa, wb = outs["mixed"]["weights_z_bar"].reshape(n, -1), outs["fp64"]["weights_z_bar"].reshape(n, -1)
        acc["weights_z_bar"] = float(np.max(np.abs(wa - wb) / (np.max(np.abs(wb), axis=1, keepdims=True) + 1e-300)))
    # the transposed product: 2 kfull nsp flops per (model, part) column, both parts
    kfull, nsp = grid.nt * grid.nz, ((grid.nspec + 63) // 64) * 64
    flops = 2.0 * kfull * grid.nspec * 2 * n
    tk = res["mixed"]["kernel_ms"]["rev_product"]
    roofs = {"rev_product": {
        "bound": "mfma", "kernel": "k_mm_bf16x3 (w_bar = basis^T y_bar)", "achieved": flops / (tk * 1e-3) / 1e12,
        "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": flops / (tk * 1e-3) / 1e12 / BF16_PEAK_TFLOPS,
        "algorithmic_per_launch": flops, "ms": tk}}
    # the reverse chain (k_grad_chain<float>): per model it reads the forward's two parts' age sums (FP32, nspec
    # each), the magnitudes, band fluxes and their cotangents (3 nbands doubles) and the parameters (33 doubles), and
    # writes the two parts' cotangents y_bar as split BF16 (2 x 2 bytes a pixel) and 16 per-model partials
    nb = bands.shape[0]
    tc = res["mixed"]["kernel_ms"]["rev_chain"]
    byts = n * (2 * grid.nspec * 4 + 2 * grid.nspec * 4 + 3 * nb * 8 + 33 * 8 + 16 * 8)
    roofs["rev_chain"] = {"bound": "hbm", "kernel": "k_grad_chain<float>", "achieved": byts / (tc * 1e-3) / 1e9,
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": byts / (tc * 1e-3) / 1e9 / HBM_PEAK_GBS,
                          "algorithmic_per_launch": byts, "ms": tc}
    dom = max(roofs, key=lambda k: roofs[k]["ms"])
    roof = dict(roofs[dom], phase=dom)
    what = ("SFH parameters (delayed tau; Simha with truncation and slope), logzsol blend over 22 Z" if sfh else
            "weights over 107 ages x 22 Z (log-normal SFH x enrichment history)")
    result = {"workload": f"batch adjoint: {n} models, {what}, 5994 px, dust, IGM, z U(0,4), LSF, chi^2 of 100 "
                          "filters; forward + reverse",
              "value": res["mixed"]["models_per_s"], "unit": "models/s (forward + gradient)", "mixed": res["mixed"],
              "fp64": res["fp64"], "speedup_vs_fp64": res["mixed"]["models_per_s"] / res["fp64"]["models_per_s"],
              "accuracy_vs_fp64_gpu": acc, "loss": loss, "roofline": roof, "rooflines": roofs, "setup_s": setup}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        m1 = 8
        t1 = time.perf_counter()
        r = O.csp_batch_grad(grid, bands, np.ascontiguousarray(prm[:, :m1]),
                             ssp_weights_z=wz[:m1] if wz is not None else None,
                             mags_bar=dmag[:m1], cosmo=(cz, cdl))
        dt = time.perf_counter() - t1
        ga, gb = outs["mixed"]["params_bar"][:m1], r["params_bar"]
        result["cpu_baseline"] = {"value": m1 / dt, "unit": "models/s (forward + gradient)", "cores": O.num_threads(),
                                  "kind": "port",
                                  "sample": f"the first {m1} models, FP64 oracle reverse pass (OpenMP, "
                                            f"{O.num_threads()} threads), {dt:.1f} s",
                                  "mixed_vs_oracle_zred_rel": float(np.max(np.abs(ga[:, 22] - gb[:, 22])) /
                                                                     (np.max(np.abs(gb[:, 22])) + 1e-300))}
    ctx.close()
    return result


def emulator_workload(args, csp_models_per_s):
    """The learned emulator on the full input set: the parametric workload's 7 parameters, the 3 Kroupa slopes and
    the dust emission / AGN parameters (duste_umin, duste_qpah, duste_gamma, fagn) -> 4 x 512 GELU -> 128 PCA + 100
    magnitudes, trained on 4M spectra that the batch generates on the SSP operator (the IMF per model, 22 Z x 107
    ages x 1500 points, 10 000 templates) with the Draine & Li and torus stages (synthetic templates of the tables'
    layout), Adam with cosine decay, on the project's own forward / backward / Adam kernels. Reports the training
    time, the validation error against the FP64 batch (16k models), the inference rate on 1M models (the fused
    kernel) against the mixed batch's (the headline batch, and the batch that computes these inputs), and the serving
    path's fallback rate with a 1 % spot-check at 0.01 mag."""
    import numpy as np
    import torch
    from libopensps_amd import emu_train as T, hip as H, synth

    t0 = time.perf_counter()
    grid = synth.ssp_grid(seed=1, nspec=5994, nt=107, nz=22)
    fset = synth.gaussian_filters(100, seed=7)
    bands, (cz, cdl) = product_tables(grid, fset)
    sp = T.FULL
    dctx = H.DeviceContext(0)
    dctx.set_grid(grid)
    dctx.set_bands(bands)
    dctx.set_cosmology(cz, cdl)
    dctx.set_options(1, 1)
    dctx.set_ssp_ops(synth.ssp_operator(grid.lam, grid.log_age, 22, npts=1500, ntpl=10000, seed=11))
    dctx.set_imf(2, 1)
    dctx.set_emission(1, 1)
    dctx.set_dust_emission(synth.DL07_QPAH, synth.DL07_UMIN, synth.dust_emission_templates(grid.lam))
    dctx.set_agn(synth.AGN_TAU, synth.agn_templates(grid.lam))
    t_setup = time.perf_counter() - t0
    emu = T.Emulator.train(dctx, (cz, cdl), n_train=args.emu_train, n_pca_fit=200_000, n_pca=128,
                           epochs=args.emu_epochs, batch=8192, lr=2e-3, hidden=512, layers=4, space=sp)
    t_total = time.perf_counter() - t0
    val = T.validate(emu, dctx, n=16384)
    # the mixed batch on these inputs (the operator with the emission stages): the rate the emulator replaces
    nb_ = 16384
    xb = torch.as_tensor(sp.sample(nb_, 778), device="cuda")
    prm_b = sp.to_prm_device(xb)
    out_b = dctx.alloc_outputs(nb_, H.MIXED, want=("mags",))
    dctx.launch(prm_b, out_b, H.MIXED, use_ops=1)
    dctx.sync()
    t1 = time.perf_counter()
    for _ in range(3):
        dctx.launch(prm_b, out_b, H.MIXED, use_ops=1)
        dctx.sync()
    batch_rate = 3 * nb_ / (time.perf_counter() - t1)
    # inference on 1M models
    n = 1 << 20
    x = sp.sample(n, 777)
    fd, dd = emu.inputs(x)
    emu.predict_raw(fd)
    torch.cuda.synchronize()
    reps = 5
    t1 = time.perf_counter()
    for _ in range(reps):
        emu.mlp.predict(fd, sync=False)
    emu.mlp.sync()
    dt = (time.perf_counter() - t1) / reps
    rate = n / dt
    flops = 2.0 * (sp.n_feat * 512 + 3 * 512 * 512 + 512 * (128 + 100)) * n
    # the serving path: 16 batches of 65536 with a 1 % spot-check at 0.01 mag
    emu.stats = {k: 0 for k in emu.stats}
    serve_x = [sp.sample(65536, 2000 + b) for b in range(16)]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for b in range(16):
        emu.predict(serve_x[b], spot=0.01, tol=0.01, seed=b)
    torch.cuda.synchronize()
    serve_rate = 16 * 65536 / (time.perf_counter() - t1)
    info = emu.train_info
    out = {"workload": f"emulator: {len(sp.names)} parameters ({', '.join(sp.names)}) -> 4 x 512 GELU -> 128 PCA "
                       "coefficients of the log-spectrum + 100 magnitudes; trained on "
                       f"{info['n_train']} spectra of the batch on the SSP operator with dust emission and the AGN "
                       f"torus (PCA from {info['n_pca_fit']})",
           "value": rate, "unit": "models/s (inference, 1M batch)", "vs_csp_mixed": rate / csp_models_per_s,
           "batch_same_inputs_models_per_s": batch_rate, "vs_batch_same_inputs": rate / batch_rate,
           "inference_ms_per_1M": 1e3 * dt * (1 << 20) / n,
           "roofline": {"bound": "mfma", "kernel": "k_emu_fused<512, 16>", "achieved": flops / dt / 1e12,
                        "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": flops / dt / 1e12 / BF16_PEAK_TFLOPS},
           "validation_vs_fp64_batch": val,
           "fallback": {"rate": emu.fallback_rate(), "spot": 0.01, "tol_mag": 0.01, **emu.stats},
           "serving_models_per_s": serve_rate,
           "serving": "16 batches of 65536 models: inputs from the host, the emulator, a 1 % spot-check against the "
                      "mixed batch and, where it fails, the whole batch recomputed by the batch (fallback included)",
           "training": {k: v for k, v in info.items() if k != "loss_history"},
           "loss_first_last": [info["loss_history"][0], info["loss_history"][-1]], "total_s": t_total,
           "setup_s": t_setup}
    emu.mlp.close()
    dctx.close()
    return out


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: the batched CSP path runs on one GPU only (on several, each GPU would take a contiguous "
              "slice of the models with its own copy of the basis and filters; no collective). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from libopensps_amd import api, synth

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    n = args.n
    t_setup = time.perf_counter()
    grid = synth.ssp_grid(seed=1, nspec=5994, nt=107, nz=22)
    fset = synth.gaussian_filters(100, seed=7)
    prm = synth.parametric_params(n, seed=5)
    ctx = api.Context()
    ctx.set_ssp_basis(grid)
    ctx.set_filters(fset)
    ctx.set_int("add_dust_emission", 0)
    ctx.set_int("add_agn_dust", 0)
    ctx.set_int("add_igm_absorption", 1)
    ctx.set_int("profile_kernels", 1)
    nb, ns = ctx.nbands, grid.nspec
    d_prm = torch.as_tensor(prm, device="cuda")
    d_mags = torch.empty((n, nb), dtype=torch.float64, device="cuda")
    d_mass = torch.empty(n, dtype=torch.float64, device="cuda")
    d_lbol = torch.empty(n, dtype=torch.float64, device="cuda")
    d_spec32 = torch.empty((n, ns), dtype=torch.float32, device="cuda")
    d_spec64 = torch.empty((n, ns), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup

    def kernel_times(reset=1):
        import ctypes as C
        ms = (C.c_double * 4)()
        cnt = (C.c_int * 4)()
        st = C.c_int()
        api.lib().opensps_context_kernel_times.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                                           C.POINTER(C.c_int)]
        api.lib().opensps_context_kernel_times(ctx.handle, ms, cnt, reset, C.byref(st))
        return [ms[i] for i in range(4)], [cnt[i] for i in range(4)]

    def timed(precision, d_spec, steps, warmup):
        for _ in range(warmup):
            ctx.compute_csp_batch_device(d_prm, d_mags, d_spec=d_spec, d_mass=d_mass, d_lbol=d_lbol,
                                         precision=precision)
        kernel_times(reset=1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.compute_csp_batch_device(d_prm, d_mags, d_spec=d_spec, d_mass=d_mass, d_lbol=d_lbol,
                                         precision=precision)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        ms, cnt = kernel_times(reset=1)
        return dt, [m / max(c, 1) for m, c in zip(ms, cnt)]

    # mixed precision: the headline
    dt_mix, kms_mix = timed(1, d_spec32, args.steps, args.warmup)
    mags_mix = d_mags.clone()
    # the same batch through the drop-in host-array call (names[] / values[] from host memory, magnitudes back to the
    # host, no spectra requested: spec = NULL), as a reference caller makes it
    ctx.set_str("kernel_precision", "mixed")
    names = synth.varied_names()
    ctx.compute_csp_batch_params(prm, names, want=("mags",))
    hsteps = max(3, args.steps // 4)
    t0 = time.perf_counter()
    for _ in range(hsteps):
        ctx.compute_csp_batch_params(prm, names, want=("mags",))
    dt_host = (time.perf_counter() - t0) / hsteps
    ctx.set_str("kernel_precision", "double")
    kernel_times(reset=1)
    value = n * args.steps / dt_mix
    # FP64 GPU path
    steps64 = max(2, args.steps // 2)
    dt_64, kms_64 = timed(0, d_spec64, steps64, 1)
    value64 = n * steps64 / dt_64
    mags_64 = d_mags.clone()

    # accuracy of the mixed path against the FP64 GPU path over the whole batch (pixels >= 1e-6 of each peak)
    ref = d_spec64
    peak = ref.abs().amax(dim=1, keepdim=True)
    mask = ref.abs() >= 1e-6 * peak
    rel = torch.where(mask, (d_spec32.double() - ref).abs() / torch.where(mask, ref.abs(), torch.ones_like(ref)),
                      torch.zeros_like(ref))
    spec_err = float(rel.max())
    fin = torch.isfinite(mags_64) & torch.isfinite(mags_mix)
    dmag = torch.where(fin, (mags_mix - mags_64).abs(), torch.zeros_like(mags_64))
    mag_err_all = float(dmag.max())
    # bands within MAG_RANGE of the model's brightest band (flux ratio 1e-16): fainter bands sit in the FP32
    # chain's subnormal range (IGM- and dust-extinguished ultraviolet at z ~ 4) and are reported apart
    bright = torch.where(fin, mags_64, torch.full_like(mags_64, float("inf"))).amin(dim=1, keepdim=True)
    inrange = fin & (mags_64 - bright <= MAG_RANGE)
    mag_err = float(torch.where(inrange, dmag, torch.zeros_like(dmag)).max())
    n_faint = int((fin & ~inrange).sum())
    nan_mismatch = int((torch.isfinite(mags_64) != torch.isfinite(mags_mix)).sum())

    # roofline of the dominant kernel of the mixed path
    nt, kpair = grid.nt, 2 * grid.nt
    dom = int(np.argmax(kms_mix))
    name = KERNELS[dom]
    if name == "age_sums":
        flops = 2.0 * ns * kpair * n            # useful MACs (the split-BF16 form issues 3 bf16 products each)
        achieved = flops / (kms_mix[dom] * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / BF16_PEAK_TFLOPS, "traffic": None, "kernel": "k_mm_bf16x3",
                "algorithmic_per_launch": flops, "issued_bf16_tflops": 3 * achieved}
    else:
        # chain: read the age sums (float, nspec per model), write the spectrum (float) and the magnitudes (double),
        # read the parameters (28 doubles)
        byts = n * (ns * 4 + ns * 4 + nb * 8 + 28 * 8)
        achieved = byts / (kms_mix[dom] * 1e-3) / 1e9
        kname = CHAIN_KERNEL if name == "chain" else name
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": kname, "algorithmic_per_launch": byts}
        from libopensps_amd.profiling import read_traffic
        tr = read_traffic(os.path.join(ROOT, "profiles", TRAFFIC_REPORT))
        if tr and kname in tr[0]:
            roof["traffic"] = tr[1]
            roof["traffic_source"] = "profiles/" + TRAFFIC_REPORT
    roof["kernel_ms"] = dict(zip(KERNELS, [round(x, 4) for x in kms_mix]))
    roof["kernel_ms_fp64"] = dict(zip(KERNELS, [round(x, 4) for x in kms_64]))

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        bands = ctx.bands()
        cosmo = O.cosmology_table(72.0, 0.27, 0.73)
        # a bounded sample: the first m1 models of the batch, repeated until about cpu_seconds of CPU work
        m0 = min(256, n)
        t0 = time.perf_counter()
        O.csp_batch(grid, bands, prm[:, :m0], add_igm=1, cosmo=cosmo, want=("mags",))
        rate = m0 / (time.perf_counter() - t0)
        m1 = int(min(n, max(m0, rate * args.cpu_seconds)))
        reps = int(max(1, min(50, round(rate * args.cpu_seconds / m1))))
        sub = np.ascontiguousarray(prm[:, :m1])
        t0 = time.perf_counter()
        for _ in range(reps):
            r = O.csp_batch(grid, bands, sub, add_igm=1, cosmo=cosmo, want=("mags", "spec"))
        dt = time.perf_counter() - t0
        hm = mags_mix[:m1].cpu().numpy()
        h64 = mags_64[:m1].cpu().numpy()
        ok = np.isfinite(r["mags"]) & np.isfinite(hm)
        br = np.min(np.where(ok, r["mags"], np.inf), axis=1, keepdims=True)
        ok_r = ok & (r["mags"] - br <= MAG_RANGE)
        cpu = {"value": m1 * reps / dt, "unit": "models/s", "cores": O.num_threads(), "kind": "port", "cpu": cpu_model(),
               "sample": f"the first {m1} models of the same batch x {reps} passes, FP64 oracle (csp_oracle.c, "
                         f"OpenMP, {O.num_threads()} threads), {dt:.1f} s",
               "mixed_vs_oracle_max_mag_err": float(np.max(np.abs(hm[ok_r] - r["mags"][ok_r]))) if ok_r.any() else None,
               "fp64_vs_oracle_max_mag_err": float(np.max(np.abs(h64[ok_r] - r["mags"][ok_r]))) if ok_r.any() else None}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "models/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt_mix / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3/f32 (f64 SFH, zero points)",
        "data": "synthetic",
        "config": {"workload": "parametric CSP, 16384 models, 107 ages x 22 Z x 5994 px, delayed-tau, logzsol blend, "
                               "power-law dust, IGM, z U(0,4), LSF U(50,300) km/s, 100 Gaussian filters",
                   "n_models": n, "nspec": ns, "nt": nt, "nz": grid.nz, "nbands": nb, "seeds": [1, 7, 5]},
        "host_api": {"models_per_s": n / dt_host, "ms_per_batch": 1e3 * dt_host,
                     "call": "opensps_context_compute_csp_batch, mixed, host names/values in, host magnitudes out, "
                             "spec NULL (parameter upload and magnitude download included)"},
        "fp64_value": value64,
        "fp64_ms_per_step": 1e3 * dt_64 / steps64,
        "speedup_vs_fp64": value / value64,
        "accuracy_vs_fp64_gpu": {"max_spec_rel_err": spec_err, "spec_pixels": "flux >= 1e-6 of the model's peak",
                                 "max_mag_err": mag_err, "mag_bands": f"within {MAG_RANGE:g} mag of the model's "
                                                                      f"brightest band",
                                 "max_mag_err_all_bands": mag_err_all, "bands_fainter": n_faint,
                                 "nan_mismatch": nan_mismatch},
        "roofline": roof,
        "cpu_baseline": cpu,
        "setup_s": t_setup,
    }
    ctx.close()
    if not args.no_weights:
        out["weight_matrix"] = weight_matrix_workload(args)
    if not args.no_operator:
        out["operator"] = operator_workload(args)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_workload(args)
    if args.adjoint_sfh:
        out["adjoint_sfh"] = adjoint_workload(args, sfh=True)
    if not args.no_emulator:
        out["emulator"] = emulator_workload(args, value)
    print(json.dumps(out))
    if args.copy_probe:
        x = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(2):
            y.copy_(x)
        torch.cuda.synchronize()
    ctx.close()


if __name__ == "__main__":
    main()
