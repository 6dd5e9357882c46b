/*
 * opensps_hip.h -- the C ABI of libopensps_hip.so: the MI355X (gfx950) kernels of OpenSPS's batched
 * composite-population path, behind plain pointers and sizes.
 *
 * This is the layer the reference's Fortran batch driver reaches through `interface ... bind(C,
 * name = "opensps_hip_*")` blocks with `type(c_ptr), value` device pointers
 * (src/physics/opensps_csp_batch.f90:260-295, 2341-2349, 3742-3757;
 * src/spectra/opensps_batch_photometry.f90:579-592, 916-930, 1182-1217). The reference's launchers
 * are one per stage and run on the default stream with file-static scratch
 * (src/accel/hip/opensps_hip_kernels.hip:2884-2885, 2988-2991, 3522-3525); here one context owns
 * its resident tables, its stream and its scratch, and one call runs the whole batch on the device.
 *
 * Conventions: arrays are column-major with the quantity fastest (as the reference's C API,
 * doc/C_API.md:484-496). Pointers documented "device" are HIP device addresses; "host" ones are
 * ordinary host memory. Every function returns a status: 0 success, 1 bad context or not set up,
 * 2 unsupported option, 3 bad argument or shape, 10 a HIP runtime error (opensps_hip_last_error
 * gives the message).
 */
#ifndef OPENSPS_HIP_H
#define OPENSPS_HIP_H

#include <stddef.h>

#include "opensps_params.h"

#ifdef __cplusplus
extern "C" {
#endif

#define OPENSPS_HIP_FP64 0   /* FP64 age sums (f64 matrix cores) and FP64 chain */
#define OPENSPS_HIP_MIXED 1  /* split-BF16 age sums (3 bf16 products, FP32 accumulation), FP32 chain */

typedef struct OpsCtx opensps_hip_ctx_t;

/* One batch on device-resident data (the arguments of compute_csp_batch, opensps_csp_batch.f90:431-457,
 * for the outputs this path computes). Outputs may be NULL except mags. spec, young and old hold
 * float values with OPENSPS_HIP_MIXED and double values with OPENSPS_HIP_FP64. */
typedef struct {
    int n;                       /* models */
    int precision;               /* OPENSPS_HIP_FP64 or OPENSPS_HIP_MIXED */
    int split_parts;             /* 1: young and old age sums apart (needed by a birth-cloud or old-star
                                    screen, or for young / old); 0: one sum per model */
    int use_ops;                 /* 1: the age sums from the SSP operator with each model's IMF
                                    (opensps_hip_set_ssp_ops; compute_csp_batch's ssp_ops argument) */
    const double *prm;           /* device (OPS_NPAR, n): prm[slot * n + m] */
    const double *ssp_weights;   /* device (nt, n) or NULL: SSP weights instead of the SFH parameters */
    double *mags;                /* device (nbands, n) AB magnitudes, NaN for masked bands */
    void *spec;                  /* device (nspec, n) L_sun/Hz per solar mass formed, or NULL */
    double *mass, *lbol;         /* device (n): surviving mass, log10 L_bol; or NULL */
    double *fband;               /* device (nbands, n): band fluxes before the distance term, or NULL */
    void *young, *old;           /* device (nspec, n): the age sums before dust (split_parts = 1), or NULL */
    int *idx;                    /* device (3, n): it_tesc, zlo (1-based), per-model status; or NULL */
    double *weights;             /* device (nt, n): the FP64 SFH weights; or NULL */
    int *op_count;               /* device (2, n): with the operator, the points of each metallicity pass with a
                                    nonzero IMF weight (the isochrone points the IMF limits select); or NULL */
    const double *ssp_weights_z; /* device (nt, nz, n) or NULL: weights over ages and metallicities (compute_csp_batch's
                                    ssp_weights_z, opensps_csp_batch.f90:616-618: age_sums_w / model_sums_zw); they
                                    replace the SFH and the per-model metallicity blend */
    const double *zred_mags;     /* device (n) or NULL: each model's magnitude redshift (compute_csp_batch's zred_mags,
                                    opensps_csp_batch.f90:1054-1061): the observed frame and the distance term at it,
                                    the IGM at zred; the reverse pass then gives zred's row the IGM's part alone
                                    (batch_backward's lzm, :4830-4832, 4928) */
    double *mdust;               /* device (n) or NULL: the dust mass of the dust emission's templates
                                    (compute_csp_batch's mdust, opensps_csp_batch.f90:1206-1211; SAFE_FLOOR for a
                                    model without emission) */
    double *emlines;             /* device (nline, n) or NULL: with nebular emission the lines' luminosities through
                                    the dust (compute_csp_batch's emlines, attenuate_lines :3365-3436), else 0 */
} opensps_hip_batch_t;

/* The reverse pass of a batch (compute_csp_batch with params_bar / weights_bar / weights_z_bar, include/opensps.h:
 * 324-348; batch_backward, opensps_csp_batch.f90:4713-5136): the forward batch fwd (its outputs are written as
 * opensps_hip_csp_batch writes them), then the derivatives of the loss whose cotangents are
 * given (device arrays, any NULL): mags_bar (nbands, n), spec_bar (nspec, n), mass_bar, lbol_bar (n). Outputs
 * (device, any NULL): params_bar (OPENSPS_HIP_NBGRAD, n) with the rows of the reference's BGRAD_* numbering
 * (opensps_csp_batch.f90:166-213) that this path computes (dust1, dust2, dust3, dust_index, dust1_index,
 * frac_obrun, frac_nodust, logzsol, zred, sigma_smooth, igm_factor and, with SFH parameters (no given weights),
 * the SFH rows tage, tau, const, fburst, tburst, sf_start, sf_trunc, sf_slope of batch_sfh_gradients
 * (opensps_csp_batch.f90:5198-5217, SFHT_ROWS :215; sfh 0, 1, 4, 5) and, with fwd.use_ops (the operator's reverse
 * pass, ssp_batch_age_sums_adjoint, opensps_ssp_batch.f90:991-1248), the IMF rows imf1, imf2 and imf3 (Kroupa) or
 * imf1, imf2 and mdave (Dave), none for Salpeter; the others are 0), weights_bar (nt, n) (with fwd.ssp_weights or
 * SFH parameters), weights_z_bar (nt, nz, n) (with fwd.ssp_weights_z, which fwd.use_ops excludes: status 2). */
#define OPENSPS_HIP_NBGRAD 44
typedef struct {
    opensps_hip_batch_t fwd;
    const double *mags_bar, *spec_bar, *mass_bar, *lbol_bar;
    double *params_bar, *weights_bar, *weights_z_bar;
} opensps_hip_grad_t;

/* ---- context ---- */
/* No reference counterpart (the reference keeps device data in the Fortran context through OpenMP
 * mappings, src/core/opensps_context.f90:1089): creates a context on HIP device `device`. */
int opensps_hip_create(int device, opensps_hip_ctx_t **ctx);
int opensps_hip_destroy(opensps_hip_ctx_t *ctx);
/* Runs the context's work on `stream` (a hipStream_t; NULL: the context's own stream). */
int opensps_hip_set_stream(opensps_hip_ctx_t *ctx, void *stream);

/* The resident SSP bases and the grids, host arrays: lam (nspec, Angstrom ascending), log_age (nt,
 * log10 yr ascending), logz (nz, log10 Z/Z_sun ascending), spec (nspec, nt, nz) L_sun/Hz, mass (nt,
 * nz) surviving mass, lbol (nt, nz) log10 L_bol. Replaces opensps_context_update_ssp_basis
 * (src/core/opensps_context.f90:757) plus the device residency of the basis (:1089). Drops the bands, which
 * index the grid: call opensps_hip_set_bands after it. */
int opensps_hip_set_grid(opensps_hip_ctx_t *ctx, int nspec, const double *lam, int nt, const double *log_age, int nz,
                         const double *logz, const double *spec, const double *mass, const double *lbol);

/* The filters, host (nspec, nbands), normalised as filters_build leaves them (src/spectra/opensps_filters.f90:395-503);
 * the supports and trapezoid coefficients of prepare_batch_grid (src/spectra/opensps_batch_photometry.f90:237-394). */
int opensps_hip_set_bands(opensps_hip_ctx_t *ctx, int nbands, const double *bands);
/* The computed bands (opensps_context_set_mag_compute, include/opensps.h:233): mask[b] != 0; NULL: all. */
int opensps_hip_set_mag_mask(opensps_hip_ctx_t *ctx, int nbands, const int *mask);
/* A fixed diffuse attenuation curve A_lambda / A_V on the grid (nspec, host; the grid's wavelengths): dust_type 2
 * (Calzetti et al. 2000) or 6 (Reddy et al. 2015) as attenuation_table_row tabulates them
 * (src/physics/opensps_csp_batch.f90:3525-3566, opensps_dust.f90:567-662); tau = dust2 curve, dust3 curve for the
 * old stars' screen (pixel_transmissions :3765-3818). NULL: the power law (dust_type 0). Call after set_grid; the
 * reverse pass with a curve is status 2. */
int opensps_hip_set_dust_curve(opensps_hip_ctx_t *ctx, int nspec, const double *curve);
/* The SFH table of sfh = 2, 3 (ctx%state%sfh_tab: opensps_context_set_sfh_tab, src/bindings/c/opensps_c_api.f90:
 * 2062-2089; load_tabular_sfh for sfh = 2): ntab forward ages in years, increasing, and SFRs (M_sun / yr), host;
 * 0 <= ntab <= 20000 (NTABMAX), else status 3. The weights of a tabular model (sfh_weights_core_pc,
 * src/physics/opensps_sfh.f90:1219-1292) are the table's bins as linear SFR segments up to its tage. */
int opensps_hip_set_sfh_tab(opensps_hip_ctx_t *ctx, int ntab, const double *age_yr, const double *sfr);
/* The Vega system (magnitudes_from_fluxes, src/spectra/opensps_batch_photometry.f90:185-232, with
 * compute_vega_mags = 1): magvega (host, nbands) Vega's AB magnitude per band (filters_zero_points,
 * src/spectra/opensps_filters.f90:536-556), vref the reference band (1-based; 0: Vega is magnitude 0 in every band;
 * k > 0: band k stays AB, is always computed, and the others are AB - magvega(b) + magvega(k)). magvega NULL or
 * nbands 0: AB magnitudes. Call after opensps_hip_set_bands (which returns to AB). */
int opensps_hip_set_vega(opensps_hip_ctx_t *ctx, int nbands, const double *magvega, int vref);
/* The cosmology table (redshift, luminosity distance in pc), n rows, host (cosmospl,
 * src/physics/opensps_cosmology.f90:288-300). */
int opensps_hip_set_cosmology(opensps_hip_ctx_t *ctx, int n, const double *z, const double *dl);
/* add_igm_absorption; smooth_fast = (smooth_velocity == 1 and smoothspec_fast == 1). */
int opensps_hip_set_options(opensps_hip_ctx_t *ctx, int add_igm, int smooth_fast);
/* add_dust_emission and add_agn_dust (the reference's defaults are 1 and 1): a model emits dust when add_dust_emission
 * is 1 and dust1 or dust2 > 0 (batch_model_indices, opensps_csp_batch.f90:1569-1570), the AGN torus when
 * add_agn_dust is 1 and fagn > 0 (agn_indices :1591-1606); a batch that needs templates the context lacks is
 * status 2. */
int opensps_hip_set_emission(opensps_hip_ctx_t *ctx, int add_dust_emission, int add_agn_dust);
/* The Draine & Li (2007) dust emission templates on the grid, host: table (nspec, nqpah, 2 numin) column-major, for
 * each (q_PAH, U_min) the spectrum of U_max = U_min, then of the power-law U distribution to U_max = 1e6
 * (dustem2_dustem, src/io/opensps_io.f90:959-1069, as interpolated there onto the grid); qpah (nqpah) and umin
 * (numin) ascending. nqpah = 0 removes them. Call after opensps_hip_set_grid. */
int opensps_hip_set_dust_emission(opensps_hip_ctx_t *ctx, int nqpah, const double *qpah, int numin, const double *umin,
                                  const double *table);
/* The AGN torus templates on the grid, host: table (nspec, ntau) column-major, tau (ntau) ascending (agndust_spec,
 * src/io/opensps_io.f90:1101-1155). ntau = 0 removes them. Call after opensps_hip_set_grid. */
int opensps_hip_set_agn(opensps_hip_ctx_t *ctx, int ntau, const double *tau, const double *table);
/* Nebular emission with the CLOUDY grids (add_nebular_batch and nebular_kernels, opensps_csp_batch.f90:2565-2998;
 * the library read as read_nebular_standard reads it, src/io/opensps_io.f90:914-937): axes logz (nnz), age (nna,
 * yr), logu (nnu), the continuum's wavelengths lam (nlam), all ascending; cont (nnz, nna, nnu, nlam) L_sun/Hz per unit
 * ionizing photon rate and lines (nnz, nna, nnu, nline) L_sun per unit rate, C order, with the lines' wavelengths.
 * The library places them on the grid as load_nebular_grid does (:749-823) and derives the nebular ages (those up to
 * the library's last age) and their photons from the grid's basis: set the grid first; setting the grid drops them.
 * nnz = 0 removes them. */
int opensps_hip_set_nebular(opensps_hip_ctx_t *ctx, int nnz, const double *logz, int nna, const double *age, int nnu,
                            const double *logu, int nlam, const double *lam, const double *cont, int nline,
                            const double *line_lam, const double *lines);
/* The context's add_neb_emission (1: on), add_neb_continuum and nebemlineinspec (defaults 0, 1, 1, as the reference's).
 * A batch with nebular emission: status 2 without tables, with use_ops or ssp_weights_z, with nebemlineinspec = 0 while
 * dust emission is on, from opensps_hip_csp_batch_noalloc, in the reverse pass, and (the per-model status) for a
 * model whose sigma_smooth differs from the first model's while the lines are in the spectrum (one profile width a
 * batch, opensps_csp_batch.f90:749-752). */
int opensps_hip_set_neb_options(opensps_hip_ctx_t *ctx, int add_neb_emission, int add_neb_continuum,
                                int nebemlineinspec);

/* The batch: compute_csp_batch's stages 1-8 (src/physics/opensps_csp_batch.f90:431-1214) on the device,
 * asynchronous on the context's stream; the status of the models is checked by opensps_hip_sync. */
int opensps_hip_csp_batch(opensps_hip_ctx_t *ctx, const opensps_hip_batch_t *args);
/* No counterpart in the reference: opensps_hip_csp_batch for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). The same kernels and results; the batch's
 * scratch comes from a fixed array of 64 MiB per device, held while the process lives, instead of the context's
 * grow-only buffers, and nothing is allocated or freed. Status 3 (nothing launched) when the batch needs more than
 * 64 MiB of scratch: about 25 KB per model in mixed precision with spectra requested (~2600 models), 100 KB per
 * model in FP64 with young and old apart (~650 models); 2 with use_ops or ssp_weights_z. Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (several contexts or streams at once) race, and
 * opensps_hip_sync must follow each call before the next one. */
int opensps_hip_csp_batch_noalloc(opensps_hip_ctx_t *ctx, const opensps_hip_batch_t *args);
/* The forward batch and its reverse pass, asynchronous on the context's stream (opensps_hip_sync waits). Replaces
 * batch_backward's stages for the outputs above (opensps_csp_batch.f90:4713-5136; backward_observed_frame :6958,
 * smooth_columns_adjoint / smooth_sigma_adjoint, opensps_batch_photometry.f90:604-700, atten_adjoint_chunk
 * :6476-6665, backward_age_sums_w :5288-5330, backward_age_sums :5226-5286, batch_mass_cotangent :5142). */
int opensps_hip_csp_batch_grad(opensps_hip_ctx_t *ctx, const opensps_hip_grad_t *args);
/* Waits for the context's stream; status 2 if a model of the last batch was unsupported (idx row 3). */
int opensps_hip_sync(opensps_hip_ctx_t *ctx);

/* ---- the SSP operator with the IMF per model ---- */
/* A caller-built operator set: the isochrone points of every (metallicity, age), host arrays. Replaces the
 * tables of ssp_opset_t (src/physics/opensps_ssp_batch.f90:99-147) that the reference builds from isochrones
 * (opensps_context_build_ssp_ops, include/opensps.h:430): here one template library serves every metallicity.
 * Points of (iz, t) are [age_ptr[iz nt + t], age_ptr[iz nt + t + 1]), sorted by initial mass within an age;
 * m1, m2: the mass bin (first = 1: the first bin of an age, whose lower edge is the IMF's lower bound); cm, cl,
 * q: surviving mass, L_bol and ionizing photons per star; age_mmax: the largest living initial mass of each
 * isochrone (the remnants); each point's nent entries (template, value; template -1: none); etpl: the
 * library (ntpl, nspec) on the context's grid. nz and nt must be the grid's. */
typedef struct {
    int nz, nt, npts, ntpl, nent;
    const int *age_ptr;          /* nz nt + 1 */
    const double *mini, *m1, *m2;
    const int *first;
    const double *cm, *cl, *q;
    const double *age_mmax;      /* nz nt */
    const int *ent_tpl;          /* npts nent */
    const double *ent_val;       /* npts nent */
    const double *etpl;          /* ntpl nspec: template e's spectrum at etpl[e nspec + j] */
} opensps_hip_ssp_ops_t;
int opensps_hip_set_ssp_ops(opensps_hip_ctx_t *ctx, const opensps_hip_ssp_ops_t *ops);
/* opensps_context_free_ssp_ops (include/opensps.h:445) */
int opensps_hip_free_ssp_ops(opensps_hip_ctx_t *ctx);
/* The context's IMF settings for the operator (imf_type: 0 Salpeter, 2 Kroupa, 4 Dave; add_stellar_remnants;
 * salp_ind, mdave; the black-hole and neutron-star progenitor limits mlim_bh, mlim_ns). Per-model slopes and
 * limits come from the slots imf1..imf3, imf_upper_limit, imf_lower_limit of the batch. */
int opensps_hip_set_imf(opensps_hip_ctx_t *ctx, int imf_type, int remnants, double salp_ind, double mdave,
                        double mlim_bh, double mlim_ns);

/* Per-kernel timing with HIP events on the context's stream (no reference counterpart; the reference times its
 * stages with rocprof / nsys traces, doc/GPU_RESULTS.md:210-224): on != 0 records events around each kernel of
 * every batch; opensps_hip_kernel_times returns the accumulated milliseconds and launch counts of [prep, group,
 * age-sum product, chain] (ms[4], launches[4]; either may be NULL) and, with reset != 0, clears them. With the
 * operator, "group" also holds the IMF weights, the fold and the mass sums, and the product is the template one. */
int opensps_hip_set_profiling(opensps_hip_ctx_t *ctx, int on);
int opensps_hip_kernel_times(opensps_hip_ctx_t *ctx, double *ms, int *launches, int reset);
/* No reference counterpart (profiling): with profiling on, the SSP operator's phases of each batch whose operator
 * launches take one round: [the IMF weights (k_op_imf, k_op_slow, k_op_weights), the fold (k_op_agesums,
 * k_op_segments, k_op_fold, k_op_sums)], accumulated milliseconds and batch counts. */
int opensps_hip_op_kernel_times(opensps_hip_ctx_t *ctx, double *ms, int *launches, int reset);
/* The same for the reverse pass's kernels: [reverse chain, transposed product, weights and params_bar with the
 * SFH tangents] (ms[3],
 * launches[3]); the forward half of a batch with gradients counts in opensps_hip_kernel_times. */
int opensps_hip_grad_kernel_times(opensps_hip_ctx_t *ctx, double *ms, int *launches, int reset);

/* ---- kernel-level entry points (the reference's per-stage launchers) ---- */
/* band_fluxes (opensps_hip_band_fluxes_sp/_dp, src/accel/hip/opensps_hip_kernels.hip:3566, 3623): device
 * bands_t (nbands, nspec), j_lo, j_hi (1-based), do_band, wt (nspec), fobs_t (n, nspec) -> fband (nbands, n).
 * precision selects FP64 or FP32 (float bands_t, wt, fobs_t) inputs; the sums are FP64. */
int opensps_hip_band_fluxes(int precision, int n, int nspec, int nbands, const void *bands_t, const int *j_lo,
                            const int *j_hi, const int *do_band, const void *wt, const void *fobs_t, double *fband,
                            void *stream);
/* observed_frame_z (opensps_hip_observed_frame, opensps_hip_kernels.hip:3503): device zc (n), lambda
 * (nspec), total (nspec, n) of the precision's type -> fobs_t (n, nspec). */
int opensps_hip_observed_frame(int precision, int n, int nspec, const double *zc, const double *lambda,
                               const void *total, void *fobs_t, void *stream);

/* Device self-test of the matrix-instruction lane layouts the age-sum products assume (as the reference's
 * opensps_hip_selftest_eft checks its error-free transforms, opensps_hip_kernels.hip:3267-3294): which = 0
 * runs one f64 16x16x4 instruction on a[64], b[64] (doubles, one per lane), which = 1 one bf16 16x16x32
 * instruction on a[64][8], b[64][8] (bf16 bit patterns); d receives each lane's four accumulators
 * (d[64][4], double or float). Host pointers. */
int opensps_hip_selftest_mfma(int which, const void *a, const void *b, void *d);

/* ---- memory (the Fortran host's device buffers) ---- */
int opensps_hip_malloc(void **ptr, size_t bytes);
int opensps_hip_free(void *ptr);
int opensps_hip_memcpy_h2d(void *dst, const void *src, size_t bytes);
int opensps_hip_memcpy_d2h(void *dst, const void *src, size_t bytes);
int opensps_hip_device_count(int *n);

/* The message of the last failure of ctx (or of the process-wide calls when ctx is NULL). */
int opensps_hip_last_error(opensps_hip_ctx_t *ctx, char *msg, int len);

#ifdef __cplusplus
}
#endif
#endif
