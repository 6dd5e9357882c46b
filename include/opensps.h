/*
 * opensps.h -- the public C ABI of libopensps.so for OpenSPS's batched composite-population path on MI355X.
 *
 * A drop-in for the reference's context API on this path (include/opensps.h of the reference; bindings in
 * src/bindings/c/opensps_c_api.f90). The names, argument meaning, array layout (column-major, quantity fastest)
 * and status codes are the reference's; each declaration cites the one it replaces. Entry points without a
 * reference counterpart are marked "No reference counterpart". The batch runs on the GPU: without a HIP device
 * a batch returns status 5, as the reference's offload builds do when their target regions would run on the host.
 */
#ifndef OPENSPS_H
#define OPENSPS_H

#ifdef __cplusplus
extern "C" {
#endif

#define OPENSPS_NBGRAD 44   /* rows of params_bar (reference include/opensps.h:15) */

#define OPENSPS_STATUS_OK 0
#define OPENSPS_ERR_UNKNOWN_INT_PARAM 101
#define OPENSPS_ERR_UNKNOWN_FLOAT_PARAM 102
#define OPENSPS_ERR_UNKNOWN_STRING_PARAM 103
#define OPENSPS_ERR_COMPUTE_UNINIT 301
#define OPENSPS_ERR_CSP_UNINIT 302

#define OPENSPS_DETECTOR_PHOTON 0
#define OPENSPS_DETECTOR_ENERGY 1
#define OPENSPS_RESAMPLE_LINEAR 0
#define OPENSPS_RESAMPLE_CONSERVING 1
#define OPENSPS_GRID_POLICY_ERROR 0
#define OPENSPS_GRID_POLICY_WARN 1
#define OPENSPS_GRID_POLICY_ALLOW 2

/* ---- driver utilities (reference include/opensps.h:109-126) ---- */
void opensps_set_debug(int flag);
void opensps_clear_error(void);
void opensps_get_last_error(int *status, char *message, int message_len);
void opensps_get_driver_version(int *major, int *minor, int *patch);
const char *opensps_get_version_string(void);
const char *opensps_get_build_id(void);
void opensps_get_nbgrad(int *nbgrad);
void opensps_get_grad_param_name(int grad_idx, char *name, int max_len, int *status);

/* ---- contexts (reference include/opensps.h:132-151) ---- */
void opensps_context_create(int *handle, int *status);
void opensps_context_destroy(int handle, int *status);
/* Set-up from a data release is not part of this build: status 301 (reference include/opensps.h:134-135). */
void opensps_context_setup(int zin, const char *isoc_type, const char *spec_type, const char *dust_type, int handle,
                           int *status);
/* No reference counterpart (the reference builds its basis at set-up and installs it with
 * opensps_context_update_ssp_basis, src/core/opensps_context.f90:757): a resident SSP basis from the caller.
 * lambda (nspec) Angstrom increasing; log_age (ntfull) log10 yr increasing; zlegend (nz) metallicity mass
 * fractions increasing; spec (nspec x ntfull x nz) L_sun/Hz per solar mass; mass, lbol (ntfull x nz) surviving
 * mass and log10 L_bol. status 1 bad handle, 3 bad sizes or grids. Resets the filters. */
void opensps_context_set_ssp_basis(int handle, int nspec, const double *lambda, int ntfull, const double *log_age,
                                   int nz, const double *zlegend, double zsol, const double *spec,
                                   const double *mass, const double *lbol, int *status);
/* Parameter setters (reference include/opensps.h:145-147). Besides the reference's keys, set_str
 * "kernel_precision" = "double" (default: FP64 age sums and chain) or "mixed" (split-BF16 age sums on the matrix
 * cores, FP32 chain), and set_int "mixed_precision" = 0 / 1 select the kernels at run time (the reference selects
 * them at build time, meson option kernel_precision). The SSP operator's IMF follows set_int "imf_type",
 * "add_stellar_remnants" and set_float "mdave"; evaluate_ssp_op's metallicity is set_int "zmet". */
void opensps_context_set_int(int handle, const char *key, int value, int *status);
void opensps_context_set_float(int handle, const char *key, double value, int *status);
void opensps_context_set_str(int handle, const char *key, const char *value, int *status);
void opensps_context_set_fast_mode(int handle, int fast_mode, int *status);   /* include/opensps.h:150 */

/* ---- dimensions and grids (reference include/opensps.h:153-175) ---- */
void opensps_context_get_dims(int handle, int *n_spec, int *n_time, int *status);
void opensps_context_get_nspec(int handle, int *n_spec, int *status);
void opensps_context_get_ntfull(int handle, int *n_time, int *status);
void opensps_context_get_nbands(int handle, int *n_bands, int *status);
void opensps_context_get_nindx(int handle, int *n_indices, int *status);
void opensps_context_get_nz(int handle, int *n_z, int *status);
void opensps_context_get_nemline(int handle, int *n_line, int *status);
void opensps_context_get_zsol(int handle, double *z_sol, int *status);
void opensps_context_get_lambda(int handle, double *lambda, int *status);
void opensps_context_get_ages(int handle, double *log_age, int *status);
void opensps_context_get_zlegend(int handle, double *z, int *status);
/* wave_eff: pivot wavelengths; mag_vega, mag_sun: the AB magnitudes of Vega and the Sun per band (the Sun's in the
 * Vega system when compute_vega_mags was 1 at set_filters), 99 without reference spectra
 * (opensps_context_set_reference_spectra) or without flux. */
void opensps_context_get_filter_data(int handle, double *wave_eff, double *mag_vega, double *mag_sun, int *status);
/* No reference counterpart (the reference reads Vega and the Sun from its data release at set-up, load_standard_sed,
 * src/io/opensps_io.f90:1516-1582): the reference spectra as tables, wavelength (Angstrom, increasing) and flux,
 * F_lambda (flambda != 0; Vega's table) or f_nu (the Sun's). They are placed on the model grid as load_standard_sed
 * places them (log-log interpolation, F_lambda times lambda^2, tiny redward of the table) whenever filters are
 * installed, and give the Vega and solar zero points and the Vega system (set_int "compute_vega_mags" = 1). status 1
 * bad handle, 3 fewer than 2 points, a non-increasing or non-positive wavelength, a negative or non-finite flux. */
void opensps_context_set_reference_spectra(int handle, int n_vega, const double *lam_vega, const double *flux_vega,
                                           int vega_flambda, int n_sun, const double *lam_sun, const double *flux_sun,
                                           int sun_flambda, int *status);
/* No reference counterpart (the reference reads its dust emission library at set-up from its data directory,
 * load_dust_emission_table, src/io/opensps_io.f90:959-1069): the Draine & Li (2007) templates (or any library of that
 * layout): qpah (nqpah), umin (numin) and lam (nlam, Angstrom) ascending, spectra nlam x 2 x numin x nqpah
 * column-major (the data library's [qpah][umin][component][wavelength]: per (q_PAH, U_min) the spectrum of U_max =
 * U_min, then of the power-law U distribution). The batch places them on the grid as the reference does (linear in
 * wavelength from 1 micron on) and, with add_dust_emission = 1, adds each dusty model's emission scaled to the light
 * its dust absorbs (energy balance; duste_qpah, duste_umin, duste_gamma per model; mdust out). Without them a dusty
 * batch with add_dust_emission = 1 is status 2. status 1 bad handle, 3 bad sizes, NULL arrays or tables not
 * ascending. */
void opensps_context_set_dust_emission_templates(int handle, int nqpah, const double *qpah, int numin,
                                                 const double *umin, int nlam, const double *lam,
                                                 const double *spectra, int *status);
/* No reference counterpart (the reference reads its CLOUDY library at set-up, read_nebular_standard,
 * src/io/opensps_io.f90:914-937, and places it on the grid, load_nebular_grid, :749-823): the tables: logz (nnz),
 * age (nna, yr), logu (nnu), lam (nlam, Angstrom) ascending; cont nnz x nna x nnu x nlam (L_sun/Hz) and lines
 * nnz x nna x nnu x nline (L_sun), per unit ionizing photon rate, C order; line_lam (nline, Angstrom). With
 * add_neb_emission = 1 each model adds the nebular continuum (add_neb_continuum = 1) and lines (into the spectrum with
 * nebemlineinspec = 1) of its young stars' ionizing photons at its gas_logz, gas_logu, escaping frac_obrun of them,
 * and emlines (get_nemline = nline rows) holds the lines' attenuated luminosities. Without the tables such a batch is
 * status 2, and so is add_neb_emission = 2 (nebular emission in the SSPs). status 1 bad handle, 3 bad sizes, NULL
 * arrays or axes not ascending. */
void opensps_context_set_nebular_tables(int handle, int nnz, const double *logz, int nna, const double *age, int nnu,
                                        const double *logu, int nlam, const double *lam, const double *cont, int nline,
                                        const double *line_lam, const double *lines, int *status);
/* No reference counterpart (the reference reads its AGN torus library at set-up, load_agn_dust_models,
 * src/io/opensps_io.f90:1101-1155): the templates: tau (ntau) and lam (nlam, Angstrom) ascending, spectra nlam x ntau
 * column-major (the data library's [tau][wavelength]). With add_agn_dust = 1 each model with fagn > 0 adds fagn L_bol
 * times its agn_tau template, attenuated by the diffuse dust, after the IGM. Without them such a batch is status 2.
 * status 1 bad handle, 3 bad sizes, NULL arrays or tables not ascending. */
void opensps_context_set_agn_templates(int handle, int ntau, const double *tau, int nlam, const double *lam,
                                       const double *spectra, int *status);

/* ---- filters (reference include/opensps.h:192-233) ---- */
void opensps_filterset_create(int *fset, int *status);
void opensps_filterset_add(int fset, const char *name, int n, const double *lambda_aa, const double *response,
                           int detector, double cal_lambda_ref, double cal_exponent, int *status);
void opensps_filterset_count(int fset, int *n, int *status);
void opensps_filterset_free(int fset, int *status);
void opensps_context_set_filters(int handle, int fset, int resample, int grid_policy, int vega_ref_band, int *status);
void opensps_context_set_mag_compute(int handle, int n_bands, const int *mask, int *status);
/* The SFH table of sfh = 2, 3 (reference include/opensps.h:229-232): ntab forward ages [Gyr], SFRs [M_sun/yr] and
 * metallicities (met may be NULL; the batch uses one metallicity, as the reference's). The reference loads sfh = 2's
 * table from a file at set-up; here both types read this table. status 1 bad handle, 3 ntab out of 1 .. 20000. */
void opensps_context_set_sfh_tab(int handle, int ntab, const double *age, const double *sfr, const double *met,
                                 int *status);
/* No reference counterpart: the context's bands on the model grid (nspec x nbands), normalised. */
void opensps_context_get_bands(int handle, double *bands, int *status);
/* No reference counterpart: the context's cosmology table, *nrow rows of redshift and luminosity distance (pc), the
 * table the batch's distance term interpolates (cosmospl); z, dl may be NULL. */
void opensps_context_get_cosmology(int handle, int *nrow, double *z, double *dl, int *status);

/* ---- the batch (reference include/opensps.h:264-316) ---- */
/* status 0, 1 bad handle / no basis, 2 unsupported (a parameter that changes the SSPs; imf1, imf2, imf3,
 * imf_lower_limit, imf_upper_limit without an SSP operator on the context, and vdmc, mdave per model; spectral
 * indices, nebular emission, dust emission or the AGN torus without their templates
 * (opensps_context_set_nebular_tables, _set_dust_emission_templates, _set_agn_templates), nebular emission in the
 * SSPs (add_neb_emission = 2) or with an SSP operator, Vega magnitudes without reference spectra,
 * dust_type other than 0, 2, 6, an SFH type other than 0 to 5), 3 unknown parameter or bad shape, 5 no HIP device,
 * 10 device error. Dust emission (add_dust_emission = 1, the default: models with dust1 or dust2 > 0; mdust out) and
 * the AGN torus (add_agn_dust = 1: models with fagn > 0) and nebular emission (add_neb_emission = 1: gas_logz,
 * gas_logu, frac_obrun per model; emlines out) run with their tables. Members of the reference's parameter set that
 * no stage here reads (pmetals, dust_clumps, uvb, mwr, gas_logn, gas_logno, gas_logco) are accepted. With an SSP
 * operator on the context (opensps_context_set_ssp_ops) every batch takes each model's IMF from it, as the reference's does
 * (src/bindings/c/opensps_c_api.f90:913-917). */
void opensps_context_compute_csp_batch(int handle, int n, int n_params, const char *const *names,
                                       const double *const *values, const double *ssp_weights, double *mags,
                                       double *spec, double *mass, double *lbol, double *mdust, double *sfr,
                                       double *emlines, double *indx, int *status);
void opensps_context_compute_csp_batch_parts(int handle, int n, int n_params, const char *const *names,
                                             const double *const *values, const double *ssp_weights, double *mags,
                                             double *spec, double *mass, double *lbol, double *mdust, double *sfr,
                                             double *emlines, double *indx, double *young, double *old,
                                             double *mass_young, double *mass_old, double *lbol_young,
                                             double *lbol_old, int *status);
/* The batch with gradients (include/opensps.h:324-364 of the reference): the forward outputs and the derivatives
 * of the loss whose cotangents are given (mags_bar nbands x n, spec_bar nspec x n, mass_bar, lbol_bar n; any
 * NULL) with respect to params_bar (NBGRAD x n), weights_bar (ntfull x n) and weights_z_bar (ntfull x nz x n).
 * ssp_weights_z (ntfull x nz x n) replaces ssp_weights and the metallicity blend. This build computes the rows of
 * dust1, dust2, dust3, dust_index, dust1_index, frac_obrun, frac_nodust, logzsol, zred, sigma_smooth, igm_factor
 * and, with SFH parameters (no weights given; sfh 0, 1, 4, 5), tage, tau, const, fburst, tburst, sf_start, sf_trunc
 * and sf_slope and, with SSP operators on the context, imf1, imf2 and imf3 (Kroupa) or imf1, imf2 and mdave (Dave)
 * (the others are 0); a batch whose dust emission or AGN torus acts is status 2 here (their reverse pass is not
 * built). zred_mags (n, or NULL): each model's magnitude redshift (the observed frame and the distance
 * term at it, the IGM at zred; zred's row is then the IGM's part); redshift_colors = 1 needs it (else status 2).
 * _rows: params_bar's leading dimension (fewer than NBGRAD: status 3, nothing written; rows beyond NBGRAD are left
 * alone). */
void opensps_context_compute_csp_batch_grad(int handle, int n, int n_params, const char *const *names,
                                            const double *const *values, const double *ssp_weights,
                                            const double *ssp_weights_z, const double *zred_mags,
                                            const double *mags_bar, const double *spec_bar, const double *mass_bar,
                                            const double *lbol_bar, double *mags, double *spec, double *mass,
                                            double *lbol, double *mdust, double *sfr, double *emlines, double *indx,
                                            double *params_bar, double *weights_bar, double *weights_z_bar,
                                            int *status);
void opensps_context_compute_csp_batch_grad_rows(int handle, int n, int n_params, const char *const *names,
                                                 const double *const *values, const double *ssp_weights,
                                                 const double *ssp_weights_z, const double *zred_mags,
                                                 const double *mags_bar, const double *spec_bar,
                                                 const double *mass_bar, const double *lbol_bar, double *mags,
                                                 double *spec, double *mass, double *lbol, double *mdust,
                                                 double *sfr, double *emlines, double *indx, double *params_bar,
                                                 int params_bar_rows, double *weights_bar, double *weights_z_bar,
                                                 int *status);
void opensps_context_get_batch_maggies(int handle, double *maggies, int *n_bands, int *n, int *status);
/* *bytes: 4 with kernel_precision "mixed", 8 with "double". */
void opensps_context_batch_spec_device(int handle, void **ptr, int *nspec, int *n, int *bytes, int *status);

/* The struct form of the batch with gradients (reference include/opensps.h:366-421, header-only there too): the
 * options and the result buffers of opensps_context_compute_csp_batch_grad_rows in two structs; params_bar_rows 0
 * means OPENSPS_NBGRAD. NULL opts, res or status: status 1 (when status is not NULL). */
typedef struct {
    int n;                          /* models */
    int n_params;                   /* entries of names / values */
    const char *const *names;       /* n_params parameter names */
    const double *const *values;    /* n_params arrays of n values */
    const double *ssp_weights;      /* ntfull x n, or NULL */
    const double *ssp_weights_z;    /* ntfull x nz x n, or NULL */
    const double *zred_mags;        /* n, or NULL */
    const double *mags_bar;         /* nbands x n, or NULL: the cotangents */
    const double *spec_bar;         /* nspec x n, or NULL */
    const double *mass_bar;         /* n, or NULL */
    const double *lbol_bar;         /* n, or NULL */
    int params_bar_rows;            /* leading dimension of params_bar; 0: OPENSPS_NBGRAD */
} opensps_batch_options_t;

typedef struct {
    double *mags, *spec, *mass, *lbol, *mdust, *sfr, *emlines, *indx;   /* forward outputs, any NULL but mags */
    double *params_bar;             /* rows x n */
    double *weights_bar;            /* ntfull x n */
    double *weights_z_bar;          /* ntfull x nz x n */
} opensps_batch_results_t;

static inline void opensps_context_compute_csp_batch_ext(int handle, const opensps_batch_options_t *opts,
                                                         opensps_batch_results_t *res, int *status) {
    int rows;
    if (opts == 0 || res == 0 || status == 0) {
        if (status != 0) *status = 1;
        return;
    }
    rows = opts->params_bar_rows > 0 ? opts->params_bar_rows : OPENSPS_NBGRAD;
    opensps_context_compute_csp_batch_grad_rows(handle, opts->n, opts->n_params, opts->names, opts->values,
                                                opts->ssp_weights, opts->ssp_weights_z, opts->zred_mags,
                                                opts->mags_bar, opts->spec_bar, opts->mass_bar, opts->lbol_bar,
                                                res->mags, res->spec, res->mass, res->lbol, res->mdust, res->sfr,
                                                res->emlines, res->indx, res->params_bar, rows, res->weights_bar,
                                                res->weights_z_bar, status);
}

/* ---- the SSP operator: the IMF per model (reference include/opensps.h:425-455) ---- */
/* Building the operator from a data release's isochrones and libraries is not part of this build: status 301 (the
 * reference's status for a context that is not set up). */
void opensps_context_build_ssp_ops(int handle, int all_metallicities, int *status);
void opensps_context_build_ssp_ops_free_mods(int handle, int all_metallicities, int *status);
/* No reference counterpart (the reference builds its operator set from isochrones, ssp_operator_build /
 * ssp_opset_build): an operator given by the caller, installed on the GPU. Every metallicity and age of the basis:
 * points of (iz, t) are [age_ptr[iz ntfull + t], age_ptr[iz ntfull + t + 1]), sorted by initial mass mini within
 * an age; m1, m2 their mass bins (first = 1: an age's first bin, whose lower edge is the IMF's lower bound); cm,
 * cl, q surviving mass, L_bol and ionizing photons per star; age_mmax (ntfull x nz) the largest living initial
 * mass of each isochrone; nent entries per point: ent_tpl (template, -1 none) and ent_val; etpl (nspec x ntpl) the
 * template library on the basis' wavelengths. status 1 bad handle / no basis, 3 bad sizes or tables, 5 no HIP
 * device, 10 device error. A new basis drops the operator. */
void opensps_context_set_ssp_ops(int handle, int nz, int ntfull, int npts, int ntpl, int nent, const int *age_ptr,
                                 const double *mini, const double *m1, const double *m2, const int *first,
                                 const double *cm, const double *cl, const double *q, const double *age_mmax,
                                 const int *ent_tpl, const double *ent_val, const double *etpl, int *status);
void opensps_context_free_ssp_ops(int handle, int *status);                   /* include/opensps.h:443 */
void opensps_context_has_ssp_ops(int handle, int *has_ops, int *status);      /* include/opensps.h:446 */
/* include/opensps.h:451-455: the SSP of every age for one IMF (imf_type < 0: the context's; imf_params [imf1, imf2,
 * imf3, vdmc, mdave] or NULL: the context's) at the context's metallicity zmet, in FP64: mass_grid (ntfull)
 * surviving mass, lbol_grid (ntfull) log10 L_bol, spec_grid (nspec x ntfull). status 303 without an operator, 3 NULL
 * outputs, 2 an IMF type other than 0, 2, 4 or (Dave's, type 4) an mdave other than the context's. */
void opensps_context_evaluate_ssp_op(int handle, int imf_type, const double *imf_params, double *mass_grid,
                                     double *lbol_grid, double *spec_grid, int *status);

/* No reference counterpart: the batch on device-resident arrays for callers that keep their data on the GPU
 * (the reference's batch_spec_device serves them on the output side). d_prm: device OPS_NPAR x n doubles in the
 * slots of opensps_params.h; d_weights: device ntfull x n or NULL; precision -1 (the context's), 0 FP64, 1
 * mixed; split 1: young and old sums apart (birth-cloud or old-star screens). Device outputs: d_mags (nbands x n,
 * required), d_spec (nspec x n, float with mixed precision, double with FP64), d_mass, d_lbol (n). */
void opensps_context_compute_csp_batch_device(int handle, int n, const double *d_prm, const double *d_weights,
                                              int precision, int split, double *d_mags, void *d_spec, double *d_mass,
                                              double *d_lbol, int *status);
/* No reference counterpart: the accumulated device time (ms) and launch counts of the batch's kernels [prep,
 * group, age-sum product, chain], from HIP events on the context's stream once set_int("profile_kernels", 1);
 * reset != 0 clears them. status 5 before the first device batch. */
void opensps_context_kernel_times(int handle, double *ms, int *launches, int reset, int *status);
/* No reference counterpart: with profile_kernels on, the SSP operator's two phases of each batch whose operator
 * launches take one round, [the IMF weights (k_op_imf, k_op_slow, k_op_weights), the ages' sums, segments, fold and
 * mass sums], accumulated ms and batch counts; reset != 0 clears them. status 5 before the first device batch. */
void opensps_context_op_kernel_times(int handle, double *ms, int *launches, int reset, int *status);
/* No reference counterpart: the context's parameter values in the slot order of opensps_params.h. */
void opensps_context_get_params(int handle, double *prm, int *status);

#ifdef __cplusplus
}
#endif
#endif
