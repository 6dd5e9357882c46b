/*
 * Test shim: calls the header-only struct entry point of include/opensps.h (opensps_context_compute_csp_batch_ext)
 * from compiled C, with the arguments of opensps_context_compute_csp_batch_grad_rows, so that a test can compare the
 * two entry points (the reference's struct form, include/opensps.h:366-421). rows 0 takes the struct's default.
 * opts_null != 0 passes NULL options (status 1).
 */
#include <stddef.h>

#include "opensps.h"

void shim_batch_ext(int handle, int n, int n_params, const char *const *names, const double *const *values,
                    const double *ssp_weights, const double *ssp_weights_z, const double *mags_bar,
                    const double *spec_bar, const double *mass_bar, const double *lbol_bar, double *mags, double *spec,
                    double *mass, double *lbol, double *params_bar, int rows, double *weights_bar,
                    double *weights_z_bar, int opts_null, int *status) {
    opensps_batch_options_t o;
    opensps_batch_results_t r;
    o.n = n;
    o.n_params = n_params;
    o.names = names;
    o.values = values;
    o.ssp_weights = ssp_weights;
    o.ssp_weights_z = ssp_weights_z;
    o.zred_mags = NULL;
    o.mags_bar = mags_bar;
    o.spec_bar = spec_bar;
    o.mass_bar = mass_bar;
    o.lbol_bar = lbol_bar;
    o.params_bar_rows = rows;
    r.mags = mags;
    r.spec = spec;
    r.mass = mass;
    r.lbol = lbol;
    r.mdust = NULL;
    r.sfr = NULL;
    r.emlines = NULL;
    r.indx = NULL;
    r.params_bar = params_bar;
    r.weights_bar = weights_bar;
    r.weights_z_bar = weights_z_bar;
    opensps_context_compute_csp_batch_ext(handle, opts_null ? NULL : &o, &r, status);
}
