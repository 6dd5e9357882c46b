/*
 * csp_oracle.c -- FP64 CPU restatement of OpenSPS's batched composite-population path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the correctness oracle of the GPU path: only the tests,
 * the smoke check and the CPU-baseline leg of the benchmark load it, and only as the checker. The
 * product never links or calls it.
 *
 * Each function restates one routine of the reference's default FP64 build (kernel_precision =
 * double, CPU) and cites it. The reference itself cannot be built here under the project's rules
 * (46 Fortran modules behind Meson plus three configure templates that have to be generated), so
 * this restatement is the oracle, pinned by the checks in tests/test_oracle_*.py: brute force on
 * tiny grids, closed forms, independent library routines (scipy) and the reference's own unit-test
 * fixtures.
 *
 * Conventions: arrays are column-major as in the reference (quantity fastest); "index" values that
 * the reference computes with find_interval stay 1-based and are converted where arrays are read.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../include/opensps_params.h"

#ifdef _OPENMP
#include <omp.h>
#endif

/* LOG_E = log10(exp(1)) (opensps_sfh.f90:61), formed as the reference forms it */
#define LOG_E (log10(exp(1.0)))

#define SAFE_FLOOR DBL_MIN                /* tiny(1.0_wp): opensps_constants.f90 */
#define C_LIGHT 2.99792458e18             /* A/s: opensps_constants.f90 */
#define N_SIGMA_RANGE 4.0                 /* opensps_smoothing.f90:28 */
#define AB_ZEROPOINT 48.60                /* opensps_batch_photometry.f90:37 */
#define N_COSMO_GRID 10000                /* opensps_cosmology.f90:43 */

/* ABS_MAG_ZEROPOINT_LOG = log10(L_sun / (4 pi (10 pc)^2)) (opensps_constants.f90:87-88) */
static double abs_mag_zeropoint_log(void) {
    const double pi = acos(-1.0);
    const double pc = (648000.0 / pi) * 1.495978707e13;
    return log10(3.828e33 / (4.0 * pi * (10.0 * pc) * (10.0 * pc)));
}

/* ---------------------------------------------------------------------------------------------
 * Interval search: find_interval_no_hint (src/math/opensps_interpolation.f90:149-198).
 * Returns the 1-based lower index i with a(i) <= v < a(i+1) for ascending a (the mirrored
 * condition for descending a), 0 below / n above the ends, 1 at v == a(1) and n-1 at v == a(n).
 * ------------------------------------------------------------------------------------------- */
int orc_find_interval(const double *a, int n, double v) {
    if (n < 2) return 1;
    int asc = a[n - 1] >= a[0];
    int lower = 0, upper = n + 1;
    while (upper - lower > 1) {
        int mid = (upper + lower) / 2;
        if (asc == (v >= a[mid - 1])) lower = mid; else upper = mid;
    }
    if (v == a[0]) return 1;
    if (v == a[n - 1]) return n - 1;
    return lower;
}

/* interpolate_linear_scalar (opensps_interpolation.f90:45-72) */
static double interp_linear(const double *x, const double *y, int n, double xo) {
    if (n < 2) return NAN;
    int i = orc_find_interval(x, n, xo);
    if (i < 1) i = 1;
    if (i > n - 1) i = n - 1;
    double slope = (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
    return y[i - 1] + slope * (xo - x[i - 1]);
}

/* mag_from_flux (src/math/opensps_special_functions.f90:34-48) */
static double mag_from_flux(double f) { return f > 0.0 ? -2.5 * log10(f) : NAN; }

/* expi: Ei(x), series for x <= 40, asymptotic above (opensps_special_functions.f90:50-122) */
double orc_expi(double x) {
    const double gam = 0.57721566490153286060651209008240243104215933593992;
    if (x == 0.0) return -INFINITY;
    if (x < 0.0) return NAN;
    double res, r;
    if (fabs(x) <= 40.0) {
        res = 1.0; r = 1.0;
        for (int k = 1; k <= 1000; k++) {
            double td = (double)k + 1.0;
            r = (r * (double)k * x) / (td * td);
            res = res + r;
            if (fabs(r) <= 1.0e-20 * fabs(res)) break;
            if (k == 1000) return NAN;
        }
        return gam + log(x) + x * res;
    }
    res = 1.0; r = 1.0;
    double inv_x = 1.0 / x;
    for (int k = 1; k <= 20; k++) { r = r * (double)k * inv_x; res = res + r; }
    return exp(x) * inv_x * res;
}

/* ---------------------------------------------------------------------------------------------
 * SFH weights (src/physics/opensps_sfh.f90)
 * ------------------------------------------------------------------------------------------- */
typedef struct {
    double tau, tage, tburst, sf_trunc, sf_slope, tq, t0, tb;
    int type, use_simha_limits;
} sfhp_t;

/* convert_sfhparams_pc (opensps_sfh.f90:1305-1355) */
static void convert_sfhparams(const double *p, double tage, sfhp_t *s) {
    int sfh = (int)p[OPS_P_SFH];
    double start = (sfh == 1 || sfh == 4 || sfh == 5) ? p[OPS_P_SF_START] * 1.0e9 : 0.0;
    s->tage = tage * 1.0e9 - start;
    s->tburst = p[OPS_P_TBURST] * 1.0e9 - start;
    s->sf_trunc = p[OPS_P_SF_TRUNC] * 1.0e9 - start;
    s->tau = p[OPS_P_TAU] * 1.0e9;
    s->sf_slope = -p[OPS_P_SF_SLOPE] / 1.0e9;
    s->tb = s->tage - s->tburst;
    if (s->sf_trunc <= 0.0 || s->sf_trunc > s->tage) s->tq = 0.0; else s->tq = s->tage - s->sf_trunc;
    if (s->sf_slope > SAFE_FLOOR) s->t0 = s->tq - 1.0 / s->sf_slope; else s->t0 = 0.0;
    if (s->t0 > s->tq || s->t0 <= 0.0) s->t0 = 0.0;
    s->use_simha_limits = 0;
    s->type = 0;
}

/* eval_indefinite_moments_log (opensps_sfh.f90:852-910), types 0, 1, 4, 5 */
static void moments_log(double log_t, const sfhp_t *s, double *m0, double *m1) {
    double t_lin = pow(10.0, log_t);
    switch (s->type) {
    case 0:
        *m0 = t_lin;
        *m1 = t_lin * (log_t - LOG_E);
        return;
    case 1: {
        if (s->tau <= SAFE_FLOOR) { *m0 = 0.0; *m1 = 0.0; return; }
        double tt = fmax(t_lin / s->tau, SAFE_FLOOR);
        double ei = orc_expi(tt);
        if (!isfinite(ei)) ei = 0.0;
        *m0 = exp(tt);
        *m1 = log_t * *m0 - LOG_E * ei;
        return;
    }
    case 4: {
        if (s->tau <= SAFE_FLOOR) { *m0 = 0.0; *m1 = 0.0; return; }
        double tt = fmax(t_lin / s->tau, SAFE_FLOOR);
        double ei = orc_expi(tt);
        if (!isfinite(ei)) ei = 0.0;
        double et = exp(tt);
        *m0 = -(t_lin - s->tage - s->tau) * et;
        double term_const = s->tau * LOG_E;
        *m1 = -((t_lin - s->tage - s->tau) * log_t - term_const) * et - (s->tage + s->tau) * LOG_E * ei;
        return;
    }
    case 5: {   /* Simha's linear part: SFR ~ 1 - slope (tq - t) in lookback time */
        const double tp = fmax(0.0, s->tage - s->sf_trunc), sf = 1.0 - s->sf_slope * tp;
        const double q = s->sf_slope * (t_lin * t_lin) / 2.0;
        *m0 = sf * t_lin + q;
        *m1 = sf * t_lin * (log_t - LOG_E) + q * (log_t - 0.5 * LOG_E);
        return;
    }
    default:
        *m0 = 0.0; *m1 = 0.0;
    }
}

/* clamp_integration_limits (opensps_sfh.f90:967-1000) */
static double clamp_limit(double tiny_logt, double log_t, const sfhp_t *s) {
    double lmin = pow(10.0, tiny_logt);
    double ts = s->use_simha_limits ? s->t0 : s->tq;
    double te = s->use_simha_limits ? s->tq : s->tage;
    double ls = log10(fmax(ts, lmin)), le = log10(fmax(te, lmin));
    return fmin(fmax(log_t, ls), le);
}

/* compute_ssp_weights (opensps_sfh.f90:97-257), logarithmic interpolation (interpolation_type 0) */
static void compute_ssp_weights(const double *tf, int nt, double tiny_logt, const sfhp_t *s, int idx_min,
                                int idx_max, double *w) {
    for (int i = 0; i < nt; i++) w[i] = 0.0;
    if (s->type == -1) {
        if (s->tb < 0.0) return;
        double log_tb = log10(fmax(s->tb, pow(10.0, tiny_logt)));
        if (log_tb <= tf[0]) { w[0] = 1.0; return; }
        int j = orc_find_interval(tf, nt, log_tb);
        if (j < 1) j = 1;
        if (j > nt - 1) j = nt - 1;
        double bdt = tf[j] - tf[j - 1];
        w[j - 1] = (tf[j] - log_tb) / bdt;
        w[j] = (log_tb - tf[j - 1]) / bdt;
        return;
    }
    double lmin = pow(10.0, tiny_logt);
    double ts = s->use_simha_limits ? s->t0 : s->tq;
    double te = s->use_simha_limits ? s->tq : s->tage;
    double lts = log10(fmax(ts, lmin)), lte = log10(fmax(te, lmin));
    int j_start = orc_find_interval(tf, nt, lts);
    int j_end = orc_find_interval(tf, nt, lte) + 1;
    int lo = idx_min > 1 ? idx_min : 1;
    if (j_start < lo) j_start = lo;
    int hi = idx_max < nt - 1 ? idx_max : nt - 1;
    if (j_end > hi) j_end = hi;
    double lim1 = fmin(fmax(tf[j_start - 1], lts), lte);
    double lim_prev = lim1, m0p, m1p;
    moments_log(lim1, s, &m0p, &m1p);
    for (int j = j_start; j <= j_end; j++) {
        double lim2 = fmin(fmax(tf[j], lts), lte);
        if (lim2 <= lim_prev) { lim_prev = lim2; continue; }
        double m0c, m1c;
        moments_log(lim2, s, &m0c, &m1c);
        double m0 = m0c - m0p, m1 = m1c - m1p;
        m0p = m0c; m1p = m1c;
        double nlo = tf[j - 1], nhi = tf[j];
        double inv_dt = 1.0 / (nhi - nlo);
        w[j - 1] = w[j - 1] + (nhi * m0 - m1) * inv_dt;
        w[j] = w[j] + (m1 - nlo * m0) * inv_dt;
        lim_prev = lim2;
    }
    if (idx_min == 0) {
        double l1 = clamp_limit(tiny_logt, tiny_logt, s);
        double l2 = clamp_limit(tiny_logt, tf[0], s);
        if (l1 < l2) {
            double a0, a1, b0, b1;
            moments_log(l2, s, &a0, &a1);
            moments_log(l1, s, &b0, &b1);
            w[0] = w[0] + (a0 - b0);
        }
    }
}

/* expm1 as the reference forms it (opensps_special_functions.f90:219-233): a cubic below |x| = 1e-5 */
static double expm1_ref(double x) {
    if (fabs(x) < 1.0e-5) return x * (1.0 + x * (0.5 + x * (1.0 / 6.0)));
    return exp(x) - 1.0;
}

/* gammainc(2, x), the regularised lower incomplete gamma (opensps_special_functions.f90:142-167) */
static double gammainc2(double x) {
    if (x < 0.0) return 0.0;
    return -expm1_ref(-x) - x * exp(-x);
}

/* The Simha SFH's linear mass fraction at age tage (Gyr): get_sfh_properties_at_age_pc's frac_linear for sfh = 5
 * (opensps_sfh.f90:327-455); log_tmax: the oldest grid age (log10 yr) */
double orc_simha_frac(const double *p, double log_tmax) {
    const double t_max = pow(10.0, log_tmax - 9.0) - p[OPS_P_SF_START];
    const double t_prime = p[OPS_P_TAGE] - p[OPS_P_SF_START];
    const double sf_trunc = p[OPS_P_SF_TRUNC], sf_start = p[OPS_P_SF_START];
    const double t_trunc = (sf_trunc < SAFE_FLOOR || sf_trunc < sf_start) ? t_max : sf_trunc - sf_start;
    const double tau = fmax(p[OPS_P_TAU], SAFE_FLOOR);
    double mass_tau = 0.0;
    if (t_prime >= 0.0) mass_tau = tau * gammainc2(fmin(t_prime, t_trunc) / tau);
    const double slope_m = -p[OPS_P_SF_SLOPE];
    const double t_zero = slope_m > 0.0 ? t_trunc + 1.0 / slope_m : t_max;
    double mass_linear = 0.0;
    if (!(t_trunc <= 0.0 || t_trunc > t_max) && t_prime > t_trunc) {
        const double sfr_trunc = (t_trunc / p[OPS_P_TAU]) * exp(-t_trunc / p[OPS_P_TAU]);
        const double t_hi = fmin(t_zero, t_prime);
        mass_linear = sfr_trunc * ((t_hi - t_trunc) - (slope_m / 2.0) * (t_hi * t_hi + t_trunc * t_trunc) +
                                   slope_m * t_hi * t_trunc);
    }
    if (t_prime >= 0.0 && mass_linear + mass_tau > SAFE_FLOOR) return mass_linear / (mass_linear + mass_tau);
    return 0.0;
}

/* The context's SFH table (opensps_context_set_sfh_tab, src/bindings/c/opensps_c_api.f90:2062-2089): forward ages
 * in years and SFRs (M_sun / yr), read by sfh = 2, 3 */
#define ORC_NTABMAX 20000
static int orc_ntab = 0;
static double orc_tab_age[ORC_NTABMAX + 1], orc_tab_sfr[ORC_NTABMAX + 1];
int orc_set_sfh_tab(int ntab, const double *age_yr, const double *sfr) {
    if (ntab < 0 || ntab > ORC_NTABMAX) return 3;
    for (int i = 0; i <= ORC_NTABMAX; i++) { orc_tab_age[i] = 0.0; orc_tab_sfr[i] = 0.0; }
    for (int i = 0; i < ntab; i++) { orc_tab_age[i] = age_yr[i]; orc_tab_sfr[i] = sfr[i]; }
    orc_ntab = ntab;
    return 0;
}

/* The tabular SFH (sfh_weights_core_pc, opensps_sfh.f90:1219-1292, one metallicity as the batch calls it): each bin
 * of the table a linear SFR segment (the type-5 moments with the bin's slope) over its lookback interval, its hat
 * weights normalised and scaled by the segment's mass */
static void tab_weights(const double *tf, int nt, double tiny_logt, double tage, double *w, double *tmp1) {
    for (int i = 0; i < nt; i++) w[i] = 0.0;
    const double lmin = pow(10.0, tiny_logt), lmax = pow(10.0, tf[nt - 1]);
    for (int j = 0; j < orc_ntab - 1; j++) {
        const double t1 = tage * 1.0e9 - orc_tab_age[j + 1], t2 = tage * 1.0e9 - orc_tab_age[j];
        if (t2 < 0.0) continue;
        const double dt = fmax(t2 - t1, SAFE_FLOOR);
        sfhp_t s;
        memset(&s, 0, sizeof s);
        s.sf_slope = -(orc_tab_sfr[j + 1] - orc_tab_sfr[j]) / dt / fmax(orc_tab_sfr[j + 1], SAFE_FLOOR);
        s.type = 5;
        s.tq = fmin(fmax(t1, lmin), lmax);
        s.tage = fmin(fmax(t2, lmin), lmax);
        s.sf_trunc = s.tage - s.tq;
        const double mass2 = orc_tab_sfr[j + 1] * (1.0 + s.sf_slope / 2.0 * (s.tage + s.tq - 2.0 * t1)) * (s.tage - s.tq);
        int imin = orc_find_interval(tf, nt, log10(fmax(t1, SAFE_FLOOR))) - 1;
        imin = imin < 0 ? 0 : (imin > nt ? nt : imin);
        int imax = orc_find_interval(tf, nt, log10(fmax(t2, SAFE_FLOOR))) + 2;
        imax = imax < 0 ? 0 : (imax > nt ? nt : imax);
        compute_ssp_weights(tf, nt, tiny_logt, &s, imin, imax, tmp1);
        double mass1 = 0.0;
        for (int i = 0; i < nt; i++) mass1 += tmp1[i];
        if (mass1 < SAFE_FLOOR) mass1 = 1.0;
        for (int i = 0; i < nt; i++) w[i] = w[i] + tmp1[i] * (mass2 / mass1);
    }
}

/* sfh_weights_core_pc (opensps_sfh.f90:1077-1294) for sfh = 0, 1, 4 (with const and burst), 5 (Simha) and 2, 3
 * (the context's table). Returns 0, or 2 for an SFH type this build does not compute. */
int orc_sfh_weights(const double *tf, int nt, double tiny_logt, const double *p, double *w, double *tmp1,
                    double *tmp2) {
    sfhp_t s;
    int sfh = (int)p[OPS_P_SFH];
    double tage = p[OPS_P_TAGE];
    for (int i = 0; i < nt; i++) { w[i] = 0.0; tmp1[i] = 0.0; tmp2[i] = 0.0; }
    if (sfh == 2 || sfh == 3) {
        tab_weights(tf, nt, tiny_logt, tage, w, tmp1);
        return 0;
    }
    convert_sfhparams(p, tage, &s);
    int imax = orc_find_interval(tf, nt, log10(s.tage)) + 2;
    if (imax < 1) imax = 1;
    if (imax > nt) imax = nt;
    if (sfh == 0) {
        s.type = -1;
        s.tb = s.tage;
        int imin = imax - 2 > 1 ? imax - 2 : 1;
        compute_ssp_weights(tf, nt, tiny_logt, &s, imin, imax, w);
        return 0;
    }
    if (sfh == 5) {   /* the delayed-tau part over [tq, tage] and the linear part over [t0, tq], mixed */
        s.type = 4;
        compute_ssp_weights(tf, nt, tiny_logt, &s, 0, imax, tmp1);
        double m1 = 0.0, m2 = 0.0;
        for (int i = 0; i < imax; i++) m1 += tmp1[i];
        s.type = 5;
        s.use_simha_limits = 1;
        compute_ssp_weights(tf, nt, tiny_logt, &s, 0, imax, tmp2);
        s.use_simha_limits = 0;
        for (int i = 0; i < imax; i++) m2 += tmp2[i];
        if (m1 < SAFE_FLOOR) m1 = 1.0;
        if (m2 < SAFE_FLOOR) m2 = 1.0;
        const double fl = orc_simha_frac(p, tf[nt - 1]);
        for (int i = 0; i < nt; i++) w[i] = (tmp1[i] / m1) * (1.0 - fl) + (tmp2[i] / m2) * fl;
        return 0;
    }
    if (sfh != 1 && sfh != 4) return 2;
    s.type = sfh;
    compute_ssp_weights(tf, nt, tiny_logt, &s, 0, imax, w);
    double mass1 = 0.0;
    for (int i = 0; i < imax; i++) mass1 += w[i];
    if (mass1 < SAFE_FLOOR) mass1 = 1.0;
    for (int i = 0; i < nt; i++) w[i] = w[i] / mass1;
    double cst = p[OPS_P_CONST], fb = p[OPS_P_FBURST];
    if (cst > 0.0 || fb > SAFE_FLOOR) {
        s.type = 0;
        compute_ssp_weights(tf, nt, tiny_logt, &s, 0, imax, tmp1);
        mass1 = 0.0;
        for (int i = 0; i < imax; i++) mass1 += tmp1[i];
        if (mass1 < SAFE_FLOOR) mass1 = 1.0;
        for (int i = 0; i < nt; i++) tmp2[i] = 0.0;
        double fbv = 0.0;
        if (s.tb >= 0.0) {
            s.type = -1;
            compute_ssp_weights(tf, nt, tiny_logt, &s, 0, imax, tmp2);
            fbv = fb;
        }
        for (int i = 0; i < nt; i++)
            w[i] = (1.0 - cst - fbv) * w[i] + cst * (tmp1[i] / mass1) + fbv * tmp2[i];
    }
    return 0;
}

/* ---------------------------------------------------------------------------------------------
 * SFH tangents (src/physics/opensps_sfh_tangent.f90): the weights' derivative with respect to one
 * SFH parameter, forward mode, sfh = 0, 1, 4 with const, fburst and truncation, and 5 (Simha), logarithmic
 * interpolation. Pinned by central differences of orc_sfh_weights (tests/test_oracle_sfh_tangent.py).
 * ------------------------------------------------------------------------------------------- */
#define ORC_GYR 1.0e9
enum { SFHT_TAGE = 1, SFHT_TAU, SFHT_CONST, SFHT_FBURST, SFHT_TBURST, SFHT_SF_START, SFHT_SF_TRUNC, SFHT_SF_SLOPE };

/* clamp_t (opensps_sfh_tangent.f90:407-422): min(max(node, lts), lte) and the derivative of the kept bound */
static void clamp_t(double node, double lts, double dlts, double lte, double dlte, double *v, double *dv) {
    *v = node;
    *dv = 0.0;
    if (lts > node) { *v = lts; *dv = dlts; }
    if (lte < *v) { *v = lte; *dv = dlte; }
}

/* moments_t (opensps_sfh_tangent.f90:426-467), logarithmic branch, types 0, 1, 4, 5: the indefinite moments at
 * log time lt and their total derivatives, with the time's (dlt) and the parameters' (dtg, dtau, dtq, ds: of tage,
 * tau, tq and the slope) tangents */
static void moments_t(double lt, double dlt, const sfhp_t *s, double dtg, double dtau, double dtq, double ds,
                      double *m0, double *dm0, double *m1, double *dm1) {
    const double ln10 = log(10.0);
    const double t = pow(10.0, lt), dt = t * ln10 * dlt;
    *m0 = 0.0; *dm0 = 0.0; *m1 = 0.0; *dm1 = 0.0;
    if ((s->type == 1 || s->type == 4) && s->tau <= SAFE_FLOOR) return;
    if (s->type == 5) {   /* Simha's linear part: slope factor 1 - s t', t' = max(0, tage - sf_trunc) */
        const double sl = s->sf_slope, tp = fmax(0.0, s->tage - s->sf_trunc);
        const double dtp = s->tage - s->sf_trunc > 0.0 ? dtq : 0.0;
        const double sf = 1.0 - sl * tp, dsf = -ds * tp - sl * dtp;
        const double h = sl * (t * t) / 2.0, dh = ds * (t * t) / 2.0 + sl * t * dt;
        *m0 = sf * t + h;
        *dm0 = dsf * t + sf * dt + dh;
        *m1 = sf * t * (lt - LOG_E) + h * (lt - 0.5 * LOG_E);
        *dm1 = dsf * t * (lt - LOG_E) + sf * (dt * (lt - LOG_E) + t * dlt) + dh * (lt - 0.5 * LOG_E) + h * dlt;
        return;
    }
    if (s->type == 0) {
        *m0 = t;
        *dm0 = dt;
        *m1 = t * (lt - LOG_E);
        *dm1 = dt * (lt - LOG_E) + t * dlt;
        return;
    }
    if (s->type != 1 && s->type != 4) return;
    const double tau = s->tau, tg = s->tage;
    double x = t / tau, dx = dt / tau - t * dtau / (tau * tau);
    if (x < SAFE_FLOOR) { x = SAFE_FLOOR; dx = 0.0; }
    const double ex = exp(x);
    double ei = orc_expi(x), dei = ex / x * dx;
    if (!isfinite(ei)) { ei = 0.0; dei = 0.0; }
    if (s->type == 1) {
        *m0 = ex;
        *dm0 = ex * dx;
        *m1 = lt * ex - LOG_E * ei;
        *dm1 = dlt * ex + lt * *dm0 - LOG_E * dei;
    } else {
        const double a = t - tg - tau, da = dt - dtg - dtau, b = a * lt - tau * LOG_E;
        *m0 = -a * ex;
        *dm0 = -(da * ex + a * ex * dx);
        *m1 = -b * ex - (tg + tau) * LOG_E * ei;
        *dm1 = -((da * lt + a * dlt - dtau * LOG_E) * ex + b * ex * dx) - (dtg + dtau) * LOG_E * ei -
               (tg + tau) * LOG_E * dei;
    }
}

/* ssp_weights_t (opensps_sfh_tangent.f90:255-358), logarithmic interpolation: compute_ssp_weights with the
 * tangents dtg, dtau, dtb, dtq of tage, tau, tb and tq (years). Segments of zero width stay (their tangent is the
 * one-sided one where a window end sits on a grid age). */
static void ssp_weights_t(const double *tf, int nt, double tiny_logt, const sfhp_t *s, double dtg, double dtau,
                          double dtb, double dtq, double dt0, double ds, int idx_min, int idx_max, double *w,
                          double *dw) {
    const double ln10 = log(10.0), lmin = pow(10.0, tiny_logt);
    for (int i = 0; i < nt; i++) { w[i] = 0.0; dw[i] = 0.0; }
    if (s->type == -1) {
        if (s->tb < 0.0) return;
        const double ltb = log10(fmax(s->tb, lmin));
        const double dltb = s->tb > lmin ? dtb / (s->tb * ln10) : 0.0;
        if (ltb <= tf[0]) { w[0] = 1.0; return; }
        int j = orc_find_interval(tf, nt, ltb);
        if (j < 1) j = 1;
        if (j > nt - 1) j = nt - 1;
        const double bdt = tf[j] - tf[j - 1];
        w[j - 1] = (tf[j] - ltb) / bdt;
        w[j] = (ltb - tf[j - 1]) / bdt;
        dw[j - 1] = -dltb / bdt;
        dw[j] = dltb / bdt;
        return;
    }
    /* the window [tq, tage] (lookback years); Simha's linear part: [t0, tq] */
    const double ts = s->use_simha_limits ? s->t0 : s->tq, dts = s->use_simha_limits ? dt0 : dtq;
    const double te = s->use_simha_limits ? s->tq : s->tage, dte = s->use_simha_limits ? dtq : dtg;
    const double lts = log10(fmax(ts, lmin)), dlts = ts > lmin ? dts / (ts * ln10) : 0.0;
    const double lte = log10(fmax(te, lmin)), dlte = te > lmin ? dte / (te * ln10) : 0.0;
    int j_start = orc_find_interval(tf, nt, lts);
    int j_end = orc_find_interval(tf, nt, lte) + 1;
    if (j_start < (idx_min > 1 ? idx_min : 1)) j_start = idx_min > 1 ? idx_min : 1;
    if (j_end > (idx_max < nt - 1 ? idx_max : nt - 1)) j_end = idx_max < nt - 1 ? idx_max : nt - 1;
    double l1, dl1;
    clamp_t(tf[j_start - 1], lts, dlts, lte, dlte, &l1, &dl1);
    double lp = l1, m0p, dm0p, m1p, dm1p;
    moments_t(l1, dl1, s, dtg, dtau, dtq, ds, &m0p, &dm0p, &m1p, &dm1p);
    for (int j = j_start; j <= j_end; j++) {
        double l2, dl2;
        clamp_t(tf[j], lts, dlts, lte, dlte, &l2, &dl2);
        if (l2 < lp) { lp = l2; continue; }
        double m0c, dm0c, m1c, dm1c;
        moments_t(l2, dl2, s, dtg, dtau, dtq, ds, &m0c, &dm0c, &m1c, &dm1c);
        const double m0 = m0c - m0p, dm0 = dm0c - dm0p, m1 = m1c - m1p, dm1 = dm1c - dm1p;
        m0p = m0c; dm0p = dm0c; m1p = m1c; dm1p = dm1c;
        const double nlo = tf[j - 1], nhi = tf[j], inv = 1.0 / (nhi - nlo);
        w[j - 1] += (nhi * m0 - m1) * inv;
        w[j] += (m1 - nlo * m0) * inv;
        dw[j - 1] += (nhi * dm0 - dm1) * inv;
        dw[j] += (dm1 - nlo * dm0) * inv;
        lp = l2;
    }
    if (idx_min == 0) {
        double la, dla, lb, dlb;
        clamp_t(tiny_logt, lts, dlts, lte, dlte, &la, &dla);
        clamp_t(tf[0], lts, dlts, lte, dlte, &lb, &dlb);
        if (la < lb) {
            double ma, dma, mb, dmb, x1, x2;
            moments_t(lb, dlb, s, dtg, dtau, dtq, ds, &mb, &dmb, &x1, &x2);
            moments_t(la, dla, s, dtg, dtau, dtq, ds, &ma, &dma, &x1, &x2);
            w[0] += mb - ma;
            dw[0] += dmb - dma;
        }
    }
}

/* simha_frac_t (opensps_sfh_tangent.f90:362-403): Simha's linear mass fraction f (orc_simha_frac) and its derivative
 * df with respect to parameter q (per Gyr, per unit of sf_slope) */
void orc_simha_frac_t(const double *p, double log_tmax, int q, double *f, double *df) {
    const double sf_start = p[OPS_P_SF_START], sf_trunc = p[OPS_P_SF_TRUNC], tau_p = p[OPS_P_TAU];
    const double tmax = pow(10.0, log_tmax - 9.0) - sf_start, tpr = p[OPS_P_TAGE] - sf_start;
    const int trunc_on = !(sf_trunc < SAFE_FLOOR || sf_trunc < sf_start);
    const double ttr = trunc_on ? sf_trunc - sf_start : tmax, tau = fmax(tau_p, SAFE_FLOOR), sm = -p[OPS_P_SF_SLOPE];
    double dtmax = 0.0, dtpr = 0.0, dttr = 0.0, dtau = 0.0, dsm = 0.0;
    switch (q) {
    case SFHT_TAGE: dtpr = 1.0; break;
    case SFHT_TAU: if (tau_p > SAFE_FLOOR) dtau = 1.0; break;
    case SFHT_SF_START: dtmax = -1.0; dtpr = -1.0; dttr = -1.0; break;
    case SFHT_SF_TRUNC: if (trunc_on) dttr = 1.0; break;
    case SFHT_SF_SLOPE: dsm = -1.0; break;
    default: break;
    }
    if (!trunc_on) dttr = dtmax;
    *f = 0.0;
    *df = 0.0;
    if (tpr < 0.0 || ttr <= 0.0 || ttr > tmax) return;
    const double mn = tpr <= ttr ? tpr : ttr, dmn = tpr <= ttr ? dtpr : dttr;
    const double x = mn / tau, dx = (dmn - x * dtau) / tau;
    const double mt = tau * gammainc2(x), dmt = dtau * gammainc2(x) + tau * x * exp(-x) * dx;
    if (tpr <= ttr) return;
    double tz, dtz;
    if (sm > 0.0) { tz = ttr + 1.0 / sm; dtz = dttr - dsm / (sm * sm); }
    else { tz = tmax; dtz = dtmax; }
    const double th = tz <= tpr ? tz : tpr, dth = tz <= tpr ? dtz : dtpr;
    const double r = ttr / tau_p, dr = (dttr - r * dtau) / tau_p;
    const double sfr = r * exp(-r), dsfr = exp(-r) * (1.0 - r) * dr;
    const double bb = (th - ttr) - (sm / 2.0) * (th * th + ttr * ttr) + sm * th * ttr;
    const double dbb = (dth - dttr) - (dsm / 2.0) * (th * th + ttr * ttr) - sm * (th * dth + ttr * dttr) +
                       dsm * th * ttr + sm * (dth * ttr + th * dttr);
    const double ml = sfr * bb, dml = dsfr * bb + sfr * dbb;
    if (ml + mt > SAFE_FLOOR) {
        *f = ml / (ml + mt);
        *df = (dml * mt - ml * dmt) / ((ml + mt) * (ml + mt));
    }
}

/* sfh_weights_tangent (opensps_sfh_tangent.f90:66-251) for sfh = 0, 1, 4, 5: the weights w and their derivative dw
 * with respect to parameter q (SFHT_*, in the units of the parameters: Gyr). s: scratch (4 nt). Returns 0, or 2 for
 * an SFH type this build does not compute. */
int orc_sfh_weights_tangent(const double *tf, int nt, double tiny_logt, const double *p, int q, double *w,
                            double *dw, double *s) {
    sfhp_t h;
    const int sfh = (int)p[OPS_P_SFH];
    for (int i = 0; i < nt; i++) { w[i] = 0.0; dw[i] = 0.0; }
    if (sfh != 0 && sfh != 1 && sfh != 4 && sfh != 5) return 2;
    convert_sfhparams(p, p[OPS_P_TAGE], &h);
    double dtg = 0.0, dtau = 0.0, dtb = 0.0, dtq = 0.0, ds = 0.0;
    switch (q) {
    case SFHT_SF_SLOPE: ds = -1.0 / ORC_GYR; break;   /* the internal slope is -sf_slope / 1e9 per year */
    case SFHT_TAGE: dtg = ORC_GYR; dtb = ORC_GYR; if (h.tq > 0.0) dtq = ORC_GYR; break;
    case SFHT_SF_TRUNC: if (h.tq > 0.0) dtq = -ORC_GYR; break;
    case SFHT_TAU: dtau = ORC_GYR; break;
    case SFHT_TBURST: dtb = -ORC_GYR; break;
    case SFHT_SF_START: if (sfh != 0) dtg = -ORC_GYR; break;
    default: break;   /* const, fburst: the mixture below */
    }
    int imax = orc_find_interval(tf, nt, log10(h.tage)) + 2;
    if (imax < 1) imax = 1;
    if (imax > nt) imax = nt;
    if (sfh == 0) {   /* a burst at tb = tage */
        h.type = -1;
        h.tb = h.tage;
        ssp_weights_t(tf, nt, tiny_logt, &h, 0.0, 0.0, dtg, 0.0, 0.0, 0.0, imax - 2 > 1 ? imax - 2 : 1, imax, w, dw);
        return 0;
    }
    double *c0 = s, *c1 = s + nt, *b0 = s + 2 * nt, *b1 = s + 3 * nt;
    if (sfh == 5) {   /* Simha: both parts normalised, mixed by the linear part's mass fraction */
        h.type = 4;
        ssp_weights_t(tf, nt, tiny_logt, &h, dtg, dtau, 0.0, dtq, 0.0, 0.0, 0, imax, c0, c1);
        h.type = 5;
        h.use_simha_limits = 1;
        /* t0 = tq - 1 / slope, where the linear SFR reaches zero (if inside (0, tq]) */
        const double dt0 = h.t0 > 0.0 ? dtq + ds / (h.sf_slope * h.sf_slope) : 0.0;
        ssp_weights_t(tf, nt, tiny_logt, &h, dtg, 0.0, 0.0, dtq, dt0, ds, 0, imax, b0, b1);
        double m1 = 0.0, dm1 = 0.0, m2 = 0.0, dm2 = 0.0;
        for (int i = 0; i < imax; i++) { m1 += c0[i]; dm1 += c1[i]; m2 += b0[i]; dm2 += b1[i]; }
        if (m1 < SAFE_FLOOR) { m1 = 1.0; dm1 = 0.0; }
        if (m2 < SAFE_FLOOR) { m2 = 1.0; dm2 = 0.0; }
        double fl, dfl;
        orc_simha_frac_t(p, tf[nt - 1], q, &fl, &dfl);
        for (int i = 0; i < nt; i++) {
            const double wt = c0[i] / m1, dwt = (c1[i] - wt * dm1) / m1;
            const double cn = b0[i] / m2, dcn = (b1[i] - cn * dm2) / m2;
            w[i] = wt * (1.0 - fl) + cn * fl;
            dw[i] = dwt * (1.0 - fl) + dcn * fl + (cn - wt) * dfl;
        }
        return 0;
    }
    h.type = sfh;
    ssp_weights_t(tf, nt, tiny_logt, &h, dtg, dtau, 0.0, dtq, 0.0, 0.0, 0, imax, w, dw);
    double mass1 = 0.0, dmass1 = 0.0;
    for (int i = 0; i < imax; i++) { mass1 += w[i]; dmass1 += dw[i]; }
    if (mass1 < SAFE_FLOOR) { mass1 = 1.0; dmass1 = 0.0; }
    /* the constant and burst components, whatever const and fburst (one-sided derivatives at 0) */
    h.type = 0;
    ssp_weights_t(tf, nt, tiny_logt, &h, dtg, 0.0, 0.0, dtq, 0.0, 0.0, 0, imax, c0, c1);
    double massc = 0.0, dmassc = 0.0;
    for (int i = 0; i < imax; i++) { massc += c0[i]; dmassc += c1[i]; }
    if (massc < SAFE_FLOOR) { massc = 1.0; dmassc = 0.0; }
    for (int i = 0; i < nt; i++) { b0[i] = 0.0; b1[i] = 0.0; }
    double fb = 0.0;
    if (h.tb >= 0.0) {
        h.type = -1;
        ssp_weights_t(tf, nt, tiny_logt, &h, 0.0, 0.0, dtb, 0.0, 0.0, 0.0, 0, imax, b0, b1);
        fb = p[OPS_P_FBURST];
    }
    double cst = p[OPS_P_CONST];
    if (!(p[OPS_P_CONST] > 0.0 || p[OPS_P_FBURST] > SAFE_FLOOR)) { cst = 0.0; fb = 0.0; }
    for (int i = 0; i < nt; i++) {
        const double wt = w[i] / mass1, dwt = (dw[i] - wt * dmass1) / mass1;
        const double cn = c0[i] / massc, dcn = (c1[i] - cn * dmassc) / massc;
        w[i] = (1.0 - cst - fb) * wt + cst * cn + fb * b0[i];
        dw[i] = (1.0 - cst - fb) * dwt + cst * dcn + fb * b1[i];
        if (q == SFHT_CONST) dw[i] += cn - wt;
        if (q == SFHT_FBURST && h.tb >= 0.0) dw[i] += b0[i] - wt;
    }
    return 0;
}

/* batch_sfh_gradients (src/physics/opensps_csp_batch.f90:5198-5217): params_bar's SFH rows of one model (0-based:
 * the reference's SFHT_ROWS = 17 18 19 20 21 22 25 26 of :215, minus 1) from d L / d w (wbar, nt). s: scratch
 * (6 nt). Returns 0, or 2 for an SFH type this build does not compute. */
static const int ORC_SFHT_ROWS[8] = {16, 17, 18, 19, 20, 21, 24, 25};
int orc_sfh_gradients(const double *tf, int nt, double tiny_logt, const double *p, const double *wbar,
                      double *pb, double *s) {
    for (int q = 1; q <= 8; q++) {
        const int st = orc_sfh_weights_tangent(tf, nt, tiny_logt, p, q, s, s + nt, s + 2 * nt);
        if (st) return st;
        double acc = 0.0;
        for (int i = 0; i < nt; i++) acc += wbar[i] * s[nt + i];
        pb[ORC_SFHT_ROWS[q - 1]] = acc;
    }
    return 0;
}

/* ---------------------------------------------------------------------------------------------
 * IGM (src/physics/opensps_cosmology.f90:65-95, 333-507)
 * ------------------------------------------------------------------------------------------- */
static const double LY_WAVE[17] = {1215.67, 1025.72, 972.537, 949.743, 937.803, 930.748, 926.226, 923.150,
                                   920.963, 919.352, 918.129, 917.181, 916.429, 915.824, 915.329, 914.919,
                                   914.576};
static const double LY_COEFF[17] = {0.0036,    0.0017,    0.0011846, 0.0009410, 0.0007960, 0.0006967,
                                    0.0006236, 0.0005665, 0.0005200, 0.0004817, 0.0004487, 0.0004200,
                                    0.0003947, 0.000372,  0.000352,  0.0003334, 0.00031644};
#define LYMAN_LIMIT 911.75
#define A_METAL 0.0017
#define MC1 0.25
#define MC2 9.4
#define MC3 0.7
#define MC4 0.023
#define MP1 0.46
#define MP2 0.18
#define MP3 1.32
#define MP4 1.68

/* igm_tau: Madau (1995) optical depth at rest wavelength w for 1 + z = opz (cosmology.f90:406-447) */
double orc_igm_tau(double w, double opz) {
    double obs = w * opz, lr = obs / LYMAN_LIMIT, tau = 0.0;
    for (int i = 0; i < 17; i++) {
        if (w < LY_WAVE[i]) {
            tau = tau + LY_COEFF[i] * pow(obs / LY_WAVE[i], 3.46);
            if (i == 0) tau = tau + A_METAL * pow(obs / LY_WAVE[i], 1.68);
        }
    }
    if (w < LYMAN_LIMIT)
        tau = tau + (MC1 * pow(lr, 3.0) * (pow(opz, MP1) - pow(lr, MP1))) +
              (MC2 * pow(lr, 1.5) * (pow(opz, MP2) - pow(lr, MP2))) -
              (MC3 * pow(lr, 3.0) * (pow(lr, -MP3) - pow(opz, -MP3))) - (MC4 * (pow(opz, MP4) - pow(lr, MP4)));
    return tau;
}

/* igm_tau_pixel_terms (cosmology.f90:458-492) */
void orc_igm_pixel_terms(double w, double *a) {
    for (int k = 0; k < 6; k++) a[k] = 0.0;
    if (w < LY_WAVE[0]) {
        double ser = 0.0;
        for (int i = 0; i < 17; i++)
            if (w < LY_WAVE[i]) ser = ser + LY_COEFF[i] * pow(LY_WAVE[i], -3.46);
        a[0] = ser * pow(w, 3.46);
        a[1] = A_METAL * pow(w / LY_WAVE[0], 1.68);
    }
    if (w < LYMAN_LIMIT) {
        double s = w / LYMAN_LIMIT;
        a[2] = MC1 * pow(s, 3.0) * (1.0 - pow(s, MP1));
        a[3] = MC2 * pow(s, 1.5) * (1.0 - pow(s, MP2));
        a[4] = -MC3 * pow(s, 3.0) * (pow(s, -MP3) - 1.0);
        a[5] = -MC4 * (1.0 - pow(s, MP4));
    }
}

/* igm_tau_z_terms (cosmology.f90:495-507) */
void orc_igm_z_terms(double opz, double *b) {
    b[0] = pow(opz, 3.46);
    b[1] = pow(opz, 1.68);
    b[2] = pow(opz, 3.0 + MP1);
    b[3] = pow(opz, 1.5 + MP2);
    b[4] = pow(opz, 3.0 - MP3);
    b[5] = pow(opz, MP4);
}

/* ---------------------------------------------------------------------------------------------
 * Cosmology table and distance term (cosmology.f90:244-305; batch_photometry.f90:784-794)
 * ------------------------------------------------------------------------------------------- */
/* get_luminosity_distance (cosmology.f90:249-279), parsecs */
double orc_lum_distance(double h0, double om0, double ol0, double z) {
    if (z <= 0.0) return 0.0;
    double hd = ((C_LIGHT * 1.0e-13) / h0) * 1.0e6;
    double dz = z / (double)(N_COSMO_GRID - 1);
    double prev = 1.0 / sqrt(om0 + ol0), sum = 0.0;
    for (int i = 2; i <= N_COSMO_GRID; i++) {
        double zc = (double)(i - 1) * dz;
        double opz = 1.0 + zc;
        double now = 1.0 / sqrt(om0 * (opz * opz * opz) + ol0);
        sum = sum + 0.5 * dz * (prev + now);
        prev = now;
    }
    return (1.0 + z) * hd * sum;
}

/* update_cosmology_table (cosmology.f90:288-300): 500 redshifts (descending) and distances */
void orc_cosmology_table(double h0, double om0, double ol0, double *cz, double *cdl) {
    for (int i = 1; i <= 500; i++) {
        double a = (double)(i - 1) / 499.0 * (1.0 - 1.0 / 1001.0) + 1.0 / 1001.0;
        cz[i - 1] = (1.0 / a) - 1.0;
        cdl[i - 1] = orc_lum_distance(h0, om0, ol0, cz[i - 1]);
    }
}

/* distance_term (batch_photometry.f90:784-794) */
double orc_distance_term(const double *cz, const double *cdl, int nc, double z) {
    if (fabs(z) <= SAFE_FLOOR) return 0.0;
    double dm = interp_linear(cz, cdl, nc, z);
    if (dm != dm || dm <= SAFE_FLOOR) return 0.0;
    return 5.0 * log10(dm / 10.0) + mag_from_flux(1.0 + z);
}

/* ---------------------------------------------------------------------------------------------
 * Velocity smoothing operator (src/spectra/opensps_smoothing.f90:553-633)
 * lo/hi are 1-based (lo == 0: pixel passes through); coef is sized nspec * (2 * wmax + 1) by the
 * caller through orc_velocity_kernel_size. Returns the number of coefficients.
 * ------------------------------------------------------------------------------------------- */
int orc_velocity_kernel(const double *lam, int n, double sigma, double minl, double maxl, int *lo, int *hi,
                        int *off, double *coef, int cap) {
    for (int i = 0; i < n; i++) { lo[i] = 0; hi[i] = -1; off[i] = 0; }
    if (n < 2 || sigma <= SAFE_FLOOR) return 0;
    double sb = sigma / (C_LIGHT * 1.0e-13);
    double inv2 = 0.5 * (1.0 / sb) * (1.0 / sb);
    double fblue = exp(-N_SIGMA_RANGE * sb), fred = exp(N_SIGMA_RANGE * sb);
    int *wlo = (int *)calloc(n, sizeof(int)), *whi = (int *)calloc(n, sizeof(int));
    double *w = (double *)malloc(sizeof(double) * n);
    int il = 1, ih = 1;
    for (int i = 1; i <= n; i++) {
        wlo[i - 1] = 0; whi[i - 1] = -1;
        double li = lam[i - 1];
        if (li < minl || li > maxl) continue;
        double vl = li * fblue, vh = li * fred;
        while (il > 1 && lam[il - 1] > vl) il--;
        while (il < n && lam[il - 1] < vl) il++;
        if (il > 1 && lam[il - 1] > vl) il--;
        if (ih < il) ih = il;
        while (ih > il && lam[ih - 1] > vh) ih--;
        while (ih < n && lam[ih - 1] < vh) ih++;
        if (ih - il + 1 < 2) continue;
        wlo[i - 1] = il; whi[i - 1] = ih;
    }
    int nt = 0;
    for (int i = 1; i <= n; i++) {
        if (wlo[i - 1] == 0) continue;
        int a = wlo[i - 1], b = whi[i - 1], nw = b - a + 1;
        double lli = log(lam[i - 1]), tw = 0.0;
        for (int m = 0; m < nw; m++) {
            double d = lli - log(lam[a - 1 + m]);
            w[m] = exp(-inv2 * (d * d)) / lam[a - 1 + m];
        }
        for (int m = 1; m <= nw; m++) {
            double dl = 0.0, dr = 0.0;
            if (m > 1) dl = lam[a + m - 2] - lam[a + m - 3];
            if (m < nw) dr = lam[a + m - 1] - lam[a + m - 2];
            w[m - 1] = 0.5 * w[m - 1] * (dl + dr);
            tw = tw + w[m - 1];
        }
        if (tw <= SAFE_FLOOR) continue;
        if (nt + nw > cap) { nt = -1; break; }
        lo[i - 1] = a; hi[i - 1] = b; off[i - 1] = nt + 1;
        for (int m = 0; m < nw; m++) coef[nt + m] = w[m] / tw;
        nt += nw;
    }
    free(wlo); free(whi); free(w);
    return nt;
}

/* smooth_kernel, FP64 sums (batch_photometry.f90:740-780): one column in place */
static void apply_velocity_kernel(int n, const int *lo, const int *hi, const int *off, const double *coef,
                                  double *spec, double *work) {
    for (int i = 0; i < n; i++) {
        int a = lo[i];
        if (a == 0) { work[i] = spec[i]; continue; }
        int o = off[i] - a;   /* 1-based: coef(off + k) for k = lo .. hi */
        double s = 0.0;
        for (int k = a; k <= hi[i]; k++) s = s + coef[o + k - 1] * spec[k - 1];
        work[i] = s;
    }
    memcpy(spec, work, sizeof(double) * n);
}

/* ---------------------------------------------------------------------------------------------
 * Photometry (src/spectra/opensps_batch_photometry.f90)
 * ------------------------------------------------------------------------------------------- */
/* observed_frame_z for one column (batch_photometry.f90:819-903): fobs(j) = max(interp, 0) / lambda */
void orc_observed_frame(const double *lam, int n, double z, const double *total, double *fobs) {
    for (int j = 0; j < n; j++) {
        double v;
        if (fabs(z) > SAFE_FLOOR) {
            double tg = lam[j] / (1.0 + z);
            int il = orc_find_interval(lam, n, tg);
            if (il < 1) il = 1;
            if (il > n - 1) il = n - 1;
            double t = (tg - lam[il - 1]) / (lam[il] - lam[il - 1]);
            v = total[il - 1] + t * (total[il] - total[il - 1]);
        } else {
            v = total[j];
        }
        if (v != v) v = 0.0;
        double wl = lam[j] > DBL_MIN ? 1.0 / lam[j] : 0.0;
        fobs[j] = fmax(v, 0.0) * wl;
    }
}

/* Band supports (prepare_batch_grid, batch_photometry.f90:293-307): 1-based j_lo, j_hi */
void orc_band_supports(const double *bands, int nspec, int nbands, int *jlo, int *jhi) {
    for (int b = 0; b < nbands; b++) {
        int ilo = nspec, ihi = 1;
        for (int j = 1; j <= nspec; j++)
            if (bands[(size_t)b * nspec + j - 1] != 0.0) {
                if (j < ilo) ilo = j;
                if (j > ihi) ihi = j;
            }
        jlo[b] = ilo - 1 > 1 ? ilo - 1 : 1;
        jhi[b] = ihi < nspec - 1 ? ihi : nspec - 1;
        if (ilo > ihi) { jlo[b] = 1; jhi[b] = 0; }
    }
}

/* band_fluxes, FP64 sums (batch_photometry.f90:1049-1171), one column, transposed bands
 * bt(nbands, nspec); wt the trapezoid half-widths */
void orc_band_fluxes_col(int nspec, int nbands, const double *bt, const int *jlo, const int *jhi,
                         const int *do_band, const double *wt, const double *fobs, double *fband) {
    for (int b = 0; b < nbands; b++) {
        double s = 0.0;
        if (do_band[b] != 0)
            for (int j = jlo[b]; j <= jhi[b]; j++)
                s = s + wt[j - 1] * (fobs[j - 1] * bt[(size_t)(j - 1) * nbands + b] +
                                     fobs[j] * bt[(size_t)j * nbands + b]);
        fband[b] = s;
    }
}

/* ---------------------------------------------------------------------------------------------
 * Filters (src/spectra/opensps_filters.f90:319-503): resample a piecewise-linear curve onto the
 * grid, clip negatives, and normalise integral T / lambda d lambda = 1 by the trapezoid rule.
 * mode 0: point interpolation (resample_point), 1: hat projection (resample_hat). detector 1 is
 * energy counting (T / lambda at the grid points). Returns 0, or 1 for a dead filter.
 * ------------------------------------------------------------------------------------------- */
static int cell_of(const double *g, int ng, double x) {
    if (x <= g[0]) return 1;
    if (x >= g[ng - 1]) return ng - 1;
    int lo = 1, hi = ng;
    while (hi - lo > 1) {
        int mid = (lo + hi) / 2;
        if (g[mid - 1] <= x) lo = mid; else hi = mid;
    }
    return lo;
}

int orc_filter_band(const double *lc, const double *rc, int nc, const double *g, int ng, int mode,
                    int detector, double *band) {
    for (int j = 0; j < ng; j++) band[j] = 0.0;
    if (nc < 2) return 1;
    if (mode == 1) {
        double *acc = (double *)calloc(ng, sizeof(double));
        for (int k = 1; k <= nc - 1; k++) {
            double x0 = fmax(lc[k - 1], g[0]), x1 = fmin(lc[k], g[ng - 1]);
            if (x1 <= x0) continue;
            int j = cell_of(g, ng, x0);
            while (j <= ng - 1) {
                if (g[j - 1] >= x1) break;
                double a = fmax(x0, g[j - 1]), b = fmin(x1, g[j]);
                if (b > a) {
                    double dx = g[j] - g[j - 1];
                    double t = (a - lc[k - 1]) / (lc[k] - lc[k - 1]);
                    double ya = rc[k - 1] + t * (rc[k] - rc[k - 1]);
                    t = (b - lc[k - 1]) / (lc[k] - lc[k - 1]);
                    double yb = rc[k - 1] + t * (rc[k] - rc[k - 1]);
                    double m = 0.5 * (a + b), ym = 0.5 * (ya + yb);
                    double ha = (g[j] - a) / dx, hb = (g[j] - b) / dx, hm = (g[j] - m) / dx;
                    acc[j - 1] = acc[j - 1] + (b - a) / 6.0 * (ya * ha + 4.0 * ym * hm + yb * hb);
                    ha = (a - g[j - 1]) / dx; hb = (b - g[j - 1]) / dx; hm = (m - g[j - 1]) / dx;
                    acc[j] = acc[j] + (b - a) / 6.0 * (ya * ha + 4.0 * ym * hm + yb * hb);
                }
                j = j + 1;
            }
        }
        band[0] = acc[0] / (0.5 * (g[1] - g[0]));
        for (int j = 2; j <= ng - 1; j++) band[j - 1] = acc[j - 1] / (0.5 * (g[j] - g[j - 2]));
        band[ng - 1] = acc[ng - 1] / (0.5 * (g[ng - 1] - g[ng - 2]));
        free(acc);
        for (int j = 0; j < ng; j++) if (band[j] < 0.0) band[j] = 0.0;
    } else {
        int si, ei;
        if (lc[0] <= g[0]) si = 1; else { si = orc_find_interval(g, ng, lc[0]); if (si < 1) si = 1; }
        if (lc[nc - 1] >= g[ng - 1]) ei = ng; else { ei = orc_find_interval(g, ng, lc[nc - 1]); if (ei > ng) ei = ng; }
        if (ei > si)
            for (int i = si; i <= ei; i++) band[i - 1] = interp_linear(lc, rc, nc, g[i - 1]);
    }
    if (detector == 1) for (int j = 0; j < ng; j++) band[j] = band[j] / g[j];
    for (int j = 0; j < ng; j++) if (band[j] < 0.0) band[j] = 0.0;
    /* integrate_trapezoid_array(grid, band / (grid + SAFE_FLOOR)) */
    double norm = 0.0;
    for (int j = 0; j < ng - 1; j++)
        norm = norm + 0.5 * (g[j + 1] - g[j]) * (band[j] / (g[j] + SAFE_FLOOR) + band[j + 1] / (g[j + 1] + SAFE_FLOOR));
    if (norm > SAFE_FLOOR) {
        for (int j = 0; j < ng; j++) band[j] = band[j] / norm;
        return 0;
    }
    for (int j = 0; j < ng; j++) band[j] = 0.0;
    return 1;
}

/* ---------------------------------------------------------------------------------------------
 * IMF integrals and bin weights (src/physics/opensps_imf.f90). The packed IMF parameters p[] of
 * this restatement: the reference's IMF_P_* slots that the power-law types read.
 * ------------------------------------------------------------------------------------------- */
#define ORC_IMF_SALP 0
#define ORC_IMF_ALPHA 1     /* 1..3: imf1, imf2, imf3 */
#define ORC_IMF_VDMC 4
#define ORC_IMF_MDAVE 5
#define ORC_IMF_LOWER_LIMIT 6
#define ORC_IMF_UPPER_LIMIT 7
#define ORC_IMF_LOWER_BOUND 8
#define ORC_IMF_MLIM_BH 9
#define ORC_IMF_MLIM_NS 10
#define ORC_IMF_NPAR 11

/* powerlaw_bin (opensps_imf.f90:754-792): int_lo^hi m^(e-1) dm given plo = lo^e, as
 * plo expm1(e dl) / e with dl = ln(hi/lo) in Kahan's log1p form; series below |e dl| = 1/4 */
double orc_powerlaw_bin(double plo, double lo, double hi, double e) {
    const double I2 = 1.0 / 2, I3 = 1.0 / 3, I4 = 1.0 / 4, I5 = 1.0 / 5, I6 = 1.0 / 6, I7 = 1.0 / 7, I8 = 1.0 / 8,
                 I9 = 1.0 / 9, I10 = 1.0 / 10, I11 = 1.0 / 11, I12 = 1.0 / 12, I13 = 1.0 / 13;
    double x = (hi - lo) / lo, u = 1.0 + x, dl;
    if (u == 1.0) dl = x;
    else dl = log(u) * x / (u - 1.0);
    if (e == 0.0) return plo * dl;
    double y = e * dl;
    if (fabs(y) < 0.25) {
        double r = 1.0 + y * I13;
        r = 1.0 + y * I12 * r; r = 1.0 + y * I11 * r; r = 1.0 + y * I10 * r; r = 1.0 + y * I9 * r;
        r = 1.0 + y * I8 * r; r = 1.0 + y * I7 * r; r = 1.0 + y * I6 * r; r = 1.0 + y * I5 * r;
        r = 1.0 + y * I4 * r; r = 1.0 + y * I3 * r; r = 1.0 + y * I2 * r;
        return plo * dl * r;
    }
    return plo * (exp(y) - 1.0) / e;
}

/* expm1_series (opensps_imf.f90:641-667) */
static double expm1_series(double y) {
    const double I2 = 1.0 / 2, I3 = 1.0 / 3, I4 = 1.0 / 4, I5 = 1.0 / 5, I6 = 1.0 / 6, I7 = 1.0 / 7, I8 = 1.0 / 8,
                 I9 = 1.0 / 9, I10 = 1.0 / 10, I11 = 1.0 / 11, I12 = 1.0 / 12, I13 = 1.0 / 13;
    if (fabs(y) < 0.25) {
        double r = 1.0 + y * I13;
        r = 1.0 + y * I12 * r; r = 1.0 + y * I11 * r; r = 1.0 + y * I10 * r; r = 1.0 + y * I9 * r;
        r = 1.0 + y * I8 * r; r = 1.0 + y * I7 * r; r = 1.0 + y * I6 * r; r = 1.0 + y * I5 * r;
        r = 1.0 + y * I4 * r; r = 1.0 + y * I3 * r; r = 1.0 + y * I2 * r;
        return y * r;
    }
    return exp(y) - 1.0;
}

/* integrate_powerlaw (opensps_imf.f90:725-748) */
static double integrate_powerlaw(double lo, double hi, double alpha, int mass_weighted, double coeff) {
    if (hi <= lo) return 0.0;
    double denom = mass_weighted ? 2.0 - alpha : 1.0 - alpha;
    return coeff * orc_powerlaw_bin(pow(lo, denom), lo, hi, denom);
}

/* imf_integral / imf_integral_p (opensps_imf.f90:86-96, 206-315) for the power-law types: 0 Salpeter,
 * 2 Kroupa (breaks 0.08, 0.5, 1), 4 Dave (break mdave). Other types integrate to 0 here (the batch
 * refuses them). */
double orc_imf_integral(int imf_type, const double *p, double m1, double m2, int mass_weighted) {
    double lo = fmin(m1, m2), hi = fmax(m1, m2), v = 0.0;
    if (hi <= lo) return 0.0;
    int t = imf_type % 10;
    if (t == 0) {
        v = integrate_powerlaw(lo, hi, p[ORC_IMF_SALP], mass_weighted, 1.0);
    } else if (t == 2) {
        double a1 = p[ORC_IMF_ALPHA], a2 = p[ORC_IMF_ALPHA + 1], a3 = p[ORC_IMF_ALPHA + 2];
        double c2 = pow(0.5, -a1 + a2), sl, sh;
        sl = fmax(lo, 0.08); sh = fmin(hi, 0.5);
        if (sh > sl) v = v + integrate_powerlaw(sl, sh, a1, mass_weighted, 1.0);
        sl = fmax(lo, 0.5); sh = fmin(hi, 1.0);
        if (sh > sl) v = v + integrate_powerlaw(sl, sh, a2, mass_weighted, c2);
        sl = fmax(lo, 1.0); sh = hi;
        if (sh > sl) v = v + integrate_powerlaw(sl, sh, a3, mass_weighted, c2);
    } else if (t == 4) {
        double a1 = p[ORC_IMF_ALPHA], a2 = p[ORC_IMF_ALPHA + 1], md = p[ORC_IMF_MDAVE], sl, sh;
        sl = fmax(lo, 0.08); sh = fmin(hi, md);
        if (sh > sl) v = v + integrate_powerlaw(sl, sh, a1, mass_weighted, 1.0);
        sl = fmax(lo, md); sh = hi;
        if (sh > sl) v = v + integrate_powerlaw(sl, sh, a2, mass_weighted, pow(md, -a1 + a2));
    }
    return v;
}

/* imf_bin_logs (opensps_imf.f90:564-579): llo = ln m1, dl = ln(m2 / m1) in Kahan's log1p form */
void orc_imf_bin_logs(double m1, double m2, double *llo, double *dl) {
    *llo = 0.0; *dl = 0.0;
    if (m1 <= 0.0 || m2 <= m1) return;
    *llo = log(m1);
    double x = (m2 - m1) / m1, u = 1.0 + x;
    *dl = (u == 1.0) ? x : log(u) * x / (u - 1.0);
}

/* imf_bin_weights_tab (opensps_imf.f90:439-559), FP64 build: the weights of a run of consecutive bins
 * (one piece of an age), stars per bin per unit initial mass; a bin inside one power-law segment is
 * c lo^e expm1(e dl) / e, lo^e carried from the previous bin when they share an edge. */
void orc_imf_bin_weights_tab(int imf_type, const double *p, double total_mass, int n, const double *mini,
                             const double *m1, const double *m2, const double *llo, const double *dl, const int *first,
                             double *w) {
    int t = imf_type % 10, cur = 0;
    double lim_lo = p[ORC_IMF_LOWER_LIMIT], lim_hi = p[ORC_IMF_UPPER_LIMIT];
    double e = 0.0, c = 0.0, d = 1.0, ce = 0.0, prev_hi = -1.0, prev_phi = 0.0;
    double inv_total = total_mass > 0.0 ? 1.0 / total_mass : 1.0;
    for (int k = 0; k < n; k++) {
        w[k] = 0.0;
        if (mini[k] < lim_lo || mini[k] > lim_hi) continue;
        double lo = first[k] == 1 ? p[ORC_IMF_LOWER_BOUND] : m1[k], hi = m2[k], val;
        if (hi <= lo) continue;
        int seg = 0;
        if (t == 0) {
            seg = 1;
        } else if (t == 2) {
            if (lo >= 0.08 && hi <= 0.5) seg = 1;
            else if (lo >= 0.5 && hi <= 1.0) seg = 2;
            else if (lo >= 1.0) seg = 3;
        } else if (t == 4) {
            if (lo >= 0.08 && hi <= p[ORC_IMF_MDAVE]) seg = 1;
            else if (lo >= p[ORC_IMF_MDAVE]) seg = 2;
        }
        if (seg > 0 && seg != cur) {
            double alpha, coeff;
            if (t == 0) {
                alpha = p[ORC_IMF_SALP]; coeff = 1.0;
            } else if (t == 2) {
                alpha = p[ORC_IMF_ALPHA + seg - 1]; coeff = 1.0;
                if (seg > 1) coeff = pow(0.5, -p[ORC_IMF_ALPHA] + p[ORC_IMF_ALPHA + 1]);
            } else {
                alpha = p[ORC_IMF_ALPHA + seg - 1]; coeff = 1.0;
                if (seg > 1) coeff = pow(p[ORC_IMF_MDAVE], -p[ORC_IMF_ALPHA] + p[ORC_IMF_ALPHA + 1]);
            }
            e = 1.0 - alpha; c = coeff; d = 1.0 - alpha;
            ce = 0.0;
            if (e != 0.0) ce = c / e;
            cur = seg;
            prev_hi = -1.0;
        }
        if (seg > 0 && cur == seg) {
            if (first[k] == 1) {
                if (fabs(d) <= 1.0e-12) val = c * log(hi / lo);
                else val = c * (pow(hi, e) - pow(lo, e)) / d;
            } else if (e == 0.0) {
                val = c * dl[k];
            } else {
                double plo = (lo == prev_hi) ? prev_phi : exp(e * llo[k]);
                double em1 = expm1_series(e * dl[k]);
                val = ce * plo * em1;
                prev_hi = hi;
                prev_phi = plo + plo * em1;
            }
        } else {
            val = orc_imf_integral(imf_type, p, lo, hi, 0);
            prev_hi = -1.0;
            if (seg == 0) cur = 0;
        }
        w[k] = val * inv_total;
    }
}

/* add_remnant_mass_p (opensps_stellar_modifications.f90:729-816): black holes (half the initial
 * mass above max(mlim_bh, the largest living mass)), neutron stars (1.4 M_sun), white dwarfs
 * (0.48 + 0.077 m) from the stars that have died, per unit initial mass */
double orc_remnant_mass(int imf_type, const double *p, double max_living_mass) {
    double lim_hi = p[ORC_IMF_UPPER_LIMIT], lbh = p[ORC_IMF_MLIM_BH], lns = p[ORC_IMF_MLIM_NS], add = 0.0;
    double norm = orc_imf_integral(imf_type, p, p[ORC_IMF_LOWER_LIMIT], lim_hi, 1);
    if (norm <= 0.0) return 0.0;
    double a = fmin(fmax(lbh, max_living_mass), lim_hi);
    if (a < lim_hi) add = add + 0.5 * orc_imf_integral(imf_type, p, a, lim_hi, 1) / norm;
    if (max_living_mass <= lbh) {
        a = fmin(fmax(lns, max_living_mass), lim_hi);
        double b = fmin(lbh, lim_hi);
        if (a < b) add = add + 1.4 * orc_imf_integral(imf_type, p, a, b, 0) / norm;
    }
    if (max_living_mass <= lns) {
        a = fmin(max_living_mass, lim_hi);
        double b = fmin(lns, lim_hi);
        if (a < b)
            add = add + (0.48 * orc_imf_integral(imf_type, p, a, b, 0)) / norm +
                  (0.077 * orc_imf_integral(imf_type, p, a, b, 1)) / norm;
    }
    return add;
}

/* ---------------------------------------------------------------------------------------------
 * The SSP operator set with the IMF per model (src/physics/opensps_ssp_batch.f90): the isochrone
 * points of every (metallicity, age) with their mass bins and per-star mass, L_bol and ionizing
 * photons; each point's template entries; the template library (ntpl x nspec).
 * ------------------------------------------------------------------------------------------- */
typedef struct {
    int nz, nt, npts, ntpl, nent;
    const int *age_ptr;       /* nz * nt + 1: points of (iz, t) are [age_ptr[iz nt + t], age_ptr[iz nt + t + 1]) */
    const double *mini, *m1, *m2;
    const int *first;
    const double *cm, *cl, *q;
    const double *age_mmax;   /* nz * nt: the largest living initial mass of each isochrone */
    const int *ent_tpl;       /* npts * nent, template of each entry (-1: none) */
    const double *ent_val;    /* npts * nent */
    const double *etpl;       /* ntpl x nspec */
} orc_ops_t;

#define ORC_PIECE 32   /* points per piece (the reference's pieces of an age, pc_ptr, <= 32 points) */

/* One model's operator age sums, ssp_batch_age_sums_z (opensps_ssp_batch.f90:527-785) with
 * imf_pieces_multi (:2913-3091) and fold_multi (:3095-3188): two passes, the operators of
 * metallicities zlo (weight 1 - dz) and zlo + 1 (weight dz; one pass with one metallicity); per age
 * the IMF weights of its points piece by piece, their mass / L_bol / photon sums (remnants added),
 * and the fold of the weighted points into template weights split at it_tesc; then the template
 * products. Outputs: young, old (nspec), smass, slum (nt, slum linear), vy, vo (ntpl scratch),
 * w (scratch, the largest age's points). */
void orc_ops_model(const orc_ops_t *op, int nspec, int imf_type, int remnants, const double *pimf, int zlo, double dz,
                   const double *wsfh, int it_tesc, double *young, double *old, double *smass, double *slum,
                   double *vy, double *vo, double *w, double *llo, double *dl, int *count) {
    const int nt = op->nt, nent = op->nent;
    int npass = op->nz > 1 ? 2 : 1;
    double tot = orc_imf_integral(imf_type, pimf, pimf[ORC_IMF_LOWER_LIMIT], pimf[ORC_IMF_UPPER_LIMIT], 1);
    for (int t = 0; t < nt; t++) { smass[t] = 0.0; slum[t] = 0.0; }
    for (int e = 0; e < op->ntpl; e++) { vy[e] = 0.0; vo[e] = 0.0; }
    if (count) { count[0] = 0; count[1] = 0; }
    for (int pass = 0; pass < npass; pass++) {
        int iz = (npass == 1) ? 0 : zlo - 1 + pass;
        double cz = (npass == 1) ? 1.0 : (pass == 0 ? 1.0 - dz : dz);
        if (cz == 0.0) continue;   /* group_models: models without weight in a pass are left out */
        for (int t = 0; t < nt; t++) {
            int k0 = op->age_ptr[iz * nt + t], k1 = op->age_ptr[iz * nt + t + 1];
            double sm = 0.0, sl = 0.0, sq = 0.0;
            for (int k = k0; k < k1; k++) orc_imf_bin_logs(op->m1[k], op->m2[k], &llo[k - k0], &dl[k - k0]);
            for (int q0 = k0; q0 < k1; q0 += ORC_PIECE) {
                int np = k1 - q0 < ORC_PIECE ? k1 - q0 : ORC_PIECE;
                orc_imf_bin_weights_tab(imf_type, pimf, tot, np, op->mini + q0, op->m1 + q0, op->m2 + q0,
                                        llo + (q0 - k0), dl + (q0 - k0), op->first + q0, w + (q0 - k0));
                double pm = 0.0, pl = 0.0, pq = 0.0;
                for (int k = q0; k < q0 + np; k++) {
                    if (count && w[k - k0] > 0.0) count[pass]++;
                    pm = pm + op->cm[k] * w[k - k0];
                    pl = pl + op->cl[k] * w[k - k0];
                    pq = pq + op->q[k] * w[k - k0];
                }
                sm = sm + pm; sl = sl + pl; sq = sq + pq;
            }
            if (remnants && k1 > k0) sm = sm + orc_remnant_mass(imf_type, pimf, op->age_mmax[iz * nt + t]);
            smass[t] = smass[t] + cz * sm;
            slum[t] = slum[t] + cz * log10(fmax(sl, SAFE_FLOOR));
            /* the fold: weighted points into the template weights, young (t <= it_tesc) or old */
            double x0 = wsfh[t] * cz;
            double *v = (t + 1 <= it_tesc) ? vy : vo;
            for (int k = k0; k < k1; k++) {
                double x = w[k - k0] * x0;
                for (int j = 0; j < nent; j++) {
                    int e = op->ent_tpl[(size_t)k * nent + j];
                    if (e >= 0) v[e] = v[e] + op->ent_val[(size_t)k * nent + j] * x;
                }
            }
        }
    }
    for (int t = 0; t < nt; t++) slum[t] = pow(10.0, slum[t]);
    /* template products (template_columns, :3247-3392) */
    for (int i = 0; i < nspec; i++) { young[i] = 0.0; old[i] = 0.0; }
    for (int e = 0; e < op->ntpl; e++) {
        const double *s = op->etpl + (size_t)e * nspec;
        double a = vy[e], b = vo[e];
        if (a != 0.0) for (int i = 0; i < nspec; i++) young[i] = young[i] + a * s[i];
        if (b != 0.0) for (int i = 0; i < nspec; i++) old[i] = old[i] + b * s[i];
    }
}

/* ---------------------------------------------------------------------------------------------
 * The IMF's derivatives along its per-model parameters (the rows imf1, imf2, imf3 / mdave of the
 * operator's reverse pass, ssp_batch_age_sums_adjoint, opensps_ssp_batch.f90:991-1248, whose
 * adj_imf_analytic takes the power-law types 2 and 4 analytically). Written here from the integrals
 * themselves: a segment xi(m) = c m^-alpha contributes int m^(s-1) dm over its part of [lo, hi]
 * (s = 1 - alpha for numbers, 2 - alpha for mass); d/d alpha of it is -int m^(s-1) ln m dm, the
 * coefficient's own derivative multiplies int m^(s-1) dm, and a break inside the interval moves two
 * limits whose terms cancel where the IMF is continuous. The moments come from Gauss-Legendre
 * quadrature in ln m (not from the reference's closed forms), so this restatement checks them.
 * Directions: Kroupa (2) imf1, imf2, imf3; Dave (4) imf1, imf2, mdave; the others none.
 * ------------------------------------------------------------------------------------------- */
#define ORC_GL_N 20
static double gl_x[ORC_GL_N], gl_w[ORC_GL_N];
static int gl_ready = 0;

/* Gauss-Legendre nodes and weights on [-1, 1]: Newton on P_n from the Chebyshev guesses */
static void gl_init(void) {
    if (gl_ready) return;
    const int n = ORC_GL_N;
    for (int i = 0; i < n; i++) {
        double x = cos(acos(-1.0) * (i + 0.75) / (n + 0.5)), dp = 1.0;
        for (int it = 0; it < 100; it++) {
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= n; k++) {
                double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (x * p1 - p0) / (x * x - 1.0);
            double dx = p1 / dp;
            x -= dx;
            if (fabs(dx) < 1e-16) break;
        }
        gl_x[i] = x;
        gl_w[i] = 2.0 / ((1.0 - x * x) * dp * dp);
    }
    gl_ready = 1;
}

/* int_lo^hi m^(s-1) (ln m)^q dm = int e^(s u) u^q du over [ln lo, ln hi], q = 0 or 1, in pieces of at most
 * 0.25 in u of ORC_GL_N nodes each */
static double moment_ln(double lo, double hi, double s, int q) {
    if (!(hi > lo) || lo <= 0.0) return 0.0;
    const double u0 = log(lo), u1 = log(hi);
    int np = (int)ceil((u1 - u0) / 0.25);
    if (np < 1) np = 1;
    const double h = (u1 - u0) / np;
    double acc = 0.0;
    for (int j = 0; j < np; j++) {
        const double a = u0 + j * h, mid = a + 0.5 * h, half = 0.5 * h;
        double part = 0.0;
        for (int i = 0; i < ORC_GL_N; i++) {
            const double u = mid + half * gl_x[i];
            part += gl_w[i] * exp(s * u) * (q ? u : 1.0);
        }
        acc += half * part;
    }
    return acc;
}

/* d/d(the type's three directions) of orc_imf_integral(imf_type, p, m1, m2, mw) */
static void imf_integral_grad(int imf_type, const double *p, double m1, double m2, int mw, double *d) {
    d[0] = 0.0; d[1] = 0.0; d[2] = 0.0;
    const double lo = fmin(m1, m2), hi = fmax(m1, m2), s0 = mw ? 2.0 : 1.0;
    if (!(hi > lo)) return;
    const int t = imf_type % 10;
    if (t == 2) {
        const double a1 = p[ORC_IMF_ALPHA], a2 = p[ORC_IMF_ALPHA + 1], a3 = p[ORC_IMF_ALPHA + 2];
        const double l2 = log(0.5), c2 = pow(0.5, -a1 + a2);
        double sl = fmax(lo, 0.08), sh = fmin(hi, 0.5);
        if (sh > sl) d[0] += -moment_ln(sl, sh, s0 - a1, 1);
        sl = fmax(lo, 0.5); sh = fmin(hi, 1.0);
        if (sh > sl) {
            const double i0 = moment_ln(sl, sh, s0 - a2, 0), i1 = moment_ln(sl, sh, s0 - a2, 1);
            d[0] += -l2 * c2 * i0;
            d[1] += l2 * c2 * i0 - c2 * i1;
        }
        sl = fmax(lo, 1.0); sh = hi;
        if (sh > sl) {
            const double i0 = moment_ln(sl, sh, s0 - a3, 0), i1 = moment_ln(sl, sh, s0 - a3, 1);
            d[0] += -l2 * c2 * i0;
            d[1] += l2 * c2 * i0;
            d[2] += -c2 * i1;
        }
    } else if (t == 4) {
        const double a1 = p[ORC_IMF_ALPHA], a2 = p[ORC_IMF_ALPHA + 1], md = p[ORC_IMF_MDAVE];
        const double lm = log(md), c = pow(md, -a1 + a2);
        double sl = fmax(lo, 0.08), sh = fmin(hi, md);
        if (sh > sl) {
            d[0] += -moment_ln(sl, sh, s0 - a1, 1);
            if (hi > md) d[2] += pow(md, s0 - a1 - 1.0);          /* the upper limit at the break */
        }
        sl = fmax(lo, md); sh = hi;
        if (sh > sl) {
            const double i0 = moment_ln(sl, sh, s0 - a2, 0), i1 = moment_ln(sl, sh, s0 - a2, 1);
            d[0] += -lm * c * i0;
            d[1] += lm * c * i0 - c * i1;
            d[2] += (a2 - a1) / md * c * i0;
            if (lo < md) d[2] -= c * pow(md, s0 - a2 - 1.0);      /* the lower limit at the break */
        }
    }
}

/* d/d(directions) of orc_remnant_mass: each term's integral over the IMF's mass normalisation */
static void remnant_mass_grad(int imf_type, const double *p, double mmax, double *d) {
    d[0] = 0.0; d[1] = 0.0; d[2] = 0.0;
    const double hi = p[ORC_IMF_UPPER_LIMIT], lbh = p[ORC_IMF_MLIM_BH], lns = p[ORC_IMF_MLIM_NS];
    const double norm = orc_imf_integral(imf_type, p, p[ORC_IMF_LOWER_LIMIT], hi, 1);
    if (norm <= 0.0) return;
    double dn[3], g[3], num = 0.0, dnum[3] = {0.0, 0.0, 0.0};
    imf_integral_grad(imf_type, p, p[ORC_IMF_LOWER_LIMIT], hi, 1, dn);
    double a = fmin(fmax(lbh, mmax), hi);
    if (a < hi) {
        num += 0.5 * orc_imf_integral(imf_type, p, a, hi, 1);
        imf_integral_grad(imf_type, p, a, hi, 1, g);
        for (int i = 0; i < 3; i++) dnum[i] += 0.5 * g[i];
    }
    if (mmax <= lbh) {
        a = fmin(fmax(lns, mmax), hi);
        const double b = fmin(lbh, hi);
        if (a < b) {
            num += 1.4 * orc_imf_integral(imf_type, p, a, b, 0);
            imf_integral_grad(imf_type, p, a, b, 0, g);
            for (int i = 0; i < 3; i++) dnum[i] += 1.4 * g[i];
        }
    }
    if (mmax <= lns) {
        a = fmin(mmax, hi);
        const double b = fmin(lns, hi);
        if (a < b) {
            num += 0.48 * orc_imf_integral(imf_type, p, a, b, 0) + 0.077 * orc_imf_integral(imf_type, p, a, b, 1);
            imf_integral_grad(imf_type, p, a, b, 0, g);
            for (int i = 0; i < 3; i++) dnum[i] += 0.48 * g[i];
            imf_integral_grad(imf_type, p, a, b, 1, g);
            for (int i = 0; i < 3; i++) dnum[i] += 0.077 * g[i];
        }
    }
    for (int i = 0; i < 3; i++) d[i] = dnum[i] / norm - num * dn[i] / (norm * norm);
}

/* The reverse of orc_ops_model for one model (ssp_batch_age_sums_adjoint, opensps_ssp_batch.f90:991-1248, with
 * adj_template_bar, adj_fold_gather, adj_lq, adj_mass_dot and adj_imf_analytic): given the cotangents yb, ob of the
 * young and old age sums, lsb of the linear L_bol sum (sum_t wsfh_t slum_t) and mb of the mass (sum_t wsfh_t
 * smass_t), and the forward's smass, slum (orc_ops_model's outputs):
 *   vyb, vob = etpl yb, etpl ob                         (the transposed template products)
 *   g_k = sum_j ent_val(k, j) v_b(ent_tpl(k, j))          (the fold's transpose; young or old by the age)
 *   wfold_t = sum_pass cz sum_k w_k g_k                   (d L / d wsfh_t through the fold)
 *   dzb = sum_pass sign sum_t wsfh_t (sum_k w_k g_k + mb sm_t + lsb slum_t ln sl_t)
 *   imfb_i = sum_pass cz sum_t wsfh_t (sum_k dw_k/di (g_k + mb cm_k + lsb slum_t cl_k / sl_t) + mb drem_t/di)
 * with w_k = I_k / total, dw_k = dI_k / total - w_k dtotal / total. Passes of zero weight are left out, as the
 * forward leaves them (group_models). Scratch: vyb, vob (ntpl), w, dw (3 x the largest age's points), llo, dl. */
void orc_ops_model_adj(const orc_ops_t *op, int nspec, int imf_type, int remnants, const double *pimf, int zlo,
                       double dz, const double *wsfh, int it_tesc, const double *yb, const double *ob, double lsb,
                       double mb, const double *slum, double *wfold, double *dzb, double *imfb, double *vyb,
                       double *vob, double *w, double *dw, double *llo, double *dl) {
    const int nt = op->nt, nent = op->nent;
    const int npass = op->nz > 1 ? 2 : 1;
    gl_init();
    const double tot = orc_imf_integral(imf_type, pimf, pimf[ORC_IMF_LOWER_LIMIT], pimf[ORC_IMF_UPPER_LIMIT], 1);
    double dtot[3];
    imf_integral_grad(imf_type, pimf, pimf[ORC_IMF_LOWER_LIMIT], pimf[ORC_IMF_UPPER_LIMIT], 1, dtot);
    const double inv_total = tot > 0.0 ? 1.0 / tot : 1.0;
    for (int e = 0; e < op->ntpl; e++) {
        const double *s = op->etpl + (size_t)e * nspec;
        double a = 0.0, b = 0.0;
        for (int i = 0; i < nspec; i++) { a += s[i] * yb[i]; b += s[i] * ob[i]; }
        vyb[e] = a;
        vob[e] = b;
    }
    for (int t = 0; t < nt; t++) wfold[t] = 0.0;
    *dzb = 0.0;
    imfb[0] = 0.0; imfb[1] = 0.0; imfb[2] = 0.0;
    for (int pass = 0; pass < npass; pass++) {
        const int iz = (npass == 1) ? 0 : zlo - 1 + pass;
        const double cz = (npass == 1) ? 1.0 : (pass == 0 ? 1.0 - dz : dz), sgn = pass == 0 ? -1.0 : 1.0;
        if (cz == 0.0) continue;
        for (int t = 0; t < nt; t++) {
            const int k0 = op->age_ptr[iz * nt + t], k1 = op->age_ptr[iz * nt + t + 1];
            const double *vb = (t + 1 <= it_tesc) ? vyb : vob;
            for (int k = k0; k < k1; k++) orc_imf_bin_logs(op->m1[k], op->m2[k], &llo[k - k0], &dl[k - k0]);
            for (int q0 = k0; q0 < k1; q0 += ORC_PIECE) {
                const int np = k1 - q0 < ORC_PIECE ? k1 - q0 : ORC_PIECE;
                orc_imf_bin_weights_tab(imf_type, pimf, tot, np, op->mini + q0, op->m1 + q0, op->m2 + q0,
                                        llo + (q0 - k0), dl + (q0 - k0), op->first + q0, w + (q0 - k0));
            }
            double sm = 0.0, sl = 0.0, fk = 0.0;
            for (int k = k0; k < k1; k++) {
                sm += op->cm[k] * w[k - k0];
                sl += op->cl[k] * w[k - k0];
            }
            /* the bins' derivatives (zero where the limits or an empty bin give no weight) */
            for (int k = k0; k < k1; k++) {
                double *d = dw + 3 * (k - k0);
                d[0] = 0.0; d[1] = 0.0; d[2] = 0.0;
                const double lo = op->first[k] == 1 ? pimf[ORC_IMF_LOWER_BOUND] : op->m1[k], hi = op->m2[k];
                if (op->mini[k] < pimf[ORC_IMF_LOWER_LIMIT] || op->mini[k] > pimf[ORC_IMF_UPPER_LIMIT] || hi <= lo)
                    continue;
                double di[3];
                imf_integral_grad(imf_type, pimf, lo, hi, 0, di);
                for (int i = 0; i < 3; i++) d[i] = di[i] * inv_total - w[k - k0] * dtot[i] * inv_total;
            }
            if (remnants && k1 > k0) sm += orc_remnant_mass(imf_type, pimf, op->age_mmax[iz * nt + t]);
            const double sl_on = sl > SAFE_FLOOR ? 1.0 : 0.0;
            const double lfac = sl_on * lsb * slum[t] / (sl > SAFE_FLOOR ? sl : 1.0);
            for (int k = k0; k < k1; k++) {
                double g = 0.0;
                for (int j = 0; j < nent; j++) {
                    const int e = op->ent_tpl[(size_t)k * nent + j];
                    if (e >= 0) g += op->ent_val[(size_t)k * nent + j] * vb[e];
                }
                fk += w[k - k0] * g;
                const double ge = g + mb * op->cm[k] + lfac * op->cl[k];
                for (int i = 0; i < 3; i++) imfb[i] += cz * wsfh[t] * ge * dw[3 * (k - k0) + i];
            }
            if (remnants && k1 > k0) {
                double dr[3];
                remnant_mass_grad(imf_type, pimf, op->age_mmax[iz * nt + t], dr);
                for (int i = 0; i < 3; i++) imfb[i] += cz * wsfh[t] * mb * dr[i];
            }
            wfold[t] += cz * fk;
            *dzb += sgn * wsfh[t] * (fk + mb * sm + lsb * slum[t] * log(fmax(sl, SAFE_FLOOR)));
        }
    }
    if (npass == 1) *dzb = 0.0;
}

/* ---------------------------------------------------------------------------------------------
 * The batch (src/physics/opensps_csp_batch.f90:431-1214, the stages that the configurations of
 * this build exercise: parametric or given SFH weights, the resident bases blended in logzsol,
 * power-law dust with birth cloud and old-star screens, per-model IGM, velocity smoothing,
 * redshift and band magnitudes).
 * ------------------------------------------------------------------------------------------- */
typedef struct {
    int nspec, nt, nz, nbands;
    const double *lambda;     /* nspec */
    const double *time_full;  /* nt, log10 yr */
    const double *logz;       /* nz, log10(Z / Z_sun) */
    const double *basis;      /* nspec x nt x nz, L_sun/Hz per solar mass */
    const double *bmass;      /* nt x nz, surviving mass */
    const double *blbol;      /* nt x nz, log10 L_bol */
    const double *bands;      /* nspec x nbands, normalised */
    const int *mag_mask;      /* nbands or NULL */
    int ncosmo;
    const double *cz, *cdl;   /* the cosmology table */
    int add_igm;              /* add_igm_absorption */
    int smooth_fast;          /* smooth_velocity == 1 and smoothspec_fast == 1 */
    double tiny_logt;
    /* the IMF per model (the operator set, or NULL: the resident bases) */
    const orc_ops_t *ops;
    int imf_type, remnants;   /* imf_type, add_stellar_remnants */
    double salp_ind, mdave, mlim_bh, mlim_ns;
    int *op_count;            /* 2 x n or NULL: points with a nonzero IMF weight per metallicity pass */
    /* weights over ages and metallicities (ssp_weights_z: nt x nz x n) or NULL; they replace the SFH and the
     * metallicity blend (age_sums_w / model_sums_zw, opensps_csp_batch.f90:2086-2127, 2269-2324) */
    const double *ssp_weights_z;
    /* a fixed diffuse attenuation curve A_lambda / A_V per pixel (dust_type 2, 6: attenuation_table_row's column 0,
     * opensps_csp_batch.f90:3525-3566) or NULL: the power law */
    const double *dcurve;
    /* each model's magnitude redshift (compute_csp_batch's zred_mags, opensps_csp_batch.f90:1054-1061) or NULL:
     * the observed frame and the distance term at it, the IGM at zred; the reverse pass then leaves the observed
     * frame's and the distance term's dz out of zred's row (batch_backward's lzm, :4830-4832, 4928) */
    const double *zred_mags;
    /* dust emission (add_dust_emission) with the Draine & Li templates on the grid (dustem2_dustem: nspec x nqpah x
     * 2 numin, column-major; qpah, umin ascending) and the AGN torus (add_agn_dust) with its templates (agndust_spec:
     * nspec x ntau; agn_tau ascending); nqpah / ntau 0: none; mdust_out (n) or NULL */
    int add_dust_emission, add_agn_dust;
    int nqpah, numin, ntau;
    const double *qpah, *umin, *dustem, *agn_tau, *agn;
    double *mdust_out;
    /* nebular emission (add_neb_emission = 1, the CLOUDY grids) with its tables on the grid as load_nebular_grid
     * leaves them (src/io/opensps_io.f90:749-823): neb_cont nspec x nnz x nna x nnu and neb_line nline x nnz x nna x
     * nnu, column-major, log10(L + 1e-95) per unit ionizing photon rate (the continuum linear in lambda onto the
     * grid); the axes ascending: gas log Z, log10 age (yr), log U; neb_lpos the lines' wavelengths. neb_cont_on:
     * add_neb_continuum; neb_lines_in_spec: nebemlineinspec. emlines_out (nline x n) or NULL */
    int add_neb, neb_cont_on, neb_lines_in_spec;
    int nnz, nna, nnu, nline;
    const double *neb_logz, *neb_lage, *neb_logu, *neb_cont, *neb_line, *neb_lpos;
    double *emlines_out;
} orc_ctx_t;

/* one_minus_trans_kp (opensps_csp_batch.f90:3925-3939), FP64: 1 - exp(-x) without cancellation */
static double one_minus_trans(double x, double t) {
    if (t == 1.0) return x;
    if (fabs(x) < 1.0) return (1.0 - t) * x / (-log(t));
    return 1.0 - t;
}

/* kadd: Kahan's compensated sum (the reference's CPU build of dust_emission_reductions, :4026-4030) */
static void kadd(double *s, double *c, double v) {
    double y = v - *c, t = *s + y;
    *c = (t - *s) - y;
    *s = t;
}

/* dl07_indices (opensps_csp_batch.f90:1611-1626) and dl07_shape (:4103-4122): the Draine & Li template at pixel i,
 * bilinear in q_PAH and U_min (intervals clamped to the grid, weights not), the two U_max spectra mixed by gamma
 * (clamped to [0, 1]), floored at the smallest normal */
static void dl07_indices(const orc_ctx_t *c, const double *p, int *iq, double *wq, int *iu, double *wu, double *gam) {
    int q = orc_find_interval(c->qpah, c->nqpah, p[OPS_P_DUSTE_QPAH]);
    if (q > c->nqpah - 1) q = c->nqpah - 1;
    if (q < 1) q = 1;
    *iq = q;
    *wq = (p[OPS_P_DUSTE_QPAH] - c->qpah[q - 1]) / (c->qpah[q] - c->qpah[q - 1]);
    int u = orc_find_interval(c->umin, c->numin, p[OPS_P_DUSTE_UMIN]);
    if (u > c->numin - 1) u = c->numin - 1;
    if (u < 1) u = 1;
    *iu = u;
    *wu = (p[OPS_P_DUSTE_UMIN] - c->umin[u - 1]) / (c->umin[u] - c->umin[u - 1]);
    *gam = fmax(0.0, fmin(p[OPS_P_DUSTE_GAMMA], 1.0));
}

/* the table's value at pixel i (0-based), q_PAH column q and spectrum column k (1-based, dustem2_dustem(i, q, k)) */
static double dustem_at(const orc_ctx_t *c, int i, int q, int k) {
    return c->dustem[((size_t)(q - 1) + (size_t)c->nqpah * (k - 1)) * c->nspec + i];
}

static double dl07_shape(const orc_ctx_t *c, int i, int iq, double wq, int iu, double wu, double gam) {
    double lo = (1 - wq) * (1 - wu) * dustem_at(c, i, iq, 2 * iu - 1) + wq * (1 - wu) * dustem_at(c, i, iq + 1, 2 * iu - 1) +
                wq * wu * dustem_at(c, i, iq + 1, 2 * (iu + 1) - 1) + (1 - wq) * wu * dustem_at(c, i, iq, 2 * (iu + 1) - 1);
    double hi = (1 - wq) * (1 - wu) * dustem_at(c, i, iq, 2 * iu) + wq * (1 - wu) * dustem_at(c, i, iq + 1, 2 * iu) +
                wq * wu * dustem_at(c, i, iq + 1, 2 * (iu + 1)) + (1 - wq) * wu * dustem_at(c, i, iq, 2 * (iu + 1));
    return fmax((1.0 - gam) * lo + gam * hi, DBL_MIN);
}

/* The model's IMF parameters (batch_imf_vectors, opensps_csp_batch.f90:1823-1853): the slopes and limits per model,
 * the rest from the context */
static void model_pimf(const orc_ctx_t *c, const double *p, double *pimf) {
    for (int i = 0; i < ORC_IMF_NPAR; i++) pimf[i] = 0.0;
    pimf[ORC_IMF_SALP] = c->salp_ind;
    pimf[ORC_IMF_ALPHA] = p[OPS_P_IMF1];
    pimf[ORC_IMF_ALPHA + 1] = p[OPS_P_IMF2];
    pimf[ORC_IMF_ALPHA + 2] = p[OPS_P_IMF3];
    pimf[ORC_IMF_MDAVE] = c->mdave;
    pimf[ORC_IMF_LOWER_LIMIT] = p[OPS_P_IMF_LOWER];
    pimf[ORC_IMF_UPPER_LIMIT] = p[OPS_P_IMF_UPPER];
    pimf[ORC_IMF_LOWER_BOUND] = p[OPS_P_IMF_LOWER];
    pimf[ORC_IMF_MLIM_BH] = c->mlim_bh;
    pimf[ORC_IMF_MLIM_NS] = c->mlim_ns;
}

/* Per-model birth-cloud age index: batch_model_indices (opensps_csp_batch.f90:1541-1576) */
static int model_it_tesc(const orc_ctx_t *c, double dust_tesc) {
    double dal = dust_tesc > SAFE_FLOOR ? dust_tesc : 7.0;
    int it = orc_find_interval(c->time_full, c->nt, dal);
    if (it > c->nt) it = c->nt;
    if (it < 1) it = 1;
    return it;
}

/* metallicity_interval (opensps_csp_batch.f90:2073-2084): weight not clamped */
static void model_zlo(const orc_ctx_t *c, double logzsol, int *zlo, double *dz) {
    int nz = c->nz;
    int z = orc_find_interval(c->logz, nz, logzsol);
    if (z > nz - 1) z = nz - 1;
    if (z < 1) z = 1;
    *zlo = z;
    *dz = (logzsol - c->logz[z - 1]) / (c->logz[z] - c->logz[z - 1]);
}

/*
 * orc_csp_batch: n models with parameters prm[slot * n + m]. ssp_weights (nt x n) replaces the SFH
 * parameters when non-NULL. Outputs (any NULL but mags): mags (nbands x n), spec (nspec x n),
 * mass, lbol (n), young / old (nspec x n, the age sums before dust), idx (3 x n: it_tesc, zlo,
 * 1-based; zlo = 1 with one basis), w_out (nt x n, the FP64 SFH weights), fband (nbands x n).
 * Returns 0, 2 (unsupported SFH type), 3 (bad arguments).
 */
/* ---------------------------------------------------------------------------------------------
 * Nebular emission with the CLOUDY grids (nebular_model = 0): nebular_setup (opensps_csp_batch.f90:1630-1770),
 * nebular_kernels (:2612-2998: the ionizing photons, the EUV rows with the escaping fraction, the continuum and the
 * line luminosities summed over the nebular ages), line_profiles (:3133-3160) and line_attenuated (:3440-3473). The
 * lines' profiles are compute_line_gaussians' (opensps_gas.f90:298-340, velocity smoothing) with the minimum width
 * of compute_nebular_kernels (opensps_initialization.f90:1377-1385), built from the first model's sigma_smooth as
 * the batch builds them (prepare_line_gaussians(ctx, psets(1))), each cut where it falls to 1e-14 of its peak.
 * ------------------------------------------------------------------------------------------- */
#define H_PLANCK 6.62607015e-27           /* erg s: opensps_constants.f90 */
#define L_SOL 3.828e33                    /* erg / s: opensps_constants.f90 */
#define NEB_PROFILE_CUT 1.0e-14           /* LINE_PROFILE_CUT (opensps_csp_batch.f90:1638) */

typedef struct {
    int nt_neb, whlylim;
    int *ia;                  /* per age 1..nt_neb: the CLOUDY age interval (1-based) */
    double *wa;               /* and its weight */
    double *q0;               /* nt_neb x nz: ionizing photons of each basis spectrum (frac_obrun = 0) */
    int *llo;                 /* per line: the pixel interval of its wavelength (1-based) */
    int *jlo, *jhi;           /* per line: the window of its profile (1-based; jlo > jhi: none) */
    double *gauss;            /* per line: the profile over its window (at goff[i]) */
    int *goff;
} orc_neb_t;

/* get_grid_indices_weights (opensps_gas.f90:394-412): the clamped interval and weight */
static void neb_index(const double *g, int ng, double v, int *idx, double *w) {
    int i = orc_find_interval(g, ng, v);
    i = i < 1 ? 1 : (i > ng - 1 ? ng - 1 : i);
    double x = (v - g[i - 1]) / (g[i] - g[i - 1]);
    *idx = i;
    *w = x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x);
}

/* zu_point (:3942-3955): a table column (stride s0 between rows of the first axis, column-major) interpolated in
 * log Z and log U at CLOUDY age a (1-based) */
static double neb_zu(const orc_ctx_t *c, const double *tab, int nrow, int row, int a, int iz, int iu, double wz,
                     double wu) {
    const size_t sz = (size_t)nrow, sa = sz * c->nnz, su = sa * c->nna;
#define NT_(z, u) tab[(size_t)(row - 1) + (size_t)((z) - 1) * sz + (size_t)(a - 1) * sa + (size_t)((u) - 1) * su]
    return (1.0 - wz) * (1.0 - wu) * NT_(iz, iu) + (1.0 - wz) * wu * NT_(iz, iu + 1) + wz * (1.0 - wu) * NT_(iz + 1, iu) +
           wz * wu * NT_(iz + 1, iu + 1);
#undef NT_
}

static void neb_free(orc_neb_t *nb) {
    free(nb->ia); free(nb->wa); free(nb->q0); free(nb->llo); free(nb->jlo); free(nb->jhi); free(nb->gauss);
    free(nb->goff);
    memset(nb, 0, sizeof(*nb));
}

/* nebular_setup's batch-wide part (the per-model intervals are formed in the model's loop), with step 1 of
 * nebular_kernels (the ionizing photons of each basis spectrum of a nebular age, the photon-number integral of
 * process_ionizing_radiation without the absorbed fraction) */
static void neb_setup(const orc_ctx_t *c, double sigma0, orc_neb_t *nb) {
    const int nspec = c->nspec, nt = c->nt, nl = c->nline;
    memset(nb, 0, sizeof(*nb));
    int ntn = orc_find_interval(c->time_full, nt, c->neb_lage[c->nna - 1]);
    nb->nt_neb = ntn < 0 ? 0 : (ntn > nt ? nt : ntn);
    nb->whlylim = orc_find_interval(c->lambda, nspec, 912.0);
    const int n1 = nb->nt_neb > 0 ? nb->nt_neb : 1;
    nb->ia = (int *)malloc(sizeof(int) * n1);
    nb->wa = (double *)malloc(sizeof(double) * n1);
    nb->q0 = (double *)malloc(sizeof(double) * n1 * c->nz);
    for (int t = 1; t <= nb->nt_neb; t++) neb_index(c->neb_lage, c->nna, c->time_full[t - 1], &nb->ia[t - 1], &nb->wa[t - 1]);
    for (int iz = 1; iz <= c->nz; iz++)
        for (int t = 1; t <= nb->nt_neb; t++) {
            const double *b = c->basis + ((size_t)(iz - 1) * nt + (t - 1)) * nspec;
            double s = 0.0;
            if (nb->whlylim >= 2) {
                for (int i = 1; i <= nb->whlylim - 1; i++) {
                    double nu0 = C_LIGHT / c->lambda[i - 1], nu1 = C_LIGHT / c->lambda[i];
                    double wt = 0.5 * fabs(nu1 - nu0);
                    s = s + wt * (b[i - 1] / nu0 + b[i] / nu1);
                }
                nb->q0[(size_t)(iz - 1) * n1 + (t - 1)] = s / H_PLANCK * L_SOL;
            } else {
                nb->q0[(size_t)(iz - 1) * n1 + (t - 1)] = 0.0;
            }
        }
    /* the lines' pixel intervals and profiles */
    nb->llo = (int *)malloc(sizeof(int) * (nl > 0 ? nl : 1));
    nb->jlo = (int *)malloc(sizeof(int) * (nl > 0 ? nl : 1));
    nb->jhi = (int *)malloc(sizeof(int) * (nl > 0 ? nl : 1));
    nb->goff = (int *)malloc(sizeof(int) * (nl + 1));
    double *g = (double *)malloc(sizeof(double) * nspec);
    const double pi = acos(-1.0), sq2pi = sqrt(2.0 * pi), kms = 1.0e13 / C_LIGHT;
    int tot = 0, cap = 1024;
    nb->gauss = (double *)malloc(sizeof(double) * cap);
    for (int i = 1; i <= nl; i++) {
        const double pos = c->neb_lpos[i - 1];
        int jc = 1, j = nspec + 1;
        while (j - jc > 1) {
            int p = (jc + j) / 2;
            if (pos >= c->lambda[p - 1]) jc = p; else j = p;
        }
        nb->llo[i - 1] = jc < 1 ? 1 : (jc > nspec - 1 ? nspec - 1 : jc);
        nb->goff[i - 1] = tot;
        nb->jlo[i - 1] = 1;
        nb->jhi[i - 1] = 0;
        if (!c->neb_lines_in_spec) continue;
        int jr = orc_find_interval(c->lambda, nspec, pos);
        jr = jr < 1 ? 1 : (jr > nspec - 1 ? nspec - 1 : jr);
        const double res = c->lambda[jr] - c->lambda[jr - 1];
        double sig = pos * sigma0 * kms;
        if (sig < res * 2.0) sig = res * 2.0;
        const double norm = (1.0 / (sq2pi * sig)) * (pos * pos / C_LIGHT);
        for (int k = 0; k < nspec; k++) {
            double x = (c->lambda[k] - pos) / sig;
            g[k] = norm * exp(-0.5 * x * x);
        }
        int jcc = orc_find_interval(c->lambda, nspec, pos);
        jcc = jcc < 1 ? 1 : (jcc > nspec ? nspec : jcc);
        int jh = jcc + 1 < nspec ? jcc + 1 : nspec;
        double cut = NEB_PROFILE_CUT * (g[jcc - 1] > g[jh - 1] ? g[jcc - 1] : g[jh - 1]);
        if (cut <= 0.0) continue;
        int lo = jcc, hi = jh;
        while (lo > 1 && g[lo - 2] > cut) lo--;
        while (hi < nspec && g[hi] > cut) hi++;
        nb->jlo[i - 1] = lo;
        nb->jhi[i - 1] = hi;
        if (tot + (hi - lo + 1) > cap) {
            while (tot + (hi - lo + 1) > cap) cap *= 2;
            nb->gauss = (double *)realloc(nb->gauss, sizeof(double) * cap);
        }
        for (int k = lo; k <= hi; k++) nb->gauss[tot++] = g[k - 1];
    }
    nb->goff[nl] = tot;
    free(g);
}

/* The nebular emission of one model on the basis path (nebular_kernels steps 2-5 for column m): the EUV rows of the
 * young and old age sums redone with the escaping fraction fesc = clamp(frac_obrun) on the nebular ages; the
 * continuum sum_t w(t) Q(t) 10^lg(pixel, t) and the lines' luminosities likewise (lg interpolated bilinearly in the
 * gas log Z and log U, linearly in log age between the CLOUDY ages; Q the blend of the bracketing metallicities'
 * photons times 1 - frac_obrun); the lines spread over their profiles. y, o: the age sums (in, out); w: the SFH
 * weights (those <= SAFE_FLOOR taken as 0); it: it_tesc; lum: 2 x nline (young, then old) out. */
static void neb_model(const orc_ctx_t *c, const orc_neb_t *nb, const double *p, const double *w, int it, int zlo,
                      double dz, int blend, double *y, double *o, double *lum) {
    const int nspec = c->nspec, nt = c->nt, nl = c->nline, ntn = nb->nt_neb;
    const int n1 = ntn > 0 ? ntn : 1;
    for (int i = 0; i < 2 * nl; i++) lum[i] = 0.0;
    if (ntn < 1) return;
    int iz = 1, iu = 1;
    double wz = 0.0, wu = 0.0;
    neb_index(c->neb_logz, c->nnz, p[OPS_P_GAS_LOGZ], &iz, &wz);
    neb_index(c->neb_logu, c->nnu, p[OPS_P_GAS_LOGU], &iu, &wu);
    double fesc = p[OPS_P_FRAC_OBRUN];
    fesc = fesc < 0.0 ? 0.0 : (fesc > 1.0 ? 1.0 : fesc);
    const double fabs_ = 1.0 - p[OPS_P_FRAC_OBRUN];
    /* 2. the EUV rows */
    const double *blo = c->basis + (size_t)(zlo - 1) * nspec * nt;
    const double *bhi = blend ? c->basis + (size_t)zlo * nspec * nt : blo;
    for (int i = 1; i <= nb->whlylim; i++) {
        double sy = 0.0, so = 0.0;
        for (int k = 1; k <= nt; k++) {
            double wk = w[k - 1] > SAFE_FLOOR ? w[k - 1] : 0.0;
            double bk = blend ? (1.0 - dz) * blo[(size_t)(k - 1) * nspec + i - 1] + dz * bhi[(size_t)(k - 1) * nspec + i - 1]
                              : (1.0 - dz) * blo[(size_t)(k - 1) * nspec + i - 1];
            if (k <= ntn) bk = bk * fesc;
            if (k <= it) sy = sy + wk * bk;
            else so = so + wk * bk;
        }
        y[i - 1] = sy;
        o[i - 1] = so;
    }
    /* the weights times Q per nebular age */
    double *cq = (double *)malloc(sizeof(double) * n1);
    for (int t = 1; t <= ntn; t++) {
        double q = blend ? (1.0 - dz) * nb->q0[(size_t)(zlo - 1) * n1 + t - 1] + dz * nb->q0[(size_t)zlo * n1 + t - 1]
                         : nb->q0[t - 1];
        q = fabs(q * fabs_);
        double wk = w[t - 1] > SAFE_FLOOR ? w[t - 1] : 0.0;
        cq[t - 1] = (wk != 0.0 && q > SAFE_FLOOR) ? wk * q : 0.0;
    }
    /* 3. the continuum */
    if (c->neb_cont_on)
        for (int j = 1; j <= nspec; j++) {
            double sy = 0.0, so = 0.0;
            for (int t = 1; t <= ntn; t++) {
                if (cq[t - 1] == 0.0) continue;
                int a = nb->ia[t - 1];
                double ra = neb_zu(c, c->neb_cont, nspec, j, a, iz, iu, wz, wu);
                double rb = neb_zu(c, c->neb_cont, nspec, j, a + 1, iz, iu, wz, wu);
                double v = cq[t - 1] * pow(10.0, (1.0 - nb->wa[t - 1]) * ra + nb->wa[t - 1] * rb);
                if (t <= it) sy = sy + v;
                else so = so + v;
            }
            y[j - 1] = y[j - 1] + sy;
            o[j - 1] = o[j - 1] + so;
        }
    /* 4. the lines' luminosities */
    for (int i = 1; i <= nl; i++) {
        double sy = 0.0, so = 0.0;
        for (int t = 1; t <= ntn; t++) {
            double wk = w[t - 1] > SAFE_FLOOR ? w[t - 1] : 0.0;
            if (wk == 0.0) continue;
            double q = blend ? (1.0 - dz) * nb->q0[(size_t)(zlo - 1) * n1 + t - 1] + dz * nb->q0[(size_t)zlo * n1 + t - 1]
                             : nb->q0[t - 1];
            q = fabs(q * fabs_);
            if (q <= SAFE_FLOOR) continue;
            int a = nb->ia[t - 1];
            double ra = neb_zu(c, c->neb_line, nl, i, a, iz, iu, wz, wu);
            double rb = neb_zu(c, c->neb_line, nl, i, a + 1, iz, iu, wz, wu);
            double lg = (1.0 - nb->wa[t - 1]) * ra + nb->wa[t - 1] * rb;
            if (t <= it) sy = sy + wk * (pow(10.0, lg) * q);
            else so = so + wk * (pow(10.0, lg) * q);
        }
        lum[i - 1] = sy;
        lum[nl + i - 1] = so;
    }
    free(cq);
    /* 5. the lines' profiles: per pixel, the lines whose window holds it, in line order (line_profiles) */
    if (c->neb_lines_in_spec)
        for (int j = 1; j <= nspec; j++) {
            double sy = 0.0, so = 0.0;
            int any = 0;
            for (int i = 1; i <= nl; i++) {
                if (j < nb->jlo[i - 1] || j > nb->jhi[i - 1]) continue;
                double g = nb->gauss[nb->goff[i - 1] + (j - nb->jlo[i - 1])];
                sy = sy + g * lum[i - 1];
                so = so + g * lum[nl + i - 1];
                any = 1;
            }
            if (any) {
                y[j - 1] = y[j - 1] + sy;
                o[j - 1] = o[j - 1] + so;
            }
        }
}

int orc_csp_batch(const orc_ctx_t *c, int n, const double *prm, const double *ssp_weights, double *mags,
                  double *spec, double *mass, double *lbol, double *young, double *old, int *idx, double *w_out,
                  double *fband_out) {
    const int nspec = c->nspec, nt = c->nt, nz = c->nz, nb = c->nbands;
    if (n < 0 || nspec < 2 || nt < 2 || nz < 1 || !mags) return 3;
    if (n == 0) return 0;
    const double amzl = abs_mag_zeropoint_log();
    int status = 0;
    int *jlo = (int *)malloc(sizeof(int) * (nb > 0 ? nb : 1)), *jhi = (int *)malloc(sizeof(int) * (nb > 0 ? nb : 1));
    int *do_band = (int *)malloc(sizeof(int) * (nb > 0 ? nb : 1));
    double *bt = (double *)malloc(sizeof(double) * (size_t)(nb > 0 ? nb : 1) * nspec);
    double *wtl = (double *)malloc(sizeof(double) * nspec), *ltab = (double *)malloc(sizeof(double) * nspec);
    orc_band_supports(c->bands, nspec, nb, jlo, jhi);
    for (int b = 0; b < nb; b++) {
        do_band[b] = c->mag_mask ? c->mag_mask[b] : 1;
        for (int j = 0; j < nspec; j++) bt[(size_t)j * nb + b] = c->bands[(size_t)b * nspec + j];
    }
    for (int j = 0; j < nspec - 1; j++) wtl[j] = 0.5 * fabs(c->lambda[j + 1] - c->lambda[j]);
    wtl[nspec - 1] = 0.0;
    for (int j = 0; j < nspec; j++) ltab[j] = log(c->lambda[j] / 5500.0);
    /* jtop: the last pixel below Lyman-alpha (apply_igm_z, opensps_csp_batch.f90:4221-4224) */
    int jtop = 0;
    for (int j = 1; j <= nspec; j++) if (c->lambda[j - 1] < 1215.67) jtop = j;
    /* z_varies: the reference takes compute_igm_transmission's direct form when every model shares
     * zred and igm_factor (opensps_csp_batch.f90:1011-1015, prepare_batch_grid) */
    int z_varies = 0;
    for (int m = 1; m < n; m++)
        if (prm[OPS_P_ZRED * n + m] != prm[OPS_P_ZRED * n] ||
            prm[OPS_P_IGM_FACTOR * n + m] != prm[OPS_P_IGM_FACTOR * n]) z_varies = 1;
    int cap = nspec * 64;
    /* nebular emission: the batch-wide set-up; not with weights over metallicities or the IMF per model (their
     * nebular paths are not restated here), nor lines in the spectrum with models of different sigma_smooth
     * (status 2, opensps_csp_batch.f90:749-752), nor lines outside the spectrum with dust emission */
    orc_neb_t neb;
    memset(&neb, 0, sizeof(neb));
    const int nl = c->add_neb ? c->nline : 0;
    if (c->add_neb) {
        int bad = c->ssp_weights_z != NULL || c->ops != NULL || (!c->neb_lines_in_spec && c->add_dust_emission == 1);
        if (c->neb_lines_in_spec)
            for (int m = 1; m < n; m++)
                if (prm[OPS_P_SIGMA_SMOOTH * n + m] != prm[OPS_P_SIGMA_SMOOTH * n]) bad = 1;
        if (bad) {
            free(jlo); free(jhi); free(do_band); free(bt); free(wtl); free(ltab);
            return 2;
        }
        neb_setup(c, prm[OPS_P_SIGMA_SMOOTH * n], &neb);
    }

    #pragma omp parallel
    {
        double *p = (double *)malloc(sizeof(double) * OPS_NPAR);
        double *w = (double *)malloc(sizeof(double) * nt), *t1 = (double *)malloc(sizeof(double) * nt);
        double *t2 = (double *)malloc(sizeof(double) * nt);
        double *y = (double *)malloc(sizeof(double) * nspec), *o = (double *)malloc(sizeof(double) * nspec);
        double *tot = (double *)malloc(sizeof(double) * nspec), *work = (double *)malloc(sizeof(double) * nspec);
        double *fobs = (double *)malloc(sizeof(double) * nspec), *fb = (double *)malloc(sizeof(double) * (nb + 1));
        int *slo = (int *)malloc(sizeof(int) * nspec), *shi = (int *)malloc(sizeof(int) * nspec);
        int *soff = (int *)malloc(sizeof(int) * nspec);
        double *scoef = (double *)malloc(sizeof(double) * cap);
        double *tau = (double *)malloc(sizeof(double) * (jtop + 1));
        double *tdv = (double *)malloc(sizeof(double) * nspec);
        double *nlum = (double *)malloc(sizeof(double) * (2 * nl + 1));
        double sm_key[3] = {NAN, NAN, NAN};
        int sm_n = 0;
        double *smass = NULL, *slum = NULL, *vy = NULL, *vo = NULL, *wpt = NULL, *llo = NULL, *dl = NULL;
        if (c->ops) {
            int maxpts = 1;
            for (int a = 0; a < c->ops->nz * nt; a++)
                if (c->ops->age_ptr[a + 1] - c->ops->age_ptr[a] > maxpts) maxpts = c->ops->age_ptr[a + 1] - c->ops->age_ptr[a];
            smass = (double *)malloc(sizeof(double) * nt); slum = (double *)malloc(sizeof(double) * nt);
            vy = (double *)malloc(sizeof(double) * c->ops->ntpl); vo = (double *)malloc(sizeof(double) * c->ops->ntpl);
            wpt = (double *)malloc(sizeof(double) * maxpts); llo = (double *)malloc(sizeof(double) * maxpts);
            dl = (double *)malloc(sizeof(double) * maxpts);
        }

        #pragma omp for schedule(dynamic, 4)
        for (int m = 0; m < n; m++) {
            for (int q = 0; q < OPS_NPAR; q++) p[q] = prm[(size_t)q * n + m];
            /* stage 1: indices */
            int it = model_it_tesc(c, p[OPS_P_DUST_TESC]);
            int zlo = 1;
            double dz = 0.0;
            int blend = nz > 1;
            if (blend) model_zlo(c, p[OPS_P_LOGZSOL], &zlo, &dz);
            if (idx) { idx[3 * m] = it; idx[3 * m + 1] = zlo; idx[3 * m + 2] = 0; }
            double lsum_m = 0.0;   /* the linear L_bol sum (the AGN's luminosity) */
            /* stage 0 checks (batch_check_models :1312-1320; plan_smoothing host modes are not built) */
            if ((!ssp_weights && !c->ssp_weights_z && p[OPS_P_TAGE] <= 0.0) || (!c->smooth_fast && p[OPS_P_SIGMA_SMOOTH] > SAFE_FLOOR)) {
                #pragma omp atomic write
                status = 2;
                continue;
            }
            /* weights over ages and metallicities: model_sums_zw and age_sums_w (metallicities outer; young
             * ages k <= it_tesc) */
            if (c->ssp_weights_z) {
                const double *wz = c->ssp_weights_z + (size_t)m * nt * nz;
                double ms = 0.0, ls = 0.0;
                for (int i = 0; i < nspec; i++) { y[i] = 0.0; o[i] = 0.0; }
                for (int zz = 0; zz < nz; zz++)
                    for (int k = 1; k <= nt; k++) {
                        double wk = wz[(size_t)zz * nt + k - 1];
                        if (wk > SAFE_FLOOR) {
                            ms = ms + wk * c->bmass[(size_t)zz * nt + k - 1];
                            ls = ls + wk * pow(10.0, c->blbol[(size_t)zz * nt + k - 1]);
                        } else {
                            wk = 0.0;
                        }
                        double *acc = k <= it ? y : o;
                        const double *bs = c->basis + ((size_t)zz * nt + k - 1) * nspec;
                        for (int i = 0; i < nspec; i++) acc[i] = acc[i] + wk * bs[i];
                    }
                if (w_out) for (int k = 0; k < nt; k++) w_out[(size_t)m * nt + k] = 0.0;
                if (mass) mass[m] = ms;
                if (lbol) lbol[m] = ls > SAFE_FLOOR ? log10(ls) : 0.0;
                lsum_m = ls;
                if (young) memcpy(young + (size_t)m * nspec, y, sizeof(double) * nspec);
                if (old) memcpy(old + (size_t)m * nspec, o, sizeof(double) * nspec);
                goto dust;
            }
            /* stage 2: SFH weights (batch_sfh_weights, :1865-1917) */
            if (ssp_weights) {
                for (int k = 0; k < nt; k++) w[k] = ssp_weights[(size_t)m * nt + k];
            } else {
                int st = orc_sfh_weights(c->time_full, nt, c->tiny_logt, p, w, t1, t2);
                if (st != 0) {
                    #pragma omp atomic write
                    status = st;
                    continue;
                }
            }
            if (w_out) for (int k = 0; k < nt; k++) w_out[(size_t)m * nt + k] = w[k];
            /* stage 3: model_sums (:2129-2176) and the age sums (age_sums_host, :2447-2511); with the IMF per
             * model, the operator's age sums and per-age sums (batch_imf_vectors :1823-1853,
             * ssp_batch_age_sums_z, model_sums MSUMS_OWN) */
            double ms = 0.0, ls = 0.0;
            if (c->ops) {
                double pimf[ORC_IMF_NPAR];
                model_pimf(c, p, pimf);
                for (int k = 0; k < nt; k++) t1[k] = w[k] > SAFE_FLOOR ? w[k] : 0.0;
                orc_ops_model(c->ops, nspec, c->imf_type, c->remnants, pimf, zlo, dz, t1, it, y, o, smass, slum, vy, vo,
                              wpt, llo, dl, c->op_count ? c->op_count + 2 * m : NULL);
                for (int k = 0; k < nt; k++)
                    if (w[k] > SAFE_FLOOR) { ms = ms + w[k] * smass[k]; ls = ls + w[k] * slum[k]; }
                goto sums_done;
            }
            double c2 = dz, c1 = 1.0 - c2;
            const double *blo = c->basis + (size_t)(zlo - 1) * nspec * nt;
            const double *bhi = blend ? c->basis + (size_t)zlo * nspec * nt : blo;
            for (int i = 0; i < nspec; i++) { y[i] = 0.0; o[i] = 0.0; }
            for (int k = 1; k <= nt; k++) {
                double wk = w[k - 1];
                if (wk > SAFE_FLOOR) {
                    double mk, lk;
                    if (blend) {
                        mk = (1.0 - dz) * c->bmass[(size_t)(zlo - 1) * nt + k - 1] + dz * c->bmass[(size_t)zlo * nt + k - 1];
                        lk = (1.0 - dz) * c->blbol[(size_t)(zlo - 1) * nt + k - 1] + dz * c->blbol[(size_t)zlo * nt + k - 1];
                        mk = wk * mk;
                        lk = wk * pow(10.0, lk);
                    } else {
                        mk = wk * c->bmass[k - 1];
                        lk = wk * pow(10.0, c->blbol[k - 1]);
                    }
                    ms = ms + mk;
                    ls = ls + lk;
                } else {
                    wk = 0.0;
                }
                double *acc = k <= it ? y : o;
                const double *bl = blo + (size_t)(k - 1) * nspec, *bh = bhi + (size_t)(k - 1) * nspec;
                if (blend)
                    for (int i = 0; i < nspec; i++) acc[i] = acc[i] + wk * (c1 * bl[i] + c2 * bh[i]);
                else
                    for (int i = 0; i < nspec; i++) acc[i] = acc[i] + wk * (c1 * bl[i]);
            }
        sums_done:
            /* stage 4: nebular emission (add_nebular_batch, :2565-2596) into the age sums */
            if (c->add_neb) neb_model(c, &neb, p, w, it, zlo, dz, blend, y, o, nlum);
            if (mass) mass[m] = ms;
            if (lbol) lbol[m] = ls > SAFE_FLOOR ? log10(ls) : 0.0;
            lsum_m = ls;
            if (young) memcpy(young + (size_t)m * nspec, y, sizeof(double) * nspec);
            if (old) memcpy(old + (size_t)m * nspec, o, sizeof(double) * nspec);
        dust:;
            /* stage 5: attenuate_batch + pixel_transmissions, dust_type 0 (:3576-3818) */
            double fob = p[OPS_P_FRAC_OBRUN], fnd = p[OPS_P_FRAC_NODUST];
            double ed = p[OPS_P_DUST_INDEX];
            int no_birth = p[OPS_P_DUST1] <= SAFE_FLOOR || 1.0 - fob <= SAFE_FLOOR;
            int no_old = fabs(p[OPS_P_DUST3]) <= SAFE_FLOOR;
            /* dust emission and the AGN torus: who emits (batch_model_indices :1569-1570, agn_indices :1591-1606) */
            const int dusty = p[OPS_P_DUST1] > SAFE_FLOOR || p[OPS_P_DUST2] > SAFE_FLOOR;
            const int emis = c->add_dust_emission == 1 && dusty;
            const int agn_on = c->add_agn_dust == 1 && p[OPS_P_FAGN] > SAFE_FLOOR;
            if ((emis && c->nqpah == 0) || (agn_on && c->ntau == 0)) {
                #pragma omp atomic write
                status = 2;
                continue;
            }
            /* the lines' luminosities through the birth cloud and the diffuse dust (attenuate_lines,
             * line_attenuated :3440-3473: T_d linear between the line's two pixels) */
            if (c->add_neb && c->emlines_out) {
                for (int i = 1; i <= nl; i++) {
                    const double pos = c->neb_lpos[i - 1];
                    const int lo = neb.llo[i - 1];
                    double td = 1.0;
                    if (!(1.0 - fnd <= SAFE_FLOOR)) {
                        double cv0 = c->dcurve ? c->dcurve[lo - 1] : (ed != 0.0 ? exp(ed * ltab[lo - 1]) : 1.0);
                        double cv1 = c->dcurve ? c->dcurve[lo] : (ed != 0.0 ? exp(ed * ltab[lo]) : 1.0);
                        double t_lo = exp(-p[OPS_P_DUST2] * cv0), t_hi = exp(-p[OPS_P_DUST2] * cv1);
                        td = t_lo + (t_hi - t_lo) / (c->lambda[lo] - c->lambda[lo - 1]) * (pos - c->lambda[lo - 1]);
                    }
                    if (td != td) td = 1.0;
                    double birth = 1.0;
                    if (!(p[OPS_P_DUST1] <= SAFE_FLOOR || 1.0 - fob <= SAFE_FLOOR))
                        birth = exp(-p[OPS_P_DUST1] * pow(pos / 5500.0, p[OPS_P_DUST1_INDEX]));
                    double ay = nlum[i - 1] * (birth * (1.0 - fob) + fob), ao = nlum[nl + i - 1];
                    if (1.0 - fnd > SAFE_FLOOR) {
                        ay = ay * (td * (1.0 - fnd) + fnd);
                        ao = ao * (td * (1.0 - fnd) + fnd);
                    }
                    c->emlines_out[(size_t)m * nl + i - 1] = ay + ao;
                }
            }
            int iq = 1, iu = 1;
            double wq = 0.0, wu = 0.0, gam = 0.0;
            if (emis) dl07_indices(c, p, &iq, &wq, &iu, &wu, &gam);
            double *absb = work;   /* (free until the smoothing) */
            for (int i = 0; i < nspec; i++) {
                double curve = 1.0;
                if (c->dcurve) curve = c->dcurve[i];
                else if (ed != 0.0) curve = curve * exp(ed * ltab[i]);
                double tau_d = p[OPS_P_DUST2] * curve, tr_d = exp(-tau_d);
                double tau_b = 0.0, tau_o = 0.0, tr_b = 1.0, tr_o = 1.0;
                if (!no_birth) { tau_b = p[OPS_P_DUST1] * exp(p[OPS_P_DUST1_INDEX] * ltab[i]); tr_b = exp(-tau_b); }
                if (!no_old) { tau_o = p[OPS_P_DUST3] * curve; tr_o = exp(-tau_o); }
                double ssum = (y[i] * tr_b * (1.0 - fob) + y[i] * fob) + o[i] * tr_o;
                if (1.0 - fnd <= SAFE_FLOOR) tot[i] = ssum;
                else tot[i] = ssum * (tr_d * (1.0 - fnd) + fnd);
                if (emis) {
                    /* the absorbed light (attenuate_batch :3640-3656) */
                    double omtb = one_minus_trans(tau_b, tr_b), omto = one_minus_trans(tau_o, tr_o);
                    double omtd = (1.0 - fnd <= SAFE_FLOOR) ? 0.0 : (1.0 - fnd) * one_minus_trans(tau_d, tr_d);
                    double ay = tr_b * (1.0 - fob) + fob;
                    absb[i] = y[i] * ((1.0 - fob) * omtb + ay * omtd) + o[i] * (omto + tr_o * omtd);
                    fobs[i] = dl07_shape(c, i, iq, wq, iu, wu, gam);   /* the template (fobs: free until stage 8) */
                    tdv[i] = tr_d;
                }
            }
            if (emis) {
                /* dust_emission_reductions (:3963-4042, the CPU build's compensated sums) and add_dust_emission
                 * (:4045-4099): the template times the diffuse transmission, normalised to the absorbed light over
                 * frequency (prepare_batch_grid's wt, src/spectra/opensps_batch_photometry.f90:262) */
                double sa = 0.0, sn = 0.0, se = 0.0, ca = 0.0, cn = 0.0, ce = 0.0;
                for (int i = 0; i < nspec - 1; i++) {
                    double wt = 0.5 * fabs(2.99792458e18 / c->lambda[i + 1] - 2.99792458e18 / c->lambda[i]);
                    kadd(&sa, &ca, wt * (absb[i] + absb[i + 1]));
                    kadd(&sn, &cn, wt * (fobs[i] + fobs[i + 1]));
                    kadd(&se, &ce, wt * (fobs[i] * tdv[i] + fobs[i + 1] * tdv[i + 1]));
                }
                double rn = se > SAFE_FLOOR ? sa / se : 0.0;
                if (rn != 0.0)
                    for (int i = 0; i < nspec; i++) tot[i] = tot[i] + fobs[i] * tdv[i] * rn;
                if (c->mdust_out) c->mdust_out[m] = sn > SAFE_FLOOR ? 3.21e-3 / (4.0 * acos(-1.0)) * (sa / sn) : SAFE_FLOOR;
            } else if (c->mdust_out) {
                c->mdust_out[m] = SAFE_FLOOR;
            }
            /* stage 6: IGM (apply_igm_z :4205-4319; uniform z: compute_igm_transmission) */
            double z = p[OPS_P_ZRED];
            if (c->add_igm && z > SAFE_FLOOR && jtop > 0) {
                double opz = 1.0 + z;
                if (z_varies) {
                    double a6[6], b6[6];
                    orc_igm_z_terms(opz, b6);
                    for (int j = 0; j < jtop; j++) {
                        orc_igm_pixel_terms(c->lambda[j], a6);
                        tau[j] = a6[0] * b6[0] + a6[1] * b6[1] + a6[2] * b6[2] + a6[3] * b6[3] + a6[4] * b6[4] + a6[5] * b6[5];
                    }
                } else {
                    for (int j = 0; j < jtop; j++) tau[j] = orc_igm_tau(c->lambda[j], opz);
                }
                double tmax = 0.0;
                int jmax = 1;
                for (int j = 1; j <= jtop; j++) if (tau[j - 1] > tmax) { tmax = tau[j - 1]; jmax = j; }
                if (jmax > 1) for (int j = 1; j <= jmax; j++) tau[j - 1] = tmax;
                for (int j = 0; j < jtop; j++) tot[j] = tot[j] * exp(-tau[j] * p[OPS_P_IGM_FACTOR]);
            }
            /* the AGN torus (add_agn_batch :4342-4433): fagn L_bol (L_bol through its logarithm) times the agn_tau
             * template, linear between its bracketing depths (agn_indices :1591-1606: weight clamped), times the
             * diffuse transmission; after the IGM */
            if (agn_on) {
                int ia = orc_find_interval(c->agn_tau, c->ntau, p[OPS_P_AGN_TAU]);
                if (ia > c->ntau - 1) ia = c->ntau - 1;
                if (ia < 1) ia = 1;
                double wta = (p[OPS_P_AGN_TAU] - c->agn_tau[ia - 1]) / (c->agn_tau[ia] - c->agn_tau[ia - 1]);
                wta = fmax(0.0, fmin(wta, 1.0));
                double lb = lsum_m > SAFE_FLOOR ? log10(lsum_m) : 0.0;
                double lagn = pow(10.0, lb) * p[OPS_P_FAGN];
                for (int i = 0; i < nspec; i++) {
                    double curve = 1.0;
                    if (c->dcurve) curve = c->dcurve[i];
                    else if (ed != 0.0) curve = curve * exp(ed * ltab[i]);
                    double tr_d = exp(-(p[OPS_P_DUST2] * curve));
                    double tmpl = (1.0 - wta) * c->agn[(size_t)(ia - 1) * nspec + i] + wta * c->agn[(size_t)ia * nspec + i];
                    tmpl = tmpl * tr_d;
                    tot[i] = tot[i] + lagn * tmpl;
                }
            }
            /* stage 7: velocity smoothing (smooth_columns / smooth_kernel FP64, build_velocity_kernel) */
            double sig = p[OPS_P_SIGMA_SMOOTH];
            if (c->smooth_fast && sig > SAFE_FLOOR) {
                if (!(sm_key[0] == sig && sm_key[1] == p[OPS_P_MIN_WAVE_SMOOTH] && sm_key[2] == p[OPS_P_MAX_WAVE_SMOOTH])) {
                    sm_n = orc_velocity_kernel(c->lambda, nspec, sig, p[OPS_P_MIN_WAVE_SMOOTH], p[OPS_P_MAX_WAVE_SMOOTH],
                                               slo, shi, soff, scoef, cap);
                    sm_key[0] = sig; sm_key[1] = p[OPS_P_MIN_WAVE_SMOOTH]; sm_key[2] = p[OPS_P_MAX_WAVE_SMOOTH];
                }
                if (sm_n > 0) apply_velocity_kernel(nspec, slo, shi, soff, scoef, tot, work);
            }
            if (spec) memcpy(spec + (size_t)m * nspec, tot, sizeof(double) * nspec);
            /* stage 8: magnitudes (compute_magnitudes_columns, observed_frame_z, band_fluxes) */
            const double zm = c->zred_mags ? c->zred_mags[m] : z;
            orc_observed_frame(c->lambda, nspec, zm, tot, fobs);
            orc_band_fluxes_col(nspec, nb, bt, jlo, jhi, do_band, wtl, fobs, fb);
            double dist = orc_distance_term(c->cz, c->cdl, c->ncosmo, zm);
            for (int b = 0; b < nb; b++) {
                double v = NAN;
                if (do_band[b] != 0) v = mag_from_flux(fb[b]) - AB_ZEROPOINT - 2.5 * amzl + dist;
                mags[(size_t)m * nb + b] = v;
                if (fband_out) fband_out[(size_t)m * nb + b] = fb[b];
            }
        }
        free(p); free(w); free(t1); free(t2); free(y); free(o); free(tot); free(work); free(fobs); free(fb);
        free(slo); free(shi); free(soff); free(scoef); free(tau); free(tdv); free(nlum);
        free(smass); free(slum); free(vy); free(vo); free(wpt); free(llo); free(dl);
    }
    free(jlo); free(jhi); free(do_band); free(bt); free(wtl); free(ltab);
    neb_free(&neb);
    return status;
}

/* ---------------------------------------------------------------------------------------------
 * The reverse pass: batch_backward (opensps_csp_batch.f90:4713-5136) for the stages above, in FP64, as
 * the transposes of the forward maps with the forward's transmissions recomputed:
 *   magnitudes -> band fluxes (mags = -2.5 log10 F + const + distance term; distance_term_dz,
 *     src/spectra/opensps_batch_photometry.f90:798-813)
 *   band fluxes -> observed frame -> the smoothed rest-frame spectrum (backward_observed_frame :6958-7136:
 *     the clamp max(v, 0) passes no gradient; dz through the interpolation fraction)
 *   velocity smoothing: d sigma (smooth_sigma_adjoint, opensps_batch_photometry.f90:627-700:
 *     dC/dsigma = C (x^2 - <x^2>) / (sigma_beta^3 c 1e-13)) and the transpose (smooth_columns_adjoint :604-625)
 *   IGM and attenuation (atten_adjoint_chunk :6476-6665: the dust parameters, frac_obrun, frac_nodust, the
 *     redshift through the IGM optical depth, capped pixels taking the cap pixel's d tau / dz, igm_factor)
 *   the age sums (backward_age_sums_w :5288-5330, backward_age_sums + age_sums_adj_slabs.inc, and the mass
 *     cotangent of batch_mass_cotangent :5142): w_bar = basis^T y_bar + L_bar L_age + M_bar M_age, the
 *     metallicity weight's cotangent into logzsol.
 * params_bar rows (0-based, the reference's BGRAD_* minus 1, :166-213): 1 dust1, 2 dust2, 3 dust3,
 * 4 dust_index, 5 dust1_index, 6 frac_obrun, 7 frac_nodust, 15 logzsol, 22 zred, 23 sigma_smooth,
 * 32 igm_factor and, where the SFH weights are computed here (no ssp_weights or ssp_weights_z), the SFH rows
 * 16 tage, 17 tau, 18 const, 19 fburst, 20 tburst, 21 sf_start, 24 sf_trunc, 25 sf_slope (orc_sfh_gradients); the
 * others are zero. The IGM optical depth is the
 * six-term form (igm_tau_pixel_terms x igm_tau_z_terms), which the reference uses whenever the
 * redshifts differ across the batch.
 * ------------------------------------------------------------------------------------------- */
#define ORC_NBGRAD 44

static const double IGM_EXP[6] = {3.46, 1.68, 3.0 + MP1, 1.5 + MP2, 3.0 - MP3, MP4};

int orc_csp_batch_grad(const orc_ctx_t *c, int n, const double *prm, const double *ssp_weights,
                       const double *mags_bar, const double *spec_bar, const double *mass_bar,
                       const double *lbol_bar, double *mags_out, double *params_bar, double *weights_bar,
                       double *weights_z_bar) {
    const int nspec = c->nspec, nt = c->nt, nz = c->nz, nb = c->nbands;
    if (n < 0 || nspec < 2 || nt < 2 || nz < 1) return 3;
    if (c->ops && c->ssp_weights_z) return 2;
    if (c->add_neb) return 2;   /* the nebular stage's reverse pass is not restated */
    if (n == 0) return 0;
    const double amzl = abs_mag_zeropoint_log(), ln10 = log(10.0);
    int status = 0;
    int *jlo = (int *)malloc(sizeof(int) * (nb > 0 ? nb : 1)), *jhi = (int *)malloc(sizeof(int) * (nb > 0 ? nb : 1));
    int *do_band = (int *)malloc(sizeof(int) * (nb > 0 ? nb : 1));
    double *bt = (double *)malloc(sizeof(double) * (size_t)(nb > 0 ? nb : 1) * nspec);
    double *wtl = (double *)malloc(sizeof(double) * nspec), *ltab = (double *)malloc(sizeof(double) * nspec);
    orc_band_supports(c->bands, nspec, nb, jlo, jhi);
    for (int b = 0; b < nb; b++) {
        do_band[b] = c->mag_mask ? c->mag_mask[b] : 1;
        for (int j = 0; j < nspec; j++) bt[(size_t)j * nb + b] = c->bands[(size_t)b * nspec + j];
    }
    for (int j = 0; j < nspec - 1; j++) wtl[j] = 0.5 * fabs(c->lambda[j + 1] - c->lambda[j]);
    wtl[nspec - 1] = 0.0;
    for (int j = 0; j < nspec; j++) ltab[j] = log(c->lambda[j] / 5500.0);
    int jtop = 0;
    for (int j = 1; j <= nspec; j++) if (c->lambda[j - 1] < 1215.67) jtop = j;
    double *pa = (double *)malloc(sizeof(double) * 6 * (jtop > 0 ? jtop : 1));
    for (int j = 0; j < jtop; j++) orc_igm_pixel_terms(c->lambda[j], pa + 6 * j);
    const int cap = nspec * 64;
    const int wz_on = c->ssp_weights_z != NULL;

    #pragma omp parallel
    {
        double *p = (double *)malloc(sizeof(double) * OPS_NPAR);
        double *w = (double *)malloc(sizeof(double) * nt), *t1 = (double *)malloc(sizeof(double) * nt);
        double *t2 = (double *)malloc(sizeof(double) * nt);
        double *y = (double *)malloc(sizeof(double) * nspec), *o = (double *)malloc(sizeof(double) * nspec);
        double *A = (double *)malloc(sizeof(double) * nspec), *S0 = (double *)malloc(sizeof(double) * nspec);
        double *S1 = (double *)malloc(sizeof(double) * nspec), *tig = (double *)malloc(sizeof(double) * nspec);
        double *tauv = (double *)malloc(sizeof(double) * nspec), *work = (double *)malloc(sizeof(double) * nspec);
        double *fobs = (double *)malloc(sizeof(double) * nspec), *fb = (double *)malloc(sizeof(double) * (nb + 1));
        double *fbb = (double *)malloc(sizeof(double) * (nb + 1)), *vb = (double *)malloc(sizeof(double) * nspec);
        double *g1 = (double *)malloc(sizeof(double) * nspec), *g0 = (double *)malloc(sizeof(double) * nspec);
        double *yb = (double *)malloc(sizeof(double) * nspec), *ob = (double *)malloc(sizeof(double) * nspec);
        int *slo = (int *)malloc(sizeof(int) * nspec), *shi = (int *)malloc(sizeof(int) * nspec);
        int *soff = (int *)malloc(sizeof(int) * nspec);
        double *scoef = (double *)malloc(sizeof(double) * cap);
        double *wbl = (double *)malloc(sizeof(double) * nt), *tsc = (double *)malloc(sizeof(double) * 6 * nt);
        double *smass = NULL, *slum = NULL, *vy = NULL, *vo = NULL, *vyb = NULL, *vob = NULL, *wpt = NULL, *dwp = NULL;
        double *llo = NULL, *dl = NULL, *wfold = NULL;
        if (c->ops) {
            int maxpts = 1;
            for (int a = 0; a < c->ops->nz * nt; a++)
                if (c->ops->age_ptr[a + 1] - c->ops->age_ptr[a] > maxpts) maxpts = c->ops->age_ptr[a + 1] - c->ops->age_ptr[a];
            const size_t ntp = (size_t)c->ops->ntpl;
            smass = (double *)malloc(sizeof(double) * nt); slum = (double *)malloc(sizeof(double) * nt);
            wfold = (double *)malloc(sizeof(double) * nt);
            vy = (double *)malloc(sizeof(double) * ntp); vo = (double *)malloc(sizeof(double) * ntp);
            vyb = (double *)malloc(sizeof(double) * ntp); vob = (double *)malloc(sizeof(double) * ntp);
            wpt = (double *)malloc(sizeof(double) * maxpts); dwp = (double *)malloc(sizeof(double) * 3 * maxpts);
            llo = (double *)malloc(sizeof(double) * maxpts); dl = (double *)malloc(sizeof(double) * maxpts);
        }

        #pragma omp for schedule(dynamic, 4)
        for (int m = 0; m < n; m++) {
            for (int q = 0; q < OPS_NPAR; q++) p[q] = prm[(size_t)q * n + m];
            const int it = model_it_tesc(c, p[OPS_P_DUST_TESC]);
            int zlo = 1;
            double dz = 0.0;
            const int blend = nz > 1 && !wz_on;
            if (blend) model_zlo(c, p[OPS_P_LOGZSOL], &zlo, &dz);
            /* (the reverse pass of the dust emission and the AGN torus is not restated: status 2 where they act) */
            const int emits = (c->add_dust_emission == 1 && (p[OPS_P_DUST1] > SAFE_FLOOR || p[OPS_P_DUST2] > SAFE_FLOOR)) ||
                              (c->add_agn_dust == 1 && p[OPS_P_FAGN] > SAFE_FLOOR);
            if (emits || (!ssp_weights && !wz_on && p[OPS_P_TAGE] <= 0.0) ||
                (!c->smooth_fast && p[OPS_P_SIGMA_SMOOTH] > SAFE_FLOOR)) {
                #pragma omp atomic write
                status = 2;
                continue;
            }
            /* ---- forward ---- */
            double ms = 0.0, ls = 0.0;
            for (int i = 0; i < nspec; i++) { y[i] = 0.0; o[i] = 0.0; }
            const double *wz = wz_on ? c->ssp_weights_z + (size_t)m * nt * nz : NULL;
            double pimf[ORC_IMF_NPAR];
            if (c->ops) {
                /* the IMF per model: the operator's age sums and per-age sums, as orc_csp_batch */
                if (ssp_weights) {
                    for (int k = 0; k < nt; k++) w[k] = ssp_weights[(size_t)m * nt + k];
                } else {
                    int st = orc_sfh_weights(c->time_full, nt, c->tiny_logt, p, w, t1, t2);
                    if (st != 0) {
                        #pragma omp atomic write
                        status = st;
                        continue;
                    }
                }
                model_pimf(c, p, pimf);
                for (int k = 0; k < nt; k++) t1[k] = w[k] > SAFE_FLOOR ? w[k] : 0.0;
                orc_ops_model(c->ops, nspec, c->imf_type, c->remnants, pimf, zlo, dz, t1, it, y, o, smass, slum, vy, vo,
                              wpt, llo, dl, NULL);
                for (int k = 0; k < nt; k++)
                    if (w[k] > SAFE_FLOOR) { ms = ms + w[k] * smass[k]; ls = ls + w[k] * slum[k]; }
            } else if (wz_on) {
                for (int zz = 0; zz < nz; zz++)
                    for (int k = 1; k <= nt; k++) {
                        double wk = wz[(size_t)zz * nt + k - 1];
                        if (wk > SAFE_FLOOR) {
                            ms = ms + wk * c->bmass[(size_t)zz * nt + k - 1];
                            ls = ls + wk * pow(10.0, c->blbol[(size_t)zz * nt + k - 1]);
                        } else {
                            wk = 0.0;
                        }
                        double *acc = k <= it ? y : o;
                        const double *bs = c->basis + ((size_t)zz * nt + k - 1) * nspec;
                        for (int i = 0; i < nspec; i++) acc[i] = acc[i] + wk * bs[i];
                    }
            } else {
                if (ssp_weights) {
                    for (int k = 0; k < nt; k++) w[k] = ssp_weights[(size_t)m * nt + k];
                } else {
                    int st = orc_sfh_weights(c->time_full, nt, c->tiny_logt, p, w, t1, t2);
                    if (st != 0) {
                        #pragma omp atomic write
                        status = st;
                        continue;
                    }
                }
                const double *blo = c->basis + (size_t)(zlo - 1) * nspec * nt;
                const double *bhi = blend ? c->basis + (size_t)zlo * nspec * nt : blo;
                for (int k = 1; k <= nt; k++) {
                    double wk = w[k - 1];
                    if (wk > SAFE_FLOOR) {
                        double mk, lk;
                        if (blend) {
                            mk = (1.0 - dz) * c->bmass[(size_t)(zlo - 1) * nt + k - 1] + dz * c->bmass[(size_t)zlo * nt + k - 1];
                            lk = (1.0 - dz) * c->blbol[(size_t)(zlo - 1) * nt + k - 1] + dz * c->blbol[(size_t)zlo * nt + k - 1];
                            ms = ms + wk * mk;
                            ls = ls + wk * pow(10.0, lk);
                        } else {
                            ms = ms + wk * c->bmass[k - 1];
                            ls = ls + wk * pow(10.0, c->blbol[k - 1]);
                        }
                    } else {
                        wk = 0.0;
                    }
                    double *acc = k <= it ? y : o;
                    const double *bl = blo + (size_t)(k - 1) * nspec, *bh = bhi + (size_t)(k - 1) * nspec;
                    if (blend)
                        for (int i = 0; i < nspec; i++) acc[i] = acc[i] + wk * ((1.0 - dz) * bl[i] + dz * bh[i]);
                    else
                        for (int i = 0; i < nspec; i++) acc[i] = acc[i] + wk * ((1.0 - dz) * bl[i]);
                }
            }
            const double fob = p[OPS_P_FRAC_OBRUN], fnd = p[OPS_P_FRAC_NODUST], ed = p[OPS_P_DUST_INDEX];
            const double dust1 = p[OPS_P_DUST1], dust2 = p[OPS_P_DUST2], dust3 = p[OPS_P_DUST3];
            const double ed1 = p[OPS_P_DUST1_INDEX];
            const int no_birth = dust1 <= SAFE_FLOOR || 1.0 - fob <= SAFE_FLOOR;
            const int no_old = fabs(dust3) <= SAFE_FLOOR;
            const int nodust = 1.0 - fnd > SAFE_FLOOR;
            for (int i = 0; i < nspec; i++) {
                double curve = ed != 0.0 ? exp(ed * ltab[i]) : 1.0;
                double tr_d = exp(-(dust2 * curve)), tr_b = 1.0, tr_o = 1.0;
                if (!no_birth) tr_b = exp(-(dust1 * exp(ed1 * ltab[i])));
                if (!no_old) tr_o = exp(-(dust3 * curve));
                double ssum = (y[i] * tr_b * (1.0 - fob) + y[i] * fob) + o[i] * tr_o;
                A[i] = nodust ? ssum * (tr_d * (1.0 - fnd) + fnd) : ssum;
                S0[i] = A[i];
                tig[i] = 1.0;
                tauv[i] = 0.0;
            }
            const double z = p[OPS_P_ZRED], opz = 1.0 + z, igmf = p[OPS_P_IGM_FACTOR];
            const int igm_on = c->add_igm && z > SAFE_FLOOR && jtop > 0;
            double b6[6], bd6[6];
            int jmax = 1;
            if (igm_on) {
                orc_igm_z_terms(opz, b6);
                for (int k = 0; k < 6; k++) bd6[k] = IGM_EXP[k] * b6[k] / opz;
                double tmax = 0.0;
                for (int j = 0; j < jtop; j++) {
                    const double *a6 = pa + 6 * j;
                    tauv[j] = a6[0] * b6[0] + a6[1] * b6[1] + a6[2] * b6[2] + a6[3] * b6[3] + a6[4] * b6[4] + a6[5] * b6[5];
                }
                for (int j = 1; j <= jtop; j++) if (tauv[j - 1] > tmax) { tmax = tauv[j - 1]; jmax = j; }
                if (jmax > 1) for (int j = 1; j <= jmax; j++) tauv[j - 1] = tmax;
                for (int j = 0; j < jtop; j++) {
                    tig[j] = exp(-tauv[j] * igmf);
                    S0[j] = A[j] * tig[j];
                }
            }
            const double sig = p[OPS_P_SIGMA_SMOOTH];
            int sm_n = 0;
            memcpy(S1, S0, sizeof(double) * nspec);
            if (c->smooth_fast && sig > SAFE_FLOOR) {
                sm_n = orc_velocity_kernel(c->lambda, nspec, sig, p[OPS_P_MIN_WAVE_SMOOTH], p[OPS_P_MAX_WAVE_SMOOTH], slo, shi,
                                           soff, scoef, cap);
                if (sm_n > 0) apply_velocity_kernel(nspec, slo, shi, soff, scoef, S1, work);
            }
            const double zm = c->zred_mags ? c->zred_mags[m] : z;   /* the magnitudes' redshift */
            orc_observed_frame(c->lambda, nspec, zm, S1, fobs);
            orc_band_fluxes_col(nspec, nb, bt, jlo, jhi, do_band, wtl, fobs, fb);
            const double dist = orc_distance_term(c->cz, c->cdl, c->ncosmo, zm);
            double *mg = mags_out ? mags_out + (size_t)m * nb : NULL;
            for (int b = 0; b < nb; b++) {
                double v = NAN;
                if (do_band[b] != 0) v = mag_from_flux(fb[b]) - AB_ZEROPOINT - 2.5 * amzl + dist;
                if (mg) mg[b] = v;
                fbb[b] = 0.0;
                if (mags_bar && v == v && fb[b] > 0.0) fbb[b] = mags_bar[(size_t)m * nb + b] * (-2.5 / ln10) / fb[b];
            }
            /* ---- reverse ---- */
            /* the distance term (distance_term_dz) */
            double dzd = 0.0;
            if (mags_bar && fabs(z) > SAFE_FLOOR && !c->zred_mags) {
                double msum = 0.0;
                for (int b = 0; b < nb; b++) if (fbb[b] != 0.0) msum = msum + mags_bar[(size_t)m * nb + b];
                const double dm = interp_linear(c->cz, c->cdl, c->ncosmo, z);
                if (msum != 0.0 && dm == dm && dm > SAFE_FLOOR) {
                    int i = orc_find_interval(c->cz, c->ncosmo, z);
                    if (i > c->ncosmo - 1) i = c->ncosmo - 1;
                    if (i < 1) i = 1;
                    const double slope = (c->cdl[i] - c->cdl[i - 1]) / (c->cz[i] - c->cz[i - 1]);
                    dzd = msum * (5.0 / ln10 * slope / dm - 2.5 / (ln10 * (1.0 + z)));
                }
            }
            /* band fluxes -> fobs -> v (the clamp passes nothing) */
            for (int j = 0; j < nspec; j++) vb[j] = 0.0;
            for (int b = 0; b < nb; b++) {
                if (fbb[b] == 0.0) continue;
                for (int j = jlo[b]; j <= jhi[b]; j++) {
                    vb[j - 1] += fbb[b] * wtl[j - 1] * bt[(size_t)(j - 1) * nb + b];
                    vb[j] += fbb[b] * wtl[j - 1] * bt[(size_t)j * nb + b];
                }
            }
            double dzo = 0.0;
            for (int j = 0; j < nspec; j++) g1[j] = spec_bar ? spec_bar[(size_t)m * nspec + j] : 0.0;
            for (int j = 0; j < nspec; j++) {
                double v, t = 0.0;
                int il = j + 1;
                const int moved = fabs(zm) > SAFE_FLOOR;
                if (moved) {
                    double tg = c->lambda[j] / (1.0 + zm);
                    il = orc_find_interval(c->lambda, nspec, tg);
                    if (il < 1) il = 1;
                    if (il > nspec - 1) il = nspec - 1;
                    t = (tg - c->lambda[il - 1]) / (c->lambda[il] - c->lambda[il - 1]);
                    v = S1[il - 1] + t * (S1[il] - S1[il - 1]);
                } else {
                    v = S1[j];
                }
                const double wl = c->lambda[j] > DBL_MIN ? 1.0 / c->lambda[j] : 0.0;
                double gv = (v == v && v > 0.0) ? vb[j] * wl : 0.0;
                if (gv == 0.0) continue;
                if (moved) {
                    if (!c->zred_mags)
                        dzo += gv * (S1[il] - S1[il - 1]) * (-c->lambda[j] / ((1.0 + z) * (1.0 + z))) /
                               (c->lambda[il] - c->lambda[il - 1]);
                    g1[il - 1] += (1.0 - t) * gv;
                    g1[il] += t * gv;
                } else {
                    g1[j] += gv;
                }
            }
            /* smoothing: d sigma, then the transpose */
            double dsg = 0.0;
            memcpy(g0, g1, sizeof(double) * nspec);
            if (sm_n > 0) {
                const double sb = sig / (C_LIGHT * 1.0e-13);
                const double kfac = 1.0 / (sb * sb * sb * C_LIGHT * 1.0e-13);
                for (int i = 0; i < nspec; i++) if (slo[i] != 0) g0[i] = 0.0;
                for (int i = 0; i < nspec; i++) {
                    const int a = slo[i];
                    if (a == 0) continue;
                    const int off = soff[i] - a;   /* coef(off + k), 1-based k */
                    const double lli = log(c->lambda[i]);
                    double x2m = 0.0;
                    for (int k = a; k <= shi[i]; k++) {
                        const double x = lli - log(c->lambda[k - 1]);
                        x2m += scoef[off + k - 1] * x * x;
                    }
                    double sx = 0.0;
                    for (int k = a; k <= shi[i]; k++) {
                        const double x = lli - log(c->lambda[k - 1]), cik = scoef[off + k - 1];
                        sx += cik * (x * x - x2m) * S0[k - 1];
                        g0[k - 1] += cik * g1[i];
                    }
                    dsg += g1[i] * sx;
                }
                dsg *= kfac;
            }
            /* IGM and attenuation */
            double d[8] = {0}, d13 = 0.0, d14 = 0.0;
            for (int i = 0; i < nspec; i++) {
                const double ga = g0[i] * tig[i];
                const double curve = ed != 0.0 ? exp(ed * ltab[i]) : 1.0;
                const double td = exp(-(dust2 * curve));
                const double pb = exp(ed1 * ltab[i]);
                const double tb = no_birth ? 1.0 : exp(-(dust1 * pb));
                const double to = no_old ? 1.0 : exp(-(dust3 * curve));
                const double dfac = nodust ? td * (1.0 - fnd) + fnd : 1.0;
                const double ay = tb * (1.0 - fob) + fob;
                yb[i] = ga * ay * dfac;
                ob[i] = ga * to * dfac;
                const double gd = ga * (y[i] * ay + o[i] * to);
                const double gtd = nodust ? gd * (1.0 - fnd) : 0.0;
                const double gtb = ga * y[i] * (1.0 - fob) * dfac;
                const double gto = ga * o[i] * dfac;
                if (igm_on && i < jtop) {
                    const int pp = (i + 1 > jmax ? i + 1 : jmax) - 1;
                    const double *a6 = pa + 6 * pp;
                    const double dtz = a6[0] * bd6[0] + a6[1] * bd6[1] + a6[2] * bd6[2] + a6[3] * bd6[3] + a6[4] * bd6[4] +
                                       a6[5] * bd6[5];
                    d13 += -ga * A[i] * igmf * dtz;
                    d14 += -ga * A[i] * tauv[i];
                }
                if (no_birth) {
                    d[1] += -gtb * pb;
                } else {
                    d[1] += -gtb * pb * tb;
                    d[5] += -gtb * dust1 * pb * ltab[i] * tb;
                }
                d[2] += -gtd * td * curve;
                d[3] += -gto * to * curve;
                d[4] += -(gtd * td * dust2 + gto * to * dust3) * curve * ltab[i];
                d[6] += ga * y[i] * (1.0 - tb) * dfac;
                if (nodust) d[7] += gd * (1.0 - td);
            }
            /* the age sums */
            const double lb = lbol_bar ? lbol_bar[m] : 0.0, mb = mass_bar ? mass_bar[m] : 0.0;
            const double lsb = ls > SAFE_FLOOR ? lb / (ls * ln10) : 0.0;
            double dzb = 0.0, imfb[3] = {0.0, 0.0, 0.0};
            if (c->ops) {
                /* the operator's reverse pass: the fold's and the per-age sums' parts of w_bar, dz, the IMF rows */
                orc_ops_model_adj(c->ops, nspec, c->imf_type, c->remnants, pimf, zlo, dz, t1, it, yb, ob, lsb, mb, slum,
                                  wfold, &dzb, imfb, vyb, vob, wpt, dwp, llo, dl);
                for (int k = 0; k < nt; k++) {
                    wbl[k] = wfold[k] + lsb * slum[k] + mb * smass[k];
                    if (weights_bar) weights_bar[(size_t)m * nt + k] = wbl[k];
                }
            } else if (wz_on) {
                for (int zz = 0; zz < nz; zz++)
                    for (int k = 1; k <= nt; k++) {
                        const double *bs = c->basis + ((size_t)zz * nt + k - 1) * nspec;
                        const double *gb = k <= it ? yb : ob;
                        double s = 0.0;
                        for (int i = 0; i < nspec; i++) s += bs[i] * gb[i];
                        if (weights_z_bar)
                            weights_z_bar[((size_t)m * nz + zz) * nt + k - 1] =
                                s + lsb * pow(10.0, c->blbol[(size_t)zz * nt + k - 1]) + mb * c->bmass[(size_t)zz * nt + k - 1];
                    }
                if (weights_bar) for (int k = 0; k < nt; k++) weights_bar[(size_t)m * nt + k] = 0.0;
            } else {
                const int zl = zlo - 1, zh = blend ? zlo : zlo - 1;
                for (int k = 1; k <= nt; k++) {
                    const double *gb = k <= it ? yb : ob;
                    const double *bl = c->basis + ((size_t)zl * nt + k - 1) * nspec, *bh = c->basis + ((size_t)zh * nt + k - 1) * nspec;
                    double sl = 0.0, sh = 0.0;
                    for (int i = 0; i < nspec; i++) { sl += bl[i] * gb[i]; sh += bh[i] * gb[i]; }
                    double mk, lk, wb;
                    if (blend) {
                        mk = (1.0 - dz) * c->bmass[(size_t)zl * nt + k - 1] + dz * c->bmass[(size_t)zh * nt + k - 1];
                        lk = pow(10.0, (1.0 - dz) * c->blbol[(size_t)zl * nt + k - 1] + dz * c->blbol[(size_t)zh * nt + k - 1]);
                        wb = (1.0 - dz) * sl + dz * sh;
                    } else {
                        mk = c->bmass[k - 1];
                        lk = pow(10.0, c->blbol[k - 1]);
                        wb = (1.0 - dz) * sl;
                    }
                    wbl[k - 1] = wb + lsb * lk + mb * mk;
                    if (weights_bar) weights_bar[(size_t)m * nt + k - 1] = wbl[k - 1];
                    if (blend) {
                        const double wk = w[k - 1] > SAFE_FLOOR ? w[k - 1] : 0.0;
                        dzb += wk * (sh - sl);
                        if (wk > 0.0) {
                            dzb += lsb * wk * lk * ln10 * (c->blbol[(size_t)zh * nt + k - 1] - c->blbol[(size_t)zl * nt + k - 1]);
                            dzb += mb * wk * (c->bmass[(size_t)zh * nt + k - 1] - c->bmass[(size_t)zl * nt + k - 1]);
                        }
                    }
                }
            }
            if (params_bar) {
                double *pb = params_bar + (size_t)m * ORC_NBGRAD;
                for (int r = 0; r < ORC_NBGRAD; r++) pb[r] = 0.0;
                for (int r = 1; r <= 7; r++) pb[r] = d[r];
                if (blend) pb[15] = dzb / (c->logz[zlo] - c->logz[zlo - 1]);
                if (c->ops) {
                    /* imf1, imf2 (BGRAD 27, 28), then imf3 (29) for Kroupa or mdave (31) for Dave */
                    const int ty = c->imf_type % 10;
                    if (ty == 2 || ty == 4) {
                        pb[26] = imfb[0];
                        pb[27] = imfb[1];
                        pb[ty == 2 ? 28 : 30] = imfb[2];
                    }
                }
                pb[22] = dzo + d13 + dzd;
                pb[23] = dsg;
                pb[32] = d14;
                /* the SFH parameters (batch_sfh_gradients, :1163-1166: only where the weights are not given) */
                if (!ssp_weights && !wz_on) {
                    int st = orc_sfh_gradients(c->time_full, nt, c->tiny_logt, p, wbl, pb, tsc);
                    if (st) {
                        #pragma omp atomic write
                        status = st;
                    }
                }
            }
        }
        free(p); free(w); free(t1); free(t2); free(y); free(o); free(A); free(S0); free(S1); free(tig); free(tauv);
        free(work); free(fobs); free(fb); free(fbb); free(vb); free(g1); free(g0); free(yb); free(ob);
        free(slo); free(shi); free(soff); free(scoef); free(wbl); free(tsc);
        free(smass); free(slum); free(wfold); free(vy); free(vo); free(vyb); free(vob); free(wpt); free(dwp);
        free(llo); free(dl);
    }
    free(jlo); free(jhi); free(do_band); free(bt); free(wtl); free(ltab); free(pa);
    return status;
}

/* Thread count the OpenMP runtime uses (for the CPU-baseline report) */
int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
