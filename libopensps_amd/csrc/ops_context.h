// The device context of the batched composite-population path: the resident SSP bases and their
// packed matrix-core operands, the wavelength and filter tables, the cosmology table, the options
// of the batch, and the scratch of the last batch. One per public context handle; never shared.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#define OPS_TILE_PX 64     // pixels per workgroup tile of the age-sum products
#define OPS_TILE_COL 64    // model columns per workgroup tile
#define OPS_KSTEP_BF16 32  // K per bf16 matrix instruction (16x16x32)
#define OPS_KSTEP_F64 4    // K per f64 matrix instruction (16x16x4)

struct OpsBuffer {
    void* p = nullptr;
    size_t bytes = 0;
};

struct OpsWPoint;   // ops_launch.h

// A fixed region the batch's scratch is carved from instead of allocated (the no-allocation entry point)
struct OpsArena {
    char* base = nullptr;
    size_t cap = 0, used = 0;
};

struct OpsCtx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;

    // ---- the grid (set_grid) ----
    int nspec = 0, nt = 0, nz = 0;
    int nsp = 0;          // nspec padded to whole pixel tiles
    int npair = 0;        // metallicity pairs (nz - 1), or 1 with one basis
    int kpair = 0;        // K of one pair's product: 2 nt (blend) or nt
    int k64 = 0, kbf = 0; // kpair padded to the f64 / bf16 instruction K
    std::vector<double> h_lam, h_logz, h_time, h_basis, h_mass, h_lbol;
    double *d_lam = nullptr, *d_loglam = nullptr, *d_ltab = nullptr, *d_wl = nullptr, *d_dlam = nullptr;
    double *d_time = nullptr, *d_logz = nullptr, *d_bmass = nullptr, *d_blbol = nullptr;
    double* d_rdlam = nullptr;    // 1 / (lam[j + 1] - lam[j]) (0 at the last pixel)
    float *d_ltab32 = nullptr, *d_wl32 = nullptr;   // FP32 copies for the FP32 chain
    float4* d_sm32 = nullptr;     // per pixel {1 / lam, dlam[j - 1] / 2, dlam[j] / 2, 0}: smoothing weights, FP32
    int nbk = 0;                  // buckets of the wavelength lookup
    double lnl0 = 0.0, bkinv = 0.0;   // bucket b starts at ln(lam) = lnl0 + b / bkinv
    int* d_bkt = nullptr;         // per bucket: the number of pixels with ln(lam) below the bucket's start
    int jtop = 0;                 // pixels below Lyman-alpha
    int jq = 4;                   // jtop rounded up to a multiple of 4: the stride of d_igm64s, d_igm32s
    double* d_igm_a = nullptr;    // (jtop, 6) IGM wavelength factors
    double* d_igm64s = nullptr;   // the same term-major (6, jq): the chain's pixel-parallel reads
    float* d_igm32s = nullptr;    // and in FP32 (the FP32 chain)
    double* d_basis64 = nullptr;  // nspec x nt x nz FP64 (for the packing)
    double* d_A64 = nullptr;      // packed f64 operands [pair][pxblk][kstep][lane]
    uint16_t *d_Ahi = nullptr, *d_Alo = nullptr;   // packed bf16 hi / lo [pair][pxblk][kstep][lane][8]

    // ---- the SSP operator with the IMF per model (set_ssp_ops) ----
    struct OpsOperator {
        bool on = false;
        int nz = 0, nt = 0, npts = 0, ntpl = 0, npz = 0;   // npz: the most points of one metallicity
        int npa_max = 1;                                   // the most points of one (metallicity, age)
        int k64 = 0, kbf = 0;                              // ntpl padded to the f64 / bf16 instruction K
        int *d_age_ptr = nullptr, *d_zoff = nullptr, *d_first = nullptr, *d_tz_ptr = nullptr;
        double *d_pt = nullptr, *d_age_mmax = nullptr;     // points: OPS_PT_N doubles each
        int2* d_te = nullptr;                              // entries by template: {point in its metallicity, age}
        int2* d_seg = nullptr;                             // segments (<= OPS_SEG entries of one template): entry range
        int *d_segoff = nullptr, *d_tseg = nullptr;        // first segment per metallicity; segments per template
        int nseg_max = 0;                                  // the most segments of one metallicity
        double* d_te_val64 = nullptr;
        float* d_te_val32 = nullptr;
        double* d_A64 = nullptr;                           // the library as the product's A operand
        uint16_t *d_Ahi = nullptr, *d_Alo = nullptr;
        // the bins off the common path for the context's IMF type (first bins, bins across a break, empty bins):
        // per (metallicity, age) CSR of point indices, rebuilt when the IMF type or mdave changes
        std::vector<int> h_age_ptr, h_first;
        std::vector<double> h_pt;                          // the points as uploaded (OPS_PT_N doubles each)
        int *d_slow_ptr = nullptr, *d_slow_pt = nullptr;
        OpsWPoint* d_wp = nullptr;                         // k_op_weights' records
        double* d_poly = nullptr;                          // k_op_weights' expm1 coefficients
        int nslow = 0;
        // the reverse pass: each point's template entries; the library as the transposed product's rows (template,
        // pixel) in FP64 (nrt x 64 rows, zero past ntpl) and, once a mixed reverse pass asks, split BF16
        int nent = 0, nrt = 0;
        int* d_pent = nullptr;
        double* d_pval = nullptr;
        double* d_At64 = nullptr;
        uint16_t *d_Ath = nullptr, *d_Atl = nullptr;
    } op;
    int imf_type = 2, remnants = 1;
    double salp_ind = 2.35, mdave = 0.5, mlim_bh = 40.0, mlim_ns = 8.5;

    // ---- filters (set_bands) ----
    int nbands = 0;
    int band_j0 = 0, band_j1 = -1;                 // 0-based pixel range the band integrals read (whole quads)
    int4* d_bmeta = nullptr;                       // per band: {coefficient offset, first pixel, pixels, 0}
    double* d_bc64 = nullptr;                      // per (band, pixel of its support) coefficient, FP64
    float* d_bc32 = nullptr;                       // the same in FP32
    int* d_mask = nullptr;                         // per band: computed (1) or NaN (0)
    int *d_bp_ptr = nullptr, *d_bp_band = nullptr; // the band coefficients by pixel over [band_j0, band_j1] (CSR)
    double* d_bp_c64 = nullptr;
    std::vector<int> h_mask;                       // the caller's mask (the device's adds the Vega reference band)
    // the Vega system (set_vega): per band Vega's AB magnitude, the 1-based reference band (0: none) and its value
    double* d_magvega = nullptr;
    int vega_on = 0, vega_ref = 0;
    // the SFH table of sfh 2, 3 (set_sfh_tab): ntab forward ages (yr), then ntab SFRs
    double* d_tab = nullptr;
    int ntab = 0;
    // the diffuse attenuation curve of dust types 2, 6 per padded pixel (set_dust_curve), FP64 and FP32
    double* d_dcurve = nullptr;
    float* d_dcurve32 = nullptr;
    int dcurve_on = 0;
    double vega_anchor = 0.0;

    // ---- dust emission and the AGN torus (set_emission, set_dust_emission, set_agn) ----
    int add_duste = 0, add_agn = 0;
    int nqpah = 0, numin = 0, ntau = 0;
    double *d_qpah = nullptr, *d_umin = nullptr, *d_dustem = nullptr;   // templates (nsp, nqpah, 2 numin)
    double *d_agn_tau = nullptr, *d_agn = nullptr;                      // templates (nsp, ntau)
    double* d_wnu = nullptr;   // per pixel the frequency trapezoid's weight wt(i - 1) + wt(i), wt(i) = |c/l(i+1) - c/l(i)| / 2

    // ---- nebular emission with the CLOUDY grids (set_nebular, set_neb_options) ----
    int add_neb = 0, neb_cont_on = 1, neb_lines_in_spec = 1;
    int nnz = 0, nna = 0, nnu = 0, nline = 0;      // the tables' axes (0: no tables)
    int nt_neb = 0, whlylim = 0;                    // the nebular ages (1..nt_neb) and the pixels below 912 A
    double *d_neb_lz = nullptr, *d_neb_lu = nullptr;             // gas log Z, log U axes
    double *d_neb_cont = nullptr, *d_neb_line = nullptr;         // (nnu, nna, nnz, nspec) / (..., nline), log10
    double *d_neb_lpos = nullptr;                                // the lines' wavelengths
    int *d_neb_ia = nullptr, *d_neb_newr = nullptr, *d_neb_llo = nullptr;   // per age / per line
    double *d_neb_wa = nullptr, *d_neb_dwa = nullptr;
    double* d_neb_q0 = nullptr;                     // nz x nt_neb: ionizing photons of the basis spectra
    std::vector<double> h_neb_lpos;
    // the lines' profiles for one width (the first model's sigma_smooth): CSR by pixel
    double neb_sigma = -1.0;
    int neb_nprof = 0;
    int *d_neb_lptr = nullptr, *d_neb_lid = nullptr;
    double* d_neb_gval = nullptr;

    // ---- cosmology ----
    int ncosmo = 0;
    double *d_cz = nullptr, *d_cdl = nullptr;

    // ---- options ----
    int add_igm = 0;       // add_igm_absorption
    int smooth_fast = 1;   // smooth_velocity == 1 and smoothspec_fast == 1
    double tiny_logt = 0.0;
    bool f64_vector = false;   // FP64 age sums with vector FMAs (OPENSPS_F64_GEMM=vector)
    bool chain_two = false;    // the two-buffer chain always (OPENSPS_CHAIN=two)
    int op_chunk = 1 << 30;    // column tiles the operator's weights, segments and fold take per round of launches

    // ---- batch scratch (grown on demand) ----
    OpsArena* arena = nullptr;   // when set, ops_reserve carves from it and never allocates
    OpsBuffer yo, wcol, w64, order, tiles, misc, spec, flag;
    OpsBuffer hdr;   // per-model constants of the chain (OPS_HDR_N doubles per model)
    OpsBuffer nebw;  // the SFH weights (n x nt) for the nebular stage when the caller does not ask for them
    OpsBuffer wsfh, opu, opwc, opsum, opseg;   // the operator path: SFH weights, point weights, age factors,
                                               // per-age sums, segment sums
    OpsBuffer opimf, opspart;                  // per-model IMF constants; the off-path bins' per-age sums and counts
    OpsBuffer opgp;                            // per (pass, tile slot, age, point group): the weights' partial sums
    // ---- weights over ages and metallicities, and the reverse pass ----
    struct Full {
        int kfull = 0, k64 = 0, kbf = 0, nrt = 0;  // K = nt nz; padded to the products' K; row tiles of 64
        double* A64 = nullptr;                      // the basis as the forward product's packed A (K = every (age, Z))
        uint16_t *Ahi = nullptr, *Alo = nullptr;
        double* At64 = nullptr;                     // the basis as the transposed product's rows (k, pixel)
        uint16_t *Ath = nullptr, *Atl = nullptr;
    } full;
    OpsBuffer glsum, gfband, gw, gout, gyb, gC;   // reverse pass: L_bol sums, band fluxes, SFH weights, per-model
                                                  // scalars, y_bar rows, the transposed product
    OpsBuffer gwb;                                // reverse pass: d L / d w for the SFH rows (no weights_bar given)
    double* lsum_out = nullptr;                   // set by the reverse pass around its forward batch
    double* op_sraw_out = nullptr;                // the same: the operator's per-pass sums before the blend
    OpsBuffer opsraw, opct, oppart, opimfd;       // the operator's reverse pass: raw per-age sums, template
                                                  // cotangents by position, per-(pass, age) sums, IMF constants
    int* d_last_flag = nullptr;   // the per-batch status flag of the last batch (read by sync)
    int* h_flag = nullptr;        // pinned host copy of that flag, written in stream order at the batch's end
    bool flag_async = false;      // h_flag holds the last batch's flag once the stream has drained

    // ---- per-kernel timing (HIP events on the context's stream) ----
    static const int NPROF = 4;     // prep, group, age-sum product, chain
    bool profiling = false;
    hipEvent_t ev[NPROF + 1] = {};
    bool ev_pending = false;
    double prof_ms[NPROF] = {0, 0, 0, 0};
    long prof_n[NPROF] = {0, 0, 0, 0};
    // the reverse pass: [reverse chain, transposed product, weights / params_bar]
    static const int NGPROF = 3;
    hipEvent_t gev[NGPROF + 1] = {};
    bool gev_pending = false;
    double gprof_ms[NGPROF] = {0, 0, 0};
    long gprof_n[NGPROF] = {0, 0, 0};
    // the operator's phases: [the IMF weights (k_op_imf, k_op_slow, k_op_weights), the fold (k_op_agesums,
    // k_op_segments, k_op_fold, k_op_sums)], timed when the operator's launches take one round
    static const int NOPROF = 2;
    hipEvent_t opev[NOPROF + 1] = {};
    bool opev_pending = false;
    double opprof_ms[NOPROF] = {0, 0};
    long opprof_n[NOPROF] = {0, 0};

    std::string last_error;
};

int ops_fail(OpsCtx* c, int code, const char* fmt, ...);
int ops_reserve(OpsCtx* c, OpsBuffer& b, size_t bytes);
