// Argument blocks of the kernels (passed by value) and their launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ops_context.h"

// Per-model constants of the chain, formed once per model by k_prep (FP64): hdr[m * OPS_HDR_N + slot]
#define OPS_HDR_RZ 0       // 1 / (1 + z)
#define OPS_HDR_LNOPZ 1    // ln(1 + z)
#define OPS_HDR_INV2 2     // 1 / (2 sigma_ln^2) of the velocity kernel
#define OPS_HDR_FBLUE 3    // exp(-4 sigma_ln), exp(+4 sigma_ln): the window edges
#define OPS_HDR_FRED 4
#define OPS_HDR_DIST 5     // distance modulus term of the magnitudes
#define OPS_HDR_TMAX 6     // the IGM optical depth's maximum below Lyman-alpha and its 1-based pixel
#define OPS_HDR_JMAX 7
#define OPS_HDR_W0 8       // smoothed pixels [w0, w1): minl <= lambda <= maxl
#define OPS_HDR_W1 9
#define OPS_HDR_B6 10      // the six (1 + z) powers of the IGM optical depth
#define OPS_HDR_DZ 16      // the metallicity blend weight dz (not clamped)
// dust emission (dl07_indices, opensps_csp_batch.f90:1611-1626) and the AGN torus (agn_indices :1591-1606), formed
// by k_emission and read by the chain with its emission stages (EM)
#define OPS_HDR_DL_IQ 17   // the q_PAH interval (0-based lower column) and its weight (not clamped)
#define OPS_HDR_DL_WQ 18
#define OPS_HDR_DL_IU 19   // the U_min interval and its weight (not clamped)
#define OPS_HDR_DL_WU 20
#define OPS_HDR_DL_GAM 21  // gamma clamped to [0, 1]
#define OPS_HDR_DL_ON 22   // 1: the model emits (add_dust_emission with dust1 or dust2 > 0)
#define OPS_HDR_DL_RN 23   // the emission's normalisation s_abs / s_esc (k_dust_sums; 0: none)
#define OPS_HDR_AGN_IT 24  // the agn_tau interval (0-based lower column) and its weight (clamped)
#define OPS_HDR_AGN_WT 25
#define OPS_HDR_AGN_L 26   // fagn L_bol (the linear L_bol sum through its logarithm; 0: no AGN)
#define OPS_HDR_LSUM 27    // the linear L_bol sum (k_prep, or k_op_sums with the operator)
#define OPS_HDR_N 28

// The wavelength lookup: the ascending grid, and buckets over ln(lambda) no wider than its finest pixel
struct LamLookup {
    int nspec, nbk;
    double lnl0, bkinv;
    const double* lam;
    const int* bkt;
};

struct PrepArgs {
    int n, nt, nz, parts, kpair, kpad, mixed, smooth_fast, ops;   // ops: the operator path (no W columns, no sums)
    double tiny_logt;
    const double *prm, *weights_in, *time, *logz, *bmass, *blbol;
    const double* wz;           // weights over ages and metallicities (nt nz per model) or NULL
    double* lsum;               // per model the L_bol sum before the log (the reverse pass), or NULL
    const double* zmags;        // per model the magnitudes' redshift (zred_mags) or NULL: zred
    double *w_out, *mass, *lbol, *w64;
    uint16_t *whi, *wlo;
    int *zlo, *it, *idx, *status, *flag;
    int *gcnt, *rank;           // per pair the models counted so far; per model its rank in its pair (k_group)
    const double* tab;          // the SFH table of sfh 2, 3: ntab forward ages (yr), then ntab SFRs (M_sun / yr)
    int ntab;
    // the chain's per-model constants
    int add_igm, jtop, ncosmo;
    LamLookup lk;
    const double *igm_a, *cz, *cdl;
    double* hdr;
};

#define OPS_GCNT_SLOTS 32
#define OPS_NTABMAX 20000   // rows of the SFH table (NTABMAX, src/core/opensps_constants.f90:147)   // counters a metallicity pair (k_prep's ranks, k_group's prefix)

struct GroupArgs {
    int n, npair, parts;
    const int *zlo, *gcnt, *rank;   // from k_prep: per pair its models, per model its rank among them
    int *order, *ntiles;
    int4* tiles;
    int* nmtiles;    // the tiles of k_mm_bf16x3 (null: none)
    int4* mtiles;
};

struct GemmArgs {
    int n, parts, nsp, npxt, npxblk, kpad;
    const int* order;
    const int* ntiles;
    const int4* tiles;
    const double *A64, *w64;
    void* yo;
};

// The split-BF16 products with LDS tiles (k_mm_bf16x3): C(row, col) = sum_k A(row, k) B(col, k), 128 rows x 128
// columns a workgroup. A in instruction order [slab][16-row block][K / 32][lane][8] (hi and lo apart); B by
// columns, K contiguous; C by columns, rows contiguous (FP32).
#define OPS_MM_ROWS 128
#define OPS_MM_COLS 128
#define OPS_MM_TS 132   // floats a column of k_mm_bf16x3's C tile takes in LDS in its epilogue (128 rows + 16 B skew)
#define OPS_MM_GROUPED 0   // columns cc of a tile {slab, first, count}: model order[cc / parts], part cc % parts
#define OPS_MM_DIRECT 1    // columns 0 .. ncols - 1 as they are

struct MMArgs {
    int mblk;                   // 16-row blocks of A (and the rows of C written: 16 mblk)
    int nrt;                    // row tiles (ceil(mblk / 8))
    int ksteps;                 // K / 32
    int mode;
    int n, parts;               // grouped: B column model * parts + part, C column part * n + model
    const int* order;
    const int* ntiles;          // grouped: the number of column tiles (device)
    const int4* tiles;          // grouped: {slab, first column, columns, 0}
    int ncols, nct;             // direct
    const uint16_t *Ah, *Al;
    size_t slab;                // A elements per slab
    const uint16_t *Bh, *Bl;
    size_t ldb;                 // B elements per column
    float* C;
    size_t ldc;                 // C elements per column
};

__global__ void k_mm_bf16x3(MMArgs a);

struct PackArgs {
    int nspec, nt, nz, npair, kpair, kpad, npxblk;
    const double* basis;
    double* A64;
    uint16_t *Ahi, *Alo;
};

// ---- the SSP operator with the IMF per model ----
#define OPS_PT_N 8         // per point: mini, m1, m2, ln m1, ln(m2/m1), surviving mass, L_bol, ionizing photons
#define OPS_PIECE 32       // points per piece of an age (the IMF weights restart their chain of shared edges)
#define OPS_TPL_BLOCK 8    // templates per wavefront of the fold's sums
#define OPS_SEG 64         // entries per segment of a template's list (one wavefront of the fold each)

// One isochrone point as k_op_weights reads it (64 bytes, one scalar load): its initial mass, ln m1, ln(m2 / m1),
// surviving mass and L_bol per star, and code = the power-law segment (1..3; 0: a bin k_op_slow takes) | 4 when
// lo^e comes down the chain from the previous common bin (same piece of 32, same segment, shared edge)
struct alignas(64) OpsWPoint {
    double mini, llo, dl, cm, cl;
    int code, pad0;
    double pad1, pad2;
};

#define OPS_OP_GPTS (4 * OPS_PIECE)   // points of one age a wavefront of k_op_weights takes (whole pieces)

struct OpWArgs {           // k_op_weights, k_op_slow, k_op_agesums
    int n, nt, npass, npz, maxtiles, imf_type, remnants;
    int tile0, utiles;     // the round's first column tile; the tile slots of u and gsum (tile - tile0)
    int ngroup;            // point groups of OPS_OP_GPTS per age
    int u_f32;             // u and wc hold floats (mixed precision) or doubles
    double salp_ind, mdave, mlim_bh, mlim_ns;
    const int* ntiles;
    const int4* tiles;
    const int* order;
    const double *prm, *hdr, *wsfh;
    const int *age_ptr, *zoff, *first;
    const double *pt, *age_mmax;
    void *u, *wc;          // per (pass, tile slot): the points' weights (npz x 64); per (pass, tile): the ages' SFH x
                           // blend factors (nt x 64)
    double *smass_p, *slum_p;   // (pass, model, age)
    int* count;                 // (2, n) points with a nonzero weight per pass, or NULL
    const OpsWPoint* wp;        // per point: k_op_weights' record for the context's IMF type
    const double* poly;         // 1 / (j + 1)!, j = 0 .. 12
    double* impc;               // per model OPS_IMFC_N: 1 / total mass integral, the segments' e, c, c / total
    const int *slow_ptr, *slow_pt;   // per (metallicity, age) the bins that leave the common path
    double* spart;              // (pass, tile, age): their sums and the remnants, 3 x 64 per wavefront
    int* scnt;                  // (pass, tile, age): their nonzero weights, 64 per wavefront
    double* gsum;               // (pass, tile slot, age, group): the common bins' mass and L_bol sums, 2 x 64
    int* gcnt;                  // (pass, tile slot, age, group): their nonzero weights, 64
    double* sraw;               // (pass, model, age) x 2: the mass and L_bol sums before the blend, or NULL
};
// The power-law segment of a bin [lo, hi] (imf_bin_weights_tab's classification, :483-498): 0 when the bin
// crosses a break or lies below 0.08 M_sun (the general integral then)
__host__ __device__ inline int bin_class(int t, double lo, double hi, double mdave) {
    if (t == 0) return 1;
    if (t == 2) {
        if (lo >= 0.08 && hi <= 0.5) return 1;
        if (lo >= 0.5 && hi <= 1.0) return 2;
        if (lo >= 1.0) return 3;
        return 0;
    }
    if (lo >= 0.08 && hi <= mdave) return 1;
    if (lo >= mdave) return 2;
    return 0;
}

#define OPS_IMFC_INV_TOTAL 0
#define OPS_IMFC_E 1
#define OPS_IMFC_C 4
#define OPS_IMFC_CT 7
#define OPS_IMFC_N 10

struct OpGArgs {           // k_op_segments: one segment's entries per wavefront
    int nt, npass, npz, maxtiles, nseg_max;
    int tile0, utiles;     // the round's first column tile; the tile slots of u and part
    const int* ntiles;
    const int4* tiles;
    const int *order, *it;
    const int* segoff;     // nz + 1: first segment of each metallicity
    const int2* seg;       // per segment its entry range
    const int2* te;        // per entry {point (local to its metallicity), age}
    const double* te_val64;
    const float* te_val32;
    const void *u, *wc;
    void* part;            // (pass, tile, segment): per lane {young, old}
};

struct OpFArgs {           // k_op_fold: per template the sums of its segments, as the product's B operand
    int ntpl, npass, maxtiles, nseg_max, kpad, parts;
    int tile0, utiles;     // the round's first column tile; the tile slots of part
    const int* ntiles;
    const int4* tiles;
    const int* order;
    const int* tseg;       // nz x (ntpl + 1): segments of each template per metallicity
    const void* part;
    uint16_t *whi, *wlo;   // the B operand of the template product (split BF16), [column][kpad]
    double* w64;           // or FP64
};

struct OpSArgs {           // k_op_sums
    int n, nt, npass;
    const double *wsfh, *smass_p, *slum_p;
    double *mass, *lbol;
    double* lsum;          // (n) the linear L_bol sum (the reverse pass's), or NULL
    double* hdr;           // the L_bol sum into the chain's header
};

// The operator's reverse pass (k_op_imf_grad, k_op_adj_t, k_op_adj_points, k_op_adj_slow, k_op_adj_out)
#define OPS_IMFD_N 9       // per model: d ln(c_seg / total) / d direction, segment-major (3 segments x 3 directions)
struct OpAdjArgs {
    int n, nt, nz, npass, ntpl, nent, imf_type, remnants;
    double salp_ind, mdave, mlim_bh, mlim_ns;
    const int* ntiles;
    const int4* tiles;
    const int *order, *it, *zlo;
    const double *prm, *hdr, *wsfh, *gout, *mass_bar, *logz;
    const double* impc;         // k_op_imf's per-model constants
    double* impd;               // per model OPS_IMFD_N
    const int *age_ptr, *first, *slow_ptr, *slow_pt;
    const double *pt, *age_mmax;
    const OpsWPoint* wp;
    const double* poly;         // 1 / (j + 1)!, j = 0 .. 12
    const int* pent;            // per point nent template indices (-1: none)
    const double* pval;         // per point nent values
    const double *smass_p, *slum_p, *sraw;
    const void* CT;             // (part, template, position): the template cotangents, kernel precision
    double* part;               // (pass, age, position) x 4: the fold's sum and the three IMF sums (scaled)
    double* gpart;              // (pass, age, block of 4 point groups, quantity, position): k_op_adj_points's sums
    int ngroup;                 // point groups of OPS_OP_GPTS per age
    double *params_bar, *weights_bar;
};

struct ChainArgs {
    int n, nspec, nsp, parts, nbands, jtop, add_igm, smooth_fast, band_j0, band_j1, sband_staged;
    int jq;   // the stride of igm64s / igm32s: jtop rounded up to a multiple of 4
    double amzl;
    LamLookup lk;
    const double* prm;
    const double* hdr;
    const void* yo;
    const double *lam, *loglam, *ltab, *wl, *dlam, *rdlam;
    const double* igm64s;   // the IGM wavelength factors term-major (6, jq), FP64 and FP32
    const float* igm32s;
    const float *ltab32, *wl32;
    const float4* sm32;
    const int4* bmeta;   // per band {coefficient offset, first pixel, pixels}, multiples of 4
    const int* mask;
    const double* bc64;
    const float* bc32;
    void *spec, *young, *old;
    void* stage;   // n x nspec staging rows of wide smoothing windows without a spectrum out (single-buffer chain)
    double *mags, *fband;
    // the diffuse attenuation curve per padded pixel (a fixed curve: dust types 2, 6) or NULL (the power law)
    const double* dcurve;
    const float* dcurve32;
    const double* zmags;   // per model the magnitudes' redshift (zred_mags) or NULL: zred
    // dust emission (the Draine & Li templates (nsp, nqpah, 2 numin)) and the AGN torus templates (nsp, ntau)
    const double *dustem, *agn;
    int nqpah, numin, ntau;
};

// k_emission: per model the dust emission's template indices and energy balance (dl07_indices, dust_emission_reductions,
// opensps_csp_batch.f90:1611-1626, 3963-4042) and the AGN torus's (agn_indices :1591-1606) into the chain's header;
// mdust; status 2 for a model that needs templates the context lacks
struct DustArgs {
    int n, nspec, nsp, parts, nqpah, numin, ntau, duste_on, agn_on;
    const double *prm, *ltab, *dcurve, *dustem, *wnu, *qpah, *umin, *agn_tau;
    const float *ltab32, *dcurve32;
    const void* yo;
    double* hdr;
    double* mdust;         // (n) or NULL
    int *status, *flag;    // per-model status and the batch's flag (as k_prep's)
};

// k_neb: per model the nebular emission with the CLOUDY grids (nebular_kernels steps 2-5, opensps_csp_batch.f90:
// 2612-2998, line_profiles :3133-3160) into the age sums, and the lines' luminosities through the dust (line_attenuated
// :3440-3473); k_neb_q0: step 1 once per grid and tables
struct NebArgs {
    int n, nspec, nsp, nt, nz, parts, nt_neb, whlylim, nline, nnz, nna, nnu, cont_on, lines_in_spec;
    double sigma0;                 // the profiles' width (the first model's sigma_smooth)
    const double *prm, *hdr, *w;   // parameters, the chain's header (dz), the SFH weights (n x nt)
    const int *zlo, *it;
    const double *basis;           // nspec x nt x nz FP64
    const double *q0;              // nz x nt_neb
    const double *lz, *lu, *cont, *line, *lpos;
    const int *ia, *newr, *llo;
    const double *wa, *dwa;
    const int *lptr, *lid;         // the profiles (CSR by pixel)
    const double* gval;
    const double *lam, *ltab, *dcurve;
    void* yo;
    double* emlines;               // (n, nline) or NULL
    int *status, *flag;
};

// ---- the reverse pass (ops_grad.hip) ----
#define OPENSPS_HIP_NBGRAD_ROWS 44   // params_bar rows (the reference's NBGRAD, opensps_csp_batch.f90:213)
#define OPS_GOUT_N 16   // per model: the reverse chain's scalars (k_grad_chain -> k_grad_out)
#define OPS_GOUT_C 0        // cscale: the model's cotangents are divided by it (FP32 range), results times it
#define OPS_GOUT_LSB 1      // dL / dL_bol-sum (scaled)
#define OPS_GOUT_D 2        // 2..8: dust1, dust2, dust3, dust_index, dust1_index, frac_obrun, frac_nodust (scaled)
#define OPS_GOUT_DZO 9      // the observed frame's dz (scaled)
#define OPS_GOUT_DZI 10     // the IGM's dz (scaled)
#define OPS_GOUT_DIGM 11    // igm_factor (scaled)
#define OPS_GOUT_DSG 12     // sigma_smooth (scaled)
#define OPS_GOUT_DZD 13     // the distance term's dz (not scaled)

struct GradArgs {
    int n, nspec, nsp, parts, nbands, jtop, add_igm, band_j0, band_j1, ncosmo;
    LamLookup lk;
    const double *prm, *hdr;
    const void* yo;
    const double *lam, *loglam, *ltab, *wl, *dlam, *rdlam, *igm_a, *cz, *cdl;
    const float4* sm32;
    const int* mask;
    const int *bp_ptr, *bp_band;   // band coefficients by pixel over [band_j0, band_j1] (CSR)
    const double* bp_c64;
    const double *mags, *fband, *lsum;
    const double *mags_bar, *spec_bar, *lbol_bar;
    void *ybh, *ybl;               // y_bar rows (model part) x nsp: bf16 hi / lo (mixed) or FP64 (ybh)
    double* gout;                  // per model OPS_GOUT_N
    const double* zmags;           // per model the magnitudes' redshift (zred_mags) or NULL: zred
};

struct GemmTArgs {                 // C(col, k) = sum_j At(k, j) Yb(col, j)
    int krows, nrt, ncols, nct, nsp, ldc;
    const double *A64, *B64;
    void* C;                       // (ncols, ldc) FP32 or FP64
};

struct GradOutArgs {
    int n, nt, nz, parts, kfull, ldc, wz;
    const int *zlo, *it;
    const double *hdr, *logz, *bmass, *blbol, *w, *mass_bar, *gout;
    const void* C;
    int mixed;
    double *params_bar, *weights_bar, *weights_z_bar;
};

template <typename T>
__global__ void k_grad_chain(GradArgs a);
__global__ void k_gemmt_f64(GemmTArgs a);
__global__ void k_grad_out(GradOutArgs a);

// params_bar's SFH rows from d L / d w (k_sfh_grad, one wavefront per model)
struct SfhGradArgs {
    int n, nt;
    double tiny_logt;
    const double *time, *prm, *wbar;
    double* params_bar;
};
__global__ void k_sfh_grad(SfhGradArgs a);
__global__ void k_vega_mags(int n, int nb, double* mags, const double* magvega, int vref, double anchor);
__global__ void k_pack_t(int nspec, int nsp, int kfull, int nrows, const double* basis, double* A64, uint16_t* Ah,
                         uint16_t* Al);

__global__ void k_prep(PrepArgs a);
__global__ void k_group(GroupArgs a);
__global__ void k_gemm_f64(GemmArgs a);
__global__ void k_gemm_f64_vec(GemmArgs a);
__global__ void k_pack_f64(PackArgs a);
__global__ void k_pack_bf16(PackArgs a);
#define OPS_CHAIN_KS 4   // smoothed pixels a thread holds in registers (single-buffer chain)
#define OPS_CHAIN_KQ 3   // observed-frame quads a thread holds in registers (single-buffer chain)
#define OPS_CHAIN_KW 2   // smoothing windows a thread looks up before the spectrum is formed
#define OPS_CHAIN_KI 3   // IGM transmissions a thread forms before the spectrum is (pixels tid + k 512)
template <typename T, bool ONE, bool EM>
__global__ void k_chain(ChainArgs a);
template <typename T>
__global__ void k_band_fluxes(int n, int nspec, int nbands, const T* bt, const int* jlo, const int* jhi, const int* do_band,
                              const T* wt, const T* fobs_t, double* fband);
template <typename T>
__global__ void k_observed_frame(int n, int nspec, const double* zc, const double* lam, const T* total, T* fobs_t);
__global__ void k_selftest_mfma(int which, const void* a, const void* b, void* d);
template <typename T>
__global__ void k_op_weights(OpWArgs a);
template <typename T>
__global__ void k_op_slow(OpWArgs a);
__global__ void k_op_imf(OpWArgs a);
__global__ void k_op_agesums(OpWArgs a);
template <typename T>
__global__ void k_op_segments(OpGArgs a);
template <typename T>
__global__ void k_op_fold(OpFArgs a);
__global__ void k_op_sums(OpSArgs a);
__global__ void k_op_imf_grad(OpAdjArgs a);
template <typename T>
__global__ void k_emission(DustArgs a);
template <typename T>
__global__ void k_neb(NebArgs a);
__global__ void k_neb_q0(int nspec, int nt, int nt_neb, int whlylim, const double* lam, const double* basis, double* q0);
template <typename T>
__global__ void k_op_adj_t(int n, int ntpl, int ldc, const int* order, const T* C, T* CT);
template <typename T>
__global__ void k_op_adj_points(OpAdjArgs a);
template <typename T>
__global__ void k_op_adj_slow(OpAdjArgs a);
__global__ void k_op_adj_out(OpAdjArgs a);
__global__ void k_op_tiles(int ncols, int width, int* ntiles, int4* tiles);
