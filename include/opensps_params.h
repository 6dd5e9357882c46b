/*
 * Per-model parameter slots of the batched composite-population path.
 *
 * A batch carries its per-model parameters as one FP64 array prm[slot * n + m] (slot-major, so that
 * the models of one parameter are contiguous and device reads of a slot are coalesced). The slots
 * are the members of the reference's PARAMS_CORE that the batch reads
 * (src/core/opensps_types.f90:33-45); the public batch call resolves each name of its names[] list
 * to one slot once per call, as the reference's batch_handle resolves each name to one word of
 * PARAMS_CORE (src/bindings/c/opensps_c_api.f90:797-853). Slots a batch does not name keep the
 * context's value for every model. OPS_P_SFH holds the integer sfh type as a double.
 */
#ifndef OPENSPS_PARAMS_H
#define OPENSPS_PARAMS_H

#define OPS_P_TAGE            0
#define OPS_P_TAU             1
#define OPS_P_CONST           2
#define OPS_P_FBURST          3
#define OPS_P_TBURST          4
#define OPS_P_SF_START        5
#define OPS_P_SF_TRUNC        6
#define OPS_P_SF_SLOPE        7
#define OPS_P_SFH             8
#define OPS_P_LOGZSOL         9
#define OPS_P_DUST1          10
#define OPS_P_DUST2          11
#define OPS_P_DUST3          12
#define OPS_P_DUST_INDEX     13
#define OPS_P_DUST1_INDEX    14
#define OPS_P_DUST_TESC      15
#define OPS_P_FRAC_OBRUN     16
#define OPS_P_FRAC_NODUST    17
#define OPS_P_ZRED           18
#define OPS_P_IGM_FACTOR     19
#define OPS_P_SIGMA_SMOOTH   20
#define OPS_P_MIN_WAVE_SMOOTH 21
#define OPS_P_MAX_WAVE_SMOOTH 22
#define OPS_P_IMF1           23
#define OPS_P_IMF2           24
#define OPS_P_IMF3           25
#define OPS_P_IMF_UPPER      26
#define OPS_P_IMF_LOWER      27
#define OPS_P_DUSTE_UMIN     28
#define OPS_P_DUSTE_QPAH     29
#define OPS_P_DUSTE_GAMMA    30
#define OPS_P_FAGN           31
#define OPS_P_AGN_TAU        32
#define OPS_P_GAS_LOGZ       33
#define OPS_P_GAS_LOGU       34
#define OPS_NPAR             35

/* Default value of every slot: the PARAMS_CORE defaults (src/core/opensps_types.f90:33-45). */
#define OPS_PARAM_DEFAULTS { \
    0.0,  /* tage */          1.0,  /* tau */          0.0, /* const */        0.0,   /* fburst */     \
    11.0, /* tburst */        0.0,  /* sf_start */     0.0, /* sf_trunc */     0.0,   /* sf_slope */   \
    0.0,  /* sfh */           0.0,  /* logzsol */      0.0, /* dust1 */        0.0,   /* dust2 */      \
    0.0,  /* dust3 */        -0.7,  /* dust_index */  -1.0, /* dust1_index */  7.0,   /* dust_tesc */  \
    0.0,  /* frac_obrun */    0.0,  /* frac_nodust */  0.0, /* zred */         1.0,   /* igm_factor */ \
    0.0,  /* sigma_smooth */  1.0e3,/* min_wave_smooth */ 1.0e4, /* max_wave_smooth */                \
    1.3,  /* imf1 */          2.3,  /* imf2 */         2.3, /* imf3 */       120.0,   /* imf_upper_limit */ \
    0.08, /* imf_lower_limit */ 1.0, /* duste_umin */  3.5, /* duste_qpah */ 0.01,   /* duste_gamma */ \
    0.0,  /* fagn */          10.0, /* agn_tau */      0.0, /* gas_logz */    -2.0    /* gas_logu */ }

#define OPS_PARAM_NAMES { \
    "tage", "tau", "const", "fburst", "tburst", "sf_start", "sf_trunc", "sf_slope", "sfh", "logzsol", \
    "dust1", "dust2", "dust3", "dust_index", "dust1_index", "dust_tesc", "frac_obrun", "frac_nodust", \
    "zred", "igm_factor", "sigma_smooth", "min_wave_smooth", "max_wave_smooth", "imf1", "imf2", "imf3", \
    "imf_upper_limit", "imf_lower_limit", "duste_umin", "duste_qpah", "duste_gamma", "fagn", "agn_tau", \
    "gas_logz", "gas_logu" }

#endif
