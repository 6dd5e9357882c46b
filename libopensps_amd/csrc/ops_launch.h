// Argument blocks of the kernels (passed by value) and their launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ops_context.h"

struct PrepArgs {
    int n, nt, nz, parts, kpair, kpad, mixed, smooth_fast;
    double tiny_logt;
    const double *prm, *weights_in, *time, *logz, *bmass, *blbol;
    double *w_out, *mass, *lbol, *w64;
    uint16_t *whi, *wlo;
    int *zlo, *it, *idx, *status, *flag;
};

struct GroupArgs {
    int n, npair, parts;
    const int* zlo;
    int *order, *ntiles;
    int4* tiles;
};

struct GemmArgs {
    int n, parts, nsp, npxt, npxblk, kpad;
    const int* order;
    const int* ntiles;
    const int4* tiles;
    const double *A64, *w64;
    const uint16_t *Ahi, *Alo, *whi, *wlo;
    void* yo;
};

struct PackArgs {
    int nspec, nt, nz, npair, kpair, kpad, npxblk;
    const double* basis;
    double* A64;
    uint16_t *Ahi, *Alo;
};

struct ChainArgs {
    int n, nspec, nsp, parts, nbands, jtop, add_igm, smooth_fast, band_j0, band_j1, ncosmo;
    double amzl;
    const double* prm;
    const void* yo;
    const double *lam, *loglam, *ltab, *wl, *dlam, *igm_a, *cz, *cdl;
    const int *bptr, *bjlo, *bnj, *mask;
    const double* bc64;
    const float* bc32;
    void *spec, *young, *old;
    double *mags, *fband;
};

template <typename T>
__device__ __forceinline__ T coef_at(const ChainArgs& a, int i);
template <>
__device__ __forceinline__ double coef_at<double>(const ChainArgs& a, int i) { return a.bc64[i]; }
template <>
__device__ __forceinline__ float coef_at<float>(const ChainArgs& a, int i) { return a.bc32[i]; }

__global__ void k_prep(PrepArgs a);
__global__ void k_group(GroupArgs a);
__global__ void k_gemm_f64(GemmArgs a);
__global__ void k_gemm_f64_vec(GemmArgs a);
__global__ void k_gemm_bf16x3(GemmArgs a);
__global__ void k_pack_f64(PackArgs a);
__global__ void k_pack_bf16(PackArgs a);
template <typename T>
__global__ void k_chain(ChainArgs a);
template <typename T>
__global__ void k_band_fluxes(int n, int nspec, int nbands, const T* bt, const int* jlo, const int* jhi, const int* do_band,
                              const T* wt, const T* fobs_t, double* fband);
template <typename T>
__global__ void k_observed_frame(int n, int nspec, const double* zc, const double* lam, const T* total, T* fobs_t);
__global__ void k_selftest_mfma(int which, const void* a, const void* b, void* d);
