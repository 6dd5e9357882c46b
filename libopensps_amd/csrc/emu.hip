// The learned emulator (include/opensps_emu.h): an MLP on the BF16 matrix cores of gfx950.
//   k_emu_gemm     C = A B^T (both operands K-contiguous bf16 rows; FP32 accumulation) with the layer epilogues:
//                  bias, GELU, the GELU derivative of the backward pass, FP32 / bf16 / transposed-bf16 outputs,
//                  the output scaling of inference. Workgroup tile 128 x 128, four wavefronts of 64 x 64
//                  (4 x 4 blocks of the 16x16x32 instruction), K tiles of 32 through double-buffered LDS.
//   k_emu_input    raw inputs -> scaled bf16 rows (and their transpose for the weight gradients)
//   k_emu_loss     the weighted squared error against the scaled targets, its gradient rows and transpose
//   k_emu_rowsum   bias gradients (row sums of the transposed pre-activation gradients)
//   k_emu_adam     Adam on the FP32 master weights, then their bf16 copies and transposes
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdlib.h>

#include <cmath>
#include <random>
#include <vector>

#include "../../include/opensps_emu.h"
#include "ops_common.h"

struct OpsCtx;
int ops_fail(OpsCtx* c, int code, const char* fmt, ...);

#define EMU_CHK(call)                                                                                 \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess) return ops_fail(nullptr, 10, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

#define EM_T 128      // workgroup tile rows / columns
#define EM_K 32       // K per LDS tile
#define EM_LDR 40     // bf16 per LDS row (32 + 8: rows 80 bytes apart)

struct EmuGemmArgs {
    int M, N, K;
    const uint16_t* A;
    int lda;
    const uint16_t* B;
    int ldb;
    const float* bias;      // N or NULL
    int act;                // 1: GELU after the bias
    const float* dpre;      // the backward pass: C *= GELU'(dpre) (M x N, row stride lddp) or NULL
    int lddp;
    float* Cf;              // FP32 out (before the activation) or NULL
    int ldcf;
    uint16_t* Cb;           // bf16 out (after the activation) or NULL
    int ldcb;
    uint16_t* Ct;           // bf16 transposed out (N x M) or NULL
    int ldct;
    const float *omu, *osd; // inference's last layer: Cf = C osd + omu, rows < mvalid, columns < nvalid
    int mvalid, nvalid;
    int ksplit;             // > 1: blockIdx.z takes K / ksplit of the sum and adds it to Cf atomically (no epilogue)
};

// GELU = x Phi(x), Phi(x) = (1 + erf(x / sqrt 2)) / 2 with erf by Abramowitz & Stegun 7.1.26 (|error| <= 1.5e-7,
// far below the bf16 activations' resolution): branch-free, one reciprocal and one exponential. Training and
// inference share it; its derivative Phi(x) + x phi(x).
__device__ __forceinline__ float emu_phi(float x) {
    const float t = __builtin_amdgcn_rcpf(fmaf(fabsf(x), 0.3275911f * 0.70710678118654752f, 1.0f));
    float p = fmaf(t, 0.5f * 1.061405429f, 0.5f * -1.453152027f);   // (a1 t + ... + a5 t^5) / 2
    p = fmaf(t, p, 0.5f * 1.421413741f);
    p = fmaf(t, p, 0.5f * -0.284496736f);
    p = fmaf(t, p, 0.5f * 0.254829592f);
    const float q = p * t * __expf(-0.5f * x * x);
    return x >= 0.0f ? 1.0f - q : q;
}
__device__ __forceinline__ float gelu(float x) { return x * emu_phi(x); }
__device__ __forceinline__ float gelu_grad(float x) { return emu_phi(x) + x * 0.39894228040143268f * __expf(-0.5f * x * x); }

__global__ __launch_bounds__(256) void k_emu_gemm(EmuGemmArgs a) {
    __shared__ __align__(16) uint16_t As[2][EM_T * EM_LDR];
    __shared__ __align__(16) uint16_t Bs[2][EM_T * EM_LDR];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, wr = wv >> 1, wc = wv & 1;
    const int row0 = blockIdx.y * EM_T, col0 = blockIdx.x * EM_T;
    const int kspan = a.ksplit > 1 ? a.K / a.ksplit : a.K, kbase = a.ksplit > 1 ? blockIdx.z * kspan : 0;
    // the loader: 2 x 16 bytes of A and of B a thread per K tile (rows tid / 4 and + 64, k chunk of 8)
    const int lr = tid >> 2, lk = (tid & 3) * 8;
    const uint16_t* pa0 = a.A + (size_t)(row0 + lr) * a.lda + kbase + lk;
    const uint16_t* pa1 = pa0 + (size_t)64 * a.lda;
    const bool vb0 = col0 + lr < a.N, vb1 = col0 + lr + 64 < a.N;
    const uint16_t* pb0 = a.B + (size_t)(vb0 ? col0 + lr : 0) * a.ldb + kbase + lk;
    const uint16_t* pb1 = a.B + (size_t)(vb1 ? col0 + lr + 64 : 0) * a.ldb + kbase + lk;
    uint4 ra0, ra1, rb0, rb1;
    auto fetch = [&](int k0) {
        ra0 = *(const uint4*)(pa0 + k0);
        ra1 = *(const uint4*)(pa1 + k0);
        rb0 = vb0 ? *(const uint4*)(pb0 + k0) : make_uint4(0, 0, 0, 0);
        rb1 = vb1 ? *(const uint4*)(pb1 + k0) : make_uint4(0, 0, 0, 0);
    };
    auto stash = [&](int buf) {
        *(uint4*)(&As[buf][lr * EM_LDR + lk]) = ra0;
        *(uint4*)(&As[buf][(lr + 64) * EM_LDR + lk]) = ra1;
        *(uint4*)(&Bs[buf][lr * EM_LDR + lk]) = rb0;
        *(uint4*)(&Bs[buf][(lr + 64) * EM_LDR + lk]) = rb1;
    };
    f32x4_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = (f32x4_t){0.f, 0.f, 0.f, 0.f};
    // column blocks of this wavefront inside N (N is a multiple of 64: a whole wavefront is in or out)
    const bool wave_on = col0 + wc * 64 < a.N;
    fetch(0);
    stash(0);
    __syncthreads();
    const int nk = kspan / EM_K;
    for (int kt = 0; kt < nk; kt++) {
        const int buf = kt & 1;
        if (kt + 1 < nk) fetch((kt + 1) * EM_K);
        if (wave_on) {
            bf16x8_t av[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; i++)
                av[i] = *(const bf16x8_t*)(&As[buf][(wr * 64 + i * 16 + (lane & 15)) * EM_LDR + 8 * (lane >> 4)]);
#pragma unroll
            for (int j = 0; j < 4; j++)
                bv[j] = *(const bf16x8_t*)(&Bs[buf][(wc * 64 + j * 16 + (lane & 15)) * EM_LDR + 8 * (lane >> 4)]);
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], bv[j], acc[i][j], 0, 0, 0);
        }
        if (kt + 1 < nk) stash(buf ^ 1);
        __syncthreads();
    }
    if (!wave_on) return;
    if (a.ksplit > 1) {   // a partial sum of K: added to the FP32 output
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int c = col0 + wc * 64 + j * 16 + (lane & 15);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int r0 = row0 + wr * 64 + i * 16 + 4 * (lane >> 4);
#pragma unroll
                for (int e = 0; e < 4; e++) atomicAdd(a.Cf + (size_t)(r0 + e) * a.ldcf + c, acc[i][j][e]);
            }
        }
        return;
    }
    // epilogue: lane l holds rows 4 (l / 16) + e of each block, column l % 16
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int c = col0 + wc * 64 + j * 16 + (lane & 15);
        const float bias = a.bias ? a.bias[c] : 0.0f;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int r0 = row0 + wr * 64 + i * 16 + 4 * (lane >> 4);
            uint16_t tb[4];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int r = r0 + e;
                float v = acc[i][j][e] + bias;
                if (a.dpre) v *= gelu_grad(a.dpre[(size_t)r * a.lddp + c]);
                if (a.Cf) {
                    if (a.omu) {
                        if (r < a.mvalid && c < a.nvalid) a.Cf[(size_t)r * a.ldcf + c] = v * a.osd[c] + a.omu[c];
                    } else {
                        a.Cf[(size_t)r * a.ldcf + c] = v;
                    }
                }
                const float post = a.act ? gelu(v) : v;
                tb[e] = bf16_rn(post);
                if (a.Cb) a.Cb[(size_t)r * a.ldcb + c] = tb[e];
            }
            if (a.Ct) {
                uint2 pk;
                pk.x = (uint32_t)tb[0] | ((uint32_t)tb[1] << 16);
                pk.y = (uint32_t)tb[2] | ((uint32_t)tb[3] << 16);
                *(uint2*)(a.Ct + (size_t)c * a.ldct + r0) = pk;
            }
        }
    }
}

// ------------------------------------------------------------------------------------------------
// k_emu_fused: inference through every layer in one kernel. A workgroup of W wavefronts (16 for hidden 256 and 512,
// else 8) takes 128 samples; their activations (128 x hidden bf16) stay in LDS from layer to layer. Wavefront w owns
// the columns [w hidden / W, (w + 1) hidden / W) of a hidden layer's output ([w 256 / W, (w + 1) 256 / W) in the last
// layer of 256) for all 128 rows: 8 row blocks x (columns / 16) blocks of the 16x16x32 instruction, the A operand from
// LDS, the weights (B) from L2 two K steps ahead (global-space loads, so that waits for the A fragments do not wait
// for them), each weight fragment serving 8 row blocks; the accumulators start from the bias. A layer's first two
// weight steps are in flight during the previous layer's epilogue (GELU, bf16 back in place; the last layer: FP32
// rows of 256 in place). hidden <= 512.
// ------------------------------------------------------------------------------------------------
#define EF_ROWS 128

// The layers' weights and biases come from a device table of pointers, which the compiler takes for flat
// (generic) addresses: flat loads also count against the LDS wait counter, so every wait for an A fragment would
// wait for the weights in flight too. Global-space views of them keep the two counters apart.
typedef const __attribute__((address_space(1))) uint16_t* gw_ptr;
typedef const __attribute__((address_space(1))) float* gf_ptr;

template <int NC>
struct EfPre {   // the weight fragments of two K steps
    bf16x8_t w[2][NC];
};

template <int NC, int K>
__device__ __forceinline__ void ef_prefetch(EfPre<4>& p, gw_ptr W, int col0) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
        for (int j = 0; j < NC; j++)
            p.w[s][j] = *(const __attribute__((address_space(1))) bf16x8_t*)(
                W + (unsigned)((col0 + j * 16 + (lane & 15)) * K + 8 * (lane >> 4) + 32 * s));
}

// the products of a layer: acc[i][j] = rows 16 i.., columns col0 + 16 j.. (NC blocks), NKS K steps, from the bias
template <int LDR, int NC, int NKS>
__device__ __forceinline__ void ef_mm(const uint16_t* act, gw_ptr W, gf_ptr bias, int col0, EfPre<4>& p,
                                      f32x4_t (&acc)[8][4]) {
    constexpr int K = 32 * NKS;
    const int lane = threadIdx.x & 63;
    unsigned wo[NC];
#pragma unroll
    for (int j = 0; j < NC; j++) wo[j] = (unsigned)((col0 + j * 16 + (lane & 15)) * K + 8 * (lane >> 4));
#pragma unroll
    for (int j = 0; j < NC; j++) {
        const float bc = bias[col0 + j * 16 + (lane & 15)];
#pragma unroll
        for (int i = 0; i < 8; i++) acc[i][j] = (f32x4_t){bc, bc, bc, bc};
    }
    const uint16_t* ap = act + (lane & 15) * LDR + 8 * (lane >> 4);
    static_assert(NKS % 2 == 0, "K steps in pairs");
#pragma unroll 1
    for (int k2 = 0; k2 < NKS; k2 += 2) {
#pragma unroll
        for (int s = 0; s < 2; s++) {   // step ks = k2 + s uses the fragments of slot s, then refills it for ks + 2
            const int ks = k2 + s;
            bf16x8_t b[NC];
#pragma unroll
            for (int j = 0; j < NC; j++) b[j] = p.w[s][j];
            if (ks + 2 < NKS) {
#pragma unroll
                for (int j = 0; j < NC; j++)
                    p.w[s][j] = *(const __attribute__((address_space(1))) bf16x8_t*)(W + (wo[j] + (ks + 2) * 32));
            }
#pragma unroll
            for (int h = 0; h < 2; h++) {   // the row blocks four at a time (their A fragments)
                bf16x8_t av[4];
#pragma unroll
                for (int i = 0; i < 4; i++) av[i] = *(const bf16x8_t*)(ap + (h * 4 + i) * 16 * LDR + ks * 32);
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < NC; j++)
                        acc[h * 4 + i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], b[j], acc[h * 4 + i][j], 0, 0, 0);
            }
        }
    }
}

// a hidden layer's epilogue: GELU, bf16 into the activations (lane l holds rows 4 (l / 16) + e of each block at
// column l % 16)
template <int LDR, int NC>
__device__ __forceinline__ void ef_epi_hidden(uint16_t* act, int col0, f32x4_t (&acc)[8][4]) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int j = 0; j < NC; j++) {
        uint16_t* o = act + 4 * (lane >> 4) * LDR + col0 + j * 16 + (lane & 15);
#pragma unroll
        for (int i = 0; i < 8; i++)
#pragma unroll
            for (int e = 0; e < 4; e++) o[(i * 16 + e) * LDR] = bf16_rn(gelu(acc[i][j][e]));
    }
}

// the last layer's epilogue: FP32 rows of 256 in place
template <int NC>
__device__ __forceinline__ void ef_epi_last(uint16_t* act, int col0, f32x4_t (&acc)[8][4]) {
    const int lane = threadIdx.x & 63;
    float* fo = (float*)act;
#pragma unroll
    for (int j = 0; j < NC; j++) {
        const int c = col0 + j * 16 + (lane & 15);
#pragma unroll
        for (int i = 0; i < 8; i++)
#pragma unroll
            for (int e = 0; e < 4; e++) fo[(i * 16 + 4 * (lane >> 4) + e) * 256 + c] = acc[i][j][e];
    }
}

struct EmuFusedArgs {
    int n, n_in, kin, hidden, layers, noutp, n_out, ldrow;
    const float *x, *imu, *isd, *omu, *osd;
    const uint16_t* const* W;   // device arrays of the layers' weight and bias pointers
    const float* const* b;
    float* y;
};

#define EF_KIN 64   // the first layer's K (the emulator's padded inputs)
#ifndef EF_W16
#define EF_W16 16
#endif

template <int HID, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_emu_fused(EmuFusedArgs a) {
    constexpr int LDR = HID + 8;
    constexpr int NCH = HID / (16 * WAVES);   // column blocks of a wavefront in a hidden layer
    constexpr int NCL = 256 / (16 * WAVES);   // and in the last layer
    constexpr int NKH = HID / 32;    // K steps of a hidden layer
    constexpr int NK0 = EF_KIN / 32;
    extern __shared__ __align__(16) uint16_t act[];
    const int row0 = blockIdx.x * EF_ROWS, wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int e = threadIdx.x; e < EF_ROWS * EF_KIN; e += blockDim.x) {
        const int r = e / EF_KIN, k = e % EF_KIN;
        float v = 0.0f;
        if (row0 + r < a.n && k < a.n_in) v = (a.x[(size_t)(row0 + r) * a.n_in + k] - a.imu[k]) / a.isd[k];
        act[r * LDR + k] = bf16_rn(v);
    }
    const int ch = wv * (HID / WAVES), cl = wv * (256 / WAVES);   // the wavefront's first column: hidden, last layer
    const int L = a.layers;
    f32x4_t acc[8][4];
    EfPre<4> ph;   // (the first NCH, or 2 in the last layer, of 4)
    ef_prefetch<NCH, EF_KIN>(ph, (gw_ptr)a.W[0], ch);
    __syncthreads();
    ef_mm<LDR, NCH, NK0>(act, (gw_ptr)a.W[0], (gf_ptr)a.b[0], ch, ph, acc);
    for (int l = 1; l < L; l++) {
        // the layer's first weight steps, then the epilogue of layer l - 1, then the layer's products
        ef_prefetch<NCH, HID>(ph, (gw_ptr)a.W[l], ch);
        __syncthreads();   // every wavefront has read the layer's input
        ef_epi_hidden<LDR, NCH>(act, ch, acc);
        __syncthreads();   // the layer's output is in place
        ef_mm<LDR, NCH, NKH>(act, (gw_ptr)a.W[l], (gf_ptr)a.b[l], ch, ph, acc);
    }
    // the last layer
    ef_prefetch<NCL, HID>(ph, (gw_ptr)a.W[L], cl);
    __syncthreads();
    ef_epi_hidden<LDR, NCH>(act, ch, acc);
    __syncthreads();
    ef_mm<LDR, NCL, NKH>(act, (gw_ptr)a.W[L], (gf_ptr)a.b[L], cl, ph, acc);
    __syncthreads();
    ef_epi_last<NCL>(act, cl, acc);
    __syncthreads();
    // the outputs, scaled, from the FP32 rows in LDS: a wavefront per row, lanes over the columns
    const float* fo = (const float*)act;
    for (int r = wv; r < EF_ROWS; r += WAVES) {
        if (row0 + r >= a.n) break;
        for (int c = lane; c < a.n_out; c += 64) a.y[(size_t)(row0 + r) * a.n_out + c] = fo[r * 256 + c] * a.osd[c] + a.omu[c];
    }
}

// raw inputs (n x n_in, FP32) -> scaled bf16 rows X (Bp x kin) and, for training, XT (kin x Bp); zero padding
__global__ void k_emu_input(int n, int n_in, int kin, int Bp, const float* x, const float* mu, const float* sd,
                            uint16_t* X, uint16_t* XT) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)Bp * kin) return;
    const int r = (int)(e / kin), k = (int)(e % kin);
    float v = 0.0f;
    if (r < n && k < n_in) v = (x[(size_t)r * n_in + k] - mu[k]) / sd[k];
    const uint16_t b = bf16_rn(v);
    X[e] = b;
    if (XT) XT[(size_t)k * Bp + r] = b;
}

// the loss against the scaled targets: dY = 2 w (y - t') / n (bf16 rows and transpose); sum of w (y - t')^2 / n
__global__ void k_emu_loss(int n, int n_out, int nout_p, int Bp, const float* y, const float* t, const float* mu,
                           const float* sd, const float* w, uint16_t* dY, uint16_t* dYT, float* loss) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (e < (size_t)Bp * nout_p) {
        const int r = (int)(e / nout_p), o = (int)(e % nout_p);
        float g = 0.0f;
        if (r < n && o < n_out) {
            const float d = y[e] - (t[(size_t)r * n_out + o] - mu[o]) / sd[o];
            g = 2.0f * w[o] * d / (float)n;
            l = w[o] * d * d / (float)n;
        }
        const uint16_t b = bf16_rn(g);
        dY[e] = b;
        dYT[(size_t)o * Bp + r] = b;
    }
    l = wave_sum(l);
    if ((threadIdx.x & 63) == 0 && l != 0.0f) atomicAdd(loss, l);
}

// bias gradients: the row sums of dPreT (rows x Bp), FP32 over the bf16 values
__global__ void k_emu_rowsum(int rows, int Bp, const uint16_t* dT, float* gb) {
    const int r = blockIdx.x;
    if (r >= rows) return;
    float s = 0.0f;
    for (int k = threadIdx.x; k < Bp; k += blockDim.x) s += bf16_to_f32(dT[(size_t)r * Bp + k]);
    s = wave_sum(s);
    __shared__ float red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) gb[r] = red[0] + red[1] + red[2] + red[3];
}

// Adam on one layer's FP32 master weights (N x K, padded: only [nv, kv] live) and biases, then the bf16 copies
struct EmuAdamArgs {
    int N, K, nv, kv;
    float *W, *b, *mW, *vW, *mb, *vb;
    const float *gW, *gb;
    uint16_t *Wb, *WTb;
    float lr, beta1, beta2, eps, c1, c2;   // c1 = 1 - beta1^t, c2 = 1 - beta2^t
};

__global__ void k_emu_adam(EmuAdamArgs a) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nw = (size_t)a.N * a.K;
    if (e < nw) {
        const int r = (int)(e / a.K), k = (int)(e % a.K);
        float w = a.W[e];
        if (r < a.nv && k < a.kv) {
            const float g = a.gW[e];
            const float m = a.beta1 * a.mW[e] + (1.0f - a.beta1) * g;
            const float v = a.beta2 * a.vW[e] + (1.0f - a.beta2) * g * g;
            a.mW[e] = m;
            a.vW[e] = v;
            w -= a.lr * (m / a.c1) / (sqrtf(v / a.c2) + a.eps);
            a.W[e] = w;
        }
        const uint16_t b = bf16_rn(w);
        a.Wb[e] = b;
        a.WTb[(size_t)k * a.N + r] = b;
    } else if (e < nw + a.N) {
        const int r = (int)(e - nw);
        if (r < a.nv) {
            const float g = a.gb[r];
            const float m = a.beta1 * a.mb[r] + (1.0f - a.beta1) * g;
            const float v = a.beta2 * a.vb[r] + (1.0f - a.beta2) * g * g;
            a.mb[r] = m;
            a.vb[r] = v;
            a.b[r] -= a.lr * (m / a.c1) / (sqrtf(v / a.c2) + a.eps);
        }
    }
}

// FP32 master weights -> bf16 copies (after set_params)
__global__ void k_emu_cast(int N, int K, const float* W, uint16_t* Wb, uint16_t* WTb) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)N * K) return;
    const int r = (int)(e / K), k = (int)(e % K);
    const uint16_t b = bf16_rn(W[e]);
    Wb[e] = b;
    WTb[(size_t)k * N + r] = b;
}

// ------------------------------------------------------------------------------------------------
// Host side
// ------------------------------------------------------------------------------------------------
struct EmuLayer {
    int N, K, nv, kv;                  // padded and live sizes
    float *W = nullptr, *b = nullptr, *mW = nullptr, *vW = nullptr, *mb = nullptr, *vb = nullptr;
    float *gW = nullptr, *gb = nullptr;
    uint16_t *Wb = nullptr, *WTb = nullptr;
};

struct OpsEmu {
    int device = 0, n_in = 0, n_out = 0, hidden = 0, layers = 0, kin = 64, noutp = 256;
    hipStream_t stream = nullptr;       // the work's stream: the emulator's own, or the caller's (set_stream)
    hipStream_t own = nullptr;
    std::vector<EmuLayer> L;            // layers + 1
    float *imu = nullptr, *isd = nullptr, *omu = nullptr, *osd = nullptr, *lw = nullptr, *dloss = nullptr;
    float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    long step = 0;
    bool per_layer = false;             // inference layer by layer (OPENSPS_EMU_PER_LAYER=1) instead of fused
    void** d_ptrs = nullptr;            // device: the layers' bf16 weight pointers [17], then their bias pointers
    // activations (grown on demand): inference ping-pong; training per-layer
    int cap_inf = 0, cap_tr = 0;
    uint16_t *X0 = nullptr, *X1 = nullptr;
    std::vector<uint16_t*> Xt, XTt;     // training: layer inputs and their transposes (layers + 1)
    std::vector<float*> pre;            // training: pre-activations (layers + 1; the last is the output)
    uint16_t *dA = nullptr, *dAT = nullptr, *dB = nullptr, *dBT = nullptr;
};

static void emu_free(void* p) {
    if (p) (void)hipFree(p);
}

static void emu_free_train(OpsEmu* e) {
    for (auto p : e->Xt) emu_free(p);
    for (auto p : e->XTt) emu_free(p);
    for (auto p : e->pre) emu_free(p);
    e->Xt.clear();
    e->XTt.clear();
    e->pre.clear();
    emu_free(e->dA);
    emu_free(e->dAT);
    emu_free(e->dB);
    emu_free(e->dBT);
    e->dA = e->dAT = e->dB = e->dBT = nullptr;
    e->cap_tr = 0;
}

static int emu_gemm(OpsEmu* e, const EmuGemmArgs& g, hipStream_t s) {
    dim3 grid((g.N + EM_T - 1) / EM_T, g.M / EM_T, g.ksplit > 1 ? g.ksplit : 1);
    k_emu_gemm<<<grid, 256, 0, s>>>(g);
    EMU_CHK(hipGetLastError());
    return 0;
}

static EmuGemmArgs gemm_args(int M, int N, int K, const uint16_t* A, int lda, const uint16_t* B, int ldb) {
    EmuGemmArgs g = {};
    g.M = M;
    g.N = N;
    g.K = K;
    g.A = A;
    g.lda = lda;
    g.B = B;
    g.ldb = ldb;
    return g;
}

extern "C" {

int opensps_emu_create(int device, int n_in, int n_out, int hidden, int layers, unsigned seed, opensps_emu_t** out) {
    if (!out) return ops_fail(nullptr, 3, "emu_create: NULL output");
    *out = nullptr;
    if (n_in < 1 || n_in > 64 || n_out < 1 || n_out > 256 || hidden < 128 || hidden > 1024 || hidden % 128 ||
        layers < 1 || layers > 16)
        return ops_fail(nullptr, 3, "emu_create: n_in 1..64, n_out 1..256, hidden a multiple of 128 up to 1024");
    EMU_CHK(hipSetDevice(device));
    OpsEmu* e = new OpsEmu();
    e->device = device;
    e->n_in = n_in;
    e->n_out = n_out;
    e->hidden = hidden;
    e->layers = layers;
    if (hipStreamCreateWithFlags(&e->own, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        return ops_fail(nullptr, 10, "emu_create: stream");
    }
    e->stream = e->own;
    const char* pl = getenv("OPENSPS_EMU_PER_LAYER");
    e->per_layer = pl && pl[0] == '1';
    std::mt19937 rng(seed);
    for (int l = 0; l <= layers; l++) {
        EmuLayer Ly;
        Ly.K = l == 0 ? e->kin : hidden;
        Ly.N = l == layers ? e->noutp : hidden;
        Ly.kv = l == 0 ? n_in : hidden;
        Ly.nv = l == layers ? n_out : hidden;
        const size_t nw = (size_t)Ly.N * Ly.K;
        float** f32[] = {&Ly.W, &Ly.mW, &Ly.vW, &Ly.gW};
        for (float** p : f32) {
            EMU_CHK(hipMalloc((void**)p, nw * sizeof(float)));
            EMU_CHK(hipMemset(*p, 0, nw * sizeof(float)));
        }
        float** f32b[] = {&Ly.b, &Ly.mb, &Ly.vb, &Ly.gb};
        for (float** p : f32b) {
            EMU_CHK(hipMalloc((void**)p, Ly.N * sizeof(float)));
            EMU_CHK(hipMemset(*p, 0, Ly.N * sizeof(float)));
        }
        EMU_CHK(hipMalloc((void**)&Ly.Wb, nw * sizeof(uint16_t)));
        EMU_CHK(hipMalloc((void**)&Ly.WTb, nw * sizeof(uint16_t)));
        // He-uniform over the live block, zeros in the padding
        std::vector<float> w(nw, 0.0f);
        const float lim = std::sqrt(6.0f / (float)Ly.kv) * (l == layers ? 0.5f : 1.0f);
        std::uniform_real_distribution<float> U(-lim, lim);
        for (int r = 0; r < Ly.nv; r++)
            for (int k = 0; k < Ly.kv; k++) w[(size_t)r * Ly.K + k] = U(rng);
        EMU_CHK(hipMemcpy(Ly.W, w.data(), nw * sizeof(float), hipMemcpyHostToDevice));
        k_emu_cast<<<(unsigned)((nw + 255) / 256), 256, 0, e->stream>>>(Ly.N, Ly.K, Ly.W, Ly.Wb, Ly.WTb);
        EMU_CHK(hipGetLastError());
        e->L.push_back(Ly);
    }
    {
        std::vector<void*> ptrs(34, nullptr);
        for (int l = 0; l <= layers; l++) {
            ptrs[l] = e->L[l].Wb;
            ptrs[17 + l] = e->L[l].b;
        }
        EMU_CHK(hipMalloc((void**)&e->d_ptrs, 34 * sizeof(void*)));
        EMU_CHK(hipMemcpy(e->d_ptrs, ptrs.data(), 34 * sizeof(void*), hipMemcpyHostToDevice));
    }
    std::vector<float> zeros(256, 0.0f), ones(256, 1.0f);
    float** vecs[] = {&e->imu, &e->isd, &e->omu, &e->osd, &e->lw};
    for (float** p : vecs) EMU_CHK(hipMalloc((void**)p, 256 * sizeof(float)));
    EMU_CHK(hipMemcpy(e->imu, zeros.data(), 256 * sizeof(float), hipMemcpyHostToDevice));
    EMU_CHK(hipMemcpy(e->isd, ones.data(), 256 * sizeof(float), hipMemcpyHostToDevice));
    EMU_CHK(hipMemcpy(e->omu, zeros.data(), 256 * sizeof(float), hipMemcpyHostToDevice));
    EMU_CHK(hipMemcpy(e->osd, ones.data(), 256 * sizeof(float), hipMemcpyHostToDevice));
    EMU_CHK(hipMemcpy(e->lw, ones.data(), 256 * sizeof(float), hipMemcpyHostToDevice));
    EMU_CHK(hipMalloc((void**)&e->dloss, sizeof(float)));
    EMU_CHK(hipStreamSynchronize(e->stream));
    *out = e;
    return 0;
}

int opensps_emu_destroy(opensps_emu_t* e) {
    if (!e) return 1;
    (void)hipStreamSynchronize(e->stream);
    for (auto& Ly : e->L) {
        void* ps[] = {Ly.W, Ly.b, Ly.mW, Ly.vW, Ly.mb, Ly.vb, Ly.gW, Ly.gb, Ly.Wb, Ly.WTb};
        for (void* p : ps) emu_free(p);
    }
    void* ps[] = {e->imu, e->isd, e->omu, e->osd, e->lw, e->dloss, e->X0, e->X1, e->d_ptrs};
    for (void* p : ps) emu_free(p);
    emu_free_train(e);
    (void)hipStreamDestroy(e->own);
    delete e;
    return 0;
}

int opensps_emu_set_stream(opensps_emu_t* e, void* stream) {
    if (!e) return 1;
    EMU_CHK(hipStreamSynchronize(e->stream));
    e->stream = (hipStream_t)stream;   // NULL: the null (default) stream
    return 0;
}

int opensps_emu_nparams(opensps_emu_t* e, long* n) {
    if (!e || !n) return 1;
    long s = 0;
    for (auto& Ly : e->L) s += (long)Ly.nv * Ly.kv + Ly.nv;
    *n = s;
    return 0;
}

int opensps_emu_set_params(opensps_emu_t* e, const float* p) {
    if (!e) return 1;
    if (!p) return ops_fail(nullptr, 3, "emu_set_params: NULL");
    EMU_CHK(hipSetDevice(e->device));
    EMU_CHK(hipStreamSynchronize(e->stream));
    size_t at = 0;
    for (auto& Ly : e->L) {
        const size_t nw = (size_t)Ly.N * Ly.K;
        std::vector<float> w(nw, 0.0f), b(Ly.N, 0.0f);
        for (int r = 0; r < Ly.nv; r++)
            for (int k = 0; k < Ly.kv; k++) w[(size_t)r * Ly.K + k] = p[at++];
        for (int r = 0; r < Ly.nv; r++) b[r] = p[at++];
        EMU_CHK(hipMemcpy(Ly.W, w.data(), nw * sizeof(float), hipMemcpyHostToDevice));
        EMU_CHK(hipMemcpy(Ly.b, b.data(), Ly.N * sizeof(float), hipMemcpyHostToDevice));
        float* z[] = {Ly.mW, Ly.vW};
        for (float* q : z) EMU_CHK(hipMemset(q, 0, nw * sizeof(float)));
        EMU_CHK(hipMemset(Ly.mb, 0, Ly.N * sizeof(float)));
        EMU_CHK(hipMemset(Ly.vb, 0, Ly.N * sizeof(float)));
        k_emu_cast<<<(unsigned)((nw + 255) / 256), 256, 0, e->stream>>>(Ly.N, Ly.K, Ly.W, Ly.Wb, Ly.WTb);
        EMU_CHK(hipGetLastError());
    }
    e->step = 0;
    EMU_CHK(hipStreamSynchronize(e->stream));
    return 0;
}

static int emu_download(OpsEmu* e, float* p, bool grads) {
    size_t at = 0;
    for (auto& Ly : e->L) {
        const size_t nw = (size_t)Ly.N * Ly.K;
        std::vector<float> w(nw), b(Ly.N);
        EMU_CHK(hipMemcpy(w.data(), grads ? Ly.gW : Ly.W, nw * sizeof(float), hipMemcpyDeviceToHost));
        EMU_CHK(hipMemcpy(b.data(), grads ? Ly.gb : Ly.b, Ly.N * sizeof(float), hipMemcpyDeviceToHost));
        for (int r = 0; r < Ly.nv; r++)
            for (int k = 0; k < Ly.kv; k++) p[at++] = w[(size_t)r * Ly.K + k];
        for (int r = 0; r < Ly.nv; r++) p[at++] = b[r];
    }
    return 0;
}

int opensps_emu_get_params(opensps_emu_t* e, float* p) {
    if (!e) return 1;
    if (!p) return ops_fail(nullptr, 3, "emu_get_params: NULL");
    EMU_CHK(hipSetDevice(e->device));
    EMU_CHK(hipStreamSynchronize(e->stream));
    return emu_download(e, p, false);
}

int opensps_emu_set_scaling(opensps_emu_t* e, const float* in_mu, const float* in_sd, const float* out_mu,
                            const float* out_sd) {
    if (!e) return 1;
    EMU_CHK(hipSetDevice(e->device));
    EMU_CHK(hipStreamSynchronize(e->stream));
    if (in_mu) EMU_CHK(hipMemcpy(e->imu, in_mu, e->n_in * sizeof(float), hipMemcpyHostToDevice));
    if (in_sd) EMU_CHK(hipMemcpy(e->isd, in_sd, e->n_in * sizeof(float), hipMemcpyHostToDevice));
    if (out_mu) EMU_CHK(hipMemcpy(e->omu, out_mu, e->n_out * sizeof(float), hipMemcpyHostToDevice));
    if (out_sd) EMU_CHK(hipMemcpy(e->osd, out_sd, e->n_out * sizeof(float), hipMemcpyHostToDevice));
    return 0;
}

int opensps_emu_set_loss_weights(opensps_emu_t* e, const float* w) {
    if (!e) return 1;
    if (!w) return ops_fail(nullptr, 3, "emu_set_loss_weights: NULL");
    EMU_CHK(hipSetDevice(e->device));
    EMU_CHK(hipStreamSynchronize(e->stream));
    EMU_CHK(hipMemcpy(e->lw, w, e->n_out * sizeof(float), hipMemcpyHostToDevice));
    return 0;
}

int opensps_emu_set_adam(opensps_emu_t* e, float beta1, float beta2, float eps) {
    if (!e) return 1;
    e->beta1 = beta1;
    e->beta2 = beta2;
    e->eps = eps;
    return 0;
}

int opensps_emu_sync(opensps_emu_t* e) {
    if (!e) return 1;
    EMU_CHK(hipStreamSynchronize(e->stream));
    return 0;
}

int opensps_emu_predict(opensps_emu_t* e, int n, const float* x, float* y, void* stream) {
    if (!e) return 1;
    if (n < 0 || (n > 0 && (!x || !y))) return ops_fail(nullptr, 3, "emu_predict: bad arguments");
    if (n == 0) return 0;
    EMU_CHK(hipSetDevice(e->device));
    hipStream_t s = stream ? (hipStream_t)stream : e->stream;
    const int Bp = ((n + EM_T - 1) / EM_T) * EM_T;
    if (Bp > e->cap_inf) {
        EMU_CHK(hipStreamSynchronize(s));
        emu_free(e->X0);
        emu_free(e->X1);
        const int w = std::max(e->hidden, e->kin);
        EMU_CHK(hipMalloc((void**)&e->X0, (size_t)Bp * w * sizeof(uint16_t)));
        EMU_CHK(hipMalloc((void**)&e->X1, (size_t)Bp * w * sizeof(uint16_t)));
        e->cap_inf = Bp;
    }
    if (e->hidden <= 512 && e->layers <= 16 && !e->per_layer) {
        // one kernel through every layer, the activations in LDS
        EmuFusedArgs f;
        f.n = n;
        f.n_in = e->n_in;
        f.kin = e->kin;
        f.hidden = e->hidden;
        f.layers = e->layers;
        f.noutp = e->noutp;
        f.n_out = e->n_out;
        f.ldrow = std::max(e->hidden, e->kin) + 8;
        f.x = x;
        f.imu = e->imu;
        f.isd = e->isd;
        f.omu = e->omu;
        f.osd = e->osd;
        f.W = (const uint16_t* const*)e->d_ptrs;
        f.b = (const float* const*)(e->d_ptrs + 17);
        f.y = y;
        const int hid = e->hidden;
        f.ldrow = hid + 8;
        // LDS: the activations, or the last layer's FP32 rows of 256 (whichever is larger)
        const size_t lds = std::max((size_t)EF_ROWS * f.ldrow * sizeof(uint16_t), (size_t)EF_ROWS * 256 * sizeof(float));
        // wavefronts a workgroup: 16 where every wavefront keeps a whole column block in each layer, else 8
        constexpr int W16 = EF_W16;
        const void* kfn = hid == 512 ? (const void*)k_emu_fused<512, W16>
                        : hid == 384 ? (const void*)k_emu_fused<384, 8>
                        : hid == 256 ? (const void*)k_emu_fused<256, W16> : (const void*)k_emu_fused<128, 8>;
        EMU_CHK(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        const unsigned grid = (unsigned)((n + EF_ROWS - 1) / EF_ROWS);
        if (hid == 512) k_emu_fused<512, W16><<<grid, 64 * W16, lds, s>>>(f);
        else if (hid == 384) k_emu_fused<384, 8><<<grid, 512, lds, s>>>(f);
        else if (hid == 256) k_emu_fused<256, W16><<<grid, 64 * W16, lds, s>>>(f);
        else k_emu_fused<128, 8><<<grid, 512, lds, s>>>(f);
        EMU_CHK(hipGetLastError());
        return 0;
    }
    const size_t nin = (size_t)Bp * e->kin;
    k_emu_input<<<(unsigned)((nin + 255) / 256), 256, 0, s>>>(n, e->n_in, e->kin, Bp, x, e->imu, e->isd, e->X0, nullptr);
    EMU_CHK(hipGetLastError());
    uint16_t *in = e->X0, *outb = e->X1;
    int st;
    for (int l = 0; l <= e->layers; l++) {
        const EmuLayer& Ly = e->L[l];
        EmuGemmArgs g = gemm_args(Bp, Ly.N, Ly.K, in, Ly.K, Ly.Wb, Ly.K);
        g.bias = Ly.b;
        if (l < e->layers) {
            g.act = 1;
            g.Cb = outb;
            g.ldcb = Ly.N;
        } else {
            g.Cf = y;
            g.ldcf = e->n_out;
            g.omu = e->omu;
            g.osd = e->osd;
            g.mvalid = n;
            g.nvalid = e->n_out;
        }
        if ((st = emu_gemm(e, g, s))) return st;
        std::swap(in, outb);
    }
    return 0;
}

// forward with the saved activations, the loss and the gradients (into the layers' gW, gb)
static int emu_forward_backward(OpsEmu* e, int n, const float* x, const float* t) {
    const int Bp = ((n + EM_T - 1) / EM_T) * EM_T;
    const int NL = e->layers + 1;
    hipStream_t s = e->stream;
    if (Bp > e->cap_tr) {
        EMU_CHK(hipStreamSynchronize(s));
        emu_free_train(e);
        for (int l = 0; l < NL; l++) {
            const EmuLayer& Ly = e->L[l];
            uint16_t *X = nullptr, *XT = nullptr;
            float* P = nullptr;
            EMU_CHK(hipMalloc((void**)&X, (size_t)Bp * Ly.K * sizeof(uint16_t)));
            EMU_CHK(hipMalloc((void**)&XT, (size_t)Bp * Ly.K * sizeof(uint16_t)));
            EMU_CHK(hipMalloc((void**)&P, (size_t)Bp * Ly.N * sizeof(float)));
            e->Xt.push_back(X);
            e->XTt.push_back(XT);
            e->pre.push_back(P);
        }
        const int w = std::max(e->hidden, e->noutp);
        uint16_t** ds[] = {&e->dA, &e->dAT, &e->dB, &e->dBT};
        for (uint16_t** p : ds) EMU_CHK(hipMalloc((void**)p, (size_t)Bp * w * sizeof(uint16_t)));
        e->cap_tr = Bp;
    }
    const size_t nin = (size_t)Bp * e->kin;
    k_emu_input<<<(unsigned)((nin + 255) / 256), 256, 0, s>>>(n, e->n_in, e->kin, Bp, x, e->imu, e->isd, e->Xt[0],
                                                              e->XTt[0]);
    EMU_CHK(hipGetLastError());
    int st;
    for (int l = 0; l < NL; l++) {
        const EmuLayer& Ly = e->L[l];
        EmuGemmArgs g = gemm_args(Bp, Ly.N, Ly.K, e->Xt[l], Ly.K, Ly.Wb, Ly.K);
        g.bias = Ly.b;
        g.Cf = e->pre[l];
        g.ldcf = Ly.N;
        if (l + 1 < NL) {
            g.act = 1;
            g.Cb = e->Xt[l + 1];
            g.ldcb = Ly.N;
            g.Ct = e->XTt[l + 1];
            g.ldct = Bp;
        }
        if ((st = emu_gemm(e, g, s))) return st;
    }
    EMU_CHK(hipMemsetAsync(e->dloss, 0, sizeof(float), s));
    const size_t no = (size_t)Bp * e->noutp;
    k_emu_loss<<<(unsigned)((no + 255) / 256), 256, 0, s>>>(n, e->n_out, e->noutp, Bp, e->pre[NL - 1], t, e->omu, e->osd,
                                                            e->lw, e->dA, e->dAT, e->dloss);
    EMU_CHK(hipGetLastError());
    // backward: dA / dAT hold the pre-activation gradient of layer l
    uint16_t *dcur = e->dA, *dcurT = e->dAT, *dnext = e->dB, *dnextT = e->dBT;
    for (int l = NL - 1; l >= 0; l--) {
        const EmuLayer& Ly = e->L[l];
        // dW = dPre^T X : (N x Bp) . (K x Bp)^T, the batch split over enough workgroups to fill the chip
        EmuGemmArgs gw = gemm_args(Ly.N, Ly.K, Bp, dcurT, Bp, e->XTt[l], Bp);
        gw.Cf = Ly.gW;
        gw.ldcf = Ly.K;
        const int tiles = (Ly.N / EM_T) * ((Ly.K + EM_T - 1) / EM_T);
        int ks = 1;
        while (ks < 64 && tiles * ks < 512 && (Bp / EM_K) % (2 * ks) == 0) ks *= 2;
        gw.ksplit = ks;
        if (ks > 1) EMU_CHK(hipMemsetAsync(Ly.gW, 0, (size_t)Ly.N * Ly.K * sizeof(float), s));
        if ((st = emu_gemm(e, gw, s))) return st;
        k_emu_rowsum<<<Ly.N, 256, 0, s>>>(Ly.N, Bp, dcurT, Ly.gb);
        EMU_CHK(hipGetLastError());
        if (l == 0) break;
        // dX = dPre W, times GELU'(pre of layer l - 1): (Bp x N) . (K x N)^T
        EmuGemmArgs gx = gemm_args(Bp, Ly.K, Ly.N, dcur, Ly.N, Ly.WTb, Ly.N);
        gx.dpre = e->pre[l - 1];
        gx.lddp = Ly.K;
        gx.Cb = dnext;
        gx.ldcb = Ly.K;
        gx.Ct = dnextT;
        gx.ldct = Bp;
        if ((st = emu_gemm(e, gx, s))) return st;
        std::swap(dcur, dnext);
        std::swap(dcurT, dnextT);
    }
    return 0;
}

int opensps_emu_gradients(opensps_emu_t* e, int n, const float* x, const float* t, float* grads, float* loss) {
    if (!e) return 1;
    if (n < 1 || !x || !t) return ops_fail(nullptr, 3, "emu_gradients: bad arguments");
    EMU_CHK(hipSetDevice(e->device));
    int st = emu_forward_backward(e, n, x, t);
    if (st) return st;
    EMU_CHK(hipStreamSynchronize(e->stream));
    if (loss) EMU_CHK(hipMemcpy(loss, e->dloss, sizeof(float), hipMemcpyDeviceToHost));
    if (grads) return emu_download(e, grads, true);
    return 0;
}

int opensps_emu_train_step(opensps_emu_t* e, int n, const float* x, const float* t, float lr, float* loss) {
    if (!e) return 1;
    if (n < 1 || !x || !t) return ops_fail(nullptr, 3, "emu_train_step: bad arguments");
    EMU_CHK(hipSetDevice(e->device));
    int st = emu_forward_backward(e, n, x, t);
    if (st) return st;
    e->step++;
    for (auto& Ly : e->L) {
        EmuAdamArgs a;
        a.N = Ly.N;
        a.K = Ly.K;
        a.nv = Ly.nv;
        a.kv = Ly.kv;
        a.W = Ly.W;
        a.b = Ly.b;
        a.mW = Ly.mW;
        a.vW = Ly.vW;
        a.mb = Ly.mb;
        a.vb = Ly.vb;
        a.gW = Ly.gW;
        a.gb = Ly.gb;
        a.Wb = Ly.Wb;
        a.WTb = Ly.WTb;
        a.lr = lr;
        a.beta1 = e->beta1;
        a.beta2 = e->beta2;
        a.eps = e->eps;
        a.c1 = 1.0f - std::pow(e->beta1, (float)e->step);
        a.c2 = 1.0f - std::pow(e->beta2, (float)e->step);
        const size_t tot = (size_t)Ly.N * Ly.K + Ly.N;
        k_emu_adam<<<(unsigned)((tot + 255) / 256), 256, 0, e->stream>>>(a);
        EMU_CHK(hipGetLastError());
    }
    if (loss) {
        EMU_CHK(hipStreamSynchronize(e->stream));
        EMU_CHK(hipMemcpy(loss, e->dloss, sizeof(float), hipMemcpyDeviceToHost));
    }
    return 0;
}

}   // extern "C"
