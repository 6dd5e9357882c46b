// Split-BF16 products on the matrix cores with LDS tiles (gfx950): the age sums of the batch (the metallicity
// pair's slab, the full basis of the weights over metallicities, the template library of the SSP operator) and the
// reverse pass's transposed product. C = A B^T in FP32 from x = hi + lo splits of both operands:
// C = A_lo B_hi + A_hi B_lo + A_hi B_hi (the lo x lo term is below FP32's resolution).
//
// Workgroup: 256 threads, a 128 x 128 tile of C; wavefront (wr, wc) owns rows 64 wr .. +64 and columns
// 64 wc .. +64 as 4 x 4 blocks of the 16x16x32 instruction. K advances 32 a step through two LDS buffers
// (one barrier a step): every thread fetches the next step's 16-byte pieces of A and B into registers, the
// wavefronts multiply from the current buffer, then the pieces go to the other buffer.
// LDS per buffer, in the instruction's operand order (lane l: row or column l % 16, K 8 (l / 16) .. +8):
// A as packed in memory (consecutive, conflict-free); B's pieces at slot 16 g + ((c % 16) xor 4 g) of their
// 16-column block, so that both the writes (four K groups of four columns a quarter-wave) and the instruction's
// reads (one K group of 16 columns a quarter-wave) touch 16 distinct 16-byte bank groups.
#include "ops_common.h"
#include "ops_launch.h"

#define MM_PIECES 512   // 16-byte pieces of one operand half (hi or lo) per K step: 8 blocks x 64 lanes

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_mm_bf16x3(MMArgs a) {
    extern __shared__ __align__(16) unsigned char mm_lds[];
    bf16x8_t* sA = (bf16x8_t*)mm_lds;        // [buffer][hi, lo][MM_PIECES]
    bf16x8_t* sB = sA + 4 * MM_PIECES;       // [buffer][hi, lo][MM_PIECES]
    float* sT = (float*)mm_lds;              // the epilogue's C tile [128 columns][OPS_MM_TS], over the operands
    int* sColB = (int*)(mm_lds + OPS_MM_COLS * OPS_MM_TS * sizeof(float));
    int* sColC = sColB + OPS_MM_COLS;

    // blockIdx -> tiles: each XCD (blockIdx % 8) takes one eighth of the row tiles, row tiles fastest, so A's rows
    // of an XCD stay in its L2 while the column tiles pass
    const int bid = blockIdx.x, xcd = bid & 7, r = bid >> 3;
    const int rtx = (a.nrt + 7) >> 3;
    const int rt = (r % rtx) * 8 + xcd, ct = r / rtx;
    const int nct = a.mode == OPS_MM_GROUPED ? *a.ntiles : a.nct;
    if (rt >= a.nrt || ct >= nct) return;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, wr = wv >> 1, wc = wv & 1;

    int slab = 0;
    if (a.mode == OPS_MM_GROUPED) {
        const int4 td = a.tiles[ct];
        slab = td.x;
        if (tid < OPS_MM_COLS) {
            const bool v = tid < td.z;
            const int cc = td.y + (v ? tid : 0);
            const int pos = cc / a.parts, part = cc - pos * a.parts;
            const int model = a.order[pos];
            sColB[tid] = model * a.parts + part;
            sColC[tid] = v ? part * a.n + model : -1;
        }
    } else if (tid < OPS_MM_COLS) {
        const int col = ct * OPS_MM_COLS + tid;
        const bool v = col < a.ncols;
        sColB[tid] = v ? col : 0;
        sColC[tid] = v ? col : -1;
    }
    __syncthreads();

    // this thread's pieces: p = tid + 256 i. A: block p / 64 of the tile (clamped to the last block of A; rows past
    // it are not stored), slot p % 64. B: column p / 4, K group p % 4.
    const int ksteps = a.ksteps;
    const bf16x8_t *gAh[2], *gAl[2], *gBh[2], *gBl[2];
    int dB[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const int p = tid + 256 * i;
        const int blk = min(rt * 8 + (p >> 6), a.mblk - 1);
        const size_t ao = (size_t)slab * (a.slab / 8) + (size_t)blk * ksteps * 64 + (p & 63);
        gAh[i] = (const bf16x8_t*)a.Ah + ao;
        gAl[i] = (const bf16x8_t*)a.Al + ao;
        const int c = p >> 2, g = p & 3;
        const size_t bo = (size_t)sColB[c] * a.ldb + 8 * g;
        gBh[i] = (const bf16x8_t*)(a.Bh + bo);
        gBl[i] = (const bf16x8_t*)(a.Bl + bo);
        dB[i] = (c >> 4) * 64 + 16 * g + ((c & 15) ^ (g << 2));
    }
    bf16x8_t fa[2][2], fb[2][2];   // [piece][hi, lo] in flight
    auto fetch = [&](int ks) {
#pragma unroll
        for (int i = 0; i < 2; i++) {
            fa[i][0] = gAh[i][(size_t)ks * 64];
            fa[i][1] = gAl[i][(size_t)ks * 64];
            fb[i][0] = gBh[i][ks * 4];
            fb[i][1] = gBl[i][ks * 4];
        }
    };
    auto stash = [&](int buf) {
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int p = tid + 256 * i;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                sA[(buf * 2 + h) * MM_PIECES + p] = fa[i][h];
                sB[(buf * 2 + h) * MM_PIECES + dB[i]] = fb[i][h];
            }
        }
    };

    f32x4_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = (f32x4_t){0.f, 0.f, 0.f, 0.f};
    const int aoff = wr * 4 * 64 + lane;
    const int boff = wc * 4 * 64 + 16 * (lane >> 4) + ((lane & 15) ^ ((lane >> 4) << 2));

    fetch(0);
    stash(0);
    __syncthreads();
    for (int ks = 0; ks < ksteps; ks++) {
        const int cur = ks & 1;
        const bool more = ks + 1 < ksteps;
        if (more) fetch(ks + 1);
        bf16x8_t ah[4], al[4], bh[4], bl[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            ah[i] = sA[(cur * 2 + 0) * MM_PIECES + aoff + i * 64];
            al[i] = sA[(cur * 2 + 1) * MM_PIECES + aoff + i * 64];
            bh[i] = sB[(cur * 2 + 0) * MM_PIECES + boff + i * 64];
            bl[i] = sB[(cur * 2 + 1) * MM_PIECES + boff + i * 64];
        }
        // the three products a term at a time over the 16 blocks: dependent instructions 16 apart
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        if (more) stash(cur ^ 1);
        __syncthreads();
    }

    // The tile goes out through LDS (the loop's last barrier has freed the operand buffers): C is column-major, a
    // column one model's part with its rows (pixels) contiguous, so each column's 128 rows leave as one 512-byte run
    // by 32 lanes instead of 64-byte pieces of 16 columns an instruction. 16x16x32 result layout: register r of lane
    // l holds row 4 (l / 16) + r, column l % 16.
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int col = wc * 64 + j * 16 + (lane & 15), row = wr * 64 + i * 16 + 4 * (lane >> 4);
            *(f32x4_t*)(sT + col * OPS_MM_TS + row) = acc[i][j];
        }
    __syncthreads();
    const int rows_valid = min(OPS_MM_ROWS, a.mblk * 16 - rt * OPS_MM_ROWS);   // whole 16-row blocks inside C
#pragma unroll 4
    for (int k = 0; k < 16; k++) {
        const int q = tid + 256 * k, col = q >> 5, part = q & 31;
        const int cc = sColC[col];
        if (cc < 0 || part * 4 >= rows_valid) continue;
        // non-temporal: C is written once and read by a later kernel; keeping it out of the caches leaves them to
        // the operands (age sums 0.223 -> 0.214 ms, and the chain that reads C after it 0.500 -> 0.491 ms)
        __builtin_nontemporal_store(*(const f32x4_t*)(sT + col * OPS_MM_TS + part * 4),
                                    (f32x4_t*)(a.C + (size_t)cc * a.ldc + rt * OPS_MM_ROWS + part * 4));
    }
}
