// nova-mi355x: the dense GEMM with a fused row argmax on the matrix cores (gfx950), shared by
// the dense-belief PBVI scoring and the batched alpha-vector policy evaluation.
//
// For each pair p = (a, o) it forms
//   C_p[i][j] = sum_k A_a[i][k] * B_p[j][k]      (i < rows, j < cols)
// and writes, per row i and tile of 256 columns, the first maximiser of C_p[i][.] over the
// tile's columns j < cols (value and index) to partVal / partJ at
// ((jtile * rows * m + i * m + a) * z + o). In dense PBVI B_p[j][k] = Gamma[j][k] * gamma
// O[a,k,o]; in the policy evaluation there is one pair and B = Gamma. Operands are zero-padded
// to whole 256-row tiles and to kpad columns (a multiple of 64), so the k loop has no bounds
// checks.
//
// dense_score_bf16 (bf16 operands, fp32 accumulation): operands are stored k-blocked,
// [k / 32][row][32], so the 256 x 32 slice of a tile that one k-stage needs is one contiguous
// 16 KB run. A workgroup of 8 waves owns a 256 x 256 tile of C. Each stage of A_a and B_p is
// copied from HBM straight into LDS (global_load_lds_dwordx4, no staging registers), with four
// stages in flight. Each wave accumulates a 128 x 64 sub-tile as 4 x 2 blocks of 32 x 32
// (v_mfma_f32_32x32x16_bf16); the fragments of the next k-half are read from LDS between the
// matrix instructions of the current one, so the matrix cores do not wait at the stage barrier.
// B_p for PBVI is formed per chunk of pairs by dense_scale_pairs, bf16(float(Gamma_bf16) * s),
// before the GEMM.
//
// dense_score<IN> (fp32 operands; also the bf16 reference path): row-major operands, 256 x
// 128-byte slices staged through registers (B scaled by s_p on the way) into one of two LDS
// buffers; v_mfma_f32_32x32x2_f32 on fp32 operands, v_mfma_f32_32x32x16_bf16 on bf16.
//
// Both number their workgroups so that each XCD takes a contiguous run of tiles, ordered in
// groups of four tile rows, and shares their A and B panels in its L2.
#ifndef NOVA_AMD_DENSE_GEMM_H
#define NOVA_AMD_DENSE_GEMM_H

#include <climits>
#include <cstdlib>
#include <cstring>

#include "nova_hip.h"

namespace nova_amd {
namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx16_t __attribute__((ext_vector_type(16)));

__device__ __forceinline__ bool better(float v, int j, float bv, int bj)
{
    // larger value wins; equal values: the smaller alpha index (the reference's first maximiser)
    return v > bv || (v == bv && j < bj);
}

constexpr int kGT = 256;                        // tile rows (of A) and columns (rows of B)
constexpr int kGThreads = 512;                  // 8 waves: 2 row halves x 4 column quarters
constexpr int kGRowBytes = 128;                 // k-slice per LDS row: 64 bf16 or 32 fp32
constexpr int kGStride = kGRowBytes + 16;       // padded LDS row (bytes)
constexpr int kGOperandBytes = kGT * kGStride;  // one operand slice
constexpr size_t kGLdsBytes = 4 * (size_t)kGOperandBytes;   // 2 buffers x 2 operands = 147,456
constexpr unsigned kGKPad = 64;                 // column padding of the dense operands
constexpr int kGConvThreads = 256;              // lanes of the conversion kernel
constexpr unsigned kGGroup = 4;                 // tile rows per group of the tile order
constexpr int kGBlk = 32;                       // k per block of the k-blocked layout
constexpr int kGStages = 4;                     // LDS stages in flight (dense_score_bf16)
constexpr int kGStageOp = kGT * kGBlk * 2;      // one operand of one stage: 16 KB
constexpr size_t kGDlLdsBytes = (size_t)kGStages * 2 * kGStageOp;   // 131,072

template <typename IN> struct DenseIn;
template <> struct DenseIn<__bf16> {
    static __device__ __forceinline__ __bf16 cvt(float x) { return (__bf16)x; }
};
template <> struct DenseIn<float> {
    static __device__ __forceinline__ float cvt(float x) { return x; }
};

// A (float, count / n rows of n) -> AI (IN, rows of kpad; the padding keeps its zeros)
template <typename IN>
__global__ __launch_bounds__(kGConvThreads) void dense_convert(unsigned n, unsigned kpad, size_t count,
                                                     const float *__restrict__ G,
                                                     IN *__restrict__ GI)
{
    for (size_t x = blockIdx.x * (size_t)kGConvThreads + threadIdx.x; x < count;
         x += (size_t)gridDim.x * kGConvThreads)
        GI[(x / n) * kpad + x % n] = DenseIn<IN>::cvt(G[x]);
}

// 16 bytes of IN scaled elementwise by s0 (and s1 for the upper four of eight bf16)
template <typename IN> __device__ __forceinline__ uint4 scale16(uint4 v, float4 s0, float4 s1);
template <> __device__ __forceinline__ uint4 scale16<float>(uint4 v, float4 s0, float4)
{
    uint4 o;
    o.x = __float_as_uint(__uint_as_float(v.x) * s0.x);
    o.y = __float_as_uint(__uint_as_float(v.y) * s0.y);
    o.z = __float_as_uint(__uint_as_float(v.z) * s0.z);
    o.w = __float_as_uint(__uint_as_float(v.w) * s0.w);
    return o;
}
__device__ __forceinline__ unsigned scale_bf16x2(unsigned v, float lo, float hi)
{
    const __bf16 bl = (__bf16)(__uint_as_float(v << 16) * lo);
    const __bf16 bh = (__bf16)(__uint_as_float(v & 0xffff0000u) * hi);
    return (unsigned)__builtin_bit_cast(unsigned short, bl) |
           ((unsigned)__builtin_bit_cast(unsigned short, bh) << 16);
}
template <> __device__ __forceinline__ uint4 scale16<__bf16>(uint4 v, float4 s0, float4 s1)
{
    return make_uint4(scale_bf16x2(v.x, s0.x, s0.y), scale_bf16x2(v.y, s0.z, s0.w),
                      scale_bf16x2(v.z, s1.x, s1.y), scale_bf16x2(v.w, s1.z, s1.w));
}

template <typename IN>
__global__ __launch_bounds__(kGThreads) void dense_score(unsigned rows, unsigned cols,
        unsigned kpad, unsigned m, unsigned z, unsigned rowTiles, unsigned colTiles,
        const IN *__restrict__ U, const IN *__restrict__ GI, const float *__restrict__ Os,
        float *__restrict__ partVal, int *__restrict__ partJ)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char glds[];
    constexpr int EPC = 16 / (int)sizeof(IN);             // elements per 16-byte chunk
    constexpr int KS = kGRowBytes / (int)sizeof(IN);      // k per step
    const unsigned ao = blockIdx.y, a = ao / z;
    // XCD-aware order: workgroups are dealt to the 8 XCDs in turn, so XCD x gets a contiguous
    // run of tiles (whole tile rows), sharing the U_a and Gamma panels in its L2
    const unsigned T = rowTiles * colTiles;
    const unsigned q = (T % 8 == 0) ? (blockIdx.x % 8) * (T / 8) + blockIdx.x / 8 : blockIdx.x;
    // tiles in groups of kGGroup tile rows, column-major inside a group: the 32 workgroups an
    // XCD runs at once then cover about 4 x 8 tiles and share 12 operand panels in its L2,
    // where a row-major run of 2 x 16 tiles would share 18
    const unsigned groupSize = kGGroup * colTiles, firstRow = (q / groupSize) * kGGroup;
    const unsigned rowsHere = min(kGGroup, rowTiles - firstRow), inGroup = q % groupSize;
    const unsigned itile = firstRow + inGroup % rowsHere, jtile = inGroup / rowsHere;
    const unsigned i0 = itile * kGT, j0 = jtile * kGT;
    const unsigned t = threadIdx.x, l = t % 64, w = t / 64, wr = w / 4, wc = w % 4;
    const size_t rpad = (size_t)rowTiles * kGT;
    const IN *Ua = U + ((size_t)a * rpad + i0) * kpad;
    const IN *Gj = GI + (size_t)j0 * kpad;
    const float *sv = Os + (size_t)ao * kpad;

    // staging: 16-byte chunk ch of rows row0 + 64 q (q < 4), for both operands
    const unsigned ch = t % 8, row0 = t / 8;
    // the next k-step's operands and gamma*O scales are all requested before the matrix work
    // (named registers, no arrays: indexable arrays would live in scratch memory)
    uint4 ra0, ra1, ra2, ra3, rb0, rb1, rb2, rb3;
    float4 sc0, sc1;
    const size_t rowOff = (size_t)row0 * kpad + ch * EPC, q64 = (size_t)64 * kpad;
    auto gload = [&](unsigned k0) {
        const IN *ua = Ua + rowOff + k0, *gj = Gj + rowOff + k0;
        ra0 = *reinterpret_cast<const uint4 *>(ua);
        ra1 = *reinterpret_cast<const uint4 *>(ua + q64);
        ra2 = *reinterpret_cast<const uint4 *>(ua + 2 * q64);
        ra3 = *reinterpret_cast<const uint4 *>(ua + 3 * q64);
        rb0 = *reinterpret_cast<const uint4 *>(gj);
        rb1 = *reinterpret_cast<const uint4 *>(gj + q64);
        rb2 = *reinterpret_cast<const uint4 *>(gj + 2 * q64);
        rb3 = *reinterpret_cast<const uint4 *>(gj + 3 * q64);
        sc0 = *reinterpret_cast<const float4 *>(sv + k0 + ch * EPC);
        sc1 = (EPC == 8) ? *reinterpret_cast<const float4 *>(sv + k0 + ch * EPC + 4) : sc0;
    };
    auto sstore = [&](unsigned buf) {
        unsigned char *A = glds + (size_t)buf * 2 * kGOperandBytes;
        unsigned char *Bm = A + kGOperandBytes;
        const unsigned o = row0 * kGStride + ch * 16, q64s = 64 * kGStride;
        *reinterpret_cast<uint4 *>(A + o) = ra0;
        *reinterpret_cast<uint4 *>(A + o + q64s) = ra1;
        *reinterpret_cast<uint4 *>(A + o + 2 * q64s) = ra2;
        *reinterpret_cast<uint4 *>(A + o + 3 * q64s) = ra3;
        *reinterpret_cast<uint4 *>(Bm + o) = scale16<IN>(rb0, sc0, sc1);
        *reinterpret_cast<uint4 *>(Bm + o + q64s) = scale16<IN>(rb1, sc0, sc1);
        *reinterpret_cast<uint4 *>(Bm + o + 2 * q64s) = scale16<IN>(rb2, sc0, sc1);
        *reinterpret_cast<uint4 *>(Bm + o + 3 * q64s) = scale16<IN>(rb3, sc0, sc1);
    };

    floatx16_t acc[4][2];
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int v = 0; v < 16; v++)
                acc[x][y][v] = 0.0f;

    const unsigned nk = kpad / KS;
    gload(0);
    sstore(0);
    __syncthreads();
    for (unsigned ks = 0; ks < nk; ks++) {
        const unsigned buf = ks % 2;
        if (ks + 1 < nk)
            gload((ks + 1) * KS);
        const unsigned char *A = glds + (size_t)buf * 2 * kGOperandBytes;
        const unsigned char *Bm = A + kGOperandBytes;
        if constexpr (sizeof(IN) == 2) {
#pragma unroll
            for (int kk = 0; kk < KS; kk += 16) {
                bf16x8_t fa[4], fb[2];
                const unsigned kb = (kk + 8 * (l / 32)) * 2;
#pragma unroll
                for (int x = 0; x < 4; x++)
                    fa[x] = *reinterpret_cast<const bf16x8_t *>(
                        A + (wr * 128 + x * 32 + l % 32) * kGStride + kb);
#pragma unroll
                for (int y = 0; y < 2; y++)
                    fb[y] = *reinterpret_cast<const bf16x8_t *>(
                        Bm + (wc * 64 + y * 32 + l % 32) * kGStride + kb);
#pragma unroll
                for (int x = 0; x < 4; x++)
#pragma unroll
                    for (int y = 0; y < 2; y++)
                        acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[x], fb[y], acc[x][y], 0, 0, 0);
            }
        } else {
#pragma unroll
            for (int kk = 0; kk < KS; kk += 2) {
                float fa[4], fb[2];
                const unsigned kb = (kk + l / 32) * 4;
#pragma unroll
                for (int x = 0; x < 4; x++)
                    fa[x] = *reinterpret_cast<const float *>(
                        A + (wr * 128 + x * 32 + l % 32) * kGStride + kb);
#pragma unroll
                for (int y = 0; y < 2; y++)
                    fb[y] = *reinterpret_cast<const float *>(
                        Bm + (wc * 64 + y * 32 + l % 32) * kGStride + kb);
#pragma unroll
                for (int x = 0; x < 4; x++)
#pragma unroll
                    for (int y = 0; y < 2; y++)
                        acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[x], fb[y], acc[x][y], 0, 0, 0);
            }
        }
        if (ks + 1 < nk)
            sstore(buf ^ 1);
        __syncthreads();
    }

    // epilogue: per row, the first maximiser over this tile's 256 alphas
    float *redV = reinterpret_cast<float *>(glds);             // [4][256]
    int *redJ = reinterpret_cast<int *>(glds + 4 * kGT * sizeof(float));
#pragma unroll
    for (int x = 0; x < 4; x++) {
#pragma unroll
        for (int v = 0; v < 16; v++) {
            float bv = kLow;
            int bj = INT_MAX;
#pragma unroll
            for (int y = 0; y < 2; y++) {
                const unsigned j = j0 + wc * 64 + y * 32 + l % 32;
                const float val = acc[x][y][v];
                if (j < cols && (val > bv || (val == bv && (int)j < bj))) {
                    bv = val;
                    bj = (int)j;
                }
            }
            for (int off = 16; off >= 1; off >>= 1) {
                const float ov = __shfl_xor(bv, off, 32);
                const int oj = __shfl_xor(bj, off, 32);
                if (better(ov, oj, bv, bj)) {
                    bv = ov;
                    bj = oj;
                }
            }
            if (l % 32 == 0) {
                const unsigned row = wr * 128 + x * 32 + (v % 4) + 4 * (l / 32) + 8 * (v / 4);
                redV[wc * kGT + row] = bv;
                redJ[wc * kGT + row] = bj;
            }
        }
    }
    __syncthreads();
    if (t < kGT) {
        const unsigned i = i0 + t;
        float v0 = redV[t];
        int jv = redJ[t];
#pragma unroll
        for (int q = 1; q < 4; q++)
            if (better(redV[q * kGT + t], redJ[q * kGT + t], v0, jv)) {
                v0 = redV[q * kGT + t];
                jv = redJ[q * kGT + t];
            }
        if (i < rows) {
            const unsigned o = ao % z;
            const size_t idx = ((size_t)jtile * rows * m + (size_t)i * m + a) * z + o;
            partVal[idx] = v0;
            partJ[idx] = jv;
        }
    }
}

// Float rows (count / n rows of n) -> bf16 k-blocked [kpad / 32][rowsPad][32]: rows >= count / n
// and columns >= n are written as zeros, so no earlier clearing is needed. Each lane forms one
// 16-byte chunk: 8 consecutive k of one row.
__global__ __launch_bounds__(kGConvThreads) void dense_convert_blk(unsigned n, unsigned kpad,
        size_t count, size_t rowsPad, const float *__restrict__ G, __bf16 *__restrict__ GB)
{
    const size_t rowsValid = count / n, chunks = rowsPad * kpad / 8;
    for (size_t c = blockIdx.x * (size_t)kGConvThreads + threadIdx.x; c < chunks;
         c += (size_t)gridDim.x * kGConvThreads) {
        // chunk c: elements 8c .. 8c + 7 in order of block kb, row, k within the block
        const size_t kb = c / (rowsPad * 4), rem = c % (rowsPad * 4);
        const size_t row = rem / 4, k0 = kb * kGBlk + (rem % 4) * 8;
        bf16x8_t o;
        if (row < rowsValid && k0 + 8 <= n) {
            const float *src = G + row * n + k0;
#pragma unroll
            for (int e = 0; e < 8; e++)
                o[e] = (__bf16)src[e];
        } else {
#pragma unroll
            for (int e = 0; e < 8; e++)
                o[e] = (row < rowsValid && k0 + e < n) ? (__bf16)G[row * n + k0 + e]
                                                       : (__bf16)0.0f;
        }
        *reinterpret_cast<bf16x8_t *>(GB + 8 * c) = o;
    }
}

// B_q = bf16(float(GB) * s_{p0 + q}) for the cnt pairs q of a chunk, all k-blocked
// ([kpad / 32][rowsPad][32] per pair): the rounding dense_score<bf16> applies in registers.
// Each lane reads 8 consecutive elements of GB (16 bytes, one k-chunk of one row) once and
// writes them scaled for every pair of the chunk, so GB is read once per chunk, not per pair.
__global__ __launch_bounds__(kGConvThreads) void dense_scale_pairs(unsigned kpad, size_t rowsPad,
        unsigned p0, unsigned cnt, const __bf16 *__restrict__ GB, const float *__restrict__ Os,
        __bf16 *__restrict__ BP)
{
    const size_t per = rowsPad * kpad, per8 = per / 8;
    for (size_t c = blockIdx.x * (size_t)kGConvThreads + threadIdx.x; c < per8;
         c += (size_t)gridDim.x * kGConvThreads) {
        // chunk c: elements 8c .. 8c + 7 of [kb][row][32]; k = kb * 32 + (8c % 32) + e
        const size_t k0 = (c / (rowsPad * 4)) * kGBlk + (c % 4) * 8;
        const bf16x8_t g = *reinterpret_cast<const bf16x8_t *>(GB + 8 * c);
        float gf[8];
#pragma unroll
        for (int e = 0; e < 8; e++)
            gf[e] = (float)g[e];
        for (unsigned q = 0; q < cnt; q++) {
            const float *s = Os + (p0 + q) * (size_t)kpad + k0;
            const float4 s0 = *reinterpret_cast<const float4 *>(s);
            const float4 s1 = *reinterpret_cast<const float4 *>(s + 4);
            const float sv[8] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
            bf16x8_t o;
#pragma unroll
            for (int e = 0; e < 8; e++)
                o[e] = (__bf16)(gf[e] * sv[e]);
            *reinterpret_cast<bf16x8_t *>(BP + q * per + 8 * c) = o;
        }
    }
}

template <int N> __device__ __forceinline__ void wait_vmem()
{
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

// The bf16 GEMM on k-blocked operands (see the top of the file). Pair q = p0 + blockIdx.y reads
// A_a, a = q / z, from U ([kpad / 32][m * rowTiles * 256][32]) and B_q at GI + blockIdx.y *
// bPairStride ([kpad / 32][colTiles * 256][32]; stride 0 shares one B between the pairs).
__global__ __launch_bounds__(kGThreads) void dense_score_bf16(unsigned rows, unsigned cols,
        unsigned kpad, unsigned m, unsigned z, unsigned rowTiles, unsigned colTiles, unsigned p0,
        size_t bPairStride, const __bf16 *__restrict__ U, const __bf16 *__restrict__ GI,
        float *__restrict__ partVal, int *__restrict__ partJ)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char glds[];
    constexpr int NST = kGStages, OPB = kGStageOp, STB = 2 * OPB;
    constexpr int RPB = 16;                     // rows per 1 KB block (one wave's copy)
    constexpr int BPW = OPB / 1024 / 8;         // 1 KB blocks per wave per operand: 2
    constexpr int LPS = 2 * BPW;                // copies per lane per stage: 4
    const unsigned ao = p0 + blockIdx.y, a = ao / z;
    const unsigned T = rowTiles * colTiles;
    const unsigned q = (T % 8 == 0) ? (blockIdx.x % 8) * (T / 8) + blockIdx.x / 8 : blockIdx.x;
    const unsigned groupSize = kGGroup * colTiles, firstRow = (q / groupSize) * kGGroup;
    const unsigned rowsHere = min(kGGroup, rowTiles - firstRow), inGroup = q % groupSize;
    const unsigned itile = firstRow + inGroup % rowsHere, jtile = inGroup / rowsHere;
    const unsigned i0 = itile * kGT, j0 = jtile * kGT;
    const unsigned t = threadIdx.x, l = t % 64, w = __builtin_amdgcn_readfirstlane(t / 64);
    const unsigned wr = w / 4, wc = w % 4;
    const size_t uRows = (size_t)m * rowTiles * kGT, gRows = (size_t)colTiles * kGT;
    const __bf16 *Ua = U + ((size_t)a * rowTiles * kGT + i0) * kGBlk;
    const __bf16 *Gj = GI + (size_t)blockIdx.y * bPairStride + (size_t)j0 * kGBlk;

    // LDS layout of an operand in a stage: row r's four 16-byte chunks (k 0-7, 8-15, 16-23,
    // 24-31) at r * 64 + 16 * (chunk ^ sw(r)), sw(r) = ((r >> 1) ^ (r >> 3)) & 3: the 8 or 16
    // lanes that the LDS serves together (consecutive rows, one chunk) hit distinct banks.
    // Lane l of wave w copies, for block b = w * BPW + p, LDS bytes [16 l, 16 l + 16) of that
    // block: row b * 16 + l / 4 and the chunk stored at position l % 4.
    unsigned off[BPW];
#pragma unroll
    for (int p = 0; p < BPW; p++) {
        const unsigned row = (w * BPW + p) * RPB + l / 4;
        const unsigned c = (l % 4) ^ (((row >> 1) ^ (row >> 3)) & 3);
        off[p] = row * kGBlk + c * 8;
    }
    auto issue = [&](unsigned s) {
        unsigned char *st = glds + (size_t)(s % NST) * STB;
        const __bf16 *us = Ua + (size_t)s * uRows * kGBlk, *gs = Gj + (size_t)s * gRows * kGBlk;
#pragma unroll
        for (int p = 0; p < BPW; p++) {
            __builtin_amdgcn_global_load_lds((const void *)(us + off[p]),
                                             (void *)(st + (w * BPW + p) * 1024), 16, 0, 0);
            __builtin_amdgcn_global_load_lds((const void *)(gs + off[p]),
                                             (void *)(st + OPB + (w * BPW + p) * 1024), 16, 0, 0);
        }
    };
    // the swizzle depends on bits 1..4 of the row, the same for the wave's 32-row blocks
    unsigned aAddr[2], bAddr[2];                // per k-half (16 k), block 0
    {
        const unsigned ra = wr * 128 + l % 32, rb = wc * 64 + l % 32;
        const unsigned sa = ((ra >> 1) ^ (ra >> 3)) & 3, sb = ((rb >> 1) ^ (rb >> 3)) & 3;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            aAddr[h] = ra * 64 + (((2 * h + l / 32) ^ sa) * 16);
            bAddr[h] = OPB + rb * 64 + (((2 * h + l / 32) ^ sb) * 16);
        }
    }

    floatx16_t acc[4][2];
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int v = 0; v < 16; v++)
                acc[x][y][v] = 0.0f;
    // eight matrix instructions on (fa, fb) interleaved with the six LDS reads of the next
    // fragments (na, nb) from stage st, k-half h
    auto step = [&](const bf16x8_t (&fa)[4], const bf16x8_t (&fb)[2], bool load,
                    const unsigned char *st, int h, bf16x8_t (&na)[4], bf16x8_t (&nb)[2]) {
#pragma unroll
        for (int i = 0; i < 8; i++) {
            acc[i / 2][i % 2] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[i / 2], fb[i % 2],
                                                                      acc[i / 2][i % 2], 0, 0, 0);
            if (load && i < 6) {
                if (i < 4)
                    na[i] = *reinterpret_cast<const bf16x8_t *>(st + aAddr[h] + i * 32 * 64);
                else
                    nb[i - 4] = *reinterpret_cast<const bf16x8_t *>(st + bAddr[h] + (i - 4) * 32 * 64);
            }
        }
        if (load) {
#pragma unroll
            for (int i = 0; i < 6; i++) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);   // one matrix instruction
                __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);   // one LDS read
            }
            __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
        }
    };

    // Pipeline: stages 0 .. NST-2 are requested up front; iteration ks works on stage ks and
    // requests stage ks + NST - 1 into the buffer of stage ks - 1. Waits are explicit counts of
    // this wave's outstanding copies (a plain barrier, not __syncthreads, whose fence would
    // wait for all of them), followed by a barrier so that every wave's copies are visible.
    const unsigned nk = kpad / kGBlk;
#pragma unroll
    for (int s = 0; s < NST - 1; s++)
        if (s < (int)nk)
            issue(s);
    {
        // stage 0 landed: at most min(NST - 2, nk - 1) later stages outstanding
        const unsigned later = min((unsigned)(NST - 2), nk - 1);
        if (later >= 2) wait_vmem<2 * LPS>();
        else if (later == 1) wait_vmem<LPS>();
        else wait_vmem<0>();
    }
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    bf16x8_t fa0[4], fb0[2], fa1[4], fb1[2];
#pragma unroll
    for (int x = 0; x < 4; x++)
        fa0[x] = *reinterpret_cast<const bf16x8_t *>(glds + aAddr[0] + x * 32 * 64);
#pragma unroll
    for (int y = 0; y < 2; y++)
        fb0[y] = *reinterpret_cast<const bf16x8_t *>(glds + bAddr[0] + y * 32 * 64);
    for (unsigned ks = 0; ks + 1 < nk; ks++) {
        step(fa0, fb0, true, glds + (size_t)(ks % NST) * STB, 1, fa1, fb1);
        // stage ks + 1 landed (at most min(NST - 3, nk - ks - 2) later stages outstanding);
        // after the barrier every wave is done with stage ks - 1, whose buffer is refilled
        const unsigned later = min((unsigned)(NST - 3), nk - ks - 2);
        if (later >= 1) wait_vmem<LPS>();
        else wait_vmem<0>();
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        if (ks + NST - 1 < nk)
            issue(ks + NST - 1);
        step(fa1, fb1, true, glds + (size_t)((ks + 1) % NST) * STB, 0, fa0, fb0);
        asm volatile("" ::: "memory");
    }
    step(fa0, fb0, true, glds + (size_t)((nk - 1) % NST) * STB, 1, fa1, fb1);
    step(fa1, fb1, false, glds, 0, fa0, fb0);
    __syncthreads();

    // epilogue: per row, the first maximiser over this tile's 256 columns
    float *redV = reinterpret_cast<float *>(glds);             // [4][256]
    int *redJ = reinterpret_cast<int *>(glds + 4 * kGT * sizeof(float));
#pragma unroll
    for (int x = 0; x < 4; x++) {
#pragma unroll
        for (int v = 0; v < 16; v++) {
            float bv = kLow;
            int bj = INT_MAX;
#pragma unroll
            for (int y = 0; y < 2; y++) {
                const unsigned j = j0 + wc * 64 + y * 32 + l % 32;
                const float val = acc[x][y][v];
                if (j < cols && (val > bv || (val == bv && (int)j < bj))) {
                    bv = val;
                    bj = (int)j;
                }
            }
            for (int o2 = 16; o2 >= 1; o2 >>= 1) {
                const float ov = __shfl_xor(bv, o2, 32);
                const int oj = __shfl_xor(bj, o2, 32);
                if (better(ov, oj, bv, bj)) {
                    bv = ov;
                    bj = oj;
                }
            }
            if (l % 32 == 0) {
                const unsigned row = wr * 128 + x * 32 + (v % 4) + 4 * (l / 32) + 8 * (v / 4);
                redV[wc * kGT + row] = bv;
                redJ[wc * kGT + row] = bj;
            }
        }
    }
    __syncthreads();
    if (t < kGT) {
        const unsigned i = i0 + t;
        float v0 = redV[t];
        int jv = redJ[t];
#pragma unroll
        for (int q2 = 1; q2 < 4; q2++)
            if (better(redV[q2 * kGT + t], redJ[q2 * kGT + t], v0, jv)) {
                v0 = redV[q2 * kGT + t];
                jv = redJ[q2 * kGT + t];
            }
        if (i < rows) {
            const unsigned o = ao % z;
            const size_t idx = ((size_t)jtile * rows * m + (size_t)i * m + a) * z + o;
            partVal[idx] = v0;
            partJ[idx] = jv;
        }
    }
}

// The bf16 reference path (dense_score<__bf16> on row-major operands) instead of
// dense_score_bf16, for comparisons: NOVA_AMD_DENSE_KERNEL=legacy.
inline bool dense_legacy_bf16()
{
    const char *e = getenv("NOVA_AMD_DENSE_KERNEL");
    return e != nullptr && strcmp(e, "legacy") == 0;
}

} // namespace
} // namespace nova_amd

#endif
