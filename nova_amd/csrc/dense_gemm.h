// nova-mi355x: the dense GEMM with a fused row argmax on the matrix cores (gfx950), shared by
// the dense-belief PBVI scoring and the batched alpha-vector policy evaluation.
//
// For each pair p = (a, o) of blockIdx.y = a * z + o it forms
//   C_p[i][j] = sum_k A_a[i][k] * s_p[k] * B[j][k]      (i < rows, j < cols)
// and writes, per row i and tile of 256 columns, the first maximiser of C_p[i][.] over the
// tile's columns j < cols (value and index) to partVal / partJ at
// ((jtile * rows * m + i * m + a) * z + o). Operands live in HBM in the input type IN (bf16 or
// float), zero-padded to whole 256-row tiles (A_a: rowTiles * 256 rows per a, B: colTiles * 256
// rows) and to kpad columns (a multiple of 64), so the k loop needs no bounds checks; s holds
// one row of kpad scales per pair.
//
// A workgroup of 8 waves owns a 256 x 256 tile of C: each k-step stages 256 x 128-byte slices
// of A_a and of B (B scaled by s_p in registers on the way) into one of two LDS buffers while
// the matrix cores work on the other; each wave accumulates a 128 x 64 sub-tile as 4 x 2 blocks
// of 32 x 32 (v_mfma_f32_32x32x16_bf16 on bf16 operands, v_mfma_f32_32x32x2_f32 on fp32
// operands, fp32 accumulation in both). Workgroups are numbered so that each XCD takes a
// contiguous run of tiles, ordered in groups of four tile rows, and shares their A and B
// panels in its L2.
#ifndef NOVA_AMD_DENSE_GEMM_H
#define NOVA_AMD_DENSE_GEMM_H

#include <climits>

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

} // namespace
} // namespace nova_amd

#endif
