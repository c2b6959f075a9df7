// nova-mi355x: point-based value iteration (PBVI) backups on one MI355X (gfx950, wave64).
//
// The backup of belief i (src/pomdp/algorithms/pomdp_pbvi_cpu.cpp:37-162):
//   score_{a,o}(j) = sum_k B[i,k] * gamma * sum_l (O[a,sp,o]*T[s_k,a,l]) * Gamma[j,sp],
//                    s_k = Z[i,k], sp = S[s_k,a,l]      (valid prefixes of Z and S only)
//   j*_{a,o}       = first argmax_j score_{a,o}(j)       (strict '>' from -1e10)
//   alpha_a(s)     = R[s,a] + sum_o gamma * sum_l (O*T) * Gamma[j*_{a,o}, sp]
//   a*             = first argmax_a sum_k alpha_a(Z[i,k]) * B[i,k];  Gamma'[i] = alpha_{a*}
//
// Per iteration (all launched on the library stream, captured as one HIP graph by execute):
//   1. gamma_transpose : Gamma (r x n) -> GT (n x r), so that the r alphas of one successor
//                        state are contiguous and a wave's loads of GT[sp, j..j+63] coalesce.
//   2. pbvi_score      : one workgroup per (belief, action, tile of 512 alphas). The valid
//                        (k, l) pairs of the belief's support are compacted into LDS with the
//                        matching O*T factors of a chunk of 16 observations; each lane owns 2
//                        alphas and accumulates 16 observation scores per alpha from one GT
//                        load per (k, l). Per-tile argmaxes are reduced with wave shuffles.
//   3. pbvi_action     : one workgroup per belief: merges the tiles (first maximiser wins),
//                        evaluates alpha_a on the belief's support in the reference order and
//                        picks the action.
//   4. pbvi_materialize: one workgroup per belief writes Gamma'[i] and pi[i].
// Two arithmetic modes share this structure:
//   exact - every product and sum rounded separately in the reference's order (no FMA), so
//           Gamma' and pi equal nova's CPU results bit for bit;
//   fast  - (default) scores accumulate with FMA from W = gamma*B*O*T; the materialisation
//           stages h(sp) = sum_o O[a*,sp,o]*Gamma[j*_o,sp] in LDS and then forms
//           R + gamma*sum_l T*h(S) per state. Results agree with the reference within float
//           rounding; alpha choices can differ only at near-ties.
// Model constant built once per initialize: OT[a][o][s][l] = O[a,S[s,a,l],o]*T[s,a,l] (0 past
// the end of the successor list). It, GT and the per-iteration argmax records live in one
// device workspace referenced by POMDPPBVIGPU::d_alphaBA.
//
// Every kernel reads the sizes and pointers of its problem from a device-resident instance
// descriptor (PbviInst) and maps its workgroup to an instance through per-stage prefix offsets,
// so one launch per stage serves any number of POMDPs of uneven size: a plain solve is a batch
// of one, and pomdp_pbvi_batch_execute_gpu captures the iterations of many POMDPs in one graph
// (each instance's results are identical to its plain solve). The scoring kernel has variants
// (alphas per lane 1, 2 or 4); a tuner times them once per size class and caches the choice.
#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <nova/amd/nova_amd.h>
#include <nova/pomdp/algorithms/pomdp_pbvi_gpu.h>

#include "dense_gemm.h"
#include "nova_hip.h"

namespace nova_amd {

// One problem of a (possibly single-element) batch, as the kernels see it.
// Pointers the kernels read out of descriptors are generic; casting them to the global address
// space lets the compiler emit global_load/global_store instead of flat accesses (which also
// wait on the LDS counter).
#define NOVA_GLOBAL __attribute__((address_space(1)))

struct PbviInst {
    unsigned n, ns, m, z, r, rz, tiles, horizon;
    unsigned rb;           // beliefs backed up (PBVI: r; Perseus: 1). r counts the alphas.
    unsigned jpt;          // alphas per lane of the scoring variant chosen for this instance
    int useLds;            // materialisation stages n floats in LDS
    float gamma;
    const int *S;
    const float *T, *O, *R;
    const int *Z;
    const float *B;
    float *G0, *G1;        // Gamma / GammaPrime
    unsigned *pi;
    float *OT, *GT, *partVal;
    int *SA;               // per-action successor rows SA[(a*n + s)*ns + l] = S[s,a,l]
    int *pairSp;           // fast mode: valid (k, l) pairs of belief i under action a,
    float *pairW;          //   pairSp/pairW[(i*m + a)*kmax + u] = S[s_k,a,l], gamma*b_k*T[s_k,a,l]
    unsigned *pairCnt;     //   pairCnt[i*m + a] = number of pairs
    unsigned kmax;         //   rz*ns (the lists exist when it is at most kUMax)
    unsigned rS;           // row stride of GT: r rounded up to a whole matrix-core tile
    float *TA, *RT, *OTr;  // TA likewise for T; RT[a*n + s] = R[s,a]; OTr[(a*z + o)*n + s'] = O[a,s',o]
    int *partJ, *jstar;
    float *aval;
    int *act;
    int *recJ;             // optional per-iteration record: recJ[(iter*r + i)*z + o] = j*
    int *recA;             // and recA[iter*r + i] = chosen action (-1: none)
    int fused;             // each iteration is one pbvi_fused launch (small instances)
    int actionMajor;       // matrix-core scorers: workgroups enumerate (action, belief) with the
                           // action slower, so that concurrent ones read one action's O rows
};

namespace {

constexpr int kPB = 256;            // lanes per workgroup
constexpr int kUMax = 512;          // max valid (k, l) pairs of one belief-action in LDS
constexpr int kKMax = 512;          // max belief support length handled in LDS
constexpr int kActMax = 16384;      // exact mode: max support length (the action kernel's LDS)
constexpr int kZC = 16;             // observations per scoring chunk
constexpr int kLdsRowFloats = 36 * 1024;  // Gamma row / h vector staged in LDS (144 KB)
constexpr unsigned kNoHorizon = 0xffffffffu;

// Instance of workgroup `id` given prefix offsets off[0..count] (off[b] <= id < off[b+1]).
__device__ __forceinline__ unsigned find_inst(const unsigned *__restrict__ off, unsigned count,
                                              unsigned id)
{
    unsigned lo = 0, hi = count - 1;
    while (lo < hi) {
        const unsigned mid = (lo + hi + 1) / 2;
        if (off[mid] <= id) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ const float *src_gamma(const PbviInst &P, unsigned iter, unsigned flip)
{
    return ((iter + flip) % 2 == 0) ? P.G0 : P.G1;
}

__device__ __forceinline__ float *dst_gamma(const PbviInst &P, unsigned iter, unsigned flip)
{
    return ((iter + flip) % 2 == 0) ? P.G1 : P.G0;
}

// --------------------------------------------------------------------- model constant

__global__ __launch_bounds__(kPB) void pbvi_build_ot(unsigned n, unsigned ns, unsigned m,
        unsigned z, const int *__restrict__ S, const float *__restrict__ T,
        const float *__restrict__ O, float *__restrict__ OT)
{
    // OT[((a*z + o)*n + s)*ns + l]
    const size_t total = (size_t)m * z * n * ns;
    for (size_t x = blockIdx.x * (size_t)kPB + threadIdx.x; x < total;
         x += (size_t)gridDim.x * kPB) {
        const unsigned l = x % ns;
        size_t y = x / ns;
        const unsigned s = y % n;
        y /= n;
        const unsigned o = y % z;
        const unsigned a = y / z;
        const size_t row = ((size_t)s * m + a) * ns;
        bool valid = true;
        for (unsigned q = 0; q <= l; q++)
            valid = valid && S[row + q] >= 0;
        float v = 0.0f;
        if (valid) {
            const unsigned sp = (unsigned)S[row + l];
            v = __fmul_rn(O[((size_t)a * n + sp) * z + o], T[row + l]);
        }
        OT[x] = v;
    }
}

// Action-major copies for the fast materialisation: one action's successor rows, rewards and
// observation planes are then contiguous over states.
__global__ __launch_bounds__(kPB) void pbvi_build_aux(unsigned n, unsigned ns, unsigned m,
        unsigned z, const int *__restrict__ S, const float *__restrict__ T,
        const float *__restrict__ R, const float *__restrict__ O, int *__restrict__ SA,
        float *__restrict__ TA, float *__restrict__ RT, float *__restrict__ OTr)
{
    const size_t stride = (size_t)gridDim.x * kPB, x0 = blockIdx.x * (size_t)kPB + threadIdx.x;
    const size_t rows = (size_t)m * n * ns;
    for (size_t x = x0; x < rows; x += stride) {
        const unsigned l = x % ns;
        const size_t y = x / ns;
        const unsigned s = y % n, a = y / n;
        const size_t src = ((size_t)s * m + a) * ns + l;
        SA[x] = S[src];
        TA[x] = T[src];
    }
    for (size_t x = x0; x < (size_t)m * n; x += stride) {
        const unsigned s = x % n, a = x / n;
        RT[x] = R[(size_t)s * m + a];
    }
    for (size_t x = x0; x < (size_t)m * z * n; x += stride) {
        const unsigned sp = x % n;
        const size_t y = x / n;
        const unsigned o = y % z, a = y / z;
        OTr[x] = O[((size_t)a * n + sp) * z + o];
    }
}

// The valid (k, l) pairs of every (belief, action) with their Gamma-independent factor
// gamma * b_k * T[s_k,a,l], in support order (fast-mode scoring reads these instead of
// compacting the support in every workgroup).
__global__ __launch_bounds__(kPB) void pbvi_build_pairs(unsigned r, unsigned m, unsigned ns,
        unsigned rz, unsigned kmax, float gamma, const int *__restrict__ S,
        const float *__restrict__ T, const int *__restrict__ Z, const float *__restrict__ B,
        int *__restrict__ pairSp, float *__restrict__ pairW, unsigned *__restrict__ pairCnt)
{
    const size_t ia = blockIdx.x * (size_t)kPB + threadIdx.x;
    if (ia >= (size_t)r * m)
        return;
    const unsigned i = ia / m, a = ia % m;
    unsigned cnt = 0;
    for (unsigned k = 0; k < rz; k++) {
        const int s = Z[(size_t)i * rz + k];
        if (s < 0)
            break;
        const float gb = gamma * B[(size_t)i * rz + k];
        const size_t row = ((size_t)s * m + a) * ns;
        for (unsigned l = 0; l < ns; l++) {
            const int sp = S[row + l];
            if (sp < 0)
                break;
            pairSp[ia * kmax + cnt] = sp;
            pairW[ia * kmax + cnt] = gb * T[row + l];
            cnt++;
        }
    }
    pairCnt[ia] = cnt;
}

// ----------------------------------------------------------------------- 1. transpose

__global__ __launch_bounds__(kPB) void pbvi_transpose(const PbviInst *__restrict__ insts,
        const unsigned *__restrict__ off, unsigned count, unsigned iter, unsigned flip)
{
    __shared__ float tile[64][65];
    if (blockIdx.x >= off[count])
        return;   // a grid larger than the plan's workgroups does nothing
    const unsigned b = find_inst(off, count, blockIdx.x);
    const PbviInst &P = insts[b];
    if (iter >= P.horizon)
        return;
    const unsigned local = blockIdx.x - off[b];
    const unsigned tilesX = (P.n + 63) / 64;
    const unsigned s0 = (local % tilesX) * 64, j0 = (local / tilesX) * 64;
    const float *G = src_gamma(P, iter, flip);
    const unsigned tx = threadIdx.x % 64, ty = threadIdx.x / 64;
    for (unsigned q = ty; q < 64; q += 4) {
        const unsigned j = j0 + q, s = s0 + tx;
        tile[q][tx] = (j < P.r && s < P.n) ? G[(size_t)j * P.n + s] : 0.0f;
    }
    __syncthreads();
    // columns r..rS-1 of each row are written as zeros
    for (unsigned q = ty; q < 64; q += 4) {
        const unsigned s = s0 + q, j = j0 + tx;
        if (s < P.n && j < P.rS)
            ((NOVA_GLOBAL float *)P.GT)[(size_t)s * P.rS + j] = tile[tx][q];
    }
}

// --------------------------------------------------------------------------- 2. score

// Compacted support of one belief for one action, in LDS.
struct SupportLds {
    int kEnd;                 // valid support length (first Z < 0)
    int uEnd;                 // number of valid (k, l) pairs
    int sK[kKMax];            // support states
    float bK[kKMax];          // belief values
    int startK[kKMax + 1];    // first pair index of support entry k
    int spU[kUMax];           // successor state of pair u
    int otU[kUMax];           // s_k*ns + l of pair u (offset into an OT plane)
};

// Fills the SupportLds of belief i and action a cooperatively (all kPB lanes).
__device__ void build_support(SupportLds &L, unsigned i, unsigned a, unsigned m, unsigned ns,
        unsigned rz, const int *__restrict__ S, const int *__restrict__ Z,
        const float *__restrict__ B, int *cnt)
{
    const unsigned t = threadIdx.x;
    if (t == 0)
        L.kEnd = (int)rz;
    __syncthreads();
    for (unsigned k = t; k < rz; k += kPB) {
        const int s = Z[(size_t)i * rz + k];
        L.sK[k] = s;
        L.bK[k] = B[(size_t)i * rz + k];
        if (s < 0)
            atomicMin(&L.kEnd, (int)k);
    }
    __syncthreads();
    const int kEnd = L.kEnd;
    if (kEnd <= 64) {
        // one wave: lane k counts entry k's valid pairs and the wave scans the counts
        if (t < 64) {
            int c = 0;
            if ((int)t < kEnd) {
                const size_t row = ((size_t)L.sK[t] * m + a) * ns;
                while (c < (int)ns && S[row + c] >= 0)
                    c++;
            }
            for (int off = 1; off < 64; off <<= 1) {
                const int y = __shfl_up(c, off);
                if ((int)t >= off)
                    c += y;
            }
            if (t == 0)
                L.startK[0] = 0;
            if ((int)t < kEnd)
                L.startK[t + 1] = c;
            if ((int)t == kEnd - 1 || (kEnd == 0 && t == 0))
                L.uEnd = (kEnd == 0) ? 0 : c;
        }
        __syncthreads();
    } else {
        for (int k = t; k < kEnd; k += kPB) {
            const size_t row = ((size_t)L.sK[k] * m + a) * ns;
            int c = 0;
            while (c < (int)ns && S[row + c] >= 0)
                c++;
            cnt[k] = c;
        }
        __syncthreads();
        // exclusive prefix sum of cnt[0..kEnd) (Hillis-Steele over startK, ping-pong via cnt)
        for (int k = t; k <= kEnd; k += kPB)
            L.startK[k] = (k == 0) ? 0 : cnt[k - 1];
        __syncthreads();
        for (int off = 1; off <= kEnd; off <<= 1) {
            for (int k = t; k <= kEnd; k += kPB)
                cnt[k] = (k >= off) ? L.startK[k] + L.startK[k - off] : L.startK[k];
            __syncthreads();
            for (int k = t; k <= kEnd; k += kPB)
                L.startK[k] = cnt[k];
            __syncthreads();
        }
        if (t == 0)
            L.uEnd = L.startK[kEnd];
        __syncthreads();
    }
    for (int x = t; x < kEnd * (int)ns; x += kPB) {
        const int k = x / (int)ns, l = x % (int)ns;
        const int u = L.startK[k] + l;
        if (u < L.startK[k + 1]) {
            const size_t row = ((size_t)L.sK[k] * m + a) * ns;
            L.spU[u] = S[row + l];
            L.otU[u] = L.sK[k] * (int)ns + l;
        }
    }
    __syncthreads();
}

// Fills the SupportLds with the support entries [k0, k1) of belief i under action a (k1 at most
// the valid length, k1 - k0 <= kKMax, and their valid pairs at most kUMax): the exact mode's
// chunks of a long support. Entry k of the chunk is support entry k0 + k.
__device__ void build_support_range(SupportLds &L, unsigned i, unsigned a, unsigned m,
        unsigned ns, unsigned rz, int k0, int k1, const int *__restrict__ S,
        const int *__restrict__ Z, const float *__restrict__ B, int *cnt)
{
    const int t = (int)threadIdx.x, kc = k1 - k0;
    __syncthreads();                  // the previous chunk's readers are done
    for (int k = t; k < kc; k += kPB) {
        const int s = Z[(size_t)i * rz + k0 + k];
        L.sK[k] = s;
        L.bK[k] = B[(size_t)i * rz + k0 + k];
        const size_t row = ((size_t)s * m + a) * ns;
        int c = 0;
        while (c < (int)ns && S[row + c] >= 0)
            c++;
        cnt[k] = c;
    }
    if (t == 0)
        L.kEnd = kc;
    __syncthreads();
    for (int k = t; k <= kc; k += kPB)
        L.startK[k] = (k == 0) ? 0 : cnt[k - 1];
    __syncthreads();
    for (int off = 1; off <= kc; off <<= 1) {
        for (int k = t; k <= kc; k += kPB)
            cnt[k] = (k >= off) ? L.startK[k] + L.startK[k - off] : L.startK[k];
        __syncthreads();
        for (int k = t; k <= kc; k += kPB)
            L.startK[k] = cnt[k];
        __syncthreads();
    }
    if (t == 0)
        L.uEnd = L.startK[kc];
    for (int x = t; x < kc * (int)ns; x += kPB) {
        const int k = x / (int)ns, l = x % (int)ns;
        const int u = L.startK[k] + l;
        if (u < L.startK[k + 1]) {
            const size_t row = ((size_t)L.sK[k] * m + a) * ns;
            L.spU[u] = S[row + l];
            L.otU[u] = L.sK[k] * (int)ns + l;
        }
    }
    __syncthreads();
}

// Valid support length of belief i (the first Z < 0, else rz), computed by all kPB lanes.
__device__ int support_length(unsigned i, unsigned rz, const int *__restrict__ Z, int *slot)
{
    if (threadIdx.x == 0)
        *slot = (int)rz;
    __syncthreads();
    for (unsigned k = threadIdx.x; k < rz; k += kPB)
        if (Z[(size_t)i * rz + k] < 0)
            atomicMin(slot, (int)k);
    __syncthreads();
    const int v = *slot;
    __syncthreads();
    return v;
}

template <bool EXACT, int JPT>
__global__ __launch_bounds__(kPB) void pbvi_score(const PbviInst *__restrict__ insts,
        const unsigned *__restrict__ off, unsigned count, unsigned iter)
{
    __shared__ SupportLds L;
    __shared__ float w[kUMax * kZC];         // [u][o] factors of the current chunk
    __shared__ int scratch[kKMax + 1];
    __shared__ float redV[kPB / 64][kZC];
    __shared__ int redJ[kPB / 64][kZC];

    if (blockIdx.x >= off[count])
        return;   // a grid larger than the plan's workgroups does nothing
    const unsigned b = find_inst(off, count, blockIdx.x);
    const PbviInst &P = insts[b];
    if (iter >= P.horizon)
        return;
    const unsigned n = P.n, ns = P.ns, m = P.m, z = P.z, r = P.r;
    const float gamma = P.gamma;
    const unsigned local = blockIdx.x - off[b];
    const unsigned rb = P.rb;
    const unsigned ia = local % (rb * m);        // belief * m + action
    const unsigned tile = local / (rb * m);
    const unsigned i = ia / m, a = ia % m;
    const unsigned t = threadIdx.x, lane = t % 64, wave = t / 64;

    int kEnd, uEnd;
    // exact mode with a support whose pairs pass the LDS: its entries are staged in chunks of
    // kc entries (at most kUMax pairs each) inside the observation loop, the sums over the
    // chunks kept in order
    const bool chunked = EXACT && (P.rz > (unsigned)kKMax || P.rz * ns > (unsigned)kUMax);
    int kFull = 0, kc = 0;
    if (EXACT && chunked) {
        kFull = support_length(i, P.rz, P.Z, &scratch[0]);
        kc = min(kKMax, kUMax / (int)ns);
        kEnd = 0;
        uEnd = 0;
    } else if (EXACT) {
        build_support(L, i, a, m, ns, P.rz, P.S, P.Z, P.B, scratch);
        kEnd = L.kEnd;
        uEnd = L.uEnd;
    } else {
        // the precomputed pair list of (i, a): successors to LDS, factors read when staged
        uEnd = (int)P.pairCnt[ia];
        kEnd = 0;
        for (int u = t; u < uEnd; u += kPB)
            L.spU[u] = P.pairSp[(size_t)ia * P.kmax + u];
        __syncthreads();
    }
    const NOVA_GLOBAL float *GT = (const NOVA_GLOBAL float *)P.GT;
    const NOVA_GLOBAL float *OT = (const NOVA_GLOBAL float *)P.OT;
    const NOVA_GLOBAL float *pw = EXACT ? nullptr : (const NOVA_GLOBAL float *)P.pairW + (size_t)ia * P.kmax;
    const NOVA_GLOBAL float *Oa = (const NOVA_GLOBAL float *)P.O + (size_t)a * n * z;

    int jq[JPT];
    bool live[JPT];
#pragma unroll
    for (int q = 0; q < JPT; q++) {
        jq[q] = (int)(tile * (kPB * JPT) + t + q * kPB);
        live[q] = jq[q] < (int)r;
    }

    for (unsigned o0 = 0; o0 < z; o0 += kZC) {
        const int zc = min((int)kZC, (int)(z - o0));
        float val[kZC][JPT];
#pragma unroll
        for (int oo = 0; oo < kZC; oo++)
#pragma unroll
            for (int q = 0; q < JPT; q++)
                val[oo][q] = 0.0f;
        // support chunks (one unless chunked)
        for (int k0 = 0; k0 < (chunked ? kFull : 1); k0 += (chunked ? kc : 1)) {
        if (chunked) {
            build_support_range(L, i, a, m, ns, P.rz, k0, min(kFull, k0 + kc), P.S, P.Z, P.B,
                                scratch);
            kEnd = L.kEnd;
            uEnd = L.uEnd;
        } else {
            __syncthreads();          // the previous observation chunk's readers of w are done
        }
        // stage the chunk's factors: exact -> O*T; fast -> (gamma*b_k*T) * O
        for (int x = t; x < uEnd * kZC; x += kPB) {
            const int u = x / kZC, oo = x % kZC;
            float v = 0.0f;
            if (oo < zc) {
                if (EXACT)
                    v = OT[((size_t)(a * z + o0 + oo) * n) * ns + L.otU[u]];
                else
                    v = pw[u] * Oa[(size_t)L.spU[u] * z + o0 + oo];
            }
            w[u * kZC + oo] = v;
        }
        __syncthreads();

        float vtk[EXACT ? kZC : 1][JPT];
#pragma unroll
        for (int oo = 0; oo < (EXACT ? kZC : 1); oo++)
#pragma unroll
            for (int q = 0; q < JPT; q++)
                vtk[oo][q] = 0.0f;
        // exact: closes support entry k (its pairs summed in the reference's order) into the
        // scores and starts the next one. Entries close right after their last pair, and an
        // entry without pairs right after the entry before it, so the scores see the same
        // operations in the same order as an entry-by-entry loop
        auto closeEntry = [&](int k) {
            const float bk = L.bK[k];
#pragma unroll
            for (int oo = 0; oo < kZC; oo++)
#pragma unroll
                for (int q = 0; q < JPT; q++) {
                    const float v = __fmul_rn(vtk[oo % (EXACT ? kZC : 1)][q], gamma);
                    val[oo][q] = __fadd_rn(val[oo][q], __fmul_rn(v, bk));
                    vtk[oo % (EXACT ? kZC : 1)][q] = 0.0f;
                }
        };
        // the pairs of all entries in order; the Gamma^T values of pair u + kPre are requested
        // while pair u is summed (a ring of kPre register sets)
        constexpr int kPre = 4;
        int k = 0;
        if (EXACT)
            while (k < kEnd && L.startK[k + 1] == 0)
                closeEntry(k++);
        if (uEnd > 0) {
            float gr[kPre][JPT];
            auto loadG = [&](int u, float (&dst)[JPT]) {
                const size_t rowOff = (size_t)L.spU[u] * P.rS;
#pragma unroll
                for (int q = 0; q < JPT; q++)
                    dst[q] = live[q] ? GT[rowOff + jq[q]] : 0.0f;
            };
#pragma unroll
            for (int d = 0; d < kPre; d++)
                loadG(min(d, uEnd - 1), gr[d]);
            for (int u0 = 0; u0 < uEnd; u0 += kPre) {
#pragma unroll
                for (int d = 0; d < kPre; d++) {
                    const int u = u0 + d;
                    if (u < uEnd) {
                        const float *wu = w + u * kZC;
#pragma unroll
                        for (int oo = 0; oo < kZC; oo++) {
                            const float f = wu[oo];
#pragma unroll
                            for (int q = 0; q < JPT; q++) {
                                if (EXACT)
                                    vtk[oo % (EXACT ? kZC : 1)][q] = __fadd_rn(
                                            vtk[oo % (EXACT ? kZC : 1)][q], __fmul_rn(f, gr[d][q]));
                                else
                                    val[oo][q] = __builtin_fmaf(f, gr[d][q], val[oo][q]);
                            }
                        }
                        if (EXACT)
                            while (k < kEnd && L.startK[k + 1] == u + 1)
                                closeEntry(k++);
                    }
                    loadG(min(u + kPre, uEnd - 1), gr[d]);
                }
            }
        }
        if (EXACT)
            while (k < kEnd)                  // entries without pairs after the last pair
                closeEntry(k++);
        }   // support chunks

        // per lane: first maximiser over its alphas (visited in increasing j)
        float bv[kZC];
        int bj[kZC];
#pragma unroll
        for (int oo = 0; oo < kZC; oo++) {
            bv[oo] = kLow;
            bj[oo] = INT_MAX;
#pragma unroll
            for (int q = 0; q < JPT; q++)
                if (live[q] && val[oo][q] > bv[oo]) {
                    bv[oo] = val[oo][q];
                    bj[oo] = jq[q];
                }
        }
        // wave reduction, then across the 4 waves
#pragma unroll
        for (int oo = 0; oo < kZC; oo++) {
            for (int off2 = 32; off2 >= 1; off2 >>= 1) {
                const float ov = __shfl_xor(bv[oo], off2);
                const int oj = __shfl_xor(bj[oo], off2);
                if (better(ov, oj, bv[oo], bj[oo])) {
                    bv[oo] = ov;
                    bj[oo] = oj;
                }
            }
            if (lane == 0) {
                redV[wave][oo] = bv[oo];
                redJ[wave][oo] = bj[oo];
            }
        }
        __syncthreads();
        if (t < (unsigned)zc) {
            float v = redV[0][t];
            int j = redJ[0][t];
            for (int wv = 1; wv < kPB / 64; wv++)
                if (better(redV[wv][t], redJ[wv][t], v, j)) {
                    v = redV[wv][t];
                    j = redJ[wv][t];
                }
            const size_t idx = ((size_t)tile * rb * m + ia) * z + o0 + t;
            P.partVal[idx] = v;
            P.partJ[idx] = j;
        }
        __syncthreads();
    }
}


// Stages the matrix-core A operand of a chunk of cPad pairs into LDS: w[u][oo] =
// pw[u] * O[a, sp_u, o0 + oo] for the uLim valid pairs and the zc valid observations of the
// block of ROWS, zero elsewhere (up to the next whole batch of the workgroup, within the CAP
// words of w). Every lane requests kWBatch elements' factors at once, with indices clamped into
// range and the invalid elements zeroed after the loads, so that staging costs a few memory
// round trips rather than one per element.
constexpr int kWBatch = 4;

template <int ROWS, int NT, int CAP>
__device__ __forceinline__ void stage_w(float *__restrict__ w, const int *__restrict__ spU,
        int cPad, int uLim, int zc, unsigned o0, unsigned z, const NOVA_GLOBAL float *pw,
        const NOVA_GLOBAL float *Oa, unsigned t)
{
    static_assert(CAP % (kWBatch * NT) == 0, "whole batches fit the staging array");
    const int total = cPad * ROWS;           // <= CAP
#pragma unroll 1
    for (int x0 = (int)t; x0 - (int)t < total; x0 += kWBatch * NT) {
        float pv[kWBatch], ov[kWBatch];
#pragma unroll
        for (int k = 0; k < kWBatch; k++) {
            const int x = x0 + k * NT;
            const int u = min(x / ROWS, uLim - 1), oo = min(x % ROWS, zc - 1);
            pv[k] = pw[u];
            ov[k] = Oa[(size_t)spU[u] * z + o0 + oo];
        }
#pragma unroll
        for (int k = 0; k < kWBatch; k++) {
            const int x = x0 + k * NT;
            w[x] = (x / ROWS < uLim && x % ROWS < zc) ? pv[k] * ov[k] : 0.0f;
        }
    }
}

// Fast-mode scoring on the matrix cores. For belief i, action a and a tile of kMTile alphas it
// forms C[o][j] = sum_u W[u][o] * GT[sp_u][j], W = gamma * b_k * O[a,sp_u,o] * T[s_k,a,l] over
// the belief's valid (k, l) pairs u, as v_mfma_f32_32x32x2_f32 products: rows are 32
// observations, columns alphas, the reduction runs over the pairs two at a time. W is staged in
// LDS in chunks of kWChunk pairs; each wave owns 128 alphas as four 32 x 32 blocks and loads
// its B operand straight from GT (two 128-byte row segments per instruction). The epilogue
// writes the same per-tile argmax partials as pbvi_score.
constexpr int kMTile = 512;          // alphas per workgroup
constexpr int kWChunk = 256;         // support pairs per staged W chunk

template <int NW, int NB, int PS, int NSET, bool BUF>
__global__ __launch_bounds__(NW * 64) void pbvi_score_mfma(const PbviInst *__restrict__ insts,
        const unsigned *__restrict__ off, unsigned count, unsigned iter)
{
    static_assert(NW * NB * 32 == kMTile, "a workgroup covers one alpha tile");
    static_assert(NSET >= 2, "at least one trip's operands in flight");
    constexpr int NT = NW * 64;
    constexpr int HALF = 2 * PS;              // pairs per trip (PS pair-steps of two pairs)
    constexpr int RING = NSET * HALF;         // pairs per turn of the ring of NSET register sets
    static_assert(kWChunk % RING == 0, "whole turns of the ring per chunk");
    static_assert(NSET * PS * NB <= 63, "a turn of the ring fits the vector-memory counter");
    __shared__ int spU[kWChunk];              // successor of each pair of the current chunk
    __shared__ float w[kWChunk * 32];         // [u][o] for the current chunk and 32 observations
    __shared__ float redV[NW][32];
    __shared__ int redJ[NW][32];

    if (blockIdx.x >= off[count])
        return;   // a grid larger than the plan's workgroups does nothing
    const unsigned b = find_inst(off, count, blockIdx.x);
    const PbviInst &P = insts[b];
    if (iter >= P.horizon)
        return;
    const unsigned n = P.n, m = P.m, z = P.z, r = P.r;
    const unsigned local = blockIdx.x - off[b];
    const unsigned rb = P.rb;
    const unsigned q = local % (rb * m);
    // ia = belief * m + action; the enumeration order of (belief, action) is the instance's
    const unsigned ia = P.actionMajor ? (q % rb) * m + q / rb : q;
    const unsigned tile = local / (rb * m);
    const unsigned a = ia % m;
    const unsigned t = threadIdx.x, lane = t % 64, wave = t / 64;

    // the precomputed pair list of (i, a), any length, staged chunk by chunk (padded with state 0
    // to whole turns of the ring, whose factors are 0)
    const int uEnd = __builtin_amdgcn_readfirstlane((int)P.pairCnt[ia]);
    const int uEndR = (uEnd + RING - 1) / RING * RING;
    const NOVA_GLOBAL int *pSp = (const NOVA_GLOBAL int *)P.pairSp + (size_t)ia * P.kmax;
    const NOVA_GLOBAL float *GT = (const NOVA_GLOBAL float *)P.GT;
    const NOVA_GLOBAL float *pw = (const NOVA_GLOBAL float *)P.pairW + (size_t)ia * P.kmax;
    const NOVA_GLOBAL float *Oa = (const NOVA_GLOBAL float *)P.O + (size_t)a * n * z;
    const unsigned jw = tile * kMTile + wave * NB * 32;  // the wave's first alpha
    // this lane's first column; GT rows are padded with zeros to whole tiles, and columns past r
    // only feed C columns the epilogue ignores
    const unsigned jl = jw + lane % 32;
    const size_t rS = P.rS;
    const __amdgpu_buffer_rsrc_t rG = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float *>((const float *)P.GT), 0, -1, kRawBufferWord3);

    for (unsigned o0 = 0; o0 < z; o0 += 32) {
        const int zc = min(32, (int)(z - o0));
        floatx16_t acc[NB];
#pragma unroll
        for (int y = 0; y < NB; y++)
#pragma unroll
            for (int v = 0; v < 16; v++)
                acc[y][v] = 0.0f;
        for (int c0 = 0; c0 < uEnd; c0 += kWChunk) {
            const int cPad = min(kWChunk, uEndR - c0);      // a multiple of RING
            __syncthreads();
            for (int u = t; u < cPad; u += NT)
                spU[u] = (c0 + u < uEnd) ? pSp[c0 + u] : 0;
            __syncthreads();
            stage_w<32, NT, kWChunk * 32>(w, spU, cPad, uEnd - c0, zc, o0, z, pw + c0, Oa, t);
            __syncthreads();
            // PS pair-steps (2 PS pairs) per trip; the B operands of a trip are requested
            // NSET - 1 trips ahead of its MFMAs, so their latency hides behind the matrix work
            float bq[NSET][PS][NB];
            auto loadB = [&](int u0, float (&dst)[PS][NB]) {
#pragma unroll
                for (int q = 0; q < PS; q++) {
                    const int uu = u0 + 2 * q + (int)(lane / 32);
                    if constexpr (BUF) {
                        // one 32-bit byte offset per row (GT is below 4 GB for these variants)
                        const unsigned o = ((unsigned)spU[uu] * (unsigned)rS + jl) * 4u;
#pragma unroll
                        for (int y = 0; y < NB; y++)
                            dst[q][y] = __uint_as_float(
                                    __builtin_amdgcn_raw_buffer_load_b32(rG, o, y * 128, 0));
                    } else {
                        const NOVA_GLOBAL float *row = GT + (size_t)spU[uu] * rS + jl;
#pragma unroll
                        for (int y = 0; y < NB; y++)
                            dst[q][y] = row[y * 32];
                    }
                }
            };
            auto mfmaTrip = [&](int u0, const float (&b)[PS][NB]) {
                float fa[PS];
#pragma unroll
                for (int q = 0; q < PS; q++)
                    fa[q] = w[(u0 + 2 * q + (int)(lane / 32)) * 32 + lane % 32];
#pragma unroll
                for (int q = 0; q < PS; q++)
#pragma unroll
                    for (int y = 0; y < NB; y++)
                        acc[y] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[q], b[q][y], acc[y], 0, 0, 0);
            };
            // a ring of NSET register sets: trip t's operands sit in set t mod NSET, requested
            // right after the MFMAs of trip t - NSET consumed that set, so NSET - 1 trips of
            // loads are in flight behind each trip's matrix work (cPad is a whole number of
            // turns). The loads carry no conditions and the last turn issues none, so the
            // compiler counts the outstanding ones exactly and nothing is left in flight after
            // the chunk. sched_barrier keeps the scheduler from moving the loads across the
            // MFMAs.
            if (jw < r) {       // a wave whose alphas all lie past r has nothing to score
#pragma unroll
                for (int s0 = 0; s0 < NSET; s0++)
                    loadB(s0 * HALF, bq[s0]);
                for (int u0 = 0; u0 < cPad - RING; u0 += RING) {
#pragma unroll
                    for (int j = 0; j < NSET; j++) {
                        __builtin_amdgcn_sched_barrier(0);
                        mfmaTrip(u0 + j * HALF, bq[j]);
                        __builtin_amdgcn_sched_barrier(0);
                        loadB(u0 + (j + NSET) * HALF, bq[j]);
                    }
                }
#pragma unroll
                for (int j = 0; j < NSET; j++)      // the last turn requests nothing more
                    mfmaTrip(cPad - RING + j * HALF, bq[j]);
            }
        }
        // per row (observation): first maximiser over the wave's alphas, then across the waves
#pragma unroll
        for (int v = 0; v < 16; v++) {
            float bv = kLow;
            int bj = INT_MAX;
#pragma unroll
            for (int y = 0; y < NB; y++) {
                const unsigned j = jw + y * 32 + lane % 32;
                if (j < r && acc[y][v] > bv) {
                    bv = acc[y][v];
                    bj = (int)j;
                }
            }
            for (int off2 = 16; off2 >= 1; off2 >>= 1) {
                const float ov = __shfl_xor(bv, off2, 32);
                const int oj = __shfl_xor(bj, off2, 32);
                if (better(ov, oj, bv, bj)) {
                    bv = ov;
                    bj = oj;
                }
            }
            if (lane % 32 == 0) {
                const int o = (v % 4) + 4 * (int)(lane / 32) + 8 * (v / 4);
                redV[wave][o] = bv;
                redJ[wave][o] = bj;
            }
        }
        __syncthreads();
        if (t < (unsigned)zc) {
            float v = redV[0][t];
            int j = redJ[0][t];
            for (int wv = 1; wv < NW; wv++)
                if (better(redV[wv][t], redJ[wv][t], v, j)) {
                    v = redV[wv][t];
                    j = redJ[wv][t];
                }
            const size_t idx = ((size_t)tile * rb * m + ia) * z + o0 + t;
            P.partVal[idx] = v;
            P.partJ[idx] = j;
        }
        __syncthreads();
    }
}

// The 32-row matrix-core scoring with the whole pair list of (i, a) staged once and every alpha
// tile of the instance swept by the same workgroup: W is built once per (belief, action) rather
// than once per tile, and each lane keeps its running first maximiser per row across the tiles,
// so the rows are reduced across lanes and waves once per sweep instead of once per tile (that
// reduction cost 7% of the n = 20,000 scorer and 20-40% at 64-128 pairs). The sweep's result is
// written as tile 0's partial, the other tiles' partials as empty. For pair lists of at most
// kWAll pairs (the tuner offers it then; longer lists use the kernel above).
constexpr int kWAll = 512;

template <int PS, int NSET, bool BUF>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void pbvi_score_mfma_all(const PbviInst *__restrict__ insts,
        const unsigned *__restrict__ off, unsigned count, unsigned iter)
{
    constexpr int NW = 8, NB = 2, NT = NW * 64;
    static_assert(NW * NB * 32 == kMTile, "a workgroup sweeps whole alpha tiles");
    static_assert(NSET >= 2, "at least one trip's operands in flight");
    constexpr int HALF = 2 * PS;              // pairs per trip
    constexpr int RING = NSET * HALF;         // pairs per turn of the ring of NSET register sets
    static_assert(kWAll % RING == 0, "whole turns of the ring");
    static_assert(NSET * PS * NB <= 63, "a turn of the ring fits the vector-memory counter");
    __shared__ int spU[kWAll];
    __shared__ float w[kWAll * 32];
    __shared__ float redV[NW][32];
    __shared__ int redJ[NW][32];

    if (blockIdx.x >= off[count])
        return;
    const unsigned b = find_inst(off, count, blockIdx.x);
    const PbviInst &P = insts[b];
    if (iter >= P.horizon)
        return;
    const unsigned n = P.n, m = P.m, z = P.z, r = P.r;
    const unsigned local = blockIdx.x - off[b];
    const unsigned rb = P.rb;
    const unsigned q = local % (rb * m);
    const unsigned ia = P.actionMajor ? (q % rb) * m + q / rb : q;   // belief * m + action
    const unsigned a = ia % m;
    const unsigned t = threadIdx.x, lane = t % 64, wave = t / 64;

    const int uEnd = __builtin_amdgcn_readfirstlane((int)P.pairCnt[ia]);
    const int uPad = (uEnd + RING - 1) / RING * RING;             // <= kWAll
    const NOVA_GLOBAL int *pSp = (const NOVA_GLOBAL int *)P.pairSp + (size_t)ia * P.kmax;
    const NOVA_GLOBAL float *GT = (const NOVA_GLOBAL float *)P.GT;
    const NOVA_GLOBAL float *pw = (const NOVA_GLOBAL float *)P.pairW + (size_t)ia * P.kmax;
    const NOVA_GLOBAL float *Oa = (const NOVA_GLOBAL float *)P.O + (size_t)a * n * z;
    const size_t rS = P.rS;
    const __amdgpu_buffer_rsrc_t rG = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float *>((const float *)P.GT), 0, -1, kRawBufferWord3);
    for (int u = t; u < uPad; u += NT)
        spU[u] = (u < uEnd) ? pSp[u] : 0;

    for (unsigned o0 = 0; o0 < z; o0 += 32) {
        const int zc = min(32, (int)(z - o0));
        __syncthreads();
        stage_w<32, NT, kWAll * 32>(w, spU, uPad, uEnd, zc, o0, z, pw, Oa, t);
        __syncthreads();
        // per lane and row: the first maximiser over the lane's alphas of every tile so far
        // (tiles in order, within a tile the lane's alphas in order, so strict > keeps the
        // first); the rows are reduced across lanes and waves once, after the last tile
        float runV[16];
        int runJ[16];
#pragma unroll
        for (int v = 0; v < 16; v++) {
            runV[v] = kLow;
            runJ[v] = INT_MAX;
        }
        for (unsigned tile = 0; tile < P.tiles; tile++) {
            const unsigned jw = tile * kMTile + wave * NB * 32;
            const unsigned jl = jw + lane % 32;
            floatx16_t acc[NB];
#pragma unroll
            for (int y = 0; y < NB; y++)
#pragma unroll
                for (int v = 0; v < 16; v++)
                    acc[y][v] = 0.0f;
            if (uPad > 0 && jw < r) {   // waves whose alphas all lie past r skip
                float bq[NSET][PS][NB];
                auto loadB = [&](int u0, float (&dst)[PS][NB]) {
#pragma unroll
                    for (int qq = 0; qq < PS; qq++) {
                        const int uu = u0 + 2 * qq + (int)(lane / 32);
                        if constexpr (BUF) {
                            // one 32-bit byte offset per row (GT is below 4 GB here)
                            const unsigned o = ((unsigned)spU[uu] * (unsigned)rS + jl) * 4u;
#pragma unroll
                            for (int y = 0; y < NB; y++)
                                dst[qq][y] = __uint_as_float(
                                        __builtin_amdgcn_raw_buffer_load_b32(rG, o, y * 128, 0));
                        } else {
                            const NOVA_GLOBAL float *row = GT + (size_t)spU[uu] * rS + jl;
#pragma unroll
                            for (int y = 0; y < NB; y++)
                                dst[qq][y] = row[y * 32];
                        }
                    }
                };
                auto mfmaTrip = [&](int u0, const float (&bb)[PS][NB]) {
                    float fa[PS];
#pragma unroll
                    for (int qq = 0; qq < PS; qq++)
                        fa[qq] = w[(u0 + 2 * qq + (int)(lane / 32)) * 32 + lane % 32];
#pragma unroll
                    for (int qq = 0; qq < PS; qq++)
#pragma unroll
                        for (int y = 0; y < NB; y++)
                            acc[y] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[qq], bb[qq][y],
                                                                          acc[y], 0, 0, 0);
                };
                // the ring of register sets of the per-tile kernel above
#pragma unroll
                for (int s0 = 0; s0 < NSET; s0++)
                    loadB(s0 * HALF, bq[s0]);
                for (int u0 = 0; u0 < uPad - RING; u0 += RING) {
#pragma unroll
                    for (int j = 0; j < NSET; j++) {
                        __builtin_amdgcn_sched_barrier(0);
                        mfmaTrip(u0 + j * HALF, bq[j]);
                        __builtin_amdgcn_sched_barrier(0);
                        loadB(u0 + (j + NSET) * HALF, bq[j]);
                    }
                }
#pragma unroll
                for (int j = 0; j < NSET; j++)      // the last turn requests nothing more
                    mfmaTrip(uPad - RING + j * HALF, bq[j]);
            }
#pragma unroll
            for (int v = 0; v < 16; v++)
#pragma unroll
                for (int y = 0; y < NB; y++) {
                    const unsigned j = jw + y * 32 + lane % 32;
                    if (j < r && acc[y][v] > runV[v]) {
                        runV[v] = acc[y][v];
                        runJ[v] = (int)j;
                    }
                }
        }
        // per row (observation): first maximiser across the lanes, then across the waves; the
        // whole sweep's result goes to tile 0's partial and the other tiles' partials get none
#pragma unroll
        for (int v = 0; v < 16; v++) {
            float bv = runV[v];
            int bj = runJ[v];
            for (int off2 = 16; off2 >= 1; off2 >>= 1) {
                const float ov = __shfl_xor(bv, off2, 32);
                const int oj = __shfl_xor(bj, off2, 32);
                if (better(ov, oj, bv, bj)) {
                    bv = ov;
                    bj = oj;
                }
            }
            if (lane % 32 == 0) {
                const int o = (v % 4) + 4 * (int)(lane / 32) + 8 * (v / 4);
                redV[wave][o] = bv;
                redJ[wave][o] = bj;
            }
        }
        __syncthreads();
        if (t < (unsigned)zc) {
            float v = redV[0][t];
            int j = redJ[0][t];
            for (int wv = 1; wv < NW; wv++)
                if (better(redV[wv][t], redJ[wv][t], v, j)) {
                    v = redV[wv][t];
                    j = redJ[wv][t];
                }
            for (unsigned tile = 0; tile < P.tiles; tile++) {
                const size_t idx = ((size_t)tile * rb * m + ia) * z + o0 + t;
                P.partVal[idx] = tile == 0 ? v : kLow;
                P.partJ[idx] = tile == 0 ? j : INT_MAX;
            }
        }
        __syncthreads();
    }
}

// The same scoring for instances with at most 16 observations, on v_mfma_f32_16x16x4_f32:
// rows are 16 observations (the 32-row kernel above would multiply half its rows by zero),
// columns 16 alphas per block, the reduction over the pairs four at a time. Each of the 8 waves
// owns 64 alphas of the kMTile tile as four 16 x 16 blocks. Operand layout of the instruction:
// lane l supplies A[l % 16][l / 16] (observation, pair) and B[l / 16][l % 16] (pair, alpha),
// and holds D[4 * (l / 16) + v][l % 16] in element v of its accumulator.
constexpr int kW16Chunk = 256;       // support pairs per staged W chunk
typedef float floatx4_t __attribute__((ext_vector_type(4)));

template <int PS, int NSET, bool BUF>
__global__ __launch_bounds__(512) void pbvi_score_mfma16(const PbviInst *__restrict__ insts,
        const unsigned *__restrict__ off, unsigned count, unsigned iter)
{
    constexpr int NW = 8, NB = 4, NT = NW * 64;
    static_assert(NW * NB * 16 == kMTile, "a workgroup covers one alpha tile");
    static_assert(NSET >= 2, "at least one trip's operands in flight");
    constexpr int HALF = 4 * PS;              // pairs per trip (PS MFMAs of four pairs)
    constexpr int RING = NSET * HALF;         // pairs per turn of the ring of NSET register sets
    static_assert(kW16Chunk % RING == 0, "whole turns of the ring per chunk");
    static_assert(NSET * PS * NB <= 63, "a turn of the ring fits the vector-memory counter");
    __shared__ int spU[kW16Chunk];
    __shared__ float w[kW16Chunk * 16];       // [u][o] for the current chunk, 16 observations
    __shared__ float redV[NW][16];
    __shared__ int redJ[NW][16];

    if (blockIdx.x >= off[count])
        return;   // a grid larger than the plan's workgroups does nothing
    const unsigned b = find_inst(off, count, blockIdx.x);
    const PbviInst &P = insts[b];
    if (iter >= P.horizon)
        return;
    const unsigned n = P.n, m = P.m, z = P.z, r = P.r;
    const unsigned local = blockIdx.x - off[b];
    const unsigned rb = P.rb;
    const unsigned q = local % (rb * m);
    const unsigned ia = P.actionMajor ? (q % rb) * m + q / rb : q;   // belief * m + action
    const unsigned tile = local / (rb * m);
    const unsigned a = ia % m;
    const unsigned t = threadIdx.x, lane = t % 64, wave = t / 64;

    const int uEnd = __builtin_amdgcn_readfirstlane((int)P.pairCnt[ia]);
    const int uEndP = (uEnd + RING - 1) / RING * RING;
    const NOVA_GLOBAL int *pSp = (const NOVA_GLOBAL int *)P.pairSp + (size_t)ia * P.kmax;
    const NOVA_GLOBAL float *GT = (const NOVA_GLOBAL float *)P.GT;
    const NOVA_GLOBAL float *pw = (const NOVA_GLOBAL float *)P.pairW + (size_t)ia * P.kmax;
    const NOVA_GLOBAL float *Oa = (const NOVA_GLOBAL float *)P.O + (size_t)a * n * z;
    const unsigned jw = tile * kMTile + wave * NB * 16;
    const unsigned jl = jw + lane % 16;
    const size_t rS = P.rS;
    const __amdgpu_buffer_rsrc_t rG = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float *>((const float *)P.GT), 0, -1, kRawBufferWord3);

    for (unsigned o0 = 0; o0 < z; o0 += 16) {
        const int zc = min(16, (int)(z - o0));
        floatx4_t acc[NB];
#pragma unroll
        for (int y = 0; y < NB; y++)
#pragma unroll
            for (int v = 0; v < 4; v++)
                acc[y][v] = 0.0f;
        for (int c0 = 0; c0 < uEnd; c0 += kW16Chunk) {
            const int cPad = min(kW16Chunk, uEndP - c0);     // a multiple of RING
            __syncthreads();
            for (int u = t; u < cPad; u += NT)
                spU[u] = (c0 + u < uEnd) ? pSp[c0 + u] : 0;
            __syncthreads();
            stage_w<16, NT, kW16Chunk * 16>(w, spU, cPad, uEnd - c0, zc, o0, z, pw + c0, Oa, t);
            __syncthreads();
            float bq[NSET][PS][NB];
            auto loadB = [&](int u0, float (&dst)[PS][NB]) {
#pragma unroll
                for (int q = 0; q < PS; q++) {
                    const int uu = u0 + 4 * q + (int)(lane / 16);
                    if constexpr (BUF) {
                        // one 32-bit byte offset per row (GT is below 4 GB for these variants)
                        const unsigned o = ((unsigned)spU[uu] * (unsigned)rS + jl) * 4u;
#pragma unroll
                        for (int y = 0; y < NB; y++)
                            dst[q][y] = __uint_as_float(
                                    __builtin_amdgcn_raw_buffer_load_b32(rG, o, y * 64, 0));
                    } else {
                        const NOVA_GLOBAL float *row = GT + (size_t)spU[uu] * rS + jl;
#pragma unroll
                        for (int y = 0; y < NB; y++)
                            dst[q][y] = row[y * 16];
                    }
                }
            };
            auto mfmaTrip = [&](int u0, const float (&bq)[PS][NB]) {
                float fa[PS];
#pragma unroll
                for (int q = 0; q < PS; q++)
                    fa[q] = w[(u0 + 4 * q + (int)(lane / 16)) * 16 + lane % 16];
#pragma unroll
                for (int q = 0; q < PS; q++)
#pragma unroll
                    for (int y = 0; y < NB; y++)
                        acc[y] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[q], bq[q][y], acc[y],
                                                                      0, 0, 0);
            };
            // the ring of register sets of the 32-row kernel above
            if (jw < r) {       // a wave whose alphas all lie past r has nothing to score
#pragma unroll
                for (int s0 = 0; s0 < NSET; s0++)
                    loadB(s0 * HALF, bq[s0]);
                for (int u0 = 0; u0 < cPad - RING; u0 += RING) {
#pragma unroll
                    for (int j = 0; j < NSET; j++) {
                        __builtin_amdgcn_sched_barrier(0);
                        mfmaTrip(u0 + j * HALF, bq[j]);
                        __builtin_amdgcn_sched_barrier(0);
                        loadB(u0 + (j + NSET) * HALF, bq[j]);
                    }
                }
#pragma unroll
                for (int j = 0; j < NSET; j++)      // the last turn requests nothing more
                    mfmaTrip(cPad - RING + j * HALF, bq[j]);
            }
        }
        // per row (observation): first maximiser over the wave's alphas, then across the waves
#pragma unroll
        for (int v = 0; v < 4; v++) {
            float bv = kLow;
            int bj = INT_MAX;
#pragma unroll
            for (int y = 0; y < NB; y++) {
                const unsigned j = jw + y * 16 + lane % 16;
                if (j < r && acc[y][v] > bv) {
                    bv = acc[y][v];
                    bj = (int)j;
                }
            }
            for (int off2 = 8; off2 >= 1; off2 >>= 1) {
                const float ov = __shfl_xor(bv, off2, 16);
                const int oj = __shfl_xor(bj, off2, 16);
                if (better(ov, oj, bv, bj)) {
                    bv = ov;
                    bj = oj;
                }
            }
            if (lane % 16 == 0) {
                const int o = 4 * (int)(lane / 16) + v;
                redV[wave][o] = bv;
                redJ[wave][o] = bj;
            }
        }
        __syncthreads();
        if (t < (unsigned)zc) {
            float v = redV[0][t];
            int j = redJ[0][t];
            for (int wv = 1; wv < NW; wv++)
                if (better(redV[wv][t], redJ[wv][t], v, j)) {
                    v = redV[wv][t];
                    j = redJ[wv][t];
                }
            const size_t idx = ((size_t)tile * rb * m + ia) * z + o0 + t;
            P.partVal[idx] = v;
            P.partJ[idx] = j;
        }
        __syncthreads();
    }
}

// The 16-row scoring with every alpha tile swept by one workgroup, as pbvi_score_mfma_all does
// with 32 rows: the pair list (at most kWAll pairs) staged once per chunk of 16 observations,
// the operands in the ring of register sets above, each lane's running first maximiser of its
// four rows kept across the tiles and the rows reduced once per sweep, the result written as
// tile 0's partial (variant 4096).
template <int PS, int NSET, bool BUF>
__global__ __launch_bounds__(512) void pbvi_score_mfma16_all(const PbviInst *__restrict__ insts,
        const unsigned *__restrict__ off, unsigned count, unsigned iter)
{
    constexpr int NW = 8, NB = 4, NT = NW * 64;
    static_assert(NW * NB * 16 == kMTile, "a workgroup sweeps whole alpha tiles");
    static_assert(NSET >= 2, "at least one trip's operands in flight");
    constexpr int HALF = 4 * PS;              // pairs per trip (PS MFMAs of four pairs)
    constexpr int RING = NSET * HALF;         // pairs per turn of the ring of NSET register sets
    static_assert(kWAll % RING == 0, "whole turns of the ring");
    static_assert(NSET * PS * NB <= 63, "a turn of the ring fits the vector-memory counter");
    __shared__ int spU[kWAll];
    __shared__ float w[kWAll * 16];           // [u][o] for 16 observations
    __shared__ float redV[NW][16];
    __shared__ int redJ[NW][16];

    if (blockIdx.x >= off[count])
        return;
    const unsigned b = find_inst(off, count, blockIdx.x);
    const PbviInst &P = insts[b];
    if (iter >= P.horizon)
        return;
    const unsigned n = P.n, m = P.m, z = P.z, r = P.r;
    const unsigned local = blockIdx.x - off[b];
    const unsigned rb = P.rb;
    const unsigned q = local % (rb * m);
    const unsigned ia = P.actionMajor ? (q % rb) * m + q / rb : q;   // belief * m + action
    const unsigned a = ia % m;
    const unsigned t = threadIdx.x, lane = t % 64, wave = t / 64;

    const int uEnd = __builtin_amdgcn_readfirstlane((int)P.pairCnt[ia]);
    const int uPad = (uEnd + RING - 1) / RING * RING;             // <= kWAll
    const NOVA_GLOBAL int *pSp = (const NOVA_GLOBAL int *)P.pairSp + (size_t)ia * P.kmax;
    const NOVA_GLOBAL float *pw = (const NOVA_GLOBAL float *)P.pairW + (size_t)ia * P.kmax;
    const NOVA_GLOBAL float *Oa = (const NOVA_GLOBAL float *)P.O + (size_t)a * n * z;
    const NOVA_GLOBAL float *GT = (const NOVA_GLOBAL float *)P.GT;
    const size_t rS = P.rS;
    const __amdgpu_buffer_rsrc_t rG = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float *>((const float *)P.GT), 0, -1, kRawBufferWord3);
    for (int u = t; u < uPad; u += NT)
        spU[u] = (u < uEnd) ? pSp[u] : 0;

    for (unsigned o0 = 0; o0 < z; o0 += 16) {
        const int zc = min(16, (int)(z - o0));
        __syncthreads();
        stage_w<16, NT, kWAll * 16>(w, spU, uPad, uEnd, zc, o0, z, pw, Oa, t);
        __syncthreads();
        float runV[4];
        int runJ[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            runV[v] = kLow;
            runJ[v] = INT_MAX;
        }
        for (unsigned tile = 0; tile < P.tiles; tile++) {
            const unsigned jw = tile * kMTile + wave * NB * 16;
            const unsigned jl = jw + lane % 16;
            floatx4_t acc[NB];
#pragma unroll
            for (int y = 0; y < NB; y++)
#pragma unroll
                for (int v = 0; v < 4; v++)
                    acc[y][v] = 0.0f;
            if (uPad > 0 && jw < r) {   // waves whose alphas all lie past r skip
                float bq[NSET][PS][NB];
                auto loadB = [&](int u0, float (&dst)[PS][NB]) {
#pragma unroll
                    for (int qq = 0; qq < PS; qq++) {
                        const int uu = u0 + 4 * qq + (int)(lane / 16);
                        if constexpr (BUF) {
                            // one 32-bit byte offset per row (GT is below 4 GB here)
                            const unsigned o = ((unsigned)spU[uu] * (unsigned)rS + jl) * 4u;
#pragma unroll
                            for (int y = 0; y < NB; y++)
                                dst[qq][y] = __uint_as_float(
                                        __builtin_amdgcn_raw_buffer_load_b32(rG, o, y * 64, 0));
                        } else {
                            const NOVA_GLOBAL float *row = GT + (size_t)spU[uu] * rS + jl;
#pragma unroll
                            for (int y = 0; y < NB; y++)
                                dst[qq][y] = row[y * 16];
                        }
                    }
                };
                auto mfmaTrip = [&](int u0, const float (&bb)[PS][NB]) {
                    float fa[PS];
#pragma unroll
                    for (int qq = 0; qq < PS; qq++)
                        fa[qq] = w[(u0 + 4 * qq + (int)(lane / 16)) * 16 + lane % 16];
#pragma unroll
                    for (int qq = 0; qq < PS; qq++)
#pragma unroll
                        for (int y = 0; y < NB; y++)
                            acc[y] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[qq], bb[qq][y],
                                                                          acc[y], 0, 0, 0);
                };
#pragma unroll
                for (int s0 = 0; s0 < NSET; s0++)
                    loadB(s0 * HALF, bq[s0]);
                for (int u0 = 0; u0 < uPad - RING; u0 += RING) {
#pragma unroll
                    for (int j = 0; j < NSET; j++) {
                        __builtin_amdgcn_sched_barrier(0);
                        mfmaTrip(u0 + j * HALF, bq[j]);
                        __builtin_amdgcn_sched_barrier(0);
                        loadB(u0 + (j + NSET) * HALF, bq[j]);
                    }
                }
#pragma unroll
                for (int j = 0; j < NSET; j++)      // the last turn requests nothing more
                    mfmaTrip(uPad - RING + j * HALF, bq[j]);
            }
            // lane's alphas in increasing order across the tiles: strict > keeps the first
#pragma unroll
            for (int v = 0; v < 4; v++)
#pragma unroll
                for (int y = 0; y < NB; y++) {
                    const unsigned j = jw + y * 16 + lane % 16;
                    if (j < r && acc[y][v] > runV[v]) {
                        runV[v] = acc[y][v];
                        runJ[v] = (int)j;
                    }
                }
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            float bv = runV[v];
            int bj = runJ[v];
            for (int off2 = 8; off2 >= 1; off2 >>= 1) {
                const float ov = __shfl_xor(bv, off2, 16);
                const int oj = __shfl_xor(bj, off2, 16);
                if (better(ov, oj, bv, bj)) {
                    bv = ov;
                    bj = oj;
                }
            }
            if (lane % 16 == 0) {
                const int o = 4 * (int)(lane / 16) + v;
                redV[wave][o] = bv;
                redJ[wave][o] = bj;
            }
        }
        __syncthreads();
        if (t < (unsigned)zc) {
            float v = redV[0][t];
            int j = redJ[0][t];
            for (int wv = 1; wv < NW; wv++)
                if (better(redV[wv][t], redJ[wv][t], v, j)) {
                    v = redV[wv][t];
                    j = redJ[wv][t];
                }
            for (unsigned tile = 0; tile < P.tiles; tile++) {
                const size_t idx = ((size_t)tile * rb * m + ia) * z + o0 + t;
                P.partVal[idx] = tile == 0 ? v : kLow;
                P.partJ[idx] = tile == 0 ? j : INT_MAX;
            }
        }
        __syncthreads();
    }
}

// -------------------------------------------------------------------------- 3. action

// EXACT: each action's alpha is formed on the belief's support in the reference's rounding order
// and dotted with b, as nova chooses. Fast: the action value is b . R[:,a] plus the observations'
// best scores (each score is already the b-weighted term of that observation).
// Exact mode's inner sum of one (support state, observation) pair, sum_l O*T * Gamma[j, S] in
// the reference's order up to the first terminator, with the NS successors and factors
// requested together (the terminator's and later entries read but not summed, index 0 in
// their place).
constexpr int kActKC = 32;           // exact action: support entries per chunk of inner sums

template <int NS>
__device__ __forceinline__ float exact_inner_sum(const int *sRow, const float *otRow,
                                                 const float *Gj)
{
    const NOVA_GLOBAL int *gs = (const NOVA_GLOBAL int *)sRow;
    const NOVA_GLOBAL float *go = (const NOVA_GLOBAL float *)otRow;
    const NOVA_GLOBAL float *gg = (const NOVA_GLOBAL float *)Gj;
    int sp[NS];
    float ov[NS], gv[NS];
#pragma unroll
    for (int l = 0; l < NS; l++) {
        sp[l] = gs[l];
        ov[l] = go[l];
    }
#pragma unroll
    for (int l = 0; l < NS; l++)
        gv[l] = gg[max(sp[l], 0)];
    float vt = 0.0f;
    bool stop = false;
#pragma unroll
    for (int l = 0; l < NS; l++) {
        stop = stop || sp[l] < 0;
        vt = stop ? vt : __fadd_rn(vt, __fmul_rn(ov[l], gv[l]));
    }
    return vt;
}

template <bool EXACT>
__global__ __launch_bounds__(kPB) void pbvi_action(const PbviInst *__restrict__ insts,
        const unsigned *__restrict__ off, unsigned count, unsigned iter, unsigned flip)
{
    __shared__ int js[kZC * 8];
    __shared__ float bestO[kZC * 8];
    __shared__ float vtL[EXACT ? kActKC * kZC * 8 : 1];   // exact: inner sums of a chunk
    __shared__ float redA[kPB / 64];                       // fast: per-wave partial sums
    __shared__ int kEndS;
    // alphaK and bS: one float per support entry each (exact mode; fast mode uses alphaK[0]),
    // dynamic LDS of 2 * max(kKMax, the longest exact support) floats (plan action_lds)
    extern __shared__ float actLds[];
    if (blockIdx.x >= off[count])
        return;   // a grid larger than the plan's workgroups does nothing
    const unsigned b = find_inst(off, count, blockIdx.x);
    const PbviInst &P = insts[b];
    if (iter >= P.horizon)
        return;
    const unsigned n = P.n, ns = P.ns, m = P.m, z = P.z, r = P.r, rz = P.rz;
    const float gamma = P.gamma;
    const float *__restrict__ G = src_gamma(P, iter, flip);
    const unsigned i = blockIdx.x - off[b], t = threadIdx.x;
    const unsigned kCap = EXACT ? max(rz, (unsigned)kKMax) : 1u;
    float *alphaK = actLds;
    float *bS = actLds + kCap;
    if (t == 0)
        kEndS = (int)rz;
    __syncthreads();
    for (unsigned k = t; k < rz; k += kPB) {
        if (EXACT)
            bS[k] = P.B[(size_t)i * rz + k];   // rz <= kActMax in exact mode
        if (P.Z[(size_t)i * rz + k] < 0)
            atomicMin(&kEndS, (int)k);
    }
    __syncthreads();
    const int kEnd = kEndS;
    float bestValue = kLow;
    int bestA = -1;
    for (unsigned a = 0; a < m; a++) {
        const size_t ia = (size_t)i * m + a;
        for (unsigned o0 = 0; o0 < z; o0 += kZC * 8) {
            // merge the alpha tiles of each observation: earlier tile wins ties
            const unsigned oc = min((unsigned)(kZC * 8), z - o0);
            for (unsigned oo = t; oo < oc; oo += kPB) {
                float v = kLow;
                int j = INT_MAX;
                for (unsigned tl = 0; tl < P.tiles; tl++) {
                    const size_t idx = ((size_t)tl * P.rb * m + ia) * z + o0 + oo;
                    if (P.partVal[idx] > v) {
                        v = P.partVal[idx];
                        j = P.partJ[idx];
                    }
                }
                if (j == INT_MAX)
                    j = 0;   // no alpha above -1e10: the reference's index is undefined
                js[oo] = j;
                bestO[oo] = v;
                P.jstar[ia * z + o0 + oo] = j;
            }
            __syncthreads();
            if (!EXACT) {
                // running action value b.R_a + sum_o best score: every lane's share, then a
                // reduction in a fixed order (the same for every run)
                float part = 0.0f;
                if (o0 == 0)
                    for (int k = (int)t; k < kEnd; k += kPB)
                        part += P.B[(size_t)i * rz + k] *
                                P.R[(size_t)P.Z[(size_t)i * rz + k] * m + a];
                for (unsigned oo = t; oo < oc; oo += kPB)
                    part += bestO[oo];
                for (int off2 = 32; off2 >= 1; off2 >>= 1)
                    part += __shfl_xor(part, off2);
                if (t % 64 == 0)
                    redA[t / 64] = part;
                __syncthreads();
                if (t == 0) {
                    float v = (o0 == 0) ? 0.0f : alphaK[0];
                    for (int w = 0; w < kPB / 64; w++)
                        v += redA[w];
                    alphaK[0] = v;
                }
                __syncthreads();
                continue;
            }
            // alpha_a on the support, observation-major accumulation as the reference: per
            // chunk of kActKC entries, the inner sums vt of every (entry, observation) pair in
            // parallel (into vtL), then each entry adds its terms gamma * vt in observation order
            for (int k0 = 0; k0 < kEnd; k0 += kActKC) {
                const int kc = min(kActKC, kEnd - k0);
                for (unsigned x = t; x < (unsigned)kc * oc; x += kPB) {
                    const unsigned kk = x / oc, oo = x % oc;
                    const int s = P.Z[(size_t)i * rz + k0 + kk];
                    const size_t row = ((size_t)s * m + a) * ns;
                    const float *Gj = G + (size_t)js[oo] * n;
                    const float *ot = P.OT + ((size_t)(a * z + o0 + oo) * n + s) * ns;
                    float vt;
                    if (ns == 8)
                        vt = exact_inner_sum<8>(P.S + row, ot, Gj);
                    else if (ns == 4)
                        vt = exact_inner_sum<4>(P.S + row, ot, Gj);
                    else {
                        vt = 0.0f;
                        for (unsigned l = 0; l < ns; l++) {
                            const int sp = P.S[row + l];
                            if (sp < 0)
                                break;
                            vt = __fadd_rn(vt, __fmul_rn(ot[l], Gj[sp]));
                        }
                    }
                    vtL[kk * (kZC * 8) + oo] = vt;
                }
                __syncthreads();
                for (int kk = t; kk < kc; kk += kPB) {
                    const int k = k0 + kk;
                    const int s = P.Z[(size_t)i * rz + k];
                    float al = (o0 == 0) ? P.R[(size_t)s * m + a] : alphaK[k];
                    for (unsigned oo = 0; oo < oc; oo++)
                        al = __fadd_rn(al, __fmul_rn(vtL[kk * (kZC * 8) + oo], gamma));
                    alphaK[k] = al;
                }
                __syncthreads();
            }
        }
        if (t == 0) {
            float v = 0.0f;
            if (EXACT) {
                for (int k = 0; k < kEnd; k++)
                    v = __fadd_rn(v, __fmul_rn(alphaK[k], bS[k]));
            } else {
                v = alphaK[0];
            }
            P.aval[ia] = v;
            if (v > bestValue) {
                bestValue = v;
                bestA = (int)a;
            }
        }
        __syncthreads();
    }
    if (t == 0)
        P.act[i] = bestA;
    if (P.recJ != nullptr) {
        // the argmax checkpoint of this iteration: the chosen action's alpha indices
        __shared__ int chosenA;
        if (t == 0) {
            chosenA = bestA;
            P.recA[(size_t)iter * r + i] = bestA;
        }
        __syncthreads();
        const int ca = chosenA;
        for (unsigned o = t; o < z; o += kPB)
            P.recJ[((size_t)iter * r + i) * z + o] =
                (ca >= 0) ? P.jstar[((size_t)i * m + ca) * z + o] : -1;
    }
}

// --------------------------------------------------------------------- 4. materialize

// Threads of the materialisation workgroup: 16 waves per belief keep enough loads in flight
// while one or two workgroups per CU hold their n-float row in LDS.
constexpr int kMB = 1024;
constexpr int kMatStates = 4;        // exact materialization: states per lane requested together

// Exact materialization of one observation's term for the lane's states t, t + kMB, ...,
// kMatStates of them at a time with all their NS successors and factors requested together
// (constant offsets from each state's row): per state the same products and sums, in the same
// order, as the general loop (entries after the first terminator take no part; its index is
// clamped to 0 for the unused read). Returns the first state it left to that loop.
template <int NS>
__device__ __forceinline__ unsigned materialize_exact_states(const int *sa, const float *ot,
        const float *__restrict__ srcv, float *out, float gamma, unsigned n, unsigned t)
{
    const NOVA_GLOBAL int *gsa = (const NOVA_GLOBAL int *)sa;
    const NOVA_GLOBAL float *got = (const NOVA_GLOBAL float *)ot;
    NOVA_GLOBAL float *gout = (NOVA_GLOBAL float *)out;
    unsigned s0 = t;
    for (; s0 + (kMatStates - 1) * kMB < n; s0 += kMatStates * kMB) {
        int spv[kMatStates][NS];
        float otv[kMatStates][NS];
#pragma unroll
        for (int q = 0; q < kMatStates; q++) {
            const size_t rw = (size_t)(s0 + q * kMB) * NS;
#pragma unroll
            for (int l = 0; l < NS; l++) {
                spv[q][l] = gsa[rw + l];
                otv[q][l] = got[rw + l];
            }
        }
#pragma unroll
        for (int q = 0; q < kMatStates; q++) {
            float vt = 0.0f;
            bool stop = false;
#pragma unroll
            for (int l = 0; l < NS; l++) {
                stop = stop || spv[q][l] < 0;
                const float g = srcv[max(spv[q][l], 0)];
                vt = stop ? vt : __fadd_rn(vt, __fmul_rn(otv[q][l], g));
            }
            const unsigned s = s0 + q * kMB;
            gout[s] = __fadd_rn(gout[s], __fmul_rn(vt, gamma));
        }
    }
    return s0;
}

// Fast materialization of the lane's states t, t + kMB, ..., kMatStates of them at a time with
// all their NS successors and factors requested together: alpha(s) = R + gamma * sum_l T h(S),
// each sum the same FMA chain, in the same order, as the general loop (entries after the first
// terminator take no part; its index is clamped to 0 for the unused read). Returns the first
// state it left to that loop.
template <int NS>
__device__ __forceinline__ unsigned materialize_fast_states(const NOVA_GLOBAL int *Sa,
        const NOVA_GLOBAL float *Ta, const NOVA_GLOBAL float *Ra, const float *row,
        NOVA_GLOBAL float *out, float gamma, unsigned n, unsigned t)
{
    unsigned s0 = t;
    for (; s0 + (kMatStates - 1) * kMB < n; s0 += kMatStates * kMB) {
        int spv[kMatStates][NS];
        float tv[kMatStates][NS];
#pragma unroll
        for (int q = 0; q < kMatStates; q++) {
            const size_t rw = (size_t)(s0 + q * kMB) * NS;
#pragma unroll
            for (int l = 0; l < NS; l++) {
                spv[q][l] = Sa[rw + l];
                tv[q][l] = Ta[rw + l];
            }
        }
#pragma unroll
        for (int q = 0; q < kMatStates; q++) {
            float acc = 0.0f;
            bool stop = false;
#pragma unroll
            for (int l = 0; l < NS; l++) {
                stop = stop || spv[q][l] < 0;
                const float h = row[max(spv[q][l], 0)];
                acc = stop ? acc : __builtin_fmaf(tv[q][l], h, acc);
            }
            const unsigned s = s0 + q * kMB;
            out[s] = __builtin_fmaf(gamma, acc, Ra[s]);
        }
    }
    return s0;
}

template <bool EXACT>
__global__ __launch_bounds__(kMB) void pbvi_materialize(const PbviInst *__restrict__ insts,
        const unsigned *__restrict__ off, unsigned count, unsigned iter, unsigned flip)
{
    extern __shared__ float row[];   // n floats when the instance stages rows in LDS
    if (blockIdx.x >= off[count])
        return;   // a grid larger than the plan's workgroups does nothing
    const unsigned b = find_inst(off, count, blockIdx.x);
    const PbviInst &P = insts[b];
    if (iter >= P.horizon)
        return;
    const unsigned n = P.n, ns = P.ns, m = P.m, z = P.z;
    const float gamma = P.gamma;
    const int useLds = P.useLds;
    const unsigned i = blockIdx.x - off[b], t = threadIdx.x;
    const int a = P.act[i];
    if (a < 0)
        return;   // no action beat -1e10: Gamma'[i] and pi[i] keep their values
    const float *__restrict__ G = src_gamma(P, iter, flip);
    float *out = dst_gamma(P, iter, flip) + (size_t)i * n;
    const int *js = P.jstar + ((size_t)i * m + a) * z;
    if (EXACT || !useLds) {
        // alpha(s) = R[s,a] + sum_o (gamma * sum_l (O*T) * Gamma[j*_o, sp]), o ascending
        for (unsigned s = t; s < n; s += kMB)
            out[s] = P.R[(size_t)s * m + a];
        for (unsigned o = 0; o < z; o++) {
            const float *Gj = G + (size_t)js[o] * n;
            if (useLds) {
                __syncthreads();
                for (unsigned s = t; s < n; s += kMB)
                    row[s] = Gj[s];
                __syncthreads();
            }
            const float *srcv = useLds ? row : Gj;
            const float *ot = P.OT + ((size_t)(a * z + o) * n) * ns;
            const int *sa = P.SA + (size_t)a * n * ns;    // the action-major copy of S
            unsigned s0 = t;
            if (useLds && ns == 8)     // row passed as it is, so its reads stay LDS reads
                s0 = materialize_exact_states<8>(sa, ot, row, out, gamma, n, t);
            else if (useLds && ns == 4)
                s0 = materialize_exact_states<4>(sa, ot, row, out, gamma, n, t);
            for (unsigned s = s0; s < n; s += kMB) {
                const size_t rw = (size_t)s * ns;
                float vt = 0.0f;
                // successors and factors four at a time (requested together), summed in order
                // up to the first terminator
                bool stop = false;
                for (unsigned l0 = 0; l0 < ns && !stop; l0 += 4) {
                    int spv[4];
                    float otv[4];
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const unsigned l = min(l0 + k, ns - 1);
                        spv[k] = l0 + k < ns ? sa[rw + l] : -1;
                        otv[k] = ot[rw + l];
                    }
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        if (!stop && spv[k] < 0)
                            stop = true;
                        if (!stop)
                            vt = __fadd_rn(vt, __fmul_rn(otv[k], srcv[spv[k]]));
                    }
                }
                out[s] = __fadd_rn(out[s], __fmul_rn(vt, gamma));
            }
        }
    } else {
        // h(sp) = sum_o O[a,sp,o] * Gamma[j*_o, sp] in LDS, then R + gamma*sum_l T*h(S), all
        // reads contiguous over states through the action-major copies
        const NOVA_GLOBAL float *Oa = (const NOVA_GLOBAL float *)P.OTr + (size_t)a * z * n;
        const NOVA_GLOBAL float *Gg = (const NOVA_GLOBAL float *)G;
        for (unsigned sp = t; sp < n; sp += kMB) {
            float h = 0.0f;
#pragma unroll 8
            for (unsigned o = 0; o < z; o++)
                h = __builtin_fmaf(Oa[(size_t)o * n + sp], Gg[(size_t)js[o] * n + sp], h);
            row[sp] = h;
        }
        __syncthreads();
        const NOVA_GLOBAL int *Sa = (const NOVA_GLOBAL int *)P.SA + (size_t)a * n * ns;
        const NOVA_GLOBAL float *Ta = (const NOVA_GLOBAL float *)P.TA + (size_t)a * n * ns;
        const NOVA_GLOBAL float *Ra = (const NOVA_GLOBAL float *)P.RT + (size_t)a * n;
        unsigned s0 = t;
        if (ns == 8)
            s0 = materialize_fast_states<8>(Sa, Ta, Ra, row, (NOVA_GLOBAL float *)out, gamma, n, t);
        else if (ns == 4)
            s0 = materialize_fast_states<4>(Sa, Ta, Ra, row, (NOVA_GLOBAL float *)out, gamma, n, t);
        for (unsigned s = s0; s < n; s += kMB) {
            const size_t rw = (size_t)s * ns;
            float acc = 0.0f;
            for (unsigned l = 0; l < ns; l++) {
                const int sp = Sa[rw + l];
                if (sp < 0)
                    break;
                acc = __builtin_fmaf(Ta[rw + l], row[sp], acc);
            }
            ((NOVA_GLOBAL float *)out)[s] = __builtin_fmaf(gamma, acc, Ra[s]);
        }
    }
    if (t == 0)
        P.pi[i] = (unsigned)a;
}

// ------------------------------------------------ small instances: one launch per iteration
//
// At nova's smallest domains an iteration's arithmetic takes far less time than launching the
// four kernels above. For instances within the bounds below one workgroup per belief runs the
// whole backup: the score of every (observation, alpha) pair, the first maximiser per
// observation, the action's alpha over all n states and its value on the support, and the
// winner's write-out. Every sum and product is rounded in nova's order (the exact mode's), so
// the kernel serves both modes: the exact mode's results are bit for bit those of the four
// kernels, and the fast mode's stay within float rounding of them, as that mode promises.
constexpr int kFN = 2048;          // max states: the alpha and the best alpha in LDS
constexpr int kFZ = 256;           // max observations
constexpr int kFS = 6144;          // max observations x alphas: the scores in LDS
constexpr int kFStage = 24576;     // max words of Gamma, O*T, S, R and the belief in LDS

// dst[x] = get(x) for x < count, eight values per lane requested before any is stored
template <typename T, typename F>
__device__ __forceinline__ void fused_stage_copy(T *dst, F get, unsigned count, unsigned t)
{
    for (unsigned x0 = t; x0 < count; x0 += 8 * kPB) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; u++)
            v[u] = get(min(x0 + u * kPB, count - 1));
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (x0 + u * kPB < count)
                dst[x0 + u * kPB] = v[u];
    }
}

__global__ __launch_bounds__(kPB) void pbvi_fused(const PbviInst *__restrict__ insts,
        const unsigned *__restrict__ off, unsigned count, unsigned iter, unsigned flip)
{
    __shared__ SupportLds L;
    __shared__ int scratch[kKMax + 1];
    __shared__ float sc[kFS];
    __shared__ float alpha[kFN], bestAlpha[kFN];
    __shared__ int js[kFZ], bestJs[kFZ];
    __shared__ float actVal;
    extern __shared__ float stage[];   // Gamma, O*T, S, R and belief i: fused_stage_words
    if (blockIdx.x >= off[count])
        return;   // a grid larger than the plan's workgroups does nothing
    const unsigned b = find_inst(off, count, blockIdx.x);
    const PbviInst &P = insts[b];
    if (iter >= P.horizon)
        return;
    const unsigned n = P.n, ns = P.ns, m = P.m, z = P.z, r = P.r;
    const float gamma = P.gamma;
    const unsigned i = blockIdx.x - off[b], t = threadIdx.x;
    // one round of loads stages everything the iteration reads: Gamma [r][n], O*T [m][z][n][ns],
    // S [m][n][ns] (action-major), R [n][m] and belief i's (Z, B)
    const unsigned rz = P.rz, zn = z * n * ns, sn = n * ns;
    // Gamma rows at an odd stride: the scoring lanes read one state of consecutive alphas, which
    // a stride of a multiple of the bank count would put in one LDS bank
    const unsigned gs = n | 1u;
    float *G = stage;
    float *OTall = G + (size_t)r * gs;
    int *Sall = reinterpret_cast<int *>(OTall + (size_t)m * zn);
    float *Rall = reinterpret_cast<float *>(Sall + (size_t)m * sn);
    int *Zi = reinterpret_cast<int *>(Rall + (size_t)n * m);
    float *Bi = reinterpret_cast<float *>(Zi + rz);
    {
        // eight loads per lane requested before any is stored (one instance's staging is the
        // kernel's latency at these sizes)
        const float *__restrict__ Gsrc = src_gamma(P, iter, flip);
        for (unsigned x0 = t; x0 < r * n; x0 += 8 * kPB) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; u++)
                v[u] = Gsrc[min(x0 + u * kPB, r * n - 1)];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const unsigned x = x0 + u * kPB;
                if (x < r * n)
                    G[(x / n) * gs + x % n] = v[u];
            }
        }
        fused_stage_copy(OTall, [&](unsigned x) { return P.OT[x]; }, m * zn, t);
        fused_stage_copy(Sall, [&](unsigned x) {
            const unsigned a = x / sn, e = x % sn;
            return P.S[((size_t)(e / ns) * m + a) * ns + e % ns];
        }, m * sn, t);
        fused_stage_copy(Rall, [&](unsigned x) { return P.R[x]; }, n * m, t);
        for (unsigned k = t; k < rz; k += kPB) {
            Zi[k] = P.Z[(size_t)i * rz + k];
            Bi[k] = P.B[(size_t)i * rz + k];
        }
    }
    float bestValue = kLow;            // the same in every lane
    int bestA = -1;
    for (unsigned a = 0; a < m; a++) {
        const float *ota = OTall + (size_t)a * zn;
        const int *Sa = Sall + (size_t)a * sn;
        // the support's pairs from the staged rows (belief 0 of one action: i = 0, m = 1,
        // a = 0); build_support reads its inputs only after its first barrier, which orders
        // the staging above before them
        build_support(L, 0, 0, 1, ns, rz, Sa, Zi, Bi, scratch);   // ends with a barrier
        const int kEnd = L.kEnd;
        // score of alpha j under observation o: sum_k ((sum_l (O*T) * Gamma[j, sp]) * gamma) * b_k
        for (unsigned x = t; x < z * r; x += kPB) {
            const unsigned o = x / r, j = x % r;
            const float *ot = ota + (size_t)o * n * ns;
            const float *Gj = G + (size_t)j * gs;
            float value = 0.0f;
            for (int k = 0; k < kEnd; k++) {
                float vt = 0.0f;
                for (int u = L.startK[k]; u < L.startK[k + 1]; u++)
                    vt = __fadd_rn(vt, __fmul_rn(ot[L.otU[u]], Gj[L.spU[u]]));
                value = __fadd_rn(value, __fmul_rn(__fmul_rn(vt, gamma), L.bK[k]));
            }
            sc[x] = value;
        }
        __syncthreads();
        // first maximiser per observation by strict '>' from -1e10 (none: alpha 0): 32 lanes
        // per observation scan the alphas j = l, l + 32, ... in order, then merge by (value,
        // lower j)
        for (unsigned o0 = 0; o0 < z; o0 += kPB / 32) {
            const unsigned o = o0 + t / 32, l = t % 32;
            float best = kLow;
            int bj = INT_MAX;
            if (o < z)
                for (unsigned j = l; j < r; j += 32)
                    if (sc[o * r + j] > best) {
                        best = sc[o * r + j];
                        bj = (int)j;
                    }
            for (int off2 = 16; off2 >= 1; off2 >>= 1) {
                const float ov = __shfl_xor(best, off2, 32);
                const int oj = __shfl_xor(bj, off2, 32);
                if (better(ov, oj, best, bj)) {
                    best = ov;
                    bj = oj;
                }
            }
            if (o < z && l == 0) {
                const int jj = (bj == INT_MAX) ? 0 : bj;
                js[o] = jj;
                P.jstar[((size_t)i * m + a) * z + o] = jj;
            }
        }
        __syncthreads();
        // alpha(s) = R[s,a] + sum_o (gamma * sum_l (O*T) * Gamma[j*_o, sp]), o ascending
        for (unsigned s = t; s < n; s += kPB) {
            const int *rw = Sa + (size_t)s * ns;
            float al = Rall[(size_t)s * m + a];
            for (unsigned o = 0; o < z; o++) {
                const float *Gj = G + (size_t)js[o] * gs;
                const float *ot = ota + ((size_t)o * n + s) * ns;
                float vt = 0.0f;
                for (unsigned l = 0; l < ns; l++) {
                    const int sp = rw[l];
                    if (sp < 0)
                        break;
                    vt = __fadd_rn(vt, __fmul_rn(ot[l], Gj[sp]));
                }
                al = __fadd_rn(al, __fmul_rn(vt, gamma));
            }
            alpha[s] = al;
        }
        __syncthreads();
        if (t == 0) {
            float v = 0.0f;
            for (int k = 0; k < kEnd; k++)
                v = __fadd_rn(v, __fmul_rn(alpha[L.sK[k]], L.bK[k]));
            P.aval[(size_t)i * m + a] = v;
            actVal = v;
        }
        __syncthreads();
        const float v = actVal;
        if (v > bestValue) {
            bestValue = v;
            bestA = (int)a;
            for (unsigned s = t; s < n; s += kPB)
                bestAlpha[s] = alpha[s];
            for (unsigned o = t; o < z; o += kPB)
                bestJs[o] = js[o];
        }
        __syncthreads();
    }
    if (t == 0) {
        P.act[i] = bestA;
        if (bestA >= 0)
            P.pi[i] = (unsigned)bestA;
        if (P.recJ != nullptr)
            P.recA[(size_t)iter * r + i] = bestA;
    }
    if (P.recJ != nullptr)
        for (unsigned o = t; o < z; o += kPB)
            P.recJ[((size_t)iter * r + i) * z + o] = (bestA >= 0) ? bestJs[o] : -1;
    if (bestA >= 0) {
        // no action beat -1e10: Gamma'[i] and pi[i] keep their values
        float *out = dst_gamma(P, iter, flip) + (size_t)i * n;
        for (unsigned s = t; s < n; s += kPB)
            out[s] = bestAlpha[s];
    }
}

// ------------------------------------------------------------ dense beliefs (MFMA)
//
// With dense beliefs (support = all n states) the scoring of observation o under action a is
//   C[i][j] = sum_s' U_a[i][s'] * (gamma O[a,s',o]) * Gamma[j][s'],   U_a = B T_a (r x n),
// a GEMM per (a, o) of r x n by n x r. U_a does not depend on Gamma and is built once per
// initialize. Operands live in HBM zero-padded to rpad = 256-multiple rows and kpad =
// 64-multiple columns, so the GEMM loop needs no bounds checks.
//   bf16 (mode 2): Gamma is converted to bf16 once per iteration; then, per chunk of (a, o)
//     pairs, dense_scale_pairs forms bf16(Gamma * gamma O[a,.,o]) and dense_score_bf16 scores
//     the chunk against U_a on the matrix cores (k-blocked operands copied straight into LDS).
//   fp32 (mode 3): dense_score<float> scales Gamma by gamma O[a,.,o] in registers on the way
//     into LDS.
// The epilogue reduces each row to its maximum over the tile's 256 alphas (first maximiser),
// the same partials the sparse kernel writes. The GEMM kernels and the operand conversions
// live in dense_gemm.h.

// U[a][i][s'] = sum over predecessors (s, l) of s' under a of T(s,a,l) b_i(s), with the dense
// belief row b_i staged in LDS from its sparse (Z, B) representation. Rows of kpad columns,
// rpad rows per action; the padding keeps its zeros. blocked: the k-blocked layout of
// dense_score_bf16, [kpad / 32][m * rpad][32], instead of [m][rpad][kpad].
template <typename IN>
__global__ __launch_bounds__(kPB) void dense_build_u(unsigned n, unsigned m, unsigned rz,
        unsigned rpad, unsigned kpad, int blocked, const int *__restrict__ Z,
        const float *__restrict__ B, const unsigned *__restrict__ pstart,
        const int *__restrict__ pS, const float *__restrict__ pT, IN *__restrict__ U)
{
    extern __shared__ float brow[];
    const unsigned a = blockIdx.y, i = blockIdx.x, t = threadIdx.x;
    for (unsigned s = t; s < n; s += kPB)
        brow[s] = 0.0f;
    __syncthreads();
    // the valid prefix of the support holds distinct states
    __shared__ int kEnd;
    if (t == 0)
        kEnd = (int)rz;
    __syncthreads();
    for (unsigned k = t; k < rz; k += kPB)
        if (Z[(size_t)i * rz + k] < 0)
            atomicMin(&kEnd, (int)k);
    __syncthreads();
    for (int k = t; k < kEnd; k += kPB)
        brow[Z[(size_t)i * rz + k]] = B[(size_t)i * rz + k];
    __syncthreads();
    const unsigned *ps = pstart + (size_t)a * (n + 1);
    const size_t row = (size_t)a * rpad + i, allRows = (size_t)m * rpad;
    for (unsigned sp = t; sp < n; sp += kPB) {
        float acc = 0.0f;
        for (unsigned q = ps[sp]; q < ps[sp + 1]; q++)
            acc += pT[q] * brow[pS[q]];
        const size_t at = blocked ? ((size_t)(sp / kGBlk) * allRows + row) * kGBlk + sp % kGBlk
                                  : row * kpad + sp;
        U[at] = DenseIn<IN>::cvt(acc);
    }
}

// Os[(a*z + o)*kpad + s'] = gamma O[a,s',o]; Rb[i*m + a] = sum_k B[i,k] R[Z[i,k], a]
__global__ __launch_bounds__(kPB) void dense_build_scales(unsigned n, unsigned kpad, unsigned m,
        unsigned z, unsigned r, unsigned rz, float gamma, const float *__restrict__ O,
        const float *__restrict__ R, const int *__restrict__ Z, const float *__restrict__ B,
        float *__restrict__ Os, float *__restrict__ Rb)
{
    const size_t total = (size_t)m * z * n;
    for (size_t x = blockIdx.x * (size_t)kPB + threadIdx.x; x < total; x += (size_t)gridDim.x * kPB) {
        const unsigned sp = x % n;
        const unsigned ao = x / n, a = ao / z, o = ao % z;
        Os[(size_t)ao * kpad + sp] = gamma * O[((size_t)a * n + sp) * z + o];
    }
    for (size_t x = blockIdx.x * (size_t)kPB + threadIdx.x; x < (size_t)r * m; x += (size_t)gridDim.x * kPB) {
        const unsigned i = x / m, a = x % m;
        float v = 0.0f;
        for (unsigned k = 0; k < rz; k++) {
            const int s = Z[(size_t)i * rz + k];
            if (s < 0)
                break;
            v += B[(size_t)i * rz + k] * R[(size_t)s * m + a];
        }
        Rb[x] = v;
    }
}


// Dense action choice: merges the alpha tiles per (a, o) (earlier tile wins ties), and picks
// the first action maximising Rb[i,a] + sum_o best score (the scores are b . alpha terms).
__global__ __launch_bounds__(kPB) void dense_action(unsigned m, unsigned z, unsigned r,
        unsigned tiles, const float *__restrict__ Rb, const float *__restrict__ partVal,
        const int *__restrict__ partJ, int *__restrict__ jstar, int *__restrict__ act)
{
    __shared__ float best[kZC * 8];
    const unsigned i = blockIdx.x, t = threadIdx.x;
    float bestValue = kLow;
    int bestA = -1;
    for (unsigned a = 0; a < m; a++) {
        const size_t ia = (size_t)i * m + a;
        float total = Rb[ia];
        for (unsigned o0 = 0; o0 < z; o0 += kZC * 8) {
            const unsigned oc = min((unsigned)(kZC * 8), z - o0);
            for (unsigned oo = t; oo < oc; oo += kPB) {
                float v = kLow;
                int j = INT_MAX;
                for (unsigned tl = 0; tl < tiles; tl++) {
                    const size_t idx = ((size_t)tl * r * m + ia) * z + o0 + oo;
                    if (partVal[idx] > v) {
                        v = partVal[idx];
                        j = partJ[idx];
                    }
                }
                best[oo] = v;
                jstar[ia * z + o0 + oo] = (j == INT_MAX) ? 0 : j;
            }
            __syncthreads();
            if (t == 0)
                for (unsigned oo = 0; oo < oc; oo++)
                    total += best[oo];
            __syncthreads();
        }
        if (t == 0 && total > bestValue) {
            bestValue = total;
            bestA = (int)a;
        }
    }
    if (t == 0)
        act[i] = bestA;
}

} // namespace

// Backup arithmetic: 0 = fast (default), 1 = exact reference order, 2 = dense beliefs with
// bf16 matrix-core scoring, 3 = dense beliefs with fp32 matrix-core scoring.
static int g_pbvi_mode = 0;

static bool dense_mode(int mode) { return mode == 2 || mode == 3; }

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Scoring variants: jpt = 1, 2, 4 alphas per lane on the vector ALUs (tiles of 256 * jpt alphas),
// jpt = 8, 16, 32 the matrix-core kernel of fast mode, jpt = 64 its 16-observation-row form,
// jpt = 128 its form that sweeps every alpha tile in one workgroup (tiles of kMTile alphas).
static unsigned tile_width(unsigned jpt) { return jpt >= 8 ? (unsigned)kMTile : kPB * jpt; }
static unsigned tiles_for(unsigned r, unsigned jpt) { return (r + tile_width(jpt) - 1) / tile_width(jpt); }

// Row stride of the transposed Gamma: r rounded up to the matrix-core tile.
static unsigned gt_stride(unsigned r) { return (r + kMTile - 1) / kMTile * kMTile; }

// Supports exact mode takes: its scorer stages the support in chunks of at most kUMax pairs
// (so ns <= kUMax), and its action kernel keeps one alpha value per support entry in LDS.
static bool pbvi_shape_ok(const nova::POMDP *p)
{
    return p->rz <= (unsigned)kActMax && p->ns <= (unsigned)kUMax;
}

// Dynamic LDS of pbvi_action for an instance: alphaK and bS over its support (exact mode), or
// the running action value (fast mode).
static size_t action_lds(const PbviInst &I, int mode)
{
    // exact mode runs only supports of at most kActMax entries (pbvi_shape_ok): a longer one
    // (dense beliefs) is never run by the exact kernel, and sizing for it would pass the LDS
    // a workgroup can have beside the kernel's static arrays
    return mode == 1 ? 2 * sizeof(float) * std::min<size_t>(std::max<size_t>(I.rz, kKMax),
                                                            kActMax)
                     : 2 * sizeof(float);
}

// Whether the fast-mode pair lists (r*m*rz*ns entries of 8 bytes) are built: always, unless they
// would pass 2^30 entries (beliefs with full support at large n belong to the dense modes).
constexpr size_t kPairCap = size_t(1) << 30;
static bool pairs_ok(const nova::POMDP *p)
{
    return (size_t)p->r * p->m * p->rz * p->ns <= kPairCap;
}

// The sparse backup can run this instance in this mode: exact mode needs the LDS-staged support,
// fast mode the pair lists (the matrix-core scorer takes any support length).
static bool sparse_shape_ok(const nova::POMDP *p, int mode)
{
    return mode == 1 ? pbvi_shape_ok(p) : pairs_ok(p);
}

// Bytes of the fast-mode pair lists.
static size_t pair_bytes(const nova::POMDP *p)
{
    if (!pairs_ok(p))
        return 0;
    const size_t lists = (size_t)p->r * p->m * p->rz * p->ns;
    return 2 * align256(sizeof(float) * lists) + align256(sizeof(unsigned) * (size_t)p->r * p->m);
}

// Bytes of one instance's workspace (tiles sized for the widest variant, jpt = 1).
static size_t inst_ws_bytes(const nova::POMDP *p)
{
    const size_t rmz = (size_t)p->r * p->m * p->z;
    const unsigned tiles = tiles_for(p->r, 1);
    return pair_bytes(p) + align256(sizeof(float) * (size_t)p->m * p->z * p->n * p->ns) +
           2 * align256(sizeof(float) * (size_t)p->m * p->n * p->ns) +
           align256(sizeof(float) * (size_t)p->m * p->n) +
           align256(sizeof(float) * (size_t)p->m * p->z * p->n) +
           align256(sizeof(float) * (size_t)p->n * gt_stride(p->r)) +
           align256(sizeof(float) * tiles * rmz) + align256(sizeof(int) * tiles * rmz) +
           align256(sizeof(int) * rmz) + align256(sizeof(float) * (size_t)p->r * p->m) +
           align256(sizeof(int) * p->r);
}

// Carves an instance's workspace (OT first) out of base.
static void inst_ws(const nova::POMDP *p, char *base, PbviInst &I)
{
    const size_t rmz = (size_t)p->r * p->m * p->z;
    const unsigned tiles = tiles_for(p->r, 1);
    char *c = base;
    I.OT = reinterpret_cast<float *>(c);
    c += align256(sizeof(float) * (size_t)p->m * p->z * p->n * p->ns);
    I.SA = reinterpret_cast<int *>(c);
    c += align256(sizeof(float) * (size_t)p->m * p->n * p->ns);
    I.TA = reinterpret_cast<float *>(c);
    c += align256(sizeof(float) * (size_t)p->m * p->n * p->ns);
    I.RT = reinterpret_cast<float *>(c);
    c += align256(sizeof(float) * (size_t)p->m * p->n);
    I.OTr = reinterpret_cast<float *>(c);
    c += align256(sizeof(float) * (size_t)p->m * p->z * p->n);
    I.GT = reinterpret_cast<float *>(c);
    I.rS = gt_stride(p->r);
    c += align256(sizeof(float) * (size_t)p->n * I.rS);
    I.kmax = p->rz * p->ns;
    if (pairs_ok(p)) {
        const size_t lists = (size_t)p->r * p->m * p->rz * p->ns;
        I.pairSp = reinterpret_cast<int *>(c);
        c += align256(sizeof(float) * lists);
        I.pairW = reinterpret_cast<float *>(c);
        c += align256(sizeof(float) * lists);
        I.pairCnt = reinterpret_cast<unsigned *>(c);
        c += align256(sizeof(unsigned) * (size_t)p->r * p->m);
    }
    I.partVal = reinterpret_cast<float *>(c);
    c += align256(sizeof(float) * tiles * rmz);
    I.partJ = reinterpret_cast<int *>(c);
    c += align256(sizeof(int) * tiles * rmz);
    I.jstar = reinterpret_cast<int *>(c);
    c += align256(sizeof(int) * rmz);
    I.aval = reinterpret_cast<float *>(c);
    c += align256(sizeof(float) * (size_t)p->r * p->m);
    I.act = reinterpret_cast<int *>(c);
}

// Whether n floats can be staged in LDS by the materialize kernels (their dynamic LDS must be
// raised above the 64 KB default once per process).
static bool pbvi_lds_rows(unsigned n)
{
    static int state = 0;   // 0 unknown, 1 ok, 2 unavailable
    if (state == 0) {
        const int bytes = kLdsRowFloats * (int)sizeof(float);
        hipError_t e1 = hipFuncSetAttribute((const void *)pbvi_materialize<true>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        hipError_t e2 = hipFuncSetAttribute((const void *)pbvi_materialize<false>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        state = (e1 == hipSuccess && e2 == hipSuccess) ? 1 : 2;
        if (state == 2)
            (void)hipGetLastError();
    }
    return state == 1 && n <= (unsigned)kLdsRowFloats;
}

static void fill_inst(const nova::POMDP *p, PbviInst &I, float *G0, float *G1, unsigned *pi,
                      char *ws, unsigned horizon)
{
    std::memset(&I, 0, sizeof(I));
    I.n = p->n; I.ns = p->ns; I.m = p->m; I.z = p->z; I.r = p->r; I.rz = p->rz;
    I.rb = p->r;
    I.horizon = horizon;
    I.gamma = p->gamma;
    I.jpt = 2;
    I.tiles = tiles_for(p->r, I.jpt);
    I.useLds = pbvi_lds_rows(p->n) ? 1 : 0;
    I.S = p->d_S; I.T = p->d_T; I.O = p->d_O; I.R = p->d_R; I.Z = p->d_Z; I.B = p->d_B;
    I.G0 = G0; I.G1 = G1; I.pi = pi;
    I.actionMajor = 1;
    if (const char *f = std::getenv("NOVA_AMD_PBVI_ORDER"))
        I.actionMajor = std::atoi(f) != 0 ? 1 : 0;
    inst_ws(p, ws, I);
}

// Words of LDS that pbvi_fused stages for an instance: Gamma, O*T, S, R and one belief.
static double fused_stage_words(const PbviInst &I)
{
    return (double)I.r * (I.n | 1u) + (double)I.m * (I.z + 1.0) * I.n * I.ns +
           (double)I.n * I.m + 2.0 * I.rz;
}

// Whether a PBVI instance runs its iterations as single pbvi_fused launches: sparse modes, the
// kernel's LDS bounds and the exact mode's support staging, and at most 256 (fast mode) or
// 1024 (exact mode) scoring multiply-adds per lane and iteration, m * ceil(z*r / 256) * rz*ns.
// Measured on nova's domains, the one launch takes 14-21 us per iteration at tiger, paint and
// shuttle (24-28 us with four kernels, fast mode) and 61 us at 4x3 (35 us fast, 83 us exact);
// at aloha and tiger_grid it is 3-5x slower than fast mode's four kernels.
// NOVA_AMD_PBVI_FUSED=0 turns it off; another value replaces the limit of both modes.
static bool fused_fits(const PbviInst &I, int mode)
{
    if (mode != 0 && mode != 1)
        return false;
    double limit = (mode == 1) ? 1024.0 : 256.0;
    if (const char *f = std::getenv("NOVA_AMD_PBVI_FUSED"))
        limit = std::atof(f);
    if (limit <= 0.0)
        return false;
    const size_t zr = (size_t)I.z * I.r;
    const double perLane = (double)I.m * (double)((zr + kPB - 1) / kPB) * I.rz * I.ns;
    const double stage = fused_stage_words(I);
    return I.rb == I.r && I.n <= (unsigned)kFN && I.z <= (unsigned)kFZ &&
           (size_t)I.z * I.r <= (size_t)kFS && I.rz <= (unsigned)kKMax &&
           (size_t)I.rz * I.ns <= (size_t)kUMax && stage <= (double)kFStage &&
           perLane <= limit;
}

// A launch plan over a set of instances: device descriptors and per-stage offsets.
struct DenseWs;

// Scoring variants a plan launches (jpt 1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048,
// 4096); PbviPlan::totSc and the offset table hold one entry per variant.
constexpr int kVariants = 13;

struct PbviPlan {
    const DenseWs *dense = nullptr;  // dense-belief workspace (plain solves in modes 2/3)
    unsigned count = 0;
    PbviInst *dInst = nullptr;       // device copy of the descriptors
    unsigned *dOff = nullptr;        // [tr | sc(jpt=1) | ... | sc(32) | be | fu], count+1 each
    unsigned totTr = 0, totSc[kVariants] = {}, totBe = 0, totFu = 0;
    size_t fuLds = 0;                // dynamic LDS of pbvi_fused: the largest fused instance's
    size_t ldsBytes = 0;
    size_t actLds = 2 * sizeof(float) * kKMax;   // dynamic LDS of pbvi_action
    unsigned maxHorizon = 0;
    std::vector<PbviInst> host;
};

static_assert(sizeof(PbviPlan::totSc) == kVariants * sizeof(unsigned),
              "one workgroup count per scoring variant");

static int jpt_slot(unsigned jpt)
{
    return jpt == 1 ? 0 : (jpt == 2 ? 1 : (jpt == 4 ? 2 : (jpt == 8 ? 3 : (jpt == 16 ? 4 :
           (jpt == 32 ? 5 : (jpt == 64 ? 6 : (jpt == 128 ? 7 : (jpt == 256 ? 8 :
           (jpt == 512 ? 9 : (jpt == 1024 ? 10 : (jpt == 2048 ? 11 : 12)))))))))));
}

// Launches the scoring kernel of variant jpt over grid g (offsets sc of that variant).
static void launch_score(unsigned jpt, int mode, dim3 g, const PbviInst *dInst,
                         const unsigned *sc, unsigned c, unsigned iter, hipStream_t st)
{
    if (mode == 1) {
        if (jpt == 1) hipLaunchKernelGGL((pbvi_score<true, 1>), g, dim3(kPB), 0, st, dInst, sc, c, iter);
        if (jpt == 2) hipLaunchKernelGGL((pbvi_score<true, 2>), g, dim3(kPB), 0, st, dInst, sc, c, iter);
        if (jpt == 4) hipLaunchKernelGGL((pbvi_score<true, 4>), g, dim3(kPB), 0, st, dInst, sc, c, iter);
    } else {
        if (jpt == 1) hipLaunchKernelGGL((pbvi_score<false, 1>), g, dim3(kPB), 0, st, dInst, sc, c, iter);
        if (jpt == 2) hipLaunchKernelGGL((pbvi_score<false, 2>), g, dim3(kPB), 0, st, dInst, sc, c, iter);
        if (jpt == 4) hipLaunchKernelGGL((pbvi_score<false, 4>), g, dim3(kPB), 0, st, dInst, sc, c, iter);
        if (jpt == 8) hipLaunchKernelGGL((pbvi_score_mfma<4, 4, 4, 2, false>), g, dim3(256), 0, st, dInst, sc, c, iter);
        if (jpt == 16) hipLaunchKernelGGL((pbvi_score_mfma<8, 2, 4, 2, false>), g, dim3(512), 0, st, dInst, sc, c, iter);
        if (jpt == 32) hipLaunchKernelGGL((pbvi_score_mfma<8, 2, 8, 2, false>), g, dim3(512), 0, st, dInst, sc, c, iter);
        if (jpt == 256) hipLaunchKernelGGL((pbvi_score_mfma<4, 4, 2, 4, true>), g, dim3(256), 0, st, dInst, sc, c, iter);
        if (jpt == 512) hipLaunchKernelGGL((pbvi_score_mfma<8, 2, 4, 4, true>), g, dim3(512), 0, st, dInst, sc, c, iter);
        if (jpt == 64) hipLaunchKernelGGL((pbvi_score_mfma16<2, 2, false>), g, dim3(512), 0, st, dInst, sc, c, iter);
        if (jpt == 1024) hipLaunchKernelGGL((pbvi_score_mfma16<2, 4, true>), g, dim3(512), 0, st, dInst, sc, c, iter);
        if (jpt == 128) hipLaunchKernelGGL((pbvi_score_mfma_all<4, 2, false>), g, dim3(512), 0, st, dInst, sc, c, iter);
        if (jpt == 2048) hipLaunchKernelGGL((pbvi_score_mfma_all<4, 4, true>), g, dim3(512), 0, st, dInst, sc, c, iter);
        if (jpt == 4096) hipLaunchKernelGGL((pbvi_score_mfma16_all<2, 4, true>), g, dim3(512), 0, st, dInst, sc, c, iter);
    }
}

// Writes descriptors and offsets into dst (device memory of plan_bytes(count)).
static size_t plan_bytes(unsigned count)
{
    return align256(sizeof(PbviInst) * count) +
           align256(sizeof(unsigned) * (kVariants + 3) * (count + 1));
}

static int plan_upload(PbviPlan &pl, void *dst, hipStream_t st, const char *fn)
{
    const unsigned c = pl.count;
    // [tr | score offsets of each variant | be | fu], c + 1 entries each; a fused instance
    // has workgroups in fu only, every other instance in the first four stages only
    std::vector<unsigned> off((kVariants + 3) * (size_t)(c + 1), 0);
    unsigned *tr = &off[0], *be = &off[(kVariants + 1) * (c + 1)];
    unsigned *fu = &off[(kVariants + 2) * (c + 1)];
    unsigned *scs[kVariants];
    for (int v = 0; v < kVariants; v++)
        scs[v] = &off[(1 + v) * (c + 1)];
    pl.ldsBytes = 0;
    pl.maxHorizon = 0;
    for (unsigned b = 0; b < c; b++) {
        PbviInst &I = pl.host[b];
        I.tiles = tiles_for(I.r, I.jpt);
        const bool four = !I.fused;
        tr[b + 1] = tr[b] + (four ? ((I.n + 63) / 64) * (I.rS / 64) : 0u);
        for (int v = 0; v < kVariants; v++)
            scs[v][b + 1] = scs[v][b] +
                            ((four && jpt_slot(I.jpt) == v)
                             ? I.rb * I.m * ((I.jpt == 128 || I.jpt == 2048 || I.jpt == 4096)
                                             ? 1u : I.tiles)
                             : 0u);
        be[b + 1] = be[b] + (four ? I.rb : 0u);
        fu[b + 1] = fu[b] + (four ? 0u : I.rb);
        if (!four)
            pl.fuLds = std::max(pl.fuLds, sizeof(float) * (size_t)fused_stage_words(I));
        if (I.useLds && four)
            pl.ldsBytes = std::max(pl.ldsBytes, sizeof(float) * (size_t)I.n);
        if (four)
            pl.actLds = std::max(pl.actLds, action_lds(I, 1));
        if (I.horizon != kNoHorizon)
            pl.maxHorizon = std::max(pl.maxHorizon, I.horizon);
    }
    pl.totTr = tr[c];
    for (int v = 0; v < kVariants; v++)
        pl.totSc[v] = scs[v][c];
    pl.totBe = be[c];
    pl.totFu = fu[c];
    hipError_t ea = hipFuncSetAttribute((const void *)pbvi_action<true>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)pl.actLds);
    if (ea == hipSuccess)
        ea = hipFuncSetAttribute((const void *)pbvi_action<false>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)pl.actLds);
    if (ea == hipSuccess && pl.totFu > 0)
        ea = hipFuncSetAttribute((const void *)pbvi_fused,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 4 * kFStage);
    if (ea != hipSuccess)
        return hip_fail(ea, fn, "Failed to set up the one-launch PBVI kernel.",
                        NOVA_ERROR_KERNEL_EXECUTION);
    pl.dInst = static_cast<PbviInst *>(dst);
    pl.dOff = reinterpret_cast<unsigned *>(static_cast<char *>(dst) +
                                           align256(sizeof(PbviInst) * c));
    hipError_t e = hipMemcpyAsync(pl.dInst, pl.host.data(), sizeof(PbviInst) * c,
                                  hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(pl.dOff, off.data(), sizeof(unsigned) * off.size(),
                           hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    return hip_fail(e, fn, "Failed to upload the PBVI launch plan.", NOVA_ERROR_MEMCPY_TO_DEVICE);
}

// The extra device buffers of a dense-belief solve (modes 2 and 3), per solver workspace.
struct DenseWs {
    const void *key = nullptr;       // POMDPPBVIGPU::d_alphaBA it belongs to
    int mode = 0;
    bool blocked = false;            // bf16 through dense_score_bf16 (k-blocked operands)
    void *mem = nullptr;
    void *U = nullptr;               // [m][rpad][kpad] in the scoring precision (blocked:
                                     // [kpad / 32][m * rpad][32])
    void *GI = nullptr;              // [rpad][kpad] Gamma in the scoring precision (blocked:
                                     // [kpad / 32][rpad][32])
    void *BP = nullptr;              // blocked: Gamma scaled by gamma*O for a chunk of pairs
    unsigned chunk = 0;              // blocked: pairs per chunk
    float *Os = nullptr;             // [m][z][kpad] gamma*O
    float *Rb = nullptr;             // [r][m] b . R[:,a]
    float *partVal = nullptr;        // [tiles][r][m][z]
    int *partJ = nullptr;
    unsigned tiles = 0, n = 0, m = 0, z = 0, r = 0, rpad = 0, kpad = 0;
};
static std::vector<DenseWs> g_dense;

static const DenseWs *dense_for(const void *key)
{
    for (auto &d : g_dense)
        if (d.key == key)
            return &d;
    return nullptr;
}

static void dense_release(const void *key, const char *fn)
{
    for (size_t x = 0; x < g_dense.size(); x++)
        if (g_dense[x].key == key) {
            device_free(g_dense[x].mem, fn, "the dense-belief workspace");
            g_dense.erase(g_dense.begin() + x);
            return;
        }
}

// Gamma conversion and the matrix-core scoring of one dense iteration; ev (nullable) gets an
// event recorded between the two.
static void dense_launch_score(const DenseWs &D, const float *G, hipStream_t st, hipEvent_t ev)
{
    const size_t rn = (size_t)D.r * D.n;
    const unsigned cg = (unsigned)std::min<size_t>((rn + kGConvThreads - 1) / kGConvThreads, 65536);
    const dim3 sg(D.tiles * D.tiles, D.m * D.z);
    if (D.blocked) {
        // Gamma once, then per chunk of (a, o) pairs B_p = bf16(Gamma * gamma O[a,.,o]) and
        // the GEMM of the chunk's pairs against U_a
        const size_t rowsPad = D.rpad, per = rowsPad * D.kpad;
        const unsigned cb = (unsigned)std::min<size_t>(
            (per / 8 + kGConvThreads - 1) / kGConvThreads, 65536);
        hipLaunchKernelGGL(dense_convert_blk, dim3(cb), dim3(kGConvThreads), 0, st, D.n, D.kpad,
                           rn, rowsPad, G, static_cast<__bf16 *>(D.GI));
        if (ev)
            hipEventRecord(ev, st);
        const unsigned pairs = D.m * D.z;
        for (unsigned p0 = 0; p0 < pairs; p0 += D.chunk) {
            const unsigned cnt = std::min(D.chunk, pairs - p0);
            const unsigned cs = (unsigned)std::min<size_t>(
                (per / 8 + kGConvThreads - 1) / kGConvThreads, 8192);
            hipLaunchKernelGGL(dense_scale_pairs, dim3(cs), dim3(kGConvThreads), 0, st, D.kpad,
                               rowsPad, p0, cnt, static_cast<const __bf16 *>(D.GI), D.Os,
                               static_cast<__bf16 *>(D.BP));
            hipLaunchKernelGGL(dense_score_bf16, dim3(D.tiles * D.tiles, cnt), dim3(kGThreads),
                               kGDlLdsBytes, st, D.r, D.r, D.kpad, D.m, D.z, D.tiles, D.tiles, p0,
                               per, static_cast<const __bf16 *>(D.U),
                               static_cast<const __bf16 *>(D.BP), D.partVal, D.partJ);
        }
    } else if (D.mode == 2) {
        hipLaunchKernelGGL(dense_convert<__bf16>, dim3(cg), dim3(kGConvThreads), 0, st, D.n,
                           D.kpad, rn, G, static_cast<__bf16 *>(D.GI));
        if (ev)
            hipEventRecord(ev, st);
        hipLaunchKernelGGL(dense_score<__bf16>, sg, dim3(kGThreads), kGLdsBytes, st, D.r, D.r,
                           D.kpad, D.m, D.z, D.tiles, D.tiles, static_cast<const __bf16 *>(D.U),
                           static_cast<const __bf16 *>(D.GI), D.Os, D.partVal, D.partJ);
    } else {
        hipLaunchKernelGGL(dense_convert<float>, dim3(cg), dim3(kGConvThreads), 0, st, D.n,
                           D.kpad, rn, G, static_cast<float *>(D.GI));
        if (ev)
            hipEventRecord(ev, st);
        hipLaunchKernelGGL(dense_score<float>, sg, dim3(kGThreads), kGLdsBytes, st, D.r, D.r,
                           D.kpad, D.m, D.z, D.tiles, D.tiles, static_cast<const float *>(D.U),
                           static_cast<const float *>(D.GI), D.Os, D.partVal, D.partJ);
    }
}

static void dense_enqueue(const PbviPlan &pl, unsigned iter, unsigned flip, hipStream_t st)
{
    const DenseWs &D = *pl.dense;
    const PbviInst &I = pl.host[0];
    const float *G = ((iter + flip) % 2 == 0) ? I.G0 : I.G1;
    dense_launch_score(D, G, st, nullptr);
    hipLaunchKernelGGL(dense_action, dim3(D.r), dim3(kPB), 0, st, D.m, D.z, D.r, D.tiles, D.Rb,
                       D.partVal, D.partJ, I.jstar, I.act);
    const unsigned *be = pl.dOff + (kVariants + 1) * (pl.count + 1);
    hipLaunchKernelGGL(pbvi_materialize<false>, dim3(pl.totBe), dim3(kMB), pl.ldsBytes, st,
                       pl.dInst, be, pl.count, iter, flip);
}

// Enqueues iteration `iter` of every instance of the plan.
static void plan_enqueue(const PbviPlan &pl, unsigned iter, unsigned flip, int mode,
                         hipStream_t st)
{
    if (pl.dense) {
        dense_enqueue(pl, iter, flip, st);
        return;
    }
    const unsigned c = pl.count;
    const unsigned *tr = pl.dOff, *be = pl.dOff + (kVariants + 1) * (c + 1);
    const unsigned *fu = pl.dOff + (kVariants + 2) * (c + 1);
    if (pl.totFu > 0)
        hipLaunchKernelGGL(pbvi_fused, dim3(pl.totFu), dim3(kPB), pl.fuLds, st, pl.dInst, fu, c,
                           iter, flip);
    if (pl.totBe == 0)
        return;
    hipLaunchKernelGGL(pbvi_transpose, dim3(pl.totTr), dim3(kPB), 0, st, pl.dInst, tr, c, iter, flip);
    static const unsigned kJpt[kVariants] = {1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048,
                                             4096};
    for (int v = 0; v < kVariants; v++) {
        if (pl.totSc[v] == 0)
            continue;
        launch_score(kJpt[v], mode, dim3(pl.totSc[v]), pl.dInst, pl.dOff + (1 + v) * (c + 1), c,
                     iter, st);
    }
    if (mode == 1)
        hipLaunchKernelGGL(pbvi_action<true>, dim3(pl.totBe), dim3(kPB), pl.actLds, st, pl.dInst, be, c, iter, flip);
    else
        hipLaunchKernelGGL(pbvi_action<false>, dim3(pl.totBe), dim3(kPB), pl.actLds, st, pl.dInst, be, c, iter, flip);
    auto matFn = (mode == 1) ? pbvi_materialize<true> : pbvi_materialize<false>;
    hipLaunchKernelGGL(matFn, dim3(pl.totBe), dim3(kMB), pl.ldsBytes, st, pl.dInst, be, c, iter,
                       flip);
}

// Captures `iters` iterations of a plan into an executable graph (on the capture stream; the
// caller launches it on the library stream).
static int plan_graph(const PbviPlan &pl, unsigned iters, unsigned flip, int mode,
                      hipStream_t, hipGraphExec_t *exec, const char *fn)
{
    hipGraph_t graph = nullptr;
    hipStream_t cs = capture_stream();
    if (cs == nullptr) {
        NOVA_ERR(fn, "Could not create the graph-capture stream.");
        return NOVA_ERROR_KERNEL_EXECUTION;
    }
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess)
        return hip_fail(e, fn, "Failed to begin graph capture.", NOVA_ERROR_KERNEL_EXECUTION);
    for (unsigned k = 0; k < iters; k++)
        plan_enqueue(pl, k, flip, mode, cs);
    e = hipStreamEndCapture(cs, &graph);
    if (e != hipSuccess)
        return hip_fail(e, fn, "Failed to capture the PBVI iterations.", NOVA_ERROR_KERNEL_EXECUTION);
    e = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) {
        *exec = nullptr;
        return hip_fail(e, fn, "Failed to instantiate the PBVI graph.", NOVA_ERROR_KERNEL_EXECUTION);
    }
    return NOVA_SUCCESS;
}

// ------------------------------------------------------------------------ the tuner
//
// The scoring kernel's variants (1, 2 or 4 alphas per lane) trade occupancy against reuse of
// the staged O*T factors. For each size class (r, belief-support pairs rz*ns and z, each
// bucketed to a power of two) the first instance seen is timed with every variant on one
// scoring launch and the fastest is remembered for the process. NOVA_AMD_PBVI_JPT forces a
// variant: 1, 2, 4 (vector ALUs), and in fast mode only the matrix-core kernels 8, 16, 32 (4
// waves; 8 waves; 8 waves with twice the pairs per trip), 256 and 512 (the 4- and 8-wave kernels
// with operands requested three trips ahead, 32 loads in flight per wave), 64 (16 observation
// rows, z <= 16), 1024 (the 16-row kernel with operands requested three trips ahead), 128
// (every alpha tile in one workgroup, pair lists of at most 512 pairs), 2048 (that one with
// operands requested three trips ahead) and 4096 (the 16-row kernel likewise sweeping every
// tile, z <= 16, pair lists of at most 512 pairs);
// NOVA_AMD_AUTOTUNE=0 disables timing (default variant 2).

struct TuneKey {
    unsigned r, u, z;
    int exact;
    bool operator==(const TuneKey &o) const
    {
        return r == o.r && u == o.u && z == o.z && exact == o.exact;
    }
};
static std::vector<std::pair<TuneKey, unsigned>> g_tuned;

static unsigned bucket(unsigned x)
{
    unsigned p = 1;
    while (p < x)
        p <<= 1;
    return p;
}

// Gamma^T of the instance spans less than 4 GB (byte offsets of 32 bits reach all of it).
static bool gt_below_4g(const PbviInst &I)
{
    return (size_t)I.n * I.rS * sizeof(float) < ((size_t)1 << 32);
}

static unsigned tune_jpt(PbviInst I, void *planMem, hipStream_t st, int mode)
{
    // the vector-ALU variants stage the support in LDS; longer supports take the matrix cores
    const bool valuOk = mode == 1 ||
                        (I.rz <= (unsigned)kKMax && I.rz * I.ns <= (unsigned)kUMax);
    if (const char *f = std::getenv("NOVA_AMD_PBVI_JPT")) {
        const int v = std::atoi(f);
        if (((v == 1 || v == 2 || v == 4) && valuOk) ||
                ((v == 8 || v == 16 || v == 32) && mode != 1) ||
                ((v == 256 || v == 512) && mode != 1 && gt_below_4g(I)) ||
                (v == 1024 && mode != 1 && I.z <= 16 && gt_below_4g(I)) ||
                (v == 2048 && mode != 1 && I.kmax <= (unsigned)kWAll && gt_below_4g(I)) ||
                (v == 64 && mode != 1 && I.z <= 16) ||
                (v == 128 && mode != 1 && I.kmax <= (unsigned)kWAll) ||
                (v == 4096 && mode != 1 && I.z <= 16 && I.kmax <= (unsigned)kWAll &&
                 gt_below_4g(I)))
            return (unsigned)v;
    }
    if (const char *f = std::getenv("NOVA_AMD_AUTOTUNE"))
        if (std::atoi(f) == 0)
            return valuOk ? 2 : 16;
    const TuneKey key{bucket(I.r), bucket(I.rz * I.ns), bucket(I.z), mode == 1 ? 1 : 0};
    for (auto &kv : g_tuned)
        if (kv.first == key)
            return kv.second;
    unsigned best = valuOk ? 2 : 16;
    float bestMs = 1e30f;
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    for (unsigned jpt : {1u, 2u, 4u, 8u, 16u, 32u, 64u, 128u, 256u, 512u, 1024u, 2048u, 4096u}) {
        if (jpt >= 8 && mode == 1)
            continue;                          // the matrix-core kernel is fast mode only
        if ((jpt == 64 || jpt == 1024 || jpt == 4096) && I.z > 16)
            continue;                          // the 16-row kernels: at most 16 observations
        if ((jpt == 128 || jpt == 2048 || jpt == 4096) && I.kmax > (unsigned)kWAll)
            continue;                          // the all-tiles kernels stage every pair once
        if ((jpt == 256 || jpt == 512 || jpt == 1024 || jpt == 2048 || jpt == 4096) &&
                !gt_below_4g(I))
            continue;                          // their buffer loads take 32-bit byte offsets
        if (jpt < 8 && !valuOk)
            continue;
        PbviPlan pl;
        pl.count = 1;
        I.jpt = jpt;
        I.horizon = kNoHorizon;
        I.fused = 0;                           // the tuner times the four-kernel path
        pl.host.assign(1, I);
        if (plan_upload(pl, planMem, st, "pbvi tuner") != NOVA_SUCCESS)
            continue;
        const unsigned tr = pl.totTr;
        hipLaunchKernelGGL(pbvi_transpose, dim3(tr), dim3(kPB), 0, st, pl.dInst, pl.dOff, 1u, 0u, 0u);
        const unsigned *sc = pl.dOff + (1 + jpt_slot(jpt)) * 2;
        dim3 g(pl.totSc[jpt_slot(jpt)]);
        float ms = 0.0f;
        for (int rep = 0; rep < 2; rep++) {   // first launch warms up
            hipEventRecord(e0, st);
            launch_score(jpt, mode, g, pl.dInst, sc, 1u, 0u, st);
            hipEventRecord(e1, st);
            hipEventSynchronize(e1);
            hipEventElapsedTime(&ms, e0, e1);
        }
        if (ms < bestMs) {
            bestMs = ms;
            best = jpt;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    g_tuned.push_back({key, best});
    return best;
}

// Builds the dense-belief workspace of a plain solver: U_a = B T_a for every action (belief
// rows staged in LDS, predecessor lists from the host S/T), gamma*O planes and b . R.
static int dense_setup(const nova::POMDP *p, const void *key, int mode, hipStream_t st,
                       const char *fn)
{
    dense_release(key, fn);
    const unsigned n = p->n, m = p->m, z = p->z, r = p->r, ns = p->ns;
    if (p->S == nullptr || p->T == nullptr || n > (unsigned)kLdsRowFloats) {
        NOVA_ERR(fn, "Dense-belief scoring needs the host S/T arrays and n <= 36864.");
        return NOVA_ERROR_INVALID_DATA;
    }
    const size_t esz = (mode == 2) ? 2 : 4;
    DenseWs D;
    D.key = key;
    D.mode = mode;
    D.n = n; D.m = m; D.z = z; D.r = r;
    D.tiles = (r + kGT - 1) / kGT;
    D.rpad = D.tiles * kGT;
    D.kpad = (n + kGKPad - 1) / kGKPad * kGKPad;
    const size_t rmz = (size_t)r * m * z;
    D.blocked = mode == 2 && !dense_legacy_bf16();
    // the scaled Gamma of as many pairs as fit about 2 GiB (at least one) per chunk;
    // NOVA_AMD_DENSE_CHUNK_PAIRS caps the pairs per chunk (for tests of the chunking)
    const size_t pairBytes = esz * (size_t)D.rpad * D.kpad;
    size_t chunk = std::max<size_t>(1, std::min<size_t>((size_t)m * z,
                                                        ((size_t)2 << 30) / pairBytes));
    if (const char *cap = getenv("NOVA_AMD_DENSE_CHUNK_PAIRS")) {
        const long v = strtol(cap, nullptr, 10);
        if (v > 0)
            chunk = std::min<size_t>(chunk, (size_t)v);
    }
    D.chunk = D.blocked ? (unsigned)std::min<size_t>(chunk, 65535) : 0u;   // grid.y limit
    const size_t bU = align256(esz * (size_t)m * D.rpad * D.kpad),
                 bG = align256(esz * (size_t)D.rpad * D.kpad),
                 bO = align256(sizeof(float) * (size_t)m * z * D.kpad),
                 bR = align256(sizeof(float) * (size_t)r * m),
                 bP = align256(sizeof(float) * D.tiles * rmz),
                 bB = align256(pairBytes * D.chunk);
    int rc = device_alloc(&D.mem, bU + bG + bO + bR + 2 * bP + bB, fn,
                          "the dense-belief workspace");
    if (rc != NOVA_SUCCESS)
        return rc;
    char *c = static_cast<char *>(D.mem);
    D.U = c; c += bU;
    D.GI = c; c += bG;
    D.Os = reinterpret_cast<float *>(c); c += bO;
    D.Rb = reinterpret_cast<float *>(c); c += bR;
    D.partVal = reinterpret_cast<float *>(c); c += bP;
    D.partJ = reinterpret_cast<int *>(c); c += bP;
    D.BP = D.blocked ? c : nullptr;
    g_dense.push_back(D);

    // predecessor lists of every action (host), then the device builds
    std::vector<unsigned> pstart((size_t)m * (n + 1), 0);
    for (unsigned s2 = 0; s2 < n; s2++)
        for (unsigned a = 0; a < m; a++)
            for (unsigned l = 0; l < ns; l++) {
                const int sp = p->S[idx3(s2, a, l, m, ns)];
                if (sp < 0)
                    break;
                pstart[(size_t)a * (n + 1) + sp + 1]++;
            }
    size_t total = 0;
    for (unsigned a = 0; a < m; a++) {
        unsigned *row = &pstart[(size_t)a * (n + 1)];
        for (unsigned sp = 0; sp < n; sp++)
            row[sp + 1] += row[sp];
        for (unsigned sp = 0; sp <= n; sp++)
            row[sp] += (unsigned)total;
        total = row[n];
    }
    std::vector<int> pS(std::max<size_t>(total, 1));
    std::vector<float> pT(std::max<size_t>(total, 1));
    {
        std::vector<unsigned> fillPos(pstart);
        for (unsigned s2 = 0; s2 < n; s2++)
            for (unsigned a = 0; a < m; a++)
                for (unsigned l = 0; l < ns; l++) {
                    const size_t x = idx3(s2, a, l, m, ns);
                    const int sp = p->S[x];
                    if (sp < 0)
                        break;
                    const unsigned q = fillPos[(size_t)a * (n + 1) + sp]++;
                    pS[q] = (int)s2;
                    pT[q] = p->T[x];
                }
    }
    void *dps = nullptr, *dpS = nullptr, *dpT = nullptr;
    rc = device_alloc(&dps, sizeof(unsigned) * pstart.size(), fn, "predecessor lists");
    if (rc == NOVA_SUCCESS)
        rc = device_alloc(&dpS, sizeof(int) * pS.size(), fn, "predecessor lists");
    if (rc == NOVA_SUCCESS)
        rc = device_alloc(&dpT, sizeof(float) * pT.size(), fn, "predecessor lists");
    hipError_t e = hipSuccess;
    if (rc == NOVA_SUCCESS) {
        e = hipMemcpy(dps, pstart.data(), sizeof(unsigned) * pstart.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(dpS, pS.data(), sizeof(int) * pS.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(dpT, pT.data(), sizeof(float) * pT.size(), hipMemcpyHostToDevice);
    }
    if (rc == NOVA_SUCCESS && e == hipSuccess) {
        // the zero padding of U, GI and Os is never written again
        const DenseWs &Dn = g_dense.back();
        e = hipMemsetAsync(Dn.mem, 0, bU + bG + bO, st);
        const int ldsBytes = (int)(sizeof(float) * n);
        const size_t lds = sizeof(float) * n;
        if (mode == 2) {
            hipFuncSetAttribute((const void *)dense_build_u<__bf16>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, ldsBytes);
            if (Dn.blocked)
                hipFuncSetAttribute((const void *)dense_score_bf16,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)kGDlLdsBytes);
            else
                hipFuncSetAttribute((const void *)dense_score<__bf16>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)kGLdsBytes);
            hipLaunchKernelGGL(dense_build_u<__bf16>, dim3(r, m), dim3(kPB), lds, st, n, m, p->rz,
                               Dn.rpad, Dn.kpad, Dn.blocked ? 1 : 0, p->d_Z, p->d_B,
                               static_cast<unsigned *>(dps), static_cast<int *>(dpS),
                               static_cast<float *>(dpT), static_cast<__bf16 *>(Dn.U));
        } else {
            hipFuncSetAttribute((const void *)dense_build_u<float>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, ldsBytes);
            hipFuncSetAttribute((const void *)dense_score<float>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)kGLdsBytes);
            hipLaunchKernelGGL(dense_build_u<float>, dim3(r, m), dim3(kPB), lds, st, n, m, p->rz,
                               Dn.rpad, Dn.kpad, 0, p->d_Z, p->d_B, static_cast<unsigned *>(dps),
                               static_cast<int *>(dpS), static_cast<float *>(dpT),
                               static_cast<float *>(Dn.U));
        }
        const size_t tot = std::max((size_t)m * z * n, (size_t)r * m);
        const unsigned grid = (unsigned)std::min<size_t>((tot + kPB - 1) / kPB, 65536);
        hipLaunchKernelGGL(dense_build_scales, dim3(grid), dim3(kPB), 0, st, n, Dn.kpad, m, z, r,
                           p->rz, p->gamma, p->d_O, p->d_R, p->d_Z, p->d_B, Dn.Os, Dn.Rb);
        if (e == hipSuccess)
            e = hipGetLastError();
        if (e == hipSuccess)
            e = hipStreamSynchronize(st);
    }
    if (dps) device_free(dps, fn, "predecessor lists");
    if (dpS) device_free(dpS, fn, "predecessor lists");
    if (dpT) device_free(dpT, fn, "predecessor lists");
    if (rc == NOVA_SUCCESS && e != hipSuccess)
        rc = hip_fail(e, fn, "Failed to build the dense-belief workspace.", NOVA_ERROR_KERNEL_EXECUTION);
    if (rc != NOVA_SUCCESS)
        dense_release(key, fn);
    return rc;
}

static bool pomdp_device_ready(const nova::POMDP *p)
{
    return p != nullptr && p->n != 0 && p->ns != 0 && p->m != 0 && p->z != 0 && p->r != 0 &&
           p->rz != 0 && p->d_S != nullptr && p->d_T != nullptr && p->d_O != nullptr &&
           p->d_R != nullptr && p->d_Z != nullptr && p->d_B != nullptr;
}

// The plain solver's workspace: [instance workspace | plan (descriptor + offsets)].
static char *plain_plan_mem(const nova::POMDP *p, nova::POMDPPBVIGPU *pb)
{
    return reinterpret_cast<char *>(pb->d_alphaBA) + align256(inst_ws_bytes(p));
}

// Cached executable graph of a plain solver's run.
struct PbviGraph {
    hipGraphExec_t exec = nullptr;
    const void *key[12] = {};
    unsigned iters = 0, flip = 0, dims[6] = {};
    float gamma = 0.0f;
    int mode = -1;
    unsigned jpt = 0;
    int fused = -1;
};
static PbviGraph g_pgraph;

// What a plain solver was initialized with, per workspace pointer: the arithmetic mode (later
// calls of nova_amd_pbvi_mode do not change it), the scoring variant the tuner chose, the
// model sizes its buffers were allocated for, and the model generation that its arrays derived
// from the model at initialize (O*T products, pair lists, dense operands) were built from.
struct PlainSolver {
    const void *ws = nullptr;
    int mode = 0;
    unsigned jpt = 2;
    unsigned dims[6] = {};
    unsigned long long generation = 0;
};
static std::vector<PlainSolver> g_plain;

static PlainSolver *plain_for(const void *ws)
{
    for (auto &S : g_plain)
        if (S.ws == ws)
            return &S;
    return nullptr;
}

static void plain_forget(const void *ws)
{
    for (size_t x = 0; x < g_plain.size(); x++)
        if (g_plain[x].ws == ws) {
            g_plain.erase(g_plain.begin() + x);
            return;
        }
}

static unsigned plain_jpt(const void *ws)
{
    const PlainSolver *S = plain_for(ws);
    return S ? S->jpt : 2u;
}

static void plain_dims(const nova::POMDP *p, unsigned (&d)[6])
{
    d[0] = p->n; d[1] = p->ns; d[2] = p->m; d[3] = p->z; d[4] = p->r; d[5] = p->rz;
}

// The solver's record if it was initialized for p's sizes (its buffers hold r x n alphas and
// r * m * rz * ns pairs of the sizes at initialize), else an error.
static int plain_sized(const nova::POMDP *p, const nova::POMDPPBVIGPU *pb, const char *fn,
                       PlainSolver **out)
{
    PlainSolver *S = plain_for(pb->d_alphaBA);
    if (S == nullptr) {
        NOVA_ERR(fn, "The solver is not initialized (pomdp_pbvi_initialize_gpu).");
        return NOVA_ERROR_INVALID_DATA;
    }
    unsigned d[6];
    plain_dims(p, d);
    if (std::memcmp(d, S->dims, sizeof(d)) != 0) {
        NOVA_ERR(fn, "The POMDP's sizes changed since pomdp_pbvi_initialize_gpu; initialize "
                     "the solver again.");
        return NOVA_ERROR_INVALID_DATA;
    }
    *out = S;
    return NOVA_SUCCESS;
}

static int build_ot(const nova::POMDP *p, const PbviInst &I, hipStream_t st);
static int dense_setup(const nova::POMDP *p, const void *key, int mode, hipStream_t st,
                       const char *fn);

// plain_sized, and the arrays derived from the model at initialize built again when model
// arrays were uploaded or released through this library since (the reference's kernels read
// the model arrays on every update).
static int plain_check(const nova::POMDP *p, nova::POMDPPBVIGPU *pb, const char *fn,
                       PlainSolver **out)
{
    PlainSolver *S = nullptr;
    int rc = plain_sized(p, pb, fn, &S);
    if (rc != NOVA_SUCCESS || S->generation == model_generation(false)) {
        *out = S;
        return rc;
    }
    hipStream_t st = stream();
    PbviInst I;
    fill_inst(p, I, pb->d_Gamma, pb->d_GammaPrime, pb->d_pi,
              reinterpret_cast<char *>(pb->d_alphaBA), kNoHorizon);
    rc = build_ot(p, I, st);
    if (rc == NOVA_SUCCESS && dense_mode(S->mode))
        rc = dense_setup(p, pb->d_alphaBA, S->mode, st, fn);
    if (rc == NOVA_SUCCESS)
        rc = hip_fail(hipStreamSynchronize(st), fn, "Failed to rebuild the model products.",
                      NOVA_ERROR_KERNEL_EXECUTION);
    if (rc == NOVA_SUCCESS)
        S->generation = model_generation(false);
    *out = S;
    return rc;
}

static int pbvi_run_iters(const nova::POMDP *p, nova::POMDPPBVIGPU *pb, unsigned iters,
                          hipStream_t st, hipEvent_t t0, hipEvent_t t1, const char *fn)
{
    PlainSolver *S = nullptr;
    int rc = plain_check(p, pb, fn, &S);
    if (rc != NOVA_SUCCESS)
        return rc;
    const int mode = S->mode;
    const unsigned flip = pb->currentHorizon % 2;
    const unsigned jpt = S->jpt;
    // every device address the captured kernels hold: the dense workspace is allocated apart
    // from d_alphaBA, so a re-initialize can move it while the other addresses repeat
    const DenseWs *dense = dense_for(pb->d_alphaBA);
    const void *key[12] = {p->d_S, p->d_T, p->d_O, p->d_R, p->d_Z, p->d_B, pb->d_Gamma,
                           pb->d_GammaPrime, pb->d_pi, pb->d_alphaBA,
                           dense ? dense->mem : nullptr, nullptr};
    const unsigned dims[6] = {p->n, p->ns, p->m, p->z, p->r, p->rz};
    // The graph's kernels read the descriptor from plan memory when they run, and other calls
    // (the tuner, profiling, recorded runs) reuse that memory: upload it on every run.
    PbviPlan pl;
    pl.count = 1;
    pl.host.resize(1);
    fill_inst(p, pl.host[0], pb->d_Gamma, pb->d_GammaPrime, pb->d_pi,
              reinterpret_cast<char *>(pb->d_alphaBA), kNoHorizon);
    pl.host[0].jpt = jpt;
    pl.host[0].fused = (dense == nullptr && fused_fits(pl.host[0], mode)) ? 1 : 0;
    const bool hit = g_pgraph.exec && g_pgraph.iters == iters && g_pgraph.flip == flip &&
                     g_pgraph.gamma == p->gamma && g_pgraph.mode == mode &&
                     g_pgraph.jpt == jpt && g_pgraph.fused == pl.host[0].fused &&
                     std::memcmp(g_pgraph.key, key, sizeof(key)) == 0 &&
                     std::memcmp(g_pgraph.dims, dims, sizeof(dims)) == 0;
    pl.dense = dense;
    if (pl.dense)
        pl.host[0].tiles = pl.dense->tiles;
    rc = plan_upload(pl, plain_plan_mem(p, pb), st, fn);
    if (rc != NOVA_SUCCESS)
        return rc;
    if (!hit) {
        if (g_pgraph.exec) {
            hipGraphExecDestroy(g_pgraph.exec);
            g_pgraph.exec = nullptr;
        }
        rc = plan_graph(pl, iters, flip, mode, st, &g_pgraph.exec, fn);
        if (rc != NOVA_SUCCESS)
            return rc;
        std::memcpy(g_pgraph.key, key, sizeof(key));
        std::memcpy(g_pgraph.dims, dims, sizeof(dims));
        g_pgraph.iters = iters;
        g_pgraph.flip = flip;
        g_pgraph.fused = pl.host[0].fused;
        g_pgraph.gamma = p->gamma;
        g_pgraph.mode = mode;
        g_pgraph.jpt = jpt;
    }
    if (t0) hipEventRecord(t0, st);
    hipError_t e = hipGraphLaunch(g_pgraph.exec, st);
    if (e != hipSuccess)
        return hip_fail(e, fn, "Failed to execute the PBVI graph.", NOVA_ERROR_KERNEL_EXECUTION);
    if (t1) hipEventRecord(t1, st);
    e = hipStreamSynchronize(st);
    if (e != hipSuccess)
        return hip_fail(e, fn, "Failed to synchronize after the PBVI iterations.",
                        NOVA_ERROR_DEVICE_SYNCHRONIZE);
    pb->currentHorizon += iters;
    return NOVA_SUCCESS;
}

// Builds the Gamma-independent parts of an instance's workspace: OT and the action-major copies.
static int build_ot(const nova::POMDP *p, const PbviInst &I, hipStream_t st)
{
    const size_t total = (size_t)p->m * p->z * p->n * p->ns;
    const unsigned grid = (unsigned)std::min<size_t>((total + kPB - 1) / kPB, 65536);
    hipLaunchKernelGGL(pbvi_build_ot, dim3(grid), dim3(kPB), 0, st, p->n, p->ns, p->m, p->z,
                       p->d_S, p->d_T, p->d_O, I.OT);
    hipLaunchKernelGGL(pbvi_build_aux, dim3(grid), dim3(kPB), 0, st, p->n, p->ns, p->m, p->z,
                       p->d_S, p->d_T, p->d_R, p->d_O, I.SA, I.TA, I.RT, I.OTr);
    if (I.pairSp != nullptr) {
        const size_t rm = (size_t)p->r * p->m;
        hipLaunchKernelGGL(pbvi_build_pairs, dim3((unsigned)((rm + kPB - 1) / kPB)), dim3(kPB), 0,
                           st, p->r, p->m, p->ns, p->rz, I.kmax, p->gamma, p->d_S, p->d_T,
                           p->d_Z, p->d_B, I.pairSp, I.pairW, I.pairCnt);
    }
    return hipGetLastError() == hipSuccess ? NOVA_SUCCESS : NOVA_ERROR_KERNEL_EXECUTION;
}

} // namespace nova_amd

namespace nova {

using namespace nova_amd;

int nova_amd_pbvi_mode(int mode)
{
    NOVA_AMD_API_LOCK();
    const int old = g_pbvi_mode;
    if (mode >= 0 && mode <= 3)
        g_pbvi_mode = mode;
    return old;
}

int pomdp_pbvi_initialize_gpu(const POMDP *pomdp, POMDPPBVIGPU *pbvi)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "pomdp_pbvi_initialize_gpu";
    if (!pomdp_device_ready(pomdp) || pbvi == nullptr || pbvi->GammaInitial == nullptr) {
        NOVA_ERR(fn, "Invalid arguments (the POMDP must be resident on the device and "
                     "GammaInitial set).");
        return NOVA_ERROR_INVALID_DATA;
    }
    if (!dense_mode(g_pbvi_mode) && !sparse_shape_ok(pomdp, g_pbvi_mode)) {
        NOVA_ERR(fn, g_pbvi_mode == 1
                 ? "Belief supports too long for exact mode (rz <= 16,384 and ns <= 512); "
                   "fast mode takes any support length."
                 : "r*m*rz*ns above 2^30: beliefs with full support use the dense modes 2/3.");
        return NOVA_ERROR_INVALID_DATA;
    }
    if (g_pgraph.exec && pbvi->d_alphaBA != nullptr && g_pgraph.key[9] == pbvi->d_alphaBA) {
        hipGraphExecDestroy(g_pgraph.exec);      // it holds the addresses freed below
        g_pgraph.exec = nullptr;
    }
    if (pbvi->d_alphaBA != nullptr) {
        dense_release(pbvi->d_alphaBA, fn);
        plain_forget(pbvi->d_alphaBA);
    }
    device_free_if_owned(pbvi->d_Gamma, fn, "Gamma");
    device_free_if_owned(pbvi->d_GammaPrime, fn, "Gamma (prime)");
    device_free_if_owned(pbvi->d_pi, fn, "pi");
    device_free_if_owned(pbvi->d_alphaBA, fn, "the PBVI workspace");
    pbvi->d_Gamma = pbvi->d_GammaPrime = nullptr;
    pbvi->d_pi = nullptr;
    pbvi->d_alphaBA = nullptr;
    pbvi->currentHorizon = 0;

    const size_t rn = (size_t)pomdp->r * pomdp->n;
    int rc = device_alloc((void **)&pbvi->d_Gamma, rn * sizeof(float), fn, "Gamma");
    if (rc == NOVA_SUCCESS)
        rc = device_alloc((void **)&pbvi->d_GammaPrime, rn * sizeof(float), fn, "Gamma (prime)");
    if (rc == NOVA_SUCCESS)
        rc = device_alloc((void **)&pbvi->d_pi, pomdp->r * sizeof(unsigned), fn, "pi");
    if (rc == NOVA_SUCCESS)
        rc = device_alloc((void **)&pbvi->d_alphaBA,
                          align256(inst_ws_bytes(pomdp)) + plan_bytes(1), fn,
                          "the PBVI workspace");
    if (rc != NOVA_SUCCESS)
        return rc;
    hipError_t e = hipMemcpy(pbvi->d_Gamma, pbvi->GammaInitial, rn * sizeof(float),
                             hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(pbvi->d_GammaPrime, pbvi->GammaInitial, rn * sizeof(float),
                      hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemset(pbvi->d_pi, 0, pomdp->r * sizeof(unsigned));
    if (e != hipSuccess)
        return hip_fail(e, fn, "Failed to copy memory from host to device for Gamma.",
                        NOVA_ERROR_MEMCPY_TO_DEVICE);
    hipStream_t st = stream();
    PbviInst I;
    fill_inst(pomdp, I, pbvi->d_Gamma, pbvi->d_GammaPrime, pbvi->d_pi,
              reinterpret_cast<char *>(pbvi->d_alphaBA), kNoHorizon);
    rc = build_ot(pomdp, I, st);
    if (rc != NOVA_SUCCESS)
        return hip_fail(hipGetLastError(), fn, "Failed to build the O*T products.", rc);
    PlainSolver rec;
    rec.ws = pbvi->d_alphaBA;
    rec.mode = g_pbvi_mode;
    plain_dims(pomdp, rec.dims);
    rec.generation = model_generation(false);
    if (dense_mode(g_pbvi_mode)) {
        rc = dense_setup(pomdp, pbvi->d_alphaBA, g_pbvi_mode, st, fn);
    } else {
        rec.jpt = tune_jpt(I, plain_plan_mem(pomdp, pbvi), st, g_pbvi_mode);
        rc = hip_fail(hipStreamSynchronize(st), fn, "Failed to prepare the PBVI workspace.",
                      NOVA_ERROR_KERNEL_EXECUTION);
    }
    if (rc == NOVA_SUCCESS)
        g_plain.push_back(rec);
    return rc;
}

int pomdp_pbvi_uninitialize_gpu(const POMDP *pomdp, POMDPPBVIGPU *pbvi)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "pomdp_pbvi_uninitialize_gpu";
    (void)pomdp;
    if (pbvi == nullptr) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    int rc = NOVA_SUCCESS;
    pbvi->currentHorizon = 0;
    if (g_pgraph.exec && g_pgraph.key[9] == pbvi->d_alphaBA) {
        hipGraphExecDestroy(g_pgraph.exec);
        g_pgraph.exec = nullptr;
    }
    if (pbvi->d_alphaBA != nullptr) {
        dense_release(pbvi->d_alphaBA, fn);
        plain_forget(pbvi->d_alphaBA);
    }
    if (device_free_if_owned(pbvi->d_Gamma, fn, "Gamma") != NOVA_SUCCESS) rc = NOVA_ERROR_DEVICE_FREE;
    if (device_free_if_owned(pbvi->d_GammaPrime, fn, "Gamma (prime)") != NOVA_SUCCESS) rc = NOVA_ERROR_DEVICE_FREE;
    if (device_free_if_owned(pbvi->d_pi, fn, "pi") != NOVA_SUCCESS) rc = NOVA_ERROR_DEVICE_FREE;
    if (device_free_if_owned(pbvi->d_alphaBA, fn, "the PBVI workspace") != NOVA_SUCCESS) rc = NOVA_ERROR_DEVICE_FREE;
    pbvi->d_Gamma = pbvi->d_GammaPrime = nullptr;
    pbvi->d_pi = nullptr;
    pbvi->d_alphaBA = nullptr;
    return rc;
}

int pomdp_pbvi_update_gpu(const POMDP *pomdp, POMDPPBVIGPU *pbvi)
{
    NOVA_AMD_API_LOCK();
    return pbvi_run_iters(pomdp, pbvi, 1, stream(), nullptr, nullptr, "pomdp_pbvi_update_gpu");
}

int pomdp_pbvi_get_policy_gpu(const POMDP *pomdp, POMDPPBVIGPU *pbvi, POMDPAlphaVectors *policy)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "pomdp_pbvi_get_policy_gpu";
    if (pomdp == nullptr || pbvi == nullptr || policy == nullptr) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    PlainSolver *S = nullptr;
    if (plain_sized(pomdp, pbvi, fn, &S) != NOVA_SUCCESS)
        return NOVA_ERROR_INVALID_DATA;
    if (pomdp_alpha_vectors_initialize(policy, pomdp->n, pomdp->m, pomdp->r) != NOVA_SUCCESS) {
        NOVA_ERR(fn, "Could not create the policy.");
        return NOVA_ERROR_POLICY_CREATION;
    }
    const float *G = (pbvi->currentHorizon % 2 == 0) ? pbvi->d_Gamma : pbvi->d_GammaPrime;
    hipError_t e = hipMemcpy(policy->Gamma, G, sizeof(float) * (size_t)pomdp->r * pomdp->n,
                             hipMemcpyDeviceToHost);
    if (e == hipSuccess)
        e = hipMemcpy(policy->pi, pbvi->d_pi, sizeof(unsigned) * pomdp->r, hipMemcpyDeviceToHost);
    return hip_fail(e, fn, "Failed to copy memory from device to host for Gamma/pi.",
                    NOVA_ERROR_MEMCPY_TO_HOST);
}

int pomdp_pbvi_execute_gpu(const POMDP *pomdp, POMDPPBVIGPU *pbvi, POMDPAlphaVectors *policy)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "pomdp_pbvi_execute_gpu";
    if (!pomdp_device_ready(pomdp) || pomdp->gamma < 0.0f || pomdp->gamma > 1.0f ||
            pomdp->horizon < 1 || pbvi == nullptr || pbvi->GammaInitial == nullptr ||
            policy == nullptr) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    if (pbvi->numThreads % 32 != 0) {
        NOVA_ERR(fn, "Invalid number of threads.");
        return NOVA_ERROR_INVALID_CUDA_PARAM;
    }
    int rc = pomdp_pbvi_initialize_gpu(pomdp, pbvi);
    if (rc != NOVA_SUCCESS)
        return rc;
    rc = pbvi_run_iters(pomdp, pbvi, pomdp->horizon, stream(), nullptr, nullptr, fn);
    if (rc == NOVA_SUCCESS)
        rc = pomdp_pbvi_get_policy_gpu(pomdp, pbvi, policy);
    const int rc2 = pomdp_pbvi_uninitialize_gpu(pomdp, pbvi);
    return (rc != NOVA_SUCCESS) ? rc : rc2;
}

int pomdp_pbvi_run_gpu(const POMDP *pomdp, POMDPPBVIGPU *pbvi, unsigned int iterations,
                       float *elapsedMs)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "pomdp_pbvi_run_gpu";
    if (!pomdp_device_ready(pomdp) || pbvi == nullptr || pbvi->d_Gamma == nullptr ||
            pbvi->d_GammaPrime == nullptr || pbvi->d_alphaBA == nullptr || iterations == 0) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    hipStream_t st = stream();
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (elapsedMs) {
        hipEventCreate(&t0);
        hipEventCreate(&t1);
    }
    int rc = pbvi_run_iters(pomdp, pbvi, iterations, st, t0, t1, fn);
    if (elapsedMs) {
        float ms = 0.0f;
        if (rc == NOVA_SUCCESS)
            hipEventElapsedTime(&ms, t0, t1);
        *elapsedMs = ms;
        hipEventDestroy(t0);
        hipEventDestroy(t1);
    }
    return rc;
}

int pomdp_pbvi_argmax_gpu(const POMDP *pomdp, const POMDPPBVIGPU *pbvi, int *jstar,
                          unsigned int *actions)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "pomdp_pbvi_argmax_gpu";
    if (pomdp == nullptr || pbvi == nullptr || pbvi->d_alphaBA == nullptr) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    PlainSolver *S = nullptr;
    if (plain_sized(pomdp, pbvi, fn, &S) != NOVA_SUCCESS)
        return NOVA_ERROR_INVALID_DATA;
    PbviInst I;
    inst_ws(pomdp, reinterpret_cast<char *>(pbvi->d_alphaBA), I);
    hipError_t e = hipSuccess;
    if (jstar)
        e = hipMemcpy(jstar, I.jstar, sizeof(int) * (size_t)pomdp->r * pomdp->m * pomdp->z,
                      hipMemcpyDeviceToHost);
    if (e == hipSuccess && actions)
        e = hipMemcpy(actions, I.act, sizeof(int) * pomdp->r, hipMemcpyDeviceToHost);
    return hip_fail(e, fn, "Failed to copy the argmax record to the host.",
                    NOVA_ERROR_MEMCPY_TO_HOST);
}

// One plain iteration with HIP events between its four stages (transpose, score, action,
// materialize); stageMs[4] receives their device times. Advances currentHorizon like
// pomdp_pbvi_update_gpu.
int pomdp_pbvi_scoring_variant_gpu(const POMDPPBVIGPU *pbvi, unsigned int *variant)
{
    NOVA_AMD_API_LOCK();
    if (pbvi == nullptr || pbvi->d_alphaBA == nullptr || variant == nullptr) {
        NOVA_ERR("pomdp_pbvi_scoring_variant_gpu", "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    *variant = dense_for(pbvi->d_alphaBA) ? 0u : plain_jpt(pbvi->d_alphaBA);
    return NOVA_SUCCESS;
}

int pomdp_pbvi_profile_gpu(const POMDP *pomdp, POMDPPBVIGPU *pbvi, float *stageMs)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "pomdp_pbvi_profile_gpu";
    if (!pomdp_device_ready(pomdp) || pbvi == nullptr || pbvi->d_alphaBA == nullptr ||
            stageMs == nullptr) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    PlainSolver *S = nullptr;
    int rc = plain_check(pomdp, pbvi, fn, &S);
    if (rc != NOVA_SUCCESS)
        return rc;
    hipStream_t st = stream();
    PbviPlan pl;
    pl.count = 1;
    pl.host.resize(1);
    fill_inst(pomdp, pl.host[0], pbvi->d_Gamma, pbvi->d_GammaPrime, pbvi->d_pi,
              reinterpret_cast<char *>(pbvi->d_alphaBA), kNoHorizon);
    pl.host[0].jpt = S->jpt;
    pl.host[0].fused = (dense_for(pbvi->d_alphaBA) == nullptr &&
                        fused_fits(pl.host[0], S->mode)) ? 1 : 0;
    rc = plan_upload(pl, plain_plan_mem(pomdp, pbvi), st, fn);
    if (rc != NOVA_SUCCESS)
        return rc;
    const unsigned flip = pbvi->currentHorizon % 2;
    const int mode = S->mode;
    const unsigned c = 1;
    const unsigned *tr = pl.dOff, *be = pl.dOff + (kVariants + 1) * (c + 1);
    hipEvent_t ev[5];
    for (auto &e : ev)
        hipEventCreate(&e);
    if (const DenseWs *Dp = dense_for(pbvi->d_alphaBA)) {
        // dense beliefs: convert, score (matrix cores), action, materialize
        const DenseWs &D = *Dp;
        const PbviInst &I = pl.host[0];
        const float *G = (flip == 0) ? I.G0 : I.G1;
        hipEventRecord(ev[0], st);
        dense_launch_score(D, G, st, ev[1]);
        hipEventRecord(ev[2], st);
        hipLaunchKernelGGL(dense_action, dim3(D.r), dim3(kPB), 0, st, D.m, D.z, D.r, D.tiles,
                           D.Rb, D.partVal, D.partJ, I.jstar, I.act);
        hipEventRecord(ev[3], st);
        hipLaunchKernelGGL(pbvi_materialize<false>, dim3(pl.totBe), dim3(kMB), pl.ldsBytes, st,
                           pl.dInst, be, c, 0u, flip);
        hipEventRecord(ev[4], st);
        hipError_t e = hipStreamSynchronize(st);
        for (int k = 0; k < 4; k++)
            hipEventElapsedTime(&stageMs[k], ev[k], ev[k + 1]);
        for (auto &x : ev)
            hipEventDestroy(x);
        if (e != hipSuccess)
            return hip_fail(e, fn, "Failed to run the profiled iteration.", NOVA_ERROR_KERNEL_EXECUTION);
        pbvi->currentHorizon++;
        return NOVA_SUCCESS;
    }
    if (pl.totFu > 0) {
        // a small instance: the whole iteration is one launch, reported as the score stage
        hipEventRecord(ev[0], st);
        hipEventRecord(ev[1], st);
        hipLaunchKernelGGL(pbvi_fused, dim3(pl.totFu), dim3(kPB), pl.fuLds, st, pl.dInst,
                           pl.dOff + (kVariants + 2) * (c + 1), c, 0u, flip);
        hipEventRecord(ev[2], st);
        hipEventRecord(ev[3], st);
        hipEventRecord(ev[4], st);
        hipError_t e = hipStreamSynchronize(st);
        for (int k = 0; k < 4; k++)
            hipEventElapsedTime(&stageMs[k], ev[k], ev[k + 1]);
        for (auto &x : ev)
            hipEventDestroy(x);
        if (e != hipSuccess)
            return hip_fail(e, fn, "Failed to run the profiled iteration.",
                            NOVA_ERROR_KERNEL_EXECUTION);
        pbvi->currentHorizon++;
        return NOVA_SUCCESS;
    }
    hipEventRecord(ev[0], st);
    hipLaunchKernelGGL(pbvi_transpose, dim3(pl.totTr), dim3(kPB), 0, st, pl.dInst, tr, c, 0u, flip);
    hipEventRecord(ev[1], st);
    const int v = jpt_slot(pl.host[0].jpt);
    const unsigned *sc = pl.dOff + (1 + v) * (c + 1);
    launch_score(pl.host[0].jpt, mode, dim3(pl.totSc[v]), pl.dInst, sc, c, 0u, st);
    hipEventRecord(ev[2], st);
    if (mode == 1)
        hipLaunchKernelGGL(pbvi_action<true>, dim3(pl.totBe), dim3(kPB), pl.actLds, st, pl.dInst, be, c, 0u, flip);
    else
        hipLaunchKernelGGL(pbvi_action<false>, dim3(pl.totBe), dim3(kPB), pl.actLds, st, pl.dInst, be, c, 0u, flip);
    hipEventRecord(ev[3], st);
    auto matFn = (mode == 1) ? pbvi_materialize<true> : pbvi_materialize<false>;
    hipLaunchKernelGGL(matFn, dim3(pl.totBe), dim3(kMB), pl.ldsBytes, st, pl.dInst, be, c, 0u, flip);
    hipEventRecord(ev[4], st);
    hipError_t e = hipStreamSynchronize(st);
    for (int k = 0; k < 4; k++)
        hipEventElapsedTime(&stageMs[k], ev[k], ev[k + 1]);
    for (auto &x : ev)
        hipEventDestroy(x);
    if (e != hipSuccess)
        return hip_fail(e, fn, "Failed to run the profiled iteration.", NOVA_ERROR_KERNEL_EXECUTION);
    pbvi->currentHorizon++;
    return NOVA_SUCCESS;
}

// Runs `iterations` plain iterations (from the current state) recording each one's argmax
// choices: d_recJ[(k*r + i)*z + o] and d_recA[k*r + i] for k < iterations (device arrays).
int pomdp_pbvi_run_recorded_gpu(const POMDP *pomdp, POMDPPBVIGPU *pbvi, unsigned int iterations,
        int *d_recJ, int *d_recA)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "pomdp_pbvi_run_recorded_gpu";
    if (!pomdp_device_ready(pomdp) || pbvi == nullptr || pbvi->d_alphaBA == nullptr ||
            iterations == 0 || d_recJ == nullptr || d_recA == nullptr ||
            dense_for(pbvi->d_alphaBA) != nullptr) {
        NOVA_ERR(fn, "Invalid arguments (recording needs a sparse-belief solver).");
        return NOVA_ERROR_INVALID_DATA;
    }
    PlainSolver *S = nullptr;
    int rc = plain_check(pomdp, pbvi, fn, &S);
    if (rc != NOVA_SUCCESS)
        return rc;
    hipStream_t st = stream();
    PbviPlan pl;
    pl.count = 1;
    pl.host.resize(1);
    fill_inst(pomdp, pl.host[0], pbvi->d_Gamma, pbvi->d_GammaPrime, pbvi->d_pi,
              reinterpret_cast<char *>(pbvi->d_alphaBA), kNoHorizon);
    pl.host[0].jpt = S->jpt;
    pl.host[0].recJ = d_recJ;
    pl.host[0].recA = d_recA;
    pl.host[0].fused = fused_fits(pl.host[0], S->mode) ? 1 : 0;
    rc = plan_upload(pl, plain_plan_mem(pomdp, pbvi), st, fn);
    if (rc != NOVA_SUCCESS)
        return rc;
    hipGraphExec_t exec = nullptr;
    rc = plan_graph(pl, iterations, pbvi->currentHorizon % 2, S->mode, st, &exec, fn);
    if (rc != NOVA_SUCCESS)
        return rc;
    hipError_t e = hipGraphLaunch(exec, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    hipGraphExecDestroy(exec);
    if (e != hipSuccess)
        return hip_fail(e, fn, "Failed to run the recorded iterations.", NOVA_ERROR_KERNEL_EXECUTION);
    pbvi->currentHorizon += iterations;
    return NOVA_SUCCESS;
}

// Batched PBVI: every POMDP must be resident on the device (pomdp_initialize_gpu) and each
// solver struct carry GammaInitial. Equivalent to pomdp_pbvi_execute_gpu on each pair, with
// all iterations of all instances captured in one graph (grouped launches per stage).
int pomdp_pbvi_batch_execute_gpu(const POMDP *const *pomdps, POMDPPBVIGPU *const *pbvis,
        unsigned int count, POMDPAlphaVectors *policies, float *elapsedMs)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "pomdp_pbvi_batch_execute_gpu";
    if (pomdps == nullptr || pbvis == nullptr || count == 0 || dense_mode(g_pbvi_mode)) {
        NOVA_ERR(fn, "Invalid arguments (batches use the sparse modes 0/1).");
        return NOVA_ERROR_INVALID_DATA;
    }
    for (unsigned b = 0; b < count; b++) {
        const POMDP *p = pomdps[b];
        if (!pomdp_device_ready(p) || p->horizon < 1 || pbvis[b] == nullptr ||
                pbvis[b]->GammaInitial == nullptr || !sparse_shape_ok(p, g_pbvi_mode)) {
            NOVA_ERR(fn, "Invalid instance (device-resident POMDP, GammaInitial and supported "
                         "belief support lengths required).");
            return NOVA_ERROR_INVALID_DATA;
        }
    }
    int rc = NOVA_SUCCESS;
    for (unsigned b = 0; b < count && rc == NOVA_SUCCESS; b++)
        rc = pomdp_pbvi_initialize_gpu(pomdps[b], pbvis[b]);
    void *planMem = nullptr;
    if (rc == NOVA_SUCCESS)
        rc = device_alloc(&planMem, plan_bytes(count), fn, "the batch plan");
    hipStream_t st = stream();
    PbviPlan pl;
    if (rc == NOVA_SUCCESS) {
        pl.count = count;
        pl.host.resize(count);
        for (unsigned b = 0; b < count; b++) {
            fill_inst(pomdps[b], pl.host[b], pbvis[b]->d_Gamma, pbvis[b]->d_GammaPrime,
                      pbvis[b]->d_pi, reinterpret_cast<char *>(pbvis[b]->d_alphaBA),
                      pomdps[b]->horizon);
            pl.host[b].jpt = plain_jpt(pbvis[b]->d_alphaBA);
            pl.host[b].fused = fused_fits(pl.host[b], g_pbvi_mode) ? 1 : 0;
        }
        rc = plan_upload(pl, planMem, st, fn);
    }
    hipGraphExec_t exec = nullptr;
    if (rc == NOVA_SUCCESS)
        rc = plan_graph(pl, pl.maxHorizon, 0, g_pbvi_mode, st, &exec, fn);
    if (rc == NOVA_SUCCESS) {
        hipEvent_t t0 = nullptr, t1 = nullptr;
        if (elapsedMs) {
            hipEventCreate(&t0);
            hipEventCreate(&t1);
            hipEventRecord(t0, st);
        }
        hipError_t e = hipGraphLaunch(exec, st);
        if (elapsedMs)
            hipEventRecord(t1, st);
        if (e == hipSuccess)
            e = hipStreamSynchronize(st);
        if (e != hipSuccess)
            rc = hip_fail(e, fn, "Failed to execute the batch graph.", NOVA_ERROR_KERNEL_EXECUTION);
        if (elapsedMs) {
            float ms = 0.0f;
            if (rc == NOVA_SUCCESS)
                hipEventElapsedTime(&ms, t0, t1);
            *elapsedMs = ms;
            hipEventDestroy(t0);
            hipEventDestroy(t1);
        }
        hipGraphExecDestroy(exec);
    }
    for (unsigned b = 0; b < count; b++) {
        if (rc == NOVA_SUCCESS) {
            pbvis[b]->currentHorizon = pomdps[b]->horizon;
            if (policies)
                rc = pomdp_pbvi_get_policy_gpu(pomdps[b], pbvis[b], &policies[b]);
        }
    }
    for (unsigned b = 0; b < count; b++)
        pomdp_pbvi_uninitialize_gpu(pomdps[b], pbvis[b]);
    if (planMem)
        device_free(planMem, fn, "the batch plan");
    return rc;
}

} // namespace nova

// ------------------------------------------------------------------------- Perseus (GPU)
//
// nova's randomized Perseus loop (src/pomdp/algorithms/pomdp_perseus_cpu.cpp:312-429) with its
// device work on the GPU: the backup of the sampled belief against the current alpha set runs
// on the PBVI kernels (one belief, rGamma alphas), and V(b) over the whole belief set is kept
// on the device (the new set only grows by one alpha per step, so V_{n+1}(b) is updated by one
// sparse dot per belief). The belief is drawn on the host with libc rand() by nova's formula,
// so with the same seed, exact mode reproduces nova's CPU results bit for bit.

namespace nova_amd {
namespace {

// v[i] = max(v[i], b_i . alpha) (first = initialise), b . alpha summed B*alpha in support order
__global__ __launch_bounds__(kPB) void perseus_vb_update(unsigned n, unsigned r, unsigned rz,
        const int *__restrict__ Z, const float *__restrict__ B, const float *__restrict__ alpha,
        int first, float *__restrict__ v)
{
    const unsigned i = blockIdx.x * kPB + threadIdx.x;
    if (i >= r)
        return;
    float d = 0.0f;
    for (unsigned k = 0; k < rz; k++) {
        const int s = Z[(size_t)i * rz + k];
        if (s < 0)
            break;
        d = __fadd_rn(d, __fmul_rn(B[(size_t)i * rz + k], alpha[s]));
    }
    if (first || v[i] < d)
        v[i] = d;
}

// V(b) over the rG alphas of G for every belief (first maximiser index in arg), -1e10 when empty
__global__ __launch_bounds__(kPB) void perseus_vb_full(unsigned n, unsigned r, unsigned rz,
        unsigned rG, const int *__restrict__ Z, const float *__restrict__ B,
        const float *__restrict__ G, float *__restrict__ v, int *__restrict__ arg)
{
    const unsigned i = blockIdx.x * kPB + threadIdx.x;
    if (i >= r)
        return;
    float best = kLow;
    int bj = 0;
    for (unsigned g = 0; g < rG; g++) {
        float d = 0.0f;
        for (unsigned k = 0; k < rz; k++) {
            const int s = Z[(size_t)i * rz + k];
            if (s < 0)
                break;
            d = __fadd_rn(d, __fmul_rn(B[(size_t)i * rz + k], G[(size_t)g * n + s]));
        }
        if (best < d) {
            best = d;
            bj = (int)g;
        }
    }
    v[i] = best;
    arg[i] = bj;
}

__global__ void fill_zero(size_t count, float *__restrict__ x, unsigned *__restrict__ a)
{
    const size_t q = blockIdx.x * (size_t)kPB + threadIdx.x;
    if (q < count)
        x[q] = 0.0f;
    if (q == 0)
        *a = 0;
}

} // namespace
} // namespace nova_amd

namespace nova {

using namespace nova_amd;

int pomdp_perseus_execute_gpu(const POMDP *pomdp, POMDPPerseusCPU *per, POMDPAlphaVectors *policy)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "pomdp_perseus_execute_gpu";
    if (!pomdp_device_ready(pomdp) || pomdp->gamma < 0.0f || pomdp->gamma > 1.0f ||
            pomdp->horizon < 1 || per == nullptr || per->GammaInitial == nullptr ||
            policy == nullptr || dense_mode(g_pbvi_mode) ||
            !sparse_shape_ok(pomdp, g_pbvi_mode)) {
        NOVA_ERR(fn, "Invalid arguments (device-resident POMDP, GammaInitial, sparse mode).");
        return NOVA_ERROR_INVALID_DATA;
    }
    const unsigned n = pomdp->n, r = pomdp->r, rz = pomdp->rz, m = pomdp->m;
    const size_t rn = (size_t)r * n;
    hipStream_t st = stream();
    // device state: the two alpha sets, the new alpha and its action, V(b) arrays, workspace
    float *GA = nullptr, *GB = nullptr, *alphaNew = nullptr, *vOld = nullptr, *vNew = nullptr;
    float *vTmp = nullptr;
    unsigned *actNew = nullptr;
    int *argOld = nullptr;
    void *ws = nullptr;
    int rc = device_alloc((void **)&GA, rn * 4, fn, "Perseus buffers");
    if (rc == NOVA_SUCCESS) rc = device_alloc((void **)&GB, rn * 4, fn, "Perseus buffers");
    if (rc == NOVA_SUCCESS) rc = device_alloc((void **)&alphaNew, (size_t)n * 4, fn, "Perseus buffers");
    if (rc == NOVA_SUCCESS) rc = device_alloc((void **)&actNew, 4, fn, "Perseus buffers");
    if (rc == NOVA_SUCCESS) rc = device_alloc((void **)&vOld, (size_t)r * 4, fn, "Perseus buffers");
    if (rc == NOVA_SUCCESS) rc = device_alloc((void **)&vNew, (size_t)r * 4, fn, "Perseus buffers");
    if (rc == NOVA_SUCCESS) rc = device_alloc((void **)&argOld, (size_t)r * 4, fn, "Perseus buffers");
    if (rc == NOVA_SUCCESS) rc = device_alloc((void **)&vTmp, (size_t)r * 4, fn, "Perseus buffers");
    if (rc == NOVA_SUCCESS)
        rc = device_alloc(&ws, align256(inst_ws_bytes(pomdp)) + plan_bytes(1), fn, "Perseus buffers");
    hipError_t e = hipSuccess;
    if (rc == NOVA_SUCCESS) {
        e = hipMemcpy(GA, per->GammaInitial, rn * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(GB, per->GammaInitial, rn * 4, hipMemcpyHostToDevice);
    }
    PbviInst I;
    if (rc == NOVA_SUCCESS && e == hipSuccess) {
        fill_inst(pomdp, I, GA, alphaNew, actNew, static_cast<char *>(ws), kNoHorizon);
        rc = build_ot(pomdp, I, st);
        if (rc == NOVA_SUCCESS)
            I.jpt = tune_jpt(I, static_cast<char *>(ws) + align256(inst_ws_bytes(pomdp)), st,
                             g_pbvi_mode);
    }
    if (rc == NOVA_SUCCESS && e != hipSuccess)
        rc = hip_fail(e, fn, "Failed to upload GammaInitial.", NOVA_ERROR_MEMCPY_TO_DEVICE);

    // host state, as POMDPPerseusCPU holds it (pomdp_perseus_cpu.cpp:183-221)
    unsigned horizon = 0, rGam = 0, rGamP = 0;
    std::vector<unsigned> piA(r, 0), piB(r, 0), BTilde(r);
    for (unsigned i = 0; i < r; i++)
        BTilde[i] = i;
    unsigned rTilde = r;
    std::vector<float> hOld(r), hNew(r);
    std::vector<int> hArg(r);
    void *planMem = static_cast<char *>(ws) + align256(inst_ws_bytes(pomdp));
    const unsigned gridB = (r + kPB - 1) / kPB;
    bool horizonStart = true;
    float *G = GA, *Gp = GB;
    unsigned *rG = &rGam, *rGp = &rGamP;
    std::vector<unsigned> *pi = &piA, *piP = &piB;
    while (rc == NOVA_SUCCESS && horizon < pomdp->horizon) {
        const bool even = (horizon % 2 == 0);
        G = even ? GA : GB;
        Gp = even ? GB : GA;
        rG = even ? &rGam : &rGamP;
        rGp = even ? &rGamP : &rGam;
        pi = even ? &piA : &piB;
        piP = even ? &piB : &piA;
        if (horizonStart) {
            // V_n(b) for every belief, fixed during this horizon; GT of the current set
            hipLaunchKernelGGL(perseus_vb_full, dim3(gridB), dim3(kPB), 0, st, n, r, rz, *rG,
                               pomdp->d_Z, pomdp->d_B, G, vOld, argOld);
            I.r = std::max(*rG, 1u);
            I.rb = 1;
            I.G0 = G;
            PbviPlan pl;
            pl.count = 1;
            pl.host.assign(1, I);
            rc = plan_upload(pl, planMem, st, fn);
            if (rc != NOVA_SUCCESS)
                break;
            hipLaunchKernelGGL(pbvi_transpose, dim3(pl.totTr), dim3(kPB), 0, st, pl.dInst,
                               pl.dOff, 1u, 0u, 0u);
            e = hipMemcpy(hOld.data(), vOld, r * 4, hipMemcpyDeviceToHost);
            if (e == hipSuccess)
                e = hipMemcpy(hArg.data(), argOld, r * 4, hipMemcpyDeviceToHost);
            if (e != hipSuccess) {
                rc = hip_fail(e, fn, "Failed to read V(b).", NOVA_ERROR_MEMCPY_TO_HOST);
                break;
            }
            horizonStart = false;
        }
        // draw a belief still to improve (nova's formula, index kept below rTilde)
        unsigned tIdx = (unsigned)((float)std::rand() / (float)RAND_MAX * (float)rTilde);
        if (tIdx >= rTilde)
            tIdx = rTilde - 1;
        const unsigned bIndex = BTilde[tIdx];
        // backup of belief bIndex against the current set: a one-belief PBVI iteration
        PbviPlan pl;
        pl.count = 1;
        PbviInst J = I;
        J.r = std::max(*rG, 1u);
        J.rb = 1;
        J.G0 = G;
        J.G1 = alphaNew;
        J.pi = actNew;
        J.Z = pomdp->d_Z + (size_t)bIndex * rz;
        J.B = pomdp->d_B + (size_t)bIndex * rz;
        if (J.pairSp != nullptr) {
            J.pairSp += (size_t)bIndex * m * J.kmax;
            J.pairW += (size_t)bIndex * m * J.kmax;
            J.pairCnt += (size_t)bIndex * m;
        }
        pl.host.assign(1, J);
        rc = plan_upload(pl, planMem, st, fn);
        if (rc != NOVA_SUCCESS)
            break;
        hipLaunchKernelGGL(fill_zero, dim3((n + kPB - 1) / kPB), dim3(kPB), 0, st, (size_t)n,
                           alphaNew, actNew);
        {
            const unsigned c = 1;
            const int v = jpt_slot(J.jpt);
            const unsigned *sc = pl.dOff + (1 + v) * (c + 1), *be = pl.dOff + (kVariants + 1) * (c + 1);
            launch_score(J.jpt, g_pbvi_mode, dim3(pl.totSc[v]), pl.dInst, sc, c, 0u, st);
            if (g_pbvi_mode == 1)
                hipLaunchKernelGGL(pbvi_action<true>, dim3(1), dim3(kPB), pl.actLds, st, pl.dInst, be, c, 0u, 0u);
            else
                hipLaunchKernelGGL(pbvi_action<false>, dim3(1), dim3(kPB), pl.actLds, st, pl.dInst, be, c, 0u, 0u);
            auto matFn = (g_pbvi_mode == 1) ? pbvi_materialize<true> : pbvi_materialize<false>;
            hipLaunchKernelGGL(matFn, dim3(1), dim3(kMB), pl.ldsBytes, st, pl.dInst, be, c, 0u, 0u);
        }
        // b . alpha_new (support order), read for bIndex only
        hipLaunchKernelGGL(perseus_vb_update, dim3(gridB), dim3(kPB), 0, st, n, r, rz,
                           pomdp->d_Z, pomdp->d_B, alphaNew, 1, vTmp);
        float bDot = 0.0f;
        unsigned aNew = 0;
        e = hipMemcpy(&bDot, vTmp + bIndex, 4, hipMemcpyDeviceToHost);
        if (e == hipSuccess)
            e = hipMemcpy(&aNew, actNew, 4, hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            rc = hip_fail(e, fn, "Failed to run the backup.", NOVA_ERROR_KERNEL_EXECUTION);
            break;
        }
        const float Vnb = (*rG > 0) ? hOld[bIndex] : kLow;
        const unsigned oldArg = (*rG > 0) ? (unsigned)hArg[bIndex] : 0u;
        float *dst = Gp + (size_t)(*rGp) * n;
        if (bDot >= Vnb) {
            e = hipMemcpyAsync(dst, alphaNew, (size_t)n * 4, hipMemcpyDeviceToDevice, st);
            (*piP)[*rGp] = aNew;
        } else {
            e = hipMemcpyAsync(dst, G + (size_t)oldArg * n, (size_t)n * 4, hipMemcpyDeviceToDevice, st);
            (*piP)[*rGp] = (*pi)[oldArg];
        }
        (*rGp)++;
        if (*rGp > r) {
            NOVA_ERR(fn, "Out of memory. Too many alpha-vectors added.");
            rc = NOVA_ERROR_OUT_OF_MEMORY;
            break;
        }
        // V_{n+1}(b) grows by the appended alpha: a running maximum per belief
        hipLaunchKernelGGL(perseus_vb_update, dim3(gridB), dim3(kPB), 0, st, n, r, rz,
                           pomdp->d_Z, pomdp->d_B, dst, (*rGp == 1) ? 1 : 0, vNew);
        e = hipMemcpy(hNew.data(), vNew, r * 4, hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            rc = hip_fail(e, fn, "Failed to read V(b).", NOVA_ERROR_MEMCPY_TO_HOST);
            break;
        }
        rTilde = 0;
        for (unsigned i = 0; i < r; i++) {
            const float vo = (*rG > 0) ? hOld[i] : kLow;
            if (hNew[i] < vo)
                BTilde[rTilde++] = i;
        }
        if (rTilde == 0) {
            horizon++;
            *rG = 0;
            rTilde = r;
            for (unsigned i = 0; i < r; i++)
                BTilde[i] = i;
            horizonStart = true;
        }
    }
    if (rc == NOVA_SUCCESS) {
        const bool even = (horizon % 2 == 0);
        const unsigned rOut = even ? rGam : rGamP;
        if (pomdp_alpha_vectors_initialize(policy, n, m, rOut) != NOVA_SUCCESS) {
            NOVA_ERR(fn, "Could not create the policy.");
            rc = NOVA_ERROR_POLICY_CREATION;
        } else {
            e = hipMemcpy(policy->Gamma, even ? GA : GB, sizeof(float) * (size_t)rOut * n,
                          hipMemcpyDeviceToHost);
            std::memcpy(policy->pi, (even ? piA : piB).data(), sizeof(unsigned) * rOut);
            if (e != hipSuccess)
                rc = hip_fail(e, fn, "Failed to copy the policy.", NOVA_ERROR_MEMCPY_TO_HOST);
        }
    }
    void *bufs[] = {GA, GB, alphaNew, actNew, vOld, vNew, vTmp, argOld, ws};
    for (void *q : bufs)
        if (q)
            device_free(q, fn, "Perseus buffers");
    return rc;
}

} // namespace nova
