// nova-mi355x: executing alpha-vector policies over many beliefs at once on one MI355X.
//
// 1. Value and action of k dense beliefs (src/pomdp/policies/pomdp_alpha_vectors.cpp:52-79 per
//    belief: V(b) = max_j Gamma_j . b over the r alphas, the action pi[j*] of the first
//    maximiser, V = -1e10 and action 0 when no alpha beats -1e10). Three arithmetic modes:
//      exact - every dot product summed over s = 0..n-1 with each product and sum rounded
//              separately, as the reference: V and the action equal nova's bit for bit. A
//              vector-ALU tile kernel: 64 beliefs x 64 alphas per workgroup, each lane a 4 x 4
//              block, both operands staged state-major in LDS 32 states at a time.
//      fp32  - the dense GEMM of dense_gemm.h on the matrix cores (v_mfma_f32_32x32x2_f32):
//              fp32 products, fp32 sums in the matrix cores' order.
//      bf16  - dense_score_bf16 of dense_gemm.h: bf16 operands in the k-blocked layout,
//              copied straight into LDS, fp32 accumulation (v_mfma_f32_32x32x16_bf16): about
//              3 significant digits per operand.
//    All modes write per-(belief, alpha tile) maxima that one merge kernel reduces in tile
//    order (the earlier tile wins ties, i.e. the first maximiser overall).
// 2. Belief updates of k (belief, action, observation) triples
//    (src/pomdp/utilities/pomdp_model_cpu.cpp:89-137 per triple): b'(s') = O[a,s',o] *
//    sum_s T[s,a,s'] b(s), normalised by its sum over s'. The sum over predecessors runs over a
//    list ordered by (s, successor slot) -- the order in which the reference's scatter adds to
//    b'(s') -- and the normaliser is summed serially over s' ascending, so b' equals nova's
//    bit for bit. Three kernels: the unnormalised b' (predecessor entries read coalesced, their
//    products summed per s' from LDS, b gathered from LDS, O from an observation-major copy
//    built once per model), the normalisers (one serial chain per triple fed through LDS), and
//    the division. A triple whose normaliser is below 1e-9 in magnitude keeps the unnormalised
//    b' and reports NOVA_WARNING_INVALID_BELIEF, as the reference returns it.
#include <algorithm>
#include <climits>
#include <cstring>
#include <mutex>
#include <vector>

#include <nova/amd/nova_amd.h>
#include <nova/pomdp/pomdp.h>

#include "dense_gemm.h"
#include "nova_hip.h"

namespace nova_amd {

namespace {

constexpr int kVT = 64;          // beliefs and alphas per exact tile
constexpr int kVS = 32;          // states per LDS step of the exact tile
constexpr int kVThreads = 256;   // 16 x 16 lanes, each a 4 x 4 block of (belief, alpha)

// Exact tile: partial first maxima of b_i . Gamma_j over the tile's alphas, j ascending, with
// strict '>' from -1e10 (alpha_vectors.cpp:64-76).
__global__ __launch_bounds__(kVThreads) void alpha_value_exact(unsigned k, unsigned n,
        unsigned r, const float *__restrict__ B, const float *__restrict__ G,
        float *__restrict__ partV, int *__restrict__ partJ)
{
    __shared__ __attribute__((aligned(16))) float bs[kVS][kVT + 4];
    __shared__ __attribute__((aligned(16))) float gs[kVS][kVT + 4];
    const unsigned i0 = blockIdx.x * kVT, j0 = blockIdx.y * kVT, t = threadIdx.x;
    const unsigned bi = t / 16, aj = t % 16;
    float acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++)
            acc[x][y] = 0.0f;
    // stage 64 rows x 32 states of each operand, transposed to state-major; entries past n
    // (and rows past k or r) are zeros, whose products add +0 and leave every sum as is. The
    // next step's values are requested into registers before the current step's products,
    // so their latency hides behind the arithmetic.
    constexpr int kPer = kVT * kVS / kVThreads;      // values per lane and operand per step
    float bpre[kPer], gpre[kPer];
    auto loadStep = [&](unsigned s0) {
#pragma unroll
        for (int u = 0; u < kPer; u++) {
            const unsigned x = t + u * kVThreads, row = x / kVS, s = x % kVS;
            const bool inS = s0 + s < n;
            bpre[u] = (inS && i0 + row < k) ? B[(size_t)(i0 + row) * n + s0 + s] : 0.0f;
            gpre[u] = (inS && j0 + row < r) ? G[(size_t)(j0 + row) * n + s0 + s] : 0.0f;
        }
    };
    loadStep(0);
    for (unsigned s0 = 0; s0 < n; s0 += kVS) {
#pragma unroll
        for (int u = 0; u < kPer; u++) {
            const unsigned x = t + u * kVThreads, row = x / kVS, s = x % kVS;
            bs[s][row] = bpre[u];
            gs[s][row] = gpre[u];
        }
        __syncthreads();
        if (s0 + kVS < n)
            loadStep(s0 + kVS);
#pragma unroll 4
        for (int s = 0; s < kVS; s++) {
            const float4 bv = *reinterpret_cast<const float4 *>(&bs[s][bi * 4]);
            const float4 gv = *reinterpret_cast<const float4 *>(&gs[s][aj * 4]);
            const float b4[4] = {bv.x, bv.y, bv.z, bv.w};
            const float g4[4] = {gv.x, gv.y, gv.z, gv.w};
#pragma unroll
            for (int x = 0; x < 4; x++)
#pragma unroll
                for (int y = 0; y < 4; y++)
                    acc[x][y] = __fadd_rn(acc[x][y], __fmul_rn(g4[y], b4[x]));
        }
        __syncthreads();
    }
#pragma unroll
    for (int x = 0; x < 4; x++) {
        float bv = kLow;
        int bj = INT_MAX;
#pragma unroll
        for (int y = 0; y < 4; y++) {
            const unsigned j = j0 + aj * 4 + y;
            if (j < r && acc[x][y] > bv) {
                bv = acc[x][y];
                bj = (int)j;
            }
        }
        // the 16 lanes of this belief row are 16 consecutive lanes of one wave
        for (int off = 8; off >= 1; off >>= 1) {
            const float ov = __shfl_xor(bv, off, 16);
            const int oj = __shfl_xor(bj, off, 16);
            if (better(ov, oj, bv, bj)) {
                bv = ov;
                bj = oj;
            }
        }
        const unsigned i = i0 + bi * 4 + x;
        if (aj == 0 && i < k) {
            partV[(size_t)blockIdx.y * k + i] = bv;
            partJ[(size_t)blockIdx.y * k + i] = bj;
        }
    }
}

// Merge of the alpha tiles per belief, tile 0 first: V, the winning alpha (or -1) and its
// action (0 when none beat -1e10, as the reference leaves a).
__global__ __launch_bounds__(256) void alpha_value_merge(unsigned k, unsigned tiles,
        const float *__restrict__ partV, const int *__restrict__ partJ,
        const unsigned *__restrict__ pi, float *__restrict__ V, unsigned *__restrict__ act,
        int *__restrict__ jOut)
{
    const unsigned i = blockIdx.x * 256 + threadIdx.x;
    if (i >= k)
        return;
    float bv = kLow;
    int bj = INT_MAX;
    for (unsigned tl = 0; tl < tiles; tl++) {
        const float v = partV[(size_t)tl * k + i];
        const int j = partJ[(size_t)tl * k + i];
        if (better(v, j, bv, bj)) {
            bv = v;
            bj = j;
        }
    }
    const bool none = (bj == INT_MAX) || !(bv > kLow);
    V[i] = none ? kLow : bv;
    act[i] = none ? 0u : pi[bj];
    if (jOut)
        jOut[i] = none ? -1 : bj;
}

__global__ __launch_bounds__(256) void fill_ones(unsigned count, float *__restrict__ x)
{
    const unsigned i = blockIdx.x * 256 + threadIdx.x;
    if (i < count)
        x[i] = 1.0f;
}

// Belief updates in three kernels, each sum in the reference's order:
//   belief_update_batch: one workgroup per triple q; b is staged in LDS. Wave w takes blocks
//     of 64 consecutive states s'. The block's predecessor entries (contiguous in the lists)
//     are read 64 at a time, coalesced, and their products T * b(s) formed into the wave's LDS
//     scratch; lane t then adds the products of s' = block + t in list order, multiplies by
//     O[a,s',o] (from the observation-major copy Ot[a][o][s']) and writes the unnormalised
//     b'(s'). When n exceeds the LDS buffer b is read from global memory.
//   belief_norms: the normaliser of triple q, summed serially over s' by one lane; 16 triples
//     per workgroup, their rows streamed through double-buffered LDS by the other waves.
//   belief_divide: b'_q / c_q for the triples whose normaliser is usable.
constexpr int kUB = 1024;
constexpr int kUpdWaves = kUB / 64;
constexpr int kUpdScratch = 1024;                // product slots per wave
constexpr int kUpdLdsFloats = 40 * 1024 - kUpdWaves * kUpdScratch;   // 160 KB in all
constexpr int kUpdBatch = 8;                     // list entries per lane requested together

// IDX: the predecessor's index type, 16 bits for n <= 65,536 (6 bytes per list entry, not 8)
template <typename IDX>
__global__ __launch_bounds__(kUB) void belief_update_batch(unsigned n, unsigned m, unsigned z,
        const unsigned *__restrict__ pstart, const IDX *__restrict__ pS,
        const float *__restrict__ pT, const float *__restrict__ Ot, const float *__restrict__ B,
        const unsigned *__restrict__ rowOf, const unsigned *__restrict__ act,
        const unsigned *__restrict__ obs, float *__restrict__ Bp, int useLds,
        const unsigned *__restrict__ order)
{
    extern __shared__ float smem[];
    // order: the triples grouped by action, so that the workgroups running together on an XCD
    // stream the same action's predecessor list and find it in their L2
    const unsigned q = order ? order[blockIdx.x] : blockIdx.x;
    const unsigned t = threadIdx.x, lane = t % 64, w = t / 64;
    const unsigned a = act[q], o = obs[q];
    if (a >= m || o >= z)
        return;                                  // belief_norms reports it
    const float *b = B + (size_t)(rowOf ? rowOf[q] : q) * n;   // rowOf: belief row of triple q
    float *out = Bp + (size_t)q * n;
    float *row = smem;                                           // n floats when useLds
    float *scr = smem + (useLds ? n : 0) + w * kUpdScratch;
    if (useLds) {
        for (unsigned s = t; s < n; s += kUB)
            row[s] = b[s];
        __syncthreads();
    }
    const float *src = useLds ? row : b;
    const unsigned *ps = pstart + (size_t)a * (n + 1);
    const float *Oao = Ot + ((size_t)a * z + o) * n;
    for (unsigned s0 = w * 64; s0 < n; s0 += kUpdWaves * 64) {
        const unsigned s1 = min(s0 + 64, n), sp = s0 + lane;
        const unsigned xa = ps[s0], xb = ps[s1];
        const unsigned lo = sp < n ? ps[sp] : xb, hi = sp < n ? ps[sp + 1] : xb;
        float u = 0.0f;
        for (unsigned xs = xa; xs < xb; xs += kUpdScratch) {
            const unsigned xe = min(xs + (unsigned)kUpdScratch, xb);
            // the lane's entries of the chunk kUpdBatch at a time, their predecessors and
            // factors requested together (indices clamped into the chunk, the extra products
            // not stored)
            for (unsigned x0 = xs + lane; x0 < xe; x0 += 64 * kUpdBatch) {
                IDX si[kUpdBatch];
                float tv[kUpdBatch];
#pragma unroll
                for (int k = 0; k < kUpdBatch; k++) {
                    const unsigned x = min(x0 + 64u * k, xe - 1);
                    si[k] = pS[x];
                    tv[k] = pT[x];
                }
#pragma unroll
                for (int k = 0; k < kUpdBatch; k++) {
                    const unsigned x = x0 + 64u * k;
                    const float pr = __fmul_rn(tv[k], src[si[k]]);
                    if (x < xe)
                        scr[x - xs] = pr;
                }
            }
            // the wave's own LDS writes, then reads by other lanes of the wave
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const unsigned e1 = min(hi, xe);
            for (unsigned x = max(lo, xs); x < e1; x++)
                u = __fadd_rn(u, scr[x - xs]);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        if (sp < n)
            out[sp] = __fmul_rn(u, Oao[sp]);
    }
}

// The k triples grouped by action (a counting sort in one workgroup; invalid actions last).
// The order within an action is whatever the atomics give: each triple's result does not
// depend on when it runs.
constexpr unsigned kOrderMaxM = 4096;

__global__ __launch_bounds__(1024) void order_by_action(unsigned k, unsigned m,
        const unsigned *__restrict__ act, unsigned *__restrict__ order)
{
    __shared__ unsigned cnt[kOrderMaxM + 1];
    const unsigned t = threadIdx.x;
    for (unsigned x = t; x <= m; x += 1024)
        cnt[x] = 0;
    __syncthreads();
    for (unsigned q = t; q < k; q += 1024)
        atomicAdd(&cnt[min(act[q], m)], 1u);
    __syncthreads();
    if (t == 0) {
        unsigned run = 0;
        for (unsigned x = 0; x <= m; x++) {
            const unsigned c = cnt[x];
            cnt[x] = run;
            run += c;
        }
    }
    __syncthreads();
    for (unsigned q = t; q < k; q += 1024)
        order[atomicAdd(&cnt[min(act[q], m)], 1u)] = q;
}

constexpr unsigned kNormT = 16;          // triples per workgroup of belief_norms
constexpr unsigned kNormC = 512;         // states per chunk
constexpr unsigned kNormLoaders = 256;   // lanes of waves 1-4, 32 values each per chunk

__global__ __launch_bounds__(64 + kNormLoaders) void belief_norms(unsigned n, unsigned m,
        unsigned z, unsigned k, const unsigned *__restrict__ act,
        const unsigned *__restrict__ obs, const float *__restrict__ Bp,
        float *__restrict__ norms, int *__restrict__ status)
{
    // waves 1-4 stream chunks of the 16 rows into one buffer (coalesced, all 32 loads of a
    // lane in flight together) while lanes 0-15 of wave 0 add the other buffer's chunk into
    // their triple's sum, in order; rows padded to 513 floats so that the 16 summing lanes
    // read distinct banks
    constexpr unsigned kPer = kNormT * kNormC / kNormLoaders;
    __shared__ float buf[2][kNormT][kNormC + 1];
    const unsigned t = threadIdx.x, q0 = blockIdx.x * kNormT;
    auto load = [&](unsigned c0, unsigned which) {
        float v[kPer];
#pragma unroll
        for (unsigned u = 0; u < kPer; u++) {
            const unsigned e = (t - 64) + kNormLoaders * u, j = e / kNormC, i = e % kNormC;
            v[u] = (q0 + j < k && c0 + i < n) ? Bp[(size_t)(q0 + j) * n + c0 + i] : 0.0f;
        }
#pragma unroll
        for (unsigned u = 0; u < kPer; u++) {
            const unsigned e = (t - 64) + kNormLoaders * u;
            buf[which][e / kNormC][e % kNormC] = v[u];
        }
    };
    if (t >= 64)
        load(0, 0);
    __syncthreads();
    float c = 0.0f;
    unsigned cur = 0;
    for (unsigned c0 = 0; c0 < n; c0 += kNormC, cur ^= 1) {
        if (t >= 64 && c0 + kNormC < n)
            load(c0 + kNormC, cur ^ 1);
        if (t < kNormT) {
            const unsigned cnt = min(kNormC, n - c0);
            for (unsigned i = 0; i < cnt; i++)
                c = __fadd_rn(c, buf[cur][t][i]);
        }
        __syncthreads();
    }
    const unsigned q = q0 + t;
    if (t >= kNormT || q >= k)
        return;
    if (act[q] >= m || obs[q] >= z) {
        status[q] = NOVA_ERROR_INVALID_DATA;
        return;
    }
    norms[q] = c;
    // a normaliser below the tolerance leaves b' unnormalised (pomdp_model_cpu.cpp:125-130)
    status[q] = ((double)fabsf(c) < NOVA_FLT_ERR_TOL) ? NOVA_WARNING_INVALID_BELIEF
                                                      : NOVA_SUCCESS;
}

__global__ __launch_bounds__(256) void belief_divide(unsigned n, const float *__restrict__ norms,
        const int *__restrict__ status, float *__restrict__ Bp)
{
    const unsigned q = blockIdx.x;
    if (status[q] != NOVA_SUCCESS)
        return;
    const float c = norms[q];
    float *out = Bp + (size_t)q * n;
    for (unsigned sp = threadIdx.x; sp < n; sp += 256)
        out[sp] = __fdiv_rn(out[sp], c);
}

__global__ __launch_bounds__(256) void observation_major(unsigned n, unsigned z, size_t count,
        const float *__restrict__ O, float *__restrict__ Ot)
{
    for (size_t x = blockIdx.x * (size_t)256 + threadIdx.x; x < count;
            x += (size_t)gridDim.x * 256) {
        const size_t sp = x % n, ao = x / n;
        const size_t a = ao / z, o = ao % z;
        Ot[x] = O[(a * n + sp) * z + o];
    }
}

// A device buffer kept between calls and grown on demand (one per use, all devices share
// the process's current device; calls are serialised by the mutex of their use).
struct Arena {
    void *mem = nullptr;
    size_t bytes = 0;
    int device = -1;
};

int arena_get(Arena &A, size_t bytes, const char *fn, void *&out)
{
    int dev = 0;
    hipGetDevice(&dev);
    if (A.mem != nullptr && (A.bytes < bytes || A.device != dev)) {
        hipDeviceSynchronize();
        device_free(A.mem, fn, "policy workspace");
        A.mem = nullptr;
        A.bytes = 0;
    }
    if (A.mem == nullptr) {
        const int rc = device_alloc(&A.mem, bytes, fn, "policy workspace");
        if (rc != NOVA_SUCCESS)
            return rc;
        A.bytes = bytes;
        A.device = dev;
    }
    out = A.mem;
    return NOVA_SUCCESS;
}

std::mutex g_valueMutex;
Arena g_valueArena;
struct DenseShape {
    unsigned k = 0, n = 0, r = 0;
    int precision = -1;
    void *mem = nullptr;
} g_valueShape;

std::mutex g_updateMutex;
Arena g_updateArena;                 // the normalisers and action order of a batch of belief updates
// predecessor lists of the last model seen, keyed by its device arrays
struct PredLists {
    const void *dS = nullptr, *dT = nullptr, *dO = nullptr;
    unsigned long long generation = 0;      // model_generation(true) when built
    unsigned n = 0, m = 0, ns = 0, z = 0;
    void *mem = nullptr;
    size_t bytes = 0;
    unsigned *pstart = nullptr;
    int *pS = nullptr;                  // predecessors when n > 65,536
    unsigned short *pS16 = nullptr;     // predecessors when n <= 65,536
    float *pT = nullptr;
    float *Ot = nullptr;      // O observation-major: Ot[a][o][s'] = O[a][s'][o]
} g_pred;

inline size_t up256(size_t x) { return (x + 255) / 256 * 256; }

} // namespace

// Predecessor lists per action of p's model, (s, T[s,a,l]) for every valid successor entry,
// grouped by successor s' and ordered by (s, l) within each group, built on the host from p's
// host arrays; and the observation-major copy of O, built on the device. Kept until another
// model (other device arrays or sizes) is passed, or S, T or O are uploaded or released again.
static int pred_lists(const nova::POMDP *p, const char *fn)
{
    if (g_pred.mem != nullptr && g_pred.dS == p->d_S && g_pred.dT == p->d_T &&
            g_pred.dO == p->d_O && g_pred.n == p->n && g_pred.m == p->m && g_pred.ns == p->ns &&
            g_pred.z == p->z && g_pred.generation == model_generation(true))
        return NOVA_SUCCESS;
    const unsigned n = p->n, m = p->m, ns = p->ns;
    std::vector<unsigned> pstart((size_t)m * (n + 1), 0);
    for (unsigned s = 0; s < n; s++)
        for (unsigned a = 0; a < m; a++)
            for (unsigned l = 0; l < ns; l++) {
                const int sp = p->S[idx3(s, a, l, m, ns)];
                if (sp < 0)
                    break;
                pstart[(size_t)a * (n + 1) + sp + 1]++;
            }
    size_t total = 0;
    for (unsigned a = 0; a < m; a++) {
        unsigned *rowp = &pstart[(size_t)a * (n + 1)];
        rowp[0] = (unsigned)total;
        for (unsigned sp = 0; sp < n; sp++)
            rowp[sp + 1] += rowp[sp];
        total = rowp[n];
    }
    std::vector<int> pS(std::max<size_t>(total, 1));
    std::vector<float> pT(std::max<size_t>(total, 1));
    std::vector<unsigned> fill(pstart);
    for (unsigned s = 0; s < n; s++)
        for (unsigned a = 0; a < m; a++)
            for (unsigned l = 0; l < ns; l++) {
                const size_t x = idx3(s, a, l, m, ns);
                const int sp = p->S[x];
                if (sp < 0)
                    break;
                const unsigned at = fill[(size_t)a * (n + 1) + sp]++;
                pS[at] = (int)s;
                pT[at] = p->T[x];
            }
    if (g_pred.mem != nullptr) {
        hipDeviceSynchronize();
        device_free(g_pred.mem, fn, "predecessor lists");
        g_pred = PredLists();
    }
    const bool narrow = n <= 65536;
    std::vector<unsigned short> pS16(narrow ? pS.size() : 0);
    for (size_t x = 0; x < pS16.size(); x++)
        pS16[x] = (unsigned short)pS[x];
    const size_t idxBytes = narrow ? sizeof(unsigned short) : sizeof(int);
    const size_t b0 = up256(sizeof(unsigned) * pstart.size()), b1 = up256(idxBytes * pS.size());
    const size_t b2 = up256(sizeof(float) * pT.size());
    const size_t otCount = (size_t)m * p->z * n;
    const size_t bytes = b0 + b1 + b2 + up256(sizeof(float) * otCount);
    void *mem = nullptr;
    int rc = device_alloc(&mem, bytes, fn, "predecessor lists");
    if (rc != NOVA_SUCCESS)
        return rc;
    char *c = static_cast<char *>(mem);
    hipError_t e = hipMemcpy(c, pstart.data(), sizeof(unsigned) * pstart.size(),
                             hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(c + b0, narrow ? (const void *)pS16.data() : (const void *)pS.data(),
                      idxBytes * pS.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(c + b0 + b1, pT.data(), sizeof(float) * pT.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        device_free(mem, fn, "predecessor lists");
        return hip_fail(e, fn, "Failed to copy the predecessor lists.",
                        NOVA_ERROR_MEMCPY_TO_DEVICE);
    }
    float *Ot = reinterpret_cast<float *>(c + b0 + b1 + b2);
    hipStream_t st = stream();
    hipLaunchKernelGGL(observation_major,
                       dim3((unsigned)std::min<size_t>((otCount + 255) / 256, 65536)), dim3(256),
                       0, st, n, p->z, otCount, p->d_O, Ot);
    e = hipGetLastError();
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        device_free(mem, fn, "predecessor lists");
        return hip_fail(e, fn, "Failed to build the observation-major copy of O.",
                        NOVA_ERROR_KERNEL_EXECUTION);
    }
    g_pred.dS = p->d_S;
    g_pred.dT = p->d_T;
    g_pred.dO = p->d_O;
    g_pred.generation = model_generation(true);
    g_pred.n = n;
    g_pred.m = m;
    g_pred.ns = ns;
    g_pred.z = p->z;
    g_pred.mem = mem;
    g_pred.bytes = bytes;
    g_pred.pstart = reinterpret_cast<unsigned *>(c);
    if (narrow)
        g_pred.pS16 = reinterpret_cast<unsigned short *>(c + b0);
    else
        g_pred.pS = reinterpret_cast<int *>(c + b0);
    g_pred.pT = reinterpret_cast<float *>(c + b0 + b1);
    g_pred.Ot = Ot;
    return NOVA_SUCCESS;
}

} // namespace nova_amd

namespace nova_amd {

// tau(b, a, o) for k triples whose beliefs are rows rowOf[q] of B (rowOf null: row q), bit for
// bit nova's pomdp_belief_update_cpu; shared by the batched entry point and the GPU PEMA.
int belief_update_rows(const nova::POMDP *pomdp, unsigned k, const float *d_B,
                       const unsigned *d_rowOf, const unsigned *d_action,
                       const unsigned *d_observation, float *d_Bp, int *d_status,
                       const char *fn)
{
    const long long bad = first_bad_successor(pomdp->S, (size_t)pomdp->n * pomdp->m, pomdp->ns,
                                              pomdp->n);
    if (bad >= 0) {
        NOVA_ERR(fn, "A successor index lies outside [0, n).");
        return NOVA_ERROR_INVALID_DATA;
    }
    std::lock_guard<std::mutex> lock(g_updateMutex);
    int rc = pred_lists(pomdp, fn);
    if (rc != NOVA_SUCCESS)
        return rc;
    void *mem = nullptr;
    const size_t normBytes = (sizeof(float) * (size_t)k + 255) / 256 * 256;
    rc = arena_get(g_updateArena, normBytes + sizeof(unsigned) * (size_t)k, fn, mem);
    if (rc != NOVA_SUCCESS)
        return rc;
    float *norms = static_cast<float *>(mem);
    unsigned *order = pomdp->m <= kOrderMaxM
                      ? reinterpret_cast<unsigned *>(static_cast<char *>(mem) + normBytes)
                      : nullptr;
    hipStream_t st = stream();
    const unsigned n = pomdp->n;
    if (order)
        hipLaunchKernelGGL(order_by_action, dim3(1), dim3(1024), 0, st, k, pomdp->m, d_action,
                           order);
    const int useLds = n <= (unsigned)kUpdLdsFloats ? 1 : 0;
    const size_t lds = sizeof(float) * ((useLds ? n : 0) + (size_t)kUpdWaves * kUpdScratch);
    if (g_pred.pS16 != nullptr) {
        hipFuncSetAttribute((const void *)belief_update_batch<unsigned short>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        hipLaunchKernelGGL(belief_update_batch<unsigned short>, dim3(k), dim3(kUB), lds, st, n,
                           pomdp->m, pomdp->z, g_pred.pstart, g_pred.pS16, g_pred.pT, g_pred.Ot,
                           d_B, d_rowOf, d_action, d_observation, d_Bp, useLds, order);
    } else {
        hipFuncSetAttribute((const void *)belief_update_batch<int>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        hipLaunchKernelGGL(belief_update_batch<int>, dim3(k), dim3(kUB), lds, st, n, pomdp->m,
                           pomdp->z, g_pred.pstart, g_pred.pS, g_pred.pT, g_pred.Ot, d_B,
                           d_rowOf, d_action, d_observation, d_Bp, useLds, order);
    }
    hipLaunchKernelGGL(belief_norms, dim3((k + kNormT - 1) / kNormT), dim3(64 + kNormLoaders), 0, st, n,
                       pomdp->m, pomdp->z, k, d_action, d_observation, d_Bp, norms, d_status);
    hipLaunchKernelGGL(belief_divide, dim3(k), dim3(256), 0, st, n, norms, d_status, d_Bp);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e != hipSuccess)
        return hip_fail(e, fn, "Kernel execution failed.", NOVA_ERROR_KERNEL_EXECUTION);
    return NOVA_SUCCESS;
}

} // namespace nova_amd

using namespace nova_amd;

extern "C" int nova_amd_alpha_vectors_value_and_action_gpu(unsigned int n, unsigned int r,
        const float *d_Gamma, const unsigned int *d_pi, unsigned int k, const float *d_B,
        int precision, float *d_V, unsigned int *d_action, int *d_alpha)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "nova_amd_alpha_vectors_value_and_action_gpu";
    if (n == 0 || r == 0 || k == 0 || d_Gamma == nullptr || d_pi == nullptr ||
            d_B == nullptr || d_V == nullptr || d_action == nullptr || precision < 0 ||
            precision > 2 || (r + kVT - 1) / kVT > 65535u) {
        NOVA_ERR(fn, "Invalid arguments (n, r, k > 0, r <= 4,194,240; device arrays; "
                     "precision 0, 1 or 2).");
        return NOVA_ERROR_INVALID_DATA;
    }
    std::lock_guard<std::mutex> lock(g_valueMutex);
    hipStream_t st = stream();
    const bool exact = precision == 0;
    const unsigned tileW = exact ? (unsigned)kVT : (unsigned)kGT;
    const unsigned colTiles = (r + tileW - 1) / tileW;
    const unsigned rowTiles = (k + kGT - 1) / kGT;
    const size_t inSize = precision == 2 ? sizeof(__bf16) : sizeof(float);
    const unsigned kpad = (n + kGKPad - 1) / kGKPad * kGKPad;
    const size_t bPart = up256((size_t)colTiles * k * sizeof(float));
    size_t bytes = 2 * bPart;
    size_t bA = 0, bG = 0, bS = 0;
    if (!exact) {
        bA = up256((size_t)rowTiles * kGT * kpad * inSize);
        bG = up256((size_t)colTiles * kGT * kpad * inSize);
        bS = up256(sizeof(float) * kpad);
        bytes += bA + bG + bS;
    }
    void *mem = nullptr;
    int rc = arena_get(g_valueArena, bytes, fn, mem);
    if (rc != NOVA_SUCCESS)
        return rc;
    char *c = static_cast<char *>(mem);
    float *partV = reinterpret_cast<float *>(c);
    int *partJ = reinterpret_cast<int *>(c + bPart);
    hipError_t e = hipSuccess;
    if (exact) {
        g_valueShape = DenseShape();   // the partials may overwrite the dense operands' padding
        hipLaunchKernelGGL(alpha_value_exact, dim3((k + kVT - 1) / kVT, colTiles),
                           dim3(kVThreads), 0, st, k, n, r, d_B, d_Gamma, partV, partJ);
    } else {
        void *A = c + 2 * bPart, *Gi = c + 2 * bPart + bA;
        float *ones = reinterpret_cast<float *>(c + 2 * bPart + bA + bG);
        // the zero padding of the row-major layouts survives between calls of one shape and
        // layout; another shape or layout clears it (the k-blocked layout is written whole)
        const bool blocked = precision == 2 && !dense_legacy_bf16();
        const int layout = blocked ? 3 : precision;
        const bool same = g_valueShape.mem == mem && g_valueShape.k == k &&
                          g_valueShape.n == n && g_valueShape.r == r &&
                          g_valueShape.precision == layout;
        if (!same) {
            e = hipMemsetAsync(A, 0, bA + bG, st);
            hipLaunchKernelGGL(fill_ones, dim3((kpad + 255) / 256), dim3(256), 0, st, kpad, ones);
            g_valueShape.mem = mem;
            g_valueShape.k = k;
            g_valueShape.n = n;
            g_valueShape.r = r;
            g_valueShape.precision = layout;
        }
        const size_t kn = (size_t)k * n, rn = (size_t)r * n;
        const unsigned ga = (unsigned)std::min<size_t>((kn + kGConvThreads - 1) / kGConvThreads,
                                                       65536);
        const unsigned gg = (unsigned)std::min<size_t>((rn + kGConvThreads - 1) / kGConvThreads,
                                                       65536);
        const dim3 sg(rowTiles * colTiles, 1);
        if (blocked) {
            // k-blocked operands, every padding element written by the conversion
            const size_t ta = (size_t)rowTiles * kGT * kpad, tg = (size_t)colTiles * kGT * kpad;
            const unsigned ca = (unsigned)std::min<size_t>(
                (ta / 8 + kGConvThreads - 1) / kGConvThreads, 65536);
            const unsigned cg = (unsigned)std::min<size_t>(
                (tg / 8 + kGConvThreads - 1) / kGConvThreads, 65536);
            hipFuncSetAttribute((const void *)dense_score_bf16,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)kGDlLdsBytes);
            hipLaunchKernelGGL(dense_convert_blk, dim3(ca), dim3(kGConvThreads), 0, st, n, kpad,
                               kn, (size_t)rowTiles * kGT, d_B, static_cast<__bf16 *>(A));
            hipLaunchKernelGGL(dense_convert_blk, dim3(cg), dim3(kGConvThreads), 0, st, n, kpad,
                               rn, (size_t)colTiles * kGT, d_Gamma, static_cast<__bf16 *>(Gi));
            hipLaunchKernelGGL(dense_score_bf16, sg, dim3(kGThreads), kGDlLdsBytes, st, k, r,
                               kpad, 1u, 1u, rowTiles, colTiles, 0u, (size_t)0,
                               static_cast<const __bf16 *>(A), static_cast<const __bf16 *>(Gi),
                               partV, partJ);
        } else if (precision == 2) {
            hipFuncSetAttribute((const void *)dense_score<__bf16>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)kGLdsBytes);
            hipLaunchKernelGGL(dense_convert<__bf16>, dim3(ga), dim3(kGConvThreads), 0, st, n,
                               kpad, kn, d_B, static_cast<__bf16 *>(A));
            hipLaunchKernelGGL(dense_convert<__bf16>, dim3(gg), dim3(kGConvThreads), 0, st, n,
                               kpad, rn, d_Gamma, static_cast<__bf16 *>(Gi));
            hipLaunchKernelGGL(dense_score<__bf16>, sg, dim3(kGThreads), kGLdsBytes, st, k, r,
                               kpad, 1u, 1u, rowTiles, colTiles, static_cast<const __bf16 *>(A),
                               static_cast<const __bf16 *>(Gi), ones, partV, partJ);
        } else {
            hipFuncSetAttribute((const void *)dense_score<float>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)kGLdsBytes);
            hipLaunchKernelGGL(dense_convert<float>, dim3(ga), dim3(kGConvThreads), 0, st, n,
                               kpad, kn, d_B, static_cast<float *>(A));
            hipLaunchKernelGGL(dense_convert<float>, dim3(gg), dim3(kGConvThreads), 0, st, n,
                               kpad, rn, d_Gamma, static_cast<float *>(Gi));
            hipLaunchKernelGGL(dense_score<float>, sg, dim3(kGThreads), kGLdsBytes, st, k, r,
                               kpad, 1u, 1u, rowTiles, colTiles, static_cast<const float *>(A),
                               static_cast<const float *>(Gi), ones, partV, partJ);
        }
    }
    hipLaunchKernelGGL(alpha_value_merge, dim3((k + 255) / 256), dim3(256), 0, st, k, colTiles,
                       partV, partJ, d_pi, d_V, d_action, d_alpha);
    if (e == hipSuccess)
        e = hipGetLastError();
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e != hipSuccess)
        return hip_fail(e, fn, "Kernel execution failed.", NOVA_ERROR_KERNEL_EXECUTION);
    return NOVA_SUCCESS;
}

extern "C" int nova_amd_belief_update_batch_gpu(const nova::POMDP *pomdp, unsigned int k,
        const float *d_B, const unsigned int *d_action, const unsigned int *d_observation,
        float *d_Bp, int *d_status)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "nova_amd_belief_update_batch_gpu";
    if (pomdp == nullptr || pomdp->n == 0 || pomdp->ns == 0 || pomdp->m == 0 ||
            pomdp->z == 0 || pomdp->S == nullptr || pomdp->T == nullptr ||
            pomdp->d_S == nullptr || pomdp->d_O == nullptr || k == 0 || d_B == nullptr ||
            d_action == nullptr || d_observation == nullptr || d_Bp == nullptr ||
            d_status == nullptr) {
        NOVA_ERR(fn, "Invalid arguments (a device-resident POMDP with its host arrays, "
                     "k > 0, device arrays).");
        return NOVA_ERROR_INVALID_DATA;
    }
    return belief_update_rows(pomdp, k, d_B, nullptr, d_action, d_observation, d_Bp, d_status,
                              fn);
}

// ---- the exact value/action without allocation
//
// The per-(belief, alpha tile) partials of the exact path live in `scratch`, a static device
// array of kNoAllocScratchBytes that every GPU the library runs on holds until the process
// ends. Host code never passes the array's name to a kernel (on the host it is only a stand-in
// for the symbol): its device address is looked up once per device with
// hipGetSymbolAddress and kept in g_scratchDev. Every partial the merge reads is written by the
// tile kernel of the same call first, so nothing in it has to start at zero. The contents
// persist between calls; calls on one GPU share the array, and overlapping ones would race
// (the library lock serialises the calls of one process).
constexpr size_t kNoAllocScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kScratchDevices = 64;
void *g_scratchDev[kScratchDevices];
hipError_t g_scratchErr[kScratchDevices];
std::once_flag g_scratchOnce[kScratchDevices];

// Device address of `scratch` on the current device (nullptr and an error message on failure).
void *scratch_address(const char *fn)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kScratchDevices) {
        hip_fail(e != hipSuccess ? e : hipErrorInvalidDevice, fn,
                 "No usable current device for the scratch array.", NOVA_ERROR_INVALID_DATA);
        return nullptr;
    }
    std::call_once(g_scratchOnce[dev], [dev]() {
        void *p = nullptr;
        g_scratchErr[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratchDev[dev] = g_scratchErr[dev] == hipSuccess ? p : nullptr;
    });
    if (g_scratchDev[dev] == nullptr)
        hip_fail(g_scratchErr[dev] != hipSuccess ? g_scratchErr[dev] : hipErrorInvalidSymbol, fn,
                 "Could not look up the device address of the scratch array.",
                 NOVA_ERROR_KERNEL_EXECUTION);
    return g_scratchDev[dev];
}
} // namespace

extern "C" int nova_amd_alpha_vectors_value_and_action_noalloc_gpu(unsigned int n,
        unsigned int r, const float *d_Gamma, const unsigned int *d_pi, unsigned int k,
        const float *d_B, float *d_V, unsigned int *d_action, int *d_alpha)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "nova_amd_alpha_vectors_value_and_action_noalloc_gpu";
    if (n == 0 || r == 0 || k == 0 || d_Gamma == nullptr || d_pi == nullptr ||
            d_B == nullptr || d_V == nullptr || d_action == nullptr ||
            (r + kVT - 1) / kVT > 65535u) {
        NOVA_ERR(fn, "Invalid arguments (n, r, k > 0, r <= 4,194,240; device arrays).");
        return NOVA_ERROR_INVALID_DATA;
    }
    const size_t colTiles = (r + kVT - 1) / kVT;
    const size_t bPart = up256(colTiles * k * sizeof(float));
    if (2 * bPart > kNoAllocScratchBytes) {
        NOVA_ERR(fn, "ceil(r / 64) * k exceeds the 16 MiB scratch array (at most 2,097,152 "
                     "(belief, 64-alpha tile) partials per call).");
        return NOVA_ERROR_INVALID_DATA;
    }
    char *c = static_cast<char *>(scratch_address(fn));
    if (c == nullptr)
        return NOVA_ERROR_KERNEL_EXECUTION;
    hipStream_t st = stream();
    float *partV = reinterpret_cast<float *>(c);
    int *partJ = reinterpret_cast<int *>(c + bPart);
    hipLaunchKernelGGL(alpha_value_exact, dim3((k + kVT - 1) / kVT, (unsigned)colTiles),
                       dim3(kVThreads), 0, st, k, n, r, d_B, d_Gamma, partV, partJ);
    hipLaunchKernelGGL(alpha_value_merge, dim3((k + 255) / 256), dim3(256), 0, st, k,
                       (unsigned)colTiles, partV, partJ, d_pi, d_V, d_action, d_alpha);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e != hipSuccess)
        return hip_fail(e, fn, "Kernel execution failed.", NOVA_ERROR_KERNEL_EXECUTION);
    return NOVA_SUCCESS;
}
