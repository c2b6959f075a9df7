// nova-mi355x: PEMA expansion with its (belief, action, observation) work on the GPU.
//
// Same result as pomdp_expand_pema_cpu bit for bit, in both alpha choices (nova_amd_pema_mode):
// src/pomdp/utilities/pomdp_expand_cpu.cpp:269-466. Per chunk of set beliefs, for every triple
// (i, a, o):
//   * Pr(o | b_i, a), summed over the support and successors in the reference's order;
//   * b' = tau(b_i, a, o) through the bit-exact batched belief update;
//   * the set belief closest to b' in L1: a vector-ALU tile kernel (64 triples x 64 set beliefs
//     per workgroup, each lane a 4 x 4 block, operands staged state-major in LDS) summing
//     |b'(s) - c(s)| over s in order, then the first minimum over the set;
//   * eps(b') against the alpha chosen for that closest belief (chosen once per set belief),
//     and Pr * eps.
// The per-action and per-belief choices (first maxima, sums over o in order) and the running
// bound over the set run on the host in float, as the reference's loop does; the winning
// successor is copied back and appended with pomdp_add_new_raw_beliefs_cpu. The cost is
// dominated by the closest-belief search: r * m * z * r * n absolute differences.
#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

#include <nova/amd/nova_amd.h>
#include <nova/pomdp/utilities/pomdp_model_cpu.h>

#include "nova_hip.h"

namespace nova_amd {

int belief_update_rows(const nova::POMDP *pomdp, unsigned k, const float *d_B,
                       const unsigned *d_rowOf, const unsigned *d_action,
                       const unsigned *d_observation, float *d_Bp, int *d_status,
                       const char *fn);
int pema_mode_current();

namespace {

constexpr int kPT = 64;          // triples and set beliefs per closest-search tile
constexpr int kPS = 32;          // states per LDS step
constexpr int kPThreads = 256;

// Dense rows of the belief set: zeros, then the valid prefix of (Z, B) in order (a repeated
// state keeps its last value, as the reference's construction does).
__global__ __launch_bounds__(256) void pema_dense(unsigned n, unsigned rz,
        const int *__restrict__ Z, const float *__restrict__ B, float *__restrict__ D)
{
    const unsigned i = blockIdx.x;
    float *row = D + (size_t)i * n;
    for (unsigned s = threadIdx.x; s < n; s += 256)
        row[s] = 0.0f;
    __syncthreads();
    if (threadIdx.x == 0)
        for (unsigned k = 0; k < rz; k++) {
            const int s = Z[(size_t)i * rz + k];
            if (s < 0)
                break;
            row[s] = B[(size_t)i * rz + k];
        }
}

// Triple q of a chunk starting at belief i0: belief i0 + q / (m z), action (q / z) % m,
// observation q % z. Pr(o | b, a) = sum over the support and successors of (T * O) * b
// (pomdp_expand_cpu.cpp:324-340).
__global__ __launch_bounds__(256) void pema_triples(unsigned K, unsigned i0, unsigned n,
        unsigned ns, unsigned m, unsigned z, unsigned rz, const int *__restrict__ S,
        const float *__restrict__ T, const float *__restrict__ O, const int *__restrict__ Z,
        const float *__restrict__ B, unsigned *__restrict__ rowOf, unsigned *__restrict__ act,
        unsigned *__restrict__ obs, float *__restrict__ pr)
{
    const unsigned q = blockIdx.x * 256 + threadIdx.x;
    if (q >= K)
        return;
    const unsigned i = i0 + q / (m * z), a = (q / z) % m, o = q % z;
    float p = 0.0f;
    for (unsigned k = 0; k < rz; k++) {
        const int s = Z[(size_t)i * rz + k];
        if (s < 0)
            break;
        const size_t row = ((size_t)s * m + a) * ns;
        for (unsigned l = 0; l < ns; l++) {
            const int sp = S[row + l];
            if (sp < 0)
                break;
            p = __fadd_rn(p, __fmul_rn(__fmul_rn(T[row + l], O[((size_t)a * n + sp) * z + o]),
                                       B[(size_t)i * rz + k]));
        }
    }
    rowOf[q] = i;
    act[q] = a;
    obs[q] = o;
    pr[q] = p;
}

__device__ __forceinline__ bool closer(float v, int j, float bv, int bj)
{
    return v < bv || (v == bv && j < bj);
}

// Partial first minima of sum_s |b'_q(s) - D_j(s)| over the tile's set beliefs, strict '<'
// from 1e10 (pomdp_expand_cpu.cpp:367-383).
__global__ __launch_bounds__(kPThreads) void pema_closest(unsigned K, unsigned n, unsigned r,
        const float *__restrict__ Bp, const float *__restrict__ D, float *__restrict__ partD,
        int *__restrict__ partJ)
{
    __shared__ __attribute__((aligned(16))) float xs[kPS][kPT + 4];
    __shared__ __attribute__((aligned(16))) float ds[kPS][kPT + 4];
    const unsigned q0 = blockIdx.x * kPT, j0 = blockIdx.y * kPT, t = threadIdx.x;
    const unsigned bi = t / 16, aj = t % 16;
    float acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++)
            acc[x][y] = 0.0f;
    for (unsigned s0 = 0; s0 < n; s0 += kPS) {
        // entries past n (and rows past K or r) are zeros: |0 - 0| adds +0
        for (unsigned x = t; x < kPT * kPS; x += kPThreads) {
            const unsigned row = x / kPS, s = x % kPS;
            const bool inS = s0 + s < n;
            xs[s][row] = (inS && q0 + row < K) ? Bp[(size_t)(q0 + row) * n + s0 + s] : 0.0f;
            ds[s][row] = (inS && j0 + row < r) ? D[(size_t)(j0 + row) * n + s0 + s] : 0.0f;
        }
        __syncthreads();
#pragma unroll 4
        for (int s = 0; s < kPS; s++) {
            const float4 xv = *reinterpret_cast<const float4 *>(&xs[s][bi * 4]);
            const float4 dv = *reinterpret_cast<const float4 *>(&ds[s][aj * 4]);
            const float x4[4] = {xv.x, xv.y, xv.z, xv.w};
            const float d4[4] = {dv.x, dv.y, dv.z, dv.w};
#pragma unroll
            for (int x = 0; x < 4; x++)
#pragma unroll
                for (int y = 0; y < 4; y++)
                    acc[x][y] = __fadd_rn(acc[x][y], fabsf(__fsub_rn(x4[x], d4[y])));
        }
        __syncthreads();
    }
#pragma unroll
    for (int x = 0; x < 4; x++) {
        float bv = kHigh;
        int bj = INT_MAX;
#pragma unroll
        for (int y = 0; y < 4; y++) {
            const unsigned j = j0 + aj * 4 + y;
            if (j < r && acc[x][y] < bv) {
                bv = acc[x][y];
                bj = (int)j;
            }
        }
        for (int off = 8; off >= 1; off >>= 1) {
            const float ov = __shfl_xor(bv, off, 16);
            const int oj = __shfl_xor(bj, off, 16);
            if (closer(ov, oj, bv, bj)) {
                bv = ov;
                bj = oj;
            }
        }
        const unsigned q = q0 + bi * 4 + x;
        if (aj == 0 && q < K) {
            partD[(size_t)blockIdx.y * K + q] = bv;
            partJ[(size_t)blockIdx.y * K + q] = bj;
        }
    }
}

// The alpha of nova's loop for each set belief c (pomdp_expand_cpu.cpp:391-402): the last
// alpha with some product alpha(s) * c(s) below -1e10, else 0.
__global__ __launch_bounds__(256) void pema_alpha_nova(unsigned n, unsigned nA,
        const float *__restrict__ G, const float *__restrict__ D, int *__restrict__ alphaOf)
{
    __shared__ int best;
    const unsigned j = blockIdx.x, t = threadIdx.x;
    if (t == 0)
        best = 0;
    __syncthreads();
    const float *c = D + (size_t)j * n;
    int mine = -1;
    for (unsigned a = t; a < nA; a += 256) {
        bool hit = false;
        for (unsigned s = 0; s < n && !hit; s++)
            hit = __fmul_rn(G[(size_t)a * n + s], c[s]) < kLow;
        if (hit)
            mine = (int)a;
    }
    if (mine >= 0)
        atomicMax(&best, mine);
    __syncthreads();
    if (t == 0)
        alphaOf[j] = best;
}

// eps(b') and Pr * eps per triple (pomdp_expand_cpu.cpp:404-423); valid = the update succeeded.
__global__ __launch_bounds__(256) void pema_eps(unsigned K, unsigned n, unsigned tiles,
        float gamma, float Rmax, float Rmin, const float *__restrict__ Bp,
        const float *__restrict__ D, const unsigned *__restrict__ rowOf,
        const int *__restrict__ status, const float *__restrict__ partD,
        const int *__restrict__ partJ, const int *__restrict__ alphaOf,
        const float *__restrict__ G, const float *__restrict__ pr, float *__restrict__ val,
        int *__restrict__ valid)
{
    const unsigned q = blockIdx.x * 256 + threadIdx.x;
    if (q >= K)
        return;
    if (status[q] != NOVA_SUCCESS) {
        valid[q] = 0;
        val[q] = 0.0f;
        return;
    }
    float bd = kHigh;
    int bj = INT_MAX;
    for (unsigned tl = 0; tl < tiles; tl++) {
        const float d = partD[(size_t)tl * K + q];
        const int j = partJ[(size_t)tl * K + q];
        if (closer(d, j, bd, bj)) {
            bd = d;
            bj = j;
        }
    }
    const unsigned closest = (bj == INT_MAX) ? 0u : (unsigned)bj;
    const float *c = D + (size_t)closest * n;
    const float *b = D + (size_t)rowOf[q] * n;
    const float *bp = Bp + (size_t)q * n;
    const float *alpha = G + (size_t)alphaOf[closest] * n;
    const float hi = __fdiv_rn(Rmax, __fsub_rn(1.0f, gamma));
    const float lo = __fdiv_rn(Rmin, __fsub_rn(1.0f, gamma));
    float eps = 0.0f;
    for (unsigned s = 0; s < n; s++) {
        const float lim = (bp[s] >= b[s]) ? hi : lo;
        eps = __fadd_rn(eps, __fmul_rn(__fsub_rn(lim, alpha[s]), __fsub_rn(bp[s], c[s])));
    }
    val[q] = __fmul_rn(pr[q], eps);
    valid[q] = 1;
}

size_t up256(size_t x) { return (x + 255) / 256 * 256; }

} // namespace

} // namespace nova_amd

using namespace nova_amd;

namespace nova {

int pomdp_expand_pema_gpu(POMDP *pomdp, const POMDPAlphaVectors *policy)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "pomdp_expand_pema_gpu";
    if (pomdp == nullptr || pomdp->n == 0 || pomdp->ns == 0 || pomdp->m == 0 ||
            pomdp->z == 0 || pomdp->r == 0 || pomdp->rz == 0 || pomdp->S == nullptr ||
            pomdp->T == nullptr || pomdp->O == nullptr || pomdp->R == nullptr ||
            pomdp->Z == nullptr || pomdp->B == nullptr || pomdp->d_S == nullptr ||
            pomdp->d_T == nullptr || pomdp->d_O == nullptr || policy == nullptr ||
            policy->r == 0 || policy->n != pomdp->n || policy->Gamma == nullptr) {
        NOVA_ERR(fn, "Invalid arguments (a POMDP whose model is resident on the device, with "
                     "its host arrays, and a policy over its states).");
        return NOVA_ERROR_INVALID_DATA;
    }
    // the kernels index with these unchecked: states of the supports and successor lists
    // must lie in [0, n)
    if (first_bad_successor(pomdp->Z, pomdp->r, pomdp->rz, pomdp->n) >= 0 ||
            first_bad_successor(pomdp->S, (size_t)pomdp->n * pomdp->m, pomdp->ns,
                                pomdp->n) >= 0) {
        NOVA_ERR(fn, "A belief-support or successor index lies outside [0, n).");
        return NOVA_ERROR_INVALID_DATA;
    }
    if ((pomdp->r + kPT - 1) / kPT > 65535u) {
        NOVA_ERR(fn, "Belief set too large (r <= 4,194,240).");
        return NOVA_ERROR_INVALID_DATA;
    }
    const unsigned n = pomdp->n, m = pomdp->m, z = pomdp->z, r = pomdp->r, rz = pomdp->rz;
    const unsigned nA = policy->r;
    const int mode = pema_mode_current();
    float Rmax = kLow, Rmin = kHigh;
    for (size_t x = 0; x < (size_t)n * m; x++) {
        if (Rmax < pomdp->R[x])
            Rmax = pomdp->R[x];
        if (Rmin > pomdp->R[x])
            Rmin = pomdp->R[x];
    }
    // beliefs per chunk: the chunk's successors take at most about 512 MB
    const size_t perBelief = (size_t)m * z * n * sizeof(float);
    const unsigned C = (unsigned)std::max<size_t>(1, std::min<size_t>(r, ((size_t)512 << 20) /
                                                                       perBelief));
    const unsigned Kmax = C * m * z, tiles = (r + kPT - 1) / kPT;
    // the belief set comes from the host arrays (the authoritative copy, which this call
    // extends): the device Z/B may lag behind it
    const size_t bD = up256(sizeof(float) * (size_t)r * n), bG = up256(sizeof(float) * nA * n);
    const size_t bA = up256(sizeof(int) * r), bBp = up256(sizeof(float) * (size_t)Kmax * n);
    const size_t bK4 = up256(4 * (size_t)Kmax), bP = up256(4 * (size_t)tiles * Kmax);
    const size_t bZ = up256(sizeof(int) * (size_t)r * rz);
    const size_t bS = up256(4 * (size_t)r) * 2 + up256(4 * (size_t)nA);
    const size_t total = bD + bG + bA + bBp + 7 * bK4 + 2 * bP + 2 * bZ + bS;
    void *mem = nullptr;
    int rc = device_alloc(&mem, total, fn, "PEMA workspace");
    if (rc != NOVA_SUCCESS)
        return rc;
    char *c = static_cast<char *>(mem);
    float *D = reinterpret_cast<float *>(c);                       c += bD;
    float *G = reinterpret_cast<float *>(c);                       c += bG;
    int *alphaOf = reinterpret_cast<int *>(c);                     c += bA;
    float *Bp = reinterpret_cast<float *>(c);                      c += bBp;
    unsigned *rowOf = reinterpret_cast<unsigned *>(c);             c += bK4;
    unsigned *act = reinterpret_cast<unsigned *>(c);               c += bK4;
    unsigned *obs = reinterpret_cast<unsigned *>(c);               c += bK4;
    float *pr = reinterpret_cast<float *>(c);                      c += bK4;
    int *status = reinterpret_cast<int *>(c);                      c += bK4;
    float *val = reinterpret_cast<float *>(c);                     c += bK4;
    int *valid = reinterpret_cast<int *>(c);                       c += bK4;
    float *partD = reinterpret_cast<float *>(c);                   c += bP;
    int *partJ = reinterpret_cast<int *>(c);                       c += bP;
    int *Zw = reinterpret_cast<int *>(c);                          c += bZ;
    float *Bw = reinterpret_cast<float *>(c);                      c += bZ;
    float *scratchV = reinterpret_cast<float *>(c);                c += up256(4 * (size_t)r);
    unsigned *scratchA = reinterpret_cast<unsigned *>(c);          c += up256(4 * (size_t)r);
    unsigned *zeroPi = reinterpret_cast<unsigned *>(c);            // one entry per alpha

    hipStream_t st = stream();
    std::vector<float> Bnew(n, 0.0f);
    hipError_t e = hipMemcpy(G, policy->Gamma, sizeof(float) * nA * n, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(Zw, pomdp->Z, sizeof(int) * (size_t)r * rz, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(Bw, pomdp->B, sizeof(float) * (size_t)r * rz, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemsetAsync(zeroPi, 0, sizeof(unsigned) * nA, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(pema_dense, dim3(r), dim3(256), 0, st, n, rz, Zw, Bw, D);
        if (mode == 0)
            hipLaunchKernelGGL(pema_alpha_nova, dim3(r), dim3(256), 0, st, n, nA, G, D,
                               alphaOf);
        e = hipStreamSynchronize(st);
    }
    if (e == hipSuccess && mode == 1) {
        // the first maximiser of alpha . c with the reference's float sums; -1 (none above
        // -1e10) stands for alpha 0
        rc = nova_amd_alpha_vectors_value_and_action_gpu(n, nA, G, zeroPi, r, D, 0, scratchV,
                                                        scratchA, alphaOf);
        if (rc == NOVA_SUCCESS) {
            std::vector<int> h(r);
            e = hipMemcpy(h.data(), alphaOf, sizeof(int) * r, hipMemcpyDeviceToHost);
            for (auto &x : h)
                x = std::max(x, 0);
            if (e == hipSuccess)
                e = hipMemcpy(alphaOf, h.data(), sizeof(int) * r, hipMemcpyHostToDevice);
        }
    }
    float bound = kLow;
    std::vector<float> hv(Kmax);
    std::vector<int> hvalid(Kmax);
    for (unsigned i0 = 0; i0 < r && e == hipSuccess && rc == NOVA_SUCCESS; i0 += C) {
        const unsigned Ci = std::min(C, r - i0), K = Ci * m * z;
        hipLaunchKernelGGL(pema_triples, dim3((K + 255) / 256), dim3(256), 0, st, K, i0, n,
                           pomdp->ns, m, z, rz, pomdp->d_S, pomdp->d_T, pomdp->d_O, Zw, Bw,
                           rowOf, act, obs, pr);
        e = hipGetLastError();
        if (e != hipSuccess)
            break;
        rc = belief_update_rows(pomdp, K, D, rowOf, act, obs, Bp, status, fn);
        if (rc != NOVA_SUCCESS)
            break;
        hipLaunchKernelGGL(pema_closest, dim3((K + kPT - 1) / kPT, tiles), dim3(kPThreads), 0,
                           st, K, n, r, Bp, D, partD, partJ);
        hipLaunchKernelGGL(pema_eps, dim3((K + 255) / 256), dim3(256), 0, st, K, n, tiles,
                           pomdp->gamma, Rmax, Rmin, Bp, D, rowOf, status, partD, partJ,
                           alphaOf, G, pr, val, valid);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpy(hv.data(), val, sizeof(float) * K, hipMemcpyDeviceToHost);
        if (e == hipSuccess)
            e = hipMemcpy(hvalid.data(), valid, sizeof(int) * K, hipMemcpyDeviceToHost);
        if (e != hipSuccess)
            break;
        // the reference's choices, in its order and float arithmetic
        for (unsigned ii = 0; ii < Ci && e == hipSuccess; ii++) {
            unsigned aStar = 0, oStarOfA = 0;
            float aStarValue = kLow;
            std::vector<unsigned> oStar(m, 0);
            for (unsigned a = 0; a < m; a++) {
                float oStarValue = kLow, aValue = 0.0f;
                for (unsigned o = 0; o < z; o++) {
                    const size_t q = ((size_t)ii * m + a) * z + o;
                    if (!hvalid[q])
                        continue;
                    if (hv[q] > oStarValue) {
                        oStar[a] = o;
                        oStarValue = hv[q];
                    }
                    aValue += hv[q];
                }
                if (aValue > aStarValue) {
                    aStar = a;
                    aStarValue = aValue;
                }
            }
            oStarOfA = oStar[aStar];
            if (aStarValue > bound) {
                const size_t q = ((size_t)ii * m + aStar) * z + oStarOfA;
                const float *src = hvalid[q] ? Bp + q * n : D + (size_t)(i0 + ii) * n;
                e = hipMemcpy(Bnew.data(), src, sizeof(float) * n, hipMemcpyDeviceToHost);
                bound = aStarValue;
            }
        }
    }
    device_free(mem, fn, "PEMA workspace");
    if (rc != NOVA_SUCCESS)
        return rc;
    if (e != hipSuccess)
        return hip_fail(e, fn, "PEMA on the device failed.", NOVA_ERROR_KERNEL_EXECUTION);
    return pomdp_add_new_raw_beliefs_cpu(pomdp, 1, Bnew.data());
}

} // namespace nova
