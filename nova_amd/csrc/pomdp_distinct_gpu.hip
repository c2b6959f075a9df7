// nova-mi355x: distinct-beliefs expansion with the scoring on the GPU
// (src/pomdp/utilities/pomdp_expand_cpu.cpp:166-266). For each belief b_i and every (a, o) with
// Pr(o | b_i, a) != 0 and a valid tau: score = min_j ||b_j - tau(b_i, a, o)||_1 (first strict
// minimum from +1e10), then the first (a, o) with the strictly largest score (from -1e10) is
// appended. Results equal nova's bit for bit:
//   * Pr(o | b, a) sums b(s) * sum_l T*O over the support in ascending state order: the dense
//     loop of the reference only adds exact zeros elsewhere;
//   * tau is formed from its (s', s, l) contributions after a stable sort by s', so each s'
//     accumulates T*b(s) in the reference's (s, l) order; the normaliser sums ascending s';
//   * the L1 distance merges the two sorted supports in ascending state order, again the
//     reference's dense sum without its exact zeros.
// One workgroup per (belief, action, observation) builds tau in LDS and scores all beliefs
// (one lane per candidate belief). Beliefs' supports are sorted once on the host.
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include <nova/amd/nova_amd.h>
#include <nova/pomdp/utilities/pomdp_model_cpu.h>

#include "nova_hip.h"

namespace nova_amd {

namespace {

constexpr int kDB = 256;
constexpr int kTauMax = 1024;    // (support x successors) contributions per tau in LDS

// Builds tau(b_i, a, o) sparsely in LDS: ts[0..cnt) ascending states, tv the values.
// Returns the count through cntOut; ok = 0 when Pr(o|b,a) == 0 or the normaliser is tiny.
__device__ void build_tau(unsigned n, unsigned ns, unsigned m, unsigned z, unsigned rz,
        const int *__restrict__ S, const float *__restrict__ T, const float *__restrict__ O,
        const int *__restrict__ Zs, const float *__restrict__ Bs, unsigned i, unsigned a,
        unsigned o, unsigned long long *keys, float *vals, int *ts, float *tv, int *cntOut,
        int *okOut, int *kEndS)
{
    const unsigned t = threadIdx.x;
    if (t == 0)
        *kEndS = (int)rz;
    __syncthreads();
    for (unsigned k = t; k < rz; k += kDB)
        if (Zs[(size_t)i * rz + k] < 0)
            atomicMin(kEndS, (int)k);
    __syncthreads();
    const int kEnd = *kEndS;
    // contributions in (s ascending, l) order: key = (s' << 32) | (k*ns + l)
    const int total = kEnd * (int)ns;
    int P2 = 1;
    while (P2 < total)
        P2 <<= 1;
    for (int x = t; x < P2; x += kDB) {
        unsigned long long key = ~0ull;
        float v = 0.0f;
        if (x < total) {
            const int k = x / (int)ns, l = x % (int)ns;
            const int s = Zs[(size_t)i * rz + k];
            const size_t row = ((size_t)s * m + a) * ns;
            bool valid = true;
            for (int q = 0; q <= l; q++)
                valid = valid && S[row + q] >= 0;
            if (valid) {
                key = ((unsigned long long)(unsigned)S[row + l] << 32) | (unsigned)x;
                v = __fmul_rn(T[row + l], Bs[(size_t)i * rz + k]);
            }
        }
        keys[x] = key;
        vals[x] = v;
    }
    __syncthreads();
    // bitonic sort of the keys (unique, so the order is the stable one)
    for (int size = 2; size <= P2; size <<= 1) {
        for (int stride = size / 2; stride > 0; stride >>= 1) {
            for (int x = t; x < P2; x += kDB) {
                const int y = x ^ stride;
                if (y > x) {
                    const bool up = (x & size) == 0;
                    if ((keys[x] > keys[y]) == up) {
                        const unsigned long long kk = keys[x];
                        keys[x] = keys[y];
                        keys[y] = kk;
                        const float vv = vals[x];
                        vals[x] = vals[y];
                        vals[y] = vv;
                    }
                }
            }
            __syncthreads();
        }
    }
    // thread 0: Pr(o|b,a) in support order, then tau's accumulation, weights and normaliser
    if (t == 0) {
        float pr = 0.0f;
        for (int k = 0; k < kEnd; k++) {
            const int s = Zs[(size_t)i * rz + k];
            const size_t row = ((size_t)s * m + a) * ns;
            float val = 0.0f;
            for (unsigned l = 0; l < ns; l++) {
                const int sp = S[row + l];
                if (sp < 0)
                    break;
                val = __fadd_rn(val, __fmul_rn(T[row + l], O[((size_t)a * n + sp) * z + o]));
            }
            pr = __fadd_rn(pr, __fmul_rn(val, Bs[(size_t)i * rz + k]));
        }
        int cnt = 0;
        float norm = 0.0f;
        int q = 0;
        while (q < total && keys[q] != ~0ull) {
            const int sp = (int)(keys[q] >> 32);
            float acc = 0.0f;
            while (q < total && keys[q] != ~0ull && (int)(keys[q] >> 32) == sp) {
                acc = __fadd_rn(acc, vals[q]);
                q++;
            }
            acc = __fmul_rn(acc, O[((size_t)a * n + sp) * z + o]);
            ts[cnt] = sp;
            tv[cnt] = acc;
            norm = __fadd_rn(norm, acc);
            cnt++;
        }
        int ok = (pr != 0.0f) && !(fabs((double)norm) < NOVA_FLT_ERR_TOL);
        if (ok)
            for (int c = 0; c < cnt; c++)
                tv[c] = __fdiv_rn(tv[c], norm);
        *cntOut = cnt;
        *okOut = ok;
    }
    __syncthreads();
}

__global__ __launch_bounds__(kDB) void distinct_score(unsigned n, unsigned ns, unsigned m,
        unsigned z, unsigned r, unsigned rz, const int *__restrict__ S,
        const float *__restrict__ T, const float *__restrict__ O, const int *__restrict__ Zs,
        const float *__restrict__ Bs, float *__restrict__ score)
{
    __shared__ unsigned long long keys[kTauMax];
    __shared__ float vals[kTauMax];
    __shared__ int ts[kTauMax];
    __shared__ float tv[kTauMax];
    __shared__ int cnt, ok, kEndS;
    __shared__ float red[kDB];
    const unsigned x = blockIdx.x;
    const unsigned o = x % z, a = (x / z) % m, i = x / (z * m);
    build_tau(n, ns, m, z, rz, S, T, O, Zs, Bs, i, a, o, keys, vals, ts, tv, &cnt, &ok, &kEndS);
    if (!ok) {
        if (threadIdx.x == 0)
            score[x] = -INFINITY;   // skipped (a, o)
        return;
    }
    float best = kHigh;
    for (unsigned j = threadIdx.x; j < r; j += kDB) {
        // ||b_j - tau||_1, merged in ascending state order
        float d = 0.0f;
        int p = 0, q = 0;
        const int *zj = Zs + (size_t)j * rz;
        const float *bj = Bs + (size_t)j * rz;
        int qEnd = (int)rz;
        for (int k = 0; k < (int)rz; k++)
            if (zj[k] < 0) { qEnd = k; break; }
        while (p < cnt || q < qEnd) {
            const int sa = (p < cnt) ? ts[p] : 0x7fffffff;
            const int sb = (q < qEnd) ? zj[q] : 0x7fffffff;
            float diff;
            if (sa == sb) {
                diff = fabsf(__fsub_rn(bj[q], tv[p]));
                p++;
                q++;
            } else if (sa < sb) {
                diff = fabsf(__fsub_rn(0.0f, tv[p]));
                p++;
            } else {
                diff = fabsf(__fsub_rn(bj[q], 0.0f));
                q++;
            }
            d = __fadd_rn(d, diff);
        }
        if (d < best)
            best = d;
    }
    red[threadIdx.x] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        float b = kHigh;
        for (int q = 0; q < kDB; q++)
            b = fminf(b, red[q]);
        score[x] = b;
    }
}

// Per belief: the first (a, o) with the strictly largest score from -1e10; writes the dense
// tau of that pair (or zeros) into Bnew row i.
__global__ __launch_bounds__(kDB) void distinct_select(unsigned n, unsigned ns, unsigned m,
        unsigned z, unsigned rz, const int *__restrict__ S, const float *__restrict__ T,
        const float *__restrict__ O, const int *__restrict__ Zs, const float *__restrict__ Bs,
        const float *__restrict__ score, float *__restrict__ Bnew, int *__restrict__ choice)
{
    __shared__ unsigned long long keys[kTauMax];
    __shared__ float vals[kTauMax];
    __shared__ int ts[kTauMax];
    __shared__ float tv[kTauMax];
    __shared__ int cnt, ok, kEndS, pick;
    const unsigned i = blockIdx.x, t = threadIdx.x;
    if (t == 0) {
        float bestV = kLow;
        int bestX = -1;
        for (unsigned x = 0; x < m * z; x++) {
            const float v = score[(size_t)i * m * z + x];
            if (v > bestV) {
                bestV = v;
                bestX = (int)x;
            }
        }
        pick = bestX;
        choice[i] = bestX;
    }
    __syncthreads();
    float *row = Bnew + (size_t)i * n;
    for (unsigned s = t; s < n; s += kDB)
        row[s] = 0.0f;
    __syncthreads();
    if (pick < 0)
        return;
    const unsigned a = (unsigned)pick / z, o = (unsigned)pick % z;
    build_tau(n, ns, m, z, rz, S, T, O, Zs, Bs, i, a, o, keys, vals, ts, tv, &cnt, &ok, &kEndS);
    for (int c = t; c < cnt; c += kDB)
        row[ts[c]] = tv[c];
}

} // namespace

} // namespace nova_amd

namespace nova {

using namespace nova_amd;

int pomdp_expand_distinct_beliefs_gpu(POMDP *pomdp)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "pomdp_expand_distinct_beliefs_gpu";
    if (pomdp == nullptr || pomdp->n == 0 || pomdp->ns == 0 || pomdp->m == 0 || pomdp->z == 0 ||
            pomdp->r == 0 || pomdp->rz == 0 || pomdp->S == nullptr || pomdp->T == nullptr ||
            pomdp->O == nullptr || pomdp->R == nullptr || pomdp->Z == nullptr ||
            pomdp->B == nullptr) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    const unsigned n = pomdp->n, ns = pomdp->ns, m = pomdp->m, z = pomdp->z, r = pomdp->r,
                   rz = pomdp->rz;
    if ((size_t)rz * ns > (size_t)kTauMax) {
        NOVA_ERR(fn, "rz*ns must not exceed 1024 for the device expansion.");
        return NOVA_ERROR_INVALID_DATA;
    }
    if (first_bad_successor(pomdp->S, (size_t)n * m, ns, n) >= 0 ||
            first_bad_successor(pomdp->Z, r, rz, n) >= 0) {
        NOVA_ERR(fn, "A successor or support index is outside [0, n).");
        return NOVA_ERROR_INVALID_DATA;
    }
    // supports sorted by state (the valid prefix; -1 padding kept at the end)
    std::vector<int> Zs((size_t)r * rz);
    std::vector<float> Bsv((size_t)r * rz);
    for (unsigned i = 0; i < r; i++) {
        unsigned k = 0;
        while (k < rz && pomdp->Z[(size_t)i * rz + k] >= 0)
            k++;
        std::vector<unsigned> ord(k);
        std::iota(ord.begin(), ord.end(), 0u);
        std::stable_sort(ord.begin(), ord.end(), [&](unsigned x, unsigned y) {
            return pomdp->Z[(size_t)i * rz + x] < pomdp->Z[(size_t)i * rz + y];
        });
        for (unsigned q = 0; q < rz; q++) {
            Zs[(size_t)i * rz + q] = (q < k) ? pomdp->Z[(size_t)i * rz + ord[q]] : -1;
            Bsv[(size_t)i * rz + q] = (q < k) ? pomdp->B[(size_t)i * rz + ord[q]] : 0.0f;
        }
    }
    void *d[8] = {};
    const size_t bytes[8] = {sizeof(int) * (size_t)n * m * ns, sizeof(float) * (size_t)n * m * ns,
                             sizeof(float) * (size_t)m * n * z, sizeof(int) * Zs.size(),
                             sizeof(float) * Bsv.size(), sizeof(float) * (size_t)r * m * z,
                             sizeof(float) * (size_t)r * n, sizeof(int) * r};
    int rc = NOVA_SUCCESS;
    for (int x = 0; x < 8 && rc == NOVA_SUCCESS; x++)
        rc = device_alloc(&d[x], bytes[x], fn, "expansion buffers");
    const void *src[5] = {pomdp->S, pomdp->T, pomdp->O, Zs.data(), Bsv.data()};
    hipError_t e = hipSuccess;
    for (int x = 0; x < 5 && rc == NOVA_SUCCESS && e == hipSuccess; x++)
        e = hipMemcpy(d[x], src[x], bytes[x], hipMemcpyHostToDevice);
    std::vector<float> Bnew;
    if (rc == NOVA_SUCCESS && e == hipSuccess) {
        hipStream_t st = stream();
        hipLaunchKernelGGL(distinct_score, dim3(r * m * z), dim3(kDB), 0, st, n, ns, m, z, r, rz,
                           static_cast<int *>(d[0]), static_cast<float *>(d[1]),
                           static_cast<float *>(d[2]), static_cast<int *>(d[3]),
                           static_cast<float *>(d[4]), static_cast<float *>(d[5]));
        hipLaunchKernelGGL(distinct_select, dim3(r), dim3(kDB), 0, st, n, ns, m, z, rz,
                           static_cast<int *>(d[0]), static_cast<float *>(d[1]),
                           static_cast<float *>(d[2]), static_cast<int *>(d[3]),
                           static_cast<float *>(d[4]), static_cast<float *>(d[5]),
                           static_cast<float *>(d[6]), static_cast<int *>(d[7]));
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipStreamSynchronize(st);
        if (e == hipSuccess) {
            Bnew.resize((size_t)r * n);
            e = hipMemcpy(Bnew.data(), d[6], bytes[6], hipMemcpyDeviceToHost);
        }
    }
    if (rc == NOVA_SUCCESS && e != hipSuccess)
        rc = hip_fail(e, fn, "Failed to run the expansion on the device.", NOVA_ERROR_KERNEL_EXECUTION);
    for (void *p : d)
        if (p)
            device_free(p, fn, "expansion buffers");
    if (rc != NOVA_SUCCESS)
        return rc;
    return pomdp_add_new_raw_beliefs_cpu(pomdp, r, Bnew.data());
}

} // namespace nova
