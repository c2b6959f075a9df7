// nova-mi355x: min-cost Bellman backups of stochastic shortest paths on one MI355X.
//
// nova runs its SSP solvers (LAO*, RTDP) on the CPU only. Their Bellman backup
// (src/mdp/algorithms/ssp_lao_star_cpu.cpp:39-68 with |V|, the sign of V marking visitation;
// src/mdp/algorithms/ssp_rtdp_cpu.cpp:39-74 with V as is) is the VI gather-reduce with costs:
//   Q(s,a) = R[s,a]; Q += T[s,a,i] * f(V[S[s,a,i]]) for i = 0.. until the first S < 0;
//   V'(s) = min over a, action 0 first and a later action only when strictly smaller.
// Here it runs as Jacobi backups of many states at once, each product and sum rounded
// separately in that order (no FMA), so every backed-up state equals nova's backup of it from
// the same V bit for bit. Two entry points:
//   * the backups of a list of states (for instance the states of an LAO* solution graph);
//   * sweeps over all states with the goals held at 0, until the largest change is at most
//     epsilon (SSP value iteration: the same fixed point LAO* and RTDP approach from s0).
// Mapping: a workgroup of 256 lanes takes 256 / m states; lane (state, action) sums its row,
// the row sums meet in LDS, and the action-0 lane of each state takes the minimum in action
// order. The sweep's largest change is a workgroup maximum folded into one word with an atomic
// max on the float's bits (non-negative floats order as their bit patterns), so it does not
// depend on the order in which workgroups finish.
#include <algorithm>
#include <cstring>

#include <nova/amd/nova_amd.h>
#include <nova/mdp/mdp.h>

#include "nova_hip.h"

namespace nova_amd {

namespace {

constexpr int kSspThreads = 256;
constexpr unsigned kSspMaxM = 256;

__device__ __forceinline__ float ssp_row(unsigned n, unsigned m, unsigned ns, unsigned s,
        unsigned a, int absV, const int *__restrict__ S, const float *__restrict__ T,
        const float *__restrict__ R, const float *__restrict__ V)
{
    const size_t row = ((size_t)s * m + a) * ns;
    float q = R[(size_t)s * m + a];
    for (unsigned i = 0; i < ns; i++) {
        const int sp = S[row + i];
        if (sp < 0)
            break;
        const float v = V[min((unsigned)sp, n - 1)];
        q = __fadd_rn(q, __fmul_rn(T[row + i], absV ? fabsf(v) : v));
    }
    return q;
}

// Backups of states[0..k) (or of 0..k-1 when states is null) from V into Vout and pi. A state
// marked in goal (when given) gets Vout = 0 and pi = 0. resid (when given) receives the
// largest |Vout(s) - V(s)| as float bits through an atomic max.
__global__ __launch_bounds__(kSspThreads) void ssp_backup_rows(unsigned n, unsigned m,
        unsigned ns, unsigned k, const unsigned *__restrict__ states,
        const unsigned char *__restrict__ goal, int absV, const int *__restrict__ S,
        const float *__restrict__ T, const float *__restrict__ R, const float *__restrict__ V,
        float *__restrict__ Vout, unsigned *__restrict__ pi, unsigned *__restrict__ resid)
{
    __shared__ float q[kSspThreads];
    __shared__ float wmax[kSspThreads / 64];
    const unsigned perBlock = kSspThreads / m;
    const unsigned t = threadIdx.x, lq = t / m, a = t % m;
    const unsigned x = blockIdx.x * perBlock + lq;
    const bool listed = lq < perBlock && x < k;
    const unsigned s = listed ? (states ? states[x] : x) : 0u;
    const bool live = listed && s < n;           // a listed index past n is skipped
    const bool isGoal = live && goal != nullptr && goal[s] != 0;
    if (live && !isGoal)
        q[t] = ssp_row(n, m, ns, s, a, absV, S, T, R, V);
    __syncthreads();
    float change = 0.0f;
    if (live && a == 0) {
        float best = 0.0f;
        unsigned bestA = 0;
        if (!isGoal) {
            best = q[t];
            for (unsigned b = 1; b < m; b++) {
                if (q[t + b] < best) {
                    best = q[t + b];
                    bestA = b;
                }
            }
        }
        if (resid)
            change = fabsf(best - V[s]);
        Vout[s] = best;
        pi[s] = bestA;
    }
    if (resid) {
        for (int off = 32; off >= 1; off >>= 1)
            change = fmaxf(change, __shfl_xor(change, off));
        if (t % 64 == 0)
            wmax[t / 64] = change;
        __syncthreads();
        if (t == 0) {
            float w = wmax[0];
            for (int i = 1; i < kSspThreads / 64; i++)
                w = fmaxf(w, wmax[i]);
            atomicMax(resid, __float_as_uint(w));
        }
    }
}

// One lane per state, actions in order: the shapes with m > 256.
__global__ __launch_bounds__(kSspThreads) void ssp_backup_states(unsigned n, unsigned m,
        unsigned ns, unsigned k, const unsigned *__restrict__ states,
        const unsigned char *__restrict__ goal, int absV, const int *__restrict__ S,
        const float *__restrict__ T, const float *__restrict__ R, const float *__restrict__ V,
        float *__restrict__ Vout, unsigned *__restrict__ pi, unsigned *__restrict__ resid)
{
    __shared__ float wmax[kSspThreads / 64];
    const unsigned t = threadIdx.x, x = blockIdx.x * kSspThreads + t;
    float change = 0.0f;
    const unsigned s = x < k ? (states ? states[x] : x) : n;
    if (s < n) {                                 // a listed index past n is skipped
        float best = 0.0f;
        unsigned bestA = 0;
        if (goal == nullptr || goal[s] == 0) {
            for (unsigned a = 0; a < m; a++) {
                const float qa = ssp_row(n, m, ns, s, a, absV, S, T, R, V);
                if (a == 0 || qa < best) {
                    best = qa;
                    bestA = a;
                }
            }
        }
        if (resid)
            change = fabsf(best - V[s]);
        Vout[s] = best;
        pi[s] = bestA;
    }
    if (resid) {
        for (int off = 32; off >= 1; off >>= 1)
            change = fmaxf(change, __shfl_xor(change, off));
        if (t % 64 == 0)
            wmax[t / 64] = change;
        __syncthreads();
        if (t == 0) {
            float w = wmax[0];
            for (int i = 1; i < kSspThreads / 64; i++)
                w = fmaxf(w, wmax[i]);
            atomicMax(resid, __float_as_uint(w));
        }
    }
}

__global__ __launch_bounds__(256) void ssp_goal_clear(unsigned n,
        unsigned char *__restrict__ mask)
{
    const unsigned x = blockIdx.x * 256 + threadIdx.x;
    if (x < n)
        mask[x] = 0;
}

__global__ __launch_bounds__(256) void ssp_goal_mark(unsigned n, unsigned ng,
        const unsigned *__restrict__ goals, unsigned char *__restrict__ mask)
{
    const unsigned x = blockIdx.x * 256 + threadIdx.x;
    if (x < ng && goals[x] < n)
        mask[goals[x]] = 1;
}

__global__ void ssp_clear_word(unsigned *__restrict__ w)
{
    if (threadIdx.x == 0)
        w[0] = 0u;
}

bool ssp_device_ready(const nova::MDP *mdp)
{
    return mdp != nullptr && mdp->n != 0 && mdp->ns != 0 && mdp->m != 0 &&
           mdp->d_S != nullptr && mdp->d_T != nullptr && mdp->d_R != nullptr;
}

void ssp_launch(const nova::MDP *mdp, unsigned k, const unsigned *states,
                const unsigned char *goal, int absV, const float *V, float *Vout, unsigned *pi,
                unsigned *resid, hipStream_t st)
{
    const unsigned n = mdp->n, m = mdp->m, ns = mdp->ns;
    if (m <= kSspMaxM) {
        const unsigned perBlock = kSspThreads / m;
        hipLaunchKernelGGL(ssp_backup_rows, dim3((k + perBlock - 1) / perBlock),
                           dim3(kSspThreads), 0, st, n, m, ns, k, states, goal, absV, mdp->d_S,
                           mdp->d_T, mdp->d_R, V, Vout, pi, resid);
    } else {
        hipLaunchKernelGGL(ssp_backup_states, dim3((k + kSspThreads - 1) / kSspThreads),
                           dim3(kSspThreads), 0, st, n, m, ns, k, states, goal, absV, mdp->d_S,
                           mdp->d_T, mdp->d_R, V, Vout, pi, resid);
    }
}

} // namespace

} // namespace nova_amd

using namespace nova_amd;

extern "C" int nova_amd_ssp_backup_gpu(const nova::MDP *mdp, unsigned int k,
        const unsigned int *d_states, int absV, const float *d_V, float *d_Vout,
        unsigned int *d_pi)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "nova_amd_ssp_backup_gpu";
    if (!ssp_device_ready(mdp) || k == 0 || d_states == nullptr || d_V == nullptr ||
            d_Vout == nullptr || d_pi == nullptr || d_V == d_Vout) {
        NOVA_ERR(fn, "Invalid arguments (a device-resident MDP, k > 0, device arrays, d_Vout "
                     "distinct from d_V).");
        return NOVA_ERROR_INVALID_DATA;
    }
    hipStream_t st = stream();
    ssp_launch(mdp, k, d_states, nullptr, absV, d_V, d_Vout, d_pi, nullptr, st);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    return hip_fail(e, fn, "Kernel execution failed.", NOVA_ERROR_KERNEL_EXECUTION);
}

extern "C" int nova_amd_ssp_sweeps_gpu(const nova::MDP *mdp, int absV, const float *d_V0,
        unsigned int maxSweeps, float epsilon, float *d_V, unsigned int *d_pi,
        unsigned int *sweepsDone, float *residual)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "nova_amd_ssp_sweeps_gpu";
    if (!ssp_device_ready(mdp) || maxSweeps == 0 || d_V0 == nullptr || d_V == nullptr ||
            d_pi == nullptr || !(epsilon >= 0.0f) || (mdp->ng > 0 && mdp->d_goals == nullptr)) {
        NOVA_ERR(fn, "Invalid arguments (a device-resident MDP with its goals, sweeps > 0, "
                     "device arrays, epsilon >= 0).");
        return NOVA_ERROR_INVALID_DATA;
    }
    const unsigned n = mdp->n;
    // workspace: the goal mask, the other V buffer, the change word
    const size_t bMask = (n + 255) / 256 * 256, bV = (size_t)n * sizeof(float);
    void *mem = nullptr;
    int rc = device_alloc(&mem, bMask + bV + 256, fn, "the sweep workspace");
    if (rc != NOVA_SUCCESS)
        return rc;
    unsigned char *mask = static_cast<unsigned char *>(mem);
    float *Vb = reinterpret_cast<float *>(static_cast<char *>(mem) + bMask);
    unsigned *word = reinterpret_cast<unsigned *>(static_cast<char *>(mem) + bMask + bV);
    hipStream_t st = stream();
    const unsigned g = (n + 255) / 256;
    hipLaunchKernelGGL(ssp_goal_clear, dim3(g), dim3(256), 0, st, n, mask);
    if (mdp->ng > 0)
        hipLaunchKernelGGL(ssp_goal_mark, dim3((mdp->ng + 255) / 256), dim3(256), 0, st, n,
                           mdp->ng, mdp->d_goals, mask);
    hipError_t e = hipMemcpyAsync(d_V, d_V0, bV, hipMemcpyDeviceToDevice, st);
    float *cur = d_V, *nxt = Vb;
    unsigned done = 0;
    float res = 0.0f;
    while (e == hipSuccess && done < maxSweeps) {
        hipLaunchKernelGGL(ssp_clear_word, dim3(1), dim3(64), 0, st, word);
        ssp_launch(mdp, n, nullptr, mask, absV, cur, nxt, d_pi, word, st);
        unsigned bitsOut = 0;
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(&bitsOut, word, sizeof(unsigned), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess)
            e = hipStreamSynchronize(st);
        std::swap(cur, nxt);
        done++;
        std::memcpy(&res, &bitsOut, sizeof(float));
        if (res <= epsilon)
            break;
    }
    if (e == hipSuccess && cur != d_V)
        e = hipMemcpyAsync(d_V, cur, bV, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    device_free(mem, fn, "the sweep workspace");
    if (sweepsDone)
        *sweepsDone = done;
    if (residual)
        *residual = res;
    return hip_fail(e, fn, "Kernel execution failed.", NOVA_ERROR_KERNEL_EXECUTION);
}
