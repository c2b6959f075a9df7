// nova-mi355x: MDP value iteration on one MI355X (gfx950, wave64).
//
// The sweep kernel reproduces src/mdp/algorithms/mdp_vi_cpu.cpp:36-63 bit for bit:
//   Q(s,a) = R[s,a]; Q += (gamma*T[s,a,i]) * V[S[s,a,i]] for i = 0.. until the first S < 0;
//   V'(s) = Q(s,0) unless a later action is strictly larger; pi(s) = that action.
// Layout and mapping (one workgroup = 256 lanes = 4 waves per tile of whole states):
//   * each (s,a) successor row of ns entries is owned by G adjacent lanes holding E = 4 (or 1)
//     consecutive entries each, so a wave reads 64*E*4 contiguous bytes of S and of T per load
//     (dwordx4, non-temporal: S/T are streamed once per sweep and must not evict V from L2);
//   * the valid prefix of each row (up to the first negative successor) is found with a wave
//     ballot before any gather, so V is only read at valid indices;
//   * the row sum runs as a serial chain across the G lanes (lane g adds its E products in
//     order, then hands the running sum to lane g+1), keeping the reference's left-to-right
//     evaluation order and its separate rounding of each product and sum (no FMA);
//   * row sums are staged in LDS; one lane per state then takes the max over actions with the
//     reference's tie rule, and V'/pi are written coalesced.
// A plain one-thread-per-state kernel covers shapes the tiled kernel does not (ns > 256,
// m > 2048, misaligned arrays).
#include <cstring>
#include <vector>

#include <nova/amd/nova_amd.h>
#include <nova/mdp/algorithms/mdp_vi_gpu.h>

#include "nova_hip.h"

namespace nova_amd {

namespace {

constexpr int kBlock = 256;
constexpr int kTileRows = 2048;   // (s,a) rows whose Q values a workgroup stages in LDS
constexpr int kUnroll = 8;        // rows in flight per lane group (8 measured ~4% faster than 4)

typedef int int4v __attribute__((ext_vector_type(4)));
typedef float float4v __attribute__((ext_vector_type(4)));

template <int E> struct Entries {
    int s[E];
    float t[E];
};

template <int E>
__device__ __forceinline__ void load_entries(const int *__restrict__ S, const float *__restrict__ T,
                                             size_t off, bool in, Entries<E> &x)
{
    if constexpr (E == 4) {
        if (in) {
            int4v sv = __builtin_nontemporal_load(reinterpret_cast<const int4v *>(S + off));
            float4v tv = __builtin_nontemporal_load(reinterpret_cast<const float4v *>(T + off));
            x.s[0] = sv.x; x.s[1] = sv.y; x.s[2] = sv.z; x.s[3] = sv.w;
            x.t[0] = tv.x; x.t[1] = tv.y; x.t[2] = tv.z; x.t[3] = tv.w;
        } else {
#pragma unroll
            for (int e = 0; e < 4; e++) { x.s[e] = -1; x.t[e] = 0.0f; }
        }
    } else {
        x.s[0] = in ? __builtin_nontemporal_load(S + off) : -1;
        x.t[0] = in ? __builtin_nontemporal_load(T + off) : 0.0f;
    }
}

template <int E, int G>
__global__ __launch_bounds__(kBlock) void vi_sweep_tiled(
        unsigned n, unsigned m, unsigned ns, unsigned statesPerTile, float gamma,
        const int *__restrict__ S, const float *__restrict__ T, const float *__restrict__ R,
        const float *__restrict__ V, float *__restrict__ Vp, unsigned *__restrict__ pi,
        int writePi)
{
    __shared__ float q[kTileRows];

    const unsigned s0 = blockIdx.x * statesPerTile;
    const unsigned s1 = min(n, s0 + statesPerTile);
    const unsigned rows = (s1 - s0) * m;
    const size_t rowBase = (size_t)s0 * m;

    const unsigned g = threadIdx.x % G;                 // lane within its row group
    const unsigned lane = threadIdx.x % 64;             // lane within the wave
    const unsigned long long groupBelow =               // lanes of my group before me
        ((1ull << g) - 1ull) << (lane - g);
    const unsigned passRows = kBlock / G;
    const bool inRow = g * E < ns;                      // my entries lie inside the row
    const unsigned nClamp = n - 1;

    for (unsigned base = threadIdx.x / G; base < rows; base += passRows * kUnroll) {
        Entries<E> x[kUnroll];
        float r0[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            const unsigned rr = base + u * passRows;
            const bool live = rr < rows;
            const size_t row = rowBase + rr;
            load_entries<E>(S, T, row * ns + g * E, live && inRow, x[u]);
            r0[u] = (live && g == 0) ? __builtin_nontemporal_load(R + row) : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            const unsigned rr = base + u * passRows;
            // valid entries of my slice: up to my first negative successor, none if an
            // earlier lane of the row already hit one
            int loc = E;
#pragma unroll
            for (int e = E - 1; e >= 0; e--)
                if (x[u].s[e] < 0) loc = e;
            const unsigned long long hit = __ballot(loc < E);
            const int valid = (hit & groupBelow) ? 0 : loc;
            float p[E];
#pragma unroll
            for (int e = 0; e < E; e++) {
                p[e] = 0.0f;
                if (e < valid) {
                    const unsigned sp = min((unsigned)x[u].s[e], nClamp);
                    p[e] = __fmul_rn(__fmul_rn(gamma, x[u].t[e]), V[sp]);
                }
            }
            // serial chain: lane 0 of the row starts from R, each lane adds its products in
            // order, then passes the running sum one lane up
            float acc = r0[u];
#pragma unroll
            for (int t = 0; t < G; t++) {
                if ((int)g == t) {
#pragma unroll
                    for (int e = 0; e < E; e++)
                        if (e < valid) acc = __fadd_rn(acc, p[e]);
                }
                if (t + 1 < G) {
                    const float up = __shfl_up(acc, 1, G);
                    if ((int)g == t + 1) acc = up;
                }
            }
            if (g == G - 1 && rr < rows)
                q[rr] = acc;
        }
    }
    __syncthreads();

    for (unsigned t = threadIdx.x; t < s1 - s0; t += kBlock) {
        const float *qs = q + (size_t)t * m;
        float best = qs[0];
        unsigned bestA = 0;
        for (unsigned a = 1; a < m; a++) {
            if (qs[a] > best) {
                best = qs[a];
                bestA = a;
            }
        }
        Vp[s0 + t] = best;
        if (writePi)
            pi[s0 + t] = bestA;
    }
}

// One thread per state, sequential: the shapes the tiled kernel does not take.
__global__ __launch_bounds__(kBlock) void vi_sweep_plain(
        unsigned n, unsigned m, unsigned ns, float gamma,
        const int *__restrict__ S, const float *__restrict__ T, const float *__restrict__ R,
        const float *__restrict__ V, float *__restrict__ Vp, unsigned *__restrict__ pi,
        int writePi)
{
    const unsigned s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n)
        return;
    float best = 0.0f;
    unsigned bestA = 0;
    for (unsigned a = 0; a < m; a++) {
        const size_t row = ((size_t)s * m + a) * ns;
        float q = R[(size_t)s * m + a];
        for (unsigned i = 0; i < ns; i++) {
            const int sp = S[row + i];
            if (sp < 0)
                break;
            q = __fadd_rn(q, __fmul_rn(__fmul_rn(gamma, T[row + i]), V[min((unsigned)sp, n - 1)]));
        }
        if (a == 0 || q > best) {
            best = q;
            bestA = a;
        }
    }
    Vp[s] = best;
    if (writePi)
        pi[s] = bestA;
}

typedef void (*TiledFn)(unsigned, unsigned, unsigned, unsigned, float, const int *,
                        const float *, const float *, const float *, float *, unsigned *, int);

template <int E> TiledFn tiled_for(unsigned G)
{
    switch (G) {
    case 1: return vi_sweep_tiled<E, 1>;
    case 2: return vi_sweep_tiled<E, 2>;
    case 4: return vi_sweep_tiled<E, 4>;
    case 8: return vi_sweep_tiled<E, 8>;
    case 16: return vi_sweep_tiled<E, 16>;
    case 32: return vi_sweep_tiled<E, 32>;
    case 64: return vi_sweep_tiled<E, 64>;
    default: return nullptr;
    }
}

unsigned pow2_at_least(unsigned x)
{
    unsigned p = 1;
    while (p < x)
        p <<= 1;
    return p;
}

} // namespace

// A sweep launch plan for one MDP shape.
struct ViPlan {
    TiledFn tiled = nullptr;
    unsigned statesPerTile = 0;
    unsigned grid = 0;
};

static ViPlan vi_plan(const nova::MDP *mdp)
{
    ViPlan p;
    const unsigned n = mdp->n, m = mdp->m, ns = mdp->ns;
    const bool aligned = ((reinterpret_cast<uintptr_t>(mdp->d_S) |
                           reinterpret_cast<uintptr_t>(mdp->d_T)) & 15u) == 0;
    if (m <= (unsigned)kTileRows) {
        if (ns % 4 == 0 && aligned && ns <= 256)
            p.tiled = tiled_for<4>(pow2_at_least(ns / 4));
        else if (ns <= 64)
            p.tiled = tiled_for<1>(pow2_at_least(ns));
    }
    if (p.tiled) {
        p.statesPerTile = kTileRows / m;
        p.grid = (n + p.statesPerTile - 1) / p.statesPerTile;
    } else {
        p.grid = (n + kBlock - 1) / kBlock;
    }
    return p;
}

// writePi: the policy buffer is shared by all sweeps and only the last sweep's greedy actions
// are ever read (src/mdp/algorithms/mdp_vi_cpu.cpp:187-191, :215-220), so a run of sweeps
// writes pi on its final sweep only.
static void vi_launch(const ViPlan &p, const nova::MDP *mdp, const float *V, float *Vp,
                      unsigned *pi, bool writePi, hipStream_t st)
{
    if (p.tiled)
        hipLaunchKernelGGL(p.tiled, dim3(p.grid), dim3(kBlock), 0, st, mdp->n, mdp->m, mdp->ns,
                           p.statesPerTile, mdp->gamma, mdp->d_S, mdp->d_T, mdp->d_R, V, Vp, pi,
                           (int)writePi);
    else
        hipLaunchKernelGGL(vi_sweep_plain, dim3(p.grid), dim3(kBlock), 0, st, mdp->n, mdp->m,
                           mdp->ns, mdp->gamma, mdp->d_S, mdp->d_T, mdp->d_R, V, Vp, pi,
                           (int)writePi);
}

// Cached executable graph of a run of sweeps (one cache entry; re-captured when any launch
// argument changes).
struct ViGraph {
    hipGraphExec_t exec = nullptr;
    const void *key[8] = {};
    unsigned sweeps = 0, parity = 0, n = 0, m = 0, ns = 0;
    float gamma = 0.0f;
};
static ViGraph g_graph;

static int vi_run_sweeps(const nova::MDP *mdp, nova::MDPVIGPU *vi, unsigned sweeps,
                         hipStream_t st, hipEvent_t t0, hipEvent_t t1, const char *fn)
{
    const ViPlan plan = vi_plan(mdp);
    const unsigned parity = vi->currentHorizon % 2;
    const void *key[8] = {mdp->d_S, mdp->d_T, mdp->d_R, vi->d_V, vi->d_VPrime, vi->d_pi,
                          (const void *)plan.tiled, nullptr};
    bool hit = g_graph.exec != nullptr && g_graph.sweeps == sweeps && g_graph.parity == parity &&
               g_graph.n == mdp->n && g_graph.m == mdp->m && g_graph.ns == mdp->ns &&
               g_graph.gamma == mdp->gamma && std::memcmp(g_graph.key, key, sizeof(key)) == 0;
    if (!hit) {
        if (g_graph.exec) {
            hipGraphExecDestroy(g_graph.exec);
            g_graph.exec = nullptr;
        }
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess)
            return hip_fail(e, fn, "Failed to begin graph capture.", NOVA_ERROR_KERNEL_EXECUTION);
        for (unsigned k = 0; k < sweeps; k++) {
            const bool even = ((parity + k) % 2) == 0;
            vi_launch(plan, mdp, even ? vi->d_V : vi->d_VPrime, even ? vi->d_VPrime : vi->d_V,
                      vi->d_pi, k + 1 == sweeps, st);
        }
        e = hipStreamEndCapture(st, &graph);
        if (e != hipSuccess)
            return hip_fail(e, fn, "Failed to capture the sweeps.", NOVA_ERROR_KERNEL_EXECUTION);
        e = hipGraphInstantiate(&g_graph.exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (e != hipSuccess) {
            g_graph.exec = nullptr;
            return hip_fail(e, fn, "Failed to instantiate the sweep graph.",
                            NOVA_ERROR_KERNEL_EXECUTION);
        }
        std::memcpy(g_graph.key, key, sizeof(key));
        g_graph.sweeps = sweeps;
        g_graph.parity = parity;
        g_graph.n = mdp->n;
        g_graph.m = mdp->m;
        g_graph.ns = mdp->ns;
        g_graph.gamma = mdp->gamma;
    }
    if (t0) hipEventRecord(t0, st);
    hipError_t e = hipGraphLaunch(g_graph.exec, st);
    if (e != hipSuccess)
        return hip_fail(e, fn, "Failed to execute the sweep graph.", NOVA_ERROR_KERNEL_EXECUTION);
    if (t1) hipEventRecord(t1, st);
    e = hipStreamSynchronize(st);
    if (e != hipSuccess)
        return hip_fail(e, fn, "Failed to synchronize after the sweeps.",
                        NOVA_ERROR_DEVICE_SYNCHRONIZE);
    vi->currentHorizon += sweeps;
    return NOVA_SUCCESS;
}

int vi_sweeps_float(const nova::MDP *mdp, nova::MDPVIGPU *vi, unsigned sweeps, hipStream_t st,
                    const char *fn)
{
    return vi_run_sweeps(mdp, vi, sweeps, st, nullptr, nullptr, fn);
}

static bool mdp_device_ready(const nova::MDP *mdp)
{
    return mdp != nullptr && mdp->n != 0 && mdp->ns != 0 && mdp->m != 0 && mdp->d_S != nullptr &&
           mdp->d_T != nullptr && mdp->d_R != nullptr;
}

} // namespace nova_amd

namespace nova {

using namespace nova_amd;

int mdp_vi_initialize_gpu(const MDP *mdp, MDPVIGPU *vi)
{
    const char *fn = "mdp_vi_initialize_gpu";
    if (mdp == nullptr || mdp->n == 0 || vi == nullptr) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    device_free_if_owned(vi->d_V, fn, "the value function");
    device_free_if_owned(vi->d_VPrime, fn, "the value function (prime)");
    device_free_if_owned(vi->d_pi, fn, "the policy (pi)");
    vi->d_V = nullptr;
    vi->d_VPrime = nullptr;
    vi->d_pi = nullptr;
    vi->currentHorizon = 0;

    const size_t n = mdp->n;
    int rc = device_alloc((void **)&vi->d_V, n * sizeof(float), fn, "the value function");
    if (rc == NOVA_SUCCESS)
        rc = device_alloc((void **)&vi->d_VPrime, n * sizeof(float), fn, "the value function (prime)");
    if (rc == NOVA_SUCCESS)
        rc = device_alloc((void **)&vi->d_pi, n * sizeof(unsigned), fn, "the policy (pi)");
    if (rc != NOVA_SUCCESS)
        return rc;

    std::vector<float> zeros;
    const float *V0 = vi->VInitial;
    if (V0 == nullptr) {
        zeros.assign(n, 0.0f);
        V0 = zeros.data();
    }
    hipError_t e = hipMemcpy(vi->d_V, V0, n * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(vi->d_VPrime, V0, n * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemset(vi->d_pi, 0, n * sizeof(unsigned));
    return hip_fail(e, fn, "Failed to copy memory from host to device for V/pi.",
                    NOVA_ERROR_MEMCPY_TO_DEVICE);
}

int mdp_vi_uninitialize_gpu(const MDP *mdp, MDPVIGPU *vi)
{
    const char *fn = "mdp_vi_uninitialize_gpu";
    if (mdp == nullptr || vi == nullptr) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    vi->currentHorizon = 0;
    int rc = NOVA_SUCCESS;
    if (device_free_if_owned(vi->d_V, fn, "the value function") != NOVA_SUCCESS)
        rc = NOVA_ERROR_DEVICE_FREE;
    if (device_free_if_owned(vi->d_VPrime, fn, "the value function (prime)") != NOVA_SUCCESS)
        rc = NOVA_ERROR_DEVICE_FREE;
    if (device_free_if_owned(vi->d_pi, fn, "the policy (pi)") != NOVA_SUCCESS)
        rc = NOVA_ERROR_DEVICE_FREE;
    vi->d_V = nullptr;
    vi->d_VPrime = nullptr;
    vi->d_pi = nullptr;
    return rc;
}

int mdp_vi_update_gpu(const MDP *mdp, MDPVIGPU *vi)
{
    const char *fn = "mdp_vi_update_gpu";
    hipStream_t st = stream();
    const ViPlan plan = vi_plan(mdp);
    const bool even = (vi->currentHorizon % 2) == 0;
    vi_launch(plan, mdp, even ? vi->d_V : vi->d_VPrime, even ? vi->d_VPrime : vi->d_V, vi->d_pi,
              true, st);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return hip_fail(e, fn, "Failed to execute the 'Bellman update' kernel.",
                        NOVA_ERROR_KERNEL_EXECUTION);
    e = hipStreamSynchronize(st);
    if (e != hipSuccess)
        return hip_fail(e, fn, "Failed to synchronize the device after 'Bellman update' kernel.",
                        NOVA_ERROR_DEVICE_SYNCHRONIZE);
    vi->currentHorizon++;
    return NOVA_SUCCESS;
}

int mdp_vi_get_policy_gpu(const MDP *mdp, MDPVIGPU *vi, MDPValueFunction *policy)
{
    const char *fn = "mdp_vi_get_policy_gpu";
    if (mdp == nullptr || vi == nullptr || policy == nullptr) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    if (mdp_value_function_initialize(policy, mdp->n, mdp->m, 0) != NOVA_SUCCESS) {
        NOVA_ERR(fn, "Could not create the policy.");
        return NOVA_ERROR_POLICY_CREATION;
    }
    const float *dV = (vi->currentHorizon % 2 == 0) ? vi->d_V : vi->d_VPrime;
    hipError_t e = hipMemcpy(policy->V, dV, sizeof(float) * mdp->n, hipMemcpyDeviceToHost);
    if (e == hipSuccess)
        e = hipMemcpy(policy->pi, vi->d_pi, sizeof(unsigned) * mdp->n, hipMemcpyDeviceToHost);
    return hip_fail(e, fn, "Failed to copy memory from device to host for V/pi.",
                    NOVA_ERROR_MEMCPY_TO_HOST);
}

int mdp_vi_execute_gpu(const MDP *mdp, MDPVIGPU *vi, MDPValueFunction *policy)
{
    const char *fn = "mdp_vi_execute_gpu";
    if (mdp == nullptr || mdp->n == 0 || mdp->ns == 0 || mdp->m == 0 || mdp->S == nullptr ||
            mdp->T == nullptr || mdp->R == nullptr || mdp->gamma < 0.0f || mdp->gamma > 1.0f ||
            mdp->horizon < 1 || vi == nullptr || policy == nullptr) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    if (vi->numThreads % 32 != 0) {
        NOVA_ERR(fn, "Invalid number of threads.");
        return NOVA_ERROR_INVALID_CUDA_PARAM;
    }
    if (!mdp_device_ready(mdp)) {
        NOVA_ERR(fn, "The MDP is not resident on the device (call mdp_initialize_gpu first).");
        return NOVA_ERROR_INVALID_DATA;
    }
    int rc = mdp_vi_initialize_gpu(mdp, vi);
    if (rc != NOVA_SUCCESS)
        return rc;
    rc = vi_run_sweeps(mdp, vi, mdp->horizon, stream(), nullptr, nullptr, fn);
    if (rc == NOVA_SUCCESS)
        rc = mdp_vi_get_policy_gpu(mdp, vi, policy);
    const int rc2 = mdp_vi_uninitialize_gpu(mdp, vi);
    return (rc != NOVA_SUCCESS) ? rc : rc2;
}

int mdp_vi_run_gpu(const MDP *mdp, MDPVIGPU *vi, const float *d_V0, unsigned int sweeps,
                   float *elapsedMs)
{
    const char *fn = "mdp_vi_run_gpu";
    if (!mdp_device_ready(mdp) || vi == nullptr || vi->d_V == nullptr ||
            vi->d_VPrime == nullptr || vi->d_pi == nullptr || sweeps == 0) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    hipStream_t st = stream();
    if (d_V0 != nullptr) {
        const size_t bytes = sizeof(float) * mdp->n;
        hipError_t e = hipMemcpyAsync(vi->d_V, d_V0, bytes, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(vi->d_VPrime, d_V0, bytes, hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess)
            return hip_fail(e, fn, "Failed to reset V from d_V0.", NOVA_ERROR_MEMCPY_TO_DEVICE);
        vi->currentHorizon = 0;
    }
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (elapsedMs) {
        hipEventCreate(&t0);
        hipEventCreate(&t1);
    }
    int rc = vi_run_sweeps(mdp, vi, sweeps, st, t0, t1, fn);
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

} // namespace nova
