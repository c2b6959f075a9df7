// nova-mi355x: value iteration to a residual threshold and its adjoint (no reference
// counterpart: nova's VI runs a fixed horizon, mdp_vi_cpu.cpp:120-131, and has no gradients).
//
// Forward (mdp_vi_solve_gpu): float sweeps with the bit-exact kernels (on the sweep layout of
// the model's size) while the float residual max|V' - V| still falls, then FP64 sweeps (V in
// double, T and R promoted) until the FP64 residual is <= tol. FP32 alone cannot resolve residuals below ulp(|V|) (7.6e-6 at |V| = 78).
// The greedy policy of the last sweep is left in d_pi.
//
// Backward (mdp_vi_adjoint_gpu): at the converged greedy policy pi, L(V*) has
//   dL/dR[s,a]   = lambda(s) [a = pi(s)]
//   dL/dT[s,a,i] = gamma * lambda(s) * V*(S[s,a,i]) [a = pi(s)]
// with (I - gamma P_pi)^T lambda = rho, rho = dL/dV*. lambda is found by the fixed-point
// iteration lambda <- rho + gamma P_pi^T lambda in FP32 over a transposed successor list (the
// pull form, built once per call with a stable radix sort, so sums run in a fixed order and
// results are reproducible); the final residual is evaluated in FP64.
#include <cstring>
#include <vector>

#include <hipcub/hipcub.hpp>

#include <nova/amd/nova_amd.h>

#include "nova_hip.h"

namespace nova_amd {

int vi_sweeps_float(const nova::MDP *mdp, nova::MDPVIGPU *vi, unsigned sweeps, hipStream_t st,
                    const char *fn);   // mdp_vi_gpu.hip
int vi_layout_acquire(const nova::MDP *mdp, nova::MDPVIGPU *vi, hipStream_t st, const char *fn,
                      bool *built);    // mdp_vi_gpu.hip
void vi_layout_release(nova::MDPVIGPU *vi, const char *fn);   // mdp_vi_gpu.hip

namespace {

constexpr int kB = 256;

// One FP64 sweep, one (s,a) row per group of G lanes (tree sum within the group), max over
// actions by one lane per state through LDS. V/Vp are doubles; S/T/R as in the model.
template <int G>
__global__ __launch_bounds__(kB) void vi_sweep_f64(unsigned n, unsigned m, unsigned ns,
        unsigned statesPerTile, double gamma, const int *__restrict__ S,
        const float *__restrict__ T, const float *__restrict__ R, const double *__restrict__ V,
        double *__restrict__ Vp, unsigned *__restrict__ pi)
{
    __shared__ double q[1024];
    const unsigned s0 = blockIdx.x * statesPerTile;
    const unsigned s1 = min(n, s0 + statesPerTile);
    const unsigned rows = (s1 - s0) * m;
    const unsigned g = threadIdx.x % G;
    for (unsigned rr = threadIdx.x / G; rr < rows; rr += kB / G) {
        const size_t row = (size_t)s0 * m + rr;
        // valid prefix: entries after the first negative successor are ignored
        double acc = 0.0;
        bool stop = false;
        for (unsigned c = 0; c < ns; c += G) {
            const unsigned i = c + g;
            const int sp = (i < ns) ? S[row * ns + i] : -1;
            const unsigned long long neg = __ballot(i < ns && sp < 0);
            const unsigned lane = threadIdx.x % 64;
            const unsigned long long groupMask = ((G == 64) ? ~0ull : ((1ull << G) - 1ull))
                                                 << (lane - g);
            const unsigned long long below = ((1ull << g) - 1ull) << (lane - g);
            const bool before = (neg & below) != 0;
            if (!stop && !before && i < ns && sp >= 0)
                acc += gamma * (double)T[row * ns + i] * V[min((unsigned)sp, n - 1)];
            stop = stop || ((neg & groupMask) != 0);
        }
        for (int off = G / 2; off >= 1; off >>= 1)
            acc += __shfl_xor(acc, off, G);
        if (g == 0)
            q[rr] = (double)R[row] + acc;
    }
    __syncthreads();
    for (unsigned t = threadIdx.x; t < s1 - s0; t += kB) {
        const double *qs = q + (size_t)t * m;
        double best = qs[0];
        unsigned bestA = 0;
        for (unsigned a = 1; a < m; a++)
            if (qs[a] > best) {
                best = qs[a];
                bestA = a;
            }
        Vp[s0 + t] = best;
        pi[s0 + t] = bestA;
    }
}

template <typename VT>
__global__ __launch_bounds__(kB) void max_abs_diff(unsigned n, const VT *__restrict__ a,
        const VT *__restrict__ b, unsigned long long *__restrict__ out)
{
    double mx = 0.0;
    for (unsigned s = blockIdx.x * kB + threadIdx.x; s < n; s += gridDim.x * kB)
        mx = fmax(mx, fabs((double)a[s] - (double)b[s]));
    for (int off = 32; off >= 1; off >>= 1)
        mx = fmax(mx, __shfl_xor(mx, off));
    if (threadIdx.x % 64 == 0)
        atomicMax(out, (unsigned long long)__double_as_longlong(mx));
}

__global__ void to_double(unsigned n, const float *__restrict__ a, double *__restrict__ b)
{
    const unsigned s = blockIdx.x * kB + threadIdx.x;
    if (s < n)
        b[s] = (double)a[s];
}

// keys/slots of the entries (s, l) of the policy's rows: key = successor (n when invalid)
__global__ void adj_keys(unsigned n, unsigned m, unsigned ns, const int *__restrict__ S,
        const unsigned *__restrict__ pi, unsigned *__restrict__ keys, unsigned *__restrict__ slot)
{
    const size_t x = blockIdx.x * (size_t)kB + threadIdx.x;
    if (x >= (size_t)n * ns)
        return;
    const unsigned s = x / ns, l = x % ns;
    const size_t row = ((size_t)s * m + pi[s]) * ns;
    bool valid = true;
    for (unsigned q = 0; q <= l; q++)
        valid = valid && S[row + q] >= 0;
    keys[x] = valid ? (unsigned)S[row + l] : n;
    slot[x] = (unsigned)x;
}

__global__ void adj_starts(unsigned n, size_t E, const unsigned *__restrict__ keys,
        unsigned *__restrict__ start)
{
    const unsigned sp = blockIdx.x * kB + threadIdx.x;
    if (sp > n)
        return;
    size_t lo = 0, hi = E;   // first position with key >= sp
    while (lo < hi) {
        const size_t mid = (lo + hi) / 2;
        if (keys[mid] < sp) lo = mid + 1; else hi = mid;
    }
    start[sp] = (unsigned)lo;
}

__global__ void adj_gather_t(unsigned n, unsigned m, unsigned ns, size_t E,
        const float *__restrict__ T, const unsigned *__restrict__ pi,
        const unsigned *__restrict__ slot, unsigned *__restrict__ src, float *__restrict__ tv)
{
    const size_t q = blockIdx.x * (size_t)kB + threadIdx.x;
    if (q >= E)
        return;
    const unsigned x = slot[q];
    const unsigned s = x / ns, l = x % ns;
    src[q] = s;
    tv[q] = T[((size_t)s * m + pi[s]) * ns + l];
}

// lambda'(sp) = rho(sp) + gamma * sum_q T_q lambda(src_q), q over the predecessors of sp
__global__ void adj_step(unsigned n, float gamma, const unsigned *__restrict__ start,
        const unsigned *__restrict__ src, const float *__restrict__ tv,
        const float *__restrict__ rho, const float *__restrict__ lam, float *__restrict__ out)
{
    const unsigned sp = blockIdx.x * kB + threadIdx.x;
    if (sp >= n)
        return;
    float acc = 0.0f;
    for (unsigned q = start[sp]; q < start[sp + 1]; q++)
        acc += tv[q] * lam[src[q]];
    out[sp] = rho[sp] + gamma * acc;
}

// FP64 residual of the adjoint equation, per state (max taken by max_abs_diff-like reduction)
__global__ void adj_resid64(unsigned n, double gamma, const unsigned *__restrict__ start,
        const unsigned *__restrict__ src, const float *__restrict__ tv,
        const float *__restrict__ rho, const float *__restrict__ lam,
        unsigned long long *__restrict__ out)
{
    double mx = 0.0;
    for (unsigned sp = blockIdx.x * kB + threadIdx.x; sp < n; sp += gridDim.x * kB) {
        double acc = 0.0;
        for (unsigned q = start[sp]; q < start[sp + 1]; q++)
            acc += (double)tv[q] * (double)lam[src[q]];
        mx = fmax(mx, fabs((double)lam[sp] - (double)rho[sp] - gamma * acc));
    }
    for (int off = 32; off >= 1; off >>= 1)
        mx = fmax(mx, __shfl_xor(mx, off));
    if (threadIdx.x % 64 == 0)
        atomicMax(out, (unsigned long long)__double_as_longlong(mx));
}

__global__ void vi_grads(unsigned n, unsigned m, unsigned ns, float gamma,
        const int *__restrict__ S, const unsigned *__restrict__ pi,
        const float *__restrict__ lam, const double *__restrict__ V, float *__restrict__ dR,
        float *__restrict__ dT)
{
    const size_t x = blockIdx.x * (size_t)kB + threadIdx.x;
    if (x >= (size_t)n * m)
        return;
    const unsigned s = x / m, a = x % m;
    const bool on = (a == pi[s]);
    if (dR)
        dR[x] = on ? lam[s] : 0.0f;
    if (dT) {
        bool valid = true;
        for (unsigned l = 0; l < ns; l++) {
            const int sp = S[x * ns + l];
            valid = valid && sp >= 0;
            dT[x * ns + l] = (on && valid) ? (float)((double)gamma * (double)lam[s] * V[sp]) : 0.0f;
        }
    }
}

double read_max(unsigned long long *d, hipStream_t st)
{
    unsigned long long h = 0;
    hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, st);
    hipStreamSynchronize(st);
    double v;
    std::memcpy(&v, &h, sizeof(v));
    return v;
}

} // namespace

static unsigned pow2_at_least_u(unsigned x)
{
    unsigned p = 1;
    while (p < x)
        p <<= 1;
    return p;
}

// One FP64 sweep Va -> Vb (pi written).
static void sweep64(const nova::MDP *mdp, const double *Va, double *Vb, unsigned *pi,
                    hipStream_t st)
{
    const unsigned G = std::min(64u, pow2_at_least_u(std::min(mdp->ns, 64u)));
    const unsigned spt = std::max(1u, 1024u / mdp->m);
    const unsigned grid = (mdp->n + spt - 1) / spt;
    const double gamma = (double)mdp->gamma;
#define NOVA_SW64(GG)                                                                        \
    hipLaunchKernelGGL(vi_sweep_f64<GG>, dim3(grid), dim3(kB), 0, st, mdp->n, mdp->m, mdp->ns, \
                       spt, gamma, mdp->d_S, mdp->d_T, mdp->d_R, Va, Vb, pi)
    switch (G) {
    case 1: NOVA_SW64(1); break;
    case 2: NOVA_SW64(2); break;
    case 4: NOVA_SW64(4); break;
    case 8: NOVA_SW64(8); break;
    case 16: NOVA_SW64(16); break;
    case 32: NOVA_SW64(32); break;
    default: NOVA_SW64(64); break;
    }
#undef NOVA_SW64
}

} // namespace nova_amd

namespace nova {

using namespace nova_amd;

int mdp_vi_solve_gpu(const MDP *mdp, MDPVIGPU *vi, double *d_V64, double tol,
                     unsigned int maxSweeps, unsigned int flags, unsigned int *sweepsDone,
                     double *residual)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "mdp_vi_solve_gpu";
    if (mdp == nullptr || mdp->n == 0 || mdp->m == 0 || mdp->ns == 0 || mdp->d_S == nullptr ||
            mdp->d_T == nullptr || mdp->d_R == nullptr || mdp->m > 1024 || vi == nullptr ||
            vi->d_V == nullptr || vi->d_VPrime == nullptr || vi->d_pi == nullptr ||
            d_V64 == nullptr || !(tol > 0.0) || maxSweeps == 0) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    const unsigned n = mdp->n;
    hipStream_t st = stream();
    unsigned long long *dmax = nullptr;
    int rc = device_alloc((void **)&dmax, sizeof(unsigned long long), fn, "a residual");
    if (rc != NOVA_SUCCESS)
        return rc;
    unsigned done = 0;
    double res = 1e300;
    const unsigned check = 8;
    const unsigned rgrid = std::min(1024u, (n + kB - 1) / kB);
    if (!(flags & 1u)) {
        // float phase, on the sweep layout of the model's size (built for this call when the
        // state has none): continue while the float residual still falls (by 0.1% per check of
        // 8 sweeps; a contraction of gamma = 0.99 alone gives 7.7%), that is until it reaches
        // the rounding floor of float V
        bool built = false;
        rc = vi_layout_acquire(mdp, vi, st, fn, &built);
        double prev = 1e300;
        while (rc == NOVA_SUCCESS && done + check <= maxSweeps) {
            rc = vi_sweeps_float(mdp, vi, check, st, fn);
            if (rc != NOVA_SUCCESS)
                break;
            done += check;
            hipMemsetAsync(dmax, 0, sizeof(*dmax), st);
            hipLaunchKernelGGL(max_abs_diff<float>, dim3(rgrid), dim3(kB), 0, st, n, vi->d_V,
                               vi->d_VPrime, dmax);
            const double r = read_max(dmax, st);
            if (r <= tol || r > 0.999 * prev)
                break;
            prev = r;
        }
        if (built)
            vi_layout_release(vi, fn);
        const float *cur = (vi->currentHorizon % 2 == 0) ? vi->d_V : vi->d_VPrime;
        hipLaunchKernelGGL(to_double, dim3((n + kB - 1) / kB), dim3(kB), 0, st, n, cur, d_V64);
    }
    // FP64 phase: d_V64 holds V; the second half is the ping-pong partner. Stops at the
    // tolerance, or when the residual has not fallen for 4 checks (FP64 rounding floor).
    double *A = d_V64, *Bv = d_V64 + n;
    double bestRes = 1e300;
    int stall = 0;
    while (rc == NOVA_SUCCESS && done < maxSweeps) {
        unsigned k = 0;
        for (; k < check && done < maxSweeps; k++, done++) {
            sweep64(mdp, A, Bv, vi->d_pi, st);
            std::swap(A, Bv);
        }
        hipMemsetAsync(dmax, 0, sizeof(*dmax), st);
        hipLaunchKernelGGL(max_abs_diff<double>, dim3(rgrid), dim3(kB), 0, st, n, A, Bv, dmax);
        res = read_max(dmax, st);
        if (res <= tol)
            break;
        if (res < 0.999 * bestRes) {
            bestRes = res;
            stall = 0;
        } else if (++stall >= 4) {
            break;
        }
    }
    if (rc == NOVA_SUCCESS && A != d_V64)
        hipMemcpyAsync(d_V64, A, sizeof(double) * n, hipMemcpyDeviceToDevice, st);
    hipError_t e = hipStreamSynchronize(st);
    device_free(dmax, fn, "a residual");
    if (rc == NOVA_SUCCESS && e != hipSuccess)
        rc = hip_fail(e, fn, "Failed to run the sweeps.", NOVA_ERROR_KERNEL_EXECUTION);
    if (sweepsDone)
        *sweepsDone = done;
    if (residual)
        *residual = res;
    return rc;
}

int mdp_vi_adjoint_gpu(const MDP *mdp, const unsigned int *d_pi, const double *d_V64,
                       const float *d_rho, float *d_lambda, float *d_dR, float *d_dT,
                       float tol, unsigned int maxIters, unsigned int *itersDone,
                       double *residual64)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "mdp_vi_adjoint_gpu";
    if (mdp == nullptr || mdp->n == 0 || mdp->m == 0 || mdp->ns == 0 || mdp->d_S == nullptr ||
            mdp->d_T == nullptr || d_pi == nullptr || d_V64 == nullptr || d_rho == nullptr ||
            d_lambda == nullptr || maxIters == 0) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    const unsigned n = mdp->n, m = mdp->m, ns = mdp->ns;
    const size_t E = (size_t)n * ns;
    hipStream_t st = stream();
    // transposed successor list of the policy's rows (stable sort: fixed order per bucket)
    unsigned *keys = nullptr, *keys2 = nullptr, *slot = nullptr, *slot2 = nullptr;
    unsigned *start = nullptr, *src = nullptr;
    float *tv = nullptr, *lamB = nullptr;
    unsigned long long *dmax = nullptr;
    void *temp = nullptr;
    size_t tempBytes = 0;
    int rc = device_alloc((void **)&keys, E * 4, fn, "adjoint buffers");
    if (rc == NOVA_SUCCESS) rc = device_alloc((void **)&keys2, E * 4, fn, "adjoint buffers");
    if (rc == NOVA_SUCCESS) rc = device_alloc((void **)&slot, E * 4, fn, "adjoint buffers");
    if (rc == NOVA_SUCCESS) rc = device_alloc((void **)&slot2, E * 4, fn, "adjoint buffers");
    if (rc == NOVA_SUCCESS) rc = device_alloc((void **)&start, (n + 1) * 4ull, fn, "adjoint buffers");
    if (rc == NOVA_SUCCESS) rc = device_alloc((void **)&src, E * 4, fn, "adjoint buffers");
    if (rc == NOVA_SUCCESS) rc = device_alloc((void **)&tv, E * 4, fn, "adjoint buffers");
    if (rc == NOVA_SUCCESS) rc = device_alloc((void **)&lamB, n * 4ull, fn, "adjoint buffers");
    if (rc == NOVA_SUCCESS) rc = device_alloc((void **)&dmax, 8, fn, "adjoint buffers");
    if (rc == NOVA_SUCCESS) {
        const unsigned gE = (unsigned)((E + kB - 1) / kB);
        hipLaunchKernelGGL(adj_keys, dim3(gE), dim3(kB), 0, st, n, m, ns, mdp->d_S, d_pi, keys, slot);
        hipcub::DeviceRadixSort::SortPairs(nullptr, tempBytes, keys, keys2, slot, slot2, (int)E,
                                           0, 32, st);
        rc = device_alloc(&temp, tempBytes, fn, "adjoint sort storage");
        if (rc == NOVA_SUCCESS) {
            hipcub::DeviceRadixSort::SortPairs(temp, tempBytes, keys, keys2, slot, slot2, (int)E,
                                               0, 32, st);
            hipLaunchKernelGGL(adj_starts, dim3((n + 1 + kB - 1) / kB), dim3(kB), 0, st, n, E,
                               keys2, start);
            hipLaunchKernelGGL(adj_gather_t, dim3(gE), dim3(kB), 0, st, n, m, ns, E, mdp->d_T,
                               d_pi, slot2, src, tv);
        }
    }
    unsigned it = 0;
    double res64 = 0.0;
    if (rc == NOVA_SUCCESS) {
        const unsigned g = (n + kB - 1) / kB;
        const unsigned rgrid = std::min(1024u, g);
        hipMemcpyAsync(d_lambda, d_rho, n * 4ull, hipMemcpyDeviceToDevice, st);
        float *a = d_lambda, *b = lamB;
        const unsigned check = 16;
        while (it < maxIters) {
            for (unsigned k = 0; k < check && it < maxIters; k++, it++) {
                hipLaunchKernelGGL(adj_step, dim3(g), dim3(kB), 0, st, n, mdp->gamma, start, src,
                                   tv, d_rho, a, b);
                std::swap(a, b);
            }
            hipMemsetAsync(dmax, 0, 8, st);
            hipLaunchKernelGGL(max_abs_diff<float>, dim3(rgrid), dim3(kB), 0, st, n, a, b, dmax);
            const double d = read_max(dmax, st);
            hipMemsetAsync(dmax, 0, 8, st);
            hipLaunchKernelGGL(max_abs_diff<float>, dim3(rgrid), dim3(kB), 0, st, n, a,
                               (const float *)d_rho, dmax);
            const double scale = read_max(dmax, st);   // max|lambda - rho|, a scale of lambda
            if (d <= (double)tol * std::max(1e-30, scale))
                break;
        }
        if (a != d_lambda)
            hipMemcpyAsync(d_lambda, a, n * 4ull, hipMemcpyDeviceToDevice, st);
        hipMemsetAsync(dmax, 0, 8, st);
        hipLaunchKernelGGL(adj_resid64, dim3(rgrid), dim3(kB), 0, st, n, (double)mdp->gamma,
                           start, src, tv, d_rho, d_lambda, dmax);
        res64 = read_max(dmax, st);
        if (d_dR || d_dT)
            hipLaunchKernelGGL(vi_grads, dim3((unsigned)(((size_t)n * m + kB - 1) / kB)),
                               dim3(kB), 0, st, n, m, ns, mdp->gamma, mdp->d_S, d_pi, d_lambda,
                               d_V64, d_dR, d_dT);
        hipError_t e = hipStreamSynchronize(st);
        if (e != hipSuccess)
            rc = hip_fail(e, fn, "Failed to run the adjoint.", NOVA_ERROR_KERNEL_EXECUTION);
    }
    void *bufs[] = {keys, keys2, slot, slot2, start, src, tv, lamB, dmax, temp};
    for (void *p : bufs)
        if (p)
            device_free(p, fn, "adjoint buffers");
    if (itersDone)
        *itersDone = it;
    if (residual64)
        *residual64 = res64;
    return rc;
}

} // namespace nova
