// nova-mi355x: value iteration to a residual threshold and its adjoint (no reference
// counterpart: nova's VI runs a fixed horizon, mdp_vi_cpu.cpp:120-131, and has no gradients).
//
// Forward (mdp_vi_solve_gpu): FP64 sweeps (V in double, T and R promoted) until the residual
// max|V' - V| is <= tol. Between blocks of 4 sweeps the iterate is moved along the slow mode
// of the error when the last sweep's step shows it (SolveChecks): at gamma = 0.99 plain
// sweeps shrink that mode by 1% each, about 1,400 sweeps to 1e-6, and the moves leave the
// dozen or two sweeps the faster modes need. Plain sweeps (flags bit 1) run first in float
// with the bit-exact kernels (on the state's sweep layout if it has one) while the float
// residual still falls; FP32 alone cannot resolve residuals below ulp(|V|) (7.6e-6 at
// |V| = 78). The greedy policy of the last sweep is left in d_pi.
//
// Backward (mdp_vi_adjoint_gpu): at the converged greedy policy pi, L(V*) has
//   dL/dR[s,a]   = lambda(s) [a = pi(s)]
//   dL/dT[s,a,i] = gamma * lambda(s) * V*(S[s,a,i]) [a = pi(s)]
// with (I - gamma P_pi)^T lambda = rho, rho = dL/dV*. lambda is found by the fixed-point
// iteration lambda <- rho + gamma P_pi^T lambda in FP64 over a transposed successor list (the
// pull form, built once per call with a stable radix sort, so sums run in a fixed order and
// results are reproducible), checked every 8 steps, with the slow mode's remaining steps added
// once the last two steps isolate it; the final residual is evaluated in FP64.
#include <cmath>
#include <cstring>
#include <vector>

#include <hipcub/hipcub.hpp>

#include <nova/amd/nova_amd.h>

#include "nova_hip.h"

namespace nova_amd {

int vi_sweeps_float(const nova::MDP *mdp, nova::MDPVIGPU *vi, unsigned sweeps, hipStream_t st,
                    const char *fn);   // mdp_vi_gpu.hip

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

// Maximum over a workgroup (every lane returns it). The reductions below take it before their
// one atomic per workgroup: an atomic per wave on one address serialises thousands of them.
__device__ __forceinline__ double block_max(double v, double *lds)
{
    for (int off = 32; off >= 1; off >>= 1)
        v = fmax(v, __shfl_xor(v, off));
    __syncthreads();
    if (threadIdx.x % 64 == 0)
        lds[threadIdx.x / 64] = v;
    __syncthreads();
    double t = lds[0];
    for (unsigned i = 1; i < kB / 64; i++)
        t = fmax(t, lds[i]);
    return t;
}

// Workgroups of the reductions (grid-stride loops over the states)
constexpr unsigned kRedBlocks = 256;

template <typename VT>
__global__ __launch_bounds__(kB) void max_abs_diff(unsigned n, const VT *__restrict__ a,
        const VT *__restrict__ b, unsigned long long *__restrict__ out)
{
    double mx = 0.0;
    for (unsigned s = blockIdx.x * kB + threadIdx.x; s < n; s += gridDim.x * kB)
        mx = fmax(mx, fabs((double)a[s] - (double)b[s]));
    __shared__ double lds[kB / 64];
    mx = block_max(mx, lds);
    if (threadIdx.x == 0)
        atomicMax(out, (unsigned long long)__double_as_longlong(mx));
}

// Order-preserving key of a double for atomicMax (negative values below positive ones), and
// its inverse.
__device__ __forceinline__ unsigned long long order_key(double x)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

double from_order_key(unsigned long long k)
{
    const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    double v;
    std::memcpy(&v, &b, sizeof(v));
    return v;
}

// The extremes of the signed sweep difference d = cur - prev: out[0] = key(max d),
// out[1] = key(max -d), both zeroed by the caller.
template <typename VT>
__global__ __launch_bounds__(kB) void diff_extremes(unsigned n, const VT *__restrict__ cur,
        const VT *__restrict__ prev, unsigned long long *__restrict__ out)
{
    double hi = -INFINITY, lo = -INFINITY;
    for (unsigned s = blockIdx.x * kB + threadIdx.x; s < n; s += gridDim.x * kB) {
        const double d = (double)cur[s] - (double)prev[s];
        hi = fmax(hi, d);
        lo = fmax(lo, -d);
    }
    __shared__ double lds[kB / 64];
    hi = block_max(hi, lds);
    lo = block_max(lo, lds);
    if (threadIdx.x == 0 && hi > -INFINITY) {
        atomicMax(&out[0], order_key(hi));
        atomicMax(&out[1], order_key(lo));
    }
}

template <typename VT>
__global__ void add_const(unsigned n, VT *__restrict__ v, double c)
{
    const unsigned s = blockIdx.x * kB + threadIdx.x;
    if (s < n)
        v[s] = (VT)((double)v[s] + c);
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

// lambda'(sp) = rho(sp) + gamma * sum_q T_q lambda(src_q), q over the predecessors of sp (FP64)
__global__ void adj_step(unsigned n, double gamma, const unsigned *__restrict__ start,
        const unsigned *__restrict__ src, const float *__restrict__ tv,
        const double *__restrict__ rho, const double *__restrict__ lam, double *__restrict__ out)
{
    const unsigned sp = blockIdx.x * kB + threadIdx.x;
    if (sp >= n)
        return;
    double acc = 0.0;
    for (unsigned q = start[sp]; q < start[sp + 1]; q++)
        acc += (double)tv[q] * lam[src[q]];
    out[sp] = rho[sp] + gamma * acc;
}

// Sums of the last two adjoint steps, d1 = lam_k - lam_{k-1} and d2 = lam_{k+1} - lam_k, in a
// fixed order (per-workgroup partials in part[2 * block + {0, 1}], then adj_sum_final), so
// that the extrapolation below, and lambda with it, is reproducible.
constexpr int kSumBlocks = 256;

__device__ __forceinline__ double block_sum(double v, double *lds)
{
    for (int off = 32; off >= 1; off >>= 1)
        v += __shfl_xor(v, off);
    const unsigned w = threadIdx.x / 64;
    __syncthreads();
    if (threadIdx.x % 64 == 0)
        lds[w] = v;
    __syncthreads();
    double t = 0.0;
    for (unsigned i = 0; i < kB / 64; i++)
        t += lds[i];
    return t;
}

__global__ __launch_bounds__(kB) void adj_sum_partials(unsigned n,
        const double *__restrict__ l0, const double *__restrict__ l1,
        const double *__restrict__ l2, double *__restrict__ part)
{
    __shared__ double lds[kB / 64];
    double s1 = 0.0, s2 = 0.0;
    for (unsigned s = blockIdx.x * kB + threadIdx.x; s < n; s += gridDim.x * kB) {
        s1 += l1[s] - l0[s];
        s2 += l2[s] - l1[s];
    }
    s1 = block_sum(s1, lds);
    s2 = block_sum(s2, lds);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s1;
        part[2 * blockIdx.x + 1] = s2;
    }
}

__global__ __launch_bounds__(kB) void adj_sum_final(unsigned blocks,
        const double *__restrict__ part, double *__restrict__ out)
{
    __shared__ double lds[kB / 64];
    double s1 = 0.0, s2 = 0.0;
    for (unsigned b = threadIdx.x; b < blocks; b += kB) {
        s1 += part[2 * b];
        s2 += part[2 * b + 1];
    }
    s1 = block_sum(s1, lds);
    s2 = block_sum(s2, lds);
    if (threadIdx.x == 0) {
        out[0] = s1;
        out[1] = s2;
    }
}

// max |d2 - q d1| over the states (the part of the last step outside the ratio-q mode)
__global__ __launch_bounds__(kB) void adj_mode_dev(unsigned n, const double *__restrict__ l0,
        const double *__restrict__ l1, const double *__restrict__ l2, double q,
        unsigned long long *__restrict__ out)
{
    double mx = 0.0;
    for (unsigned s = blockIdx.x * kB + threadIdx.x; s < n; s += gridDim.x * kB) {
        const double d1 = l1[s] - l0[s], d2 = l2[s] - l1[s];
        mx = fmax(mx, fabs(d2 - q * d1));
    }
    __shared__ double lds[kB / 64];
    mx = block_max(mx, lds);
    if (threadIdx.x == 0)
        atomicMax(out, (unsigned long long)__double_as_longlong(mx));
}

// out = l2 + (l2 - l1) f: the sum of the remaining steps of a mode that shrinks by q per step,
// f = q / (1 - q)
__global__ void adj_extrapolate(unsigned n, const double *__restrict__ l1,
        const double *__restrict__ l2, double f, double *__restrict__ out)
{
    const unsigned s = blockIdx.x * kB + threadIdx.x;
    if (s < n)
        out[s] = l2[s] + (l2[s] - l1[s]) * f;
}

__global__ void to_float(unsigned n, const double *__restrict__ a, float *__restrict__ b)
{
    const unsigned s = blockIdx.x * kB + threadIdx.x;
    if (s < n)
        b[s] = (float)a[s];
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
    __shared__ double lds[kB / 64];
    mx = block_max(mx, lds);
    if (threadIdx.x == 0)
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

// Sweep-difference kernels of the solve's checks. d = cur - prv is the last sweep's step;
// dprev (doubles) holds the step of the check before.
template <typename VT>
__global__ void diff_store(unsigned n, const VT *__restrict__ cur, const VT *__restrict__ prv,
        double *__restrict__ dprev)
{
    const unsigned s = blockIdx.x * kB + threadIdx.x;
    if (s < n)
        dprev[s] = (double)cur[s] - (double)prv[s];
}

template <typename VT>
__global__ __launch_bounds__(kB) void ratio_sum_partials(unsigned n, const VT *__restrict__ cur,
        const VT *__restrict__ prv, const double *__restrict__ dprev, double *__restrict__ part)
{
    __shared__ double lds[kB / 64];
    double s1 = 0.0, s2 = 0.0;
    for (unsigned s = blockIdx.x * kB + threadIdx.x; s < n; s += gridDim.x * kB) {
        s1 += dprev[s];
        s2 += (double)cur[s] - (double)prv[s];
    }
    s1 = block_sum(s1, lds);
    s2 = block_sum(s2, lds);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s1;
        part[2 * blockIdx.x + 1] = s2;
    }
}

template <typename VT>
__global__ __launch_bounds__(kB) void ratio_dev(unsigned n, const VT *__restrict__ cur,
        const VT *__restrict__ prv, const double *__restrict__ dprev, double q,
        unsigned long long *__restrict__ out)
{
    double mx = 0.0;
    for (unsigned s = blockIdx.x * kB + threadIdx.x; s < n; s += gridDim.x * kB)
        mx = fmax(mx, fabs(((double)cur[s] - (double)prv[s]) - q * dprev[s]));
    __shared__ double lds[kB / 64];
    mx = block_max(mx, lds);
    if (threadIdx.x == 0)
        atomicMax(out, (unsigned long long)__double_as_longlong(mx));
}

template <typename VT>
__global__ void ratio_extrapolate(unsigned n, VT *__restrict__ cur, const VT *__restrict__ prv,
        double f)
{
    const unsigned s = blockIdx.x * kB + threadIdx.x;
    if (s < n)
        cur[s] = (VT)((double)cur[s] + ((double)cur[s] - (double)prv[s]) * f);
}

// The checks of the residual-stopped solve and what they may do to the iterate between blocks
// of sweeps. measure() gives the residual max|d| of the block's last sweep, d = V_k - V_{k-1}
// = T V_{k-1} - V_{k-1}; accelerate() then moves V_k towards the fixed point when the step
// shows which way:
// * d of one sign and nearly uniform (its spread at most a quarter of its smallest magnitude):
//   with every row of T summing to 1, T(V + c) = TV + gamma c, so V* lies within
//   [V_k + g dmin, V_k + g dmax], g = gamma / (1 - gamma) (MacQueen's bounds), and V_k moves
//   to the middle, which bounds its error by g (dmax - dmin) / 2, never more than V_k's own
//   bound g max(|dmin|, |dmax|). The error then lay along the constant vector, the mode that
//   sweeps shrink by gamma alone; the other modes had already died out.
// * otherwise (rows summing below 1, as when successor lists are cut short), once the policy
//   has settled the sweeps are linear and the slowest mode of gamma P_pi is left: d is then
//   qb = sum(d) / sum(dprev) times the step of the check before, b sweeps earlier (block), up
//   to a tenth of max|d|, and the remaining steps of that mode sum to d q / (1 - q),
//   q = qb^(1/b).
// Either move is followed by at least one more check of plain sweeps, and the solve stops on
// the residual alone, so a move can cost sweeps but never accuracy. The sums run in a fixed
// order, so the solve is reproducible.
struct SolveChecks {
    unsigned n = 0;
    double gamma = 0.0;
    unsigned long long *dext = nullptr;   // [max d key, max -d key, scratch max]
    double *dprev = nullptr;              // n doubles
    double *part = nullptr;               // 2 kSumBlocks partials + 2 sums
    bool havePrev = false;
    unsigned block = 8;                   // sweeps between two checks
    double dmin = 0.0, dmax = 0.0;

    int alloc(const char *fn)
    {
        int rc = device_alloc((void **)&dext, 3 * sizeof(unsigned long long), fn, "a residual");
        if (rc == NOVA_SUCCESS)
            rc = device_alloc((void **)&dprev, sizeof(double) * n, fn, "the solve's checks");
        if (rc == NOVA_SUCCESS)
            rc = device_alloc((void **)&part, sizeof(double) * (2 * kSumBlocks + 2), fn,
                              "the solve's checks");
        return rc;
    }
    void release(const char *fn)
    {
        void *bufs[] = {dext, dprev, part};
        for (void *p : bufs)
            if (p)
                device_free(p, fn, "the solve's checks");
    }
    unsigned grid() const { return std::min(kRedBlocks, (n + kB - 1) / kB); }

    template <typename VT>
    double measure(const VT *cur, const VT *prv, hipStream_t st)
    {
        hipMemsetAsync(dext, 0, 2 * sizeof(unsigned long long), st);
        hipLaunchKernelGGL(diff_extremes<VT>, dim3(grid()), dim3(kB), 0, st, n, cur, prv, dext);
        unsigned long long h[2] = {0, 0};
        hipMemcpyAsync(h, dext, sizeof(h), hipMemcpyDeviceToHost, st);
        hipStreamSynchronize(st);
        dmax = from_order_key(h[0]);
        dmin = -from_order_key(h[1]);
        return std::max(std::fabs(dmin), std::fabs(dmax));
    }

    template <typename VT>
    void accelerate(VT *cur, const VT *prv, hipStream_t st)
    {
        const unsigned eg = (n + kB - 1) / kB;
        const double res = std::max(std::fabs(dmin), std::fabs(dmax));
        const bool oneSign = (dmin > 0.0 && dmax > 0.0) || (dmin < 0.0 && dmax < 0.0);
        if (oneSign && gamma < 1.0 &&
                dmax - dmin <= 0.25 * std::min(std::fabs(dmin), std::fabs(dmax))) {
            hipLaunchKernelGGL(add_const<VT>, dim3(eg), dim3(kB), 0, st, n, cur,
                               gamma / (1.0 - gamma) * 0.5 * (dmin + dmax));
            havePrev = false;
            return;
        }
        if (havePrev) {
            const unsigned sb = std::min((unsigned)kSumBlocks, eg);
            hipLaunchKernelGGL(ratio_sum_partials<VT>, dim3(sb), dim3(kB), 0, st, n,
                               (const VT *)cur, prv, (const double *)dprev, part);
            hipLaunchKernelGGL(adj_sum_final, dim3(1), dim3(kB), 0, st, sb, part,
                               part + 2 * kSumBlocks);
            double sums[2] = {0.0, 0.0};
            hipMemcpyAsync(sums, part + 2 * kSumBlocks, sizeof(sums), hipMemcpyDeviceToHost, st);
            hipStreamSynchronize(st);
            const double qb = (sums[0] != 0.0) ? sums[1] / sums[0] : 0.0;
            if (qb > 0.0 && qb < 1.0) {
                hipMemsetAsync(dext + 2, 0, sizeof(unsigned long long), st);
                hipLaunchKernelGGL(ratio_dev<VT>, dim3(grid()), dim3(kB), 0, st, n,
                                   (const VT *)cur, prv, (const double *)dprev, qb, dext + 2);
                if (read_max(dext + 2, st) <= 0.1 * res) {
                    const double q = std::pow(qb, 1.0 / (double)block);
                    hipLaunchKernelGGL(ratio_extrapolate<VT>, dim3(eg), dim3(kB), 0, st, n, cur,
                                       prv, q / (1.0 - q));
                    havePrev = false;
                    return;
                }
            }
        }
        hipLaunchKernelGGL(diff_store<VT>, dim3(eg), dim3(kB), 0, st, n, (const VT *)cur, prv,
                           dprev);
        havePrev = true;
    }
};

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
    SolveChecks chk;
    chk.n = n;
    chk.gamma = (double)mdp->gamma;
    int rc = chk.alloc(fn);
    if (rc != NOVA_SUCCESS) {
        chk.release(fn);
        return rc;
    }
    unsigned done = 0;
    double res = 1e300;
    const unsigned eg = (n + kB - 1) / kB;
    const bool shifts = !(flags & 2u);     // bit 1: plain sweeps, no moves between checks
    // With the moves, a solve is a few dozen sweeps, all in FP64 (at n = 200,000 an FP64 sweep
    // costs no more than a float one without a sweep layout, and FP64 keeps the steps the moves
    // read free of float rounding), checked every 4. Plain sweeps first run in float on the
    // state's sweep layout, checked every 8.
    const unsigned check = shifts ? 4 : 8;
    chk.block = check;
    if (!(flags & 1u) && shifts) {
        hipLaunchKernelGGL(to_double, dim3(eg), dim3(kB), 0, st, n, (const float *)vi->d_V,
                           d_V64);
    } else if (!(flags & 1u)) {
        // float phase: continue while the float residual still falls (by 0.1% per check of 8
        // sweeps), that is until it reaches the rounding floor of float V
        double prev = 1e300;
        while (rc == NOVA_SUCCESS && done + check <= maxSweeps) {
            rc = vi_sweeps_float(mdp, vi, check, st, fn);
            if (rc != NOVA_SUCCESS)
                break;
            done += check;
            const bool even = vi->currentHorizon % 2 == 0;
            float *cur = even ? vi->d_V : vi->d_VPrime;
            const float *prv = even ? vi->d_VPrime : vi->d_V;
            const double r = chk.measure<float>(cur, prv, st);
            if (r <= tol || r > 0.999 * prev)
                break;
            prev = r;
        }
        const float *cur = (vi->currentHorizon % 2 == 0) ? vi->d_V : vi->d_VPrime;
        hipLaunchKernelGGL(to_double, dim3(eg), dim3(kB), 0, st, n, cur, d_V64);
    }
    // FP64 phase: d_V64 holds V; the second half is the ping-pong partner. Stops at the
    // tolerance, or when the residual has not fallen for 4 checks (FP64 rounding floor).
    double *A = d_V64, *Bv = d_V64 + n;
    double bestRes = 1e300;
    int stall = 0;
    chk.havePrev = false;
    while (rc == NOVA_SUCCESS && done < maxSweeps) {
        unsigned k = 0;
        for (; k < check && done < maxSweeps; k++, done++) {
            sweep64(mdp, A, Bv, vi->d_pi, st);
            std::swap(A, Bv);
        }
        res = chk.measure<double>(A, Bv, st);
        if (res <= tol)
            break;
        if (res < 0.999 * bestRes) {
            bestRes = res;
            stall = 0;
        } else if (++stall >= 4) {
            break;
        }
        if (shifts && done < maxSweeps)
            chk.accelerate<double>(A, Bv, st);
    }
    if (rc == NOVA_SUCCESS && A != d_V64)
        hipMemcpyAsync(d_V64, A, sizeof(double) * n, hipMemcpyDeviceToDevice, st);
    hipError_t e = hipStreamSynchronize(st);
    chk.release(fn);
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
    float *tv = nullptr;
    double *lam64 = nullptr, *part = nullptr;   // lam64: rho and three iterates, n doubles each
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
    if (rc == NOVA_SUCCESS)
        rc = device_alloc((void **)&lam64, 4ull * n * sizeof(double), fn, "adjoint buffers");
    if (rc == NOVA_SUCCESS)
        rc = device_alloc((void **)&part, sizeof(double) * (2 * kSumBlocks + 2), fn,
                          "adjoint buffers");
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
        const unsigned rgrid = std::min(kRedBlocks, g);
        // FP64 iterates: the moves below read differences of consecutive steps, which FP32
        // rounding (about 6e-8 of lambda) would bury before the tolerance is reached
        double *rho64 = lam64;
        hipLaunchKernelGGL(to_double, dim3(g), dim3(kB), 0, st, n, d_rho, rho64);
        hipMemcpyAsync(lam64 + n, rho64, n * sizeof(double), hipMemcpyDeviceToDevice, st);
        // cur = lambda_{k+1}, prv = lambda_k, old = lambda_{k-1}
        double *cur = lam64 + n, *prv = lam64 + 2 * n, *old = lam64 + 3 * n;
        const unsigned check = 8;
        const unsigned sb = std::min((unsigned)kSumBlocks, g);
        while (it < maxIters) {
            unsigned k = 0;
            for (; k < check && it < maxIters; k++, it++) {
                hipLaunchKernelGGL(adj_step, dim3(g), dim3(kB), 0, st, n, (double)mdp->gamma,
                                   start, src, tv, (const double *)rho64, (const double *)cur,
                                   old);
                double *t3 = old;
                old = prv;
                prv = cur;
                cur = t3;
            }
            hipMemsetAsync(dmax, 0, 8, st);
            hipLaunchKernelGGL(max_abs_diff<double>, dim3(rgrid), dim3(kB), 0, st, n, cur, prv,
                               dmax);
            const double d = read_max(dmax, st);
            hipMemsetAsync(dmax, 0, 8, st);
            hipLaunchKernelGGL(max_abs_diff<double>, dim3(rgrid), dim3(kB), 0, st, n, cur,
                               (const double *)rho64, dmax);
            const double scale = read_max(dmax, st);   // max|lambda - rho|, a scale of lambda
            if (d <= (double)tol * std::max(1e-30, scale) || it >= maxIters || k < 2)
                break;
            // The error of the iteration lambda <- rho + gamma P^T lambda is a sum of modes that
            // shrink by gamma times the eigenvalues of P^T per step; with P's rows summing to 1
            // the slowest is the stationary distribution's, at gamma, and the others die out
            // much faster. Once the last step d2 is q = sum(d2) / sum(d1) times the step d1
            // before it, up to a tenth of max|d2|, what is left of the error is that one mode,
            // and its remaining steps sum to d2 q / (1 - q): add them.
            hipLaunchKernelGGL(adj_sum_partials, dim3(sb), dim3(kB), 0, st, n,
                               (const double *)old, (const double *)prv, (const double *)cur,
                               part);
            hipLaunchKernelGGL(adj_sum_final, dim3(1), dim3(kB), 0, st, sb, part,
                               part + 2 * kSumBlocks);
            double sums[2] = {0.0, 0.0};
            hipMemcpyAsync(sums, part + 2 * kSumBlocks, sizeof(sums), hipMemcpyDeviceToHost, st);
            hipStreamSynchronize(st);
            if (!(sums[0] != 0.0))
                continue;
            const double q = sums[1] / sums[0];
            if (!(q > 0.0 && q < 1.0))
                continue;
            hipMemsetAsync(dmax, 0, 8, st);
            hipLaunchKernelGGL(adj_mode_dev, dim3(rgrid), dim3(kB), 0, st, n,
                               (const double *)old, (const double *)prv, (const double *)cur, q,
                               dmax);
            if (read_max(dmax, st) <= 0.1 * d) {
                hipLaunchKernelGGL(adj_extrapolate, dim3(g), dim3(kB), 0, st, n,
                                   (const double *)prv, (const double *)cur, q / (1.0 - q), old);
                std::swap(cur, old);
            }
        }
        hipLaunchKernelGGL(to_float, dim3(g), dim3(kB), 0, st, n, (const double *)cur, d_lambda);
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
    void *bufs[] = {keys, keys2, slot, slot2, start, src, tv, lam64, part, dmax, temp};
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
