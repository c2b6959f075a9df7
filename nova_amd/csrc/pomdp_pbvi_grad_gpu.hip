// nova-mi355x: gradient of the PBVI objective L = sum_b max_{alpha in Gamma_H} alpha . b with
// respect to R by replaying recorded argmax choices (no reference counterpart).
//
// With the choices of every iteration fixed (recorded by pomdp_pbvi_run_recorded_gpu: the
// action a_i^t chosen for alpha i at iteration t and its alpha indices j*_{i,o}^t), the PBVI
// iteration is linear in R:
//   alpha_{t+1,i} = R[:, a_i] + sum_o gamma P_{a_i,o} alpha_{t, j*_{i,o}},
//   P_{a,o}[s, s'] = sum_l [S(s,a,l) = s'] T(s,a,l) O(a,s',o),
// and alpha_0 = GammaInitial does not depend on R. Backward, with mu_H[j] = sum of the beliefs
// whose maximiser at the end is alpha j:
//   dL/dR[:, a]  += sum_{i : a_i^t = a} mu_{t+1}[i]
//   mu_t[j]       = sum_{(i,o) : j*_{i,o}^t = j} gamma O(a_i,.,o) * (T_{a_i}^T mu_{t+1}[i])
//                   (+ mu_{t+1}[j] when no action was chosen for j: its alpha was kept)
// Every sum runs over lists built on the host in a fixed order, so results are reproducible.
#include <algorithm>
#include <cstring>
#include <vector>

#include <nova/amd/nova_amd.h>

#include "nova_hip.h"

namespace nova_amd {

namespace {

constexpr int kB = 256;

// values[i] = max_j Gamma[j] . b_i over the sparse support, first maximiser from -1e10.
__global__ __launch_bounds__(kB) void belief_values(unsigned n, unsigned rz, unsigned rG,
        const int *__restrict__ Z, const float *__restrict__ B, const float *__restrict__ G,
        float *__restrict__ values, int *__restrict__ arg)
{
    __shared__ float bv[kB];
    __shared__ int bj[kB];
    const unsigned i = blockIdx.x, t = threadIdx.x;
    float best = kLow;
    int bestJ = 0x7fffffff;
    for (unsigned j = t; j < rG; j += kB) {
        float v = 0.0f;
        for (unsigned k = 0; k < rz; k++) {
            const int s = Z[(size_t)i * rz + k];
            if (s < 0)
                break;
            v += B[(size_t)i * rz + k] * G[(size_t)j * n + s];
        }
        if (v > best) {
            best = v;
            bestJ = (int)j;
        }
    }
    bv[t] = best;
    bj[t] = bestJ;
    __syncthreads();
    if (t == 0) {
        float b = kLow;
        int jj = 0x7fffffff;
        for (int q = 0; q < kB; q++)
            if (bv[q] > b || (bv[q] == b && bj[q] < jj)) {
                b = bv[q];
                jj = bj[q];
            }
        values[i] = b;
        arg[i] = (jj == 0x7fffffff) ? 0 : jj;
    }
}

// mu[j] = sum of the beliefs listed for j (beliefs in list order, support entries in parallel)
__global__ __launch_bounds__(kB) void mu_init(unsigned n, unsigned rz, const int *__restrict__ Z,
        const float *__restrict__ B, const unsigned *__restrict__ start,
        const unsigned *__restrict__ list, float *__restrict__ mu)
{
    const unsigned j = blockIdx.x, t = threadIdx.x;
    float *row = mu + (size_t)j * n;
    for (unsigned s = t; s < n; s += kB)
        row[s] = 0.0f;
    __syncthreads();
    for (unsigned q = start[j]; q < start[j + 1]; q++) {
        const unsigned i = list[q];
        for (unsigned k = t; k < rz; k += kB) {
            const int s = Z[(size_t)i * rz + k];
            bool valid = s >= 0;
            for (unsigned k2 = 0; k2 < k && valid; k2++)
                valid = Z[(size_t)i * rz + k2] >= 0;
            if (valid)
                row[s] += B[(size_t)i * rz + k];
        }
        __syncthreads();
    }
}

// u[i][s'] = sum over predecessors (s, l) of s' under a_i of T(s,a_i,l) mu[i][s]
__global__ __launch_bounds__(kB) void pred_gather(unsigned n, unsigned r,
        const int *__restrict__ act, const unsigned *__restrict__ pstart,
        const int *__restrict__ pS, const float *__restrict__ pT, const float *__restrict__ mu,
        float *__restrict__ u)
{
    const unsigned i = blockIdx.y;
    const unsigned sp = blockIdx.x * kB + threadIdx.x;
    if (sp >= n)
        return;
    const int a = act[i];
    float acc = 0.0f;
    if (a >= 0) {
        const unsigned *ps = pstart + (size_t)a * (n + 1);
        const float *mi = mu + (size_t)i * n;
        for (unsigned q = ps[sp]; q < ps[sp + 1]; q++)
            acc += pT[q] * mi[pS[q]];
    }
    u[(size_t)i * n + sp] = acc;
}

// dR[s][a] += sum over alphas i with a_i = a of mu[i][s]  (lists per action)
__global__ __launch_bounds__(kB) void dr_accumulate(unsigned n, unsigned m,
        const unsigned *__restrict__ astart, const unsigned *__restrict__ alist,
        const float *__restrict__ mu, float *__restrict__ dR)
{
    const unsigned a = blockIdx.y;
    const unsigned s = blockIdx.x * kB + threadIdx.x;
    if (s >= n)
        return;
    float acc = dR[(size_t)s * m + a];
    for (unsigned q = astart[a]; q < astart[a + 1]; q++)
        acc += mu[(size_t)alist[q] * n + s];
    dR[(size_t)s * m + a] = acc;
}

// OTr[(a z + o) n + s'] = O[(a n + s') z + o]: mu_partial reads O with the states contiguous
__global__ void o_transpose(unsigned n, unsigned m, unsigned z, const float *__restrict__ O,
        float *__restrict__ OTr)
{
    const size_t x = blockIdx.x * (size_t)kB + threadIdx.x;
    if (x >= (size_t)m * n * z)
        return;
    const unsigned o = x % z;
    const size_t as = x / z;                    // a * n + s'
    const unsigned a = (unsigned)(as / n), sp = (unsigned)(as % n);
    OTr[((size_t)a * z + o) * n + sp] = O[x];
}

// The (i, o) pairs listed for an alpha j, cut into segments of at most kSeg pairs so that a
// much-chosen alpha is summed by many workgroups: partial[seg][s'] = sum over the segment's
// (i, o) of gamma O(a_i, s', o) u[i][s'], then mu_combine adds the segments of each alpha in
// order. The segments are fixed by the lists, so the sums are reproducible.
constexpr unsigned kSeg = 32;

__global__ void mu_partial(unsigned n, unsigned z, float gamma, const int *__restrict__ act,
        unsigned segBase, const unsigned *__restrict__ segQ, const unsigned *__restrict__ jlist,
        const float *__restrict__ OTr, const float *__restrict__ u, float *__restrict__ partial)
{
    const unsigned seg = segBase + blockIdx.y;
    const unsigned sp = blockIdx.x * kB + threadIdx.x;
    if (sp >= n)
        return;
    float acc = 0.0f;
    for (unsigned q = segQ[2 * seg]; q < segQ[2 * seg + 1]; q++) {
        const unsigned io = jlist[q];
        const unsigned i = io / z, o = io % z;
        const int a = act[i];
        acc += gamma * OTr[((size_t)a * z + o) * n + sp] * u[(size_t)i * n + sp];
    }
    partial[(size_t)blockIdx.y * n + sp] = acc;
}

// mu_t[j][s'] = the segments' sums of j in order (+ mu_{t+1}[j][s'] when no action was chosen
// for j: its alpha was kept)
__global__ void mu_combine(unsigned n, const int *__restrict__ act,
        const unsigned *__restrict__ sstart, unsigned segBase, const float *__restrict__ partial,
        const float *__restrict__ muNext, float *__restrict__ mu)
{
    const unsigned j = blockIdx.y;
    const unsigned sp = blockIdx.x * kB + threadIdx.x;
    if (sp >= n)
        return;
    float acc = (act[j] < 0) ? muNext[(size_t)j * n + sp] : 0.0f;
    for (unsigned seg = sstart[j]; seg < sstart[j + 1]; seg++)
        acc += partial[(size_t)(seg - segBase) * n + sp];
    mu[(size_t)j * n + sp] = acc;
}

__global__ void zero_floats(size_t count, float *__restrict__ x)
{
    const size_t q = blockIdx.x * (size_t)kB + threadIdx.x;
    if (q < count)
        x[q] = 0.0f;
}

} // namespace

} // namespace nova_amd

namespace nova {

using namespace nova_amd;

int pomdp_pbvi_belief_values_gpu(const POMDP *pomdp, const float *d_Gamma, unsigned int rGamma,
        float *d_values, int *d_arg)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "pomdp_pbvi_belief_values_gpu";
    if (pomdp == nullptr || pomdp->d_Z == nullptr || pomdp->d_B == nullptr || d_Gamma == nullptr ||
            rGamma == 0 || d_values == nullptr || d_arg == nullptr) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    hipStream_t st = stream();
    hipLaunchKernelGGL(belief_values, dim3(pomdp->r), dim3(kB), 0, st, pomdp->n, pomdp->rz,
                       rGamma, pomdp->d_Z, pomdp->d_B, d_Gamma, d_values, d_arg);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    return hip_fail(e, fn, "Failed to evaluate the beliefs.", NOVA_ERROR_KERNEL_EXECUTION);
}

int pomdp_pbvi_replay_grad_gpu(const POMDP *pomdp, unsigned int iterations, const int *d_recJ,
        const int *d_recA, const int *d_finalArg, float *d_dR)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "pomdp_pbvi_replay_grad_gpu";
    if (pomdp == nullptr || pomdp->S == nullptr || pomdp->T == nullptr ||
            pomdp->d_O == nullptr || pomdp->d_Z == nullptr || pomdp->d_B == nullptr ||
            d_recJ == nullptr || d_recA == nullptr || d_finalArg == nullptr || d_dR == nullptr ||
            iterations == 0) {
        NOVA_ERR(fn, "Invalid arguments (host S/T and device O/Z/B are required).");
        return NOVA_ERROR_INVALID_DATA;
    }
    const unsigned n = pomdp->n, m = pomdp->m, z = pomdp->z, r = pomdp->r, ns = pomdp->ns;
    const unsigned H = iterations;
    // ---- host: records, inverse lists, predecessor lists
    std::vector<int> recJ((size_t)H * r * z), recA((size_t)H * r), fin(r);
    hipError_t e = hipMemcpy(recJ.data(), d_recJ, sizeof(int) * recJ.size(), hipMemcpyDeviceToHost);
    if (e == hipSuccess)
        e = hipMemcpy(recA.data(), d_recA, sizeof(int) * recA.size(), hipMemcpyDeviceToHost);
    if (e == hipSuccess)
        e = hipMemcpy(fin.data(), d_finalArg, sizeof(int) * r, hipMemcpyDeviceToHost);
    if (e != hipSuccess)
        return hip_fail(e, fn, "Failed to read the records.", NOVA_ERROR_MEMCPY_TO_HOST);
    for (int j : fin)
        if (j < 0 || (unsigned)j >= r) {
            NOVA_ERR(fn, "A final argmax index is outside [0, r).");
            return NOVA_ERROR_INVALID_DATA;
        }
    for (size_t q = 0; q < recJ.size(); q++)
        if (recJ[q] >= (int)r || (recA[q / z] >= 0 && recJ[q] < 0) || recA[q / z] >= (int)m) {
            NOVA_ERR(fn, "A recorded choice is out of range.");
            return NOVA_ERROR_INVALID_DATA;
        }
    // final maximiser lists: beliefs per alpha
    std::vector<unsigned> fstart(r + 1, 0), flist(r);
    for (unsigned i = 0; i < r; i++)
        fstart[fin[i] + 1]++;
    for (unsigned j = 0; j < r; j++)
        fstart[j + 1] += fstart[j];
    {
        std::vector<unsigned> pos(fstart.begin(), fstart.end() - 1);
        for (unsigned i = 0; i < r; i++)
            flist[pos[fin[i]]++] = i;
    }
    // per iteration: (i, o) pairs per target alpha, alphas per action
    std::vector<unsigned> jstart((size_t)H * (r + 1), 0), astart((size_t)H * (m + 1), 0);
    std::vector<unsigned> jlist, alist;
    jlist.reserve((size_t)H * r * z);
    alist.reserve((size_t)H * r);
    for (unsigned t = 0; t < H; t++) {
        unsigned *js = &jstart[(size_t)t * (r + 1)], *as = &astart[(size_t)t * (m + 1)];
        const size_t base = jlist.size(), abase = alist.size();
        for (unsigned i = 0; i < r; i++) {
            const int a = recA[(size_t)t * r + i];
            if (a < 0)
                continue;
            as[a + 1]++;
            for (unsigned o = 0; o < z; o++)
                js[recJ[((size_t)t * r + i) * z + o] + 1]++;
        }
        for (unsigned j = 0; j < r; j++)
            js[j + 1] += js[j];
        for (unsigned a = 0; a < m; a++)
            as[a + 1] += as[a];
        jlist.resize(base + js[r]);
        alist.resize(abase + as[m]);
        std::vector<unsigned> jp(js, js + r), ap(as, as + m);
        for (unsigned i = 0; i < r; i++) {
            const int a = recA[(size_t)t * r + i];
            if (a < 0)
                continue;
            alist[abase + ap[a]++] = i;
            for (unsigned o = 0; o < z; o++)
                jlist[base + jp[recJ[((size_t)t * r + i) * z + o]]++] = i * z + o;
        }
        for (unsigned j = 0; j <= r; j++)
            js[j] += (unsigned)base;
        for (unsigned a = 0; a <= m; a++)
            as[a] += (unsigned)abase;
    }
    // segments of at most kSeg pairs of each alpha's list: sstart[t (r + 1) + j] is the first
    // segment of (t, j), segQ[2 seg .. 2 seg + 1] its pair range in jlist
    std::vector<unsigned> sstart((size_t)H * (r + 1)), segQ;
    unsigned maxSegs = 1;
    for (unsigned t = 0; t < H; t++) {
        const unsigned *js = &jstart[(size_t)t * (r + 1)];
        const unsigned first = (unsigned)(segQ.size() / 2);
        for (unsigned j = 0; j < r; j++) {
            sstart[(size_t)t * (r + 1) + j] = (unsigned)(segQ.size() / 2);
            for (unsigned q = js[j]; q < js[j + 1]; q += kSeg) {
                segQ.push_back(q);
                segQ.push_back(std::min(q + kSeg, js[j + 1]));
            }
        }
        const unsigned end = (unsigned)(segQ.size() / 2);
        sstart[(size_t)t * (r + 1) + r] = end;
        maxSegs = std::max(maxSegs, end - first);
    }
    if (segQ.empty())
        segQ.assign(2, 0u);
    // predecessor lists per action
    std::vector<unsigned> pstart((size_t)m * (n + 1), 0);
    for (unsigned s = 0; s < n; s++)
        for (unsigned a = 0; a < m; a++)
            for (unsigned l = 0; l < ns; l++) {
                const int sp = pomdp->S[idx3(s, a, l, m, ns)];
                if (sp < 0)
                    break;
                if ((unsigned)sp >= n) {
                    NOVA_ERR(fn, "A successor index is outside [0, n).");
                    return NOVA_ERROR_INVALID_DATA;
                }
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
        for (unsigned s = 0; s < n; s++)
            for (unsigned a = 0; a < m; a++)
                for (unsigned l = 0; l < ns; l++) {
                    const size_t x = idx3(s, a, l, m, ns);
                    const int sp = pomdp->S[x];
                    if (sp < 0)
                        break;
                    const unsigned q = fillPos[(size_t)a * (n + 1) + sp]++;
                    pS[q] = (int)s;
                    pT[q] = pomdp->T[x];
                }
    }
    // ---- device buffers
    const size_t rn = (size_t)r * n;
    // [mu, mu', u | fstart, flist, sstart, jlist, astart, alist, pstart, pS, pT, segQ (uploaded)
    //  | partial sums of the segments, O transposed]
    constexpr int kBufs = 15;
    void *d[kBufs] = {};
    const size_t bytes[kBufs] = {sizeof(float) * rn, sizeof(float) * rn, sizeof(float) * rn,
                              sizeof(unsigned) * fstart.size(), sizeof(unsigned) * flist.size(),
                              sizeof(unsigned) * sstart.size(), sizeof(unsigned) * std::max<size_t>(jlist.size(), 1),
                              sizeof(unsigned) * astart.size(), sizeof(unsigned) * std::max<size_t>(alist.size(), 1),
                              sizeof(unsigned) * pstart.size(), sizeof(int) * pS.size(),
                              sizeof(float) * pT.size(), sizeof(unsigned) * segQ.size(),
                              sizeof(float) * maxSegs * (size_t)n, sizeof(float) * (size_t)m * z * n};
    int rc = NOVA_SUCCESS;
    for (int x = 0; x < kBufs && rc == NOVA_SUCCESS; x++)
        rc = device_alloc(&d[x], bytes[x], fn, "replay buffers");
    const void *src[kBufs] = {nullptr, nullptr, nullptr, fstart.data(), flist.data(),
                              sstart.data(), jlist.data(), astart.data(), alist.data(),
                              pstart.data(), pS.data(), pT.data(), segQ.data(), nullptr, nullptr};
    for (int x = 3; x < 13 && rc == NOVA_SUCCESS; x++) {
        const size_t b = (x == 6) ? sizeof(unsigned) * jlist.size()
                         : (x == 8) ? sizeof(unsigned) * alist.size() : bytes[x];
        if (b > 0 && hipMemcpy(d[x], src[x], b, hipMemcpyHostToDevice) != hipSuccess)
            rc = NOVA_ERROR_MEMCPY_TO_DEVICE;
    }
    if (rc == NOVA_SUCCESS) {
        hipStream_t st = stream();
        float *muA = static_cast<float *>(d[0]), *muB = static_cast<float *>(d[1]);
        float *u = static_cast<float *>(d[2]);
        const unsigned *dfs = static_cast<unsigned *>(d[3]), *dfl = static_cast<unsigned *>(d[4]);
        const unsigned *dss = static_cast<unsigned *>(d[5]), *djl = static_cast<unsigned *>(d[6]);
        const unsigned *das = static_cast<unsigned *>(d[7]), *dal = static_cast<unsigned *>(d[8]);
        const unsigned *dps = static_cast<unsigned *>(d[9]);
        const int *dpS = static_cast<int *>(d[10]);
        const float *dpT = static_cast<float *>(d[11]);
        const unsigned *dsq = static_cast<unsigned *>(d[12]);
        float *partial = static_cast<float *>(d[13]), *OTr = static_cast<float *>(d[14]);
        const size_t mzn = (size_t)m * z * n;
        hipLaunchKernelGGL(o_transpose, dim3((unsigned)((mzn + kB - 1) / kB)), dim3(kB), 0, st, n,
                           m, z, pomdp->d_O, OTr);
        const size_t nm = (size_t)n * m;
        hipLaunchKernelGGL(zero_floats, dim3((unsigned)((nm + kB - 1) / kB)), dim3(kB), 0, st, nm, d_dR);
        hipLaunchKernelGGL(mu_init, dim3(r), dim3(kB), 0, st, n, pomdp->rz, pomdp->d_Z,
                           pomdp->d_B, dfs, dfl, muA);
        const dim3 g((n + kB - 1) / kB, r), ga((n + kB - 1) / kB, m);
        for (int t = (int)H - 1; t >= 0; t--) {
            const int *act = d_recA + (size_t)t * r;
            hipLaunchKernelGGL(dr_accumulate, ga, dim3(kB), 0, st, n, m,
                               das + (size_t)t * (m + 1), dal, muA, d_dR);
            hipLaunchKernelGGL(pred_gather, g, dim3(kB), 0, st, n, r, act, dps, dpS, dpT, muA, u);
            const unsigned *ssT = &sstart[(size_t)t * (r + 1)];
            const unsigned segBase = ssT[0], segs = ssT[r] - ssT[0];
            // in slices of at most 65,535 segments (the grid's y extent)
            for (unsigned y0 = 0; y0 < segs; y0 += 65535u)
                hipLaunchKernelGGL(mu_partial, dim3((n + kB - 1) / kB, std::min(65535u, segs - y0)),
                                   dim3(kB), 0, st, n, z, pomdp->gamma, act, segBase + y0, dsq,
                                   djl, OTr, u, partial + (size_t)y0 * n);
            hipLaunchKernelGGL(mu_combine, g, dim3(kB), 0, st, n, act,
                               dss + (size_t)t * (r + 1), segBase, partial, muA, muB);
            std::swap(muA, muB);
        }
        e = hipStreamSynchronize(st);
        if (e != hipSuccess)
            rc = hip_fail(e, fn, "Failed to run the replay.", NOVA_ERROR_KERNEL_EXECUTION);
    } else if (rc == NOVA_ERROR_MEMCPY_TO_DEVICE) {
        NOVA_ERR(fn, "Failed to copy the replay lists to the device.");
    }
    for (void *p : d)
        if (p)
            device_free(p, fn, "replay buffers");
    return rc;
}

} // namespace nova
