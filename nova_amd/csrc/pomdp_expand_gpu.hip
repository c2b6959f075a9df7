// nova-mi355x: random belief expansion with the trajectories stepped on the GPU.
//
// Follows src/pomdp/utilities/pomdp_expand_gpu.cu:137-220 (host code in the reference): row 0
// of the result is belief 0 of the set; then, while rows remain, a trajectory of random length
// h in [0, horizon] starts again from belief 0 and appends every belief it visits, each step
// taking a random action and an observation drawn by cumulative Pr(o | b, a). The random draws
// (h per trajectory, then action and target per step) are taken on the host with libc rand()
// in the reference's order, so a seeded call consumes the same stream. Deviation: like the
// host variant (pomdp_expand_cpu.cpp:127-137) an observation is only chosen when its own
// probability is positive; the reference's device-side variant omits that guard and can then
// append an all-zero belief.
//
// Device work: one workgroup per trajectory. Each step forms u(s') = sum over predecessors
// (s, l) of T[s,a,l] * b(s) by a gather over a predecessor list (built once on the host, fixed
// order, so results are reproducible), Pr(o) = sum_s' O[a,s',o] u(s') by block reductions,
// picks o, and writes b' = O[a,.,o] u / Pr(o) (left unnormalised when |Pr(o)| < 1e-9, as
// nova's belief update does).
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <vector>

#include <nova/amd/nova_amd.h>
#include <nova/pomdp/utilities/pomdp_expand_gpu.h>

#include "nova_hip.h"

namespace nova_amd {

namespace {

constexpr int kEB = 256;
constexpr int kZMax = 64;                 // observations reduced per pass
constexpr int kLdsFloats = 36 * 1024;     // dense belief + prediction staged in LDS

__device__ float block_sum(float v, float *red)
{
    for (int off = 32; off >= 1; off >>= 1)
        v += __shfl_xor(v, off);
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    __syncthreads();
    if (lane == 0)
        red[wave] = v;
    __syncthreads();
    float s = 0.0f;
    for (int w = 0; w < kEB / 64; w++)
        s += red[w];
    return s;
}

__global__ __launch_bounds__(kEB) void expand_trajectories(unsigned n, unsigned m, unsigned z,
        int useLds, const unsigned *__restrict__ predStart, const int *__restrict__ predS,
        const float *__restrict__ predT, const float *__restrict__ O,
        const float *__restrict__ b0, const unsigned *__restrict__ trajLen,
        const unsigned *__restrict__ trajOff, const unsigned *__restrict__ stepAct,
        const float *__restrict__ stepTarget, float *__restrict__ scratch,
        float *__restrict__ Bnew, unsigned *__restrict__ nnzOut)
{
    extern __shared__ float lds[];
    __shared__ float red[kEB / 64];
    __shared__ float pr[kZMax];
    __shared__ int chosen;
    __shared__ float norm;
    const unsigned k = blockIdx.x, t = threadIdx.x;
    float *b = useLds ? lds : scratch + (size_t)k * 2 * n;
    float *u = b + n;
    for (unsigned s = t; s < n; s += kEB)
        b[s] = b0[s];
    __syncthreads();
    const unsigned h = trajLen[k], off = trajOff[k];
    for (unsigned step = 0; step < h; step++) {
        const unsigned a = stepAct[off + step];
        const float target = stepTarget[off + step];
        const unsigned *ps = predStart + (size_t)a * (n + 1);
        for (unsigned sp = t; sp < n; sp += kEB) {
            float acc = 0.0f;
            for (unsigned q = ps[sp]; q < ps[sp + 1]; q++)
                acc += predT[q] * b[predS[q]];
            u[sp] = acc;
        }
        __syncthreads();
        // Pr(o) for every o, kZMax observations per pass
        for (unsigned o0 = 0; o0 < z; o0 += kZMax) {
            const unsigned oc = min((unsigned)kZMax, z - o0);
            for (unsigned oo = 0; oo < oc; oo++) {
                float part = 0.0f;
                for (unsigned sp = t; sp < n; sp += kEB)
                    part += O[((size_t)a * n + sp) * z + o0 + oo] * u[sp];
                const float tot = block_sum(part, red);
                if (t == 0)
                    pr[oo] = tot;
            }
            __syncthreads();
            if (t == 0) {
                if (o0 == 0) {
                    chosen = -1;
                    norm = 0.0f;   // running cumulative sum between passes
                }
                if (chosen < 0) {
                    float cum = norm;
                    for (unsigned oo = 0; oo < oc; oo++) {
                        cum += pr[oo];
                        if (cum >= target && pr[oo] > 0.0f) {
                            chosen = (int)(o0 + oo);
                            break;
                        }
                    }
                    norm = cum;
                }
            }
            __syncthreads();
        }
        if (t == 0) {
            if (chosen < 0)
                chosen = 0;
        }
        __syncthreads();
        const unsigned o = (unsigned)chosen;
        // normaliser of the chosen observation
        float part = 0.0f;
        for (unsigned sp = t; sp < n; sp += kEB)
            part += O[((size_t)a * n + sp) * z + o] * u[sp];
        const float nc = block_sum(part, red);
        const bool ok = fabsf(nc) >= (float)NOVA_FLT_ERR_TOL;
        unsigned cnt = 0;
        float *out = Bnew + (size_t)(off + step) * n;
        for (unsigned sp = t; sp < n; sp += kEB) {
            float v = O[((size_t)a * n + sp) * z + o] * u[sp];
            if (ok)
                v /= nc;
            b[sp] = v;
            out[sp] = v;
            cnt += (v > 0.0f) ? 1u : 0u;
        }
        const float c = block_sum((float)cnt, red);
        if (t == 0)
            nnzOut[off + step] = (unsigned)c;
        __syncthreads();
    }
}

} // namespace

// Host side of the random expansion; rows 1.. of Bnew are filled on the device.
static int expand_random_device(const nova::POMDP *p, unsigned numDesired, unsigned &maxNnz,
                                float *&Bnew, const char *fn)
{
    const unsigned n = p->n, m = p->m, z = p->z, ns = p->ns;
    maxNnz = 0;
    // draws, in the reference's order
    std::vector<unsigned> len, off, act;
    std::vector<float> target;
    unsigned filled = 1;
    while (filled < numDesired) {
        unsigned h = (unsigned)((float)std::rand() / (float)RAND_MAX * (float)(p->horizon + 1));
        unsigned use = 0;
        const unsigned start = filled - 1;
        for (unsigned t = 0; t < h && filled < numDesired; t++) {
            unsigned a = (unsigned)((float)std::rand() / (float)RAND_MAX * (float)m);
            if (a >= m) a = m - 1;
            act.push_back(a);
            target.push_back((float)std::rand() / (float)RAND_MAX);
            use++;
            filled++;
        }
        if (use > 0) {
            len.push_back(use);
            off.push_back(start);
        }
    }
    const unsigned steps = numDesired - 1, trajs = (unsigned)len.size();

    // predecessor lists per action: (s, T[s,a,l]) for every valid successor entry, by s'
    std::vector<unsigned> pstart((size_t)m * (n + 1), 0);
    for (unsigned s = 0; s < n; s++)
        for (unsigned a = 0; a < m; a++)
            for (unsigned l = 0; l < ns; l++) {
                const int sp = p->S[idx3(s, a, l, m, ns)];
                if (sp < 0)
                    break;
                pstart[(size_t)a * (n + 1) + sp + 1]++;
            }
    for (unsigned a = 0; a < m; a++)
        for (unsigned sp = 0; sp < n; sp++)
            pstart[(size_t)a * (n + 1) + sp + 1] += pstart[(size_t)a * (n + 1) + sp];
    size_t total = 0;
    for (unsigned a = 0; a < m; a++) {
        unsigned *row = &pstart[(size_t)a * (n + 1)];
        for (unsigned sp = 0; sp <= n; sp++)
            row[sp] += (unsigned)total;
        total = row[n];
    }
    std::vector<int> pS(total > 0 ? total : 1);
    std::vector<float> pT(total > 0 ? total : 1);
    {
        std::vector<unsigned> fillPos(pstart);
        for (unsigned s = 0; s < n; s++)
            for (unsigned a = 0; a < m; a++)
                for (unsigned l = 0; l < ns; l++) {
                    const size_t x = idx3(s, a, l, m, ns);
                    const int sp = p->S[x];
                    if (sp < 0)
                        break;
                    const unsigned q = fillPos[(size_t)a * (n + 1) + sp]++;
                    pS[q] = (int)s;
                    pT[q] = p->T[x];
                }
    }
    std::vector<float> b0(n, 0.0f);
    for (unsigned k = 0; k < p->rz; k++) {
        const int s = p->Z[k];
        if (s < 0)
            break;
        b0[s] = p->B[k];
    }

    Bnew = new float[(size_t)numDesired * n];
    std::memcpy(Bnew, b0.data(), sizeof(float) * n);
    if (steps == 0 || trajs == 0)
        return NOVA_SUCCESS;

    const bool useLds = 2 * n <= (unsigned)kLdsFloats;
    void *d[11] = {};
    const size_t bytes[11] = {
        sizeof(unsigned) * pstart.size(), sizeof(int) * pS.size(), sizeof(float) * pT.size(),
        p->d_O ? 0 : sizeof(float) * (size_t)m * n * z, sizeof(float) * n,
        sizeof(unsigned) * trajs, sizeof(unsigned) * trajs, sizeof(unsigned) * steps,
        sizeof(float) * steps, useLds ? 4 : sizeof(float) * 2 * (size_t)n * trajs,
        sizeof(float) * (size_t)steps * n + sizeof(unsigned) * steps};
    int rc = NOVA_SUCCESS;
    for (int x = 0; x < 11 && rc == NOVA_SUCCESS; x++)
        if (bytes[x] > 0)
            rc = device_alloc(&d[x], bytes[x], fn, "expansion buffers");
    hipError_t e = hipSuccess;
    const float *dO = p->d_O ? p->d_O : static_cast<const float *>(d[3]);
    if (rc == NOVA_SUCCESS) {
        const void *src[9] = {pstart.data(), pS.data(), pT.data(), p->O, b0.data(), len.data(),
                              off.data(), act.data(), target.data()};
        for (int x = 0; x < 9 && e == hipSuccess; x++)
            if (bytes[x] > 0)
                e = hipMemcpy(d[x], src[x], bytes[x], hipMemcpyHostToDevice);
        if (e != hipSuccess)
            rc = hip_fail(e, fn, "Failed to copy the expansion inputs to the device.",
                          NOVA_ERROR_MEMCPY_TO_DEVICE);
    }
    if (rc == NOVA_SUCCESS) {
        float *dB = static_cast<float *>(d[10]);
        unsigned *dN = reinterpret_cast<unsigned *>(dB + (size_t)steps * n);
        const size_t lds = useLds ? sizeof(float) * 2 * n : 0;
        if (useLds)
            hipFuncSetAttribute((const void *)expand_trajectories,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                kLdsFloats * (int)sizeof(float));
        hipStream_t st = stream();
        hipLaunchKernelGGL(expand_trajectories, dim3(trajs), dim3(kEB), lds, st, n, m, z,
                           useLds ? 1 : 0, static_cast<unsigned *>(d[0]),
                           static_cast<int *>(d[1]), static_cast<float *>(d[2]), dO,
                           static_cast<float *>(d[4]), static_cast<unsigned *>(d[5]),
                           static_cast<unsigned *>(d[6]), static_cast<unsigned *>(d[7]),
                           static_cast<float *>(d[8]), static_cast<float *>(d[9]), dB, dN);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipStreamSynchronize(st);
        if (e == hipSuccess)
            e = hipMemcpy(Bnew + n, dB, sizeof(float) * (size_t)steps * n, hipMemcpyDeviceToHost);
        std::vector<unsigned> nnz(steps);
        if (e == hipSuccess)
            e = hipMemcpy(nnz.data(), dN, sizeof(unsigned) * steps, hipMemcpyDeviceToHost);
        if (e != hipSuccess)
            rc = hip_fail(e, fn, "Failed to run the trajectories.", NOVA_ERROR_KERNEL_EXECUTION);
        for (unsigned x = 0; x < steps && rc == NOVA_SUCCESS; x++)
            if (nnz[x] > maxNnz)
                maxNnz = nnz[x];
    }
    for (int x = 0; x < 11; x++)
        if (d[x])
            device_free(d[x], fn, "expansion buffers");
    if (rc != NOVA_SUCCESS) {
        delete[] Bnew;
        Bnew = nullptr;
    }
    return rc;
}

} // namespace nova_amd

namespace nova {

static bool expand_args_ok(const POMDP *p, unsigned num, float *Bnew)
{
    return p != nullptr && p->n != 0 && p->ns != 0 && p->m != 0 && p->z != 0 && p->r != 0 &&
           p->rz != 0 && p->S != nullptr && p->T != nullptr && p->O != nullptr &&
           p->R != nullptr && p->Z != nullptr && p->B != nullptr && p->horizon >= 1 &&
           num > 1 && Bnew == nullptr;
}

int pomdp_expand_random_gpu(const POMDP *pomdp, unsigned int numThreads,
        unsigned int numDesiredBeliefPoints, unsigned int &maxNonZeroValues, float *&Bnew)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "pomdp_expand_random_gpu";
    (void)numThreads;
    if (!expand_args_ok(pomdp, numDesiredBeliefPoints, Bnew)) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    if (nova_amd::first_bad_successor(pomdp->S, (size_t)pomdp->n * pomdp->m, pomdp->ns,
                                      pomdp->n) >= 0) {
        NOVA_ERR(fn, "A successor index is outside [0, n).");
        return NOVA_ERROR_INVALID_DATA;
    }
    std::srand((unsigned)std::time(nullptr));
    return nova_amd::expand_random_device(pomdp, numDesiredBeliefPoints, maxNonZeroValues, Bnew,
                                          fn);
}

int pomdp_expand_random_seeded_gpu(const POMDP *pomdp, unsigned int seed,
        unsigned int numDesiredBeliefPoints, unsigned int *maxNonZeroValues, float **Bnew)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "pomdp_expand_random_seeded_gpu";
    if (maxNonZeroValues == nullptr || Bnew == nullptr ||
            !expand_args_ok(pomdp, numDesiredBeliefPoints, *Bnew)) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    if (nova_amd::first_bad_successor(pomdp->S, (size_t)pomdp->n * pomdp->m, pomdp->ns,
                                      pomdp->n) >= 0) {
        NOVA_ERR(fn, "A successor index is outside [0, n).");
        return NOVA_ERROR_INVALID_DATA;
    }
    std::srand(seed);
    return nova_amd::expand_random_device(pomdp, numDesiredBeliefPoints, *maxNonZeroValues,
                                          *Bnew, fn);
}

int nova_amd_free_floats(float *p)
{
    delete[] p;
    return NOVA_SUCCESS;
}

} // namespace nova
