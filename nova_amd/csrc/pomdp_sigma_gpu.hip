// nova-mi355x: the sigma-approximation of a device-resident belief set on one MI355X.
//
// nova's pomdp_sigma_cpu (src/pomdp/utilities/pomdp_sigma_cpu.cpp:50-118) keeps each belief's k
// largest entries, renormalised by their sum sigma_b, and reports sigma = min(1, min_b sigma_b).
// Here one 256-lane workgroup takes one belief: its (value, state, position) triples are sorted
// in LDS by a bitonic network, value descending, and entries of equal value by their position
// in the row (nova's std::sort leaves the order of equal values to the library's algorithm, so
// where a row holds equal values at the cut the kept states can differ from nova's; with
// distinct values both keep the same entries in the same order). Lane 0 then sums the kept
// values in that order (sigma_b, as nova's loop) and every kept entry is divided by it with
// IEEE division, so B' and Z' equal nova's bit for bit for rows of distinct values. Entries past
// the row's rz count as value 0, as in the host restatement (nova reads past its arrays there).
#include <algorithm>
#include <vector>

#include <nova/amd/nova_amd.h>
#include <nova/pomdp/pomdp.h>

#include "nova_hip.h"

namespace nova_amd {

namespace {

constexpr int kSigThreads = 256;
constexpr unsigned kSigMaxRz = 8192;

__global__ __launch_bounds__(kSigThreads) void sigma_rows(unsigned rz, unsigned cap, unsigned k,
        const int *__restrict__ Z, const float *__restrict__ B, int *__restrict__ Znew,
        float *__restrict__ Bnew, float *__restrict__ sigmaRow)
{
    extern __shared__ unsigned char sigLds[];
    float *val = reinterpret_cast<float *>(sigLds);
    int *st = reinterpret_cast<int *>(val + cap);
    int *pos = st + cap;
    __shared__ float sigmab;
    const unsigned i = blockIdx.x, t = threadIdx.x;
    for (unsigned x = t; x < cap; x += kSigThreads) {
        const bool in = x < rz;
        val[x] = in ? B[(size_t)i * rz + x] : -__builtin_inff();
        st[x] = in ? Z[(size_t)i * rz + x] : -1;
        pos[x] = (int)x;
    }
    __syncthreads();
    // bitonic sort: "x before y" when val[x] > val[y], or equal values and pos[x] < pos[y]
    for (unsigned size = 2; size <= cap; size <<= 1) {
        for (unsigned stride = size >> 1; stride > 0; stride >>= 1) {
            for (unsigned x = t; x < cap; x += kSigThreads) {
                const unsigned y = x ^ stride;
                if (y > x) {
                    const bool up = (x & size) == 0;
                    const bool xFirst = val[x] > val[y] || (val[x] == val[y] && pos[x] < pos[y]);
                    if (xFirst != up) {
                        const float v = val[x];
                        val[x] = val[y];
                        val[y] = v;
                        const int s = st[x];
                        st[x] = st[y];
                        st[y] = s;
                        const int p = pos[x];
                        pos[x] = pos[y];
                        pos[y] = p;
                    }
                }
            }
            __syncthreads();
        }
    }
    if (t == 0) {
        float s = 0.0f;
        const unsigned kk = min(k, rz);
        for (unsigned x = 0; x < kk; x++)
            s = __fadd_rn(s, val[x]);
        sigmab = s;
        sigmaRow[i] = s;
    }
    __syncthreads();
    const float sb = sigmab;
    for (unsigned x = t; x < k; x += kSigThreads) {
        const bool keep = x < rz && val[x] > 0.0f;
        Bnew[(size_t)i * k + x] = keep ? val[x] / sb : 0.0f;
        Znew[(size_t)i * k + x] = keep ? st[x] : -1;
    }
}

} // namespace

} // namespace nova_amd

using namespace nova_amd;

namespace nova {

int pomdp_sigma_gpu(POMDP *pomdp, unsigned int numDesiredNonZeroValues, float *sigma)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "pomdp_sigma_gpu";
    if (pomdp == nullptr || pomdp->r == 0 || pomdp->rz == 0 || pomdp->d_Z == nullptr ||
            pomdp->d_B == nullptr || numDesiredNonZeroValues == 0 || sigma == nullptr ||
            pomdp->rz > kSigMaxRz) {
        NOVA_ERR(fn, "Invalid arguments (a POMDP with its beliefs on the device, rz <= 8,192, "
                     "k > 0).");
        return NOVA_ERROR_INVALID_DATA;
    }
    const unsigned r = pomdp->r, rz = pomdp->rz, k = numDesiredNonZeroValues;
    unsigned cap = 1;
    while (cap < rz)
        cap <<= 1;
    void *zNew = nullptr, *bNew = nullptr, *sRow = nullptr;
    int rc = device_alloc(&zNew, sizeof(int) * (size_t)r * k, fn, "the nonzero beliefs");
    if (rc == NOVA_SUCCESS)
        rc = device_alloc(&bNew, sizeof(float) * (size_t)r * k, fn, "the belief points");
    if (rc == NOVA_SUCCESS)
        rc = device_alloc(&sRow, sizeof(float) * (size_t)r, fn, "the kept masses");
    std::vector<float> rows(r);
    std::vector<int> hz;
    std::vector<float> hb;
    hipError_t e = hipSuccess;
    if (rc == NOVA_SUCCESS) {
        hipStream_t st = stream();
        const size_t lds = (size_t)cap * (sizeof(float) + 2 * sizeof(int));
        e = hipFuncSetAttribute((const void *)sigma_rows,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(sigma_rows, dim3(r), dim3(kSigThreads), lds, st, rz, cap, k,
                               pomdp->d_Z, pomdp->d_B, static_cast<int *>(zNew),
                               static_cast<float *>(bNew), static_cast<float *>(sRow));
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(rows.data(), sRow, sizeof(float) * r, hipMemcpyDeviceToHost, st);
        hz.resize((size_t)r * k);
        hb.resize((size_t)r * k);
        if (e == hipSuccess && pomdp->Z != nullptr)
            e = hipMemcpyAsync(hz.data(), zNew, sizeof(int) * hz.size(), hipMemcpyDeviceToHost,
                               st);
        if (e == hipSuccess && pomdp->B != nullptr)
            e = hipMemcpyAsync(hb.data(), bNew, sizeof(float) * hb.size(),
                               hipMemcpyDeviceToHost, st);
        if (e == hipSuccess)
            e = hipStreamSynchronize(st);
        rc = hip_fail(e, fn, "Kernel execution failed.", NOVA_ERROR_KERNEL_EXECUTION);
    }
    device_free(sRow, fn, "the kept masses");
    if (rc != NOVA_SUCCESS) {
        device_free(zNew, fn, "the nonzero beliefs");
        device_free(bNew, fn, "the belief points");
        return rc;
    }
    float s = 1.0f;
    for (unsigned i = 0; i < r; i++)
        if (rows[i] < s)
            s = rows[i];
    *sigma = s;
    // the new set replaces the device copy (released as pomdp_uninitialize_*_gpu would) and,
    // when the POMDP keeps host arrays, the host copy (as pomdp_sigma_cpu replaces them)
    model_changed(false);
    device_free_if_owned(pomdp->d_Z, fn, "the nonzero beliefs");
    device_free_if_owned(pomdp->d_B, fn, "the belief points");
    pomdp->d_Z = static_cast<int *>(zNew);
    pomdp->d_B = static_cast<float *>(bNew);
    if (pomdp->Z != nullptr && pomdp->B != nullptr) {
        int *Zh = tracked_new<int>((size_t)r * k);
        float *Bh = tracked_new<float>((size_t)r * k);
        std::copy(hz.begin(), hz.end(), Zh);
        std::copy(hb.begin(), hb.end(), Bh);
        release(pomdp->Z);
        release(pomdp->B);
        pomdp->Z = Zh;
        pomdp->B = Bh;
    }
    pomdp->rz = k;
    return NOVA_SUCCESS;
}

} // namespace nova
