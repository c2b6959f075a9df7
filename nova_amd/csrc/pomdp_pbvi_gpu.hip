// nova-mi355x: point-based value iteration (PBVI) backups on one MI355X (gfx950, wave64).
//
// The backup of belief i (src/pomdp/algorithms/pomdp_pbvi_cpu.cpp:37-162):
//   score_{a,o}(j) = sum_k B[i,k] * gamma * sum_l (O[a,sp,o]*T[s_k,a,l]) * Gamma[j,sp],
//                    s_k = Z[i,k], sp = S[s_k,a,l]      (valid prefixes of Z and S only)
//   j*_{a,o}       = first argmax_j score_{a,o}(j)       (strict '>' from -1e10)
//   alpha_a(s)     = R[s,a] + sum_o gamma * sum_l (O*T) * Gamma[j*_{a,o}, sp]
//   a*             = first argmax_a sum_k alpha_a(Z[i,k]) * B[i,k];  Gamma'[i] = alpha_{a*}
//
// Per iteration (all launched on the library stream, captured as one HIP graph by execute):
//   1. gamma_transpose : Gamma (r x n) -> GT (n x r), so that the r alphas of one successor
//                        state are contiguous and a wave's loads of GT[sp, j..j+63] coalesce.
//   2. pbvi_score      : one workgroup per (belief, action, tile of 512 alphas). The valid
//                        (k, l) pairs of the belief's support are compacted into LDS with the
//                        matching O*T factors of a chunk of 16 observations; each lane owns 2
//                        alphas and accumulates 16 observation scores per alpha from one GT
//                        load per (k, l). Per-tile argmaxes are reduced with wave shuffles.
//   3. pbvi_action     : one workgroup per belief: merges the tiles (first maximiser wins),
//                        evaluates alpha_a on the belief's support in the reference order and
//                        picks the action.
//   4. pbvi_materialize: one workgroup per belief writes Gamma'[i] and pi[i].
// Two arithmetic modes share this structure:
//   exact - every product and sum rounded separately in the reference's order (no FMA), so
//           Gamma' and pi equal nova's CPU results bit for bit;
//   fast  - (default) scores accumulate with FMA from W = gamma*B*O*T; the materialisation
//           stages h(sp) = sum_o O[a*,sp,o]*Gamma[j*_o,sp] in LDS and then forms
//           R + gamma*sum_l T*h(S) per state. Results agree with the reference within float
//           rounding; alpha choices can differ only at near-ties.
// Model constant built once per initialize: OT[a][o][s][l] = O[a,S[s,a,l],o]*T[s,a,l] (0 past
// the end of the successor list). It, GT and the per-iteration argmax records live in one
// device workspace referenced by POMDPPBVIGPU::d_alphaBA.
#include <climits>
#include <cstring>
#include <vector>

#include <nova/amd/nova_amd.h>
#include <nova/pomdp/algorithms/pomdp_pbvi_gpu.h>

#include "nova_hip.h"

namespace nova_amd {

namespace {

constexpr int kPB = 256;            // lanes per workgroup
constexpr int kUMax = 512;          // max valid (k, l) pairs of one belief-action in LDS
constexpr int kKMax = 512;          // max belief support length handled in LDS
constexpr int kZC = 16;             // observations per scoring chunk
constexpr int kJPT = 2;             // alphas per lane
constexpr int kJT = kPB * kJPT;     // alphas per scoring tile
constexpr int kLdsRowFloats = 36 * 1024;  // Gamma row / h vector staged in LDS (144 KB)

__device__ __forceinline__ bool better(float v, int j, float bv, int bj)
{
    // larger value wins; equal values: the smaller alpha index (the reference's first maximiser)
    return v > bv || (v == bv && j < bj);
}

// --------------------------------------------------------------------- model constant

__global__ __launch_bounds__(kPB) void pbvi_build_ot(unsigned n, unsigned ns, unsigned m,
        unsigned z, const int *__restrict__ S, const float *__restrict__ T,
        const float *__restrict__ O, float *__restrict__ OT)
{
    // OT[((a*z + o)*n + s)*ns + l]
    const size_t total = (size_t)m * z * n * ns;
    for (size_t x = blockIdx.x * (size_t)kPB + threadIdx.x; x < total;
         x += (size_t)gridDim.x * kPB) {
        const unsigned l = x % ns;
        size_t y = x / ns;
        const unsigned s = y % n;
        y /= n;
        const unsigned o = y % z;
        const unsigned a = y / z;
        const size_t row = ((size_t)s * m + a) * ns;
        bool valid = true;
        for (unsigned q = 0; q <= l; q++)
            valid = valid && S[row + q] >= 0;
        float v = 0.0f;
        if (valid) {
            const unsigned sp = (unsigned)S[row + l];
            v = __fmul_rn(O[((size_t)a * n + sp) * z + o], T[row + l]);
        }
        OT[x] = v;
    }
}

// ----------------------------------------------------------------------- 1. transpose

__global__ __launch_bounds__(kPB) void gamma_transpose(unsigned r, unsigned n,
        const float *__restrict__ G, float *__restrict__ GT)
{
    __shared__ float tile[64][65];
    const unsigned j0 = blockIdx.y * 64, s0 = blockIdx.x * 64;
    const unsigned tx = threadIdx.x % 64, ty = threadIdx.x / 64;
    for (unsigned q = ty; q < 64; q += 4) {
        const unsigned j = j0 + q, s = s0 + tx;
        tile[q][tx] = (j < r && s < n) ? G[(size_t)j * n + s] : 0.0f;
    }
    __syncthreads();
    for (unsigned q = ty; q < 64; q += 4) {
        const unsigned s = s0 + q, j = j0 + tx;
        if (s < n && j < r)
            GT[(size_t)s * r + j] = tile[tx][q];
    }
}

// --------------------------------------------------------------------------- 2. score

// Compacted support of one belief for one action, in LDS.
struct SupportLds {
    int kEnd;                 // valid support length (first Z < 0)
    int uEnd;                 // number of valid (k, l) pairs
    int sK[kKMax];            // support states
    float bK[kKMax];          // belief values
    int startK[kKMax + 1];    // first pair index of support entry k
    int spU[kUMax];           // successor state of pair u
    int otU[kUMax];           // s_k*ns + l of pair u (offset into an OT plane)
};

// Fills the SupportLds of belief i and action a cooperatively (all kPB lanes).
__device__ void build_support(SupportLds &L, unsigned i, unsigned a, unsigned m, unsigned ns,
        unsigned rz, const int *__restrict__ S, const int *__restrict__ Z,
        const float *__restrict__ B, int *cnt)
{
    const unsigned t = threadIdx.x;
    if (t == 0)
        L.kEnd = (int)rz;
    __syncthreads();
    for (unsigned k = t; k < rz; k += kPB) {
        const int s = Z[(size_t)i * rz + k];
        L.sK[k] = s;
        L.bK[k] = B[(size_t)i * rz + k];
        if (s < 0)
            atomicMin(&L.kEnd, (int)k);
    }
    __syncthreads();
    const int kEnd = L.kEnd;
    for (int k = t; k < kEnd; k += kPB) {
        const size_t row = ((size_t)L.sK[k] * m + a) * ns;
        int c = 0;
        while (c < (int)ns && S[row + c] >= 0)
            c++;
        cnt[k] = c;
    }
    __syncthreads();
    // exclusive prefix sum of cnt[0..kEnd) (Hillis-Steele over startK, ping-pong via cnt)
    for (int k = t; k <= kEnd; k += kPB)
        L.startK[k] = (k == 0) ? 0 : cnt[k - 1];
    __syncthreads();
    for (int off = 1; off <= kEnd; off <<= 1) {
        for (int k = t; k <= kEnd; k += kPB)
            cnt[k] = (k >= off) ? L.startK[k] + L.startK[k - off] : L.startK[k];
        __syncthreads();
        for (int k = t; k <= kEnd; k += kPB)
            L.startK[k] = cnt[k];
        __syncthreads();
    }
    if (t == 0)
        L.uEnd = L.startK[kEnd];
    for (int x = t; x < kEnd * (int)ns; x += kPB) {
        const int k = x / (int)ns, l = x % (int)ns;
        const int u = L.startK[k] + l;
        if (u < L.startK[k + 1]) {
            const size_t row = ((size_t)L.sK[k] * m + a) * ns;
            L.spU[u] = S[row + l];
            L.otU[u] = L.sK[k] * (int)ns + l;
        }
    }
    __syncthreads();
}

template <bool EXACT>
__global__ __launch_bounds__(kPB) void pbvi_score(unsigned n, unsigned ns, unsigned m,
        unsigned z, unsigned r, unsigned rz, float gamma, const int *__restrict__ S,
        const int *__restrict__ Z, const float *__restrict__ B, const float *__restrict__ OT,
        const float *__restrict__ GT, float *__restrict__ partVal, int *__restrict__ partJ)
{
    __shared__ SupportLds L;
    __shared__ float w[kUMax * kZC];         // [u][o] factors of the current chunk
    __shared__ int scratch[kKMax + 1];
    __shared__ float redV[kPB / 64][kZC];
    __shared__ int redJ[kPB / 64][kZC];

    const unsigned ia = blockIdx.x;              // belief * m + action
    const unsigned i = ia / m, a = ia % m;
    const unsigned tile = blockIdx.y;
    const unsigned t = threadIdx.x, lane = t % 64, wave = t / 64;

    build_support(L, i, a, m, ns, rz, S, Z, B, scratch);
    const int kEnd = L.kEnd, uEnd = L.uEnd;

    int jq[kJPT];
    bool live[kJPT];
#pragma unroll
    for (int q = 0; q < kJPT; q++) {
        jq[q] = (int)(tile * kJT + t + q * kPB);
        live[q] = jq[q] < (int)r;
    }

    for (unsigned o0 = 0; o0 < z; o0 += kZC) {
        const int zc = min((int)kZC, (int)(z - o0));
        // stage the chunk's factors: exact -> O*T; fast -> gamma*B_k*O*T
        for (int x = t; x < uEnd * kZC; x += kPB) {
            const int u = x / kZC, oo = x % kZC;
            float v = 0.0f;
            if (oo < zc) {
                v = OT[((size_t)(a * z + o0 + oo) * n) * ns + L.otU[u]];
                if (!EXACT) {
                    // the support entry k of pair u: startK[k] <= u < startK[k+1]
                    int lo = 0, hi = kEnd - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) / 2;
                        if (L.startK[mid] <= u) lo = mid; else hi = mid - 1;
                    }
                    v = gamma * L.bK[lo] * v;
                }
            }
            w[u * kZC + oo] = v;
        }
        __syncthreads();

        float val[kZC][kJPT];
        float vtk[EXACT ? kZC : 1][kJPT];
#pragma unroll
        for (int oo = 0; oo < kZC; oo++)
#pragma unroll
            for (int q = 0; q < kJPT; q++)
                val[oo][q] = 0.0f;

        for (int k = 0; k < kEnd; k++) {
            if (EXACT) {
#pragma unroll
                for (int oo = 0; oo < kZC; oo++)
#pragma unroll
                    for (int q = 0; q < kJPT; q++)
                        vtk[oo][q] = 0.0f;
            }
            const int u1 = L.startK[k + 1];
            for (int u = L.startK[k]; u < u1; u++) {
                const size_t rowOff = (size_t)L.spU[u] * r;
                float g[kJPT];
#pragma unroll
                for (int q = 0; q < kJPT; q++)
                    g[q] = live[q] ? GT[rowOff + jq[q]] : 0.0f;
                const float *wu = w + u * kZC;
#pragma unroll
                for (int oo = 0; oo < kZC; oo++) {
                    const float f = wu[oo];
#pragma unroll
                    for (int q = 0; q < kJPT; q++) {
                        if (EXACT)
                            vtk[oo % (EXACT ? kZC : 1)][q] =
                                __fadd_rn(vtk[oo % (EXACT ? kZC : 1)][q], __fmul_rn(f, g[q]));
                        else
                            val[oo][q] = __builtin_fmaf(f, g[q], val[oo][q]);
                    }
                }
            }
            if (EXACT) {
                const float bk = L.bK[k];
#pragma unroll
                for (int oo = 0; oo < kZC; oo++)
#pragma unroll
                    for (int q = 0; q < kJPT; q++) {
                        const float v = __fmul_rn(vtk[oo % (EXACT ? kZC : 1)][q], gamma);
                        val[oo][q] = __fadd_rn(val[oo][q], __fmul_rn(v, bk));
                    }
            }
        }

        // per lane: first maximiser over its alphas (visited in increasing j)
        float bv[kZC];
        int bj[kZC];
#pragma unroll
        for (int oo = 0; oo < kZC; oo++) {
            bv[oo] = kLow;
            bj[oo] = INT_MAX;
#pragma unroll
            for (int q = 0; q < kJPT; q++)
                if (live[q] && val[oo][q] > bv[oo]) {
                    bv[oo] = val[oo][q];
                    bj[oo] = jq[q];
                }
        }
        // wave reduction, then across the 4 waves
#pragma unroll
        for (int oo = 0; oo < kZC; oo++) {
            for (int off = 32; off >= 1; off >>= 1) {
                const float ov = __shfl_xor(bv[oo], off);
                const int oj = __shfl_xor(bj[oo], off);
                if (better(ov, oj, bv[oo], bj[oo])) {
                    bv[oo] = ov;
                    bj[oo] = oj;
                }
            }
            if (lane == 0) {
                redV[wave][oo] = bv[oo];
                redJ[wave][oo] = bj[oo];
            }
        }
        __syncthreads();
        if (t < (unsigned)zc) {
            float v = redV[0][t];
            int j = redJ[0][t];
            for (int wv = 1; wv < kPB / 64; wv++)
                if (better(redV[wv][t], redJ[wv][t], v, j)) {
                    v = redV[wv][t];
                    j = redJ[wv][t];
                }
            const size_t idx = ((size_t)tile * r * m + ia) * z + o0 + t;
            partVal[idx] = v;
            partJ[idx] = j;
        }
        __syncthreads();
    }
}

// -------------------------------------------------------------------------- 3. action

template <bool EXACT>
__global__ __launch_bounds__(kPB) void pbvi_action(unsigned n, unsigned ns, unsigned m,
        unsigned z, unsigned r, unsigned rz, unsigned tiles, float gamma,
        const int *__restrict__ S, const float *__restrict__ R, const int *__restrict__ Z,
        const float *__restrict__ B, const float *__restrict__ OT,
        const float *__restrict__ G, const float *__restrict__ partVal,
        const int *__restrict__ partJ, int *__restrict__ jstar, float *__restrict__ aval,
        int *__restrict__ act)
{
    __shared__ int js[kZC * 8];
    __shared__ float alphaK[kKMax];
    __shared__ float bS[kKMax];
    __shared__ int kEndS;
    const unsigned i = blockIdx.x, t = threadIdx.x;
    if (t == 0)
        kEndS = (int)rz;
    __syncthreads();
    for (unsigned k = t; k < rz; k += kPB) {
        bS[k] = B[(size_t)i * rz + k];
        if (Z[(size_t)i * rz + k] < 0)
            atomicMin(&kEndS, (int)k);
    }
    __syncthreads();
    const int kEnd = kEndS;
    float bestValue = kLow;
    int bestA = -1;
    for (unsigned a = 0; a < m; a++) {
        const size_t ia = (size_t)i * m + a;
        for (unsigned o0 = 0; o0 < z; o0 += kZC * 8) {
            // merge the alpha tiles of each observation: earlier tile wins ties
            const unsigned oc = min((unsigned)(kZC * 8), z - o0);
            for (unsigned oo = t; oo < oc; oo += kPB) {
                float v = kLow;
                int j = INT_MAX;
                for (unsigned tl = 0; tl < tiles; tl++) {
                    const size_t idx = ((size_t)tl * r * m + ia) * z + o0 + oo;
                    if (partVal[idx] > v) {
                        v = partVal[idx];
                        j = partJ[idx];
                    }
                }
                if (j == INT_MAX)
                    j = 0;   // no alpha above -1e10: the reference's index is undefined
                js[oo] = j;
                jstar[ia * z + o0 + oo] = j;
            }
            __syncthreads();
            // alpha_a on the support, observation-major accumulation as the reference
            for (int k = t; k < kEnd; k += kPB) {
                const int s = Z[(size_t)i * rz + k];
                const size_t row = ((size_t)s * m + a) * ns;
                float al = (o0 == 0) ? R[(size_t)s * m + a] : alphaK[k];
                for (unsigned oo = 0; oo < oc; oo++) {
                    const float *Gj = G + (size_t)js[oo] * n;
                    const float *ot = OT + ((size_t)(a * z + o0 + oo) * n + s) * ns;
                    float vt = 0.0f;
                    for (unsigned l = 0; l < ns; l++) {
                        const int sp = S[row + l];
                        if (sp < 0)
                            break;
                        vt = __fadd_rn(vt, __fmul_rn(ot[l], Gj[sp]));
                    }
                    al = __fadd_rn(al, __fmul_rn(vt, gamma));
                }
                alphaK[k] = al;
            }
            __syncthreads();
        }
        if (t == 0) {
            float v = 0.0f;
            for (int k = 0; k < kEnd; k++)
                v = __fadd_rn(v, __fmul_rn(alphaK[k], bS[k]));
            aval[ia] = v;
            if (v > bestValue) {
                bestValue = v;
                bestA = (int)a;
            }
        }
        __syncthreads();
    }
    if (t == 0)
        act[i] = bestA;
}

// --------------------------------------------------------------------- 4. materialize

template <bool EXACT>
__global__ __launch_bounds__(kPB) void pbvi_materialize(unsigned n, unsigned ns, unsigned m,
        unsigned z, float gamma, int useLds, const int *__restrict__ S,
        const float *__restrict__ T, const float *__restrict__ O, const float *__restrict__ R,
        const float *__restrict__ OT, const float *__restrict__ G,
        const int *__restrict__ jstar, const int *__restrict__ act,
        float *__restrict__ Gp, unsigned *__restrict__ pi)
{
    extern __shared__ float row[];   // kLdsRowFloats when useLds
    const unsigned i = blockIdx.x, t = threadIdx.x;
    const int a = act[i];
    if (a < 0)
        return;   // no action beat -1e10: Gamma'[i] and pi[i] keep their values
    const int *js = jstar + ((size_t)i * m + a) * z;
    float *out = Gp + (size_t)i * n;
    if (EXACT || !useLds) {
        // alpha(s) = R[s,a] + sum_o (gamma * sum_l (O*T) * Gamma[j*_o, sp]), o ascending
        for (unsigned s = t; s < n; s += kPB)
            out[s] = R[(size_t)s * m + a];
        for (unsigned o = 0; o < z; o++) {
            const float *Gj = G + (size_t)js[o] * n;
            if (useLds) {
                __syncthreads();
                for (unsigned s = t; s < n; s += kPB)
                    row[s] = Gj[s];
                __syncthreads();
            }
            const float *src = useLds ? row : Gj;
            const float *ot = OT + ((size_t)(a * z + o) * n) * ns;
            for (unsigned s = t; s < n; s += kPB) {
                const size_t rw = ((size_t)s * m + a) * ns;
                float vt = 0.0f;
                for (unsigned l = 0; l < ns; l++) {
                    const int sp = S[rw + l];
                    if (sp < 0)
                        break;
                    vt = __fadd_rn(vt, __fmul_rn(ot[(size_t)s * ns + l], src[sp]));
                }
                out[s] = __fadd_rn(out[s], __fmul_rn(vt, gamma));
            }
        }
    } else {
        // h(sp) = sum_o O[a,sp,o] * Gamma[j*_o, sp] in LDS, then R + gamma*sum_l T*h(S)
        for (unsigned sp = t; sp < n; sp += kPB) {
            const float *Os = O + ((size_t)a * n + sp) * z;
            float h = 0.0f;
            for (unsigned o = 0; o < z; o++)
                h = __builtin_fmaf(Os[o], G[(size_t)js[o] * n + sp], h);
            row[sp] = h;
        }
        __syncthreads();
        for (unsigned s = t; s < n; s += kPB) {
            const size_t rw = ((size_t)s * m + a) * ns;
            float acc = 0.0f;
            for (unsigned l = 0; l < ns; l++) {
                const int sp = S[rw + l];
                if (sp < 0)
                    break;
                acc = __builtin_fmaf(T[rw + l], row[sp], acc);
            }
            out[s] = __builtin_fmaf(gamma, acc, R[(size_t)s * m + a]);
        }
    }
    if (t == 0)
        pi[i] = (unsigned)a;
}

} // namespace

// Scoring and materialisation arithmetic: 0 = fast (default), 1 = exact reference order.
static int g_pbvi_mode = 0;

// Device workspace carved out of one allocation (POMDPPBVIGPU::d_alphaBA).
struct PbviWs {
    float *OT;
    float *GT;
    float *partVal;
    int *partJ;
    int *jstar;
    float *aval;
    int *act;
    unsigned tiles;
};

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

static size_t pbvi_ws_bytes(const nova::POMDP *p, unsigned tiles)
{
    const size_t rmz = (size_t)p->r * p->m * p->z;
    return align256(sizeof(float) * (size_t)p->m * p->z * p->n * p->ns) +
           align256(sizeof(float) * (size_t)p->n * p->r) +
           align256(sizeof(float) * tiles * rmz) + align256(sizeof(int) * tiles * rmz) +
           align256(sizeof(int) * rmz) + align256(sizeof(float) * (size_t)p->r * p->m) +
           align256(sizeof(int) * p->r);
}

static PbviWs pbvi_ws(const nova::POMDP *p, void *base)
{
    PbviWs w;
    w.tiles = (p->r + kJT - 1) / kJT;
    const size_t rmz = (size_t)p->r * p->m * p->z;
    char *c = static_cast<char *>(base);
    w.OT = reinterpret_cast<float *>(c);
    c += align256(sizeof(float) * (size_t)p->m * p->z * p->n * p->ns);
    w.GT = reinterpret_cast<float *>(c);
    c += align256(sizeof(float) * (size_t)p->n * p->r);
    w.partVal = reinterpret_cast<float *>(c);
    c += align256(sizeof(float) * w.tiles * rmz);
    w.partJ = reinterpret_cast<int *>(c);
    c += align256(sizeof(int) * w.tiles * rmz);
    w.jstar = reinterpret_cast<int *>(c);
    c += align256(sizeof(int) * rmz);
    w.aval = reinterpret_cast<float *>(c);
    c += align256(sizeof(float) * (size_t)p->r * p->m);
    w.act = reinterpret_cast<int *>(c);
    return w;
}

static bool pbvi_shape_ok(const nova::POMDP *p)
{
    return p->rz <= (unsigned)kKMax && (size_t)p->rz * p->ns <= (size_t)kUMax;
}

// Whether a Gamma row / h vector of n floats can be staged in LDS by the materialize kernels
// (their dynamic LDS must be raised above the 64 KB default once per process).
static bool pbvi_lds_rows(unsigned n)
{
    static int state = 0;   // 0 unknown, 1 ok, 2 unavailable
    if (state == 0) {
        const int bytes = kLdsRowFloats * (int)sizeof(float);
        hipError_t e1 = hipFuncSetAttribute((const void *)pbvi_materialize<true>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        hipError_t e2 = hipFuncSetAttribute((const void *)pbvi_materialize<false>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        state = (e1 == hipSuccess && e2 == hipSuccess) ? 1 : 2;
        if (state == 2)
            (void)hipGetLastError();
    }
    return state == 1 && n <= (unsigned)kLdsRowFloats;
}

// Enqueues one PBVI iteration G -> Gp on stream st.
static void pbvi_enqueue(const nova::POMDP *p, const PbviWs &w, const float *G, float *Gp,
                         unsigned *pi, int mode, hipStream_t st)
{
    const unsigned n = p->n, r = p->r, m = p->m;
    dim3 tg((n + 63) / 64, (r + 63) / 64);
    hipLaunchKernelGGL(gamma_transpose, tg, dim3(kPB), 0, st, r, n, G, w.GT);
    dim3 sg(r * m, w.tiles);
    if (mode == 1)
        hipLaunchKernelGGL(pbvi_score<true>, sg, dim3(kPB), 0, st, n, p->ns, m, p->z, r, p->rz,
                           p->gamma, p->d_S, p->d_Z, p->d_B, w.OT, w.GT, w.partVal, w.partJ);
    else
        hipLaunchKernelGGL(pbvi_score<false>, sg, dim3(kPB), 0, st, n, p->ns, m, p->z, r, p->rz,
                           p->gamma, p->d_S, p->d_Z, p->d_B, w.OT, w.GT, w.partVal, w.partJ);
    auto actFn = (mode == 1) ? pbvi_action<true> : pbvi_action<false>;
    hipLaunchKernelGGL(actFn, dim3(r), dim3(kPB), 0, st, n, p->ns, m, p->z, r, p->rz, w.tiles,
                       p->gamma, p->d_S, p->d_R, p->d_Z, p->d_B, w.OT, G, w.partVal, w.partJ,
                       w.jstar, w.aval, w.act);
    const int useLds = pbvi_lds_rows(n) ? 1 : 0;
    const size_t lds = useLds ? sizeof(float) * n : 0;
    auto matFn = (mode == 1) ? pbvi_materialize<true> : pbvi_materialize<false>;
    hipLaunchKernelGGL(matFn, dim3(r), dim3(kPB), lds, st, n, p->ns, m, p->z, p->gamma, useLds,
                       p->d_S, p->d_T, p->d_O, p->d_R, w.OT, G, w.jstar, w.act, Gp, pi);
}

static bool pomdp_device_ready(const nova::POMDP *p)
{
    return p != nullptr && p->n != 0 && p->ns != 0 && p->m != 0 && p->z != 0 && p->r != 0 &&
           p->rz != 0 && p->d_S != nullptr && p->d_T != nullptr && p->d_O != nullptr &&
           p->d_R != nullptr && p->d_Z != nullptr && p->d_B != nullptr;
}

// Cached executable graph of `iters` iterations.
struct PbviGraph {
    hipGraphExec_t exec = nullptr;
    const void *key[12] = {};
    unsigned iters = 0, parity = 0, dims[6] = {};
    float gamma = 0.0f;
    int mode = -1;
};
static PbviGraph g_pgraph;

static int pbvi_run_iters(const nova::POMDP *p, nova::POMDPPBVIGPU *pb, unsigned iters,
                          hipStream_t st, hipEvent_t t0, hipEvent_t t1, const char *fn)
{
    const PbviWs w = pbvi_ws(p, pb->d_alphaBA);
    const unsigned parity = pb->currentHorizon % 2;
    const void *key[12] = {p->d_S, p->d_T, p->d_O, p->d_R, p->d_Z, p->d_B, pb->d_Gamma,
                           pb->d_GammaPrime, pb->d_pi, pb->d_alphaBA, nullptr, nullptr};
    const unsigned dims[6] = {p->n, p->ns, p->m, p->z, p->r, p->rz};
    const bool hit = g_pgraph.exec && g_pgraph.iters == iters && g_pgraph.parity == parity &&
                     g_pgraph.gamma == p->gamma && g_pgraph.mode == g_pbvi_mode &&
                     std::memcmp(g_pgraph.key, key, sizeof(key)) == 0 &&
                     std::memcmp(g_pgraph.dims, dims, sizeof(dims)) == 0;
    if (!hit) {
        if (g_pgraph.exec) {
            hipGraphExecDestroy(g_pgraph.exec);
            g_pgraph.exec = nullptr;
        }
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess)
            return hip_fail(e, fn, "Failed to begin graph capture.", NOVA_ERROR_KERNEL_EXECUTION);
        for (unsigned k = 0; k < iters; k++) {
            const bool even = ((parity + k) % 2) == 0;
            pbvi_enqueue(p, w, even ? pb->d_Gamma : pb->d_GammaPrime,
                         even ? pb->d_GammaPrime : pb->d_Gamma, pb->d_pi, g_pbvi_mode, st);
        }
        e = hipStreamEndCapture(st, &graph);
        if (e != hipSuccess)
            return hip_fail(e, fn, "Failed to capture the PBVI iterations.",
                            NOVA_ERROR_KERNEL_EXECUTION);
        e = hipGraphInstantiate(&g_pgraph.exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (e != hipSuccess) {
            g_pgraph.exec = nullptr;
            return hip_fail(e, fn, "Failed to instantiate the PBVI graph.",
                            NOVA_ERROR_KERNEL_EXECUTION);
        }
        std::memcpy(g_pgraph.key, key, sizeof(key));
        std::memcpy(g_pgraph.dims, dims, sizeof(dims));
        g_pgraph.iters = iters;
        g_pgraph.parity = parity;
        g_pgraph.gamma = p->gamma;
        g_pgraph.mode = g_pbvi_mode;
    }
    if (t0) hipEventRecord(t0, st);
    hipError_t e = hipGraphLaunch(g_pgraph.exec, st);
    if (e != hipSuccess)
        return hip_fail(e, fn, "Failed to execute the PBVI graph.", NOVA_ERROR_KERNEL_EXECUTION);
    if (t1) hipEventRecord(t1, st);
    e = hipStreamSynchronize(st);
    if (e != hipSuccess)
        return hip_fail(e, fn, "Failed to synchronize after the PBVI iterations.",
                        NOVA_ERROR_DEVICE_SYNCHRONIZE);
    pb->currentHorizon += iters;
    return NOVA_SUCCESS;
}

} // namespace nova_amd

namespace nova {

using namespace nova_amd;

int nova_amd_pbvi_mode(int mode)
{
    const int old = g_pbvi_mode;
    if (mode == 0 || mode == 1)
        g_pbvi_mode = mode;
    return old;
}

int pomdp_pbvi_initialize_gpu(const POMDP *pomdp, POMDPPBVIGPU *pbvi)
{
    const char *fn = "pomdp_pbvi_initialize_gpu";
    if (!pomdp_device_ready(pomdp) || pbvi == nullptr || pbvi->GammaInitial == nullptr) {
        NOVA_ERR(fn, "Invalid arguments (the POMDP must be resident on the device and "
                     "GammaInitial set).");
        return NOVA_ERROR_INVALID_DATA;
    }
    if (!pbvi_shape_ok(pomdp)) {
        NOVA_ERR(fn, "Belief supports too long for this implementation (rz <= 512 and "
                     "rz*ns <= 512 required).");
        return NOVA_ERROR_INVALID_DATA;
    }
    device_free_if_owned(pbvi->d_Gamma, fn, "Gamma");
    device_free_if_owned(pbvi->d_GammaPrime, fn, "Gamma (prime)");
    device_free_if_owned(pbvi->d_pi, fn, "pi");
    device_free_if_owned(pbvi->d_alphaBA, fn, "the PBVI workspace");
    pbvi->d_Gamma = pbvi->d_GammaPrime = nullptr;
    pbvi->d_pi = nullptr;
    pbvi->d_alphaBA = nullptr;
    pbvi->currentHorizon = 0;

    const size_t rn = (size_t)pomdp->r * pomdp->n;
    const unsigned tiles = (pomdp->r + kJT - 1) / kJT;
    int rc = device_alloc((void **)&pbvi->d_Gamma, rn * sizeof(float), fn, "Gamma");
    if (rc == NOVA_SUCCESS)
        rc = device_alloc((void **)&pbvi->d_GammaPrime, rn * sizeof(float), fn, "Gamma (prime)");
    if (rc == NOVA_SUCCESS)
        rc = device_alloc((void **)&pbvi->d_pi, pomdp->r * sizeof(unsigned), fn, "pi");
    if (rc == NOVA_SUCCESS)
        rc = device_alloc((void **)&pbvi->d_alphaBA, pbvi_ws_bytes(pomdp, tiles), fn,
                          "the PBVI workspace");
    if (rc != NOVA_SUCCESS)
        return rc;
    hipError_t e = hipMemcpy(pbvi->d_Gamma, pbvi->GammaInitial, rn * sizeof(float),
                             hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(pbvi->d_GammaPrime, pbvi->GammaInitial, rn * sizeof(float),
                      hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemset(pbvi->d_pi, 0, pomdp->r * sizeof(unsigned));
    if (e != hipSuccess)
        return hip_fail(e, fn, "Failed to copy memory from host to device for Gamma.",
                        NOVA_ERROR_MEMCPY_TO_DEVICE);
    const PbviWs w = pbvi_ws(pomdp, pbvi->d_alphaBA);
    const size_t total = (size_t)pomdp->m * pomdp->z * pomdp->n * pomdp->ns;
    const unsigned grid = (unsigned)std::min<size_t>((total + kPB - 1) / kPB, 65536);
    hipStream_t st = stream();
    hipLaunchKernelGGL(pbvi_build_ot, dim3(grid), dim3(kPB), 0, st, pomdp->n, pomdp->ns,
                       pomdp->m, pomdp->z, pomdp->d_S, pomdp->d_T, pomdp->d_O, w.OT);
    e = hipGetLastError();
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    return hip_fail(e, fn, "Failed to build the observation-transition products.",
                    NOVA_ERROR_KERNEL_EXECUTION);
}

int pomdp_pbvi_uninitialize_gpu(const POMDP *pomdp, POMDPPBVIGPU *pbvi)
{
    const char *fn = "pomdp_pbvi_uninitialize_gpu";
    (void)pomdp;
    if (pbvi == nullptr) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    int rc = NOVA_SUCCESS;
    pbvi->currentHorizon = 0;
    if (device_free_if_owned(pbvi->d_Gamma, fn, "Gamma") != NOVA_SUCCESS) rc = NOVA_ERROR_DEVICE_FREE;
    if (device_free_if_owned(pbvi->d_GammaPrime, fn, "Gamma (prime)") != NOVA_SUCCESS) rc = NOVA_ERROR_DEVICE_FREE;
    if (device_free_if_owned(pbvi->d_pi, fn, "pi") != NOVA_SUCCESS) rc = NOVA_ERROR_DEVICE_FREE;
    if (device_free_if_owned(pbvi->d_alphaBA, fn, "the PBVI workspace") != NOVA_SUCCESS) rc = NOVA_ERROR_DEVICE_FREE;
    pbvi->d_Gamma = pbvi->d_GammaPrime = nullptr;
    pbvi->d_pi = nullptr;
    pbvi->d_alphaBA = nullptr;
    return rc;
}

int pomdp_pbvi_update_gpu(const POMDP *pomdp, POMDPPBVIGPU *pbvi)
{
    return pbvi_run_iters(pomdp, pbvi, 1, stream(), nullptr, nullptr, "pomdp_pbvi_update_gpu");
}

int pomdp_pbvi_get_policy_gpu(const POMDP *pomdp, POMDPPBVIGPU *pbvi, POMDPAlphaVectors *policy)
{
    const char *fn = "pomdp_pbvi_get_policy_gpu";
    if (pomdp == nullptr || pbvi == nullptr || policy == nullptr) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    if (pomdp_alpha_vectors_initialize(policy, pomdp->n, pomdp->m, pomdp->r) != NOVA_SUCCESS) {
        NOVA_ERR(fn, "Could not create the policy.");
        return NOVA_ERROR_POLICY_CREATION;
    }
    const float *G = (pbvi->currentHorizon % 2 == 0) ? pbvi->d_Gamma : pbvi->d_GammaPrime;
    hipError_t e = hipMemcpy(policy->Gamma, G, sizeof(float) * (size_t)pomdp->r * pomdp->n,
                             hipMemcpyDeviceToHost);
    if (e == hipSuccess)
        e = hipMemcpy(policy->pi, pbvi->d_pi, sizeof(unsigned) * pomdp->r, hipMemcpyDeviceToHost);
    return hip_fail(e, fn, "Failed to copy memory from device to host for Gamma/pi.",
                    NOVA_ERROR_MEMCPY_TO_HOST);
}

int pomdp_pbvi_execute_gpu(const POMDP *pomdp, POMDPPBVIGPU *pbvi, POMDPAlphaVectors *policy)
{
    const char *fn = "pomdp_pbvi_execute_gpu";
    if (!pomdp_device_ready(pomdp) || pomdp->gamma < 0.0f || pomdp->gamma > 1.0f ||
            pomdp->horizon < 1 || pbvi == nullptr || pbvi->GammaInitial == nullptr ||
            policy == nullptr) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    if (pbvi->numThreads % 32 != 0) {
        NOVA_ERR(fn, "Invalid number of threads.");
        return NOVA_ERROR_INVALID_CUDA_PARAM;
    }
    int rc = pomdp_pbvi_initialize_gpu(pomdp, pbvi);
    if (rc != NOVA_SUCCESS)
        return rc;
    rc = pbvi_run_iters(pomdp, pbvi, pomdp->horizon, stream(), nullptr, nullptr, fn);
    if (rc == NOVA_SUCCESS)
        rc = pomdp_pbvi_get_policy_gpu(pomdp, pbvi, policy);
    const int rc2 = pomdp_pbvi_uninitialize_gpu(pomdp, pbvi);
    return (rc != NOVA_SUCCESS) ? rc : rc2;
}

int pomdp_pbvi_run_gpu(const POMDP *pomdp, POMDPPBVIGPU *pbvi, unsigned int iterations,
                       float *elapsedMs)
{
    const char *fn = "pomdp_pbvi_run_gpu";
    if (!pomdp_device_ready(pomdp) || pbvi == nullptr || pbvi->d_Gamma == nullptr ||
            pbvi->d_GammaPrime == nullptr || pbvi->d_alphaBA == nullptr || iterations == 0) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    hipStream_t st = stream();
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (elapsedMs) {
        hipEventCreate(&t0);
        hipEventCreate(&t1);
    }
    int rc = pbvi_run_iters(pomdp, pbvi, iterations, st, t0, t1, fn);
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

int pomdp_pbvi_argmax_gpu(const POMDP *pomdp, const POMDPPBVIGPU *pbvi, int *jstar,
                          unsigned int *actions)
{
    const char *fn = "pomdp_pbvi_argmax_gpu";
    if (pomdp == nullptr || pbvi == nullptr || pbvi->d_alphaBA == nullptr) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    const PbviWs w = pbvi_ws(pomdp, pbvi->d_alphaBA);
    hipError_t e = hipSuccess;
    if (jstar)
        e = hipMemcpy(jstar, w.jstar, sizeof(int) * (size_t)pomdp->r * pomdp->m * pomdp->z,
                      hipMemcpyDeviceToHost);
    if (e == hipSuccess && actions)
        e = hipMemcpy(actions, w.act, sizeof(int) * pomdp->r, hipMemcpyDeviceToHost);
    return hip_fail(e, fn, "Failed to copy the argmax record to the host.",
                    NOVA_ERROR_MEMCPY_TO_HOST);
}

} // namespace nova
