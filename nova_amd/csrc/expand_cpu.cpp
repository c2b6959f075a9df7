// nova-mi355x: belief-set expansion on the host (random trajectories, distinct beliefs, PEMA).
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <vector>

#include <nova/pomdp/utilities/pomdp_expand_cpu.h>
#include <nova/pomdp/utilities/pomdp_model_cpu.h>
#include <nova/amd/nova_amd.h>

#include "nova_common.h"

namespace nova_amd {

// Dense copy of sparse belief i (src/pomdp/utilities/pomdp_expand_cpu.cpp:39-54).
static void dense_belief(const nova::POMDP *p, unsigned i, float *b)
{
    std::memset(b, 0, sizeof(float) * p->n);
    const int *Zi = p->Z + (size_t)i * p->rz;
    for (unsigned k = 0; k < p->rz && Zi[k] >= 0; k++)
        b[Zi[k]] = p->B[(size_t)i * p->rz + k];
}

// Pr(o | b, a), summed per state then weighted (pomdp_expand_cpu.cpp:57-79).
static float prob_obs(const nova::POMDP *p, const float *b, unsigned a, unsigned o)
{
    const float *Oa = p->O + (size_t)a * p->n * p->z;
    float pr = 0.0f;
    for (unsigned s = 0; s < p->n; s++) {
        const size_t row = idx3(s, a, 0, p->m, p->ns);
        float val = 0.0f;
        for (unsigned l = 0; l < p->ns; l++) {
            const int sp = p->S[row + l];
            if (sp < 0)
                break;
            val += p->T[row + l] * Oa[(size_t)sp * p->z + o];
        }
        pr += val * b[s];
    }
    return pr;
}

// (unsigned)(rand()/RAND_MAX * k) as the reference draws indices, kept below k: the float
// quotient reaches 1.0 for the top 64 values of rand().
static unsigned draw_index(unsigned k)
{
    unsigned v = (unsigned)((float)std::rand() / (float)RAND_MAX * (float)k);
    return (v >= k && k > 0) ? k - 1 : v;
}

static bool expand_args_ok(const nova::POMDP *p)
{
    return p != nullptr && p->n != 0 && p->ns != 0 && p->m != 0 && p->z != 0 && p->r != 0 &&
           p->rz != 0 && p->S != nullptr && p->T != nullptr && p->O != nullptr &&
           p->R != nullptr && p->Z != nullptr && p->B != nullptr;
}

} // namespace nova_amd

namespace nova {

// Random trajectories (src/pomdp/utilities/pomdp_expand_cpu.cpp:82-163): start from a random
// existing belief, walk a random number h <= horizon of steps with a random action and an
// observation drawn by cumulative Pr(o|b,a) (first o whose running sum reaches the target and
// whose own probability is positive, else o = 0), append every visited belief.
int pomdp_expand_random_cpu(POMDP *pomdp, unsigned int numBeliefsToAdd)
{
    if (!nova_amd::expand_args_ok(pomdp) || pomdp->horizon < 1 || numBeliefsToAdd == 0) {
        NOVA_ERR("pomdp_expand_random_cpu", "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    std::srand((unsigned)std::time(nullptr));
    const unsigned n = pomdp->n;
    std::vector<float> b(n), Bnew((size_t)numBeliefsToAdd * n, 0.0f);
    unsigned added = 0;
    while (added < numBeliefsToAdd) {
        nova_amd::dense_belief(pomdp, nova_amd::draw_index(pomdp->r), b.data());
        const unsigned h = (unsigned)((float)std::rand() / (float)RAND_MAX *
                                      (float)(pomdp->horizon + 1));
        for (unsigned t = 0; t < h; t++) {
            const unsigned a = nova_amd::draw_index(pomdp->m);
            const float target = (float)std::rand() / (float)RAND_MAX;
            float cum = 0.0f;
            unsigned o = 0;
            for (unsigned op = 0; op < pomdp->z; op++) {
                const float pr = nova_amd::prob_obs(pomdp, b.data(), a, op);
                cum += pr;
                if (cum >= target && pr > 0.0f) {
                    o = op;
                    break;
                }
            }
            float *bp = nullptr;
            nova_amd::belief_update_host(pomdp, b.data(), a, o, bp, true);
            std::memcpy(b.data(), bp, sizeof(float) * n);
            delete[] bp;
            std::memcpy(&Bnew[(size_t)added * n], b.data(), sizeof(float) * n);
            if (++added >= numBeliefsToAdd)
                break;
        }
    }
    return pomdp_add_new_raw_beliefs_cpu(pomdp, numBeliefsToAdd, Bnew.data());
}

// Distinct beliefs (src/pomdp/utilities/pomdp_expand_cpu.cpp:166-266): for each belief, the
// successor tau(b,a,o) over (a, o) with Pr(o|b,a) != 0 maximising min_j ||b_j - tau||_1
// (dense sum in state order; strict '<' from 1e10 for the min, strict '>' from -1e10 for the
// max). When no (a, o) qualifies the reference appends an uninitialised buffer; this appends
// zeros (an empty support).
int pomdp_expand_distinct_beliefs_cpu(POMDP *pomdp)
{
    if (!nova_amd::expand_args_ok(pomdp)) {
        NOVA_ERR("pomdp_expand_distinct_beliefs_cpu", "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    const unsigned n = pomdp->n, r = pomdp->r;
    std::vector<float> Bnew((size_t)r * n, 0.0f), b(n), bt(n);
    for (unsigned i = 0; i < r; i++) {
        nova_amd::dense_belief(pomdp, i, b.data());
        float starValue = nova_amd::kLow;
        float *out = &Bnew[(size_t)i * n];
        for (unsigned a = 0; a < pomdp->m; a++) {
            for (unsigned o = 0; o < pomdp->z; o++) {
                if (nova_amd::prob_obs(pomdp, b.data(), a, o) == 0.0f)
                    continue;
                float *bp = nullptr;
                const int rc = nova_amd::belief_update_host(pomdp, b.data(), a, o, bp, true);
                if (rc != NOVA_SUCCESS) {
                    delete[] bp;
                    continue;
                }
                float dmin = nova_amd::kHigh;
                for (unsigned j = 0; j < r; j++) {
                    nova_amd::dense_belief(pomdp, j, bt.data());
                    float d = 0.0f;
                    for (unsigned s = 0; s < n; s++)
                        d += std::fabs(bt[s] - bp[s]);
                    if (d < dmin)
                        dmin = d;
                }
                if (dmin > starValue) {
                    std::memcpy(out, bp, sizeof(float) * n);
                    starValue = dmin;
                }
                delete[] bp;
            }
        }
    }
    return pomdp_add_new_raw_beliefs_cpu(pomdp, r, Bnew.data());
}

// PEMA: adds the one successor belief whose expected approximation error is largest
// (src/pomdp/utilities/pomdp_expand_cpu.cpp:269-466). For every belief b of the set, action a
// and observation o with a valid successor b' = tau(b, a, o): Pr(o | b, a), the set belief c
// closest to b' in L1 (first minimum), an alpha chosen for c, and
//   eps(b') = sum_s ((b'(s) >= b(s) ? Rmax : Rmin) / (1 - gamma) - alpha(s)) * (b'(s) - c(s));
// per action the observation with the largest Pr * eps and the sum of Pr * eps over o; per
// belief the action with the largest sum; over the set the first belief with the largest
// value, whose successor under its action and observation is appended (b itself when that
// update is invalid). Every sum runs in the reference's order with float rounding.
//
// The alpha choice: nova's loop (:391-402) assigns instead of accumulating and compares the
// wrong way, so it picks alpha 0 unless some single product alpha(s) * c(s) lies below -1e10
// (then the last such alpha). That is the default here, for results equal to nova's;
// nova_amd_pema_mode(1) selects the first maximiser of alpha . c instead.
static std::atomic<int> g_pemaMode{0};

int nova_amd_pema_mode(int mode)
{
    if (mode != 0 && mode != 1) {
        NOVA_ERR("nova_amd_pema_mode", "Mode must be 0 (nova's alpha choice) or 1 (argmax).");
        return NOVA_ERROR_INVALID_DATA;
    }
    g_pemaMode.store(mode);
    return NOVA_SUCCESS;
}

} // namespace nova

namespace nova_amd {
// The current alpha choice of PEMA, for the device variant.
int pema_mode_current() { return nova::g_pemaMode.load(); }
} // namespace nova_amd

namespace nova {

// The alpha used for the closest belief c (see above).
static unsigned pema_alpha(const POMDPAlphaVectors *policy, const float *c, int mode)
{
    const unsigned n = policy->n;
    unsigned best = 0;
    if (mode == 0) {
        for (unsigned j = 0; j < policy->r; j++)
            for (unsigned s = 0; s < n; s++)
                if (policy->Gamma[(size_t)j * n + s] * c[s] < nova_amd::kLow)
                    best = j;
        return best;
    }
    float bestV = nova_amd::kLow;
    for (unsigned j = 0; j < policy->r; j++) {
        float v = 0.0f;
        for (unsigned s = 0; s < n; s++)
            v += policy->Gamma[(size_t)j * n + s] * c[s];
        if (v > bestV) {
            bestV = v;
            best = j;
        }
    }
    return best;
}

int pomdp_expand_pema_cpu(POMDP *pomdp, const POMDPAlphaVectors *policy)
{
    const char *fn = "pomdp_expand_pema_cpu";
    if (!nova_amd::expand_args_ok(pomdp) || policy == nullptr || policy->r == 0 ||
            policy->n != pomdp->n || policy->Gamma == nullptr) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    const unsigned n = pomdp->n, m = pomdp->m, z = pomdp->z, r = pomdp->r, rz = pomdp->rz;
    const unsigned ns = pomdp->ns;
    const int mode = g_pemaMode.load();
    float Rmax = nova_amd::kLow, Rmin = nova_amd::kHigh;
    for (size_t x = 0; x < (size_t)n * m; x++) {
        if (Rmax < pomdp->R[x])
            Rmax = pomdp->R[x];
        if (Rmin > pomdp->R[x])
            Rmin = pomdp->R[x];
    }
    // dense copies of the whole set, built once (the reference rebuilds them per comparison)
    std::vector<float> dense((size_t)r * n);
    for (unsigned j = 0; j < r; j++)
        nova_amd::dense_belief(pomdp, j, dense.data() + (size_t)j * n);
    std::vector<float> Bnew(n, 0.0f);
    std::vector<unsigned> oStar(m);
    std::vector<float> oStarValue(m);
    float bound = nova_amd::kLow;
    for (unsigned i = 0; i < r; i++) {
        const float *b = dense.data() + (size_t)i * n;
        unsigned aStar = 0;
        float aStarValue = nova_amd::kLow;
        for (unsigned a = 0; a < m; a++) {
            oStar[a] = 0;
            oStarValue[a] = nova_amd::kLow;
        }
        for (unsigned a = 0; a < m; a++) {
            float aValue = 0.0f;
            for (unsigned o = 0; o < z; o++) {
                float pr = 0.0f;
                for (unsigned k = 0; k < rz; k++) {
                    const int s = pomdp->Z[(size_t)i * rz + k];
                    if (s < 0)
                        break;
                    const size_t row = nova_amd::idx3((unsigned)s, a, 0, m, ns);
                    for (unsigned l = 0; l < ns; l++) {
                        const int sp = pomdp->S[row + l];
                        if (sp < 0)
                            break;
                        pr += pomdp->T[row + l] * pomdp->O[((size_t)a * n + sp) * z + o] *
                              pomdp->B[(size_t)i * rz + k];
                    }
                }
                float *bp = nullptr;
                if (nova_amd::belief_update_host(pomdp, b, a, o, bp, true) != NOVA_SUCCESS) {
                    delete[] bp;
                    continue;
                }
                // with no set belief nearer than 1e10 the reference's copy stays
                // uninitialised; index 0 stands in for it
                unsigned closest = 0;
                float closestDist = nova_amd::kHigh;
                for (unsigned j = 0; j < r; j++) {
                    const float *c = dense.data() + (size_t)j * n;
                    float d = 0.0f;
                    for (unsigned s = 0; s < n; s++)
                        d += std::fabs(bp[s] - c[s]);
                    if (d < closestDist) {
                        closest = j;
                        closestDist = d;
                    }
                }
                const float *c = dense.data() + (size_t)closest * n;
                const float *alpha = policy->Gamma + (size_t)pema_alpha(policy, c, mode) * n;
                float eps = 0.0f;
                for (unsigned s = 0; s < n; s++) {
                    const float lim = (bp[s] >= b[s] ? Rmax : Rmin) / (1.0f - pomdp->gamma);
                    eps += (lim - alpha[s]) * (bp[s] - c[s]);
                }
                if (pr * eps > oStarValue[a]) {
                    oStar[a] = o;
                    oStarValue[a] = pr * eps;
                }
                aValue += pr * eps;
                delete[] bp;
            }
            if (aValue > aStarValue) {
                aStar = a;
                aStarValue = aValue;
            }
        }
        if (aStarValue > bound) {
            float *bp = nullptr;
            if (nova_amd::belief_update_host(pomdp, b, aStar, oStar[aStar], bp, true) == NOVA_SUCCESS)
                std::memcpy(Bnew.data(), bp, sizeof(float) * n);
            else
                std::memcpy(Bnew.data(), b, sizeof(float) * n);
            delete[] bp;
            bound = aStarValue;
        }
    }
    return pomdp_add_new_raw_beliefs_cpu(pomdp, 1, Bnew.data());
}

} // namespace nova
