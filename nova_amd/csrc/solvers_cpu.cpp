// nova-mi355x: the host solvers of the drop-in ABI (value iteration, PBVI, Perseus).
// They reproduce nova's CPU arithmetic bit for bit (same loop and evaluation order, strict
// comparisons, -1e10 sentinels); this file is compiled with -ffp-contract=off and no
// fast-math so that every product and sum is rounded exactly as in the reference build.
#include <cstdlib>
#include <cstring>
#include <vector>

#include <nova/mdp/algorithms/mdp_vi_cpu.h>
#include <nova/pomdp/algorithms/pomdp_pbvi_cpu.h>
#include <nova/pomdp/algorithms/pomdp_perseus_cpu.h>

#include "nova_common.h"

namespace nova_amd {

// One Bellman sweep V -> Vp (src/mdp/algorithms/mdp_vi_cpu.cpp:36-63).
static void vi_sweep(const nova::MDP *mdp, const float *V, float *Vp, unsigned *pi)
{
    const unsigned n = mdp->n, m = mdp->m, ns = mdp->ns;
    const float gamma = mdp->gamma;
    for (unsigned s = 0; s < n; s++) {
        float best = kLow;
        unsigned bestA = 0;
        for (unsigned a = 0; a < m; a++) {
            const int *Sr = mdp->S + idx3(s, a, 0, m, ns);
            const float *Tr = mdp->T + idx3(s, a, 0, m, ns);
            float q = mdp->R[(size_t)s * m + a];
            for (unsigned i = 0; i < ns && Sr[i] >= 0; i++)
                q += gamma * Tr[i] * V[Sr[i]];
            if (a == 0 || q > best) {
                best = q;
                bestA = a;
            }
        }
        Vp[s] = best;
        pi[s] = bestA;
    }
}

// Value of one action's alpha vector on one belief, through the observation-wise argmax over
// the alpha set (src/pomdp/algorithms/pomdp_pbvi_cpu.cpp:37-103 and :126-150).
struct BackupScratch {
    std::vector<float> alpha;
};

// (O*T)*Gamma summed over the successor list, then times gamma (pomdp_pbvi_cpu.cpp:63-72).
static inline float project(const nova::POMDP *p, const float *G, unsigned s, unsigned a,
        unsigned o)
{
    const size_t row = idx3(s, a, 0, p->m, p->ns);
    const float *Oa = p->O + (size_t)a * p->n * p->z;
    float v = 0.0f;
    for (unsigned l = 0; l < p->ns; l++) {
        const int sp = p->S[row + l];
        if (sp < 0)
            break;
        v += Oa[(size_t)sp * p->z + o] * p->T[row + l] * G[sp];
    }
    v *= p->gamma;
    return v;
}

// Backup of belief i against rG alphas; the winner goes to out/aOut. With no candidate alpha
// (rG == 0) index 0 is used, where the reference reads an uninitialised index
// (pomdp_perseus_cpu.cpp:75-117). Returns whether any action beat -1e10.
static bool backup_belief(const nova::POMDP *p, const float *Gamma, unsigned rG, unsigned i,
        float *alpha, float *out, unsigned *aOut)
{
    const unsigned n = p->n, rz = p->rz;
    const int *Zi = p->Z + (size_t)i * rz;
    const float *Bi = p->B + (size_t)i * rz;
    float bestValue = kLow;
    bool found = false;
    for (unsigned a = 0; a < p->m; a++) {
        for (unsigned s = 0; s < n; s++)
            alpha[s] = p->R[(size_t)s * p->m + a];
        for (unsigned o = 0; o < p->z; o++) {
            float best = kLow;
            unsigned bestj = 0;
            for (unsigned j = 0; j < rG; j++) {
                const float *Gj = Gamma + (size_t)j * n;
                float value = 0.0f;
                for (unsigned k = 0; k < rz && Zi[k] >= 0; k++)
                    value += project(p, Gj, (unsigned)Zi[k], a, o) * Bi[k];
                if (value > best) {
                    best = value;
                    bestj = j;
                }
            }
            const float *Gb = Gamma + (size_t)bestj * n;
            for (unsigned s = 0; s < n; s++)
                alpha[s] += project(p, Gb, s, a, o);
        }
        float value = 0.0f;
        for (unsigned k = 0; k < rz && Zi[k] >= 0; k++)
            value += alpha[Zi[k]] * Bi[k];
        if (value > bestValue) {
            std::memcpy(out, alpha, sizeof(float) * n);
            *aOut = a;
            bestValue = value;
            found = true;
        }
    }
    return found;
}

// sum_k B[i,k] * alpha[Z[i,k]] over the support (pomdp_perseus_cpu.cpp:37-49).
static float sparse_dot(const nova::POMDP *p, unsigned i, const float *alpha)
{
    const int *Zi = p->Z + (size_t)i * p->rz;
    const float *Bi = p->B + (size_t)i * p->rz;
    float v = 0.0f;
    for (unsigned k = 0; k < p->rz && Zi[k] >= 0; k++)
        v += Bi[k] * alpha[Zi[k]];
    return v;
}

// max_g b_i . Gamma_g, first maximiser from -1e10 (pomdp_perseus_cpu.cpp:52-68).
static float sparse_vb(const nova::POMDP *p, unsigned i, const float *Gamma, unsigned rG,
        unsigned *arg)
{
    float best = kLow;
    for (unsigned g = 0; g < rG; g++) {
        const float v = sparse_dot(p, i, Gamma + (size_t)g * p->n);
        if (best < v) {
            best = v;
            if (arg)
                *arg = g;
        }
    }
    return best;
}

} // namespace nova_amd

namespace nova {

using nova_amd::kLow;

// ---------------------------------------------------------------- value iteration (host)

int mdp_vi_initialize_cpu(const MDP *mdp, MDPVICPU *vi)
{
    if (mdp == nullptr || mdp->n == 0 || vi == nullptr) {
        NOVA_ERR("mdp_vi_initialize_cpu", "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    nova_amd::release_if_owned(vi->V);
    nova_amd::release_if_owned(vi->VPrime);
    nova_amd::release_if_owned(vi->pi);
    vi->currentHorizon = 0;
    vi->V = nova_amd::tracked_new<float>(mdp->n);
    vi->VPrime = nova_amd::tracked_new<float>(mdp->n);
    vi->pi = nova_amd::tracked_new<unsigned int>(mdp->n);
    for (unsigned s = 0; s < mdp->n; s++) {
        const float v0 = (vi->VInitial != nullptr) ? vi->VInitial[s] : 0.0f;
        vi->V[s] = v0;
        vi->VPrime[s] = v0;
    }
    return NOVA_SUCCESS;
}

int mdp_vi_uninitialize_cpu(const MDP *mdp, MDPVICPU *vi)
{
    if (mdp == nullptr || vi == nullptr) {
        NOVA_ERR("mdp_vi_uninitialize_cpu", "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    vi->currentHorizon = 0;
    nova_amd::release(vi->V);
    nova_amd::release(vi->VPrime);
    nova_amd::release(vi->pi);
    return NOVA_SUCCESS;
}

int mdp_vi_update_cpu(const MDP *mdp, MDPVICPU *vi)
{
    if (vi->currentHorizon % 2 == 0)
        nova_amd::vi_sweep(mdp, vi->V, vi->VPrime, vi->pi);
    else
        nova_amd::vi_sweep(mdp, vi->VPrime, vi->V, vi->pi);
    vi->currentHorizon++;
    return NOVA_SUCCESS;
}

int mdp_vi_get_policy_cpu(const MDP *mdp, MDPVICPU *vi, MDPValueFunction *policy)
{
    if (mdp == nullptr || vi == nullptr || policy == nullptr) {
        NOVA_ERR("mdp_vi_get_policy_cpu", "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    if (mdp_value_function_initialize(policy, mdp->n, mdp->m, 0) != NOVA_SUCCESS) {
        NOVA_ERR("mdp_vi_get_policy_cpu", "Could not create the policy.");
        return NOVA_ERROR_POLICY_CREATION;
    }
    const float *V = (vi->currentHorizon % 2 == 0) ? vi->V : vi->VPrime;
    std::memcpy(policy->V, V, sizeof(float) * mdp->n);
    std::memcpy(policy->pi, vi->pi, sizeof(unsigned int) * mdp->n);
    return NOVA_SUCCESS;
}

int mdp_vi_execute_cpu(const MDP *mdp, MDPVICPU *vi, MDPValueFunction *policy)
{
    if (mdp == nullptr || mdp->n == 0 || mdp->ns == 0 || mdp->m == 0 || mdp->S == nullptr ||
            mdp->T == nullptr || mdp->R == nullptr || mdp->gamma < 0.0f || mdp->gamma > 1.0f ||
            mdp->horizon < 1 || vi == nullptr || policy == nullptr) {
        NOVA_ERR("mdp_vi_execute_cpu", "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    int rc = mdp_vi_initialize_cpu(mdp, vi);
    if (rc != NOVA_SUCCESS)
        return rc;
    while (vi->currentHorizon < mdp->horizon)
        mdp_vi_update_cpu(mdp, vi);
    rc = mdp_vi_get_policy_cpu(mdp, vi, policy);
    const int rc2 = mdp_vi_uninitialize_cpu(mdp, vi);
    return (rc != NOVA_SUCCESS) ? rc : rc2;
}

// ---------------------------------------------------------------- PBVI (host)

int pomdp_pbvi_initialize_cpu(const POMDP *pomdp, POMDPPBVICPU *pbvi)
{
    if (pomdp == nullptr || pbvi == nullptr || pbvi->GammaInitial == nullptr ||
            pomdp->r == 0 || pomdp->n == 0) {
        NOVA_ERR("pomdp_pbvi_initialize_cpu", "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    nova_amd::release_if_owned(pbvi->Gamma);
    nova_amd::release_if_owned(pbvi->GammaPrime);
    nova_amd::release_if_owned(pbvi->pi);
    const size_t rn = (size_t)pomdp->r * pomdp->n;
    pbvi->currentHorizon = 0;
    pbvi->Gamma = nova_amd::tracked_new<float>(rn);
    pbvi->GammaPrime = nova_amd::tracked_new<float>(rn);
    pbvi->pi = nova_amd::tracked_new<unsigned int>(pomdp->r);
    std::memcpy(pbvi->Gamma, pbvi->GammaInitial, sizeof(float) * rn);
    std::memcpy(pbvi->GammaPrime, pbvi->GammaInitial, sizeof(float) * rn);
    return NOVA_SUCCESS;
}

int pomdp_pbvi_uninitialize_cpu(const POMDP *pomdp, POMDPPBVICPU *pbvi)
{
    if (pbvi == nullptr) {
        NOVA_ERR("pomdp_pbvi_uninitialize_cpu", "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    (void)pomdp;
    pbvi->currentHorizon = 0;
    nova_amd::release(pbvi->Gamma);
    nova_amd::release(pbvi->GammaPrime);
    nova_amd::release(pbvi->pi);
    return NOVA_SUCCESS;
}

// One PBVI iteration (src/pomdp/algorithms/pomdp_pbvi_cpu.cpp:106-162, :257-273).
int pomdp_pbvi_update_cpu(const POMDP *pomdp, POMDPPBVICPU *pbvi)
{
    const bool even = (pbvi->currentHorizon % 2 == 0);
    const float *G = even ? pbvi->Gamma : pbvi->GammaPrime;
    float *Gp = even ? pbvi->GammaPrime : pbvi->Gamma;
    std::vector<float> alpha(pomdp->n);
    for (unsigned i = 0; i < pomdp->r; i++)
        nova_amd::backup_belief(pomdp, G, pomdp->r, i, alpha.data(),
                                Gp + (size_t)i * pomdp->n, &pbvi->pi[i]);
    pbvi->currentHorizon++;
    return NOVA_SUCCESS;
}

int pomdp_pbvi_get_policy_cpu(const POMDP *pomdp, POMDPPBVICPU *pbvi, POMDPAlphaVectors *policy)
{
    if (pomdp == nullptr || pbvi == nullptr || policy == nullptr) {
        NOVA_ERR("pomdp_pbvi_get_policy_cpu", "Invalid arguments. Policy must be undefined.");
        return NOVA_ERROR_INVALID_DATA;
    }
    if (pomdp_alpha_vectors_initialize(policy, pomdp->n, pomdp->m, pomdp->r) != NOVA_SUCCESS) {
        NOVA_ERR("pomdp_pbvi_get_policy_cpu", "Could not create the policy.");
        return NOVA_ERROR_POLICY_CREATION;
    }
    const float *G = (pbvi->currentHorizon % 2 == 0) ? pbvi->Gamma : pbvi->GammaPrime;
    std::memcpy(policy->Gamma, G, sizeof(float) * (size_t)pomdp->r * pomdp->n);
    std::memcpy(policy->pi, pbvi->pi, sizeof(unsigned int) * pomdp->r);
    return NOVA_SUCCESS;
}

static bool pomdp_host_valid(const POMDP *p)
{
    return p != nullptr && p->n != 0 && p->ns != 0 && p->m != 0 && p->z != 0 && p->r != 0 &&
           p->rz != 0 && p->S != nullptr && p->T != nullptr && p->O != nullptr &&
           p->R != nullptr && p->Z != nullptr && p->B != nullptr && p->gamma >= 0.0f &&
           p->gamma <= 1.0f && p->horizon >= 1;
}

int pomdp_pbvi_execute_cpu(const POMDP *pomdp, POMDPPBVICPU *pbvi, POMDPAlphaVectors *policy)
{
    if (!pomdp_host_valid(pomdp) || pbvi == nullptr || pbvi->GammaInitial == nullptr ||
            policy == nullptr) {
        NOVA_ERR("pomdp_pbvi_execute_cpu", "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    int rc = pomdp_pbvi_initialize_cpu(pomdp, pbvi);
    if (rc != NOVA_SUCCESS)
        return rc;
    while (pbvi->currentHorizon < pomdp->horizon)
        pomdp_pbvi_update_cpu(pomdp, pbvi);
    rc = pomdp_pbvi_get_policy_cpu(pomdp, pbvi, policy);
    if (rc != NOVA_SUCCESS)
        return rc;
    return pomdp_pbvi_uninitialize_cpu(pomdp, pbvi);
}

// ---------------------------------------------------------------- Perseus (host)

int pomdp_perseus_initialize_cpu(const POMDP *pomdp, POMDPPerseusCPU *per)
{
    if (pomdp == nullptr || per == nullptr || per->GammaInitial == nullptr || pomdp->r == 0) {
        NOVA_ERR("pomdp_perseus_initialize_cpu", "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    nova_amd::release_if_owned(per->Gamma);
    nova_amd::release_if_owned(per->GammaPrime);
    nova_amd::release_if_owned(per->pi);
    nova_amd::release_if_owned(per->piPrime);
    nova_amd::release_if_owned(per->BTilde);
    const size_t rn = (size_t)pomdp->r * pomdp->n;
    per->currentHorizon = 0;
    per->Gamma = nova_amd::tracked_new<float>(rn);
    per->GammaPrime = nova_amd::tracked_new<float>(rn);
    per->pi = nova_amd::tracked_new<unsigned int>(pomdp->r);
    per->piPrime = nova_amd::tracked_new<unsigned int>(pomdp->r);
    std::memcpy(per->Gamma, per->GammaInitial, sizeof(float) * rn);
    std::memcpy(per->GammaPrime, per->GammaInitial, sizeof(float) * rn);
    // V_0 and V_1 start empty; every belief needs improving (pomdp_perseus_cpu.cpp:183-221).
    per->rGamma = 0;
    per->rGammaPrime = 0;
    per->rTilde = pomdp->r;
    per->BTilde = nova_amd::tracked_new<unsigned int>(pomdp->r);
    for (unsigned i = 0; i < pomdp->r; i++)
        per->BTilde[i] = i;
    return NOVA_SUCCESS;
}

int pomdp_perseus_uninitialize_cpu(const POMDP *pomdp, POMDPPerseusCPU *per)
{
    if (per == nullptr) {
        NOVA_ERR("pomdp_perseus_uninitialize_cpu", "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    (void)pomdp;
    per->currentHorizon = 0;
    nova_amd::release(per->Gamma);
    nova_amd::release(per->GammaPrime);
    nova_amd::release(per->pi);
    nova_amd::release(per->piPrime);
    nova_amd::release(per->BTilde);
    per->rGamma = per->rGammaPrime = per->rTilde = 0;
    return NOVA_SUCCESS;
}

// One randomized Perseus step (src/pomdp/algorithms/pomdp_perseus_cpu.cpp:312-429). The
// belief is drawn from BTilde with libc rand() by the reference's formula; the index is
// clamped below rTilde (the reference can index one past the end when rand() == RAND_MAX).
int pomdp_perseus_update_cpu(const POMDP *pomdp, POMDPPerseusCPU *per)
{
    const bool even = (per->currentHorizon % 2 == 0);
    float *G = even ? per->Gamma : per->GammaPrime;
    float *Gp = even ? per->GammaPrime : per->Gamma;
    unsigned *rG = even ? &per->rGamma : &per->rGammaPrime;
    unsigned *rGp = even ? &per->rGammaPrime : &per->rGamma;
    unsigned *pi = even ? per->pi : per->piPrime;
    unsigned *piP = even ? per->piPrime : per->pi;
    const unsigned n = pomdp->n;

    unsigned t = (unsigned)((float)std::rand() / (float)RAND_MAX * (float)per->rTilde);
    if (t >= per->rTilde)
        t = per->rTilde - 1;
    const unsigned bIndex = per->BTilde[t];

    std::vector<float> scratch(n), alpha(n, 0.0f);
    unsigned aNew = 0;
    nova_amd::backup_belief(pomdp, G, *rG, bIndex, scratch.data(), alpha.data(), &aNew);
    const float bDotAlpha = nova_amd::sparse_dot(pomdp, bIndex, alpha.data());
    unsigned oldArg = 0;
    const float Vnb = nova_amd::sparse_vb(pomdp, bIndex, G, *rG, &oldArg);
    float *dst = Gp + (size_t)(*rGp) * n;
    if (bDotAlpha >= Vnb) {
        std::memcpy(dst, alpha.data(), sizeof(float) * n);
        piP[*rGp] = aNew;
    } else {
        std::memcpy(dst, G + (size_t)oldArg * n, sizeof(float) * n);
        piP[*rGp] = pi[oldArg];
    }
    (*rGp)++;
    if (*rGp > pomdp->r) {
        NOVA_ERR("pomdp_perseus_update_cpu", "Out of memory. Too many alpha-vectors added.");
        return NOVA_ERROR_OUT_OF_MEMORY;
    }

    per->rTilde = 0;
    for (unsigned i = 0; i < pomdp->r; i++) {
        const float vOld = nova_amd::sparse_vb(pomdp, i, G, *rG, nullptr);
        const float vNew = nova_amd::sparse_vb(pomdp, i, Gp, *rGp, nullptr);
        if (vNew < vOld)
            per->BTilde[per->rTilde++] = i;
    }
    if (per->rTilde == 0) {
        per->currentHorizon++;
        *rG = 0;
        per->rTilde = pomdp->r;
        for (unsigned i = 0; i < pomdp->r; i++)
            per->BTilde[i] = i;
        return NOVA_CONVERGED;
    }
    return NOVA_SUCCESS;
}

int pomdp_perseus_get_policy_cpu(const POMDP *pomdp, POMDPPerseusCPU *per,
        POMDPAlphaVectors *policy)
{
    if (pomdp == nullptr || per == nullptr || policy == nullptr) {
        NOVA_ERR("pomdp_perseus_get_policy_cpu", "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    const bool even = (per->currentHorizon % 2 == 0);
    const unsigned rG = even ? per->rGamma : per->rGammaPrime;
    if (pomdp_alpha_vectors_initialize(policy, pomdp->n, pomdp->m, rG) != NOVA_SUCCESS) {
        NOVA_ERR("pomdp_perseus_get_policy_cpu", "Could not create the policy.");
        return NOVA_ERROR_POLICY_CREATION;
    }
    std::memcpy(policy->Gamma, even ? per->Gamma : per->GammaPrime,
                sizeof(float) * (size_t)rG * pomdp->n);
    std::memcpy(policy->pi, even ? per->pi : per->piPrime, sizeof(unsigned int) * rG);
    return NOVA_SUCCESS;
}

int pomdp_perseus_execute_cpu(const POMDP *pomdp, POMDPPerseusCPU *per, POMDPAlphaVectors *policy)
{
    if (!pomdp_host_valid(pomdp) || per == nullptr || per->GammaInitial == nullptr ||
            policy == nullptr) {
        NOVA_ERR("pomdp_perseus_execute_cpu", "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    int rc = pomdp_perseus_initialize_cpu(pomdp, per);
    if (rc != NOVA_SUCCESS)
        return rc;
    while (per->currentHorizon < pomdp->horizon) {
        rc = NOVA_SUCCESS;
        while (rc != NOVA_CONVERGED) {
            rc = pomdp_perseus_update_cpu(pomdp, per);
            if (rc != NOVA_CONVERGED && rc != NOVA_SUCCESS)
                return rc;
        }
    }
    rc = pomdp_perseus_get_policy_cpu(pomdp, per, policy);
    if (rc != NOVA_SUCCESS)
        return rc;
    return pomdp_perseus_uninitialize_cpu(pomdp, per);
}

} // namespace nova
