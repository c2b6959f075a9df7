// nova-mi355x: host-side model allocation and the dense belief utilities.
#include <cmath>
#include <cstring>

#include <nova/mdp/utilities/mdp_model_cpu.h>
#include <nova/pomdp/utilities/pomdp_model_cpu.h>

#include "nova_common.h"

namespace nova {

// src/mdp/utilities/mdp_model_cpu.cpp:35-82: validation, S = -1, T = R = 0, goals = 0.
int mdp_initialize_cpu(MDP *mdp, unsigned int n, unsigned int ns, unsigned int m, float gamma,
        unsigned int horizon, float epsilon, unsigned int s0, unsigned int ng)
{
    if (mdp == nullptr || mdp->goals != nullptr || mdp->S != nullptr || mdp->T != nullptr ||
            mdp->R != nullptr || n == 0 || ns == 0 || n < ns || m == 0 || gamma < 0.0f ||
            gamma > 1.0f || horizon == 0 || epsilon < 0.0f || s0 >= n || ng >= n) {
        NOVA_ERR("mdp_initialize_cpu", "Invalid input.");
        return NOVA_ERROR_INVALID_DATA;
    }
    mdp->n = n;
    mdp->ns = ns;
    mdp->m = m;
    mdp->gamma = gamma;
    mdp->horizon = horizon;
    mdp->epsilon = epsilon;
    mdp->s0 = s0;
    mdp->ng = ng;

    const size_t nmns = (size_t)n * m * ns;
    mdp->S = new int[nmns];
    mdp->T = new float[nmns];
    for (size_t x = 0; x < nmns; x++) {
        mdp->S[x] = -1;
        mdp->T[x] = 0.0f;
    }
    mdp->R = new float[(size_t)n * m]();
    mdp->goals = (ng > 0) ? new unsigned int[ng]() : nullptr;

    mdp->d_goals = nullptr;
    mdp->d_S = nullptr;
    mdp->d_T = nullptr;
    mdp->d_R = nullptr;
    return NOVA_SUCCESS;
}

int mdp_uninitialize_cpu(MDP *mdp)
{
    if (mdp == nullptr) {
        NOVA_ERR("mdp_uninitialize_cpu", "Invalid input.");
        return NOVA_ERROR_INVALID_DATA;
    }
    mdp->n = mdp->ns = mdp->m = 0;
    mdp->gamma = 0.0f;
    mdp->horizon = 0;
    mdp->epsilon = 0.0f;
    mdp->s0 = mdp->ng = 0;
    delete[] mdp->goals;
    delete[] mdp->S;
    delete[] mdp->T;
    delete[] mdp->R;
    mdp->goals = nullptr;
    mdp->S = nullptr;
    mdp->T = nullptr;
    mdp->R = nullptr;
    return NOVA_SUCCESS;
}

// src/pomdp/utilities/pomdp_model_cpu.cpp:34-86. Z starts at 0 (not -1), as in the reference
// (:73-75): a freshly initialised belief set lists state 0 in every slot.
int pomdp_initialize_cpu(POMDP *pomdp, unsigned int n, unsigned int ns, unsigned int m,
        unsigned int z, unsigned int r, unsigned int rz, float gamma, unsigned int horizon)
{
    if (pomdp == nullptr || pomdp->S != nullptr || pomdp->T != nullptr || pomdp->O != nullptr ||
            pomdp->R != nullptr || pomdp->Z != nullptr || pomdp->B != nullptr || n == 0 ||
            ns == 0 || n < ns || m == 0 || z == 0 || r == 0 || rz == 0 || gamma < 0.0f ||
            gamma > 1.0f || horizon == 0) {
        NOVA_ERR("pomdp_initialize_cpu", "Invalid input.");
        return NOVA_ERROR_INVALID_DATA;
    }
    pomdp->n = n;
    pomdp->ns = ns;
    pomdp->m = m;
    pomdp->z = z;
    pomdp->r = r;
    pomdp->rz = rz;
    pomdp->gamma = gamma;
    pomdp->horizon = horizon;

    const size_t nmns = (size_t)n * m * ns;
    pomdp->S = new int[nmns];
    pomdp->T = new float[nmns];
    for (size_t x = 0; x < nmns; x++) {
        pomdp->S[x] = -1;
        pomdp->T[x] = 0.0f;
    }
    pomdp->O = new float[(size_t)m * n * z]();
    pomdp->R = new float[(size_t)n * m]();
    pomdp->Z = new int[(size_t)r * rz]();
    pomdp->B = new float[(size_t)r * rz]();

    pomdp->d_S = nullptr;
    pomdp->d_T = nullptr;
    pomdp->d_O = nullptr;
    pomdp->d_R = nullptr;
    pomdp->d_Z = nullptr;
    pomdp->d_B = nullptr;
    return NOVA_SUCCESS;
}

// tau(b, a, o) into a fresh new[n] buffer. src/pomdp/utilities/pomdp_model_cpu.cpp:89-137:
// scatter T*b(s) along the successor lists, weight by O[a,sp,o] while summing the normaliser
// in state order, warn (code 10, unnormalised bp kept) when |normaliser| < 1e-9, else divide.
int pomdp_belief_update_cpu(const POMDP *pomdp, const float *b, unsigned int a, unsigned int o,
        float *&bp)
{
    return nova_amd::belief_update_host(pomdp, b, a, o, bp, false);
}

} // namespace nova

namespace nova_amd {

// The body of pomdp_belief_update_cpu. quiet: the impossible-observation warning is returned
// without its stderr line (the expansions call it for every (action, observation) and skip
// those, so the line says nothing to their callers); results are the same either way.
int belief_update_host(const nova::POMDP *pomdp, const float *b, unsigned int a, unsigned int o,
        float *&bp, bool quiet)
{
    if (pomdp == nullptr || pomdp->n == 0 || pomdp->ns == 0 || pomdp->m == 0 || pomdp->z == 0 ||
            pomdp->S == nullptr || pomdp->T == nullptr || pomdp->O == nullptr ||
            pomdp->R == nullptr || bp != nullptr) {
        NOVA_ERR("pomdp_belief_update_cpu", "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    const unsigned n = pomdp->n, ns = pomdp->ns, m = pomdp->m, z = pomdp->z;
    float *out = new float[n]();
    for (unsigned s = 0; s < n; s++) {
        const size_t row = nova_amd::idx3(s, a, 0, m, ns);
        for (unsigned l = 0; l < ns; l++) {
            const int sp = pomdp->S[row + l];
            if (sp < 0)
                break;
            out[sp] += pomdp->T[row + l] * b[s];
        }
    }
    const float *Oa = pomdp->O + (size_t)a * n * z;
    float norm = 0.0f;
    for (unsigned sp = 0; sp < n; sp++) {
        out[sp] *= Oa[(size_t)sp * z + o];
        norm += out[sp];
    }
    bp = out;
    if (std::fabs(norm) < NOVA_FLT_ERR_TOL) {
        if (!quiet)
            NOVA_ERR("pomdp_belief_update_cpu",
                     "Computed belief is invalid. The observation is impossible given the belief and action.");
        return NOVA_WARNING_INVALID_BELIEF;
    }
    for (unsigned sp = 0; sp < n; sp++)
        out[sp] /= norm;
    return NOVA_SUCCESS;
}

} // namespace nova_amd

namespace nova {

// Appends dense beliefs to the sparse set. src/pomdp/utilities/pomdp_model_cpu.cpp:157-227:
// rz grows to the largest count of entries > 0; old rows are padded with -1/0; new rows list
// their positive entries in ascending state order, padded with -1/0. (The reference falls off
// the end without a return statement; this returns NOVA_SUCCESS.)
int pomdp_add_new_raw_beliefs_cpu(POMDP *pomdp, unsigned int numBeliefsToAdd, float *Bnew)
{
    if (pomdp == nullptr || pomdp->n == 0 || pomdp->ns == 0 || pomdp->m == 0 || pomdp->z == 0 ||
            numBeliefsToAdd == 0 || Bnew == nullptr) {
        NOVA_ERR("pomdp_add_new_beliefs_raw_cpu", "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    const unsigned n = pomdp->n, r0 = pomdp->r, rz0 = pomdp->rz;
    const unsigned r1 = r0 + numBeliefsToAdd;
    unsigned rz1 = rz0;
    for (unsigned i = 0; i < numBeliefsToAdd; i++) {
        unsigned cnt = 0;
        for (unsigned s = 0; s < n; s++)
            cnt += (Bnew[(size_t)i * n + s] > 0.0f) ? 1u : 0u;
        if (cnt > rz1)
            rz1 = cnt;
    }
    int *Z1 = new int[(size_t)r1 * rz1];
    float *B1 = new float[(size_t)r1 * rz1];
    for (unsigned i = 0; i < r0; i++) {
        for (unsigned k = 0; k < rz1; k++) {
            const bool old = k < rz0;
            Z1[(size_t)i * rz1 + k] = old ? pomdp->Z[(size_t)i * rz0 + k] : -1;
            B1[(size_t)i * rz1 + k] = old ? pomdp->B[(size_t)i * rz0 + k] : 0.0f;
        }
    }
    for (unsigned i = 0; i < numBeliefsToAdd; i++) {
        const float *src = Bnew + (size_t)i * n;
        int *zr = Z1 + (size_t)(r0 + i) * rz1;
        float *br = B1 + (size_t)(r0 + i) * rz1;
        unsigned k = 0;
        for (unsigned s = 0; s < n; s++) {
            if (src[s] > 0.0f) {
                zr[k] = (int)s;
                br[k] = src[s];
                k++;
            }
        }
        for (; k < rz1; k++) {
            zr[k] = -1;
            br[k] = 0.0f;
        }
    }
    delete[] pomdp->Z;
    delete[] pomdp->B;
    pomdp->r = r1;
    pomdp->rz = rz1;
    pomdp->Z = Z1;
    pomdp->B = B1;
    return NOVA_SUCCESS;
}

int pomdp_uninitialize_cpu(POMDP *pomdp)
{
    if (pomdp == nullptr) {
        NOVA_ERR("pomdp_uninitialize_cpu", "Invalid input.");
        return NOVA_ERROR_INVALID_DATA;
    }
    pomdp->n = pomdp->ns = pomdp->m = pomdp->z = pomdp->r = pomdp->rz = 0;
    pomdp->gamma = 0.0f;
    pomdp->horizon = 0;
    delete[] pomdp->S;
    delete[] pomdp->T;
    delete[] pomdp->O;
    delete[] pomdp->R;
    delete[] pomdp->Z;
    delete[] pomdp->B;
    pomdp->S = nullptr;
    pomdp->T = nullptr;
    pomdp->O = nullptr;
    pomdp->R = nullptr;
    pomdp->Z = nullptr;
    pomdp->B = nullptr;
    return NOVA_SUCCESS;
}

} // namespace nova
