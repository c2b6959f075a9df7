// nova-mi355x: output containers (MDPValueFunction, POMDPAlphaVectors).
// Ownership contract of src/mdp/policies/mdp_value_function.cpp:34-87 and
// src/pomdp/policies/pomdp_alpha_vectors.cpp:34-107: the library allocates with new[], the
// caller releases through *_uninitialize.
#include <nova/mdp/policies/mdp_value_function.h>
#include <nova/pomdp/policies/pomdp_alpha_vectors.h>

#include "nova_common.h"

namespace nova {

int mdp_value_function_initialize(MDPValueFunction *policy, unsigned int n, unsigned int m,
        unsigned int r)
{
    if (policy == nullptr || policy->S != nullptr || policy->V != nullptr ||
            policy->pi != nullptr || n == 0 || m == 0) {
        NOVA_ERR("mdp_value_function_initialize", "Invalid input.");
        return NOVA_ERROR_INVALID_DATA;
    }
    policy->n = n;
    policy->m = m;
    policy->r = r;
    const unsigned len = (r > 0) ? r : n;
    policy->S = (r > 0) ? new unsigned int[r] : nullptr;
    policy->V = new float[len];
    policy->pi = new unsigned int[len];
    return NOVA_SUCCESS;
}

int mdp_value_function_uninitialize(MDPValueFunction *policy)
{
    if (policy == nullptr) {
        NOVA_ERR("mdp_value_function_uninitialize", "Invalid input.");
        return NOVA_ERROR_INVALID_DATA;
    }
    policy->n = policy->m = policy->r = 0;
    delete[] policy->S;
    delete[] policy->V;
    delete[] policy->pi;
    policy->S = nullptr;
    policy->V = nullptr;
    policy->pi = nullptr;
    return NOVA_SUCCESS;
}

int pomdp_alpha_vectors_initialize(POMDPAlphaVectors *policy, unsigned int n, unsigned int m,
        unsigned int r)
{
    if (policy == nullptr || policy->Gamma != nullptr || policy->pi != nullptr ||
            n == 0 || m == 0 || r == 0) {
        NOVA_ERR("pomdp_alpha_vectors_initialize", "Invalid input.");
        return NOVA_ERROR_INVALID_DATA;
    }
    policy->n = n;
    policy->m = m;
    policy->r = r;
    policy->Gamma = new float[(size_t)r * n];
    policy->pi = new unsigned int[r];
    return NOVA_SUCCESS;
}

// Dense alpha . b per vector, accumulated in state order; the first strictly larger value
// from -1e10 wins (pomdp_alpha_vectors.cpp:52-79).
int pomdp_alpha_vectors_value_and_action(const POMDPAlphaVectors *policy, const float *b,
        float &Vb, unsigned int &a)
{
    if (policy == nullptr || policy->n == 0 || policy->m == 0 || policy->r == 0 ||
            policy->Gamma == nullptr || policy->pi == nullptr || b == nullptr) {
        NOVA_ERR("pomdp_alpha_vectors_value_and_action", "Invalid input.");
        return NOVA_ERROR_INVALID_DATA;
    }
    float best = nova_amd::kLow;
    unsigned act = 0;
    for (unsigned i = 0; i < policy->r; i++) {
        const float *row = policy->Gamma + (size_t)i * policy->n;
        float v = 0.0f;
        for (unsigned s = 0; s < policy->n; s++)
            v += row[s] * b[s];
        if (best < v) {
            best = v;
            act = policy->pi[i];
        }
    }
    Vb = best;
    a = act;
    return NOVA_SUCCESS;
}

int pomdp_alpha_vectors_uninitialize(POMDPAlphaVectors *policy)
{
    if (policy == nullptr) {
        NOVA_ERR("pomdp_alpha_vectors_uninitialize", "Invalid input.");
        return NOVA_ERROR_INVALID_DATA;
    }
    policy->n = policy->m = policy->r = 0;
    delete[] policy->Gamma;
    delete[] policy->pi;
    policy->Gamma = nullptr;
    policy->pi = nullptr;
    return NOVA_SUCCESS;
}

} // namespace nova
