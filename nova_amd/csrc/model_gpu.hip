// nova-mi355x: HBM residency of the model arrays (hipMalloc + copy into the d_* fields).
// Contract of src/mdp/utilities/mdp_model_gpu.cu:34-259 and
// src/pomdp/utilities/pomdp_model_gpu.cu:34-353: each helper validates, allocates and
// uploads one array; the aggregate initialize/uninitialize return the sum of the helpers'
// codes. Additionally, successor and belief-support indices are checked on the host before
// upload, because the kernels gather with them unchecked.
#include <nova/mdp/utilities/mdp_model_gpu.h>
#include <nova/pomdp/utilities/pomdp_model_gpu.h>

#include <atomic>

#include "nova_hip.h"

namespace nova_amd {

static std::atomic<unsigned long long> g_genTransitions{1}, g_genAll{1};

unsigned long long model_generation(bool transitionsOnly)
{
    return transitionsOnly ? g_genTransitions.load() : g_genAll.load();
}

void model_changed(bool transitions)
{
    if (transitions)
        g_genTransitions.fetch_add(1);
    g_genAll.fetch_add(1);
}

// transitions: the array is S, T or O
// Uploads and releases take the library's lock: their synchronous copies run on the legacy null
// stream and must not overlap a graph capture by another thread's library call.
template <typename T>
static int upload(T *&dst, const T *src, size_t count, const char *fn, const char *what,
                  bool transitions)
{
    NOVA_AMD_API_LOCK();
    model_changed(transitions);
    device_free_if_owned(dst, fn, what);
    dst = nullptr;
    void *p = nullptr;
    int rc = device_alloc(&p, count * sizeof(T), fn, what);
    if (rc != NOVA_SUCCESS)
        return rc;
    hipError_t e = hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        device_free(p, fn, what);
        return hip_fail(e, fn, "Failed to copy memory from host to device.",
                        NOVA_ERROR_MEMCPY_TO_DEVICE);
    }
    dst = static_cast<T *>(p);
    return NOVA_SUCCESS;
}

// Releases a d_* field. Only memory this library allocated is freed: a caller may point a
// d_* field at device memory it owns (for instance a PyTorch tensor) and keep ownership.
template <typename T> static int release_device(T *&p, const char *fn, const char *what,
                                                bool transitions)
{
    NOVA_AMD_API_LOCK();
    model_changed(transitions);
    int rc = device_free_if_owned(p, fn, what);
    p = nullptr;
    return rc;
}

} // namespace nova_amd

namespace nova {

using namespace nova_amd;

// ------------------------------------------------------------------------------ MDP

int mdp_initialize_successors_gpu(MDP *mdp)
{
    const char *fn = "mdp_initialize_successors_gpu";
    if (mdp == nullptr || mdp->n == 0 || mdp->m == 0 || mdp->ns == 0 || mdp->S == nullptr) {
        NOVA_ERR(fn, "Invalid input.");
        return NOVA_ERROR_INVALID_DATA;
    }
    if (first_bad_successor(mdp->S, (size_t)mdp->n * mdp->m, mdp->ns, mdp->n) >= 0) {
        NOVA_ERR(fn, "A successor index is outside [0, n).");
        return NOVA_ERROR_INVALID_DATA;
    }
    return upload(mdp->d_S, mdp->S, (size_t)mdp->n * mdp->m * mdp->ns, fn,
                  "the successor states", true);
}

int mdp_uninitialize_successors_gpu(MDP *mdp)
{
    if (mdp == nullptr) {
        NOVA_ERR("mdp_uninitialize_successors_gpu", "Invalid input.");
        return NOVA_ERROR_INVALID_DATA;
    }
    return release_device(mdp->d_S, "mdp_uninitialize_successors_gpu", "the successor states",
                          true);
}

int mdp_initialize_state_transitions_gpu(MDP *mdp)
{
    const char *fn = "mdp_initialize_state_transitions_gpu";
    if (mdp == nullptr || mdp->n == 0 || mdp->m == 0 || mdp->ns == 0 || mdp->T == nullptr) {
        NOVA_ERR(fn, "Invalid input.");
        return NOVA_ERROR_INVALID_DATA;
    }
    return upload(mdp->d_T, mdp->T, (size_t)mdp->n * mdp->m * mdp->ns, fn,
                  "the state transitions", true);
}

int mdp_uninitialize_state_transitions_gpu(MDP *mdp)
{
    if (mdp == nullptr) {
        NOVA_ERR("mdp_uninitialize_state_transitions_gpu", "Invalid input.");
        return NOVA_ERROR_INVALID_DATA;
    }
    return release_device(mdp->d_T, "mdp_uninitialize_state_transitions_gpu",
                          "the state transitions", true);
}

int mdp_initialize_rewards_gpu(MDP *mdp)
{
    const char *fn = "mdp_initialize_rewards_gpu";
    if (mdp == nullptr || mdp->n == 0 || mdp->m == 0 || mdp->R == nullptr) {
        NOVA_ERR(fn, "Invalid input.");
        return NOVA_ERROR_INVALID_DATA;
    }
    return upload(mdp->d_R, mdp->R, (size_t)mdp->n * mdp->m, fn, "the rewards", false);
}

int mdp_uninitialize_rewards_gpu(MDP *mdp)
{
    if (mdp == nullptr) {
        NOVA_ERR("mdp_uninitialize_rewards_gpu", "Invalid input.");
        return NOVA_ERROR_INVALID_DATA;
    }
    return release_device(mdp->d_R, "mdp_uninitialize_rewards_gpu", "the rewards", false);
}

int mdp_initialize_goals_gpu(MDP *mdp)
{
    const char *fn = "mdp_initialize_goals_gpu";
    if (mdp == nullptr || mdp->ng == 0 || mdp->goals == nullptr) {
        NOVA_ERR(fn, "Invalid input.");
        return NOVA_ERROR_INVALID_DATA;
    }
    return upload(mdp->d_goals, mdp->goals, (size_t)mdp->ng, fn, "the goals", false);
}

int mdp_uninitialize_goals_gpu(MDP *mdp)
{
    if (mdp == nullptr) {
        NOVA_ERR("mdp_uninitialize_goals_gpu", "Invalid input.");
        return NOVA_ERROR_INVALID_DATA;
    }
    return release_device(mdp->d_goals, "mdp_uninitialize_goals_gpu", "the goals", false);
}

int mdp_initialize_gpu(MDP *mdp)
{
    if (mdp == nullptr) {
        NOVA_ERR("mdp_initialize_gpu", "Invalid input.");
        return NOVA_ERROR_INVALID_DATA;
    }
    int rc = mdp_initialize_successors_gpu(mdp);
    rc += mdp_initialize_state_transitions_gpu(mdp);
    rc += mdp_initialize_rewards_gpu(mdp);
    if (mdp->ng > 0)
        rc += mdp_initialize_goals_gpu(mdp);
    return rc;
}

int mdp_uninitialize_gpu(MDP *mdp)
{
    if (mdp == nullptr) {
        NOVA_ERR("mdp_uninitialize_gpu", "Invalid input.");
        return NOVA_ERROR_INVALID_DATA;
    }
    int rc = mdp_uninitialize_successors_gpu(mdp);
    rc += mdp_uninitialize_state_transitions_gpu(mdp);
    rc += mdp_uninitialize_rewards_gpu(mdp);
    if (mdp->ng > 0)
        rc += mdp_uninitialize_goals_gpu(mdp);
    return rc;
}

// ---------------------------------------------------------------------------- POMDP

#define NOVA_POMDP_HELPERS(NAME, FIELD, DFIELD, COUNT, WHAT, TRANSITIONS, EXTRA_CHECK)     \
    int pomdp_initialize_##NAME##_gpu(POMDP *pomdp)                                         \
    {                                                                                       \
        const char *fn = "pomdp_initialize_" #NAME "_gpu";                                  \
        if (pomdp == nullptr || pomdp->n == 0 || pomdp->m == 0 || pomdp->FIELD == nullptr) { \
            NOVA_ERR(fn, "Invalid input.");                                                 \
            return NOVA_ERROR_INVALID_DATA;                                                 \
        }                                                                                   \
        EXTRA_CHECK                                                                         \
        return upload(pomdp->DFIELD, pomdp->FIELD, (size_t)(COUNT), fn, WHAT, TRANSITIONS); \
    }                                                                                       \
    int pomdp_uninitialize_##NAME##_gpu(POMDP *pomdp)                                       \
    {                                                                                       \
        if (pomdp == nullptr) {                                                             \
            NOVA_ERR("pomdp_uninitialize_" #NAME "_gpu", "Invalid input.");                 \
            return NOVA_ERROR_INVALID_DATA;                                                 \
        }                                                                                   \
        return release_device(pomdp->DFIELD, "pomdp_uninitialize_" #NAME "_gpu", WHAT,       \
                              TRANSITIONS);                                                 \
    }

NOVA_POMDP_HELPERS(successors, S, d_S, (size_t)pomdp->n * pomdp->m * pomdp->ns,
    "the successor states", true,
    if (pomdp->ns == 0 ||
            first_bad_successor(pomdp->S, (size_t)pomdp->n * pomdp->m, pomdp->ns, pomdp->n) >= 0) {
        NOVA_ERR(fn, "A successor index is outside [0, n).");
        return NOVA_ERROR_INVALID_DATA;
    })
NOVA_POMDP_HELPERS(state_transitions, T, d_T, (size_t)pomdp->n * pomdp->m * pomdp->ns,
    "the state transitions", true, )
NOVA_POMDP_HELPERS(observation_transitions, O, d_O, (size_t)pomdp->m * pomdp->n * pomdp->z,
    "the observation transitions", true, )
NOVA_POMDP_HELPERS(rewards, R, d_R, (size_t)pomdp->n * pomdp->m, "the rewards", false, )
NOVA_POMDP_HELPERS(nonzero_beliefs, Z, d_Z, (size_t)pomdp->r * pomdp->rz,
    "the belief supports", false,
    if (pomdp->r == 0 || pomdp->rz == 0 ||
            first_bad_successor(pomdp->Z, pomdp->r, pomdp->rz, pomdp->n) >= 0) {
        NOVA_ERR(fn, "A belief support index is outside [0, n).");
        return NOVA_ERROR_INVALID_DATA;
    })
NOVA_POMDP_HELPERS(belief_points, B, d_B, (size_t)pomdp->r * pomdp->rz, "the belief values", false, )

int pomdp_initialize_gpu(POMDP *pomdp)
{
    if (pomdp == nullptr) {
        NOVA_ERR("pomdp_initialize_gpu", "Invalid input.");
        return NOVA_ERROR_INVALID_DATA;
    }
    int rc = pomdp_initialize_successors_gpu(pomdp);
    rc += pomdp_initialize_state_transitions_gpu(pomdp);
    rc += pomdp_initialize_observation_transitions_gpu(pomdp);
    rc += pomdp_initialize_rewards_gpu(pomdp);
    if (pomdp->r > 0) {
        rc += pomdp_initialize_nonzero_beliefs_gpu(pomdp);
        rc += pomdp_initialize_belief_points_gpu(pomdp);
    }
    return rc;
}

int pomdp_uninitialize_gpu(POMDP *pomdp)
{
    if (pomdp == nullptr) {
        NOVA_ERR("pomdp_uninitialize_gpu", "Invalid input.");
        return NOVA_ERROR_INVALID_DATA;
    }
    int rc = pomdp_uninitialize_successors_gpu(pomdp);
    rc += pomdp_uninitialize_state_transitions_gpu(pomdp);
    rc += pomdp_uninitialize_observation_transitions_gpu(pomdp);
    rc += pomdp_uninitialize_rewards_gpu(pomdp);
    if (pomdp->r > 0) {
        rc += pomdp_uninitialize_nonzero_beliefs_gpu(pomdp);
        rc += pomdp_uninitialize_belief_points_gpu(pomdp);
    }
    return rc;
}

} // namespace nova
