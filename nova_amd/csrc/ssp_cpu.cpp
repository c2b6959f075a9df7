// nova-mi355x: the stochastic-shortest-path solvers of the drop-in ABI, LAO* and RTDP, on the
// host (the reference has them on the CPU only: src/mdp/algorithms/ssp_lao_star_cpu.cpp,
// src/mdp/algorithms/ssp_rtdp_cpu.cpp). R holds positive costs, goals are absorbing, and the
// backup takes the minimum over actions. Both are sequential searches from s0 whose Bellman
// backups touch one state at a time, so they stay on the host. Compiled with
// -ffp-contract=off and no fast-math so every product and sum rounds as nova's build.
//
// State kept in the ABI structs, as nova keeps it (callers may read it between updates):
//   LAO*: pi[s] == m marks a state never expanded; a negative sign bit on V[s] marks a state
//         not yet visited by the current pass, and the backup reads |V| for that reason.
//   RTDP: pi[s] == m marks a state never expanded.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <nova/mdp/algorithms/ssp_lao_star_cpu.h>
#include <nova/mdp/algorithms/ssp_rtdp_cpu.h>

#include "nova_common.h"

namespace nova_amd {

namespace {

constexpr float kCostCap = 1e10f;   // include/nova/constants.h:33 (FLT_MAX = +1e+10)

// Goal membership for one pass (goals is a short list in nova's files).
std::vector<unsigned char> goal_mask(const nova::MDP *mdp)
{
    std::vector<unsigned char> g(mdp->n, 0);
    for (unsigned i = 0; i < mdp->ng; i++)
        if (mdp->goals[i] < mdp->n)
            g[mdp->goals[i]] = 1;
    return g;
}

// min over actions of R[s,a] + sum_i T * f(V[sp]) along the successor list (stopping at the
// first negative entry); action 0 first, a later one only when strictly cheaper. ABS selects
// f = |.| (LAO*: the sign of V is a visitation mark) or the identity (RTDP).
// ssp_lao_star_cpu.cpp:39-68; ssp_rtdp_cpu.cpp:39-74.
template <bool ABS>
void min_cost_backup(const nova::MDP *mdp, unsigned s, float *V, unsigned *pi)
{
    const unsigned m = mdp->m, ns = mdp->ns;
    float best = kCostCap;
    for (unsigned a = 0; a < m; a++) {
        const int *Sr = mdp->S + idx3(s, a, 0, m, ns);
        const float *Tr = mdp->T + idx3(s, a, 0, m, ns);
        float q = mdp->R[(size_t)s * m + a];
        for (unsigned i = 0; i < ns && Sr[i] >= 0; i++)
            q += Tr[i] * (ABS ? std::fabs(V[Sr[i]]) : V[Sr[i]]);
        if (a == 0 || q < best) {
            best = q;
            pi[s] = a;
        }
    }
    V[s] = best;
}

bool valid_mdp(const nova::MDP *mdp)
{
    return mdp != nullptr && mdp->n != 0 && mdp->ns != 0 && mdp->m != 0 &&
           mdp->S != nullptr && mdp->T != nullptr && mdp->R != nullptr && mdp->horizon >= 1 &&
           mdp->ng >= 1 && mdp->goals != nullptr && mdp->s0 < mdp->n;
}

// Fills policy with the states selected by keep(s): S, V and pi over those states.
template <typename Keep>
int collect_policy(const nova::MDP *mdp, const float *V, const unsigned *pi,
                   nova::MDPValueFunction *policy, const char *fn, Keep keep)
{
    unsigned r = 0;
    for (unsigned s = 0; s < mdp->n; s++)
        r += keep(s) ? 1u : 0u;
    if (nova::mdp_value_function_initialize(policy, mdp->n, mdp->m, r) != NOVA_SUCCESS) {
        NOVA_ERR(fn, "Could not create the policy.");
        return NOVA_ERROR_POLICY_CREATION;
    }
    unsigned k = 0;
    for (unsigned s = 0; s < mdp->n; s++) {
        if (!keep(s))
            continue;
        policy->S[k] = s;
        policy->V[k] = V[s];
        policy->pi[k] = pi[s];
        k++;
    }
    return NOVA_SUCCESS;
}

// ------------------------------------------------------------------------------- LAO*

struct LaoPassResult {
    unsigned expanded = 0;     // states backed up for the first time (expand pass)
    bool tip = false;          // an unexpanded state on the greedy solution (check pass)
    bool changed = false;      // an action changed in the reverse-order backups
    float residual = 0.0f;     // largest | |V'| - |V| | of those backups
};

// One pass over the greedy partial solution from s0 (ssp_lao_star_cpu.cpp:154-237 expand,
// :240-323 check). Every state is first marked unvisited; a depth-first search pops the
// fringe, skips visited states, marks and records the state, stops at goals, and at a never
// expanded state either backs it up (expand) or notes the tip (check); an expanded state
// pushes its greedy successors that are still unvisited. The recorded states are then backed
// up in reverse order. The fringe is a growable stack: nova's is a fixed n entries, which a
// state pushed by several parents can overrun.
LaoPassResult lao_pass(const nova::MDP *mdp, nova::SSPLAOStarCPU *lao, bool check)
{
    const unsigned n = mdp->n, m = mdp->m, ns = mdp->ns;
    float *V = lao->V;
    unsigned *pi = lao->pi;
    const std::vector<unsigned char> goal = goal_mask(mdp);
    for (unsigned s = 0; s < n; s++)
        V[s] = -std::fabs(V[s]);

    LaoPassResult out;
    std::vector<unsigned> fringe{mdp->s0};
    std::vector<unsigned> visited;
    while (!fringe.empty()) {
        const unsigned s = fringe.back();
        fringe.pop_back();
        if (!std::signbit(V[s]))
            continue;
        V[s] = std::fabs(V[s]);
        visited.push_back(s);
        if (goal[s])
            continue;
        if (pi[s] >= m) {
            if (check) {
                out.tip = true;
            } else {
                out.expanded++;
                min_cost_backup<true>(mdp, s, V, pi);
            }
            continue;
        }
        const int *Sr = mdp->S + idx3(s, pi[s], 0, m, ns);
        for (unsigned i = 0; i < ns && Sr[i] >= 0; i++)
            if (std::signbit(V[Sr[i]]))
                fringe.push_back((unsigned)Sr[i]);
    }
    for (auto it = visited.rbegin(); it != visited.rend(); ++it) {
        const unsigned s = *it;
        const float before = V[s];
        const unsigned a = pi[s];
        min_cost_backup<true>(mdp, s, V, pi);
        out.residual = std::max(out.residual, std::fabs(std::fabs(V[s]) - std::fabs(before)));
        out.changed = out.changed || (a != pi[s]);
    }
    lao->currentHorizon++;
    return out;
}

void lao_free(nova::SSPLAOStarCPU *lao)
{
    release_if_owned(lao->V);
    release_if_owned(lao->pi);
}

// ------------------------------------------------------------------------------- RTDP

// Every action a certain self-loop of positive cost (ssp_rtdp_cpu.cpp:107-121).
bool dead_end(const nova::MDP *mdp, unsigned s)
{
    for (unsigned a = 0; a < mdp->m; a++) {
        const size_t row = idx3(s, a, 0, mdp->m, mdp->ns);
        if (!((unsigned)mdp->S[row] == s && mdp->T[row] == 1.0f &&
              mdp->R[(size_t)s * mdp->m + a] > 0.0f))
            return false;
    }
    return true;
}

// Successor of (s, a) for the draw rand()/RAND_MAX: the first entry whose running sum of T
// reaches the draw, the first entry when none does (ssp_rtdp_cpu.cpp:83-104).
unsigned sample_successor(const nova::MDP *mdp, unsigned s, unsigned a)
{
    const size_t row = idx3(s, a, 0, mdp->m, mdp->ns);
    const float draw = (float)std::rand() / (float)RAND_MAX;
    float run = 0.0f;
    for (unsigned i = 0; i < mdp->ns && mdp->S[row + i] >= 0; i++) {
        run += mdp->T[row + i];
        if (run >= draw)
            return (unsigned)mdp->S[row + i];
    }
    return (unsigned)mdp->S[row];
}

void rtdp_free(nova::SSPRTDPCPU *rtdp)
{
    release_if_owned(rtdp->V);
    release_if_owned(rtdp->pi);
}

} // namespace

} // namespace nova_amd

namespace nova {

using namespace nova_amd;

// ---- LAO* (ssp_lao_star_cpu.cpp:326-542)

int ssp_lao_star_initialize_cpu(const MDP *mdp, SSPLAOStarCPU *lao)
{
    if (mdp == nullptr || mdp->n == 0 || lao == nullptr) {
        NOVA_ERR("ssp_lao_star_initialize_cpu", "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    lao_free(lao);   // a second initialize does not leak the first one's arrays
    lao->currentHorizon = 0;
    lao->V = tracked_new<float>(mdp->n);
    lao->pi = tracked_new<unsigned int>(mdp->n);
    for (unsigned s = 0; s < mdp->n; s++) {
        lao->V[s] = lao->VInitial != nullptr ? lao->VInitial[s] : 0.0f;
        lao->pi[s] = 0;
    }
    return NOVA_SUCCESS;
}

// Expand passes while the horizon allows and the last pass expanded something, then check
// passes while the horizon allows and neither convergence (no action changed and the residual
// is below epsilon) nor an unexpanded tip was found. NOVA_CONVERGED when the horizon is used
// up or a check converged without a tip (ssp_lao_star_cpu.cpp:440-481). nova's expand step
// has no return statement, so its update fails or never returns depending on the build; this
// one returns its evident result.
int ssp_lao_star_update_cpu(const MDP *mdp, SSPLAOStarCPU *lao)
{
    if (mdp == nullptr || lao == nullptr || lao->V == nullptr || lao->pi == nullptr ||
            mdp->s0 >= mdp->n) {
        NOVA_ERR("ssp_lao_star_update_cpu", "Invalid arguments (initialize first).");
        return NOVA_ERROR_INVALID_DATA;
    }
    unsigned expanded = 1;
    while (lao->currentHorizon < mdp->horizon && expanded != 0)
        expanded = lao_pass(mdp, lao, false).expanded;
    bool converged = false, tip = false;
    while (lao->currentHorizon < mdp->horizon && !converged && !tip) {
        const LaoPassResult c = lao_pass(mdp, lao, true);
        converged = !c.changed && c.residual < mdp->epsilon;
        tip = c.tip;
    }
    if (lao->currentHorizon >= mdp->horizon || (converged && !tip))
        return NOVA_CONVERGED;
    return NOVA_SUCCESS;
}

// The states visited by the last pass (the final greedy solution from s0) with V and pi.
int ssp_lao_star_get_policy_cpu(const MDP *mdp, SSPLAOStarCPU *lao, MDPValueFunction *policy)
{
    if (mdp == nullptr || lao == nullptr || lao->V == nullptr || lao->pi == nullptr ||
            policy == nullptr) {
        NOVA_ERR("ssp_lao_star_get_policy_cpu", "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    const float *V = lao->V;
    return collect_policy(mdp, lao->V, lao->pi, policy, "ssp_lao_star_get_policy_cpu",
                          [V](unsigned s) { return !std::signbit(V[s]); });
}

int ssp_lao_star_uninitialize_cpu(const MDP *mdp, SSPLAOStarCPU *lao)
{
    if (mdp == nullptr || lao == nullptr) {
        NOVA_ERR("ssp_lao_star_uninitialize_cpu", "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    lao->currentHorizon = 0;
    release(lao->V);      // as nova's uninitialize: a non-null array is the solver's
    release(lao->pi);
    return NOVA_SUCCESS;
}

int ssp_lao_star_execute_cpu(const MDP *mdp, SSPLAOStarCPU *lao, MDPValueFunction *policy)
{
    const char *fn = "ssp_lao_star_execute_cpu";
    if (!valid_mdp(mdp) || mdp->epsilon < 0.0f || lao == nullptr || policy == nullptr) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    int rc = ssp_lao_star_initialize_cpu(mdp, lao);
    if (rc != NOVA_SUCCESS)
        return rc;
    for (unsigned s = 0; s < mdp->n; s++)
        lao->pi[s] = mdp->m;                 // every state starts never expanded
    do {
        rc = ssp_lao_star_update_cpu(mdp, lao);
    } while (rc == NOVA_SUCCESS);
    if (rc == NOVA_CONVERGED)
        rc = ssp_lao_star_get_policy_cpu(mdp, lao, policy);
    else
        NOVA_ERR(fn, "Failed to perform the update step of LAO*.");
    ssp_lao_star_uninitialize_cpu(mdp, lao);
    return rc;
}

// ---- RTDP (ssp_rtdp_cpu.cpp:124-342)

int ssp_rtdp_initialize_cpu(const MDP *mdp, SSPRTDPCPU *rtdp)
{
    if (mdp == nullptr || mdp->n == 0 || rtdp == nullptr) {
        NOVA_ERR("ssp_rtdp_initialize_cpu", "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    rtdp_free(rtdp);
    rtdp->currentTrial = 0;
    rtdp->currentHorizon = 0;
    rtdp->V = tracked_new<float>(mdp->n);
    rtdp->pi = tracked_new<unsigned int>(mdp->n);
    for (unsigned s = 0; s < mdp->n; s++) {
        rtdp->V[s] = rtdp->VInitial != nullptr ? rtdp->VInitial[s] : 0.0f;
        rtdp->pi[s] = mdp->m;
    }
    return NOVA_SUCCESS;
}

// One trial from s0: back up the current state, follow its greedy action to a sampled
// successor, until a goal, a dead end or `horizon` steps; a final goal gets V = 0, a final
// dead end V = 1e10, and the final state is marked expanded with pi = 0
// (ssp_rtdp_cpu.cpp:260-300).
int ssp_rtdp_update_cpu(const MDP *mdp, SSPRTDPCPU *rtdp)
{
    if (mdp == nullptr || rtdp == nullptr || rtdp->V == nullptr || rtdp->pi == nullptr ||
            mdp->s0 >= mdp->n) {
        NOVA_ERR("ssp_rtdp_update_cpu", "Invalid arguments (initialize first).");
        return NOVA_ERROR_INVALID_DATA;
    }
    const std::vector<unsigned char> goal = goal_mask(mdp);
    unsigned s = mdp->s0;
    bool atGoal = goal[s] != 0, atDeadEnd = dead_end(mdp, s);
    rtdp->currentHorizon = 0;
    while (!atGoal && !atDeadEnd && rtdp->currentHorizon < mdp->horizon) {
        min_cost_backup<false>(mdp, s, rtdp->V, rtdp->pi);
        s = sample_successor(mdp, s, rtdp->pi[s]);
        atGoal = goal[s] != 0;
        atDeadEnd = dead_end(mdp, s);
        rtdp->currentHorizon++;
    }
    if (atGoal)
        rtdp->V[s] = 0.0f;
    if (atDeadEnd)
        rtdp->V[s] = kCostCap;
    rtdp->pi[s] = 0;
    return NOVA_SUCCESS;
}

// The expanded states (pi < m) with V and pi.
int ssp_rtdp_get_policy_cpu(const MDP *mdp, SSPRTDPCPU *rtdp, MDPValueFunction *policy)
{
    if (mdp == nullptr || rtdp == nullptr || rtdp->V == nullptr || rtdp->pi == nullptr ||
            policy == nullptr) {
        NOVA_ERR("ssp_rtdp_get_policy_cpu", "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    const unsigned *pi = rtdp->pi;
    const unsigned m = mdp->m;
    return collect_policy(mdp, rtdp->V, rtdp->pi, policy, "ssp_rtdp_get_policy_cpu",
                          [pi, m](unsigned s) { return pi[s] < m; });
}

int ssp_rtdp_uninitialize_cpu(const MDP *mdp, SSPRTDPCPU *rtdp)
{
    if (mdp == nullptr || rtdp == nullptr) {
        NOVA_ERR("ssp_rtdp_uninitialize_cpu", "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    rtdp->currentHorizon = 0;
    rtdp->currentTrial = 0;
    release(rtdp->V);
    release(rtdp->pi);
    return NOVA_SUCCESS;
}

int ssp_rtdp_execute_cpu(const MDP *mdp, SSPRTDPCPU *rtdp, MDPValueFunction *policy)
{
    const char *fn = "ssp_rtdp_execute_cpu";
    if (!valid_mdp(mdp) || rtdp == nullptr || rtdp->trials < 1 || policy == nullptr) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    int rc = ssp_rtdp_initialize_cpu(mdp, rtdp);
    if (rc != NOVA_SUCCESS)
        return rc;
    for (rtdp->currentTrial = 0; rtdp->currentTrial < rtdp->trials && rc == NOVA_SUCCESS;
            rtdp->currentTrial++)
        rc = ssp_rtdp_update_cpu(mdp, rtdp);
    if (rc == NOVA_SUCCESS)
        rc = ssp_rtdp_get_policy_cpu(mdp, rtdp, policy);
    ssp_rtdp_uninitialize_cpu(mdp, rtdp);
    return rc;
}

} // namespace nova
