/*
 * TEST INFRASTRUCTURE ONLY - the oracle (see oracle_impl.h). Built into oracle/liboracle.so by
 * oracle/Makefile; loaded by tests/, the smoke check and the benchmark's CPU-baseline leg,
 * never by the shipped library.
 *
 * Compile with -ffp-contract=off: the float twin must round every product and sum the way
 * the reference's plain x86-64 build does.
 */
#include <math.h>
#include <stddef.h>
#include <string.h>

#define REAL float
#define SFX(x) x##_f32
#include "oracle_impl.h"
#undef REAL
#undef SFX
#undef ORC_SENT_LO
#undef ORC_SENT_HI

#define REAL double
#define SFX(x) x##_f64
#include "oracle_impl.h"
#undef REAL
#undef SFX

/* ---- FP64-only references for the differentiable solve (no reference counterpart) ----
 * The reference has no residual-stopped VI and no gradients; these follow the mathematics:
 * V* = max_a R + gamma*P_a V*; at the greedy policy pi, L = rho.V* has
 * dL/dR[s,a] = lambda(s)[a = pi(s)] and dL/dT[s,a,i] = gamma*lambda(s)*V*(S[s,a,i])[a = pi(s)]
 * with (I - gamma P_pi)^T lambda = rho. tests/test_oracle.py pins them against exact policy
 * iteration, a dense solve of the adjoint system and central differences of the exact V*. */

/* Sweeps V <- max_a(R + gamma*T.V) in FP64 (Jacobi, one buffer pair, same tie rule as the
 * sweep) until max|V' - V| <= tol or maxIter sweeps. Returns the sweeps done; pi is the greedy
 * action of the last sweep; V holds the last iterate. work: n doubles. */
unsigned orc_vi_solve_f64(unsigned n, unsigned ns, unsigned m, float gamma,
        const int *S, const float *T, const float *R, double *V, double *work,
        unsigned *pi, unsigned maxIter, double tol, double *resid)
{
    unsigned it = 0;
    double res = 0;
    while (it < maxIter) {
        orc_vi_sweep_f64(n, ns, m, gamma, S, T, R, V, work, pi);
        res = 0;
        for (unsigned s = 0; s < n; s++) {
            double d = fabs(work[s] - V[s]);
            if (d > res) res = d;
            V[s] = work[s];
        }
        it++;
        if (res <= tol)
            break;
    }
    if (resid) *resid = res;
    return it;
}

/* lambda = rho + gamma * P_pi^T lambda by fixed-point iteration (a gamma-contraction in the
 * weighted-by-row-sum norm for stochastic T). Returns iterations; work: n doubles. */
unsigned orc_adjoint_f64(unsigned n, unsigned ns, unsigned m, float gamma,
        const int *S, const float *T, const unsigned *pi, const double *rho,
        double *lambda, double *work, unsigned maxIter, double tol)
{
    const double g = (double)gamma;
    for (unsigned s = 0; s < n; s++)
        lambda[s] = rho[s];
    unsigned it = 0;
    while (it < maxIter) {
        for (unsigned s = 0; s < n; s++)
            work[s] = rho[s];
        for (unsigned s = 0; s < n; s++) {
            const size_t row = ((size_t)s * m + pi[s]) * ns;
            for (unsigned i = 0; i < ns; i++) {
                const int sp = S[row + i];
                if (sp < 0)
                    break;
                work[sp] += g * (double)T[row + i] * lambda[s];
            }
        }
        double res = 0;
        for (unsigned s = 0; s < n; s++) {
            double d = fabs(work[s] - lambda[s]);
            if (d > res) res = d;
            lambda[s] = work[s];
        }
        it++;
        if (res <= tol)
            break;
    }
    return it;
}

/* dL/dR (n*m) and dL/dT (n*m*ns) from lambda, V and the greedy policy. */
void orc_vi_grads_f64(unsigned n, unsigned ns, unsigned m, float gamma, const int *S,
        const unsigned *pi, const double *lambda, const double *V, double *dR, double *dT)
{
    const double g = (double)gamma;
    memset(dR, 0, sizeof(double) * (size_t)n * m);
    memset(dT, 0, sizeof(double) * (size_t)n * m * ns);
    for (unsigned s = 0; s < n; s++) {
        const unsigned a = pi[s];
        dR[(size_t)s * m + a] = lambda[s];
        const size_t row = ((size_t)s * m + a) * ns;
        for (unsigned i = 0; i < ns; i++) {
            const int sp = S[row + i];
            if (sp < 0)
                break;
            dT[row + i] = g * lambda[s] * V[sp];
        }
    }
}

/* ---- Stochastic shortest paths (float, as the reference; no FP64 twin) ----
 * R holds positive costs; goals are absorbing. Both solvers keep V and pi per state, as the
 * reference's SSPLAOStarCPU / SSPRTDPCPU structs expose them.
 * How each is pinned (tests/test_ssp.py):
 *   orc_ssp_rtdp      bit-for-bit vs nova's build on the same libc rand() stream (grid worlds,
 *                     random SSPs with and without dead ends); converges to V*(s0)
 *   orc_ssp_lao_star  nova's build of it cannot run (see below): a closed form; the exact
 *                     shortest path of a deterministic maze; optimal values and actions on its
 *                     solution graph vs orc_ssp_vi_f64 with a tight epsilon
 *   orc_ssp_vi_f64    a closed form; its greedy policy's linear system solved densely;
 *                     Bellman optimality of every action */

#define ORC_SSP_HI 1e10f   /* include/nova/constants.h:33 (FLT_MAX = +1e+10) */

static int orc_ssp_is_goal(unsigned ng, const unsigned *goals, unsigned s)
{
    for (unsigned i = 0; i < ng; i++)
        if (goals[i] == s)
            return 1;
    return 0;
}

/* Q(s,a) = R[s,a] + sum_i T * |V[sp]| (LAO*, the sign of V marks visitation) or + T * V[sp]
 * (RTDP) until the first negative successor; V(s) = min over a, action 0 taken first and a
 * later one only when strictly smaller. src/mdp/algorithms/ssp_lao_star_cpu.cpp:39-68;
 * src/mdp/algorithms/ssp_rtdp_cpu.cpp:39-74. */
static void orc_ssp_backup(unsigned ns, unsigned m, const int *S, const float *T,
        const float *R, unsigned s, int useAbs, float *V, unsigned *pi)
{
    float best = ORC_SSP_HI;
    for (unsigned a = 0; a < m; a++) {
        const size_t row = ((size_t)s * m + a) * ns;
        float q = R[(size_t)s * m + a];
        for (unsigned i = 0; i < ns; i++) {
            const int sp = S[row + i];
            if (sp < 0)
                break;
            q += T[row + i] * (useAbs ? fabsf(V[sp]) : V[sp]);
        }
        if (a == 0 || q < best) {
            best = q;
            pi[s] = a;
        }
    }
    V[s] = best;
}

/* One LAO* pass from s0 (ssp_lao_star_cpu.cpp:154-237 with check = 0, :240-323 with check = 1):
 * every V is first marked unvisited (negative sign); a depth-first search over the greedy
 * partial solution pops the fringe stack, skips visited states, marks and records the state,
 * stops at goals, and at an unexpanded state (pi == m) either backs it up and counts it
 * (expand) or notes the tip (check); expanded states push their greedy successors that are
 * unvisited. The recorded states are then backed up in reverse order; check also tracks the
 * largest | |V'| - |V| | and whether any action changed. The fringe stack grows as needed
 * (the reference's is sized n and can overflow when several parents push one state; inputs
 * that stay within n behave identically). Returns the number of newly expanded states
 * (expand) or sets *converged / *tip (check). stack: room for n*ns + 1 entries; order: n. */
static unsigned orc_lao_pass(unsigned n, unsigned ns, unsigned m, const int *S, const float *T,
        const float *R, unsigned s0, unsigned ng, const unsigned *goals, float epsilon,
        float *V, unsigned *pi, unsigned *stack, unsigned *order, int check,
        int *converged, int *tip)
{
    for (unsigned s = 0; s < n; s++)
        V[s] = -fabsf(V[s]);
    size_t top = 0;
    unsigned cnt = 0, fresh = 0;
    stack[top++] = s0;
    int tipFound = 0;
    while (top > 0) {
        const unsigned s = stack[--top];
        if (!signbit(V[s]))
            continue;
        V[s] = fabsf(V[s]);
        order[cnt++] = s;
        if (orc_ssp_is_goal(ng, goals, s))
            continue;
        if (pi[s] >= m) {
            if (check) {
                tipFound = 1;
            } else {
                fresh++;
                orc_ssp_backup(ns, m, S, T, R, s, 1, V, pi);
            }
        } else {
            const size_t row = ((size_t)s * m + pi[s]) * ns;
            for (unsigned i = 0; i < ns; i++) {
                const int sp = S[row + i];
                if (sp < 0)
                    break;
                if (signbit(V[sp]))
                    stack[top++] = (unsigned)sp;
            }
        }
    }
    int changed = 0;
    float residual = 0.0f;
    while (cnt > 0) {
        const unsigned s = order[--cnt];
        const float before = V[s];
        const unsigned a = pi[s];
        orc_ssp_backup(ns, m, S, T, R, s, 1, V, pi);
        const float d = fabsf(fabsf(V[s]) - fabsf(before));
        if (d > residual)
            residual = d;
        if (a != pi[s])
            changed = 1;
    }
    if (check) {
        *converged = (!changed && residual < epsilon);
        *tip = tipFound;
    }
    return fresh;
}

/* LAO* to convergence (ssp_lao_star_cpu.cpp:326-481): V from V0 (zeros when NULL), pi = m
 * ("never expanded"); then update rounds until converged: expand passes while the horizon
 * allows and the last pass expanded something, then check passes while the horizon allows
 * and neither convergence nor an unexpanded tip was found; done when the horizon is used up
 * or a check converged without a tip. Each pass counts one step of the horizon. The
 * reference's expand step (ssp_lao_star_cpu.cpp:154-237) falls off its end without a return
 * value, so its update fails or never returns depending on the build; this restatement takes
 * the evident intent (success). Returns the passes used. */
unsigned orc_ssp_lao_star(unsigned n, unsigned ns, unsigned m, const int *S, const float *T,
        const float *R, unsigned s0, unsigned ng, const unsigned *goals, unsigned horizon,
        float epsilon, const float *V0, float *V, unsigned *pi, unsigned *stack,
        unsigned *order)
{
    for (unsigned s = 0; s < n; s++) {
        V[s] = V0 ? V0[s] : 0.0f;
        pi[s] = m;
    }
    unsigned h = 0;
    for (;;) {
        unsigned fresh = 1;
        while (h < horizon && fresh != 0) {
            fresh = orc_lao_pass(n, ns, m, S, T, R, s0, ng, goals, epsilon, V, pi, stack,
                                 order, 0, NULL, NULL);
            h++;
        }
        int converged = 0, tip = 0;
        while (h < horizon && !converged && !tip) {
            orc_lao_pass(n, ns, m, S, T, R, s0, ng, goals, epsilon, V, pi, stack, order, 1,
                         &converged, &tip);
            h++;
        }
        if (h >= horizon || (converged && !tip))
            return h;
    }
}

/* RTDP (src/mdp/algorithms/ssp_rtdp_cpu.cpp:124-300): V from V0 (zeros when NULL), pi = m;
 * `trials` trials from s0, each backing up the current state, taking its greedy action and
 * sampling the successor with rnd() / rnd_max against the running sum of T (the first entry
 * when no sum reaches the draw), until a goal, a dead end (every action a certain self-loop
 * of positive cost) or `horizon` steps; a final goal gets V = 0, a final dead end V = 1e10,
 * and the final state is marked expanded with pi = 0. rnd is the libc rand() the reference
 * calls, so a seeded srand() reproduces its trials. */
typedef int (*orc_rand_fn)(void);

static int orc_ssp_dead_end(unsigned ns, unsigned m, const int *S, const float *T,
        const float *R, unsigned s)
{
    for (unsigned a = 0; a < m; a++) {
        const size_t row = ((size_t)s * m + a) * ns;
        if (!((unsigned)S[row] == s && T[row] == 1.0f && R[(size_t)s * m + a] > 0.0f))
            return 0;
    }
    return 1;
}

void orc_ssp_rtdp(unsigned n, unsigned ns, unsigned m, const int *S, const float *T,
        const float *R, unsigned s0, unsigned ng, const unsigned *goals, unsigned horizon,
        unsigned trials, const float *V0, orc_rand_fn rnd, int rnd_max, float *V,
        unsigned *pi)
{
    for (unsigned s = 0; s < n; s++) {
        V[s] = V0 ? V0[s] : 0.0f;
        pi[s] = m;
    }
    for (unsigned t = 0; t < trials; t++) {
        unsigned s = s0;
        int goal = orc_ssp_is_goal(ng, goals, s);
        int dead = orc_ssp_dead_end(ns, m, S, T, R, s);
        for (unsigned h = 0; !goal && !dead && h < horizon; h++) {
            orc_ssp_backup(ns, m, S, T, R, s, 0, V, pi);
            const size_t row = ((size_t)s * m + pi[s]) * ns;
            unsigned next = (unsigned)S[row];
            const float draw = (float)rnd() / (float)rnd_max;
            float run = 0.0f;
            for (unsigned i = 0; i < ns; i++) {
                if (S[row + i] < 0)
                    break;
                run += T[row + i];
                if (run >= draw) {
                    next = (unsigned)S[row + i];
                    break;
                }
            }
            s = next;
            goal = orc_ssp_is_goal(ng, goals, s);
            dead = orc_ssp_dead_end(ns, m, S, T, R, s);
        }
        if (goal)
            V[s] = 0.0f;
        if (dead)
            V[s] = ORC_SSP_HI;
        pi[s] = 0;
    }
}

/* Jacobi min-cost backups in float (no reference counterpart as a sweep; each backup is
 * src/mdp/algorithms/ssp_lao_star_cpu.cpp:39-68 with useAbs, ssp_rtdp_cpu.cpp:39-74 without,
 * and is pinned bit for bit against those functions of nova's build): for the k listed states
 * (every state when states is NULL), Vout[s] and pi[s] from the backup of V; with goals given,
 * a goal state gets Vout = 0 and pi = 0 instead. V is only read, so the order of the list does
 * not matter. */
void orc_ssp_jacobi_f32(unsigned n, unsigned ns, unsigned m, const int *S, const float *T,
        const float *R, unsigned ng, const unsigned *goals, int useAbs, const unsigned *states,
        unsigned k, const float *V, float *Vout, unsigned *pi)
{
    const unsigned count = states ? k : n;
    for (unsigned x = 0; x < count; x++) {
        const unsigned s = states ? states[x] : x;
        if (goals && orc_ssp_is_goal(ng, goals, s)) {
            Vout[s] = 0.0f;
            pi[s] = 0;
            continue;
        }
        float best = ORC_SSP_HI;
        unsigned bestA = 0;
        for (unsigned a = 0; a < m; a++) {
            const size_t row = ((size_t)s * m + a) * ns;
            float q = R[(size_t)s * m + a];
            for (unsigned i = 0; i < ns; i++) {
                const int sp = S[row + i];
                if (sp < 0)
                    break;
                q += T[row + i] * (useAbs ? fabsf(V[sp]) : V[sp]);
            }
            if (a == 0 || q < best) {
                best = q;
                bestA = a;
            }
        }
        Vout[s] = best;
        pi[s] = bestA;
    }
}

/* Optimal SSP values in FP64 (no reference counterpart; pins the two solvers above):
 * V(g) = 0 at goals, V(s) = min_a R[s,a] + sum_i T * V(S[s,a,i]) elsewhere, by Jacobi sweeps
 * from V = 0 (a lower bound for positive costs, so the sweeps rise monotonically to V*) until
 * max |V' - V| <= tol or maxIter sweeps. pi is the greedy action (first minimiser). Returns
 * the sweeps done; *resid the last change. work: n doubles. */
unsigned orc_ssp_vi_f64(unsigned n, unsigned ns, unsigned m, const int *S, const float *T,
        const float *R, unsigned ng, const unsigned *goals, double *V, double *work,
        unsigned *pi, unsigned maxIter, double tol, double *resid)
{
    for (unsigned s = 0; s < n; s++)
        V[s] = 0.0;
    unsigned it = 0;
    double res = 0.0;
    while (it < maxIter) {
        res = 0.0;
        for (unsigned s = 0; s < n; s++) {
            if (orc_ssp_is_goal(ng, goals, s)) {
                work[s] = 0.0;
                pi[s] = 0;
                continue;
            }
            double best = 0.0;
            for (unsigned a = 0; a < m; a++) {
                const size_t row = ((size_t)s * m + a) * ns;
                double q = (double)R[(size_t)s * m + a];
                for (unsigned i = 0; i < ns; i++) {
                    const int sp = S[row + i];
                    if (sp < 0)
                        break;
                    q += (double)T[row + i] * V[sp];
                }
                if (a == 0 || q < best) {
                    best = q;
                    pi[s] = a;
                }
            }
            work[s] = best;
        }
        for (unsigned s = 0; s < n; s++) {
            const double d = fabs(work[s] - V[s]);
            if (d > res)
                res = d;
            V[s] = work[s];
        }
        it++;
        if (res <= tol)
            break;
    }
    if (resid)
        *resid = res;
    return it;
}
