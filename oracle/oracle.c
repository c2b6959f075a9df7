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
