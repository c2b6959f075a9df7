/*
 * TEST INFRASTRUCTURE ONLY - the oracle. Nothing in the shipped library links or calls this
 * code; only tests/, the smoke check and the benchmark's CPU-baseline leg load it.
 *
 * Plain-C restatement of nova's CPU Bellman-backup path, written once over REAL and included
 * twice by oracle.c: REAL=float reproduces the reference's float arithmetic (same loop order,
 * same evaluation order, strict comparisons, the -1e10/+1e10 sentinels, no FMA contraction);
 * REAL=double is the FP64 twin used to state the error of every float result. Model arrays
 * (T, O, R, B) stay float in both, so the twin solves the same problem in higher precision.
 *
 * Every function cites the reference file:line it follows. Where the reference has undefined
 * behaviour the restatement picks a defined value and says so at that line.
 *
 * How each function is pinned (tests/test_oracle.py):
 *   orc_vi_sweep, orc_vi_run      nova's known answers; bit-for-bit vs nova's build (seeded,
 *                                 ragged); FP64 twin vs a dense numpy Bellman operator
 *   orc_pbvi_backup, orc_pbvi_run bit-for-bit vs nova's build; nova's closed forms; FP64 twin
 *                                 vs a dense enumeration of every argmax
 *   orc_belief_update, orc_prob_obs  bit-for-bit vs nova's build; normalisation invariants
 *   orc_expand_distinct           bit-for-bit vs nova's build
 *   orc_alpha_value               nova's known answers
 *   orc_perseus_step, orc_sparse_vb  driven by nova's loop on the same libc rand() stream,
 *                                 bit-for-bit vs nova's build
 */

#define ORC_SENT_LO ((REAL)-1e10)   /* include/nova/constants.h:34 (FLT_MIN = -1e+10) */
#define ORC_SENT_HI ((REAL)1e10)    /* include/nova/constants.h:33 (FLT_MAX = +1e+10) */

/* One Bellman sweep. src/mdp/algorithms/mdp_vi_cpu.cpp:36-63: Q = R[s,a]; Q += (gamma*T)*V[sp]
 * over the successor list until the first negative entry; action 0 is taken unconditionally,
 * a later action only when strictly better. */
void SFX(orc_vi_sweep)(unsigned n, unsigned ns, unsigned m, float gamma_f,
        const int *S, const float *T, const float *R, const REAL *V, REAL *Vp, unsigned *pi)
{
    const REAL gamma = (REAL)gamma_f;
    for (unsigned s = 0; s < n; s++) {
        REAL best = ORC_SENT_LO;
        for (unsigned a = 0; a < m; a++) {
            const size_t row = ((size_t)s * m + a) * ns;
            REAL q = (REAL)R[(size_t)s * m + a];
            for (unsigned i = 0; i < ns; i++) {
                const int sp = S[row + i];
                if (sp < 0)
                    break;
                q += gamma * (REAL)T[row + i] * V[sp];
            }
            if (a == 0 || q > best) {
                best = q;
                pi[s] = a;
            }
        }
        Vp[s] = best;
    }
}

/* `horizon` sweeps with the ping-pong of mdp_vi_cpu.cpp:184-195 and the read-out of
 * mdp_vi_cpu.cpp:198-223: the returned V is the buffer selected by the parity of the final
 * horizon (i.e. V_h), and pi is the single shared buffer written by the last sweep (the
 * greedy action with respect to V_{h-1}). V0 == NULL means zeros (mdp_vi_cpu.cpp:83-95).
 * work must hold 2n REAL. */
void SFX(orc_vi_run)(unsigned n, unsigned ns, unsigned m, float gamma, unsigned horizon,
        const int *S, const float *T, const float *R, const float *V0,
        REAL *work, REAL *Vout, unsigned *piOut)
{
    REAL *V = work, *Vp = work + n;
    for (unsigned s = 0; s < n; s++) {
        V[s] = V0 ? (REAL)V0[s] : (REAL)0;
        Vp[s] = V[s];
        piOut[s] = 0;
    }
    for (unsigned h = 0; h < horizon; h++) {
        if (h % 2 == 0)
            SFX(orc_vi_sweep)(n, ns, m, gamma, S, T, R, V, Vp, piOut);
        else
            SFX(orc_vi_sweep)(n, ns, m, gamma, S, T, R, Vp, V, piOut);
    }
    const REAL *fin = (horizon % 2 == 0) ? V : Vp;
    for (unsigned s = 0; s < n; s++)
        Vout[s] = fin[s];
}

/* gamma * sum_l (O[a,sp,o] * T[s,a,l]) * G[sp] for one state s, as the inner loops of
 * src/pomdp/algorithms/pomdp_pbvi_cpu.cpp:63-72 and :90-98 evaluate it:
 * Vtk accumulates (O*T)*Gamma left to right, then Vtk *= gamma. */
static REAL SFX(orc_project)(unsigned n, unsigned ns, unsigned m, unsigned z, REAL gamma,
        const int *S, const float *T, const float *O, const REAL *G,
        unsigned s, unsigned a, unsigned o)
{
    const size_t row = ((size_t)s * m + a) * ns;
    REAL v = 0;
    for (unsigned l = 0; l < ns; l++) {
        const int sp = S[row + l];
        if (sp < 0)
            break;
        v += (REAL)O[(size_t)a * n * z + (size_t)sp * z + o] * (REAL)T[row + l] * G[sp];
    }
    v *= gamma;
    return v;
}

/* The PBVI backup of one belief i against the r0 alpha vectors in Gamma (row j at
 * Gamma[j*n]). src/pomdp/algorithms/pomdp_pbvi_cpu.cpp:37-103 (best alpha per observation by
 * strict '>' from -1e10, then materialisation over all states) and :106-162 (alpha = R[:,a]
 * plus the per-observation terms, its value on the belief support, strict '>' over actions
 * from -1e10). The winner goes to out[n] / *aOut; returns 0 when no action beats -1e10 (the
 * reference then leaves the destination untouched).
 * When no alpha scores above -1e10 (or r0 == 0) the reference's index is uninitialised
 * (pomdp_pbvi_cpu.cpp:44, 79-82; pomdp_perseus_cpu.cpp:75-117); the restatement uses 0.
 * Optional diagnostics for this belief (NULL to skip): jstar[a*z+o] the chosen alpha,
 * gap[a*z+o] best score minus the runner-up (1e30 when r0 < 2), aval[a] each action's value.
 * alpha is scratch of n REAL. */
static int SFX(orc_backup_belief)(unsigned n, unsigned ns, unsigned m, unsigned z, unsigned rz,
        REAL gamma, const int *S, const float *T, const float *O, const float *R,
        const int *Z, const float *B, const REAL *Gamma, unsigned r0, unsigned i,
        REAL *out, unsigned *aOut, REAL *alpha, int *jstar, REAL *gap, REAL *aval)
{
    REAL bestValue = ORC_SENT_LO;
    int found = 0;
    for (unsigned a = 0; a < m; a++) {
        for (unsigned s = 0; s < n; s++)
            alpha[s] = (REAL)R[(size_t)s * m + a];
        for (unsigned o = 0; o < z; o++) {
            REAL best = ORC_SENT_LO, second = ORC_SENT_LO;
            unsigned bestj = 0;
            for (unsigned j = 0; j < r0; j++) {
                REAL value = 0;
                for (unsigned k = 0; k < rz; k++) {
                    const int s = Z[(size_t)i * rz + k];
                    if (s < 0)
                        break;
                    REAL Vtk = SFX(orc_project)(n, ns, m, z, gamma, S, T, O,
                                                Gamma + (size_t)j * n, (unsigned)s, a, o);
                    value += Vtk * (REAL)B[(size_t)i * rz + k];
                }
                if (value > best) {
                    second = best;
                    best = value;
                    bestj = j;
                } else if (value > second) {
                    second = value;
                }
            }
            if (jstar)
                jstar[a * z + o] = (int)bestj;
            if (gap)
                gap[a * z + o] = (r0 > 1) ? best - second : (REAL)1e30;
            for (unsigned s = 0; s < n; s++)
                alpha[s] += SFX(orc_project)(n, ns, m, z, gamma, S, T, O,
                                             Gamma + (size_t)bestj * n, s, a, o);
        }
        REAL value = 0;
        for (unsigned k = 0; k < rz; k++) {
            const int s = Z[(size_t)i * rz + k];
            if (s < 0)
                break;
            value += alpha[s] * (REAL)B[(size_t)i * rz + k];
        }
        if (aval)
            aval[a] = value;
        if (value > bestValue) {
            for (unsigned s = 0; s < n; s++)
                out[s] = alpha[s];
            *aOut = a;
            bestValue = value;
            found = 1;
        }
    }
    return found;
}

/* The PBVI backup of beliefs [i0, i1): GammaPrime row i and pi[i] for each. Diagnostics are
 * indexed [(i-i0)*m*z + a*z + o] (jstar, gap) and [(i-i0)*m + a] (aval). alpha is scratch of
 * n REAL; the winner is written straight into GammaPrime. */
void SFX(orc_pbvi_backup)(unsigned n, unsigned ns, unsigned m, unsigned z, unsigned rz,
        float gamma_f, const int *S, const float *T, const float *O, const float *R,
        const int *Z, const float *B, const REAL *Gamma, unsigned r0,
        unsigned i0, unsigned i1, REAL *GammaPrime, unsigned *pi,
        REAL *alpha, int *jstar, REAL *gap, REAL *aval)
{
    const REAL gamma = (REAL)gamma_f;
    for (unsigned i = i0; i < i1; i++) {
        const size_t d = (size_t)(i - i0);
        SFX(orc_backup_belief)(n, ns, m, z, rz, gamma, S, T, O, R, Z, B, Gamma, r0, i,
                               GammaPrime + (size_t)i * n, &pi[i], alpha,
                               jstar ? jstar + d * m * z : 0, gap ? gap + d * m * z : 0,
                               aval ? aval + d * m : 0);
    }
}

/* `horizon` PBVI iterations over all r beliefs with the ping-pong and read-out of
 * pomdp_pbvi_cpu.cpp:165-183 (Gamma = GammaPrime = GammaInitial, pi = 0), :257-273 (even
 * horizon reads Gamma, odd reads GammaPrime) and :276-300 (result buffer by parity, pi shared).
 * work must hold 2*r*n + n REAL. */
void SFX(orc_pbvi_run)(unsigned n, unsigned ns, unsigned m, unsigned z, unsigned r, unsigned rz,
        float gamma, unsigned horizon, const int *S, const float *T, const float *O,
        const float *R, const int *Z, const float *B, const float *GammaInitial,
        REAL *work, REAL *GammaOut, unsigned *piOut)
{
    const size_t rn = (size_t)r * n;
    REAL *G = work, *Gp = work + rn, *alpha = work + 2 * rn;
    for (size_t x = 0; x < rn; x++) {
        G[x] = (REAL)GammaInitial[x];
        Gp[x] = G[x];
    }
    for (unsigned i = 0; i < r; i++)
        piOut[i] = 0;
    for (unsigned h = 0; h < horizon; h++) {
        if (h % 2 == 0)
            SFX(orc_pbvi_backup)(n, ns, m, z, rz, gamma, S, T, O, R, Z, B, G, r, 0, r, Gp,
                                 piOut, alpha, 0, 0, 0);
        else
            SFX(orc_pbvi_backup)(n, ns, m, z, rz, gamma, S, T, O, R, Z, B, Gp, r, 0, r, G,
                                 piOut, alpha, 0, 0, 0);
    }
    const REAL *fin = (horizon % 2 == 0) ? G : Gp;
    for (size_t x = 0; x < rn; x++)
        GammaOut[x] = fin[x];
}

/* tau(b, a, o) on a dense belief. src/pomdp/utilities/pomdp_model_cpu.cpp:89-137:
 * bp[sp] += T*b[s] over the successor lists, bp[sp] *= O[a,sp,o] while summing the normaliser
 * in state order, NOVA_WARNING_INVALID_BELIEF (10) when |normaliser| < 1e-9 (bp then holds the
 * unnormalised product), else bp[sp] /= normaliser. */
int SFX(orc_belief_update)(unsigned n, unsigned ns, unsigned m, unsigned z,
        const int *S, const float *T, const float *O, const REAL *b,
        unsigned a, unsigned o, REAL *bp)
{
    for (unsigned sp = 0; sp < n; sp++)
        bp[sp] = 0;
    for (unsigned s = 0; s < n; s++) {
        const size_t row = ((size_t)s * m + a) * ns;
        for (unsigned l = 0; l < ns; l++) {
            const int sp = S[row + l];
            if (sp < 0)
                break;
            bp[sp] += (REAL)T[row + l] * b[s];
        }
    }
    REAL norm = 0;
    for (unsigned sp = 0; sp < n; sp++) {
        bp[sp] *= (REAL)O[(size_t)a * n * z + (size_t)sp * z + o];
        norm += bp[sp];
    }
    if (fabs((double)norm) < 1e-9)
        return 10;
    for (unsigned sp = 0; sp < n; sp++)
        bp[sp] /= norm;
    return 0;
}

/* Pr(o | b, a) = sum_s b(s) * sum_l T*O, accumulated per state first.
 * src/pomdp/utilities/pomdp_expand_cpu.cpp:57-79. */
REAL SFX(orc_prob_obs)(unsigned n, unsigned ns, unsigned m, unsigned z,
        const int *S, const float *T, const float *O, const REAL *b, unsigned a, unsigned o)
{
    REAL pr = 0;
    for (unsigned s = 0; s < n; s++) {
        const size_t row = ((size_t)s * m + a) * ns;
        REAL val = 0;
        for (unsigned l = 0; l < ns; l++) {
            const int sp = S[row + l];
            if (sp < 0)
                break;
            val += (REAL)T[row + l] * (REAL)O[(size_t)a * n * z + (size_t)sp * z + o];
        }
        pr += val * b[s];
    }
    return pr;
}

/* Dense copy of sparse belief i. src/pomdp/utilities/pomdp_expand_cpu.cpp:39-54. */
void SFX(orc_construct_belief)(unsigned n, unsigned rz, const int *Z, const float *B,
        unsigned i, REAL *b)
{
    for (unsigned s = 0; s < n; s++)
        b[s] = 0;
    for (unsigned k = 0; k < rz; k++) {
        const int s = Z[(size_t)i * rz + k];
        if (s < 0)
            break;
        b[s] = (REAL)B[(size_t)i * rz + k];
    }
}

/* Distinct-beliefs selection. src/pomdp/utilities/pomdp_expand_cpu.cpp:166-266: for each
 * belief i, over (a, o) in order with Pr(o|b,a) != 0 and a valid tau, score tau by
 * min_j ||b_j - tau||_1 (strict '<' from +1e10, sum in state order) and keep the first
 * strictly largest score (from -1e10). Writes the r chosen beliefs densely into Bnew (r x n).
 * When no (a, o) qualifies the reference copies an uninitialised buffer (:178-180, :258);
 * the restatement writes zeros. Returns in bestA/bestO (may be NULL) the chosen pair or -1.
 * work must hold 3n REAL. */
void SFX(orc_expand_distinct)(unsigned n, unsigned ns, unsigned m, unsigned z, unsigned r,
        unsigned rz, const int *S, const float *T, const float *O, const int *Z, const float *B,
        REAL *work, REAL *Bnew, int *bestA, int *bestO)
{
    REAL *b = work, *bp = work + n, *bt = work + 2 * n;
    for (unsigned i = 0; i < r; i++) {
        SFX(orc_construct_belief)(n, rz, Z, B, i, b);
        REAL *out = Bnew + (size_t)i * n;
        for (unsigned s = 0; s < n; s++)
            out[s] = 0;
        REAL starValue = ORC_SENT_LO;
        int sa = -1, so = -1;
        for (unsigned a = 0; a < m; a++) {
            for (unsigned o = 0; o < z; o++) {
                if (SFX(orc_prob_obs)(n, ns, m, z, S, T, O, b, a, o) == 0)
                    continue;
                if (SFX(orc_belief_update)(n, ns, m, z, S, T, O, b, a, o, bp) != 0)
                    continue;
                REAL jStar = ORC_SENT_HI;
                for (unsigned j = 0; j < r; j++) {
                    SFX(orc_construct_belief)(n, rz, Z, B, j, bt);
                    REAL d = 0;
                    for (unsigned s = 0; s < n; s++)
                        d += (REAL)fabs((double)(bt[s] - bp[s]));
                    if (d < jStar)
                        jStar = d;
                }
                if (jStar > starValue) {
                    for (unsigned s = 0; s < n; s++)
                        out[s] = bp[s];
                    starValue = jStar;
                    sa = (int)a;
                    so = (int)o;
                }
            }
        }
        if (bestA) bestA[i] = sa;
        if (bestO) bestO[i] = so;
    }
}

/* V(b) = max_i Gamma_i . b for a dense b, the first maximiser winning from -1e10, and its
 * action. src/pomdp/policies/pomdp_alpha_vectors.cpp:52-79. Returns the argmax index (or -1). */
int SFX(orc_alpha_value)(unsigned n, unsigned r, const REAL *Gamma, const unsigned *pi,
        const REAL *b, REAL *Vb, unsigned *a)
{
    int arg = -1;
    *Vb = ORC_SENT_LO;
    *a = 0;
    for (unsigned i = 0; i < r; i++) {
        REAL v = 0;
        for (unsigned s = 0; s < n; s++)
            v += Gamma[(size_t)i * n + s] * b[s];
        if (*Vb < v) {
            *Vb = v;
            *a = pi[i];
            arg = (int)i;
        }
    }
    return arg;
}

/* max over the rG alphas of b_i . alpha on the sparse support (sum B*alpha in support order),
 * first maximiser from -1e10; index 0 when rG == 0.
 * src/pomdp/algorithms/pomdp_perseus_cpu.cpp:37-68. */
REAL SFX(orc_sparse_vb)(unsigned n, unsigned rz, const int *Z, const float *B, unsigned i,
        const REAL *Gamma, unsigned rG, unsigned *arg)
{
    REAL best = ORC_SENT_LO;
    if (arg) *arg = 0;
    for (unsigned g = 0; g < rG; g++) {
        REAL v = 0;
        for (unsigned k = 0; k < rz; k++) {
            const int s = Z[(size_t)i * rz + k];
            if (s < 0)
                break;
            v += (REAL)B[(size_t)i * rz + k] * Gamma[(size_t)g * n + s];
        }
        if (best < v) {
            best = v;
            if (arg) *arg = g;
        }
    }
    return best;
}

/* One Perseus step for belief bIndex (the caller draws it; the reference draws it with
 * rand(), pomdp_perseus_cpu.cpp:339-341). Backup against the rG alphas of Gamma exactly as
 * the PBVI backup (pomdp_perseus_cpu.cpp:71-180; with rG == 0 the reference's argmax index
 * is uninitialised and the restatement uses row 0 of Gamma), keep the new alpha when
 * b.alpha_new >= V(b) else copy the old maximiser (:352-376), append it to GammaPrime at row
 * *rGp, then rebuild BTilde = {i : V'(b_i) < V(b_i)} (:387-402).
 * Returns 11 (NOVA_ERROR_OUT_OF_MEMORY) when more than r alphas would exist, else 0.
 * alpha must hold 2n REAL. */
int SFX(orc_perseus_step)(unsigned n, unsigned ns, unsigned m, unsigned z, unsigned r,
        unsigned rz, float gamma, const int *S, const float *T, const float *O, const float *R,
        const int *Z, const float *B, const REAL *Gamma, unsigned rG, const unsigned *pi,
        REAL *GammaPrime, unsigned *rGp, unsigned *piPrime, unsigned bIndex,
        unsigned *BTilde, unsigned *rTilde, REAL *alpha)
{
    REAL *alphaBest = alpha + n;
    unsigned aBest = 0;
    for (unsigned s = 0; s < n; s++)
        alphaBest[s] = 0;
    SFX(orc_backup_belief)(n, ns, m, z, rz, (REAL)gamma, S, T, O, R, Z, B, Gamma, rG, bIndex,
                           alphaBest, &aBest, alpha, 0, 0, 0);
    REAL bDotAlpha = 0;
    for (unsigned k = 0; k < rz; k++) {
        const int s = Z[(size_t)bIndex * rz + k];
        if (s < 0)
            break;
        bDotAlpha += (REAL)B[(size_t)bIndex * rz + k] * alphaBest[s];
    }
    unsigned oldArg = 0;
    REAL Vnb = SFX(orc_sparse_vb)(n, rz, Z, B, bIndex, Gamma, rG, &oldArg);
    REAL *dst = GammaPrime + (size_t)(*rGp) * n;
    if (bDotAlpha >= Vnb) {
        for (unsigned s = 0; s < n; s++)
            dst[s] = alphaBest[s];
        piPrime[*rGp] = aBest;
    } else {
        for (unsigned s = 0; s < n; s++)
            dst[s] = Gamma[(size_t)oldArg * n + s];
        piPrime[*rGp] = pi[oldArg];
    }
    (*rGp)++;
    if (*rGp > r)
        return 11;
    *rTilde = 0;
    for (unsigned i = 0; i < r; i++) {
        REAL vOld = SFX(orc_sparse_vb)(n, rz, Z, B, i, Gamma, rG, 0);
        REAL vNew = SFX(orc_sparse_vb)(n, rz, Z, B, i, GammaPrime, *rGp, 0);
        if (vNew < vOld)
            BTilde[(*rTilde)++] = i;
    }
    return 0;
}
