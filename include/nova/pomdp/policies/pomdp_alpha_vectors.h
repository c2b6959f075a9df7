// nova-mi355x: POMDP alpha-vector policy container
// (replaces include/nova/pomdp/policies/pomdp_alpha_vectors.h:39-74).
// x86-64 sizeof 32: n@0 m@4 r@8 Gamma@16 pi@24. Gamma is r x n row-major (row i at Gamma[i*n]).
#ifndef NOVA_POMDP_ALPHA_VECTORS_H
#define NOVA_POMDP_ALPHA_VECTORS_H

namespace nova {

typedef struct NovaPOMDPAlphaVectors {
    unsigned int n;
    unsigned int m;
    unsigned int r;
    float *Gamma;
    unsigned int *pi;
} POMDPAlphaVectors;

// Allocates Gamma (r*n) and pi (r); rejects non-null arrays or zero sizes
// (pomdp_alpha_vectors.cpp:34-49).
extern "C" int pomdp_alpha_vectors_initialize(POMDPAlphaVectors *policy,
        unsigned int n, unsigned int m, unsigned int r);

// Vb = max_i Gamma_i . b over a dense belief b[n] (first maximiser wins, start at -1e10),
// a = pi[argmax] (pomdp_alpha_vectors.cpp:52-79).
extern "C" int pomdp_alpha_vectors_value_and_action(const POMDPAlphaVectors *policy,
        const float *b, float &Vb, unsigned int &a);

extern "C" int pomdp_alpha_vectors_uninitialize(POMDPAlphaVectors *policy);

} // namespace nova

#endif
