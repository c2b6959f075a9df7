// nova-mi355x: the POMDP model container (replaces include/nova/pomdp/pomdp.h:61-85).
//
// Layout is an interface fact: x86-64 sizeof 128, offsets n@0 ns@4 m@8 z@12 r@16 rz@20
// gamma@24 horizon@28 S@32 T@40 O@48 R@56 Z@64 B@72 d_S@80 d_T@88 d_O@96 d_R@104
// d_Z@112 d_B@120.
//
//   S, T           as in the MDP: S[s*m*ns + a*ns + l], -1 ends a successor list.
//   O[a*n*z + sp*z + o]   Pr(o | a, sp).
//   R[s*m + a]     reward.
//   Z[i*rz + k]    k-th state in the support of belief i (-1 ends the list).
//   B[i*rz + k]    b_i(Z[i*rz + k]).
#ifndef NOVA_POMDP_H
#define NOVA_POMDP_H

namespace nova {

typedef struct NovaPOMDP {
    unsigned int n;
    unsigned int ns;
    unsigned int m;
    unsigned int z;
    unsigned int r;
    unsigned int rz;

    float gamma;
    unsigned int horizon;

    int *S;
    float *T;
    float *O;
    float *R;
    int *Z;
    float *B;

    int *d_S;
    float *d_T;
    float *d_O;
    float *d_R;
    int *d_Z;
    float *d_B;
} POMDP;

} // namespace nova

#endif // NOVA_POMDP_H
