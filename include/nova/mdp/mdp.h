// nova-mi355x: the MDP model container (replaces include/nova/mdp/mdp.h:53-75).
//
// Layout is an interface fact and must not change: x86-64 sizeof 96, offsets
// n@0 ns@4 m@8 gamma@12 horizon@16 epsilon@20 s0@24 ng@28 goals@32 S@40 T@48 R@56
// d_goals@64 d_S@72 d_T@80 d_R@88.
//
//   S[s*m*ns + a*ns + i]  successor state (int32); the first negative entry ends the list.
//   T[s*m*ns + a*ns + i]  Pr(S[..] | s, a) (float32), same index as S.
//   R[s*m + a]            reward (float32).
//   d_*                   device (HBM) mirrors, owned by mdp_initialize_*_gpu.
#ifndef NOVA_MDP_H
#define NOVA_MDP_H

namespace nova {

typedef struct NovaMDP {
    unsigned int n;
    unsigned int ns;
    unsigned int m;

    float gamma;
    unsigned int horizon;
    float epsilon;

    unsigned int s0;
    unsigned int ng;
    unsigned int *goals;

    int *S;
    float *T;
    float *R;

    unsigned int *d_goals;

    int *d_S;
    float *d_T;
    float *d_R;
} MDP;

} // namespace nova

#endif // NOVA_MDP_H
