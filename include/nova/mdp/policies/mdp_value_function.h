// nova-mi355x: MDP value-function output container
// (replaces include/nova/mdp/policies/mdp_value_function.h:48-73).
// x86-64 sizeof 40: n@0 m@4 r@8 S@16 V@24 pi@32. The library allocates V/pi (and S when
// r > 0) with new[]; the caller releases them through mdp_value_function_uninitialize.
#ifndef NOVA_MDP_VALUE_FUNCTION_H
#define NOVA_MDP_VALUE_FUNCTION_H

namespace nova {

typedef struct NovaMDPValueFunction {
    unsigned int n;
    unsigned int m;
    unsigned int r;
    unsigned int *S;
    float *V;
    unsigned int *pi;
} MDPValueFunction;

// Allocates V and pi (length n when r == 0, else length r, plus S of length r).
// Rejects a null policy, non-null arrays, n == 0 or m == 0 (mdp_value_function.cpp:36-40).
extern "C" int mdp_value_function_initialize(MDPValueFunction *policy,
        unsigned int n, unsigned int m, unsigned int r);

// Frees the arrays and zeroes the sizes (mdp_value_function.cpp:60-87).
extern "C" int mdp_value_function_uninitialize(MDPValueFunction *policy);

} // namespace nova

#endif
