// nova-mi355x: shared helpers of the host-side library code (no device code here).
#ifndef NOVA_AMD_COMMON_H
#define NOVA_AMD_COMMON_H

#include <cstddef>
#include <cstdio>

#include <nova/constants.h>
#include <nova/error_codes.h>

// Every failing entry point prints "Error[function]: message" on stderr and returns a code,
// the convention of the reference (e.g. src/mdp/algorithms/mdp_vi_cpu.cpp:69).
#define NOVA_ERR(fn, msg) std::fprintf(stderr, "Error[%s]: %s\n", fn, msg)

namespace nova {
struct NovaPOMDP;
typedef NovaPOMDP POMDP;
}

namespace nova_amd {

// Sentinels as float, the type the reference compares against (include/nova/constants.h:33-35).
constexpr float kLow = (float)NOVA_FLT_MIN;
constexpr float kHigh = (float)NOVA_FLT_MAX;

// Buffers this library allocated for the caller-owned solver structs. initialize() may be
// called on a struct whose pointers are uninitialised (the reference's own C++ tests do this)
// or on one still holding buffers from an earlier initialize() (its Python wrapper does this,
// src/mdp/algorithms/mdp_vi_cpu.cpp:112 then leaks them). Only pointers recorded here are
// released by a repeated initialize().
void registry_track(const void *p);
bool registry_untrack(const void *p);   // true when p was tracked (and is now forgotten)

// new[] that records the allocation.
template <typename T> T *tracked_new(size_t count)
{
    T *p = new T[count]();
    registry_track(p);
    return p;
}

// delete[] of a pointer this library allocated; garbage or foreign pointers are left alone.
template <typename T> void release_if_owned(T *&p)
{
    if (p != nullptr && registry_untrack(p))
        delete[] p;
    p = nullptr;
}

// delete[] as the reference's uninitialize does (non-null means owned), forgetting it.
template <typename T> void release(T *&p)
{
    if (p != nullptr) {
        registry_untrack(p);
        delete[] p;
    }
    p = nullptr;
}

// pomdp_belief_update_cpu without its stderr line for an impossible observation when quiet
// (same results and codes); the expansions skip such observations.
int belief_update_host(const nova::POMDP *pomdp, const float *b, unsigned int a, unsigned int o,
                       float *&bp, bool quiet);

inline size_t idx3(unsigned s, unsigned a, unsigned i, unsigned m, unsigned ns)
{
    return ((size_t)s * m + a) * ns + i;
}

} // namespace nova_amd

#endif
