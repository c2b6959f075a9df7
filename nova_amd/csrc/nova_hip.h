// nova-mi355x: HIP runtime helpers shared by the device-side translation units.
#ifndef NOVA_AMD_HIP_H
#define NOVA_AMD_HIP_H

#include <mutex>

#include <hip/hip_runtime.h>

#include "nova_common.h"

namespace nova_amd {

// Word 3 of a raw buffer resource (32-bit dword data, no format conversion) for
// __builtin_amdgcn_make_buffer_rsrc: buffer loads take a 32-bit byte offset in one register.
constexpr int kRawBufferWord3 = 0x00020000;

// The library's stream on the current device: created on first use with the default
// priority and as a blocking stream, so it orders after work on the legacy null stream (the
// stream PyTorch and plain HIP callers use by default). Every ABI call is synchronous at
// return: it synchronises this stream before returning.
hipStream_t stream();

// A non-blocking stream of the current device on which the library only captures graphs (it
// never executes work): a capture there cannot be invalidated by other threads' work on the
// legacy null stream, with which the library stream synchronises. The captured graphs are
// launched on stream().
hipStream_t capture_stream();

// Serialises the library's GPU entry points, which share per-process state: the library
// stream, cached layouts, graphs, workspaces and solver records. Recursive, because entry
// points call one another (execute runs initialize, update and get-policy).
std::recursive_mutex &api_mutex();
#define NOVA_AMD_API_LOCK() \
    const std::lock_guard<std::recursive_mutex> novaAmdApiLock_(nova_amd::api_mutex())

// Reports a HIP failure as "Error[fn]: msg (hip error text)" and returns code, or returns
// NOVA_SUCCESS when e == hipSuccess.
int hip_fail(hipError_t e, const char *fn, const char *msg, int code);

// hipMalloc + registry tracking; device pointers allocated here can be released again by a
// repeated initialize (see registry_track).
int device_alloc(void **p, size_t bytes, const char *fn, const char *what);
// hipFree of a pointer when this library allocated it (or unconditionally when force).
int device_free(void *p, const char *fn, const char *what);
int device_free_if_owned(void *p, const char *fn, const char *what);

// Process-wide counts of the model-array uploads and releases made through this library
// (the *_initialize_*_gpu / *_uninitialize_*_gpu of MDP and POMDP arrays): of the transition
// arrays S, T and O only (transitionsOnly), or of any array. Data derived from device-resident
// model arrays and cached by address compares them, so that arrays re-uploaded at a reused
// address are not mistaken for the ones it was derived from.
unsigned long long model_generation(bool transitionsOnly);
void model_changed(bool transitions);

// Checks that every successor index in the valid prefix of each list lies in [0, limit):
// the kernels gather V[sp]/Gamma[sp] without further checks. Returns the first bad flat index
// or -1.
long long first_bad_successor(const int *S, size_t rows, unsigned len, unsigned limit);

} // namespace nova_amd

#endif
