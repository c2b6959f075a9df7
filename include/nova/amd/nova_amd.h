// nova-mi355x: entry points added beside nova's ABI. None of them has a counterpart in the
// reference; they do not change any of its 75 signatures.
#ifndef NOVA_AMD_H
#define NOVA_AMD_H

#include <nova/mdp/mdp.h>
#include <nova/mdp/algorithms/mdp_vi_gpu.h>

namespace nova {

// Runs `sweeps` Bellman sweeps on the device continuing from vi's state (the V/VPrime
// ping-pong by parity of currentHorizon, exactly as `sweeps` calls of mdp_vi_update_gpu), as
// one captured HIP graph on the library stream, then synchronises. When d_V0 is non-null it
// is first copied (device to device, n floats) into d_V and d_VPrime and currentHorizon is
// reset to 0. elapsedMs (nullable) receives the device time of the sweeps alone, measured
// with HIP events on the stream the sweeps run on.
extern "C" int mdp_vi_run_gpu(const MDP *mdp, MDPVIGPU *vi, const float *d_V0,
        unsigned int sweeps, float *elapsedMs);

// Library build facts for callers and tests: writes up to len-1 characters of a
// NUL-terminated description (target architecture, kernel variants) into buf.
extern "C" int nova_amd_build_info(char *buf, unsigned int len);

} // namespace nova

#endif
