// nova-mi355x: sigma-approximation (replaces include/nova/pomdp/utilities/pomdp_sigma_cpu.h:43).
// Preprocessing of B, outside the backup path: exported for ctypes import and returns
// NOVA_ERROR_INVALID_DATA with a message on stderr.
#ifndef NOVA_POMDP_SIGMA_CPU_H
#define NOVA_POMDP_SIGMA_CPU_H

#include <nova/pomdp/pomdp.h>

namespace nova {

extern "C" int pomdp_sigma_cpu(POMDP *pomdp, unsigned int numDesiredNonZeroValues, float &sigma);

} // namespace nova

#endif
