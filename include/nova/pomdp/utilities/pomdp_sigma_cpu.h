// nova-mi355x: sigma-approximation (replaces include/nova/pomdp/utilities/pomdp_sigma_cpu.h:43):
// keeps each belief's numDesiredNonZeroValues largest entries, renormalised, and reports the
// smallest kept mass as sigma. Host arrays only; re-run pomdp_initialize_gpu afterwards.
#ifndef NOVA_POMDP_SIGMA_CPU_H
#define NOVA_POMDP_SIGMA_CPU_H

#include <nova/pomdp/pomdp.h>

namespace nova {

extern "C" int pomdp_sigma_cpu(POMDP *pomdp, unsigned int numDesiredNonZeroValues, float &sigma);

} // namespace nova

#endif
