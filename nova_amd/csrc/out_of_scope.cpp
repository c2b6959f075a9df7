// nova-mi355x: exported entry point outside the Bellman-backup path this library provides.
// The reference's ctypes layer binds every symbol at import,
// so it must exist; it reports that it is not provided and returns NOVA_ERROR_INVALID_DATA.
#include <nova/pomdp/utilities/pomdp_sigma_cpu.h>

#include "nova_common.h"

#define NOVA_NOT_PROVIDED(fn)                                                    \
    do {                                                                         \
        NOVA_ERR(fn, "Not provided by this library (outside the Bellman-backup path)."); \
        return NOVA_ERROR_INVALID_DATA;                                          \
    } while (0)

namespace nova {

int pomdp_sigma_cpu(POMDP *, unsigned int, float &) { NOVA_NOT_PROVIDED("pomdp_sigma_cpu"); }

} // namespace nova
