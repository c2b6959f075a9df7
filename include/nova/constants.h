// nova-mi355x: numeric sentinels shared by every solver.
//
// Replaces include/nova/constants.h:33-35 of the reference. The reference defines the
// macros FLT_MAX = 1e+10, FLT_MIN = -1e+10 and FLT_ERR_TOL = 1e-9, which shadow <cfloat>.
// Here the values live under NOVA_-prefixed names; the legacy spellings are available for
// callers that rely on them by defining NOVA_LEGACY_FLT_MACROS before including this file.
#ifndef NOVA_CONSTANTS_H
#define NOVA_CONSTANTS_H

#define NOVA_FLT_MAX      1e+10
#define NOVA_FLT_MIN      -1e+10
#define NOVA_FLT_ERR_TOL  1e-9

#ifdef NOVA_LEGACY_FLT_MACROS
#undef FLT_MAX
#undef FLT_MIN
#define FLT_MAX NOVA_FLT_MAX
#define FLT_MIN NOVA_FLT_MIN
#define FLT_ERR_TOL NOVA_FLT_ERR_TOL
#endif

#endif // NOVA_CONSTANTS_H
