// Shared host-side helpers of libchimera_hip: status codes, last-error text,
// HIP error checking and argument validation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/chimera_hip.h"

namespace chimera {

void set_error(const char *fmt, ...);

#define CH_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::chimera::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #call,           \
                                 hipGetErrorString(e_));                                \
            return (e_ == hipErrorOutOfMemory) ? CHIMERA_ENOMEM : CHIMERA_EHIP;         \
        }                                                                               \
    } while (0)

#define CH_REQUIRE(cond, ...)                                                           \
    do {                                                                                \
        if (!(cond)) {                                                                  \
            ::chimera::set_error(__VA_ARGS__);                                          \
            return CHIMERA_EINVAL;                                                      \
        }                                                                               \
    } while (0)

#define CH_TRY(expr)                                                                    \
    do {                                                                                \
        int rc_ = (expr);                                                               \
        if (rc_ != CHIMERA_OK) return rc_;                                              \
    } while (0)

inline bool bc_flag_ok(int f) { return f >= 1 && f <= 4; }

// Per-axis periodicity.  Only the periodic / non-periodic distinction changes the
// stage update (fixed and inflow fall back to zero-gradient without an ambient state,
// src/mhd_bc.f90:62-67, 76-81, 92-97, 106-111); a mixed periodic / non-periodic pair
// on one axis is rejected because the halo of a wrapped tile would be ambiguous.
int check_bc(const chimera_bc_t *bc, int *px, int *py);
int check_scheme(const chimera_scheme_t *s);

// Launch-error check after a kernel launch.
inline int launch_check(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("launch of %s failed: %s", what, hipGetErrorString(e));
        return CHIMERA_EHIP;
    }
    return CHIMERA_OK;
}

}  // namespace chimera
