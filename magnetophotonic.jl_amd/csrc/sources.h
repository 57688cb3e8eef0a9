// Step-driver side of the source injection (sources.hip): the soft and plane sources read the
// field from one buffer set and write it to another (D in place), for the ping-pong step.
// With in == out they are mp_inject_soft_3d / mp_inject_plane_3d.
#pragma once
#include "common.h"

namespace mp {
int inject_soft_pp(const mp_yee *in, const mp_yee *out, int32_t component, int64_t i, int64_t j, int64_t k,
                   double value);
int inject_plane_pp(const mp_yee *in, const mp_yee *out, int32_t component, int32_t axis, int64_t i0, double w0,
                    int64_t i1, double w1, double value);
}  // namespace mp
