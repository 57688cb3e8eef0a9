// Step-driver side of the Yee updates (yee.hip): the ping-pong step, in which one buffer
// set of E and H is read and the other written, so the interior H and E updates run as one
// fused launch.  Each step of a pair swaps the sets; the results equal the in-place
// mp_update_H_3d / mp_update_E_3d sequence bit for bit.
#pragma once
#include "common.h"

namespace mp {
// Whether the ping-pong step applies: FP32 fields and arithmetic, the vector kernels, an
// interior box with a CPML-shell split, box rows of at most 256 float4 groups.
bool yee_fused_ok(const mp_yee *g, const mp_cpml *c);
// H on the CPML shell from in's E and H into out's H.
int yee_fused_H_shell(const mp_yee *in, const mp_yee *out, mp_cpml *c, double dt, double mu0);
// H then E on the interior box: in's E and H (and the shell H already in out) into out's H
// and E; D in place.
int yee_fused_interior(const mp_yee *in, const mp_yee *out, const mp_cpml *c, double dt, double mu0);
// E on the CPML shell from out's (new) H into out's E; D in place.
int yee_fused_E_shell(const mp_yee *out, mp_cpml *c, double dt);
// Copy the cells no update writes (E on the low planes, H on the high planes; E only with
// e_only) from one set to the other.  H there changes only under an H source, E also under
// the source and the film list work, so a ping-pong step first carries E there over.
int yee_sync_boundary(const mp_yee *from, const mp_yee *to, bool e_only);
// Whether the next in-place H updates split the rest of their CPML shell by x (the box's x
// range without x-slab terms in a leaner kernel, then the x edges).  The step driver turns
// it on while the film update runs beside the H update: there the lean kernel fits two waves
// per SIMD where the general one fits one.  Alone, one launch is faster.
void yee_shell_split_x(bool on);
}  // namespace mp
