// qqcd_loops.h -- the disconnected-loop operators of the reference
// (loopops / spacesum) on the device (qqcd_loops.hip).
#pragma once
#include "qqcd_internal.h"
#include "qqcd_noise.h"

namespace qqcd {

// Loop operators of source b and sink x (FP64 spinors, b[par], x[par] on
// checkerboard par) with the FP64 18-real links U, summed per timeslice:
//   d_out[((t * 5 + m) * 9 + op) * 2 + ri] (+)= scale * (cur_scale for the four
//   point-split currents) * sum over the sites of timeslice t,
// op = J_1..J_4, scalar, V_1..V_4; m = momentum 1..5 (nmom = 1 leaves m > 0
// at zero); io = 1-based origin of the momentum factors (null: 1,1,1).
// d_part holds 2 * nt * loopsum_gx(g) * 58 doubles.  J (optional, two FP64
// spinors) receives the per-site operators in components 0..8.  d_out may be
// null when only J is wanted.
void loopsum(const Geom &g, const Gauge &U, const Spinor b[2], const Spinor x[2], int ksub, const int *io, int nmom,
             double scale, double cur_scale, double *d_out, int accumulate, double *d_part, Spinor *J,
             hipStream_t s);
int loopsum_gx(const Geom &g);

}  // namespace qqcd
