// qqcd_noise.h -- noise sources and local trace estimators (qqcd_noise.hip).
#pragma once
#include "qqcd_internal.h"

namespace qqcd {

// The 16 local Dirac matrices as (column, phase) per row.
struct Gamma16 {
  int col[16][4];
  double re[16][4], im[16][4];
};
Gamma16 gamma16();

// Z4 noise of right-hand side `rhs` on checkerboard par (FP64 spinor)
void z4_source(const Geom &g, Spinor &out, int par, uint64_t seed, long rhs, double scale, hipStream_t s);
// kind 0: Z4, 1: Z2
void noise_source(const Geom &g, Spinor &out, int par, uint64_t seed, long rhs, double scale, int kind,
                  hipStream_t s);

// d_out[(t*16 + n)*2 + {0,1}] (+)= sum over the sites of checkerboard-field
// timeslice t of conj(eta) Gamma_n x; d_part holds trace16_gx(g) * nt * 32 doubles
void trace16(const Geom &g, const Spinor &eta, const Spinor &x, double *d_out, int accumulate, double *d_part,
             hipStream_t s);
int trace16_gx(const Geom &g);

}  // namespace qqcd
