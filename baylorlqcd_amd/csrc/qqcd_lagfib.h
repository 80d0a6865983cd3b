// qqcd_lagfib.h -- the reference's lagged-Fibonacci random numbers (qqcd_lagfib.hip).
#pragma once
#include <cstdint>

#include "qqcd_internal.h"

namespace qqcd {

// iwlagfib(0:520), jrlagfib, krlagfib of qqcd-test/cfgsprops/glue/lagfib.f90
struct LagState {
  uint32_t w[521];
  int jr = 0, kr = 0;
};

// initlagfib(myid, iseed) (lagfib.f90:80-216), on the host
void lagfib_init(int myid, int iseed, LagState &st);
// throws unless 0 <= jr < 521 and kr = jr - 32 mod 521 (the only cursor pair rnds keeps)
void lagfib_check(const LagState &st);
// the next m raw draws (the integers rnds scales by fnorm) into d_out (device),
// st advanced past them; synchronises s
void lagfib_generate(LagState &st, long m, uint32_t *d_out, hipStream_t s);
// rn[t] = draw t (a signed 32-bit integer) * fnorm, as rnds returns it
void lagfib_to_double(const uint32_t *d_y, long m, double *d_rn, hipStream_t s);
// Z4 (kind 0) or Z2 (kind 1) noise of checkerboard par (FP64 spinor) from the
// 192 nvhalf draws z4source / z2source spend on that checkerboard
void lag_noise_source(const Geom &g, Spinor &out, int par, const uint32_t *d_draws, int kind, double scale,
                      hipStream_t s);
// the rnds scale factor: the single-precision literal 0.465661e-9 held in a double
constexpr double LAG_FNORM = (double)0.465661e-9f;
// kernel time (ms) of generating m draws in one launch of chunks of `chunk`
// draws (0: the length lagfib_generate picks); m is cut to 2048 chunks
double bench_lagfib(long m, int niter, long chunk, hipStream_t s);
// drop the cached jump-ahead tables (device memory)
void lagfib_release_device();
// draws of one qqcd_rnds_noalloc call at most (its static device workspace)
constexpr long LAG_NOALLOC_MAX = 1L << 20;
// the device address of that workspace on device dev (current), taken once
void lagfib_noalloc_setup(int dev);
// rnds(m, rn) into host rn through the workspace of device dev: no allocation
void rnds_noalloc(LagState &st, long m, double *rn, int dev, hipStream_t s);

}  // namespace qqcd
