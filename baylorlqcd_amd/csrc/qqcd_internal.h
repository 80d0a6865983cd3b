// qqcd_internal.h -- shared definitions of the MI355X solve-path library.
//
// Device data layout (all fields resident in HBM):
//   * even-odd: a field lives on one global checkerboard (parity p = 0 for
//     x+y+z+t even -- the reference's gblclr=1 -- and p = 1 for odd);
//   * site order inside a checkerboard: cb = (x + nx*(y + ny*(z + nz*t))) / 2,
//     so consecutive lanes of a wavefront hold x-neighbouring sites;
//   * spinors are site-fastest structure-of-arrays: component k = 3*spin +
//     colour (12 complex components) is a contiguous run of `stride` complex
//     numbers, so each of the 12 loads of a wavefront is one contiguous
//     64-lane transaction.  FP16 spinors carry one float norm per site;
//   * links per (parity, mu): 9 (18-real) or 6 (12-real: rows 0 and 1, row 2
//     rebuilt as conj(row0 x row1)) complex component runs of `stride`.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

namespace qqcd {

enum Prec { PREC_DOUBLE = 0, PREC_SINGLE = 1, PREC_HALF = 2 };
enum Recon { RECON_18 = 18, RECON_12 = 12 };

typedef _Float16 f16;

template <typename T> struct cx { T re, im; };

template <typename T> __host__ __device__ inline cx<T> mk(T a, T b) { return cx<T>{a, b}; }
template <typename T> __host__ __device__ inline cx<T> operator+(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T> __host__ __device__ inline cx<T> operator-(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T> __host__ __device__ inline cx<T> operator*(T s, cx<T> a) { return {s * a.re, s * a.im}; }
template <typename T> __host__ __device__ inline cx<T> cmul(cx<T> a, cx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// conj(a) * b
template <typename T> __host__ __device__ inline cx<T> cjmul(cx<T> a, cx<T> b) {
  return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}
// i * a
template <typename T> __host__ __device__ inline cx<T> mul_i(cx<T> a) { return {-a.im, a.re}; }

// Lattice geometry of one process / one GPU.
struct Geom {
  int n[4];      // nx, ny, nz, nt (all even, nx divisible by 4 for the reference layout)
  long V;        // sites
  long Vh;       // sites per checkerboard
  long stride;   // allocation stride of one component run (>= Vh, multiple of 64)
  int bc[4];     // 1 periodic, -1 antiperiodic, 0 fixed (hop across the edge scaled by bc)
  double c1[4];  // coact(3,1,mu)
  double c2[4];  // coact(3,2,mu)
};

// A spinor field on one checkerboard, any precision.
struct Spinor {
  Prec prec = PREC_DOUBLE;
  long stride = 0;
  void *data = nullptr;   // 12 component runs
  float *norm = nullptr;  // FP16 only: per-site scale
  size_t bytes() const;
};

// Gauge field on the device: links[parity][mu] component runs.
struct Gauge {
  Prec prec = PREC_DOUBLE;
  int recon = 18;
  long stride = 0;
  void *data = nullptr;
  size_t bytes() const;
};

// Several right-hand sides on one checkerboard, RHS-innermost: slots are
// grouped in blocks of nb = 2^lognb, and element (slot r, component k, site
// j) sits at ((r/nb * 12 + k) * stride + j) * nb + r%nb, so the nb lanes of
// one site are adjacent and share every link load.  FP16 fields carry one
// float norm per (site, slot) at (r/nb * stride + j) * nb + r%nb.
struct MultiSpinor {
  Prec prec = PREC_SINGLE;
  int cap = 0;  // right-hand-side slots
  int nb = 1;
  int lognb = 0;
  long stride = 0;
  void *data = nullptr;
  float *norm = nullptr;
  int nblk() const { return (cap + nb - 1) / nb; }
  long blk_elems() const { return 12L * stride * nb; }
  long elems() const { return blk_elems() * nblk(); }
  size_t bytes() const;
};

// Launch shape of the multi-RHS hopping kernel (picked by the autotuner).
struct MultiTune {
  int order = 1;  // 0: RHS block outermost; 1: innermost
  int grid = 0;   // persistent workgroups (0: occupancy x CUs)
  // kernel 0: hops one after the other, tiles dealt to the XCDs per sweep of
  // the grid; 1: pipelined hops (depth 2), each XCD owning one contiguous
  // range of tiles for the whole launch and walking its sites in blocks of
  // bz z-planes swept in t (bz 0: the largest of 8, 4, 2 dividing nz); 2: the
  // same with one workgroup per tile, dispatched in order, instead of a
  // persistent grid; 3 / 4: FP32 spinors two right-hand sides per lane, one
  // workgroup per tile, two / one hops' loads in flight (FP64 spinors: as 1)
  int kernel = 3;
  int bz = 0;
  int nt = 0;  // kernel 1: links and output streamed past the L2 (measured slower here)
  int lds = 0;  // kernel 2: dynamic LDS per workgroup (unused; caps the workgroups per CU)
};

// Launch shape of the single-right-hand-side hopping kernel (projector form):
// depth 0 = hops computed one after the other; 1/2/4 = the pipelined kernel
// with that many hops' loads in flight; order/bz = site traversal (0: the
// checkerboard index, 1: blocks of bz z-planes swept in t; bz 0 picks the
// largest of 16, 8, 4, 2 dividing nz); nt = links and output streamed past
// the L2.  The default is the fastest shape measured at 32^3x64 in all six
// storage formats.
struct DslashTune {
  int order = 1;
  int bz = 0;
  int depth = 2;
  int nt = 1;
};
void set_dslash_tune(const DslashTune &t);
// overrides of the multi-RHS launch shape for every call (-1: the caller's)
void set_multi_override(int kernel, int bz, int nt);
void set_multi_lds(int bytes);
DslashTune get_dslash_tune();

// Coefficients of the local spin matrix applied in a Dslash epilogue:
//   out = a0*x + a1*(i g5 x) + b0*D + b1*(i g5 D)
// with D the hopping sum and x a spinor on the output checkerboard.
struct Epilogue {
  double a0 = 0, a1 = 0, b0 = 1, b1 = 0;
};

// Per-direction hopping coefficients as the kernel consumes them.  In
// projector form (|c1| == |c2|) the forward hop is c1 (1 - sig g_mu) U psi(x+mu)
// and the backward hop c1 (1 + sig g_mu) U^dag psi(x-mu), sig = s*c2/c1.
// Otherwise each hop is split into its (1+g) and (1-g) projections.
struct HopCoef {
  float dummy;
  double fw_p[4], fw_m[4];  // forward:  fw_p (1+g) + fw_m (1-g)
  double bw_p[4], bw_m[4];  // backward: bw_p (1+g) + bw_m (1-g)
  int proj;                 // 1 if every direction is a single projector
};
HopCoef make_hopcoef(const Geom &g, int idag);

// error handling: every entry point returns 0 or a negative code and stores
// a message retrievable with qqcd_last_error().
void set_error(const std::string &msg);
#define QQCD_HIP(call)                                                          \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess) {                                                     \
      throw std::runtime_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
    }                                                                           \
  } while (0)

// ---- kernels (qqcd_dslash.hip) ----
// out32 (FP16 spinors only): the result also as an FP32 field, in the same
// launch where the pipelined kernel runs; only32 leaves out unwritten there
// (out is then scratch for the other kernels).
void dslash(const Geom &g, const Gauge &U, const Spinor &in, Spinor &out, int par_out, int idag,
            const Epilogue &ep, const Spinor *x, hipStream_t s, Spinor *out32 = nullptr, bool only32 = false);

// Hopping operator on the active slots [0, *d_nact) of a multi-RHS field
// (d_nact in device memory); same epilogue as dslash.  With slot_n2 (device,
// one value per slot) the epilogue's result is scaled by 1/sqrt(slot_n2[r])
// (0 where that is not positive), and with side (an FP16 field without
// norms) the epilogue also writes side = side_hs x / sqrt(slot_n2[r]); both
// need the two-slot FP32 kernel (dslash_multi_can_scale) and x.
void dslash_multi(const Geom &g, const Gauge &U, const MultiSpinor &in, MultiSpinor &out, int par_out, int idag,
                  const Epilogue &ep, const MultiSpinor *x, const int *d_nact, const MultiTune &t, hipStream_t s,
                  const double *slot_n2 = nullptr, MultiSpinor *side = nullptr, double side_hs = 0.0);
// whether dslash_multi on these fields takes the two-slot FP32 kernel
bool dslash_multi_can_scale(const Geom &g, Prec p, int nb, const MultiTune &t);

// ---- layout (qqcd_layout.hip) ----
long ref_nvhalf(const Geom &g);
long ref_ntotal(const Geom &g);
void ref_to_spinor(const Geom &g, const double *dref, Spinor &out, int par, hipStream_t s);
void spinor_to_ref(const Geom &g, const Spinor &in, double *dref, int par, hipStream_t s);
void ref_to_gauge(const Geom &g, const double *dref, Gauge &U, hipStream_t s);
void canon_to_gauge(const Geom &g, const double *dcanon, Gauge &U, hipStream_t s);
// out = scale * in across precisions (FP16 fields carry their per-site norm)
void spinor_copy_convert(const Geom &g, const Spinor &in, Spinor &out, hipStream_t s, double scale = 1.0);
// out = scale * in / sqrt(*dn2) (0 when *dn2 == 0), dn2 in device memory
void spinor_convert_rnorm(const Geom &g, const Spinor &in, Spinor &out, const double *dn2, hipStream_t s,
                          double scale = 1.0);

}  // namespace qqcd
