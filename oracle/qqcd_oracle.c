/*
 * qqcd_oracle.c -- CPU oracle for the Wilson / twisted-mass solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP library is
 * compared against; it is never linked into, loaded by or called from the
 * product library.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load liboracle.so.
 *
 * It is a plain-C restatement, written from the mathematics, of the
 * operators in the reference Fortran (BaylorLQCD "qqcd"):
 *
 *   - the blocked / checkerboarded storage order of spinors and links
 *     (qqcd-test/cfgsprops/glue/lattice.f90:140-354, siteindex/setblk/
 *     gblchker; closed form in qqcd-test/cfgsprops/quark/quark.f90:82-105);
 *   - the hopping operator H_eo / H_oe of Hsingle
 *     (qqcd-test/cfgsprops/quark/diracops.f90:554-599 with mulbac :950,
 *     mulfor :1163, mv/mdv :1441-1530, boundary signs in setvector /
 *     slidevector :1641-1747);
 *   - the even-odd Schur operators Hdouble (:264-346) and Hdbletm
 *     (:435-550), including the signed kappa(2) < 1e-8 fallback (:467);
 *   - gammamult (:105-195) and vecdot (qqcd-test/cfgsprops/quark/
 *     inverters.f90:13485-13530);
 *   - the source preparation and odd-site reconstruction of fermprop
 *     (qqcd-test/cfgsprops/quark/quark.f90:570-617, 786-819) and the
 *     full-system operator M = G - kappa*H that defines the true residual.
 *
 * Internally the oracle works on the canonical lattice order
 *   site s = x + nx*(y + ny*(z + nz*t)),
 *   spinor psi[(s*4 + spin)*3 + colour]  (complex double),
 *   link   U[(s*4 + mu)*9 + row*3 + col] (complex double),
 * and converts from and to the reference order with the index map below.
 * The gamma matrices are held as explicit 4x4 complex tables and applied
 * by matrix multiplication: no spin projection, no fused sign patterns,
 * so that a sign or index slip in the GPU kernel's hand-derived projectors
 * cannot be mirrored here.
 *
 * Parity status: every function here is pinned by golden vectors written by
 * the reference Fortran itself (compiled from source by oracle/ref/Makefile)
 * and by closed-form checks (free-field plane waves, gamma5-hermiticity,
 * gauge covariance, dense brute force on 4^4); see tests/.
 */
#include <complex.h>
#include <stdint.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

/* ------------------------------------------------------------------ */
/* gamma matrices, qqcd-test/cfgsprops/quark/diracops.f90:61-71        */
/* index [mu][row][col], mu = 0..3 for gamma_1..gamma_4, 4 for gamma_5 */
/* ------------------------------------------------------------------ */
static cplx G[5][4][4];
static int G_ready = 0;

static void init_gamma(void)
{
    if (G_ready) return;
    memset(G, 0, sizeof(G));
    /* gamma_1 = antidiag(1,1,1,1) */
    G[0][0][3] = 1; G[0][1][2] = 1; G[0][2][1] = 1; G[0][3][0] = 1;
    /* gamma_2 */
    G[1][0][3] = -I; G[1][1][2] = I; G[1][2][1] = -I; G[1][3][0] = I;
    /* gamma_3 */
    G[2][0][2] = 1; G[2][1][3] = -1; G[2][2][0] = 1; G[2][3][1] = -1;
    /* gamma_4 = diag(1,1,-1,-1) */
    G[3][0][0] = 1; G[3][1][1] = 1; G[3][2][2] = -1; G[3][3][3] = -1;
    /* gamma_5 = gamma_1 gamma_2 gamma_3 gamma_4, computed, not typed in */
    cplx a[4][4], b[4][4];
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) {
            cplx s = 0;
            for (int k = 0; k < 4; k++) s += G[0][r][k] * G[1][k][c];
            a[r][c] = s;
        }
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) {
            cplx s = 0;
            for (int k = 0; k < 4; k++) s += a[r][k] * G[2][k][c];
            b[r][c] = s;
        }
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) {
            cplx s = 0;
            for (int k = 0; k < 4; k++) s += b[r][k] * G[3][k][c];
            G[4][r][c] = s;
        }
    G_ready = 1;
}

void orc_gamma_table(double *out /* [5][4][4][2] */)
{
    init_gamma();
    for (int m = 0; m < 5; m++)
        for (int r = 0; r < 4; r++)
            for (int c = 0; c < 4; c++) {
                out[((m * 4 + r) * 4 + c) * 2 + 0] = creal(G[m][r][c]);
                out[((m * 4 + r) * 4 + c) * 2 + 1] = cimag(G[m][r][c]);
            }
}

/* ------------------------------------------------------------------ */
/* geometry                                                            */
/* ------------------------------------------------------------------ */
static inline long vol(const int n[4]) { return (long)n[0] * n[1] * n[2] * n[3]; }

static inline long site_of(const int n[4], const int x[4])
{
    return x[0] + (long)n[0] * (x[1] + (long)n[1] * (x[2] + (long)n[2] * x[3]));
}

static inline void coords_of(const int n[4], long s, int x[4])
{
    x[0] = (int)(s % n[0]); s /= n[0];
    x[1] = (int)(s % n[1]); s /= n[1];
    x[2] = (int)(s % n[2]); s /= n[2];
    x[3] = (int)s;
}

/* nvhalf = V/32 and ntotal = nvhalf + nbmax for one process,
 * qqcd-test/cfgsprops/basics.f90:20-30 */
long orc_ref_nvhalf(const int n[4]) { return vol(n) / 32; }

long orc_ref_ntotal(const int n[4])
{
    long nvh = orc_ref_nvhalf(n), nb = 0;
    for (int mu = 0; mu < 4; mu++) {
        long b = 2 * nvh / n[mu];
        if (b > nb) nb = b;
    }
    return nvh + nb;
}

/* Reference storage slot of global site x (0-based coordinates).
 * Block bit b = x mod 2, in-block coordinate s = x/2 (lattice.f90:188-191);
 * block ibl = 1 + bx + 2by + 4bz + 8bt (setblk, lattice.f90:312-330);
 * in-block parity ieo = 1 + (sum s) mod 2 (lattice.f90:197-203);
 * site i = 1 + floor(L/2), L = sx + nsx(sy + nsy(sz + nsz st)) (isg,
 * lattice.f90:195-208; closed form quark.f90:82-87);
 * ibleo = ceil(ibl/2) and gblclr = 1 for even sum(b) (gblchker,
 * lattice.f90:349-352).  All outputs 1-based like the Fortran. */
void orc_ref_site(const int n[4], const int x[4], int *i, int *ieo, int *ibl,
                  int *ibleo, int *gblclr)
{
    int b[4], s[4], ns[4];
    for (int mu = 0; mu < 4; mu++) {
        b[mu] = x[mu] & 1;
        s[mu] = x[mu] >> 1;
        ns[mu] = n[mu] / 2;
    }
    long L = s[0] + (long)ns[0] * (s[1] + (long)ns[1] * (s[2] + (long)ns[2] * s[3]));
    *i = (int)(1 + L / 2);
    *ieo = 1 + ((s[0] + s[1] + s[2] + s[3]) & 1);
    *ibl = 1 + b[0] + 2 * b[1] + 4 * b[2] + 8 * b[3];
    *ibleo = (*ibl + 1) / 2;
    *gblclr = 1 + ((b[0] + b[1] + b[2] + b[3]) & 1);
}

/* offset of element (a, i, d, ieo, ibleo) of a (6,ntotal,4,2,8) array */
static inline long spin_off(long ntotal, int a, int i, int d, int ieo, int ibleo)
{
    return (a - 1) + 6 * ((i - 1) + ntotal * ((d - 1) + 4 * ((ieo - 1) + 2 * (long)(ibleo - 1))));
}

/* offset of element (a, i, mu, ieo, ibl) of a (18,ntotal,4,2,16) array */
static inline long link_off(long ntotal, int a, int i, int mu, int ieo, int ibl)
{
    return (a - 1) + 18 * ((i - 1) + ntotal * ((mu - 1) + 4 * ((ieo - 1) + 2 * (long)(ibl - 1))));
}

/* Canonical spinor -> reference be/bo arrays.  be holds the gblclr=1 sites
 * (global parity even), bo the gblclr=2 sites; either may be NULL.  Every
 * element of the output arrays (halo rows included) is written. */
void orc_spinor_canon_to_ref(const int n[4], const double *psi_ri, double *be, double *bo)
{
    const cplx *psi = (const cplx *)psi_ri;
    long ntot = orc_ref_ntotal(n), V = vol(n);
    long len = 6 * ntot * 4 * 2 * 8;
    if (be) memset(be, 0, sizeof(double) * len);
    if (bo) memset(bo, 0, sizeof(double) * len);
    for (long s = 0; s < V; s++) {
        int x[4], i, ieo, ibl, ibleo, gc;
        coords_of(n, s, x);
        orc_ref_site(n, x, &i, &ieo, &ibl, &ibleo, &gc);
        double *dst = gc == 1 ? be : bo;
        if (!dst) continue;
        for (int d = 1; d <= 4; d++)
            for (int c = 1; c <= 3; c++) {
                cplx v = psi[(s * 4 + (d - 1)) * 3 + (c - 1)];
                dst[spin_off(ntot, 2 * c - 1, i, d, ieo, ibleo)] = creal(v);
                dst[spin_off(ntot, 2 * c, i, d, ieo, ibleo)] = cimag(v);
            }
    }
}

/* Reference be/bo -> canonical spinor (rows 1..nvhalf only; a NULL input
 * leaves zeros on that parity). */
void orc_spinor_ref_to_canon(const int n[4], const double *be, const double *bo, double *psi_ri)
{
    cplx *psi = (cplx *)psi_ri;
    long ntot = orc_ref_ntotal(n), V = vol(n);
    for (long s = 0; s < V; s++) {
        int x[4], i, ieo, ibl, ibleo, gc;
        coords_of(n, s, x);
        orc_ref_site(n, x, &i, &ieo, &ibl, &ibleo, &gc);
        const double *src = gc == 1 ? be : bo;
        for (int d = 1; d <= 4; d++)
            for (int c = 1; c <= 3; c++) {
                cplx v = 0;
                if (src)
                    v = src[spin_off(ntot, 2 * c - 1, i, d, ieo, ibleo)] +
                        I * src[spin_off(ntot, 2 * c, i, d, ieo, ibleo)];
                psi[(s * 4 + (d - 1)) * 3 + (c - 1)] = v;
            }
    }
}

/* Link layout: u(a,i,mu,ieo,ibl), row-major 3x3, U_rc = u(6r+2c+1) + i u(6r+2c+2)
 * (0-based r,c), qqcd-test/cfgsprops/cfgsprops.f90:113-117. */
void orc_links_ref_to_canon(const int n[4], const double *u, double *U_ri)
{
    cplx *U = (cplx *)U_ri;
    long ntot = orc_ref_ntotal(n), V = vol(n);
    for (long s = 0; s < V; s++) {
        int x[4], i, ieo, ibl, ibleo, gc;
        coords_of(n, s, x);
        orc_ref_site(n, x, &i, &ieo, &ibl, &ibleo, &gc);
        for (int mu = 1; mu <= 4; mu++)
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++)
                    U[(s * 4 + (mu - 1)) * 9 + r * 3 + c] =
                        u[link_off(ntot, 6 * r + 2 * c + 1, i, mu, ieo, ibl)] +
                        I * u[link_off(ntot, 6 * r + 2 * c + 2, i, mu, ieo, ibl)];
    }
}

void orc_links_canon_to_ref(const int n[4], const double *U_ri, double *u)
{
    const cplx *U = (const cplx *)U_ri;
    long ntot = orc_ref_ntotal(n), V = vol(n);
    memset(u, 0, sizeof(double) * 18 * ntot * 4 * 2 * 16);
    for (long s = 0; s < V; s++) {
        int x[4], i, ieo, ibl, ibleo, gc;
        coords_of(n, s, x);
        orc_ref_site(n, x, &i, &ieo, &ibl, &ibleo, &gc);
        for (int mu = 1; mu <= 4; mu++)
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) {
                    cplx v = U[(s * 4 + (mu - 1)) * 9 + r * 3 + c];
                    u[link_off(ntot, 6 * r + 2 * c + 1, i, mu, ieo, ibl)] = creal(v);
                    u[link_off(ntot, 6 * r + 2 * c + 2, i, mu, ieo, ibl)] = cimag(v);
                }
    }
}

/* ------------------------------------------------------------------ */
/* hopping operator                                                    */
/* ------------------------------------------------------------------ */

/* out(x) = sum_mu [ (c1 - sg c2 gamma_mu) U_mu(x) in(x+mu) bc_fwd
 *                 + (c1 + sg c2 gamma_mu) U_mu(x-mu)^dag in(x-mu) bc_bwd ]
 * for every site x of global parity par_out (0 = even sum, 1 = odd); sg is
 * +1 for idag=0 and -1 for idag=1 (mulbac/mulfor, diracops.f90:1015-1154,
 * 1247-1386).  A hop that crosses the lattice edge in direction mu carries
 * the factor bc[mu] (setvector/slidevector, diracops.f90:1684-1694,
 * 1730-1740).  Sites of the other parity are left untouched. */
void orc_hop(const int n[4], const double *U_ri, const int bc[4], const double c1[4],
             const double c2[4], int idag, int par_out, const double *in_ri, double *out_ri)
{
    init_gamma();
    const cplx *U = (const cplx *)U_ri;
    const cplx *in = (const cplx *)in_ri;
    cplx *out = (cplx *)out_ri;
    long V = vol(n);
    double sg = idag ? -1.0 : 1.0;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < V; s++) {
        int x[4];
        coords_of(n, s, x);
        if (((x[0] + x[1] + x[2] + x[3]) & 1) != par_out) continue;
        cplx acc[4][3];
        memset(acc, 0, sizeof(acc));
        for (int mu = 0; mu < 4; mu++) {
            /* forward neighbour */
            int y[4] = {x[0], x[1], x[2], x[3]};
            double f = 1.0;
            y[mu] += 1;
            if (y[mu] == n[mu]) { y[mu] = 0; f = bc[mu]; }
            long sf = site_of(n, y);
            /* backward neighbour */
            int z[4] = {x[0], x[1], x[2], x[3]};
            double b = 1.0;
            z[mu] -= 1;
            if (z[mu] < 0) { z[mu] = n[mu] - 1; b = bc[mu]; }
            long sb = site_of(n, z);
            const cplx *Uf = U + (s * 4 + mu) * 9;
            const cplx *Ub = U + (sb * 4 + mu) * 9;
            cplx w1[4][3], w2[4][3];
            for (int d = 0; d < 4; d++)
                for (int r = 0; r < 3; r++) {
                    cplx a1 = 0, a2 = 0;
                    for (int c = 0; c < 3; c++) {
                        a1 += Uf[r * 3 + c] * in[(sf * 4 + d) * 3 + c];
                        a2 += conj(Ub[c * 3 + r]) * in[(sb * 4 + d) * 3 + c];
                    }
                    w1[d][r] = f * a1;
                    w2[d][r] = b * a2;
                }
            for (int d = 0; d < 4; d++)
                for (int r = 0; r < 3; r++) {
                    cplx a = c1[mu] * (w1[d][r] + w2[d][r]);
                    for (int e = 0; e < 4; e++)
                        a += sg * c2[mu] * G[mu][d][e] * (w2[e][r] - w1[e][r]);
                    acc[d][r] += a;
                }
        }
        for (int d = 0; d < 4; d++)
            for (int r = 0; r < 3; r++) out[(s * 4 + d) * 3 + r] = acc[d][r];
    }
}

/* out = (i gamma_5) in on the sites of parity par (par < 0: all sites). */
static void igam5(const int n[4], int par, const cplx *in, cplx *out)
{
    init_gamma();
    long V = vol(n);
#pragma omp parallel for schedule(static)
    for (long s = 0; s < V; s++) {
        int x[4];
        coords_of(n, s, x);
        if (par >= 0 && ((x[0] + x[1] + x[2] + x[3]) & 1) != par) continue;
        for (int r = 0; r < 3; r++) {
            cplx v[4];
            for (int d = 0; d < 4; d++) {
                cplx a = 0;
                for (int e = 0; e < 4; e++) a += I * G[4][d][e] * in[(s * 4 + e) * 3 + r];
                v[d] = a;
            }
            for (int d = 0; d < 4; d++) out[(s * 4 + d) * 3 + r] = v[d];
        }
    }
}

/* out = gamma_mu in (mu = 1..5) on all sites: gammamult, diracops.f90:105-195 */
void orc_gammamult(const int n[4], int mu, const double *in_ri, double *out_ri)
{
    init_gamma();
    const cplx *in = (const cplx *)in_ri;
    cplx *out = (cplx *)out_ri;
    long V = vol(n);
    for (long s = 0; s < V; s++)
        for (int r = 0; r < 3; r++) {
            cplx v[4];
            for (int d = 0; d < 4; d++) {
                cplx a = 0;
                for (int e = 0; e < 4; e++) a += G[mu - 1][d][e] * in[(s * 4 + e) * 3 + r];
                v[d] = a;
            }
            for (int d = 0; d < 4; d++) out[(s * 4 + d) * 3 + r] = v[d];
        }
}

typedef struct {
    int n[4];
    int bc[4];
    double c1[4], c2[4];
    const cplx *U;
} orc_ctx;

static void scale_par(const int n[4], int par, cplx a, const cplx *x, cplx b, const cplx *y, cplx *out)
{
    long V = vol(n);
    for (long s = 0; s < V; s++) {
        int xx[4];
        coords_of(n, s, xx);
        if (((xx[0] + xx[1] + xx[2] + xx[3]) & 1) != par) continue;
        for (int k = 0; k < 12; k++) out[s * 12 + k] = a * x[s * 12 + k] + (y ? b * y[s * 12 + k] : 0);
    }
}

/* Even-odd Schur operator on the even sites of canonical arrays:
 *   Wilson (kappa(2) < 1e-8, a signed test, diracops.f90:467-477; Hdouble
 *   :330-344):  out_e = in_e/kappa^2 - H_eo H_oe in_e
 *   twisted (diracops.f90:483-514), a = 2 kappa mu:
 *     out_e = (1 + a i g5) in_e / kappa^2
 *             - H_eo (1 - a i g5) H_oe in_e / (1 + a^2)
 * in and out must be distinct whole-lattice arrays; odd sites of out are
 * left untouched, tmp (whole lattice) is scratch. */
void orc_schur(const int n[4], const double *U_ri, const int bc[4], const double c1[4],
               const double c2[4], int idag, const double kappa[2], const double *in_ri,
               double *out_ri, double *tmp_ri, double *tmp2_ri)
{
    const cplx *in = (const cplx *)in_ri;
    cplx *out = (cplx *)out_ri, *tmp = (cplx *)tmp_ri, *tmp2 = (cplx *)tmp2_ri;
    double k = kappa[0], mu = kappa[1];
    orc_hop(n, U_ri, bc, c1, c2, idag, 1, in_ri, tmp_ri);
    if (mu < 1e-8) {
        orc_hop(n, U_ri, bc, c1, c2, idag, 0, tmp_ri, out_ri);
        scale_par(n, 0, 1.0 / (k * k), in, -1.0, out, out);
        return;
    }
    double a = 2.0 * k * mu;
    igam5(n, 1, tmp, tmp2);
    scale_par(n, 1, 1.0, tmp, -a, tmp2, tmp);
    orc_hop(n, U_ri, bc, c1, c2, idag, 0, tmp_ri, out_ri);
    igam5(n, 0, in, tmp2);
    long V = vol(n);
    double f3 = 1.0 / (1.0 + a * a);
    for (long s = 0; s < V; s++) {
        int xx[4];
        coords_of(n, s, xx);
        if (((xx[0] + xx[1] + xx[2] + xx[3]) & 1) != 0) continue;
        for (int q = 0; q < 12; q++)
            out[s * 12 + q] = in[s * 12 + q] / (k * k) + (2.0 * mu / k) * tmp2[s * 12 + q] - f3 * out[s * 12 + q];
    }
}

/* Full-lattice operator M x = G x - kappa H x with G = 1 + 2 kappa mu i g5
 * (twisted Wilson, no clover): the system fermprop solves
 * (diracops.f90:10-33, quark.f90:600-617, 797-819).  Used for the true
 * full-system residual ||b - M x|| / ||b||. */
void orc_full_m(const int n[4], const double *U_ri, const int bc[4], const double c1[4],
                const double c2[4], const double kappa[2], const double *x_ri, double *out_ri,
                double *tmp_ri)
{
    const cplx *x = (const cplx *)x_ri;
    cplx *out = (cplx *)out_ri, *tmp = (cplx *)tmp_ri;
    orc_hop(n, U_ri, bc, c1, c2, 0, 0, x_ri, out_ri);
    orc_hop(n, U_ri, bc, c1, c2, 0, 1, x_ri, out_ri);
    double k = kappa[0], a = 2.0 * kappa[0] * kappa[1];
    igam5(n, -1, x, tmp);
    long V = vol(n);
    for (long q = 0; q < V * 12; q++) out[q] = x[q] + a * tmp[q] - k * out[q];
}

/* Complex dot sum conj(a) b over whole canonical arrays, inverters.f90:13485 */
void orc_vecdot(long nc, const double *a_ri, const double *b_ri, double out[2])
{
    const cplx *a = (const cplx *)a_ri, *b = (const cplx *)b_ri;
    cplx s = 0;
    for (long q = 0; q < nc; q++) s += conj(a[q]) * b[q];
    out[0] = creal(s);
    out[1] = cimag(s);
}

/* ------------------------------------------------------------------ */
/* reference-layout wrappers (argument meaning of the Fortran)         */
/* ------------------------------------------------------------------ */

/* coact(4,4,4) Fortran array: c1[mu] = coact(3,1,mu), c2[mu] = coact(3,2,mu) */
static void coact_split(const double *coact, double c1[4], double c2[4])
{
    for (int mu = 0; mu < 4; mu++) {
        c1[mu] = coact[2 + 4 * (0 + 4 * mu)];
        c2[mu] = coact[2 + 4 * (1 + 4 * mu)];
    }
}

/* Hsingle(gout,u,gin,idag,coact,bc,gblclr,...), diracops.f90:554-599:
 * gout = H_eo gin for gblclr=1, H_oe gin for gblclr=2; all of gout is
 * zeroed first (:589), only rows 1..nvhalf carry results. */
void orc_hsingle_ref(const int n[4], const double *u, const double *coact, const int bc[4],
                     int gblclr, int idag, const double *gin, double *gout)
{
    long V = vol(n);
    double c1[4], c2[4];
    coact_split(coact, c1, c2);
    double *U = malloc(sizeof(double) * 2 * 36 * V);
    double *a = malloc(sizeof(double) * 24 * V);
    double *b = calloc(24 * V, sizeof(double));
    orc_links_ref_to_canon(n, u, U);
    if (gblclr == 1) orc_spinor_ref_to_canon(n, NULL, gin, a);
    else orc_spinor_ref_to_canon(n, gin, NULL, a);
    orc_hop(n, U, bc, c1, c2, idag, gblclr - 1, a, b);
    if (gblclr == 1) orc_spinor_canon_to_ref(n, b, gout, NULL);
    else orc_spinor_canon_to_ref(n, b, NULL, gout);
    free(U); free(a); free(b);
}

/* Hdbletm(he,u,GeeGooinv,ge,idag,coact,kappa,iflag=-1,...) and, through
 * its kappa(2) < 1e-8 branch, Hdouble(he,u,ge,idag,coact,kappa(1),...). */
void orc_hdbletm_ref(const int n[4], const double *u, const double *coact, const int bc[4],
                     int idag, const double kappa[2], const double *ge, double *he)
{
    long V = vol(n);
    double c1[4], c2[4];
    coact_split(coact, c1, c2);
    double *U = malloc(sizeof(double) * 2 * 36 * V);
    double *a = malloc(sizeof(double) * 24 * V);
    double *b = calloc(24 * V, sizeof(double));
    double *t1 = calloc(24 * V, sizeof(double));
    double *t2 = calloc(24 * V, sizeof(double));
    orc_links_ref_to_canon(n, u, U);
    orc_spinor_ref_to_canon(n, ge, NULL, a);
    orc_schur(n, U, bc, c1, c2, idag, kappa, a, b, t1, t2);
    orc_spinor_canon_to_ref(n, b, he, NULL);
    free(U); free(a); free(b); free(t1); free(t2);
}

/* ------------------------------------------------------------------------
 * Z4 noise (the library's counter-based restatement of z4source,
 * qqcd-test/cfgsprops/quark/disconloops.f90:4518-4569): component k =
 * 3*spin + colour of site s of right-hand side rhs takes the uniform number
 *   u = (mix(seed + G*(1 + (rhs*V + s)*12 + k)) >> 11) * 2^-53,
 * mix = the splitmix64 finaliser, G = 0x9E3779B97F4A7C15, and the value
 * 1 (u > 0.75), -1 (0.5 < u < 0.75), i (0.25 < u < 0.5), -i (otherwise),
 * times scale.  Canonical output psi[(s*4 + spin)*3 + colour].
 * ---------------------------------------------------------------------- */
static uint64_t orc_mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* kind 0: Z4 (as above); kind 1: Z2, 1 for u > 0.5 else -1 in the real part
 * (z2source, qqcd-test/cfgsprops/quark/disconloops.f90:4583-4649) */
void orc_noise(const int n[4], uint64_t seed, long rhs, double scale, int kind, double *psi_ri);

void orc_z4(const int n[4], uint64_t seed, long rhs, double scale, double *psi_ri)
{
    orc_noise(n, seed, rhs, scale, 0, psi_ri);
}

void orc_noise(const int n[4], uint64_t seed, long rhs, double scale, int kind, double *psi_ri)
{
    const long V = vol(n);
    for (long s = 0; s < V; s++)
        for (int k = 0; k < 12; k++) {
            const uint64_t c = ((uint64_t)rhs * (uint64_t)V + (uint64_t)s) * 12ull + (uint64_t)k;
            const double u = (double)(orc_mix(seed + 0x9E3779B97F4A7C15ull * (1ull + c)) >> 11) *
                             (1.0 / 9007199254740992.0);
            double re = 0, im = 0;
            if (kind == 1) re = u > 0.5 ? 1 : -1;
            else if (u > 0.75) re = 1;
            else if (0.5 < u && u < 0.75) re = -1;
            else if (0.25 < u && u < 0.5) im = 1;
            else im = -1;
            psi_ri[2 * (s * 12 + k)] = scale * re;
            psi_ri[2 * (s * 12 + k) + 1] = scale * im;
        }
}
