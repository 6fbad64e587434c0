/*
 * CPU oracle for the 27-point mixed-grid Helmholtz operator with PML.
 *
 * TEST INFRASTRUCTURE. This file is the checker, never the product: only
 * tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py may
 * load the library built from it.  The product path (the HIP library) never
 * links or calls it.
 *
 * It restates, from the mathematics and with a different organisation (one
 * coefficient rule K(d) evaluated per neighbour instead of 27 hand-written
 * lines), the reference stencil of SLIM-release-apps
 * tools/algorithms/3DFreqModeling/Helm3d_27pt.h:
 *
 *   stencil weights W1..WM4 .............. Helm3d_27pt.h:194-200
 *   h-dependent constants ................ Helm3d_27pt.h:639-671 (compute_coef_consts)
 *   PML profile gamma / xi_lo / xi_hi ..... Helm3d_27pt.h:326-390
 *   xi sampled at the 1-based index i+1 .. Helm3d_27pt.h:394-471
 *   forward coefficients ................. Helm3d_27pt.h:680-714
 *   adjoint coefficients (xi taken at the
 *     neighbour, wn at the centre, conj) . Helm3d_27pt.h:718-822
 *   zero-padded neighbourhoods ........... Helm3d_27pt.h:263-317, 483-631
 *   driver (one output per point) ........ Helm3d_27pt.h:827-875
 *   compile-time variants ADJ/DERIV/WNCMPLX Helm3d_27pt.h:4-16, 202-206
 *
 * Memory order is x fastest: idx = i + j*nx + k*nx*ny (Helm3d_27pt.h:21-23).
 * Complex vectors cross this interface as split real/imaginary arrays, as at
 * the MEX boundary (Helm3dmvp_mex.c:67-98).
 *
 * Kaczmarz sweeps (Helm3dkaczswp_mex.c:60-117) are restated at the end of the
 * file; pinned by tests/test_oracle_kaczmarz.py.
 *
 * Pinned by: tests/test_oracle_stencil.py compares every variant with the
 * reference header compiled in place (oracle/_ref) on seeded inputs, checks
 * <Hx,y> = <x,H^H y> for (fwd, adj) and (deriv, adj-deriv), the DERIV linearity
 * in wn, and the interior constant-coefficient closed form.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

/* Stencil weights exactly as the C literals of Helm3d_27pt.h:194-200
 * (the MATLAB fractions differ by ~1e-15; the MEX path is the parity target). */
static const double W1 = 1.8395262e-5;
static const double W2 = 0.296692333333333;
static const double W3 = 0.027476150000000;
static const double WM1 = 0.49649658;
static const double WM2 = 0.075168750000000;
static const double WM3 = 0.004373916666667;
static const double WM4 = 5.690375e-07;

typedef struct {
    double wn_coef;   /* centre mass weight */
    double c[3];      /* pml{x,y,z}_coef: face weight along each axis (negative) */
    double xy, xz, yz;/* cross weights W2/(2 h_a^2 + 2 h_b^2) */
    double w3a2;      /* 2 * 3 W3 / (4 |h|^2) */
} hconst;

/* Helm3d_27pt.h:639-671, same evaluation order. */
static hconst make_hconst(const double *h)
{
    hconst k;
    double hx2 = h[0] * h[0], hy2 = h[1] * h[1], hz2 = h[2] * h[2];
    double hxy = hx2 + hy2, hxz = hx2 + hz2, hyz = hy2 + hz2;
    double hxyz = hx2 + hy2 + hz2;
    double w3a = (W3) * 3 / (4 * hxyz);
    k.w3a2 = 2 * w3a;
    k.wn_coef = -(W1 + 3 * W2 + 16 * w3a * hxyz / 3 + WM1 - 1);
    k.c[0] = -(W1 / hx2 + W2 / hx2 + W2 / hxz + W2 / hxy + 8 * w3a);
    k.c[1] = -(W1 / hy2 + W2 / hy2 + W2 / hyz + W2 / hxy + 8 * w3a);
    k.c[2] = -(W1 / hz2 + W2 / hz2 + W2 / hxz + W2 / hyz + 8 * w3a);
    k.xz = W2 / (2 * hxz);
    k.xy = W2 / (2 * hxy);
    k.yz = W2 / (2 * hyz);
    return k;
}

/* PML damping profile, Helm3d_27pt.h:326-334.  x is 1-based. */
static double gam_lo(double x, double n, double nlo)
{
    return cos(M_PI * ((x - 1) * n / (2 * (n + 1) * nlo)));
}
static double gam_hi(double x, double n, double nhi)
{
    return cos(M_PI * ((1 - (x - 1) / (n + 1)) * n / (2 * nhi)));
}

/* xi_lo(x) = 2 / [(2 - g(x+1)(g(x+1)+g(x))) + i(3 g(x+1) + g(x))],
 * Helm3d_27pt.h:346-367.  The interior test uses the LOWER count at both ends
 * (the reference's own behaviour, kept for parity). */
static cplx xi_lower(double x, double n, double nlo, double nhi)
{
    double g0, g1;
    if (x <= nlo) {
        g1 = (x == nlo) ? 0 : gam_lo(x + 1, n, nlo);
        g0 = gam_lo(x, n, nlo);
    } else if (x <= n + 2 - nlo) {
        g0 = 0;
        g1 = (x == n + 2 - nlo) ? gam_hi(x + 1, n, nhi) : 0;
    } else {
        g0 = gam_hi(x, n, nhi);
        g1 = gam_hi(x + 1, n, nhi);
    }
    return (2.0 + 0.0 * I) / ((2.0 - g1 * (g1 + g0)) + (3.0 * g1 + g0) * I);
}

/* xi_hi(x) = 2 / [(2 - g(x+1)(g(x+1)+g(x+2))) + i(3 g(x+1) + g(x+2))],
 * Helm3d_27pt.h:369-390. */
static cplx xi_upper(double x, double n, double nlo, double nhi)
{
    double g1, g2;
    if (x <= nlo) {
        g1 = (x == nlo) ? 0 : gam_lo(x + 1, n, nlo);
        g2 = (x >= nlo - 1) ? 0 : gam_lo(x + 2, n, nlo);
    } else if (x <= n + 2 - nlo) {
        g2 = (x >= n + 1 - nlo) ? gam_hi(x + 2, n, nhi) : 0;
        g1 = (x == n + 2 - nlo) ? gam_hi(x + 1, n, nhi) : 0;
    } else {
        g2 = gam_hi(x + 2, n, nhi);
        g1 = gam_hi(x + 1, n, nhi);
    }
    return (2.0 + 0.0 * I) / ((2.0 - g1 * (g1 + g2)) + (3.0 * g1 + g2) * I);
}

/* One axis' xi values for 0-based indices -1..n (the adjoint reads the
 * neighbour's value, Helm3d_27pt.h:397-402, 437-463).  Slot t holds index t-1. */
typedef struct { cplx *lo, *hi; } axis_tab;

static axis_tab make_axis(int n, int nlo, int nhi)
{
    axis_tab a;
    a.lo = (cplx *)malloc(sizeof(cplx) * (size_t)(n + 2));
    a.hi = (cplx *)malloc(sizeof(cplx) * (size_t)(n + 2));
    for (int t = 0; t < n + 2; t++) {
        double x = (double)t; /* 1-based position of 0-based index t-1 */
        a.lo[t] = xi_lower(x, n, nlo, nhi);
        a.hi[t] = xi_upper(x, n, nlo, nhi);
    }
    return a;
}
static void free_axis(axis_tab *a) { free(a->lo); free(a->hi); }

/* Stiffness (non-mass) part of the coefficient that couples a point whose
 * PML values are (lo/hi per axis) to its neighbour at offset d.
 * Helm3d_27pt.h:680-714 (the non_deriv_mode terms), written as one rule:
 *   centre:      -(c_x xi_x + c_y xi_y + c_z xi_z)
 *   face  a:      c_a xi_a^s + sum_{b != a} w_ab xi_b
 *   edge  (a,b): -w_ab (xi_a^s + xi_b^s) + w3a2 xi_c
 *   corner:      -w3a2 (xi_x^s + xi_y^s + xi_z^s)
 * where s = lo for a -1 offset and hi for +1, and xi = lo + hi. */
static cplx stencil_K(const hconst *k, const int d[3], const cplx lo[3], const cplx hi[3])
{
    cplx full[3], side[3];
    int nz = 0, on[3];
    for (int a = 0; a < 3; a++) {
        full[a] = lo[a] + hi[a];
        side[a] = d[a] < 0 ? lo[a] : hi[a];
        on[a] = d[a] != 0;
        nz += on[a];
    }
    /* w_ab indexed by the pair of axes */
    double w[3][3] = {{0, k->xy, k->xz}, {k->xy, 0, k->yz}, {k->xz, k->yz, 0}};
    cplx K = 0;
    if (nz == 0) {
        K = -(k->c[0] * full[0] + k->c[1] * full[1] + k->c[2] * full[2]);
    } else if (nz == 1) {
        int a = on[0] ? 0 : (on[1] ? 1 : 2);
        K = k->c[a] * side[a];
        for (int b = 0; b < 3; b++)
            if (b != a) K += w[a][b] * full[b];
    } else if (nz == 2) {
        int c = !on[0] ? 0 : (!on[1] ? 1 : 2);
        int a = (c == 0) ? 1 : 0, b = (c == 2) ? 1 : 2;
        K = -w[a][b] * (side[a] + side[b]) + k->w3a2 * full[c];
    } else {
        K = -k->w3a2 * (side[0] + side[1] + side[2]);
    }
    return K;
}

/* Mass-lumping weight by the number of non-zero offsets, Helm3d_27pt.h:680-714
 * (wn_coef at the centre, -WM2 faces, -WM3 edges, -WM4 corners). */
static double mass_mu(const hconst *k, int nnz)
{
    switch (nnz) {
    case 0: return k->wn_coef;
    case 1: return -WM2;
    case 2: return -WM3;
    default: return -WM4;
    }
}

/* Generic operator.  adj: y = H^H x; deriv: stiffness dropped, wn -> dwn;
 * wnc: wn complex (wni read).  Signature otherwise that of do_Hmvp
 * (Helm3d_27pt.h:827). */
static void hmvp(int adj, int deriv, int wnc,
                 const double *wnr, const double *wni, const double *h,
                 const double *n, const double *npml, double *yr, double *yi,
                 const double *xr, const double *xi, int nthreads)
{
    const int nx = (int)n[0], ny = (int)n[1], nz = (int)n[2];
    const hconst hc = make_hconst(h);
    axis_tab ax[3];
    ax[0] = make_axis(nx, (int)npml[0], (int)npml[1]);
    ax[1] = make_axis(ny, (int)npml[2], (int)npml[3]);
    ax[2] = make_axis(nz, (int)npml[4], (int)npml[5]);
    const long sy = nx, sz = (long)nx * ny;
    if (nthreads < 1) nthreads = 1;

#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int k = 0; k < nz; k++) {
        for (int j = 0; j < ny; j++) {
            for (int i = 0; i < nx; i++) {
                const int p[3] = {i, j, k};
                const long ip = i + j * sy + k * sz;
                cplx acc = 0;
                for (int dz = -1; dz <= 1; dz++) {
                    for (int dy = -1; dy <= 1; dy++) {
                        for (int dx = -1; dx <= 1; dx++) {
                            const int d[3] = {dx, dy, dz};
                            int q[3], inside = 1;
                            for (int a = 0; a < 3; a++) {
                                q[a] = p[a] + d[a];
                                if (q[a] < 0 || q[a] >= (int)n[a]) inside = 0;
                            }
                            if (!inside) continue; /* zero padding */
                            const long iq = q[0] + q[1] * sy + q[2] * sz;
                            const int nnz = (dx != 0) + (dy != 0) + (dz != 0);
                            cplx c;
                            if (!adj) {
                                /* row p: xi at p, wn at the neighbour */
                                cplx w = wnc ? wnr[iq] + wni[iq] * I : wnr[iq];
                                c = mass_mu(&hc, nnz) * w;
                                if (!deriv) {
                                    cplx lo[3], hi[3];
                                    for (int a = 0; a < 3; a++) {
                                        lo[a] = ax[a].lo[p[a] + 1];
                                        hi[a] = ax[a].hi[p[a] + 1];
                                    }
                                    c += stencil_K(&hc, d, lo, hi);
                                }
                            } else {
                                /* column p of H, conjugated: the coupling of row q
                                 * to p (offset -d), xi at q, wn at p */
                                cplx w = wnc ? wnr[ip] + wni[ip] * I : wnr[ip];
                                c = mass_mu(&hc, nnz) * w;
                                if (!deriv) {
                                    cplx lo[3], hi[3];
                                    int md[3] = {-dx, -dy, -dz};
                                    for (int a = 0; a < 3; a++) {
                                        lo[a] = ax[a].lo[q[a] + 1];
                                        hi[a] = ax[a].hi[q[a] + 1];
                                    }
                                    c += stencil_K(&hc, md, lo, hi);
                                }
                                c = conj(c);
                            }
                            acc += c * (xr[iq] + (xi ? xi[iq] : 0.0) * I);
                        }
                    }
                }
                yr[ip] = creal(acc);
                yi[ip] = cimag(acc);
            }
        }
    }
    for (int a = 0; a < 3; a++) free_axis(&ax[a]);
}

/* The six variants built as MEX binaries (ibin/install_MEX:88-105), each with
 * do_Hmvp's signature. */
#define ORACLE_VARIANT(name, adj, deriv, wnc)                                         \
    void name(const double *wnr, const double *wni, const double *h, const double *n, \
              const double *npml, double *yr, double *yi, const double *xr,           \
              const double *xi, int nthreads)                                         \
    {                                                                                 \
        hmvp(adj, deriv, wnc, wnr, wni, h, n, npml, yr, yi, xr, xi, nthreads);        \
    }
ORACLE_VARIANT(oracle_Hmvp_forw, 0, 0, 0)
ORACLE_VARIANT(oracle_Hmvp_adj, 1, 0, 0)
ORACLE_VARIANT(oracle_Hmvp_forw_deriv, 0, 1, 0)
ORACLE_VARIANT(oracle_Hmvp_adj_deriv, 1, 1, 0)
ORACLE_VARIANT(oracle_Hmvp_forw_wnc, 0, 0, 1)
ORACLE_VARIANT(oracle_Hmvp_adj_wnc, 1, 0, 1)

/* PML tables as seen by the oracle (for table-level checks of the product):
 * out_lo/out_hi receive n+2 complex values (interleaved re,im) for 0-based
 * indices -1..n. */
void oracle_pml_axis(int n, int nlo, int nhi, double *out_lo, double *out_hi)
{
    axis_tab a = make_axis(n, nlo, nhi);
    for (int t = 0; t < n + 2; t++) {
        out_lo[2 * t] = creal(a.lo[t]); out_lo[2 * t + 1] = cimag(a.lo[t]);
        out_hi[2 * t] = creal(a.hi[t]); out_hi[2 * t + 1] = cimag(a.hi[t]);
    }
    free_axis(&a);
}

/* h-dependent constants: wn_coef, c_x, c_y, c_z, xy, xz, yz, w3a2. */
void oracle_hconst(const double *h, double *out8)
{
    hconst k = make_hconst(h);
    out8[0] = k.wn_coef; out8[1] = k.c[0]; out8[2] = k.c[1]; out8[3] = k.c[2];
    out8[4] = k.xy; out8[5] = k.xz; out8[6] = k.yz; out8[7] = k.w3a2;
}

/* ------------------------------------------------------------------------
 * Kaczmarz row projections (Helm3dkaczswp_mex.c:60-117).
 *
 * Row p of H (adj = 0) or of H^H (adj = 1): the same coefficients as the
 * matvec above, listed for the in-grid neighbours in the order dz, dy, dx =
 * -1..1 (x fastest).  A projection with relaxation omega:
 *   d = omega (b_p - a.x) / |a|^2 ;  x_q += d conj(a_q)   (:96-117)
 * A double sweep visits every row forward and then backward
 * (xyzloop_updown, Helm3d_27pt.h:63-108).  Two orders:
 *   lex: k, j, i ascending, then descending (the reference's sweep);
 *   mc:  27 colours c = (i mod 3) + 3 (j mod 3) + 9 (k mod 3) ascending, then
 *        descending; rows of one colour have disjoint 3x3x3 supports, so
 *        their order inside the colour does not matter (the GPU form).
 * wn complex (wnc) is honoured for both orders.
 * ------------------------------------------------------------------------ */

static int row_of(int adj, int wnc, const double *wnr, const double *wni, const hconst *hc,
                  const axis_tab ax[3], const int n[3], const int p[3], cplx coef[27], long idx[27])
{
    const long sy = n[0], sz = (long)n[0] * n[1];
    const long ip = p[0] + p[1] * sy + p[2] * sz;
    int cnt = 0;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                const int d[3] = {dx, dy, dz};
                int q[3], inside = 1;
                for (int a = 0; a < 3; a++) {
                    q[a] = p[a] + d[a];
                    if (q[a] < 0 || q[a] >= n[a]) inside = 0;
                }
                if (!inside) continue;
                const long iq = q[0] + q[1] * sy + q[2] * sz;
                const int nnz = (dx != 0) + (dy != 0) + (dz != 0);
                cplx c, lo[3], hi[3];
                if (!adj) {
                    cplx w = wnc ? wnr[iq] + wni[iq] * I : wnr[iq];
                    for (int a = 0; a < 3; a++) {
                        lo[a] = ax[a].lo[p[a] + 1];
                        hi[a] = ax[a].hi[p[a] + 1];
                    }
                    c = mass_mu(hc, nnz) * w + stencil_K(hc, d, lo, hi);
                } else {
                    cplx w = wnc ? wnr[ip] + wni[ip] * I : wnr[ip];
                    int md[3] = {-dx, -dy, -dz};
                    for (int a = 0; a < 3; a++) {
                        lo[a] = ax[a].lo[q[a] + 1];
                        hi[a] = ax[a].hi[q[a] + 1];
                    }
                    c = conj(mass_mu(hc, nnz) * w + stencil_K(hc, md, lo, hi));
                }
                coef[cnt] = c;
                idx[cnt] = iq;
                cnt++;
            }
    return cnt;
}

static void project(int adj, int wnc, const double *wnr, const double *wni, const hconst *hc,
                    const axis_tab ax[3], const int n[3], const int p[3], double omega, double *xr,
                    double *xi, const double *br, const double *bi)
{
    cplx coef[27];
    long idx[27];
    const int cnt = row_of(adj, wnc, wnr, wni, hc, ax, n, p, coef, idx);
    const long ip = p[0] + p[1] * (long)n[0] + p[2] * (long)n[0] * n[1];
    cplx atx = 0;
    double nrm = 0;
    for (int t = 0; t < cnt; t++) {
        atx += coef[t] * (xr[idx[t]] + xi[idx[t]] * I);
        nrm += creal(conj(coef[t]) * coef[t]);
    }
    cplx d = omega * ((br[ip] + (bi ? bi[ip] : 0.0) * I) - atx) / nrm;
    for (int t = 0; t < cnt; t++) {
        cplx u = d * conj(coef[t]);
        xr[idx[t]] += creal(u);
        xi[idx[t]] += cimag(u);
    }
}

/* variant: bit 0 adjoint, bit 2 complex wn.  x (xr, xi) is swept in place. */
static void kacz_sweep(int order_mc, int variant, const double *wnr, const double *wni, const double *h,
                       const double *nd, const double *npml, double omega, double *xr, double *xi,
                       const double *br, const double *bi)
{
    const int adj = variant & 1, wnc = (variant & 4) != 0;
    const int n[3] = {(int)nd[0], (int)nd[1], (int)nd[2]};
    const hconst hc = make_hconst(h);
    axis_tab ax[3];
    ax[0] = make_axis(n[0], (int)npml[0], (int)npml[1]);
    ax[1] = make_axis(n[1], (int)npml[2], (int)npml[3]);
    ax[2] = make_axis(n[2], (int)npml[4], (int)npml[5]);
    for (int pass = 0; pass < 2; pass++) {
        if (!order_mc) {
            for (long t = 0; t < (long)n[0] * n[1] * n[2]; t++) {
                const long s = pass == 0 ? t : (long)n[0] * n[1] * n[2] - 1 - t;
                const int p[3] = {(int)(s % n[0]), (int)((s / n[0]) % n[1]), (int)(s / ((long)n[0] * n[1]))};
                project(adj, wnc, wnr, wni, &hc, ax, n, p, omega, xr, xi, br, bi);
            }
        } else {
            for (int cc = 0; cc < 27; cc++) {
                const int c = pass == 0 ? cc : 26 - cc;
                const int ci = c % 3, cj = (c / 3) % 3, ck = c / 9;
                for (int k = ck; k < n[2]; k += 3)
                    for (int j = cj; j < n[1]; j += 3)
                        for (int i = ci; i < n[0]; i += 3) {
                            const int p[3] = {i, j, k};
                            project(adj, wnc, wnr, wni, &hc, ax, n, p, omega, xr, xi, br, bi);
                        }
            }
        }
    }
    for (int a = 0; a < 3; a++) free_axis(&ax[a]);
}

void oracle_kacz_sweep_lex(int variant, const double *wnr, const double *wni, const double *h, const double *n,
                           const double *npml, double omega, double *xr, double *xi, const double *br,
                           const double *bi)
{
    kacz_sweep(0, variant, wnr, wni, h, n, npml, omega, xr, xi, br, bi);
}

void oracle_kacz_sweep_mc(int variant, const double *wnr, const double *wni, const double *h, const double *n,
                          const double *npml, double omega, double *xr, double *xi, const double *br,
                          const double *bi)
{
    kacz_sweep(1, variant, wnr, wni, h, n, npml, omega, xr, xi, br, bi);
}

/* Row p of H or H^H as (count, coefficients interleaved re/im, flat indices). */
int oracle_row(int variant, const double *wnr, const double *wni, const double *h, const double *nd,
               const double *npml, int i, int j, int k, double *coef_out, long *idx_out)
{
    const int n[3] = {(int)nd[0], (int)nd[1], (int)nd[2]};
    const hconst hc = make_hconst(h);
    axis_tab ax[3];
    ax[0] = make_axis(n[0], (int)npml[0], (int)npml[1]);
    ax[1] = make_axis(n[1], (int)npml[2], (int)npml[3]);
    ax[2] = make_axis(n[2], (int)npml[4], (int)npml[5]);
    cplx coef[27];
    const int p[3] = {i, j, k};
    const int cnt = row_of(variant & 1, (variant & 4) != 0, wnr, wni, &hc, ax, n, p, coef, idx_out);
    for (int t = 0; t < cnt; t++) {
        coef_out[2 * t] = creal(coef[t]);
        coef_out[2 * t + 1] = cimag(coef[t]);
    }
    for (int a = 0; a < 3; a++) free_axis(&ax[a]);
    return cnt;
}
