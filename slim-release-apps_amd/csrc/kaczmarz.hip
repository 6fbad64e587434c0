// Kaczmarz row projections of the 27-point operator on gfx950.
//
// Reference semantics: do_sweep of Helm3dkaczswp_mex.c:60-117 -- for every
// row p of A (A = H, or H^H for the adjoint sweep),
//   d = omega (b_p - a_p . x) / |a_p|^2,   x_q += d conj(a_pq)  for the 27 q,
// visiting the rows forward and then backward (xyzloop_updown,
// Helm3d_27pt.h:63-108).  The lexicographic order is sequential; here the
// rows are ordered by 27 colours c = (i mod 3) + 3 (j mod 3) + 9 (k mod 3):
// rows of one colour touch disjoint 3x3x3 blocks, so one launch projects all
// of them at once, and colours 0..26 followed by 26..0 form the symmetric
// double sweep (its linear part Q keeps I - Q Hermitian, as CGMN requires).
// The iterates differ from the lexicographic sweep; the fixed point (A x = b)
// and the CGMN solution do not.
//
// Row coefficients are formed on the fly from the per-axis PML tables and wn
// (Helm3d_27pt.h:680-714 forward, :718-822 adjoint).  Complex wn is honoured
// (the reference's complex-wavenumber sweep gateway is built with a macro the
// header does not test and falls back to real wn; not reproduced).
#include "helm27.hpp"

namespace helm3d {

namespace {

constexpr int KB = 256;

__device__ inline cpx<double> tab(const cpx<double>* t, int slot) { return t[slot]; }

// Stiffness coefficient coupling a point with PML values (lo, hi per axis)
// to its neighbour at offset d (Helm3d_27pt.h:680-714, non-mass terms):
//   centre -(c_x xi_x + c_y xi_y + c_z xi_z); face a: c_a xi_a^s + sum_{b!=a} w_ab xi_b;
//   edge (a,b): -w_ab (xi_a^s + xi_b^s) + w3a2 xi_c; corner -w3a2 sum xi^s.
__device__ inline cpx<double> stiff(const OpDesc<double>& op, int dx, int dy, int dz, const cpx<double> lo[3],
                                    const cpx<double> hi[3])
{
    const int d[3] = {dx, dy, dz};
    cpx<double> full[3], side[3];
    int nnz = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        full[a] = lo[a] + hi[a];
        side[a] = d[a] < 0 ? lo[a] : hi[a];
        nnz += d[a] != 0;
    }
    const double w[3][3] = {{0.0, op.xy, op.xz}, {op.xy, 0.0, op.yz}, {op.xz, op.yz, 0.0}};
    cpx<double> K{0.0, 0.0};
    if (nnz == 0) {
        K = -(op.c[0] * full[0] + op.c[1] * full[1] + op.c[2] * full[2]);
    } else if (nnz == 1) {
        const int a = d[0] ? 0 : (d[1] ? 1 : 2);
        K = op.c[a] * side[a];
#pragma unroll
        for (int b = 0; b < 3; ++b)
            if (b != a) K = rfma(K, w[a][b], full[b]);
    } else if (nnz == 2) {
        const int c = !d[0] ? 0 : (!d[1] ? 1 : 2);
        const int a = c == 0 ? 1 : 0, b = c == 2 ? 1 : 2;
        K = (-w[a][b]) * (side[a] + side[b]);
        K = rfma(K, op.w3a2, full[c]);
    } else {
        K = (-op.w3a2) * (side[0] + side[1] + side[2]);
    }
    return K;
}

__device__ inline double mass_mu(const OpDesc<double>& op, int nnz)
{
    return nnz == 0 ? op.wn_coef : (nnz == 1 ? -op.wm2 : (nnz == 2 ? -op.wm3 : -op.wm4));
}

__device__ inline cpx<double> wn_at(const OpDesc<double>& op, long q)
{
    if (op.wnc) return op.wnc[q];
    return cpx<double>{op.wn ? op.wn[q] : op.wn_const, 0.0};
}

// One thread per row of colour (ci, cj, ck) on a 3-D grid of 64 x 4 blocks
// (row index x along the lanes, one y row per wave, one z plane per grid
// layer).  The 27 coefficients and neighbour values are formed once and
// kept in registers: the 27 loads of a row issue together, and the update
// writes back x_q + d conj(a_pq) from the values already read (rows of a
// colour have disjoint supports, so no other row touches them).
template <bool ADJ>
__global__ __launch_bounds__(KB) void kacz_color_kernel(OpDesc<double> op, double omega, cpx<double>* __restrict__ x,
                                                         const cpx<double>* __restrict__ b, int ci, int cj, int ck,
                                                         int mx, int my)
{
    const int ti = int(blockIdx.x) * 64 + int(threadIdx.x & 63);
    const int tj = __builtin_amdgcn_readfirstlane(int(blockIdx.y) * 4 + int(threadIdx.x >> 6));
    const int tk = blockIdx.z;
    if (tj >= my) return;
    const bool act = ti < mx;  // lanes past the last row compute a copy and write nothing
    const int i = ci + 3 * min(ti, mx - 1), j = cj + 3 * tj, k = ck + 3 * tk;
    const int nx = op.nx, ny = op.ny, nz = op.nz;
    const long sy = nx, sz = long(nx) * ny;
    const long ip = i + j * sy + k * sz;
    // PML values at coordinate + d (d = -1, 0, 1) per axis; the forward row
    // uses only its own (d = 0), the adjoint row its neighbours'
    cpx<double> lo[3][3], hi[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (!ADJ && d != 1) continue;
        lo[0][d] = tab(op.xlo, i + d); hi[0][d] = tab(op.xhi, i + d);
        lo[1][d] = tab(op.ylo, j + d); hi[1][d] = tab(op.yhi, j + d);
        lo[2][d] = tab(op.zlo, k + d); hi[2][d] = tab(op.zhi, k + d);
    }
    const cpx<double> wp = wn_at(op, ip);
    cpx<double> cf[27], xv[27];
    long qn[27];
    bool in[27];
#pragma unroll
    for (int n = 0; n < 27; ++n) {
        const int dx = n % 3 - 1, dy = (n / 3) % 3 - 1, dz = n / 9 - 1;
        const int a = i + dx, bb = j + dy, c = k + dz;
        in[n] = a >= 0 && a < nx && bb >= 0 && bb < ny && c >= 0 && c < nz;
        qn[n] = in[n] ? a + bb * sy + c * sz : ip;
        const int nnz = (dx != 0) + (dy != 0) + (dz != 0);
        cpx<double> v;
        if constexpr (!ADJ) {
            const cpx<double> l3[3] = {lo[0][1], lo[1][1], lo[2][1]}, h3[3] = {hi[0][1], hi[1][1], hi[2][1]};
            v = cmul(cpx<double>{mass_mu(op, nnz), 0.0}, wn_at(op, qn[n])) + stiff(op, dx, dy, dz, l3, h3);
        } else {
            // column p of H at row q, conjugated: xi of q, wn of p
            const cpx<double> l3[3] = {lo[0][dx + 1], lo[1][dy + 1], lo[2][dz + 1]};
            const cpx<double> h3[3] = {hi[0][dx + 1], hi[1][dy + 1], hi[2][dz + 1]};
            v = conjc(mass_mu(op, nnz) * wp + stiff(op, -dx, -dy, -dz, l3, h3));
        }
        cf[n] = in[n] ? v : cpx<double>{0.0, 0.0};
        xv[n] = x[qn[n]];
    }
    // residual of the row and |a|^2 (neighbours in the order dz, dy, dx)
    double atr = 0.0, ati = 0.0, nrm = 0.0;
#pragma unroll
    for (int n = 0; n < 27; ++n) {
        atr += cf[n].re * xv[n].re - cf[n].im * xv[n].im;
        ati += cf[n].re * xv[n].im + cf[n].im * xv[n].re;
        nrm += cf[n].re * cf[n].re + cf[n].im * cf[n].im;
    }
    const cpx<double> bp = b ? b[ip] : cpx<double>{0.0, 0.0};
    const double dr = omega * (bp.re - atr) / nrm, di = omega * (bp.im - ati) / nrm;
    // x_q += d conj(a_q)
#pragma unroll
    for (int n = 0; n < 27; ++n) {
        if (act && in[n] && (cf[n].re != 0.0 || cf[n].im != 0.0)) {
            cpx<double> v = xv[n];
            v.re += dr * cf[n].re + di * cf[n].im;
            v.im += di * cf[n].re - dr * cf[n].im;
            x[qn[n]] = v;
        }
    }
}

}  // namespace

hipError_t launch_kacz_sweep(bool adj, const OpDesc<double>& op, double omega, cpx<double>* x,
                             const cpx<double>* b, hipStream_t s)
{
    for (int pass = 0; pass < 2; ++pass) {
        for (int cc = 0; cc < 27; ++cc) {
            const int c = pass == 0 ? cc : 26 - cc;
            const int ci = c % 3, cj = (c / 3) % 3, ck = c / 9;
            const int mx = (op.nx - ci + 2) / 3, my = (op.ny - cj + 2) / 3, mz = (op.nz - ck + 2) / 3;
            if (mx <= 0 || my <= 0 || mz <= 0) continue;
            const dim3 grid(unsigned((mx + 63) / 64), unsigned((my + 3) / 4), unsigned(mz));
            if (adj)
                hipLaunchKernelGGL(kacz_color_kernel<true>, grid, dim3(KB), 0, s, op, omega, x, b, ci, cj, ck, mx, my);
            else
                hipLaunchKernelGGL(kacz_color_kernel<false>, grid, dim3(KB), 0, s, op, omega, x, b, ci, cj, ck, mx, my);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

}  // namespace helm3d
