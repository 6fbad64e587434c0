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

template <bool ADJ>
__global__ __launch_bounds__(KB) void kacz_color_kernel(OpDesc<double> op, double omega, cpx<double>* __restrict__ x,
                                                         const cpx<double>* __restrict__ b, int ci, int cj, int ck,
                                                         int mx, int my, int mz)
{
    const long t = long(blockIdx.x) * KB + threadIdx.x;
    const long rows = long(mx) * my * mz;
    if (t >= rows) return;
    const int ti = int(t % mx), tj = int((t / mx) % my), tk = int(t / (long(mx) * my));
    const int i = ci + 3 * ti, j = cj + 3 * tj, k = ck + 3 * tk;
    const int nx = op.nx, ny = op.ny, nz = op.nz;
    const long sy = nx, sz = long(nx) * ny;
    const long ip = i + j * sy + k * sz;
    cpx<double> lo[3], hi[3];
    if constexpr (!ADJ) {
        lo[0] = tab(op.xlo, i + 1); hi[0] = tab(op.xhi, i + 1);
        lo[1] = tab(op.ylo, j + 1); hi[1] = tab(op.yhi, j + 1);
        lo[2] = tab(op.zlo, k + 1); hi[2] = tab(op.zhi, k + 1);
    }
    const cpx<double> wp = wn_at(op, ip);
    // coefficient of neighbour (dx,dy,dz) and its index (zero outside the grid)
    auto coef_at = [&](int dx, int dy, int dz, long& q) -> cpx<double> {
        const int a = i + dx, bb = j + dy, c = k + dz;
        const bool in = a >= 0 && a < nx && bb >= 0 && bb < ny && c >= 0 && c < nz;
        q = in ? a + bb * sy + c * sz : ip;
        if (!in) return cpx<double>{0.0, 0.0};
        const int nnz = (dx != 0) + (dy != 0) + (dz != 0);
        if constexpr (!ADJ) {
            return cmul(cpx<double>{mass_mu(op, nnz), 0.0}, wn_at(op, q)) + stiff(op, dx, dy, dz, lo, hi);
        } else {
            // column p of H at row q, conjugated: xi of q, wn of p
            cpx<double> ql[3], qh[3];
            ql[0] = tab(op.xlo, a + 1); qh[0] = tab(op.xhi, a + 1);
            ql[1] = tab(op.ylo, bb + 1); qh[1] = tab(op.yhi, bb + 1);
            ql[2] = tab(op.zlo, c + 1); qh[2] = tab(op.zhi, c + 1);
            return conjc(mass_mu(op, nnz) * wp + stiff(op, -dx, -dy, -dz, ql, qh));
        }
    };
    // residual of the row and |a|^2 (neighbours in the order dz, dy, dx)
    double atr = 0.0, ati = 0.0, nrm = 0.0;
#pragma unroll 1
    for (int dz = -1; dz <= 1; ++dz)
#pragma unroll 1
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                long q;
                const cpx<double> cf = coef_at(dx, dy, dz, q);
                const cpx<double> xv = x[q];
                atr += cf.re * xv.re - cf.im * xv.im;
                ati += cf.re * xv.im + cf.im * xv.re;
                nrm += cf.re * cf.re + cf.im * cf.im;
            }
    const cpx<double> bp = b ? b[ip] : cpx<double>{0.0, 0.0};
    const double dr = omega * (bp.re - atr) / nrm, di = omega * (bp.im - ati) / nrm;
    // x_q += d conj(a_q), coefficients formed again (registers over ALU)
#pragma unroll 1
    for (int dz = -1; dz <= 1; ++dz)
#pragma unroll 1
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                long q;
                const cpx<double> cf = coef_at(dx, dy, dz, q);
                if (cf.re != 0.0 || cf.im != 0.0) {
                    cpx<double> v = x[q];
                    v.re += dr * cf.re + di * cf.im;
                    v.im += di * cf.re - dr * cf.im;
                    x[q] = v;
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
            const long rows = long(mx) * my * mz;
            if (rows <= 0) continue;
            const dim3 grid(unsigned((rows + KB - 1) / KB));
            if (adj)
                hipLaunchKernelGGL(kacz_color_kernel<true>, grid, dim3(KB), 0, s, op, omega, x, b, ci, cj, ck, mx, my,
                                   mz);
            else
                hipLaunchKernelGGL(kacz_color_kernel<false>, grid, dim3(KB), 0, s, op, omega, x, b, ci, cj, ck, mx,
                                   my, mz);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

}  // namespace helm3d
