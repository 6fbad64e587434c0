// 27-point Helmholtz stencil with PML on gfx950: forward, adjoint, derivative
// and complex-wavenumber forms in one z-marching kernel template.
//
// Reference semantics: do_Hmvp of SLIM-release-apps
// tools/algorithms/3DFreqModeling/Helm3d_27pt.h:827-875 with the coefficients
// of :680-714 (forward) and :718-822 (adjoint), the DERIV / WNCMPLX variants of
// :4-16, 202-206, and zero padding outside the grid (:263-317).
//
// Formulation.  Every coefficient is linear in the PML values of one point, so
// the row of H at p = (i,j,k) splits into
//   * per-column complex coefficient sets F (in-plane neighbours) and G
//     (neighbours in planes k+-1) that carry the x/y PML of the column and are
//     fixed for the whole z-march,
//   * a real 9-point in-plane stencil Qz that carries the z PML, weighted per
//     plane by xi_z^lo, xi_z^hi, xi_z (uniform across the block),
//   * the mass-lumping stencil (wn_coef, -WM2, -WM3, -WM4 by |d|_1) acting on
//     z = wn .* x (forward) or on x with conj(wn) at the centre (adjoint).
// The adjoint is the exact conjugate transpose: F/G take the neighbour's x/y
// PML values and the z weights come from the neighbouring planes.
//
// Tiling.  A 256-thread block owns a 64 x 4 column tile and marches over a z
// chunk.  Each plane (with a one-point x/y halo) is staged once in LDS
// (double-buffered, one barrier per plane); the next plane's global loads are
// issued before the barrier so HBM latency overlaps the current plane.  Block
// ids are remapped so consecutive tiles of the same chunk land on the same XCD
// and share halo lines in its L2.
#include "helm27.hpp"

namespace helm3d {

namespace {

constexpr int TX = 64;
constexpr int TY = 4;
constexpr int NT = TX * TY;
constexpr int HX = TX + 2;
constexpr int HY = TY + 2;
constexpr int NS = HX * HY;
static_assert(NS <= 2 * NT, "halo tile must fit two load slots per thread");

template <typename R>
__device__ inline cpx<R> zero() { return cpx<R>{R(0), R(0)}; }

// Per-column coefficient sets.  For neighbour offset (dx,dy) (a = dx+1,
// b = dy+1) with side values xs[a], ys[b] (xs[1], ys[1] = full xi):
//   F (dz = 0):  centre -(c_x xi_x + c_y xi_y); x faces c_x xi_x^s + w_xy xi_y;
//                y faces c_y xi_y^s + w_xy xi_x; diagonals -w_xy (xi_x^s + xi_y^s)
//   G (dz=+-1):  centre w_xz xi_x + w_yz xi_y; x faces -w_xz xi_x^s + w3a2 xi_y;
//                y faces -w_yz xi_y^s + w3a2 xi_x; diagonals -w3a2 (xi_x^s + xi_y^s)
// i.e. the non-xi_z part of the coefficients of Helm3d_27pt.h:680-714.
template <typename R>
__device__ inline void column_coefs(const OpDesc<R>& op, const cpx<R> xs[3], const cpx<R> ys[3],
                                    cpx<R> F[3][3], cpx<R> G[3][3])
{
    const cpx<R> xf = xs[1], yf = ys[1];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            if (a == 1 && b == 1) {
                F[a][b] = -(op.c[0] * xf + op.c[1] * yf);
                G[a][b] = op.xz * xf + op.yz * yf;
            } else if (b == 1) {
                F[a][b] = op.c[0] * xs[a] + op.xy * yf;
                G[a][b] = (-op.xz) * xs[a] + op.w3a2 * yf;
            } else if (a == 1) {
                F[a][b] = op.c[1] * ys[b] + op.xy * xf;
                G[a][b] = (-op.yz) * ys[b] + op.w3a2 * xf;
            } else {
                F[a][b] = (-op.xy) * (xs[a] + ys[b]);
                G[a][b] = (-op.w3a2) * (xs[a] + ys[b]);
            }
        }
    }
}

template <typename R, bool ADJ, bool DERIV, bool WNC>
__global__ __launch_bounds__(NT) void helm27_kernel(StencilLaunch<R> L)
{
    const OpDesc<R>& op = L.op;
    __shared__ cpx<R> su[2][HY][HX];            // input vector plane (+halo)
    __shared__ cpx<R> sm[ADJ ? 1 : 2][HY][HX];  // forward: wn .* x plane (+halo)

    const int nx = op.nx, ny = op.ny, nz = op.nz;

    // XCD-aware block remap: the first 8*floor(G/8) blocks are dealt so that
    // each XCD (round-robin by block id) receives a contiguous logical range.
    const int G = gridDim.x, lin = blockIdx.x;
    const int G8 = G & ~7;
    const int logical = lin < G8 ? (lin & 7) * (G8 >> 3) + (lin >> 3) : lin;
    const int ntile = L.tiles_x * L.tiles_y;
    const int chunk = logical / ntile;
    const int tile = logical - chunk * ntile;
    const int tyt = tile / L.tiles_x;
    const int i0 = (tile - tyt * L.tiles_x) * TX;
    const int j0 = tyt * TY;
    const int k0 = chunk * L.kchunk;
    const int k1 = min(nz, k0 + L.kchunk);

    const int tid = threadIdx.x;
    const int tx = tid & (TX - 1), ty = tid / TX;
    const int i = i0 + tx, j = j0 + ty;
    const bool active = (i < nx) && (j < ny);
    const int ic = min(i, nx - 1), jc = min(j, ny - 1);  // clamped for table reads

    // ---- per-column coefficients (fixed for the whole march)
    cpx<R> F[3][3], Gc[3][3];
    if constexpr (!DERIV) {
        cpx<R> xs[3], ys[3];
        if constexpr (!ADJ) {
            xs[0] = op.xlo[ic + 1]; xs[2] = op.xhi[ic + 1]; xs[1] = xs[0] + xs[2];
            ys[0] = op.ylo[jc + 1]; ys[2] = op.yhi[jc + 1]; ys[1] = ys[0] + ys[2];
        } else {
            // neighbour i-1 couples through its hi value, i+1 through its lo value
            xs[0] = op.xhi[ic]; xs[2] = op.xlo[ic + 2]; xs[1] = op.xlo[ic + 1] + op.xhi[ic + 1];
            ys[0] = op.yhi[jc]; ys[2] = op.ylo[jc + 2]; ys[1] = op.ylo[jc + 1] + op.yhi[jc + 1];
        }
        column_coefs(op, xs, ys, F, Gc);
        if constexpr (ADJ) {
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) { F[a][b] = conjc(F[a][b]); Gc[a][b] = conjc(Gc[a][b]); }
        }
    }

    // ---- load slots: the (TX+2) x (TY+2) halo tile, two slots per thread
    const long plane = long(nx) * ny;
    const int s0 = tid, s1 = tid + NT;
    const int s0x = s0 % HX, s0y = s0 / HX;
    const int s1x = s1 % HX, s1y = s1 / HX;
    const bool has1 = s1 < NS;
    const int g0i = i0 - 1 + s0x, g0j = j0 - 1 + s0y;
    const int g1i = i0 - 1 + s1x, g1j = j0 - 1 + s1y;
    const bool in0 = g0i >= 0 && g0i < nx && g0j >= 0 && g0j < ny;
    const bool in1 = has1 && g1i >= 0 && g1i < nx && g1j >= 0 && g1j < ny;
    const long off0 = in0 ? g0i + long(g0j) * nx : 0;
    const long off1 = in1 ? g1i + long(g1j) * nx : 0;
    const long offc = long(ic) + long(jc) * nx;

    cpx<R> px0 = zero<R>(), px1 = zero<R>();
    cpx<R> pw0 = zero<R>(), pw1 = zero<R>();  // wn at the slots (forward), re only if real
    cpx<R> pwc = zero<R>();                    // adjoint: wn at own column, plane being output next
    double nrm = 0.0;                          // ||y||^2 of this thread's outputs

    auto load_wn = [&](long idx) -> cpx<R> {
        if constexpr (WNC) return op.wnc[idx];
        else return cpx<R>{op.wn ? op.wn[idx] : op.wn_const, R(0)};
    };
    // store y = alpha*out + beta*b, or run the gradient epilogue
    auto emit = [&](int kc, cpx<R> out) {
        const long idx = long(kc) * plane + offc;
        if (L.gg) {
            const int qi = i - L.p0[0], qj = j - L.p0[1], qk = kc - L.p0[2];
            if (qi >= 0 && qi < L.pn[0] && qj >= 0 && qj < L.pn[1] && qk >= 0 && qk < L.pn[2]) {
                const cpx<R> u = L.gu[idx];
                const double v = double(u.re) * double(out.re) + double(u.im) * double(out.im);
                const long g = qi + long(qj) * L.pn[0] + long(qk) * L.pn[0] * L.pn[1];
                L.gg[g] += L.gscale * v;
            }
            return;
        }
        cpx<R> r = L.alpha * out;
        if (L.b) r = rfma(r, L.beta, L.b[idx]);
        L.y[idx] = r;
        nrm += double(r.re) * double(r.re) + double(r.im) * double(r.im);
    };
    auto fetch = [&](int kp) {
        const bool inz = kp >= 0 && kp < nz;
        const long base = long(kp) * plane;
        px0 = (inz && in0) ? L.x[base + off0] : zero<R>();
        px1 = (inz && in1) ? L.x[base + off1] : zero<R>();
        if constexpr (!ADJ) {
            pw0 = (inz && in0) ? load_wn(base + off0) : zero<R>();
            pw1 = (inz && in1) ? load_wn(base + off1) : zero<R>();
        }
    };

    // pipeline state (see the header comment of this file)
    cpx<R> Qprev = zero<R>(), Qcur = zero<R>(), Acc = zero<R>(), Pend = zero<R>();
    cpx<R> MAcc = zero<R>(), MOcur = zero<R>();

    fetch(k0 - 1);
    int t = 0;
    for (int kp = k0 - 1; kp <= k1; ++kp, ++t) {
        const int buf = t & 1;
        // stage plane kp
        su[buf][s0y][s0x] = px0;
        if constexpr (!ADJ) sm[buf][s0y][s0x] = WNC ? cmul(pw0, px0) : pw0.re * px0;
        if (has1) {
            su[buf][s1y][s1x] = px1;
            if constexpr (!ADJ) sm[buf][s1y][s1x] = WNC ? cmul(pw1, px1) : pw1.re * px1;
        }
        if constexpr (ADJ) {
            // wn at own column of plane kp (that plane is output next iteration)
            pwc = (kp >= 0 && kp < nz) ? load_wn(long(kp) * plane + offc) : zero<R>();
        }
        if (kp < k1) fetch(kp + 1);
        __syncthreads();

        // ---- plane quantities from the 3x3 neighbourhood
        cpx<R> u[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) u[a][b] = su[buf][ty + b][tx + a];

        cpx<R> Qn = zero<R>(), Fn = zero<R>(), Gn = zero<R>();
        if constexpr (!DERIV) {
            const cpx<R> sxp = u[0][1] + u[2][1], syp = u[1][0] + u[1][2];
            const cpx<R> dg = (u[0][0] + u[2][0]) + (u[0][2] + u[2][2]);
            Qn = op.c[2] * u[1][1];
            Qn = rfma(Qn, -op.xz, sxp);
            Qn = rfma(Qn, -op.yz, syp);
            Qn = rfma(Qn, -op.w3a2, dg);
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    Fn = cfma(Fn, F[a][b], u[a][b]);
                    Gn = cfma(Gn, Gc[a][b], u[a][b]);
                }
        }
        cpx<R> m[3][3];
        if constexpr (!ADJ) {
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) m[a][b] = sm[buf][ty + b][tx + a];
        } else {
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) m[a][b] = u[a][b];
        }
        const cpx<R> mf = (m[0][1] + m[2][1]) + (m[1][0] + m[1][2]);
        const cpx<R> md = (m[0][0] + m[2][0]) + (m[0][2] + m[2][2]);
        cpx<R> Mc = op.wn_coef * m[1][1];
        Mc = rfma(Mc, -op.wm2, mf);
        Mc = rfma(Mc, -op.wm3, md);
        cpx<R> Mo = (-op.wm2) * m[1][1];
        Mo = rfma(Mo, -op.wm3, mf);
        Mo = rfma(Mo, -op.wm4, md);

        if constexpr (!ADJ) {
            // y(kc) = wlo Q(kc-1) + whi Q(kc+1) - (wlo+whi) Q(kc) + GM(kc-1) + GM(kc+1) + FM(kc)
            const cpx<R> GMn = Gn + Mo;
            if (kp >= k0 + 1) {
                const int kc = kp - 1;
                cpx<R> out = Acc + GMn;
                if constexpr (!DERIV) {
                    const cpx<R> wl = op.zlo[kc + 1], wh = op.zhi[kc + 1];
                    out = cfma(out, wl, Qprev - Qcur);
                    out = cfma(out, wh, Qn - Qcur);
                }
                if (active) emit(kc, out);
            }
            Qprev = Qcur;
            Qcur = Qn;
            Acc = Pend + (Fn + Mc);
            Pend = GMn;
        } else {
            // contributions of plane kp to outputs kp+1 (Sdn), kp-1 (Sup), kp (S0)
            cpx<R> Sdn = Gn, Sup = Gn, S0 = Fn;
            if constexpr (!DERIV) {
                const cpx<R> zl = op.zlo[kp + 1], zh = op.zhi[kp + 1];
                Sdn = cfma(Sdn, conjc(zh), Qn);
                Sup = cfma(Sup, conjc(zl), Qn);
                S0 = cfma(S0, -conjc(zl + zh), Qn);
            }
            if (kp >= k0 + 1) {
                const int kc = kp - 1;
                cpx<R> out = Acc + Sup;
                out = cfma(out, conjc(Qprev), MAcc + Mo);  // Qprev holds wn(kc)
                if (active) emit(kc, out);
            }
            Acc = Pend + S0;
            MAcc = MOcur + Mc;
            Pend = Sdn;
            MOcur = Mo;
            Qprev = pwc;  // wn of plane kp, used when kp is output
        }
    }

    // optional per-block partial of ||y||^2 (deterministic two-stage reduction)
    if (L.norm_part) {
        for (int o = 32; o > 0; o >>= 1) nrm += __shfl_xor(nrm, o, 64);
        __shared__ double red[NT / 64];
        if ((tid & 63) == 0) red[tid >> 6] = nrm;
        __syncthreads();
        if (tid == 0) {
            double s = 0;
            for (int w = 0; w < NT / 64; ++w) s += red[w];
            L.norm_part[blockIdx.x] = s;
        }
    }
}

}  // namespace

template <typename R>
StencilGrid stencil_grid(int nx, int ny, int nz, int kchunk)
{
    StencilGrid g;
    g.tiles_x = (nx + TX - 1) / TX;
    g.tiles_y = (ny + TY - 1) / TY;
    if (kchunk <= 0) {
        // enough blocks to fill 256 CUs several times, chunks of at least 8 planes
        const long tiles = long(g.tiles_x) * g.tiles_y;
        int nch = int((2048 + tiles - 1) / tiles);
        nch = std::max(1, std::min(nch, (nz + 7) / 8));
        kchunk = (nz + nch - 1) / nch;
    }
    g.kchunk = std::max(1, std::min(kchunk, nz));
    g.nchunks = (nz + g.kchunk - 1) / g.kchunk;
    g.blocks = g.tiles_x * g.tiles_y * g.nchunks;
    return g;
}

template <typename R>
hipError_t launch_stencil(int variant, StencilLaunch<R> L, const StencilGrid& g, hipStream_t s)
{
    L.tiles_x = g.tiles_x;
    L.tiles_y = g.tiles_y;
    L.kchunk = g.kchunk;
    const dim3 grid(g.blocks), block(NT);
    const bool adj = variant & VAR_ADJ, der = variant & VAR_DERIV, wnc = variant & VAR_WNC;
    if (!adj && !der && !wnc) hipLaunchKernelGGL((helm27_kernel<R, false, false, false>), grid, block, 0, s, L);
    else if (!adj && !der && wnc) hipLaunchKernelGGL((helm27_kernel<R, false, false, true>), grid, block, 0, s, L);
    else if (!adj && der && !wnc) hipLaunchKernelGGL((helm27_kernel<R, false, true, false>), grid, block, 0, s, L);
    else if (!adj && der && wnc) hipLaunchKernelGGL((helm27_kernel<R, false, true, true>), grid, block, 0, s, L);
    else if (adj && !der && !wnc) hipLaunchKernelGGL((helm27_kernel<R, true, false, false>), grid, block, 0, s, L);
    else if (adj && !der && wnc) hipLaunchKernelGGL((helm27_kernel<R, true, false, true>), grid, block, 0, s, L);
    else if (adj && der && !wnc) hipLaunchKernelGGL((helm27_kernel<R, true, true, false>), grid, block, 0, s, L);
    else hipLaunchKernelGGL((helm27_kernel<R, true, true, true>), grid, block, 0, s, L);
    return hipGetLastError();
}

template StencilGrid stencil_grid<double>(int, int, int, int);
template StencilGrid stencil_grid<float>(int, int, int, int);
template hipError_t launch_stencil<double>(int, StencilLaunch<double>, const StencilGrid&, hipStream_t);
template hipError_t launch_stencil<float>(int, StencilLaunch<float>, const StencilGrid&, hipStream_t);

}  // namespace helm3d
