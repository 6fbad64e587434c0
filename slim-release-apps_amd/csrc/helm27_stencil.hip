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
#include <type_traits>

#include "helm27.hpp"

namespace helm3d {

namespace {

constexpr int TX = 64;  // one wavefront along x
constexpr int HX = TX + 2;
constexpr int kMaxChunk = 254;  // z planes per block (LDS holds their PML weights)

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

// TY: rows of the column tile (block = 64 x TY threads); PF: planes of global
// loads kept in flight ahead of the plane being computed (1 or 2).
//
// Memory pipeline: every global load of the loop goes into one of PF register
// sets that are refilled right after being staged into LDS, so no register is
// ever copied while its load is in flight and the only vmcnt waits are those
// of the set being staged (PF planes after it was requested).  The per-plane z
// PML weights are staged in LDS once per block, and the own-point operands of
// the epilogue (adjoint wn, addend b, gradient U and g) travel in the same set
// as the halo slots of the next plane.
// EPI: 0 stores y = alpha*H x, 1 stores y = alpha*H x + beta*b, 2 runs the
// gradient epilogue (accumulates Re(conj(U) . out) into g).
template <typename R, bool ADJ, bool DERIV, bool WNC, int TY, int PF, int EPI>
__global__ __launch_bounds__(TX* TY) void helm27_kernel(
    StencilLaunch<R> L)
{
    constexpr int NT = TX * TY;
    constexpr int HY = TY + 2;
    constexpr int NS = HX * HY;
    static_assert(NS <= 2 * NT, "halo tile must fit two load slots per thread");
    static_assert(PF == 1 || PF == 2, "prefetch depth 1 or 2");
    const OpDesc<R>& op = L.op;
    __shared__ cpx<R> su[2][HY][HX];            // input vector plane (+halo)
    __shared__ cpx<R> sm[ADJ ? 1 : 2][HY][HX];  // forward: wn .* x plane (+halo)
    __shared__ cpx<R> szl[kMaxChunk + 4], szh[kMaxChunk + 4];  // z PML slots k0 .. k1+3

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
    // A wave is one row j of 64 columns: it takes the interior path when every
    // column and the row have xi_lo = xi_hi = 1 at index-1, index, index+1.
    const int jw = __builtin_amdgcn_readfirstlane(j);
    const bool interior = i0 >= op.int_lo[0] && i0 + TX - 1 <= op.int_hi[0] && jw >= op.int_lo[1] &&
                          jw <= op.int_hi[1];

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
        // z weights of this chunk: slots k0 .. k1+1 (planes k0-1 .. k1) plus
        // two padding entries read by the trailing pipeline step
        for (int q = tid; q < k1 - k0 + 4; q += NT) {
            const int sl = min(k0 + q, nz + 1);
            szl[q] = op.zlo[sl];
            szh[q] = op.zhi[sl];
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
    // gradient epilogue: position of the own column in the physical box
    const int qi = i - L.p0[0], qj = j - L.p0[1];
    const bool gcol = EPI == 2 && active && qi >= 0 && qi < L.pn[0] && qj >= 0 && qj < L.pn[1];
    const long gcol_off = gcol ? qi + long(qj) * L.pn[0] : 0;
    const long gplane = long(L.pn[0]) * L.pn[1];
    const R* wnp = op.wn ? op.wn : reinterpret_cast<const R*>(L.x);  // real-wn loads (forward)

    double nrm = 0.0;  // ||y||^2 of this thread's outputs
    const int sink_slot = int((long(blockIdx.x) * NT + tid) % kSinkSlots);

    // One register set: halo slots of plane q and the own-point operands of
    // plane q-1 (the plane output when q is staged).  Every load is issued
    // unconditionally from a clamped, in-bounds address and validity is
    // applied when the set is consumed, so the number of loads in flight is
    // fixed and the compiler waits only for the set it stages.
    struct Set {
        cpx<R> x0, x1;  // slots
        R w0, w1;       // forward, real wn at the slots
        cpx<R> c0, c1;  // forward, complex wn at the slots
        cpx<R> wc;      // adjoint: wn at own point of plane q-1
        cpx<R> bb;      // addend b at own point of plane q-1
        cpx<R> gu;      // gradient: U at own point of plane q-1
        double gv;      // gradient: g at own point of plane q-1
    };
    auto fetch = [&](int q, Set& S) {
        const long base = long(min(max(q, 0), nz - 1)) * plane;
        S.x0 = L.x[base + off0];
        S.x1 = L.x[base + off1];
        if constexpr (!ADJ) {
            if constexpr (WNC) {
                S.c0 = op.wnc[base + off0];
                S.c1 = op.wnc[base + off1];
            } else {
                S.w0 = wnp[base + off0];
                S.w1 = wnp[base + off1];
            }
        }
        const long own = long(min(max(q - 1, 0), nz - 1)) * plane + offc;
        if constexpr (ADJ) {
            if constexpr (WNC) S.wc = op.wnc[own];
            else S.wc = cpx<R>{wnp[own], R(0)};
        }
        if constexpr (EPI == 1) S.bb = L.b[own];
        if constexpr (EPI == 2) {
            S.gu = L.gu[own];
            const int qk = min(max(q - 1 - L.p0[2], 0), L.pn[2] - 1);
            S.gv = L.gg[gcol_off + long(qk) * gplane];
        }
    };

    // pipeline state (see the header comment of this file)
    cpx<R> Qprev = zero<R>(), Qcur = zero<R>(), Acc = zero<R>(), Pend = zero<R>();
    cpx<R> MAcc = zero<R>(), MOcur = zero<R>();

    // Stage set S (plane q) into LDS, refill S with plane q+PF, compute plane
    // q and emit plane q-1.
    auto step = [&](int q, int t, Set& S) {
        const int buf = t & 1;
        const bool inz = q >= 0 && q < nz;
        const bool v0 = inz && in0, v1 = inz && in1;
        const cpx<R> x0 = v0 ? S.x0 : zero<R>(), x1 = v1 ? S.x1 : zero<R>();
        su[buf][s0y][s0x] = x0;
        if (has1) su[buf][s1y][s1x] = x1;
        if constexpr (!ADJ) {
            if constexpr (WNC) {
                sm[buf][s0y][s0x] = cmul(S.c0, x0);
                if (has1) sm[buf][s1y][s1x] = cmul(S.c1, x1);
            } else {
                const R w0 = op.wn ? S.w0 : op.wn_const, w1 = op.wn ? S.w1 : op.wn_const;
                sm[buf][s0y][s0x] = w0 * x0;
                if (has1) sm[buf][s1y][s1x] = w1 * x1;
            }
        }
        // own-point operands of plane q-1 leave the set before the refill
        cpx<R> wc = S.wc;
        if constexpr (ADJ && !WNC) wc.re = op.wn ? wc.re : op.wn_const;
        const cpx<R> bb = S.bb, gu = S.gu;
        const double gv = S.gv;
        fetch(q + PF, S);  // unconditional: a fixed number of loads per step
        __syncthreads();

        // ---- mass-lumping sums (forward: on wn .* x, adjoint: on x)
        cpx<R> Mc, Mo;
        {
            cpx<R> m[3][3];
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) m[a][b] = ADJ ? su[buf][ty + b][tx + a] : sm[buf][ty + b][tx + a];
            const cpx<R> mf = (m[0][1] + m[2][1]) + (m[1][0] + m[1][2]);
            const cpx<R> md = (m[0][0] + m[2][0]) + (m[0][2] + m[2][2]);
            Mc = op.wn_coef * m[1][1];
            Mc = rfma(Mc, -op.wm2, mf);
            Mc = rfma(Mc, -op.wm3, md);
            Mo = (-op.wm2) * m[1][1];
            Mo = rfma(Mo, -op.wm3, mf);
            Mo = rfma(Mo, -op.wm4, md);
        }

        // ---- stiffness: Qz (z PML, applied per plane) and F, G (x/y PML)
        cpx<R> Qn = zero<R>(), Fn = zero<R>(), Gn = zero<R>();
        if constexpr (!DERIV) {
            cpx<R> u[3][3];
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) u[a][b] = su[buf][ty + b][tx + a];
            // pair sums along y (per column a) and along x (per row b)
            const cpx<R> sy0 = u[0][0] + u[0][2], sy1 = u[1][0] + u[1][2], sy2 = u[2][0] + u[2][2];
            const cpx<R> sx0 = u[0][0] + u[2][0], sx1 = u[0][1] + u[2][1], sx2 = u[0][2] + u[2][2];
            const cpx<R> dg = sy0 + sy2;
            Qn = op.c[2] * u[1][1];
            Qn = rfma(Qn, -op.xz, sx1);
            Qn = rfma(Qn, -op.yz, sy1);
            Qn = rfma(Qn, -op.w3a2, dg);
            if (interior) {
                // x/y PML absent around this wave: F, G are real constant
                // stencils on the sums already formed for Qz
                Fn = op.fi[0] * u[1][1];
                Fn = rfma(Fn, op.fi[1], sx1);
                Fn = rfma(Fn, op.fi[2], sy1);
                Fn = rfma(Fn, op.fi[3], dg);
                Gn = op.gi[0] * u[1][1];
                Gn = rfma(Gn, op.gi[1], sx1);
                Gn = rfma(Gn, op.gi[2], sy1);
                Gn = rfma(Gn, op.gi[3], dg);
            } else {
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int b = 0; b < 3; ++b) {
                        Fn = cfma(Fn, F[a][b], u[a][b]);
                        Gn = cfma(Gn, Gc[a][b], u[a][b]);
                    }
            }
        }

        cpx<R> out = zero<R>();
        const bool emit = q >= k0 + 1 && q <= k1;
        const int kc = q - 1;
        if constexpr (!ADJ) {
            // y(kc) = wlo Q(kc-1) + whi Q(kc+1) - (wlo+whi) Q(kc) + GM(kc-1) + GM(kc+1) + FM(kc)
            const cpx<R> GMn = Gn + Mo;
            if (emit) {
                out = Acc + GMn;
                if constexpr (!DERIV) {
                    const cpx<R> wl = szl[kc + 1 - k0], wh = szh[kc + 1 - k0];
                    out = cfma(out, wl, Qprev - Qcur);
                    out = cfma(out, wh, Qn - Qcur);
                }
            }
            Qprev = Qcur;
            Qcur = Qn;
            Acc = Pend + (Fn + Mc);
            Pend = GMn;
        } else {
            // contributions of plane q to outputs q+1 (Sdn), q-1 (Sup), q (S0)
            cpx<R> Sdn = Gn, Sup = Gn, S0 = Fn;
            if constexpr (!DERIV) {
                const cpx<R> zl = szl[q + 1 - k0], zh = szh[q + 1 - k0];
                Sdn = cfma(Sdn, conjc(zh), Qn);
                Sup = cfma(Sup, conjc(zl), Qn);
                S0 = cfma(S0, -conjc(zl + zh), Qn);
            }
            if (emit) {
                out = Acc + Sup;
                out = cfma(out, conjc(wc), MAcc + Mo);
            }
            Acc = Pend + S0;
            MAcc = MOcur + Mc;
            Pend = Sdn;
            MOcur = Mo;
        }
        // Exactly one store per step and lane: steps or lanes without an
        // output write to the sink instead (keeps the vmcnt bookkeeping exact).
        const bool live = emit && active;
        if constexpr (EPI == 2) {
            const int qk = kc - L.p0[2];
            const bool g = live && gcol && qk >= 0 && qk < L.pn[2];
            const double v = double(gu.re) * double(out.re) + double(gu.im) * double(out.im);
            double* dst = g ? &L.gg[gcol_off + long(qk) * gplane] : reinterpret_cast<double*>(&L.sink[sink_slot]);
            *dst = gv + L.gscale * v;
        } else {
            cpx<R> r = L.alpha * out;
            if constexpr (EPI == 1) r = rfma(r, L.beta, bb);
            cpx<R>* dst = live ? &L.y[long(kc) * plane + offc] : &L.sink[sink_slot];
            *dst = r;
            if (live) nrm += double(r.re) * double(r.re) + double(r.im) * double(r.im);
        }
    };

    // Steps for planes k0-1 .. k1 (rounded up to a multiple of PF; a trailing
    // padding step computes but emits nothing).  No memory operation in the
    // loop body is conditional on the step, so the compiler can keep PF sets
    // of loads in flight across the back-edge.
    Set A{}, B{};
    fetch(k0 - 1, A);
    if constexpr (PF == 2) fetch(k0, B);
    const int ntrips = (k1 - k0 + 2 + PF - 1) / PF;
    for (int tr = 0; tr < ntrips; ++tr) {
        const int q = k0 - 1 + tr * PF;
        step(q, tr * PF, A);
        if constexpr (PF == 2) step(q + 1, tr * PF + 1, B);
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
StencilGrid stencil_grid(int nx, int ny, int nz, int kchunk, int ty, int pf)
{
    StencilGrid g;
    g.ty = (ty == 8) ? 8 : 4;
    g.pf = (pf == 2) ? 2 : 1;
    g.tiles_x = (nx + TX - 1) / TX;
    g.tiles_y = (ny + g.ty - 1) / g.ty;
    if (kchunk <= 0) {
        // about two blocks per CU (256 CUs) with long z marches: measured best
        // on 240x240x140 (fewer, longer marches amortise the two pipeline
        // fill planes and keep halo planes in L2); chunks of >= 8 planes
        const long tiles = long(g.tiles_x) * g.tiles_y;
        int nch = int((480 + tiles - 1) / tiles);
        nch = std::max(1, std::min(nch, (nz + 7) / 8));
        kchunk = (nz + nch - 1) / nch;
    }
    g.kchunk = std::max(1, std::min(std::min(kchunk, nz), kMaxChunk));
    g.nchunks = (nz + g.kchunk - 1) / g.kchunk;
    g.blocks = g.tiles_x * g.tiles_y * g.nchunks;
    return g;
}

namespace {
template <typename R, bool A, bool D, bool W, int EPI>
hipError_t launch_shape(const StencilLaunch<R>& L, const StencilGrid& g, hipStream_t s)
{
    const dim3 grid(g.blocks);
    if (g.ty == 8) {
        if (g.pf == 2) hipLaunchKernelGGL((helm27_kernel<R, A, D, W, 8, 2, EPI>), grid, dim3(TX * 8), 0, s, L);
        else hipLaunchKernelGGL((helm27_kernel<R, A, D, W, 8, 1, EPI>), grid, dim3(TX * 8), 0, s, L);
    } else {
        if (g.pf == 2) hipLaunchKernelGGL((helm27_kernel<R, A, D, W, 4, 2, EPI>), grid, dim3(TX * 4), 0, s, L);
        else hipLaunchKernelGGL((helm27_kernel<R, A, D, W, 4, 1, EPI>), grid, dim3(TX * 4), 0, s, L);
    }
    return hipGetLastError();
}

// Epilogue by operand presence: the addend form exists for the wavenumber
// operators (residuals), the gradient form for the FP64 adjoint derivative.
template <typename R, bool A, bool D, bool W>
hipError_t launch_epi(const StencilLaunch<R>& L, const StencilGrid& g, hipStream_t s)
{
    if constexpr (std::is_same<R, double>::value && A && D && !W) {
        if (L.gg) return launch_shape<R, A, D, W, 2>(L, g, s);
    }
    if constexpr (!D) {
        if (L.b) return launch_shape<R, A, D, W, 1>(L, g, s);
    }
    if (L.gg || L.b) return hipErrorInvalidValue;
    return launch_shape<R, A, D, W, 0>(L, g, s);
}
}  // namespace

template <typename R>
hipError_t launch_stencil(int variant, StencilLaunch<R> L, const StencilGrid& g, hipStream_t s)
{
    L.tiles_x = g.tiles_x;
    L.tiles_y = g.tiles_y;
    L.kchunk = g.kchunk;
    const bool adj = variant & VAR_ADJ, der = variant & VAR_DERIV, wnc = variant & VAR_WNC;
    if (!adj && !der && !wnc) return launch_epi<R, false, false, false>(L, g, s);
    if (!adj && !der && wnc) return launch_epi<R, false, false, true>(L, g, s);
    if (!adj && der && !wnc) return launch_epi<R, false, true, false>(L, g, s);
    if (!adj && der && wnc) return launch_epi<R, false, true, true>(L, g, s);
    if (adj && !der && !wnc) return launch_epi<R, true, false, false>(L, g, s);
    if (adj && !der && wnc) return launch_epi<R, true, false, true>(L, g, s);
    if (adj && der && !wnc) return launch_epi<R, true, true, false>(L, g, s);
    return launch_epi<R, true, true, true>(L, g, s);
}

template StencilGrid stencil_grid<double>(int, int, int, int, int, int);
template StencilGrid stencil_grid<float>(int, int, int, int, int, int);
template hipError_t launch_stencil<double>(int, StencilLaunch<double>, const StencilGrid&, hipStream_t);
template hipError_t launch_stencil<float>(int, StencilLaunch<float>, const StencilGrid&, hipStream_t);

}  // namespace helm3d
