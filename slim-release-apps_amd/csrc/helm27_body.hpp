// Device body of the 27-point stencil (included by the kernel files).
//
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
// Tiling.  A block of 64 x TY / RY threads owns a 64 x TY column tile and
// marches over a z chunk.  Each plane (with a one-point x/y halo) is staged
// once in LDS (double-buffered, one barrier per plane); the loads of the
// planes ahead are issued before the barrier so HBM latency overlaps the
// current plane.  Block
// ids are remapped so consecutive tiles of the same chunk land on the same XCD
// and share halo lines in its L2.
#pragma once

#include <type_traits>

#include "bicgstab.hpp"
#include "helm27.hpp"

namespace helm3d {

namespace {

constexpr int TX = 64;  // one wavefront along x
constexpr int HX = TX + 2;
constexpr int kMaxChunk = 254;  // z planes per block (LDS holds their PML weights)

template <typename R>
__device__ inline cpx<R> zero() { return cpx<R>{R(0), R(0)}; }

// Output stores carry the non-temporal hint: no output is read again by the
// launch that writes it, and streaming the outputs past the caches leaves
// them to the halo planes of the inputs (the 240x240x140 matvec: 4-7 %
// faster in FP64 and complex64, forward and adjoint).
template <typename R>
__device__ __forceinline__ void st_out(cpx<R>* p, cpx<R> v)
{
    typedef R r2 __attribute__((ext_vector_type(2)));
    r2 w;
    w.x = v.re;
    w.y = v.im;
    __builtin_nontemporal_store(w, reinterpret_cast<r2*>(p));
}

// FP64 complex scalars of the fused BiCGStab forms (device slots, see
// bicgstab.hpp); a zero denominator gives zero (the breakdown guard of the
// BiCGStab kernels).
struct Cd {
    double re, im;
};
__device__ inline Cd cd_ld(const double* s, int off) { return Cd{s[off], s[off + 1]}; }
__device__ inline Cd cd_mul(Cd a, Cd b) { return Cd{fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re)}; }
__device__ inline Cd cd_div(Cd a, Cd b)
{
    const double d = fma(b.re, b.re, b.im * b.im);
    if (d == 0.0) return Cd{0.0, 0.0};
    return Cd{fma(a.re, b.re, a.im * b.im) / d, fma(a.im, b.re, -(a.re * b.im)) / d};
}

// TY: rows of the column tile; RY: rows per thread (1 or 2: a block is
// 64 x TY / RY threads, one wave per RY consecutive rows, and the 3 x (RY+2)
// window of a plane is read once for the RY outputs); PF: planes of global
// loads kept in flight ahead of the plane being computed (1 to 4).
//
// Memory pipeline: every global load of the loop goes into one of PF register
// sets that are refilled right after being staged into LDS, so no register is
// ever copied while its load is in flight and the only vmcnt waits are those
// of the set being staged (PF planes after it was requested).  The per-plane z
// PML weights are staged in LDS once per block, and the own-point operands of
// the epilogue (adjoint wn, addend b, gradient U and g) travel in the same set
// as the halo slots of the next plane.
// EPI: 0 stores y = alpha*H x, 1 stores y = alpha*H x + beta*b, 2 runs the
// gradient epilogue (accumulates Re(conj(U) . out) into g).  3 and 4 are the
// two operator applications of a BiCGStab iteration with every vector update
// fused in (scalars read from the device slots L.sc), so an iteration is two
// launches and s is never stored:
//   3: from r, p, v, t of the previous iteration and its alpha, omega:
//      s = r - alpha v, r' = s - omega t, x = p' = r' + beta (p - omega v)
//      formed while staging, z += alpha p + omega s, p' and r' stored at the
//      own point, y = v' = H x, partial (rhat, v'); rho' = (rhat, s) -
//      omega (rhat, t) and beta come from the partials of the previous
//      EPI 4 launch
//   4: x = s = r - alpha v formed while staging, y = t = H s, partials
//      (t, s), ||t||^2, (rhat, s) and (rhat, t)
// The body runs one block `lin` of a launch of `G` blocks described by L.
template <typename R, bool ADJ, bool DERIV, bool WNC, int TY, int PF, int EPI, int RY = 1>
__device__ __forceinline__ void helm27_body(const StencilLaunch<R>& L, const int lin, const int G)
{
    static_assert(RY == 1 || RY == 2, "one or two rows per thread");
    static_assert(TY % RY == 0, "rows per thread divide the tile rows");
    constexpr int NT = TX * TY / RY;     // threads
    constexpr int HY = TY + 2;
    constexpr int NS = HX * HY;          // halo tile slots
    constexpr int NSL = (NS + NT - 1) / NT;  // load slots per thread
    static_assert(PF >= 1 && PF <= 4, "prefetch depth 1 to 4");
    if (L.skip && *L.skip) return;  // uniform over the launch
    const OpDesc<R>& op = L.op;
    __shared__ cpx<R> su[2][HY][HX];            // input vector plane (+halo)
    // forward mass operand: wn .* x (WNC, DERIV) or the real wn itself, the
    // product being formed from the x plane already read for the stiffness
    constexpr bool kWnSep = !ADJ && !WNC && !DERIV;
    __shared__ cpx<R> sm[(ADJ || kWnSep) ? 1 : 2][HY][HX];
    __shared__ R sw[kWnSep ? 2 : 1][HY][HX];
    __shared__ cpx<R> szl[kMaxChunk + 2 + PF], szh[kMaxChunk + 2 + PF];  // z PML slots k0 .. k1+PF+1
    // EPI 3: r' and the z increment of the staged points, read back at the
    // own point
    constexpr int kE3 = EPI == 3 ? 1 : 0;
    __shared__ cpx<R> sr[1 + kE3][1 + kE3 * (HY - 1)][1 + kE3 * (HX - 1)];
    __shared__ cpx<R> sd[1 + kE3][1 + kE3 * (HY - 1)][1 + kE3 * (HX - 1)];

    const int nx = op.nx, ny = op.ny, nz = op.nz;

    // XCD-aware block remap: the first 8*floor(G/8) blocks are dealt so that
    // each XCD (round-robin by block id) receives a contiguous logical range.
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
    const int tx = tid & (TX - 1), tw = tid / TX;  // lane, wave
    const int tr0 = tw * RY;                       // first tile row of this thread
    const int i = i0 + tx;
    const int ic = min(i, nx - 1);  // clamped for table reads
    // A wave is RY rows of 64 columns.  Per row it takes one of two paths
    // (uniform over the wave, and the same for an output point whatever the
    // launch shape): interior when every column and the row have
    // xi_lo = xi_hi = 1 at index-1, index, index+1, the general path
    // otherwise.
    const int jw = __builtin_amdgcn_readfirstlane(j0 + tr0);
    const bool xint = i0 >= op.int_lo[0] && i0 + TX - 1 <= op.int_hi[0];
    bool yint[RY], active[RY];
    long offc[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        yint[r] = jw + r >= op.int_lo[1] && jw + r <= op.int_hi[1];
        active[r] = (i < nx) && (jw + r < ny);
        offc[r] = long(ic) + long(min(jw + r, ny - 1)) * nx;
    }

    // ---- per-column coefficients (fixed for the whole march)
    // Every F and G coefficient is a real combination of the six x/y PML
    // values of the column, xs[0], xs[2], ys[0], ys[2] (neighbour sides) and
    // xs[1], ys[1] (the column's own xi), so F.u and G.u are applied in
    // factored form: six complex multipliers times real-weighted sums of the
    // plane's 3x3 neighbourhood (24 registers instead of 18 complex
    // coefficients).  Adjoint: side values of the neighbours, conjugated.
    // Forward, the face multipliers are the sums of the side ones and are not
    // formed.  The y multipliers are uniform over the wave (rows jw ..
    // jw+RY-1): read at wave-uniform indices so they live in scalar registers.
    cpx<R> px0{}, px2{}, pxf{};
    cpx<R> py0[RY], py2[RY], pyf[RY];
    if constexpr (!DERIV) {
        if constexpr (!ADJ) {
            px0 = op.xlo[ic + 1]; px2 = op.xhi[ic + 1];  // face multiplier: px0 + px2
        } else {
            // neighbour i-1 couples through its hi value, i+1 through its lo value
            px0 = conjc(op.xhi[ic]); px2 = conjc(op.xlo[ic + 2]); pxf = conjc(op.xlo[ic + 1] + op.xhi[ic + 1]);
        }
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const int jcw = min(jw + r, ny - 1);
            if constexpr (!ADJ) {
                py0[r] = op.ylo[jcw + 1]; py2[r] = op.yhi[jcw + 1];
            } else {
                py0[r] = conjc(op.yhi[jcw]); py2[r] = conjc(op.ylo[jcw + 2]);
                pyf[r] = conjc(op.ylo[jcw + 1] + op.yhi[jcw + 1]);
            }
        }
        // z weights of this chunk: slots k0 .. k1+1 (planes k0-1 .. k1) plus
        // the padding entries read by the trailing pipeline steps
        for (int q = tid; q < k1 - k0 + 2 + PF; q += NT) {
            const int sl = min(k0 + q, nz + 1);
            szl[q] = op.zlo[sl];
            szh[q] = op.zhi[sl];
        }
    }

    // ---- load slots: the (TX+2) x (TY+2) halo tile, NSL slots per thread
    const long plane = long(nx) * ny;
    int sxs[NSL], sys[NSL];
    bool sin_[NSL], shas[NSL];
    long soff[NSL];
#pragma unroll
    for (int t = 0; t < NSL; ++t) {
        const int sl = tid + t * NT;
        sxs[t] = sl % HX;
        sys[t] = sl / HX;
        shas[t] = (t + 1) * NT <= NS || sl < NS;
        const int gi = i0 - 1 + sxs[t], gj = j0 - 1 + sys[t];
        sin_[t] = shas[t] && gi >= 0 && gi < nx && gj >= 0 && gj < ny;
        soff[t] = sin_[t] ? gi + long(gj) * nx : 0;
    }
    // gradient epilogue: position of the own column in the physical box
    const int qi = i - L.p0[0];
    bool gcol[RY];
    long gcol_off[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        const int qj = jw + r - L.p0[1];
        gcol[r] = EPI == 2 && active[r] && qi >= 0 && qi < L.pn[0] && qj >= 0 && qj < L.pn[1];
        gcol_off[r] = gcol[r] ? qi + long(qj) * L.pn[0] : 0;
    }
    const long gplane = long(L.pn[0]) * L.pn[1];
    // real-wn loads (forward); with a constant wn they read a valid dummy address
    const R* wnp = op.wn ? op.wn : reinterpret_cast<const R*>((EPI == 3 || EPI == 4) ? L.f0 : L.x);

    // block partial sums: ||y||^2 (EPI 0/1), (rhat, v) (EPI 3), (t, s),
    // ||t||^2, (rhat, s), (rhat, t) (EPI 4)
    constexpr int NV = EPI == 3 ? 2 : (EPI == 4 ? 7 : 1);
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] = 0.0;
    const int sink_slot = int((long(lin) * NT + tid) % kSinkSlots);
    const int sink_slot2 = (sink_slot + kSinkSlots / 2) % kSinkSlots;
    const int sink_slot3 = (sink_slot + kSinkSlots / 4) % kSinkSlots;
    const int sink_slot4 = (sink_slot + 3 * (kSinkSlots / 4)) % kSinkSlots;

    // BiCGStab coefficients.  EPI 3: s = r + na v, r' = s + nw t,
    // p' = r' + bt p + nbw v, dz = za p + zw s (na = -alpha, nw = -omega,
    // bt = beta, nbw = -beta omega, za = alpha, zw = omega; all zero on the
    // first iteration, whose p, v, t are zero and r = b).  EPI 4: s = r + ca v
    // (ca = -alpha).
    cpx<R> ca = zero<R>();
    cpx<R> na = zero<R>(), nw = zero<R>(), bt = zero<R>(), nbw = zero<R>(), za = zero<R>(), zw = zero<R>();
    if constexpr (EPI == 3 || EPI == 4) {
        double* sc = L.sc;
        if constexpr (EPI == 3) {
            if (!L.first) {
                const Cd a = cd_ld(sc, BS_ALPHA);
                const double tn = sc[BS_TT];
                const Cd w = tn > 0.0 ? Cd{sc[BS_TS] / tn, sc[BS_TS + 1] / tn} : Cd{0.0, 0.0};
                const Cd wrt = cd_mul(w, cd_ld(sc, BS_RT));
                const Cd rho = Cd{sc[BS_RS] - wrt.re, sc[BS_RS + 1] - wrt.im};
                const Cd beta = cd_mul(cd_div(rho, cd_ld(sc, L.rho_old)), cd_div(a, w));
                const Cd bw = cd_mul(beta, w);
                if (lin == 0 && tid == 0) {  // no block of this launch reads these slots
                    sc[L.rho_new] = rho.re;
                    sc[L.rho_new + 1] = rho.im;
                    sc[BS_OMEGA] = w.re;
                    sc[BS_OMEGA + 1] = w.im;
                }
                na = cpx<R>{R(-a.re), R(-a.im)};
                nw = cpx<R>{R(-w.re), R(-w.im)};
                bt = cpx<R>{R(beta.re), R(beta.im)};
                nbw = cpx<R>{R(-bw.re), R(-bw.im)};
                za = cpx<R>{R(a.re), R(a.im)};
                zw = cpx<R>{R(w.re), R(w.im)};
            }
        } else {
            const Cd a = cd_div(cd_ld(sc, L.rho_new), cd_ld(sc, BS_RV));
            if (lin == 0 && tid == 0) {
                L.sc[BS_ALPHA] = a.re;
                L.sc[BS_ALPHA + 1] = a.im;
            }
            ca = cpx<R>{R(-a.re), R(-a.im)};
        }
    }
    // EPI 3/4: staged x at the own points of the previous plane; EPI 3: r'
    // and dz there
    cpx<R> Xlast[RY], Rlast[RY], Dlast[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) Xlast[r] = Rlast[r] = Dlast[r] = zero<R>();
    // output scale: alpha, or alpha / sqrt(*anrm) with the norm held on the device
    const R alpha = (EPI <= 1 && L.anrm) ? R(double(L.alpha) / sqrt(*L.anrm)) : L.alpha;

    // One register set: halo slots of plane q and the own-point operands of
    // plane q-1 (the plane output when q is staged).  Every load is issued
    // unconditionally from a clamped, in-bounds address and validity is
    // applied when the set is consumed, so the number of loads in flight is
    // fixed and the compiler waits only for the set it stages.
    struct Set {
        cpx<R> x[NSL];   // slots (EPI 3/4: r)
        cpx<R> e[NSL];   // EPI 3: p, EPI 4: v at the slots
        cpx<R> f[NSL];   // EPI 3: v at the slots
        cpx<R> g[NSL];   // EPI 3: t at the slots
        cpx<R> zo[RY];   // EPI 3: z at own points of plane q-1
        R w[NSL];        // forward, real wn at the slots
        cpx<R> c[NSL];   // forward, complex wn at the slots
        cpx<R> wc[RY];   // adjoint: wn at own points of plane q-1
        cpx<R> bb[RY];   // addend b (EPI 1) / rhat (EPI 3/4) at own points of plane q-1
        cpx<R> gu[RY];   // gradient: U at own points of plane q-1
        double gv[RY];   // gradient: g at own points of plane q-1
    };
    auto fetch = [&](int q, Set& S) {
        const long base = long(min(max(q, 0), nz - 1)) * plane;
#pragma unroll
        for (int t = 0; t < NSL; ++t) {
            if constexpr (EPI == 3 || EPI == 4) {
                S.x[t] = L.f0[base + soff[t]];
                S.e[t] = L.f1[base + soff[t]];
                if constexpr (EPI == 3) {
                    S.f[t] = L.f2[base + soff[t]];
                    S.g[t] = L.f3[base + soff[t]];
                }
            } else {
                S.x[t] = L.x[base + soff[t]];
            }
            if constexpr (!ADJ) {
                if constexpr (WNC) S.c[t] = op.wnc[base + soff[t]];
                else S.w[t] = wnp[base + soff[t]];
            }
        }
        const long ownk = long(min(max(q - 1, 0), nz - 1)) * plane;
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const long own = ownk + offc[r];
            if constexpr (ADJ) {
                if constexpr (WNC) S.wc[r] = op.wnc[own];
                else S.wc[r] = cpx<R>{wnp[own], R(0)};
            }
            if constexpr (EPI == 1) S.bb[r] = L.b[own];
            if constexpr (EPI == 3 || EPI == 4) S.bb[r] = L.rh[own];
            if constexpr (EPI == 3) S.zo[r] = L.zz[own];
            if constexpr (EPI == 2) {
                S.gu[r] = L.gu[own];
                const int qk = min(max(q - 1 - L.p0[2], 0), L.pn[2] - 1);
                S.gv[r] = L.gg[gcol_off[r] + long(qk) * gplane];
            }
        }
    };

    // pipeline state per output row (see the header comment of this file)
    cpx<R> Qprev[RY], Qcur[RY], Acc[RY], Pend[RY], MAcc[RY], MOcur[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        Qprev[r] = zero<R>();
        Qcur[r] = zero<R>();
        Acc[r] = zero<R>();
        Pend[r] = zero<R>();
        MAcc[r] = zero<R>();
        MOcur[r] = zero<R>();
    }

    // Stage set S (plane q) into LDS, refill S with plane q+PF, compute plane
    // q and emit plane q-1.
    auto step = [&](int q, int tstep, Set& S) {
        const int buf = tstep & 1;
        const bool inz = q >= 0 && q < nz;
        cpx<R> xs[NSL];
#pragma unroll
        for (int t = 0; t < NSL; ++t) {
            const bool v = inz && sin_[t];
            cpx<R> xv = S.x[t];
            if constexpr (EPI == 3) {
                const cpx<R> sv = cfma(S.x[t], na, S.f[t]);
                const cpx<R> rv = cfma(sv, nw, S.g[t]);
                xv = cfma(cfma(rv, bt, S.e[t]), nbw, S.f[t]);
                if (shas[t]) {
                    sr[buf][sys[t]][sxs[t]] = rv;
                    sd[buf][sys[t]][sxs[t]] = cfma(cmul(za, S.e[t]), zw, sv);
                }
            } else if constexpr (EPI == 4) {
                xv = cfma(xv, ca, S.e[t]);
            }
            xs[t] = v ? xv : zero<R>();
            if (shas[t]) su[buf][sys[t]][sxs[t]] = xs[t];
            if constexpr (!ADJ) {
                if constexpr (WNC) {
                    if (shas[t]) sm[buf][sys[t]][sxs[t]] = cmul(S.c[t], xs[t]);
                } else {
                    const R wv = op.wn ? S.w[t] : op.wn_const;
                    if constexpr (kWnSep) {
                        if (shas[t]) sw[buf][sys[t]][sxs[t]] = wv;
                    } else {
                        if (shas[t]) sm[buf][sys[t]][sxs[t]] = wv * xs[t];
                    }
                }
            }
        }
        // own-point operands of plane q-1 leave the set before the refill
        cpx<R> wc[RY], bb[RY], gu[RY];
        double gv[RY];
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            wc[r] = S.wc[r];
            if constexpr (ADJ && !WNC) wc[r].re = op.wn ? wc[r].re : op.wn_const;
            bb[r] = S.bb[r];
            gu[r] = S.gu[r];
            gv[r] = S.gv[r];
        }
        cpx<R> zo[RY];
#pragma unroll
        for (int r = 0; r < RY; ++r) zo[r] = S.zo[r];
        fetch(q + PF, S);  // unconditional: a fixed number of loads per step
        __syncthreads();

        // ---- the plane's 3 x (RY+2) window (stiffness operand) and its
        // mass operand; b indexes the rows tr0-1 .. tr0+RY of the tile
        constexpr int WB = RY + 2;
        // u: the window; m1: the mass operand of the middle column; xsu, xsm:
        // row sums along x of u and of the mass operand, shared by the RY
        // outputs (the side columns of the mass operand are needed only
        // through these sums)
        cpx<R> u[3][WB], m1[WB], xsu[WB], xsm[WB];
#pragma unroll
        for (int b = 0; b < WB; ++b) {
            cpx<R> uv[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                uv[a] = su[buf][tr0 + b][tx + a];
                if constexpr (!DERIV) u[a][b] = uv[a];
            }
            if constexpr (!DERIV) xsu[b] = uv[0] + uv[2];
            if constexpr (kWnSep) {
                // wn .* x formed here: w0 x0 + w2 x2 as one product and one fma
                const R w0 = sw[buf][tr0 + b][tx], w1 = sw[buf][tr0 + b][tx + 1], w2 = sw[buf][tr0 + b][tx + 2];
                xsm[b] = rfma(w2 * uv[2], w0, uv[0]);
                m1[b] = w1 * uv[1];
            } else if constexpr (ADJ) {
                xsm[b] = uv[0] + uv[2];
                m1[b] = uv[1];
            } else {
                xsm[b] = sm[buf][tr0 + b][tx] + sm[buf][tr0 + b][tx + 2];
                m1[b] = sm[buf][tr0 + b][tx + 1];
            }
        }

        const bool emit = q >= k0 + 1 && q <= k1;
        const int kc = q - 1;
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            // ---- mass-lumping sums (forward: on wn .* x, adjoint: on x)
            const cpx<R> mf = xsm[r + 1] + (m1[r] + m1[r + 2]);
            const cpx<R> md = xsm[r] + xsm[r + 2];
            cpx<R> Mc = ADJ ? rfma(MOcur[r], op.wn_coef, m1[r + 1]) : op.wn_coef * m1[r + 1];
            Mc = rfma(Mc, -op.wm2, mf);
            Mc = rfma(Mc, -op.wm3, md);
            cpx<R> Mo = (-op.wm2) * m1[r + 1];
            Mo = rfma(Mo, -op.wm3, mf);
            Mo = rfma(Mo, -op.wm4, md);

            // ---- stiffness: Qz (z PML, applied per plane) and F, G (x/y PML)
            // Forward: F and G accumulate onto the mass terms of the same
            // planes (FM = F + Mc, GM = G + Mo).  Adjoint: the mass terms are
            // weighted by conj(wn) at the output and stay apart; F starts
            // from the pending plane-below sum and Mc from the pending mass
            // sum, so both need no separate add.
            cpx<R> Qn = zero<R>(), Fn = zero<R>(), Gn = zero<R>();
            if constexpr (!ADJ) {
                Fn = Mc;
                Gn = Mo;
            } else {
                Fn = Pend[r];
            }
            if constexpr (!DERIV) {
                const cpx<R> u11 = u[1][r + 1];
                // pair sums along y (column a) and along x (row b)
                const cpx<R> sy1 = u[1][r] + u[1][r + 2];
                const cpx<R> sx0 = xsu[r], sx1 = xsu[r + 1], sx2 = xsu[r + 2];
                const cpx<R> dg = sx0 + sx2;
                Qn = op.c[2] * u11;
                Qn = rfma(Qn, -op.xz, sx1);
                Qn = rfma(Qn, -op.yz, sy1);
                Qn = rfma(Qn, -op.w3a2, dg);
                const R cx = op.c[0], cy = op.c[1], wxy = op.xy, wxz = op.xz, wyz = op.yz, w3 = op.w3a2;
                if (xint && yint[r]) {
                    // x/y PML absent around this wave: F, G are real constant
                    // stencils on the sums already formed for Qz
                    Fn = rfma(Fn, op.fi[0], u11);
                    Fn = rfma(Fn, op.fi[1], sx1);
                    Fn = rfma(Fn, op.fi[2], sy1);
                    Fn = rfma(Fn, op.fi[3], dg);
                    Gn = rfma(Gn, op.gi[0], u11);
                    Gn = rfma(Gn, op.gi[1], sx1);
                    Gn = rfma(Gn, op.gi[2], sy1);
                    Gn = rfma(Gn, op.gi[3], dg);
                } else if constexpr (!ADJ) {
                    // Forward: the face multipliers are the sums of the side
                    // ones (pxf = px0 + px2, pyf = py0 + py2), so each face
                    // term folds into the two side terms, which then act on
                    // differences against the centre row / column shared by
                    // F and G:
                    //   F = px0 (cx d0u + w_xy d0s) + px2 (cx d2u + w_xy d2s)
                    //     + py0 (cy e0u + w_xy e0s) + py2 (cy e2u + w_xy e2s)
                    //   G = px0 (-w_xz d0u + w3a2 d0s) + ... + py2 (-w_yz e2u + w3a2 e2s)
                    // (the terms listed in the adjoint branch below with
                    // pxf = px0 + px2 and pyf = py0 + py2, regrouped).
                    const cpx<R> sy0 = u[0][r] + u[0][r + 2], sy2 = u[2][r] + u[2][r + 2];
                    const cpx<R> d0u = u[0][r + 1] - u11, d2u = u[2][r + 1] - u11;
                    const cpx<R> d0s = sy1 - sy0, d2s = sy1 - sy2;
                    const cpx<R> e0u = u[1][r] - u11, e2u = u[1][r + 2] - u11;
                    const cpx<R> e0s = sx1 - sx0, e2s = sx1 - sx2;
                    Fn = cfma(Fn, px0, rfma(cx * d0u, wxy, d0s));
                    Gn = cfma(Gn, px0, rfma((-wxz) * d0u, w3, d0s));
                    Fn = cfma(Fn, px2, rfma(cx * d2u, wxy, d2s));
                    Gn = cfma(Gn, px2, rfma((-wxz) * d2u, w3, d2s));
                    Fn = cfma(Fn, py0[r], rfma(cy * e0u, wxy, e0s));
                    Gn = cfma(Gn, py0[r], rfma((-wyz) * e0u, w3, e0s));
                    Fn = cfma(Fn, py2[r], rfma(cy * e2u, wxy, e2s));
                    Gn = cfma(Gn, py2[r], rfma((-wyz) * e2u, w3, e2s));
                } else {
                    // F: centre -(cx xf + cy yf); x faces cx xs + w_xy yf; y faces
                    //    cy ys + w_xy xf; diagonals -w_xy (xs + ys)
                    // G: centre w_xz xf + w_yz yf; x faces -w_xz xs + w3a2 yf;
                    //    y faces -w_yz ys + w3a2 xf; diagonals -w3a2 (xs + ys)
                    // (the non-xi_z part of Helm3d_27pt.h:680-714), grouped by
                    // multiplier; sy0/sy2: column sums, sx0/sx2: row sums of the
                    // diagonal pairs.  Adjoint: the face multipliers are the
                    // column's own values, not sums of the side ones.
                    const cpx<R> sy0 = u[0][r] + u[0][r + 2], sy2 = u[2][r] + u[2][r + 2];
                    // accumulated multiplier by multiplier (few live temporaries)
                    Fn = cfma(Fn, px0, rfma(cx * u[0][r + 1], -wxy, sy0));
                    Gn = cmul(px0, rfma((-wxz) * u[0][r + 1], -w3, sy0));
                    Fn = cfma(Fn, px2, rfma(cx * u[2][r + 1], -wxy, sy2));
                    Gn = cfma(Gn, px2, rfma((-wxz) * u[2][r + 1], -w3, sy2));
                    Fn = cfma(Fn, py0[r], rfma(cy * u[1][r], -wxy, sx0));
                    Gn = cfma(Gn, py0[r], rfma((-wyz) * u[1][r], -w3, sx0));
                    Fn = cfma(Fn, py2[r], rfma(cy * u[1][r + 2], -wxy, sx2));
                    Gn = cfma(Gn, py2[r], rfma((-wyz) * u[1][r + 2], -w3, sx2));
                    Fn = cfma(Fn, pyf[r], rfma(wxy * sx1, -cy, u11));
                    Gn = cfma(Gn, pyf[r], rfma(w3 * sx1, wyz, u11));
                    Fn = cfma(Fn, pxf, rfma(wxy * sy1, -cx, u11));
                    Gn = cfma(Gn, pxf, rfma(w3 * sy1, wxz, u11));
                }
            }

            cpx<R> out = zero<R>();
            if constexpr (!ADJ) {
                // y(kc) = wlo Q(kc-1) + whi Q(kc+1) - (wlo+whi) Q(kc) + GM(kc-1) + GM(kc+1) + FM(kc)
                const cpx<R> GMn = Gn;  // G + Mo
                if (emit) {
                    out = Acc[r] + GMn;
                    if constexpr (!DERIV) {
                        const cpx<R> wl = szl[kc + 1 - k0], wh = szh[kc + 1 - k0];
                        out = cfma(out, wl, Qprev[r] - Qcur[r]);
                        out = cfma(out, wh, Qn - Qcur[r]);
                    }
                }
                Qprev[r] = Qcur[r];
                Qcur[r] = Qn;
                Acc[r] = Pend[r] + Fn;  // GM(kc-1) + F + Mc
                Pend[r] = GMn;
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
                    out = Acc[r] + Sup;
                    out = cfma(out, conjc(wc[r]), MAcc[r] + Mo);
                }
                Acc[r] = S0;
                MAcc[r] = Mc;
                Pend[r] = Sdn;
                MOcur[r] = Mo;
            }
            // Exactly one store per step, lane and row (four for EPI 3):
            // steps or lanes without an output write to the sink instead
            // (keeps the vmcnt bookkeeping exact).
            const bool live = emit && active[r];
            const int sk = r == 0 ? sink_slot : sink_slot2;
            if constexpr (EPI == 3) {
                const long od = long(kc) * plane + offc[r];
                cpx<R>* dy = live ? &L.y[od] : &L.sink[sink_slot];
                cpx<R>* d2 = live ? &L.o2[od] : &L.sink[sink_slot2];
                cpx<R>* d3 = live ? &L.o3[od] : &L.sink[sink_slot3];
                cpx<R>* d4 = live ? &L.zz[od] : &L.sink[sink_slot4];
                st_out(dy, out);
                st_out(d2, Xlast[r]);
                st_out(d3, Rlast[r]);
                st_out(d4, zo[r] + Dlast[r]);
                if (live) {
                    // (rhat, v) += conj(rhat) v
                    acc[0] += double(bb[r].re) * double(out.re) + double(bb[r].im) * double(out.im);
                    acc[1] += double(bb[r].re) * double(out.im) - double(bb[r].im) * double(out.re);
                }
                Xlast[r] = su[buf][tr0 + r + 1][tx + 1];  // own point of plane q
                Rlast[r] = sr[buf][tr0 + r + 1][tx + 1];
                Dlast[r] = sd[buf][tr0 + r + 1][tx + 1];
            } else if constexpr (EPI == 4) {
                cpx<R>* dy = live ? &L.y[long(kc) * plane + offc[r]] : &L.sink[sk];
                st_out(dy, out);
                if (live) {
                    // (t, s) += conj(t) s ; ||t||^2 ; (rhat, s) ; (rhat, t)
                    const double tr_ = out.re, ti = out.im, sr_ = Xlast[r].re, si = Xlast[r].im, hr = bb[r].re,
                                 hi = bb[r].im;
                    acc[0] += tr_ * sr_ + ti * si;
                    acc[1] += tr_ * si - ti * sr_;
                    acc[2] += tr_ * tr_ + ti * ti;
                    acc[3] += hr * sr_ + hi * si;
                    acc[4] += hr * si - hi * sr_;
                    acc[5] += hr * tr_ + hi * ti;
                    acc[6] += hr * ti - hi * tr_;
                }
                Xlast[r] = su[buf][tr0 + r + 1][tx + 1];  // own point of plane q
            } else if constexpr (EPI == 2) {
                const int qk = kc - L.p0[2];
                const bool g = live && gcol[r] && qk >= 0 && qk < L.pn[2];
                const double v = fma(double(gu[r].re), double(out.re), double(gu[r].im) * double(out.im));
                double* dst = g ? &L.gg[gcol_off[r] + long(qk) * gplane] : reinterpret_cast<double*>(&L.sink[sk]);
                *dst = fma(L.gscale, v, gv[r]);
            } else {
                cpx<R> res = alpha * out;
                if constexpr (EPI == 1) res = rfma(res, L.beta, bb[r]);
                cpx<R>* dst = live ? &L.y[long(kc) * plane + offc[r]] : &L.sink[sk];
                st_out(dst, res);
                if (live) acc[0] += double(res.re) * double(res.re) + double(res.im) * double(res.im);
            }
        }
    };

    // Steps for planes k0-1 .. k1 (rounded up to a multiple of PF; a trailing
    // padding step computes but emits nothing).  No memory operation in the
    // loop body is conditional on the step, so the compiler can keep PF sets
    // of loads in flight across the back-edge.
    Set S[PF];
#pragma unroll
    for (int p = 0; p < PF; ++p) {
        // keep each set's loads ahead of the next set's: the loop then waits
        // only for the set it stages (the wait count at the loop header is
        // merged from this prologue and the back edge)
        if (p > 0) __builtin_amdgcn_sched_barrier(0);
        fetch(k0 - 1 + p, S[p]);
    }
    const int ntrips = (k1 - k0 + 2 + PF - 1) / PF;
    for (int trp = 0; trp < ntrips; ++trp) {
        const int q = k0 - 1 + trp * PF;
#pragma unroll
        for (int p = 0; p < PF; ++p) step(q + p, trp * PF + p, S[p]);
    }

    // optional per-block partials (deterministic two-stage reduction):
    // part[q*G + lin] for q < NV
    if (L.norm_part) {
        __shared__ double red[NV][NT / 64];
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double v = acc[q];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            if ((tid & 63) == 0) red[q][tid >> 6] = v;
        }
        __syncthreads();
        if (tid < NV) {
            double s = 0;
            for (int w = 0; w < NT / 64; ++w) s += red[tid][w];
            L.norm_part[long(tid) * G + lin] = s;
        }
    }
}


}  // namespace
}  // namespace helm3d
