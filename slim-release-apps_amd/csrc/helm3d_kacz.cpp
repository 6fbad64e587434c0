// Kaczmarz sweeps and the CGMN solver on them (C ABI).
//
//   helm3d_kacz_sweep   device form of Helm3d_kacz_sweep{,_wnc,_adj}_mex
//                       (Helm3dkaczswp_mex.c:60-117, 27-colour order)
//   helm3d_kaczswp      the same on host split-complex arrays, MEX argument
//                       checks of Helm3dkaczswp_mex.c:138-176
//   helm3d_solve_cgmn   CGMN.m on the Kaczmarz-preconditioned system of
//                       linearsolve.m:52-58 (omega = 1.5, one double sweep):
//                       b' = SWP(0, b), A' y = y - SWP(y, 0), CG from x = 0
#include "capi_internal.hpp"

namespace helm3d {
hipError_t launch_kacz_sweep(bool adj, const OpDesc<double>& op, double omega, cpx<double>* x,
                             const cpx<double>* b, hipStream_t s);
}

namespace {

int sweep(helm3d_ctx* c, int mode, double omega, cpx<double>* x, const cpx<double>* b, int n)
{
    for (int s = 0; s < n; ++s) HIPCHK(launch_kacz_sweep(mode != 1, c->op64, omega, x, b, c->stream));
    return HELM3D_OK;
}

int check_ctx(helm3d_ctx* c)
{
    CHK(bind(c));
    if (!c->has_model) return fail(HELM3D_ERR_NOMODEL, "no model set");
    return HELM3D_OK;
}

}  // namespace

namespace {

// ||b - A x|| / ||b|| with the operator, and the solve's info (both
// Kaczmarz-preconditioned solvers); returns NOTCONV above tol.
int finish_kacz_solve(helm3d_ctx* c, int mode, const cpx<double>* b, const cpx<double>* x, cpx<double>* t, long k,
                      double rel, double tol, helm3d_solve_info* info)
{
    const long N = c->N;
    const hipStream_t st = c->stream;
    const int nblk = krylov_blocks(N);
    StencilLaunch<double> L{};
    L.op = c->op64;
    L.x = x;
    L.y = t;
    L.b = b;
    L.alpha = -1.0;
    L.beta = 1.0;
    L.norm_part = c->part;
    L.sink = static_cast<cpx<double>*>(c->sink);
    HIPCHK(launch_stencil<double>((mode != 1 ? VAR_ADJ : VAR_FORW) | (c->wnc ? VAR_WNC : 0), L, c->grid, st));
    HIPCHK(launch_reduce(c->part, c->grid.blocks, 1, c->small, st));
    HIPCHK(launch_norm2<double>(b, N, c->part, nblk, st));
    HIPCHK(launch_reduce(c->part, nblk, 1, c->small + 1, st));
    HIPCHK(hipMemcpyAsync(c->small_host, c->small, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const double rr = c->small_host[0], b2 = c->small_host[1];
    if (info) {
        info->iters = k;
        info->inner_iters = 2 * k + 1;  // double sweeps
        info->cycles = k;
        info->relres = b2 > 0 ? std::sqrt(rr / b2) : 0.0;
        info->quasi_relres = rel;
        info->converged = !(rel > tol);
    }
    return !(rel > tol) ? HELM3D_OK : fail(HELM3D_ERR_NOTCONV, "maxit reached above tolerance");
}

}  // namespace

extern "C" {

int helm3d_kacz_sweep(helm3d_ctx* c, int mode, double omega, int nsweeps, void* x, const void* b)
{
    CHK(check_ctx(c));
    if (!x) return fail(HELM3D_ERR_ARG, "null vector");
    if (nsweeps < 0) return fail(HELM3D_ERR_ARG, "nsweeps must be >= 0");
    return sweep(c, mode, omega, static_cast<cpx<double>*>(x), static_cast<const cpx<double>*>(b), nsweeps);
}

int helm3d_kaczswp(int variant, const double* wnr, const double* wni, const double* h, const double* n,
                   long long n_len, const double* npml, double omega, const double* xr, const double* xi,
                   long long x_rows, long long x_cols, const double* br, const double* bi, long long b_rows,
                   long long b_cols, double* yr, double* yi)
{
    if (n_len != 3 || !n) return fail(HELM3D_ERR_NSIZE, "n must be a 3-length vector");
    int ni[3];
    for (int a = 0; a < 3; ++a) {
        if (!(n[a] >= 1) || n[a] != std::floor(n[a]) || n[a] > 2147483647.0)
            return fail(HELM3D_ERR_NSIZE, "n must be a 3-length vector");
        ni[a] = int(n[a]);
    }
    const long long N = (long long)ni[0] * ni[1] * ni[2];
    if (x_rows != N) return fail(HELM3D_ERR_XSIZE, "x must have nx*ny*nz elements");
    if (b_rows != N) return fail(HELM3D_ERR_BSIZE, "b must have nx*ny*nz elements");
    if (x_cols < 0) return fail(HELM3D_ERR_XSIZE, "x must have a non-negative number of columns");
    // column s of y sweeps column s of x against column s of b
    // (Helm3dkaczswp_mex.c:159,182-186)
    if (b_cols != x_cols) return fail(HELM3D_ERR_BSIZE, "b must have as many columns as x");
    if (variant != HELM3D_FORW && variant != HELM3D_ADJ && variant != (HELM3D_FORW | HELM3D_WNC) &&
        variant != (HELM3D_ADJ | HELM3D_WNC))
        return fail(HELM3D_ERR_ARG, "bad variant");
    const bool wnc = variant & HELM3D_WNC;
    if (!wnr || !h || !npml || (wnc && !wni)) return fail(HELM3D_ERR_ARG, "null argument");
    if (x_cols > 0 && (!xr || !br || !yr || !yi)) return fail(HELM3D_ERR_ARG, "null argument");
    int np[6];
    for (int a = 0; a < 6; ++a) {
        if (!(npml[a] >= 0)) return fail(HELM3D_ERR_ARG, "npml must be non-negative");
        np[a] = int(npml[a]);
    }
    if (x_cols == 0) return HELM3D_OK;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    helm3d_ctx* c = nullptr;
    CHK(helm3d_create(dev, nullptr, &c));
    struct Guard {
        helm3d_ctx* c;
        void* p = nullptr;
        ~Guard()
        {
            if (p) (void)hipFree(p);
            helm3d_destroy(c);
        }
    } g{c};
    // device staging: wn (complex when wnc), x, b (interleaved), planar scratch
    const size_t cb = size_t(N) * sizeof(cpx<double>);
    HIPCHK(hipMalloc(&g.p, 4 * cb));
    char* base = static_cast<char*>(g.p);
    cpx<double>* d_wn = reinterpret_cast<cpx<double>*>(base);
    cpx<double>* d_x = reinterpret_cast<cpx<double>*>(base + cb);
    cpx<double>* d_b = reinterpret_cast<cpx<double>*>(base + 2 * cb);
    double* d_pl = reinterpret_cast<double*>(base + 3 * cb);
    const hipStream_t st = c->stream;
    auto up = [&](const double* re, const double* im, cpx<double>* dst) -> int {
        HIPCHK(hipMemcpyAsync(d_pl, re, size_t(N) * sizeof(double), hipMemcpyHostToDevice, st));
        if (im) HIPCHK(hipMemcpyAsync(d_pl + N, im, size_t(N) * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(launch_interleave(d_pl, im ? d_pl + N : nullptr, dst, N, st));
        return HELM3D_OK;
    };
    if (wnc) {
        CHK(up(wnr, wni, d_wn));
        CHK(helm3d_set_model(c, ni, h, np, d_wn, 1));
    } else {
        HIPCHK(hipMemcpyAsync(d_wn, wnr, size_t(N) * sizeof(double), hipMemcpyHostToDevice, st));
        CHK(helm3d_set_model(c, ni, h, np, d_wn, 0));
    }
    for (long long q = 0; q < x_cols; ++q) {
        const size_t o = size_t(q) * size_t(N);
        CHK(up(xr + o, xi ? xi + o : nullptr, d_x));
        CHK(up(br + o, bi ? bi + o : nullptr, d_b));
        CHK(sweep(c, (variant & HELM3D_ADJ) ? 2 : 1, omega, d_x, d_b, 1));
        HIPCHK(launch_deinterleave(d_x, d_pl, d_pl + N, N, st));
        HIPCHK(hipMemcpyAsync(yr + o, d_pl, size_t(N) * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(yi + o, d_pl + N, size_t(N) * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));  // the staging buffers serve the next column
    }
    return HELM3D_OK;
}

int helm3d_solve_cgmn(helm3d_ctx* c, int mode, const void* bv, void* xv, const helm3d_solve_opts* opts,
                      helm3d_solve_info* info)
{
    CHK(check_ctx(c));
    if (!bv || !xv) return fail(HELM3D_ERR_ARG, "null vector");
    helm3d_solve_opts o;
    helm3d_solve_opts_default(&o);
    if (opts) o = *opts;
    const double omega = 1.5;  // linearsolve.m:53
    const long N = c->N;
    const long ld = slot_ld(N);
    CHK(ensure_vecs(c->T64, N, 5));
    CHK(ensure_work(c));
    cpx<double>* r = c->T64.p;
    cpx<double>* p = r + ld;
    cpx<double>* Ap = r + 2 * ld;
    cpx<double>* t = r + 3 * ld;
    cpx<double>* bp = r + 4 * ld;
    const cpx<double>* b = static_cast<const cpx<double>*>(bv);
    cpx<double>* x = static_cast<cpx<double>*>(xv);
    const hipStream_t st = c->stream;
    const int nblk = krylov_blocks(N);
    const cpx<double>* none = nullptr;
    auto norm2 = [&](const cpx<double>* v, double* out) -> int {
        HIPCHK(launch_norm2<double>(v, N, c->part, nblk, st));
        HIPCHK(launch_reduce(c->part, nblk, 1, c->small, st));
        HIPCHK(hipMemcpyAsync(c->small_host, c->small, sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        *out = c->small_host[0];
        return HELM3D_OK;
    };
    // b' = SWP(0, b)
    HIPCHK((launch_axpby<double, double>(0.0, 0.0, none, 0.0, 0.0, bp, N, false, st)));
    CHK(sweep(c, mode, omega, bp, b, 1));
    // x = 0, r = p = b'
    HIPCHK((launch_axpby<double, double>(0.0, 0.0, none, 0.0, 0.0, x, N, false, st)));
    HIPCHK((launch_axpby<double, double>(1.0, 0.0, bp, 0.0, 0.0, r, N, false, st)));
    HIPCHK((launch_axpby<double, double>(1.0, 0.0, bp, 0.0, 0.0, p, N, false, st)));
    double nb2 = 0;
    CHK(norm2(bp, &nb2));
    const double normb = std::sqrt(nb2);
    double normr = normb;
    long k = 0;
    double rel = normb > 0 ? 1.0 : 0.0;
    if (normb > 0) {
        while (true) {
            // Ap = p - SWP(p, 0)
            HIPCHK((launch_axpby<double, double>(1.0, 0.0, p, 0.0, 0.0, t, N, false, st)));
            CHK(sweep(c, mode, omega, t, nullptr, 1));
            HIPCHK((launch_axpby<double, double>(1.0, 0.0, p, -1.0, 0.0, t, N, true, st)));
            std::swap(Ap, t);
            // pAp = p^H Ap
            HIPCHK(launch_dots<double>(p, ld, 1, Ap, N, c->part, nblk, st));
            HIPCHK(launch_reduce(c->part, nblk, 2, c->small, st));
            HIPCHK(hipMemcpyAsync(c->small_host, c->small, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            const std::complex<double> pAp(c->small_host[0], c->small_host[1]);
            const std::complex<double> alpha = (normr * normr) / pAp;
            HIPCHK((launch_axpby<double, double>(alpha.real(), alpha.imag(), p, 1.0, 0.0, x, N, true, st)));
            HIPCHK((launch_axpby<double, double>(-alpha.real(), -alpha.imag(), Ap, 1.0, 0.0, r, N, true, st)));
            const double normr_old = normr;
            double nr2 = 0;
            CHK(norm2(r, &nr2));
            normr = std::sqrt(nr2);
            rel = normr / normb;
            ++k;
            if (k >= o.maxit || !(rel > o.tol)) break;  // CGMN.m: stop at hst(end) <= tol
            const double beta = (normr / normr_old) * (normr / normr_old);
            HIPCHK((launch_axpby<double, double>(1.0, 0.0, r, beta, 0.0, p, N, true, st)));
        }
    } else {
        HIPCHK((launch_axpby<double, double>(0.0, 0.0, none, 0.0, 0.0, x, N, false, st)));
    }
    return finish_kacz_solve(c, mode, b, x, t, k, rel, o.tol, info);
}

int helm3d_solve_crmn(helm3d_ctx* c, int mode, const void* bv, void* xv, const helm3d_solve_opts* opts,
                      helm3d_solve_info* info)
{
    CHK(check_ctx(c));
    if (!bv || !xv) return fail(HELM3D_ERR_ARG, "null vector");
    helm3d_solve_opts o;
    helm3d_solve_opts_default(&o);
    if (opts) o = *opts;
    const double omega = 1.5;  // linearsolve.m:53
    const long N = c->N;
    const long ld = slot_ld(N);
    CHK(ensure_vecs(c->T64, N, 6));
    CHK(ensure_work(c));
    cpx<double>* r = c->T64.p;
    cpx<double>* p = r + ld;
    cpx<double>* Ap = r + 2 * ld;
    cpx<double>* Ar = r + 3 * ld;
    cpx<double>* t = r + 4 * ld;
    cpx<double>* bp = r + 5 * ld;
    const cpx<double>* b = static_cast<const cpx<double>*>(bv);
    cpx<double>* x = static_cast<cpx<double>*>(xv);
    const hipStream_t st = c->stream;
    const int nblk = krylov_blocks(N);
    const cpx<double>* none = nullptr;
    // host copies of a few reduced values
    auto fetch = [&](int count) -> int {
        HIPCHK(hipMemcpyAsync(c->small_host, c->small, count * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        return HELM3D_OK;
    };
    // out = in - SWP(in, 0): A' of linearsolve.m:55
    auto apply = [&](const cpx<double>* in, cpx<double>* out) -> int {
        HIPCHK((launch_axpby<double, double>(1.0, 0.0, in, 0.0, 0.0, out, N, false, st)));
        CHK(sweep(c, mode, omega, out, nullptr, 1));
        HIPCHK((launch_axpby<double, double>(1.0, 0.0, in, -1.0, 0.0, out, N, true, st)));
        return HELM3D_OK;
    };
    // (u, v) = u^H v and ||v||^2 into small[0..2], small[2]
    auto dot_norm = [&](const cpx<double>* u, const cpx<double>* v) -> int {
        HIPCHK(launch_dots<double>(u, ld, 1, v, N, c->part, nblk, st));
        HIPCHK(launch_reduce(c->part, nblk, 2, c->small, st));
        HIPCHK(launch_norm2<double>(v, N, c->part, nblk, st));
        HIPCHK(launch_reduce(c->part, nblk, 1, c->small + 2, st));
        return HELM3D_OK;
    };
    // b' = SWP(0, b); x = 0, r = p = b', Ar = Ap = A' r (CRMN.m:46-57)
    HIPCHK((launch_axpby<double, double>(0.0, 0.0, none, 0.0, 0.0, bp, N, false, st)));
    CHK(sweep(c, mode, omega, bp, b, 1));
    HIPCHK((launch_axpby<double, double>(0.0, 0.0, none, 0.0, 0.0, x, N, false, st)));
    HIPCHK((launch_axpby<double, double>(1.0, 0.0, bp, 0.0, 0.0, r, N, false, st)));
    HIPCHK(launch_norm2<double>(bp, N, c->part, nblk, st));
    HIPCHK(launch_reduce(c->part, nblk, 1, c->small + 3, st));
    CHK(apply(r, Ar));
    HIPCHK((launch_axpby<double, double>(1.0, 0.0, r, 0.0, 0.0, p, N, false, st)));
    HIPCHK((launch_axpby<double, double>(1.0, 0.0, Ar, 0.0, 0.0, Ap, N, false, st)));
    CHK(dot_norm(r, Ar));  // rAr; small[2] = ||Ar||^2 = ||Ap||^2
    CHK(fetch(4));
    const double normb = std::sqrt(c->small_host[3]);
    std::complex<double> rAr(c->small_host[0], c->small_host[1]);
    double nAp2 = c->small_host[2];
    long k = 0;
    double rel = normb > 0 ? 1.0 : 0.0;
    if (normb > 0) {
        while (true) {
            // CRMN.m:68-88
            const std::complex<double> gamma = rAr / nAp2;
            HIPCHK((launch_axpby<double, double>(gamma.real(), gamma.imag(), p, 1.0, 0.0, x, N, true, st)));
            HIPCHK((launch_axpby<double, double>(-gamma.real(), -gamma.imag(), Ap, 1.0, 0.0, r, N, true, st)));
            HIPCHK(launch_norm2<double>(r, N, c->part, nblk, st));
            HIPCHK(launch_reduce(c->part, nblk, 1, c->small, st));
            CHK(fetch(1));
            rel = std::sqrt(c->small_host[0]) / normb;
            ++k;
            if (k >= o.maxit || !(rel > o.tol)) break;  // stop at hst(end) <= tol
            // CRMN.m:94-106
            CHK(apply(r, Ar));
            HIPCHK(launch_dots<double>(r, ld, 1, Ar, N, c->part, nblk, st));
            HIPCHK(launch_reduce(c->part, nblk, 2, c->small, st));
            CHK(fetch(2));
            const std::complex<double> old = rAr;
            rAr = std::complex<double>(c->small_host[0], c->small_host[1]);
            const std::complex<double> beta = rAr / old;
            HIPCHK((launch_axpby<double, double>(1.0, 0.0, r, beta.real(), beta.imag(), p, N, true, st)));
            HIPCHK((launch_axpby<double, double>(1.0, 0.0, Ar, beta.real(), beta.imag(), Ap, N, true, st)));
            HIPCHK(launch_norm2<double>(Ap, N, c->part, nblk, st));
            HIPCHK(launch_reduce(c->part, nblk, 1, c->small, st));
            CHK(fetch(1));
            nAp2 = c->small_host[0];
        }
    }
    return finish_kacz_solve(c, mode, b, x, t, k, rel, o.tol, info);
}

}  // extern "C"
