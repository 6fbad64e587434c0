// Host-array entry points for gateways that hold MATLAB's split real /
// imaginary double arrays: load a model into a context and solve for one or
// many right-hand-side columns, with every device allocation, copy and
// launch checked here so a MEX gateway needs no HIP calls of its own.
//
// A solve-level gateway replaces linearsolve.m (:20-48, 94-97, 139-141) with
// the FGMRES solver: one column runs FGMRES.m (helm3d_solve), several columns
// FGMRES_MT.m (the multi-RHS branch of FGMRES.m:35-37, helm3d_batch_solve
// with HELM3D_BATCH_PLAIN).
#include "capi_internal.hpp"

namespace {

// Device scratch released on every return path.
struct DevBuf {
    void* p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

int parse_n(const double* n, long long n_len, int ni[3], long long* N)
{
    if (n_len != 3 || !n) return fail(HELM3D_ERR_NSIZE, "n must be a 3-length vector");
    for (int a = 0; a < 3; ++a) {
        if (!(n[a] >= 1) || n[a] != std::floor(n[a]) || n[a] > 2147483647.0)
            return fail(HELM3D_ERR_NSIZE, "n must be a 3-length vector");
        ni[a] = int(n[a]);
    }
    *N = (long long)ni[0] * ni[1] * ni[2];
    return HELM3D_OK;
}

}  // namespace

extern "C" {

int helm3d_set_model_host(helm3d_ctx* c, const double* n, long long n_len, const double* h, const double* npml,
                          const double* wnr, const double* wni)
{
    CHK(bind(c));
    int ni[3];
    long long N = 0;
    CHK(parse_n(n, n_len, ni, &N));
    if (!h || !npml || !wnr) return fail(HELM3D_ERR_ARG, "null argument");
    int np[6];
    for (int a = 0; a < 6; ++a) {
        if (!(npml[a] >= 0) || npml[a] > 2147483647.0) return fail(HELM3D_ERR_ARG, "npml must be non-negative");
        np[a] = int(npml[a]);
    }
    const size_t nb = size_t(N) * sizeof(double);
    DevBuf d;
    HIPCHK(hipMalloc(&d.p, (wni ? 4 : 1) * nb));
    double* pl = static_cast<double*>(d.p);
    HIPCHK(hipMemcpyAsync(pl, wnr, nb, hipMemcpyHostToDevice, c->stream));
    if (wni) {
        HIPCHK(hipMemcpyAsync(pl + N, wni, nb, hipMemcpyHostToDevice, c->stream));
        cpx<double>* z = reinterpret_cast<cpx<double>*>(pl + 2 * N);
        HIPCHK(launch_interleave(pl, pl + N, z, N, c->stream));
        CHK(helm3d_set_model(c, ni, h, np, z, 1));
    } else {
        CHK(helm3d_set_model(c, ni, h, np, pl, 0));
    }
    // the context holds its own copies once its stream has drained
    HIPCHK(hipStreamSynchronize(c->stream));
    return HELM3D_OK;
}

int helm3d_solve_host(helm3d_ctx* c, int mode, const double* br, const double* bi, long long b_rows,
                      long long b_cols, double* xr, double* xi, const helm3d_solve_opts* opts,
                      helm3d_solve_info* info)
{
    CHK(bind(c));
    if (!c->has_model) return fail(HELM3D_ERR_NOMODEL, "no model set");
    const long long N = c->N;
    if (b_rows != N) return fail(HELM3D_ERR_BSIZE, "b must have nx*ny*nz rows");
    if (b_cols < 0 || b_cols > HELM3D_BATCH_MAX_ITEMS)
        return fail(HELM3D_ERR_ARG, "b must have between 0 and HELM3D_BATCH_MAX_ITEMS columns");
    if (b_cols == 0) return HELM3D_OK;
    if (!br || !xr || !xi) return fail(HELM3D_ERR_ARG, "null argument");
    helm3d_solve_opts o;
    helm3d_solve_opts_default(&o);
    if (opts) o = *opts;
    if (o.use_x0) return fail(HELM3D_ERR_ARG, "host solves start from x = 0 (linearsolve.m:21-23)");
    const long ld = slot_ld(N);
    const size_t cb = size_t(ld) * sizeof(cpx<double>);
    // per column: b and x interleaved; one planar staging pair
    DevBuf d;
    HIPCHK(hipMalloc(&d.p, size_t(2 * b_cols + 1) * cb));
    cpx<double>* B = static_cast<cpx<double>*>(d.p);
    cpx<double>* X = B + size_t(b_cols) * ld;
    double* pl = reinterpret_cast<double*>(X + size_t(b_cols) * ld);
    for (long long q = 0; q < b_cols; ++q) {
        HIPCHK(hipMemcpyAsync(pl, br + q * N, size_t(N) * sizeof(double), hipMemcpyHostToDevice, c->stream));
        if (bi)
            HIPCHK(hipMemcpyAsync(pl + N, bi + q * N, size_t(N) * sizeof(double), hipMemcpyHostToDevice,
                                  c->stream));
        HIPCHK(launch_interleave(pl, bi ? pl + N : nullptr, B + q * ld, N, c->stream));
    }
    int status = HELM3D_OK;
    std::vector<helm3d_solve_info> inf(static_cast<size_t>(b_cols));
    if (b_cols == 1) {
        status = helm3d_solve(c, mode, B, X, &o, inf.data(), nullptr, 0, nullptr);
    } else {
        helm3d_batch* bt = nullptr;
        CHK(helm3d_batch_create(c->device, c->stream, &bt));
        std::vector<helm3d_ctx*> cs(static_cast<size_t>(b_cols), c);
        std::vector<const void*> bp(static_cast<size_t>(b_cols));
        std::vector<void*> xp(static_cast<size_t>(b_cols));
        for (long long q = 0; q < b_cols; ++q) {
            bp[size_t(q)] = B + q * ld;
            xp[size_t(q)] = X + q * ld;
        }
        status = helm3d_batch_solve(bt, int(b_cols), cs.data(), mode, bp.data(), xp.data(), HELM3D_BATCH_PLAIN, &o,
                                    inf.data());
        const std::string msg = g_err;
        helm3d_batch_destroy(bt);
        g_err = msg;
    }
    if (status != HELM3D_OK && status != HELM3D_ERR_NOTCONV) return status;
    for (long long q = 0; q < b_cols; ++q) {
        HIPCHK(launch_deinterleave(X + q * ld, pl, pl + N, N, c->stream));
        HIPCHK(hipMemcpyAsync(xr + q * N, pl, size_t(N) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(xi + q * N, pl + N, size_t(N) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        // the staging pair is reused by the next column
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    if (info)
        for (long long q = 0; q < b_cols; ++q) info[q] = inf[size_t(q)];
    return status;
}

}  // extern "C"
