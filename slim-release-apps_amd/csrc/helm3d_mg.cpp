// Multigrid-preconditioned FGMRES: the ML_GMRES preconditioner of the
// reference (ML_GMRES.m:34-72, opMultigrid.m:71-162, construct_multigrid.m,
// GMG_V.m) on device-resident level operators.
//
// Levels l = 0 (the operator being solved) .. L-1 (coarsest), each with two
// contexts holding the same operator: A_l (its workspace serves the outer
// solve at l = 0 and the coarse-grid solves at l >= 1) and S_l (the
// smoother's workspace).  With ks = (ks_outer, ks_inner), kc = (kc_outer,
// kc_inner) the recursive V-cycle of opMultigrid.construct (:73-110) is
//   V_l(b):  x = S_l(b, 0)                      FGMRES(ks_inner) x ks_outer, tol 1e-6
//            r = b - A_l x
//            c = R_l r                          R = 1/8 P^T
//            y = C_{l+1}(c)                     FGMRES(kc_inner) x kc_outer, tol 0.5,
//                                               preconditioned by V_{l+1} (l+1 < L-1)
//                                               or by FGMRES(kc) on A_{L-1} (l+1 = L-1)
//            x = x + P_l y ;  x = S_l(b, x)     (GMG_V.m:23-31, one cycle)
// and the outer solve is FGMRES on A_0 preconditioned by V_0.  The mixed form
// (helm3d_solve_mg_mixed) runs the whole V-cycle -- smoothers, residuals,
// transfers, coarse solves -- in complex64 with each context's FP32 operator,
// under the FP64 outer FGMRES.
#include "capi_internal.hpp"
#include "mg.hpp"

namespace {

struct MgLevel {
    helm3d_ctx* A = nullptr;
    helm3d_ctx* S = nullptr;
    long N = 0;
    cpx<double>* vb = nullptr;    // V-cycle right-hand side
    cpx<double>* vr = nullptr;    // residual
    cpx<double>* crhs = nullptr;  // coarse-solve right-hand side (levels >= 1)
    cpx<double>* cx = nullptr;    // coarse-solve solution (levels >= 1)
    MgAxes T{};                   // transfer to level + 1
};

// Transfer tables of one axis: LinInterp1D(linspace(0,1,nc), linspace(0,1,nf)).
struct AxisTab {
    std::vector<int> I, rp, fi, pfi;
    std::vector<double> wa, wb, fw, pfw;
    int pw = 0;
};

AxisTab axis_tab(int nf, int nc)
{
    AxisTab t;
    t.I.resize(nf);
    t.wa.resize(nf);
    t.wb.resize(nf);
    const double hc = nc > 1 ? 1.0 / (nc - 1) : 1.0;
    for (int i = 0; i < nf; ++i) {
        const double x = nf > 1 ? double(i) / (nf - 1) : 0.0;
        // nearest coarse point, stepped left when it lies right of x
        int I = nc > 1 ? int(std::lround(x / hc)) : 0;
        I = std::max(0, std::min(I, nc - 1));
        if (I > 0 && I * hc > x) --I;
        if (I >= nc - 1) {
            t.I[i] = nc - 1;
            t.wa[i] = 1.0;
            t.wb[i] = 0.0;
        } else {
            const double x0 = I * hc, x1 = (I + 1) * hc;
            t.I[i] = I;
            t.wa[i] = (x1 - x) / (x1 - x0);
            t.wb[i] = (x - x0) / (x1 - x0);
        }
    }
    // transposed lists
    std::vector<std::vector<std::pair<int, double>>> col(nc);
    for (int i = 0; i < nf; ++i) {
        col[t.I[i]].push_back({i, t.wa[i]});
        if (t.I[i] + 1 < nc) col[t.I[i] + 1].push_back({i, t.wb[i]});
    }
    t.rp.assign(1, 0);
    for (int c = 0; c < nc; ++c) {
        for (auto& e : col[c]) {
            t.fi.push_back(e.first);
            t.fw.push_back(e.second);
        }
        t.rp.push_back(int(t.fi.size()));
    }
    // padded lists: every coarse index gets pw entries
    for (int c = 0; c < nc; ++c) t.pw = std::max(t.pw, t.rp[c + 1] - t.rp[c]);
    t.pfi.assign(size_t(nc) * t.pw, 0);
    t.pfw.assign(size_t(nc) * t.pw, 0.0);
    for (int c = 0; c < nc; ++c)
        for (int e = 0; e < t.pw; ++e) {
            const int q = t.rp[c] + e;
            const bool real = q < t.rp[c + 1];
            t.pfi[size_t(c) * t.pw + e] = real ? t.fi[q] : (t.rp[c] < t.rp[c + 1] ? t.fi[t.rp[c]] : 0);
            t.pfw[size_t(c) * t.pw + e] = real ? t.fw[q] : 0.0;
        }
    return t;
}

}  // namespace

struct helm3d_mg {
    int L = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<MgLevel> lev;
    void* vec_buf = nullptr;
    void* tab_buf = nullptr;
    ~helm3d_mg()
    {
        (void)hipSetDevice(device);
        if (vec_buf) (void)hipFree(vec_buf);
        if (tab_buf) (void)hipFree(tab_buf);
    }
};

namespace {

struct MgParams {
    int ks_outer = 3, ks_inner = 5, kc_outer = 3, kc_inner = 5;
    double smooth_tol = 1e-6, coarse_tol = 0.5;
};

// Level buffers viewed in the V-cycle's precision (a complex64 vector fits in
// a complex128 slot).
template <typename R>
cpx<R>* as(cpx<double>* p)
{
    return reinterpret_cast<cpx<R>*>(p);
}

template <typename R>
int vcycle(helm3d_mg* mg, const MgParams& k, int l, int mode, const cpx<R>* b, cpx<R>* x);

// Precondition with V_l: z = V_l(s * v)
template <typename R>
struct VCycle : PrecondT<R> {
    helm3d_mg* mg;
    const MgParams* k;
    int l, mode;
    int apply(const cpx<R>* v, double s, cpx<R>* z) override
    {
        MgLevel& L = mg->lev[l];
        HIPCHK((launch_axpby<R, R>(s, 0.0, v, 0.0, 0.0, as<R>(L.vb), L.N, false, mg->stream)));
        return vcycle<R>(mg, *k, l, mode, as<R>(L.vb), z);
    }
};

// Precondition the coarsest solve with another FGMRES on the coarsest level
// (construct: ls_opts.precond = LinearSolver(H{end}, coarse options)).
template <typename R>
struct CoarseInner : PrecondT<R> {
    helm3d_mg* mg;
    const MgParams* k;
    int mode;
    int apply(const cpx<R>* v, double s, cpx<R>* z) override
    {
        MgLevel& L = mg->lev[mg->L - 1];
        HIPCHK((launch_axpby<R, R>(s, 0.0, v, 0.0, 0.0, as<R>(L.vb), L.N, false, mg->stream)));
        GmresResult r;
        return gmres_engine<R>(L.S, model_variant(L.S, mode != 1), as<R>(L.vb), z, false, k->coarse_tol,
                               k->kc_inner, k->kc_outer, 0, nullptr, &r, nullptr, false);
    }
};

// Mixed precision: the FP64 outer FGMRES hands s * v to a complex64 V-cycle
// and takes its result back in complex128.
struct MixedTop : Precond {
    helm3d_mg* mg;
    const MgParams* k;
    int mode;
    int apply(const cpx<double>* v, double s, cpx<double>* z) override
    {
        MgLevel& L = mg->lev[0];
        cpx<float>* b32 = as<float>(L.vb);
        cpx<float>* x32 = as<float>(L.cx);
        HIPCHK((launch_axpby<double, float>(s, 0.0, v, 0.0, 0.0, b32, L.N, false, mg->stream)));
        CHK(vcycle<float>(mg, *k, 0, mode, b32, x32));
        HIPCHK((launch_axpby<float, double>(1.0, 0.0, x32, 0.0, 0.0, z, L.N, false, mg->stream)));
        return HELM3D_OK;
    }
};

template <typename R>
int residual(MgLevel& L, int mode, const cpx<R>* b, const cpx<R>* x, cpx<R>* r)
{
    StencilLaunch<R> S{};
    if constexpr (std::is_same<R, double>::value)
        S.op = L.A->op64;
    else
        S.op = L.A->op32;
    S.x = x;
    S.y = r;
    S.b = b;
    S.alpha = R(-1);
    S.beta = R(1);
    S.sink = static_cast<cpx<R>*>(L.A->sink);
    HIPCHK(launch_stencil<R>(model_variant(L.A, mode != 1), S, L.A->grid, L.A->stream));
    return HELM3D_OK;
}

template <typename R>
int vcycle(helm3d_mg* mg, const MgParams& k, int l, int mode, const cpx<R>* b, cpx<R>* x)
{
    MgLevel& L = mg->lev[l];
    MgLevel& C = mg->lev[l + 1];
    GmresResult r;
    // pre-smoothing from zero
    CHK(gmres_engine<R>(L.S, model_variant(L.S, mode != 1), b, x, false, k.smooth_tol, k.ks_inner, k.ks_outer, 0,
                        nullptr, &r, nullptr, false));
    CHK(residual<R>(L, mode, b, x, as<R>(L.vr)));
    HIPCHK(launch_restrict<R>(L.T, 0.125, as<R>(L.vr), as<R>(C.crhs), mg->stream));
    // coarse-grid solve, preconditioned by the next level
    if (l + 1 < mg->L - 1) {
        VCycle<R> P;
        P.mg = mg;
        P.k = &k;
        P.l = l + 1;
        P.mode = mode;
        CHK(gmres_engine<R>(C.A, model_variant(C.A, mode != 1), as<R>(C.crhs), as<R>(C.cx), false, k.coarse_tol,
                            k.kc_inner, k.kc_outer, 0, &P, &r, nullptr, false));
    } else {
        CoarseInner<R> P;
        P.mg = mg;
        P.k = &k;
        P.mode = mode;
        CHK(gmres_engine<R>(C.A, model_variant(C.A, mode != 1), as<R>(C.crhs), as<R>(C.cx), false, k.coarse_tol,
                            k.kc_inner, k.kc_outer, 0, &P, &r, nullptr, false));
    }
    HIPCHK(launch_prolong_add<R>(L.T, as<R>(C.cx), x, mg->stream));
    // post-smoothing from the corrected x
    CHK(gmres_engine<R>(L.S, model_variant(L.S, mode != 1), b, x, true, k.smooth_tol, k.ks_inner, k.ks_outer, 0,
                        nullptr, &r, nullptr, false));
    return HELM3D_OK;
}

int solve_mg(helm3d_mg* mg, bool mixed, int mode, const void* b, void* x, const helm3d_solve_opts* opts, const int* k,
             helm3d_solve_info* info)
{
    if (!mg || !b || !x) return fail(HELM3D_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(mg->device));
    helm3d_solve_opts o;
    helm3d_solve_opts_default(&o);
    if (opts) o = *opts;
    MgParams prm;
    if (k) {
        if (k[0] < 1 || k[1] < 1 || k[2] < 1 || k[3] < 1 || k[1] >= kMaxBasis || k[3] >= kMaxBasis)
            return fail(HELM3D_ERR_ARG, "bad multigrid cycle sizes");
        prm.ks_outer = k[0];
        prm.ks_inner = k[1];
        prm.kc_outer = k[2];
        prm.kc_inner = k[3];
    }
    for (auto& L : mg->lev) {
        L.A->stream = mg->stream;
        L.S->stream = mg->stream;
    }
    MgLevel& F = mg->lev[0];
    VCycle<double> P64;
    MixedTop P32;
    P64.mg = P32.mg = mg;
    P64.k = P32.k = &prm;
    P64.l = 0;
    P64.mode = P32.mode = mode;
    GmresResult r;
    CHK(gmres_engine<double>(F.A, model_variant(F.A, mode != 1), static_cast<const cpx<double>*>(b),
                             static_cast<cpx<double>*>(x), o.use_x0 != 0, o.tol, o.maxinnerit, long(o.maxit), 0,
                             mixed ? static_cast<Precond*>(&P32) : static_cast<Precond*>(&P64), &r, nullptr));
    if (info) {
        info->iters = r.iters;
        info->inner_iters = 0;
        info->cycles = r.cycles;
        info->relres = r.relres;
        info->quasi_relres = r.quasi;
        info->converged = r.converged;
    }
    return r.converged ? HELM3D_OK : fail(HELM3D_ERR_NOTCONV, "maxit reached above tolerance");
}

}  // namespace

extern "C" {

int helm3d_mg_create(int nlevels, helm3d_ctx* const* A, helm3d_ctx* const* S, helm3d_mg** out)
{
    if (!out) return fail(HELM3D_ERR_ARG, "null out");
    *out = nullptr;
    if (nlevels < 2 || nlevels > 8 || !A || !S) return fail(HELM3D_ERR_ARG, "need 2..8 levels");
    for (int l = 0; l < nlevels; ++l) {
        if (!A[l] || !S[l] || !A[l]->has_model || !S[l]->has_model)
            return fail(HELM3D_ERR_NOMODEL, "level without a model");
        for (int a = 0; a < 3; ++a)
            if (A[l]->n[a] != S[l]->n[a]) return fail(HELM3D_ERR_ARG, "smoothing context grid differs");
        if (A[l]->device != A[0]->device || S[l]->device != A[0]->device)
            return fail(HELM3D_ERR_ARG, "levels on different devices");
    }
    CHK(bind(A[0]));
    helm3d_mg* mg = new helm3d_mg();
    mg->L = nlevels;
    mg->device = A[0]->device;
    mg->stream = A[0]->stream;
    mg->lev.resize(nlevels);
    size_t vec_elems = 0;
    for (int l = 0; l < nlevels; ++l) {
        MgLevel& L = mg->lev[l];
        L.A = A[l];
        L.S = S[l];
        L.N = A[l]->N;
        // one stream for every context of the hierarchy
        A[l]->stream = mg->stream;
        S[l]->stream = mg->stream;
        vec_elems += 4 * size_t(slot_ld(L.N));
    }
    std::vector<AxisTab> tabs;
    size_t ints = 0, dbls = 0;
    for (int l = 0; l + 1 < nlevels; ++l)
        for (int a = 0; a < 3; ++a) {
            tabs.push_back(axis_tab(A[l]->n[a], A[l + 1]->n[a]));
            const AxisTab& t = tabs.back();
            ints += t.I.size() + t.rp.size() + t.fi.size() + t.pfi.size();
            dbls += t.wa.size() + t.wb.size() + t.fw.size() + t.pfw.size();
        }
    auto cleanup = [&](int rc) {
        delete mg;
        return rc;
    };
    if (hipMalloc(&mg->vec_buf, vec_elems * sizeof(cpx<double>)) != hipSuccess)
        return cleanup(fail(HELM3D_ERR_NOMEM, "multigrid vectors"));
    if (hipMalloc(&mg->tab_buf, dbls * sizeof(double) + ints * sizeof(int) + 64) != hipSuccess)
        return cleanup(fail(HELM3D_ERR_NOMEM, "multigrid tables"));
    cpx<double>* v = static_cast<cpx<double>*>(mg->vec_buf);
    for (int l = 0; l < nlevels; ++l) {
        MgLevel& L = mg->lev[l];
        const long ld = slot_ld(L.N);
        L.vb = v;
        L.vr = v + ld;
        L.crhs = v + 2 * ld;
        L.cx = v + 3 * ld;
        v += 4 * ld;
    }
    // tables: doubles first, then ints
    std::vector<double> hd;
    std::vector<int> hi;
    hd.reserve(dbls);
    hi.reserve(ints);
    struct Off {
        size_t I, wa, wb, rp, fi, fw, pfi, pfw;
    };
    std::vector<Off> offs;
    for (const AxisTab& t : tabs) {
        Off o;
        o.wa = hd.size();
        hd.insert(hd.end(), t.wa.begin(), t.wa.end());
        o.wb = hd.size();
        hd.insert(hd.end(), t.wb.begin(), t.wb.end());
        o.fw = hd.size();
        hd.insert(hd.end(), t.fw.begin(), t.fw.end());
        o.I = hi.size();
        hi.insert(hi.end(), t.I.begin(), t.I.end());
        o.rp = hi.size();
        hi.insert(hi.end(), t.rp.begin(), t.rp.end());
        o.fi = hi.size();
        hi.insert(hi.end(), t.fi.begin(), t.fi.end());
        o.pfw = hd.size();
        hd.insert(hd.end(), t.pfw.begin(), t.pfw.end());
        o.pfi = hi.size();
        hi.insert(hi.end(), t.pfi.begin(), t.pfi.end());
        offs.push_back(o);
    }
    double* dd = static_cast<double*>(mg->tab_buf);
    int* di = reinterpret_cast<int*>(dd + hd.size());
    if (hipMemcpy(dd, hd.data(), hd.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(di, hi.data(), hi.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(HELM3D_ERR_HIP, "multigrid table upload"));
    for (int l = 0; l + 1 < nlevels; ++l) {
        MgAxes& T = mg->lev[l].T;
        for (int a = 0; a < 3; ++a) {
            const Off& o = offs[size_t(l) * 3 + a];
            T.nf[a] = A[l]->n[a];
            T.nc[a] = A[l + 1]->n[a];
            T.ax[a] = MgAxis{di + o.I,  dd + o.wa,  dd + o.wb,  di + o.rp,
                             di + o.fi, dd + o.fw, di + o.pfi, dd + o.pfw, tabs[size_t(l) * 3 + a].pw};
        }
    }
    *out = mg;
    return HELM3D_OK;
}

int helm3d_mg_destroy(helm3d_mg* mg)
{
    delete mg;
    return HELM3D_OK;
}

int helm3d_mg_set_stream(helm3d_mg* mg, void* stream)
{
    if (!mg) return fail(HELM3D_ERR_ARG, "null hierarchy");
    mg->stream = static_cast<hipStream_t>(stream);
    for (auto& L : mg->lev) {
        L.A->stream = mg->stream;
        L.S->stream = mg->stream;
    }
    return HELM3D_OK;
}

int helm3d_mg_transfer(helm3d_mg* mg, int level, int direction, const void* in, void* out)
{
    if (!mg || level < 0 || level + 1 >= mg->L || !in || !out) return fail(HELM3D_ERR_ARG, "bad argument");
    HIPCHK(hipSetDevice(mg->device));
    const MgAxes& T = mg->lev[level].T;
    if (direction == 0) {
        HIPCHK(launch_prolong_add<double>(T, static_cast<const cpx<double>*>(in), static_cast<cpx<double>*>(out),
                                          mg->stream));
    } else {
        HIPCHK(launch_restrict<double>(T, 0.125, static_cast<const cpx<double>*>(in), static_cast<cpx<double>*>(out),
                                       mg->stream));
    }
    return HELM3D_OK;
}

int helm3d_solve_mg(helm3d_mg* mg, int mode, const void* b, void* x, const helm3d_solve_opts* opts, const int* k,
                    helm3d_solve_info* info)
{
    return solve_mg(mg, false, mode, b, x, opts, k, info);
}

int helm3d_solve_mg_mixed(helm3d_mg* mg, int mode, const void* b, void* x, const helm3d_solve_opts* opts,
                          const int* k, helm3d_solve_info* info)
{
    return solve_mg(mg, true, mode, b, x, opts, k, info);
}

}  // extern "C"
