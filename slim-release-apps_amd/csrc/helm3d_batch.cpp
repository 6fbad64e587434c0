// Batched solves: many right-hand sides, possibly on different operators and
// grids (sources x frequencies x models), advanced together so that every
// kernel type runs as one launch over all items per Krylov step.
//
// Outer method: the restarted (flexible) GMRES of FGMRES.m:33-128 per item,
// in lockstep as FGMRES_MT.m:40-135 does for several right-hand sides: one
// cycle loop and one Arnoldi step loop for all items, per-item Hessenberg
// matrices and least squares on the host, one device-to-host copy and one
// synchronisation per step for the whole batch.
//   HELM3D_BATCH_PLAIN  complex128 GMRES(m), identity preconditioner; the
//                       step loop runs while the largest quasi-residual is
//                       above tol and the cycle loop while the largest true
//                       residual is (FGMRES_MT.m:78,92).
//   HELM3D_BATCH_MIXED  FP64 flexible GMRES whose preconditioner is a
//                       fixed number of complex64 BiCGStab iterations; an
//                       item leaves the step loop when its own quasi-residual
//                       is below tol and leaves the batch when its true
//                       residual is.
// The inner BiCGStab uses the fused stencil forms (epilogues 3 and 4 of the
// stencil body): per iteration three streaming passes instead of seven, with
// every scalar on the device; its whole sequence for all items is captured
// once into a HIP graph and replayed per outer step.
//
// The basis is kept unnormalised on the device: ||u_c||^2 of every basis
// vector is reduced into the item's small buffer and the update kernel forms
// a_c = d_c / ||u_c||^2 itself, so no coefficient crosses to the device
// during a cycle.
#include "batch.hpp"
#include "capi_internal.hpp"

namespace {

constexpr int kSmallStride = 4 * kMaxBasis + 8;  // doubles per item
constexpr int kDots = 0;                          // small[0 .. 2k): dots of the step
constexpr int kNrm = 2 * kMaxBasis;               // small[kNrm + c]: ||u_c||^2
constexpr int kSc = 2 * kMaxBasis;                // sc stride per item (BS_SLOTS used)

// complex64 workspace slots of the inner BiCGStab
// (B is also the shadow residual rhat and the residual before the first
// iteration: the inner solve starts from z = 0 and never writes it)
enum : int { S32_B = 0, S32_Z, S32_R0, S32_R1, S32_P0, S32_P1, S32_V0, S32_V1, S32_T, S32_COUNT };

template <typename T>
int upload(T* dev, const std::vector<T>& host, hipStream_t s)
{
    if (host.empty()) return HELM3D_OK;
    HIPCHK(hipMemcpyAsync(dev, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return HELM3D_OK;
}

}  // namespace

struct helm3d_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t cap_stream = nullptr;
    void* ws64 = nullptr;
    size_t ws64_cap = 0;
    void* ws32 = nullptr;
    size_t ws32_cap = 0;
    double* part = nullptr;
    size_t part_cap = 0;
    double* small = nullptr;
    size_t small_cap = 0;
    double* sc = nullptr;
    size_t sc_cap = 0;
    double* coef = nullptr;  // device: per item 2*kMaxBasis combine coefficients
    size_t coef_cap = 0;
    double* small_host = nullptr;  // pinned
    size_t small_host_cap = 0;
    double* coef_host = nullptr;  // pinned
    size_t coef_host_cap = 0;
    int* act_host = nullptr;  // pinned active flags
    size_t act_host_cap = 0;
    void* sink = nullptr;
    size_t sink_cap = 0;
    void* desc = nullptr;  // device descriptors (items, stencil launches, prefix sums, active flags)
    size_t desc_cap = 0;
    hipGraphExec_t graph = nullptr;
    ~helm3d_batch()
    {
        (void)hipSetDevice(device);
        if (graph) (void)hipGraphExecDestroy(graph);
        void* bufs[] = {ws64, ws32, part, small, sc, coef, sink, desc};
        for (void* b : bufs)
            if (b) (void)hipFree(b);
        if (small_host) (void)hipHostFree(small_host);
        if (coef_host) (void)hipHostFree(coef_host);
        if (act_host) (void)hipHostFree(act_host);
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
    }
};

namespace {

template <typename T>
int host_alloc(T** p, size_t* cap, size_t count)
{
    if (*cap >= count && *p) return HELM3D_OK;
    if (*p) HIPCHK(hipHostFree(*p));
    *p = nullptr;
    *cap = 0;
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(p), count * sizeof(T)));
    *cap = count;
    return HELM3D_OK;
}

struct ItemState {
    double normr0 = 0, beta = 0, relres = 1, quasi = 1;
    long iters = 0, inner = 0, cycles = 0;
    int k = 0;  // Arnoldi steps done in the current cycle
    bool done = false;      // left the batch (converged, zero RHS or frozen)
    bool stepping = false;  // takes part in the current step
    std::vector<std::complex<double>> H, y;
    std::vector<double> sig;
};

struct Layout {
    int n = 0, m = 0;
    bool mixed = false;
    int variant = 0;
    int ty = 4, pf = 2, ry = 1;
    std::vector<long> N, ld, off64, off32;
    std::vector<int> nblk, sblk, first;
    std::vector<StencilGrid> grid;
    int nblk_max = 1, sblk_total = 0;
    long total_points = 0;  // sum of the items' grid sizes
    // device descriptor regions (byte offsets into b->desc)
    size_t o_i64 = 0, o_i32 = 0, o_st64 = 0, o_res64 = 0, o_k1 = 0, o_k2 = 0, o_k1f = 0, o_first = 0, o_active = 0;
    size_t total = 0;
};

template <typename T>
T* at(helm3d_batch* b, size_t off) { return reinterpret_cast<T*>(static_cast<char*>(b->desc) + off); }

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

template <typename R>
StencilLaunch<R> base_launch(const OpDesc<R>& op, const StencilGrid& g, void* sink)
{
    StencilLaunch<R> L{};
    L.op = op;
    L.alpha = R(1);
    L.sink = static_cast<cpx<R>*>(sink);
    L.tiles_x = g.tiles_x;
    L.tiles_y = g.tiles_y;
    L.kchunk = g.kchunk;
    return L;
}

}  // namespace

extern "C" {

int helm3d_batch_create(int device, void* stream, helm3d_batch** out)
{
    if (!out) return fail(HELM3D_ERR_ARG, "null out");
    *out = nullptr;
    int dev = device;
    if (dev < 0) HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipSetDevice(dev));
    helm3d_batch* b = new helm3d_batch();
    b->device = dev;
    b->stream = static_cast<hipStream_t>(stream);
    *out = b;
    return HELM3D_OK;
}

int helm3d_batch_destroy(helm3d_batch* b)
{
    delete b;
    return HELM3D_OK;
}

int helm3d_batch_set_stream(helm3d_batch* b, void* stream)
{
    if (!b) return fail(HELM3D_ERR_ARG, "null batch");
    b->stream = static_cast<hipStream_t>(stream);
    return HELM3D_OK;
}

int helm3d_batch_item_bytes(const helm3d_ctx* c, int method, int maxinnerit, size_t* bytes)
{
    if (!bytes) return fail(HELM3D_ERR_ARG, "null bytes");
    *bytes = 0;
    if (!c || !c->has_model) return fail(HELM3D_ERR_NOMODEL, "context without a model");
    if (method != HELM3D_BATCH_PLAIN && method != HELM3D_BATCH_MIXED) return fail(HELM3D_ERR_ARG, "bad method");
    if (maxinnerit < 1 || maxinnerit > kMaxBasis - 1) return fail(HELM3D_ERR_ARG, "maxinnerit must be in [1, 31]");
    const bool mixed = method == HELM3D_BATCH_MIXED;
    const size_t ld = size_t(slot_ld(c->N));
    const size_t slots64 = size_t(maxinnerit + 1) + (mixed ? maxinnerit : 0);
    const size_t slots32 = mixed ? S32_COUNT : 0;
    const auto g = stencil_grid<double>(c->n[0], c->n[1], c->n[2], c->grid.kchunk, c->grid.ty, c->grid.pf);
    const size_t part = size_t(std::max(krylov_blocks(c->N), g.blocks)) * 2 * kMaxBasis * sizeof(double);
    *bytes = ld * (slots64 * sizeof(cpx<double>) + slots32 * sizeof(cpx<float>)) + part +
             sizeof(double) * (kSmallStride + kSc + 2 * kMaxBasis) + 4096;
    return HELM3D_OK;
}

int helm3d_batch_solve(helm3d_batch* B, int n, helm3d_ctx* const* ctxs, int mode, const void* const* bs,
                       void* const* xs, int method, const helm3d_solve_opts* opts, helm3d_solve_info* info)
{
    using C = std::complex<double>;
    if (!B) return fail(HELM3D_ERR_ARG, "null batch");
    if (n < 1 || n > kBatchMaxItems) return fail(HELM3D_ERR_ARG, "item count must be in [1, 1024]");
    if (!ctxs || !bs || !xs) return fail(HELM3D_ERR_ARG, "null argument");
    if (method != HELM3D_BATCH_PLAIN && method != HELM3D_BATCH_MIXED) return fail(HELM3D_ERR_ARG, "bad method");
    helm3d_solve_opts o;
    helm3d_solve_opts_default(&o);
    if (opts) o = *opts;
    const int m = o.maxinnerit;
    if (m < 1 || m > kMaxBasis - 1) return fail(HELM3D_ERR_ARG, "maxinnerit must be in [1, 31]");
    if (o.use_x0) return fail(HELM3D_ERR_ARG, "batched solves start from x = 0");
    HIPCHK(hipSetDevice(B->device));
    const bool wnc = ctxs[0] ? ctxs[0]->wnc : false;
    for (int i = 0; i < n; ++i) {
        helm3d_ctx* c = ctxs[i];
        if (!c || !c->has_model) return fail(HELM3D_ERR_NOMODEL, "item without a model");
        if (c->device != B->device) return fail(HELM3D_ERR_ARG, "items on another device");
        if (c->wnc != wnc) return fail(HELM3D_ERR_ARG, "all items must have real wn, or all complex wn");
        if (!bs[i] || !xs[i]) return fail(HELM3D_ERR_ARG, "null vector");
    }
    const hipStream_t st = B->stream;

    // ---- layout
    Layout Lo;
    Lo.n = n;
    Lo.m = m;
    Lo.mixed = method == HELM3D_BATCH_MIXED;
    Lo.variant = (mode != 1 ? VAR_ADJ : VAR_FORW) | (wnc ? VAR_WNC : 0);
    Lo.ty = ctxs[0]->grid.ty;
    Lo.pf = ctxs[0]->grid.pf;
    Lo.ry = ctxs[0]->grid.ry;
    const int slots64 = (m + 1) + (Lo.mixed ? m : 0);
    const int slots32 = Lo.mixed ? S32_COUNT : 0;
    long tot64 = 0, tot32 = 0;
    size_t part_tot = 0;
    Lo.first.push_back(0);
    for (int i = 0; i < n; ++i) {
        helm3d_ctx* c = ctxs[i];
        Lo.N.push_back(c->N);
        Lo.total_points += c->N;
        Lo.ld.push_back(slot_ld(c->N));
        Lo.off64.push_back(tot64);
        Lo.off32.push_back(tot32);
        tot64 += Lo.ld.back() * slots64;
        tot32 += Lo.ld.back() * slots32;
        Lo.grid.push_back(stencil_grid<double>(c->n[0], c->n[1], c->n[2], c->grid.kchunk, Lo.ty, Lo.pf));
        Lo.nblk.push_back(krylov_blocks(c->N));
        Lo.sblk.push_back(Lo.grid.back().blocks);
        Lo.nblk_max = std::max(Lo.nblk_max, Lo.nblk.back());
        Lo.sblk_total += Lo.sblk.back();
        Lo.first.push_back(Lo.sblk_total);
        part_tot += size_t(std::max(Lo.nblk.back(), Lo.sblk.back())) * 2 * kMaxBasis;
    }
    CHK(ensure(&B->ws64, &B->ws64_cap, size_t(tot64) * sizeof(cpx<double>)));
    if (Lo.mixed) CHK(ensure(&B->ws32, &B->ws32_cap, size_t(tot32) * sizeof(cpx<float>)));
    CHK(ensure_t(&B->part, &B->part_cap, part_tot));
    CHK(ensure_t(&B->small, &B->small_cap, size_t(n) * kSmallStride));
    CHK(ensure_t(&B->sc, &B->sc_cap, size_t(n) * kSc));
    CHK(ensure_t(&B->coef, &B->coef_cap, size_t(n) * 2 * kMaxBasis));
    CHK(host_alloc(&B->small_host, &B->small_host_cap, size_t(n) * kSmallStride));
    CHK(host_alloc(&B->coef_host, &B->coef_host_cap, size_t(n) * 2 * kMaxBasis));
    CHK(host_alloc(&B->act_host, &B->act_host_cap, size_t(n)));
    CHK(ensure(&B->sink, &B->sink_cap, size_t(kSinkSlots) * sizeof(cpx<double>)));

    // descriptor regions
    size_t off = 0;
    auto region = [&](size_t bytes) {
        const size_t o = off;
        off = align256(off + bytes);
        return o;
    };
    Lo.o_i64 = region(sizeof(VecItem<double>) * n);
    Lo.o_i32 = region(sizeof(VecItem<float>) * n);
    Lo.o_st64 = region(sizeof(StencilLaunch<double>) * n * (m + 1));  // step k = 1..m at index k
    Lo.o_res64 = region(sizeof(StencilLaunch<double>) * n);
    Lo.o_k1 = region(sizeof(StencilLaunch<float>) * n * 2);
    Lo.o_k2 = region(sizeof(StencilLaunch<float>) * n * 2);
    Lo.o_k1f = region(sizeof(StencilLaunch<float>) * n);
    Lo.o_first = region(sizeof(int) * (n + 1));
    Lo.o_active = region(sizeof(int) * n);
    Lo.total = off;
    // the captured inner graph refers to the descriptors: rebuild it whenever
    // they are rewritten
    if (B->graph) {
        HIPCHK(hipGraphExecDestroy(B->graph));
        B->graph = nullptr;
    }
    CHK(ensure(&B->desc, &B->desc_cap, Lo.total));

    // ---- descriptors
    std::vector<VecItem<double>> i64(n);
    std::vector<VecItem<float>> i32(n);
    std::vector<StencilLaunch<double>> st64(size_t(n) * (m + 1)), res64(n);
    std::vector<StencilLaunch<float>> k1(size_t(n) * 2), k2(size_t(n) * 2);
    // first iteration: beta = alpha = omega = 0 and r = b
    std::vector<StencilLaunch<float>> k1first(Lo.mixed ? n : 0);
    size_t poff = 0;
    for (int i = 0; i < n; ++i) {
        helm3d_ctx* c = ctxs[i];
        cpx<double>* w64 = static_cast<cpx<double>*>(B->ws64) + Lo.off64[i];
        const long ld = Lo.ld[i];
        VecItem<double>& I = i64[i];
        I.ws = w64;
        I.ld = ld;
        I.n = Lo.N[i];
        I.nblk = Lo.nblk[i];
        I.sblk = Lo.sblk[i];
        I.part = B->part + poff;
        I.small = B->small + size_t(i) * kSmallStride;
        I.sc = B->sc + size_t(i) * kSc;
        I.b = static_cast<const cpx<double>*>(bs[i]);
        I.x = static_cast<cpx<double>*>(xs[i]);
        poff += size_t(std::max(Lo.nblk[i], Lo.sblk[i])) * 2 * kMaxBasis;
        auto V = [&](int s) { return w64 + long(s) * ld; };
        auto Z = [&](int s) { return w64 + long(m + 1 + s) * ld; };
        for (int k = 1; k <= m; ++k) {
            StencilLaunch<double> L = base_launch<double>(c->op64, Lo.grid[i], B->sink);
            L.x = Lo.mixed ? Z(k - 1) : V(k - 1);
            L.y = V(k);
            st64[size_t(k) * n + i] = L;
        }
        StencilLaunch<double> R0 = base_launch<double>(c->op64, Lo.grid[i], B->sink);
        R0.x = I.x;
        R0.y = V(0);
        R0.b = I.b;
        R0.alpha = -1.0;
        R0.beta = 1.0;
        R0.norm_part = I.part;
        res64[i] = R0;
        if (Lo.mixed) {
            cpx<float>* w32 = static_cast<cpx<float>*>(B->ws32) + Lo.off32[i];
            auto S = [&](int s) { return w32 + long(s) * ld; };
            VecItem<float>& J = i32[i];
            J = VecItem<float>{S(0), ld, Lo.N[i], Lo.nblk[i], Lo.sblk[i], I.part, I.small, I.sc, I.b, I.x};
            for (int par = 0; par < 2; ++par) {
                const int cur = par, nxt = par ^ 1;
                // iteration i (parity cur = i & 1) reads r, p, v of
                // iteration i-1 from the cur slots and t, and writes the
                // nxt slots; z is updated in place
                StencilLaunch<float> A = base_launch<float>(c->op32, Lo.grid[i], B->sink);
                A.f0 = S(cur == 0 ? S32_R0 : S32_R1);
                A.f1 = S(S32_P0 + cur);
                A.f2 = S(S32_V0 + cur);
                A.f3 = S(S32_T);
                A.o2 = S(S32_P0 + nxt);
                A.o3 = S(nxt == 0 ? S32_R0 : S32_R1);
                A.zz = S(S32_Z);
                A.y = S(S32_V0 + nxt);
                A.rh = S(S32_B);
                A.sc = I.sc;
                A.rho_new = cur == 0 ? BS_RHO0 : BS_RHO1;
                A.rho_old = cur == 0 ? BS_RHO1 : BS_RHO0;
                A.norm_part = I.part;
                k1[size_t(par) * n + i] = A;
                if (par == 0) {
                    A.first = 1;
                    A.f0 = S(S32_B);
                    k1first[i] = A;
                }
                StencilLaunch<float> Bk = base_launch<float>(c->op32, Lo.grid[i], B->sink);
                Bk.f0 = S(nxt == 0 ? S32_R0 : S32_R1);
                Bk.f1 = S(S32_V0 + nxt);
                Bk.y = S(S32_T);
                Bk.rh = S(S32_B);
                Bk.sc = I.sc;
                Bk.rho_new = cur == 0 ? BS_RHO0 : BS_RHO1;
                Bk.norm_part = I.part;
                k2[size_t(par) * n + i] = Bk;
            }
        }
    }
    int* active = B->act_host;
    for (int i = 0; i < n; ++i) active[i] = 1;
    CHK(upload(at<VecItem<double>>(B, Lo.o_i64), i64, st));
    if (Lo.mixed) CHK(upload(at<VecItem<float>>(B, Lo.o_i32), i32, st));
    CHK(upload(at<StencilLaunch<double>>(B, Lo.o_st64), st64, st));
    CHK(upload(at<StencilLaunch<double>>(B, Lo.o_res64), res64, st));
    if (Lo.mixed) {
        CHK(upload(at<StencilLaunch<float>>(B, Lo.o_k1), k1, st));
        CHK(upload(at<StencilLaunch<float>>(B, Lo.o_k2), k2, st));
    }
    CHK(upload(at<int>(B, Lo.o_first), Lo.first, st));
    HIPCHK(hipMemcpyAsync(at<int>(B, Lo.o_active), active, size_t(n) * sizeof(int), hipMemcpyHostToDevice, st));
    StencilLaunch<float>* d_k1first = at<StencilLaunch<float>>(B, Lo.o_k1f);
    if (Lo.mixed) CHK(upload(d_k1first, k1first, st));

    const VecItem<double>* d_i64 = at<VecItem<double>>(B, Lo.o_i64);
    const VecItem<float>* d_i32 = at<VecItem<float>>(B, Lo.o_i32);
    const int* d_first = at<int>(B, Lo.o_first);
    int* d_active = at<int>(B, Lo.o_active);
    const int nbm = Lo.nblk_max;

    // ---- inner solve sequence (fixed iterations), all items: two stencil
    // launches and two reductions per BiCGStab iteration (every vector
    // update is fused into the stencil launches), then the last z update
    const long inner_its = std::max<long>(1, o.inner_maxit);
    auto inner_seq = [&](hipStream_t s) -> int {
        HIPCHK(launch_bbicg_init<float>(d_i32, d_active, n, nbm, S32_B, S32_P0, S32_V0, S32_T, S32_Z, s));
        HIPCHK(launch_breduce<float>(d_i32, d_active, n, 3, BS_RHO0, BT_SC, false, s));
        for (long it = 0; it < inner_its; ++it) {
            const int par = int(it & 1);
            const StencilLaunch<float>* d1 =
                it == 0 ? d_k1first : at<StencilLaunch<float>>(B, Lo.o_k1) + size_t(par) * n;
            HIPCHK(launch_stencil_batch<float>(Lo.variant, 3, d1, d_first, n, Lo.sblk_total, d_active, Lo.ty,
                                               Lo.pf, s, Lo.ry));
            HIPCHK(launch_breduce<float>(d_i32, d_active, n, 2, BS_RV, BT_SC, true, s));
            HIPCHK(launch_stencil_batch<float>(Lo.variant, 4, at<StencilLaunch<float>>(B, Lo.o_k2) + size_t(par) * n,
                                               d_first, n, Lo.sblk_total, d_active, Lo.ty, Lo.pf, s, Lo.ry));
            HIPCHK(launch_breduce<float>(d_i32, d_active, n, 7, BS_TS, BT_SC, true, s));
        }
        const int last = int(inner_its & 1);  // slots of the last iteration's p, r, v
        HIPCHK(launch_bbicg_fin<float>(d_i32, d_active, n, nbm, S32_P0 + last, last == 0 ? S32_R0 : S32_R1,
                                       S32_V0 + last, S32_Z, s));
        return HELM3D_OK;
    };
    auto run_inner = [&]() -> int {
        if (!B->graph) {
            if (!B->cap_stream) HIPCHK(hipStreamCreateWithFlags(&B->cap_stream, hipStreamNonBlocking));
            // descriptors must be in place before the captured kernels run
            HIPCHK(hipStreamSynchronize(st));
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(B->cap_stream, hipStreamCaptureModeThreadLocal));
            const int rc = inner_seq(B->cap_stream);
            const hipError_t e = hipStreamEndCapture(B->cap_stream, &g);
            if (rc != HELM3D_OK) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            HIPCHK(e);
            const hipError_t ei = hipGraphInstantiate(&B->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIPCHK(ei);
        }
        HIPCHK(hipGraphLaunch(B->graph, st));
        return HELM3D_OK;
    };

    // ---- host state
    std::vector<ItemState> S(n);
    for (auto& s : S) {
        s.H.assign(size_t(m + 1) * m, C(0));
        s.sig.assign(m + 1, 0.0);
    }
    // The pinned flags are rewritten only once the copy of their previous
    // contents has run (every step ends with a synchronisation; the cycle
    // end synchronises here when two updates follow each other).
    bool act_pending = true;
    auto set_active = [&](auto pred) -> int {
        std::vector<int> want(n);
        bool changed = false;
        for (int i = 0; i < n; ++i) {
            want[i] = pred(i) ? 1 : 0;
            changed = changed || want[i] != active[i];
        }
        if (!changed) return HELM3D_OK;
        if (act_pending) HIPCHK(hipStreamSynchronize(st));
        for (int i = 0; i < n; ++i) active[i] = want[i];
        HIPCHK(hipMemcpyAsync(d_active, active, size_t(n) * sizeof(int), hipMemcpyHostToDevice, st));
        act_pending = true;
        return HELM3D_OK;
    };
    auto fetch = [&]() -> int {
        HIPCHK(hipMemcpyAsync(B->small_host, B->small, size_t(n) * kSmallStride * sizeof(double),
                              hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        act_pending = false;
        return HELM3D_OK;
    };
    auto hs = [&](int i, int q) { return B->small_host[size_t(i) * kSmallStride + q]; };

    // r0 = b, x = 0 (FGMRES.m:57-62); ||r0||^2 -> small[kNrm]
    HIPCHK(launch_bstart(d_i64, d_active, n, nbm, 0, st));
    HIPCHK(launch_bnorm2<double>(d_i64, d_active, n, nbm, 0, st));
    HIPCHK(launch_breduce<double>(d_i64, d_active, n, 1, kNrm, BT_SMALL, false, st));
    CHK(fetch());
    for (int i = 0; i < n; ++i) {
        S[i].normr0 = std::sqrt(hs(i, kNrm));
        S[i].beta = S[i].normr0;
        if (S[i].normr0 == 0.0) {  // zero right-hand side: x = 0 (opAbstractMatrix.m:118)
            S[i].done = true;
            S[i].relres = 0.0;
            S[i].quasi = 0.0;
        }
    }
    long it_counter = 1;
    const double tol = o.tol;
    auto any_open = [&]() {
        for (auto& s : S)
            if (!s.done) return true;
        return false;
    };
    while (any_open() && (it_counter / m) < o.maxit) {
        for (auto& s : S) {
            if (s.done) continue;
            std::fill(s.H.begin(), s.H.end(), C(0));
            s.sig[0] = 1.0 / s.beta;
            s.k = 0;
            s.stepping = true;
            s.quasi = 1.0;
        }
        int k = 1;
        while (k <= m) {
            CHK(set_active([&](int i) { return !S[i].done && S[i].stepping; }));
            bool any = false;
            for (auto& s : S) any = any || (!s.done && s.stepping);
            if (!any) break;
            if (Lo.mixed) {
                // z_{k-1} = M^{-1} (v_{k-1}), v = u / ||u||  (FGMRES.m:94)
                HIPCHK(launch_bdown_norm(d_i64, d_i32, d_active, n, nbm, kNrm + k - 1, k - 1, S32_B, st));
                CHK(run_inner());
                HIPCHK(launch_bup(d_i32, d_i64, d_active, n, nbm, S32_Z, m + 1 + (k - 1), st));
            }
            HIPCHK(launch_stencil_batch<double>(Lo.variant, 0, at<StencilLaunch<double>>(B, Lo.o_st64) + size_t(k) * n,
                                                d_first, n, Lo.sblk_total, d_active, Lo.ty, Lo.pf, st, Lo.ry));
            // classical Gram-Schmidt (FGMRES.m:100-103) on the unnormalised basis
            // streaming basis loads once the slots read per pass exceed the last-level cache
            const bool nt = double(k) * double(Lo.total_points) * sizeof(cpx<double>) > 256.0 * (1 << 20);
            HIPCHK(launch_bdots<double>(d_i64, d_active, n, nbm, k, 0, k, st, nt));
            HIPCHK(launch_breduce<double>(d_i64, d_active, n, 2 * k, kDots, BT_SMALL, false, st));
            HIPCHK(launch_bupdate_nrm<double>(d_i64, d_active, n, nbm, k, 0, k, kNrm, st, nt));
            HIPCHK(launch_breduce<double>(d_i64, d_active, n, 1, kNrm + k, BT_SMALL, false, st));
            CHK(fetch());
            double qmax = 0.0;
            for (int i = 0; i < n; ++i) {
                ItemState& s = S[i];
                if (s.done || !s.stepping) continue;
                // w was formed from v_{k-1} (mixed: true-scaled z) or from the raw
                // u_{k-1} (plain): h_q = sig_q * d_q (* sig_{k-1} for plain)
                const double fw = Lo.mixed ? 1.0 : s.sig[k - 1];
                for (int q = 0; q < k; ++q)
                    s.H[q + (k - 1) * (m + 1)] = s.sig[q] * fw * C(hs(i, kDots + 2 * q), hs(i, kDots + 2 * q + 1));
                const double un = std::sqrt(hs(i, kNrm + k));
                s.H[k + (k - 1) * (m + 1)] = fw * un;
                s.sig[k] = un > 0.0 ? 1.0 / un : 0.0;
                s.quasi = hess_lsq(k, m + 1, s.H, s.beta, s.y) / s.normr0;
                s.k = k;
                ++s.iters;
                if (Lo.mixed) s.inner += 2 * inner_its;
                // breakdown (w in the span) ends this item's cycle
                if (un == 0.0 || !(s.quasi == s.quasi)) s.stepping = false;
                qmax = std::max(qmax, s.quasi);
                if (Lo.mixed && s.quasi <= tol) s.stepping = false;
            }
            ++it_counter;
            ++k;
            if (!Lo.mixed && !(qmax > tol)) break;  // FGMRES_MT.m:92, hst(end) = max over items
        }
        // x += V y (plain) or Z y (mixed), FGMRES.m:115-119
        int kmax = 0;
        for (auto& s : S)
            if (!s.done) kmax = std::max(kmax, s.k);
        CHK(set_active([&](int i) { return !S[i].done && S[i].k > 0; }));
        for (int i = 0; i < n; ++i) {
            ItemState& s = S[i];
            double* a = B->coef_host + size_t(i) * 2 * kMaxBasis;
            std::fill(a, a + 2 * kMaxBasis, 0.0);
            if (s.done) continue;
            for (int q = 0; q < s.k; ++q) {
                const C v = Lo.mixed ? s.y[q] : s.y[q] * s.sig[q];
                a[2 * q] = v.real();
                a[2 * q + 1] = v.imag();
            }
        }
        HIPCHK(hipMemcpyAsync(B->coef, B->coef_host, size_t(n) * 2 * kMaxBasis * sizeof(double),
                              hipMemcpyHostToDevice, st));
        if (kmax > 0)
            HIPCHK(launch_bcombine(d_i64, d_active, n, nbm, kmax, Lo.mixed ? m + 1 : 0, B->coef, st,
                                   double(kmax) * double(Lo.total_points) * sizeof(cpx<double>) > 256.0 * (1 << 20)));
        // true residual r = b - A x into slot 0 and its norm (FGMRES.m:120-124)
        CHK(set_active([&](int i) { return !S[i].done; }));
        HIPCHK(launch_stencil_batch<double>(Lo.variant, 1, at<StencilLaunch<double>>(B, Lo.o_res64), d_first, n,
                                            Lo.sblk_total, d_active, Lo.ty, Lo.pf, st, Lo.ry));
        HIPCHK(launch_breduce<double>(d_i64, d_active, n, 1, kNrm, BT_SMALL, true, st));
        CHK(fetch());
        bool all_below = true;
        for (int i = 0; i < n; ++i) {
            ItemState& s = S[i];
            if (s.done) continue;
            const double rn = std::sqrt(hs(i, kNrm));
            s.relres = rn / s.normr0;
            s.beta = rn;
            ++s.cycles;
            if (!(s.relres < tol)) all_below = false;
            if (Lo.mixed && s.relres < tol) s.done = true;
            if (rn == 0.0) s.done = true;
        }
        if (!Lo.mixed && all_below) {
            for (auto& s : S) s.done = true;
        }
    }
    HIPCHK(hipStreamSynchronize(st));
    bool all_conv = true;
    for (int i = 0; i < n; ++i) {
        const ItemState& s = S[i];
        const bool conv = s.relres < tol;
        all_conv = all_conv && conv;
        if (info) {
            info[i].iters = s.iters;
            info[i].inner_iters = s.inner;
            info[i].cycles = s.cycles;
            info[i].relres = s.relres;
            info[i].quasi_relres = s.quasi;
            info[i].converged = conv;
        }
    }
    return all_conv ? HELM3D_OK : fail(HELM3D_ERR_NOTCONV, "maxit reached above tolerance");
}

}  // extern "C"
