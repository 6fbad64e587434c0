// C ABI of the MI355X Helmholtz library (declared in include/helm3d.h).
//
// Host side of the hot path: operator contexts (device-resident model, PML
// tables, Krylov workspace), the drop-in matvec of the MEX boundary, and the
// restarted GMRES driver of FGMRES.m with its FP64 Hessenberg least squares.
#include <algorithm>
#include <type_traits>

#include "capi_internal.hpp"

namespace helm3d {
thread_local std::string g_err;
}  // namespace helm3d

namespace {

// Fold completed event pairs into the per-kind totals (stream must be idle).
int prof_fold(helm3d_ctx* c)
{
    for (const auto& pr : c->ev_pending) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev_pool[pr.second], c->ev_pool[pr.second + 1]));
        c->prof_ms[pr.first] += ms;
        c->prof_n[pr.first] += 1;
    }
    c->ev_pending.clear();
    c->ev_used = 0;
    return HELM3D_OK;
}

// Record the start of a timed launch of class `kind`; returns the slot.
int prof_begin(helm3d_ctx* c, int kind, long* slot)
{
    *slot = -1;
    if (!c->prof) return HELM3D_OK;
    if (c->ev_used + 2 > c->ev_pool.size()) {
        if (c->ev_pool.size() >= 8192) {
            HIPCHK(hipStreamSynchronize(c->stream));
            CHK(prof_fold(c));
        } else {
            for (int q = 0; q < 256; ++q) {
                hipEvent_t e;
                HIPCHK(hipEventCreate(&e));
                c->ev_pool.push_back(e);
            }
        }
    }
    *slot = long(c->ev_used);
    c->ev_pending.push_back({kind, c->ev_used});
    c->ev_used += 2;
    HIPCHK(hipEventRecord(c->ev_pool[*slot], c->stream));
    return HELM3D_OK;
}
int prof_end(helm3d_ctx* c, long slot)
{
    if (slot < 0) return HELM3D_OK;
    HIPCHK(hipEventRecord(c->ev_pool[slot + 1], c->stream));
    return HELM3D_OK;
}

// Wrap a launch expression returning hipError_t with profiling of `kind`.
#define TIMED(kind, expr)                   \
    do {                                    \
        long slot_;                         \
        CHK(prof_begin(c, (kind), &slot_)); \
        HIPCHK(expr);                       \
        CHK(prof_end(c, slot_));            \
    } while (0)

template <typename R>
OpDesc<R>& op_of(helm3d_ctx* c);
template <>
OpDesc<double>& op_of<double>(helm3d_ctx* c) { return c->op64; }
template <>
OpDesc<float>& op_of<float>(helm3d_ctx* c) { return c->op32; }

template <typename R>
Vecs<R>& basis_of(helm3d_ctx* c);
template <>
Vecs<double>& basis_of<double>(helm3d_ctx* c) { return c->V64; }
template <>
Vecs<float>& basis_of<float>(helm3d_ctx* c) { return c->V32; }

template <typename R>
Vecs<R>& zbasis_of(helm3d_ctx* c);
template <>
Vecs<double>& zbasis_of<double>(helm3d_ctx* c) { return c->Z64; }
template <>
Vecs<float>& zbasis_of<float>(helm3d_ctx* c) { return c->Z32; }



// Stencil launch on the context's operator; y = alpha*Hx + beta*b.
template <typename R>
int stencil(helm3d_ctx* c, int variant, const OpDesc<R>& op, const cpx<R>* x, cpx<R>* y, R alpha,
            const cpx<R>* b, R beta, double* norm_part, const int* skip = nullptr)
{
    StencilLaunch<R> L{};
    L.skip = skip;
    L.op = op;
    L.x = x;
    L.y = y;
    L.b = b;
    L.alpha = alpha;
    L.beta = beta;
    L.norm_part = norm_part;
    L.sink = static_cast<cpx<R>*>(c->sink);
    TIMED(HELM3D_PROF_STENCIL, launch_stencil<R>(variant, L, c->grid, c->stream));
    return HELM3D_OK;
}

// Reduce `nq` partial columns (part[q*nblk + b]) into c->small[off..off+nq)
// (and into copy[0..nq) when given).
int reduce_to_small(helm3d_ctx* c, int nblk, int nq, int off, double* copy = nullptr)
{
    TIMED(HELM3D_PROF_REDUCE, launch_reduce(c->part, nblk, nq, c->small + off, c->stream, copy));
    return HELM3D_OK;
}
int fetch_small(helm3d_ctx* c, int count)
{
    HIPCHK(hipMemcpyAsync(c->small_host, c->small, count * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->prof) CHK(prof_fold(c));
    return HELM3D_OK;
}

}  // namespace

namespace helm3d {

// Launch the captured Arnoldi cycle of precision R on the context stream,
// capturing it first (on the capture stream, through `record`) when no graph
// exists for the current launch arguments: the operator descriptor, launch
// shape, cycle length and every buffer address the steps use.
template <typename R, typename Rec>
int cycle_graph(helm3d_ctx* c, int variant, int m, const cpx<R>* V, long ld, int nblk, const int* skip,
                Rec&& record)
{
    const int p = std::is_same<R, double>::value ? 0 : 1;
    std::vector<unsigned char> key;
    auto put = [&](const void* a, size_t n) {
        const auto* b = static_cast<const unsigned char*>(a);
        key.insert(key.end(), b, b + n);
    };
    const OpDesc<R>& op = op_of<R>(c);
    const void* ptrs[] = {V, c->part, c->nrm, c->gs, c->ring_dev, c->sink, skip};
    const long ints[] = {variant, m, ld, nblk, c->N, c->grid.tiles_x, c->grid.tiles_y, c->grid.kchunk,
                         c->grid.blocks, c->grid.ty, c->grid.pf, c->grid.ry};
    put(&op, sizeof(op));
    put(ptrs, sizeof(ptrs));
    put(ints, sizeof(ints));
    if (!c->cyc_graph[p] || c->cyc_key[p] != key) {
        if (c->cyc_graph[p]) {
            HIPCHK(hipGraphExecDestroy(c->cyc_graph[p]));
            c->cyc_graph[p] = nullptr;
        }
        if (!c->cap_stream) HIPCHK(hipStreamCreateWithFlags(&c->cap_stream, hipStreamNonBlocking));
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(c->cap_stream, hipStreamCaptureModeThreadLocal));
        const int st = record(c->cap_stream);
        const hipError_t e = hipStreamEndCapture(c->cap_stream, &g);
        if (st != HELM3D_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        HIPCHK(e);
        const hipError_t ei = hipGraphInstantiate(&c->cyc_graph[p], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(ei);
        c->cyc_key[p] = std::move(key);
    }
    HIPCHK(hipGraphLaunch(c->cyc_graph[p], c->stream));
    return HELM3D_OK;
}

// Restarted GMRES of FGMRES.m:33-128 on precision R.  The basis is kept
// unnormalised (slot c raw, scale sig[c]); P makes it the flexible variant of
// FGMRES.m:96-98,118 with Z stored true-scaled.
//
// Without P the Arnoldi steps run ahead of the host: the squared norms of the
// basis slots stay on the device (c->nrm), where the next matvec takes its
// input scale and the update pass its coefficients, so steps are enqueued
// before the host reads the dot products of earlier ones (written into pinned
// rows, one per step).  On large grids one step runs ahead (the device is
// busy anyway); on grids of at most kDeepLookahead points -- coarse multigrid
// levels, where a step is a few short kernels -- the whole cycle is enqueued
// at once and the host meets the device once per cycle instead of once per
// step.  The host builds the Hessenberg columns and quasi-residuals in order;
// steps enqueued past the stopping point are discarded (their slots are never
// combined), so the arithmetic is that of the one-step-at-a-time loop.
//
// Calls without P that want only x (multigrid smoothers and coarsest-level
// solves, at most kMaxDeviceCycles cycles) never meet the host: every cycle
// the loop bound allows is enqueued, and device kernels take the decisions --
// the least squares per cycle (Givens rotations as the columns come, the
// quasi-residual |g_k| / ||r0|| of the same least squares), the true residual
// per restart, a zero initial residual -- with the kernels of the cycles
// after the stop reading a device flag and returning.
template <typename R>
int gmres_engine(helm3d_ctx* c, int variant, const cpx<R>* b, cpx<R>* x, bool use_x0, double tol, int m,
                 long maxit, long iter_cap, PrecondT<R>* P, GmresResult* res, std::vector<double>* hst,
                 bool need_result)
{
    using C = std::complex<double>;
    if (m < 1 || m > kMaxBasis - 1) return fail(HELM3D_ERR_ARG, "maxinnerit must be in [1, 31]");
    // Result-free calls without a preconditioner (multigrid smoothers and
    // coarsest solves) run every decision on the device: no host round trip.
    const bool device_driven =
        !P && !need_result && !hst && iter_cap <= 0 && tol < 1.0 && maxit >= 1 && maxit <= kMaxDeviceCycles;
    const long N = c->N;
    const long ld = slot_ld(N);
    OpDesc<R>& op = op_of<R>(c);
    Vecs<R>& Vb = basis_of<R>(c);
    CHK(ensure_vecs(Vb, N, m + 1));
    CHK(ensure_work(c));
    cpx<R>* V = Vb.p;
    cpx<R>* Z = nullptr;
    if (P) {
        Vecs<R>& Zb = zbasis_of<R>(c);
        CHK(ensure_vecs(Zb, N, m));
        Z = Zb.p;
    }
    const int nblk = krylov_blocks(N);
    // Gram-Schmidt passes: their own block counts on long complex128 vectors
    // (equal to nblk on short ones; complex64 keeps nblk); the solution update
    // x += V y streams like the update pass and takes its count
    constexpr bool kWide = std::is_same<R, double>::value;
    const int nblk_d = kWide ? dots_blocks(N) : nblk;
    const int nblk_u = kWide ? update_blocks(N) : nblk;
    auto slot = [&](int s) { return V + size_t(s) * ld; };
    GmresDevState* st = c->gst;
    const int* skip = device_driven ? &st->skip : nullptr;

    // r0 = b - A x0, or b with x = 0 (FGMRES.m:57-62); ||r0||^2 into nrm[0]
    // and the pinned mirror
    if (use_x0) {
        CHK(stencil<R>(c, variant, op, x, slot(0), R(-1), b, R(1), c->part));
        TIMED(HELM3D_PROF_REDUCE, launch_reduce(c->part, c->grid.blocks, 1, c->nrm, c->stream, c->small_host_dev));
    } else {
        HIPCHK(launch_start<R>(b, slot(0), x, N, c->part, nblk, c->stream));
        TIMED(HELM3D_PROF_REDUCE, launch_reduce(c->part, nblk, 1, c->nrm, c->stream, c->small_host_dev));
    }
    // enqueue Arnoldi step j: slot(j) = A slot(j-1)/||slot(j-1)||, dots into
    // row j, update with device scales, ||slot(j)||^2 into row j and nrm[j]
    auto enqueue_on = [&](int j, hipStream_t s, bool record) -> int {
        double* row = c->gs + size_t(j) * kGsRow;
        StencilLaunch<R> L{};
        L.op = op;
        L.x = slot(j - 1);
        L.y = slot(j);
        L.alpha = R(1);
        L.beta = R(0);
        L.anrm = c->nrm + (j - 1);
        L.sink = static_cast<cpx<R>*>(c->sink);
        L.skip = skip;
        TIMED(HELM3D_PROF_STENCIL, launch_stencil<R>(variant, L, c->grid, s));
        // reduced dots into row (the update's coefficients) and the
        // pinned ring row, ||slot(j)||^2 into nrm[j] and the ring row
        double* hrow = device_driven ? row : c->ring_dev + size_t(j) * kGsRow;
        if (nblk <= kInBlockDots) {
            // short dots columns: summed by every block of the update
            // pass (partials one column up, clear of the norm partials)
            TIMED(HELM3D_PROF_DOTS, launch_dots<R>(V, ld, j, slot(j), N, c->part + nblk, nblk, s, skip));
            TIMED(HELM3D_PROF_UPDATE, launch_update<R>(V, ld, j, nullptr, CoefPack{}, slot(j), N, c->part,
                                                       nblk, s, c->nrm, c->part + nblk, hrow, skip));
        } else {
            TIMED(HELM3D_PROF_DOTS, launch_dots<R>(V, ld, j, slot(j), N, c->part, nblk_d, s, skip));
            TIMED(HELM3D_PROF_REDUCE, launch_reduce(c->part, nblk_d, 2 * j, row, s, hrow, skip));
            TIMED(HELM3D_PROF_UPDATE, launch_update<R>(V, ld, j, row, CoefPack{}, slot(j), N, c->part, nblk_u, s,
                                                       c->nrm, nullptr, nullptr, skip));
        }
        const int nu = nblk <= kInBlockDots ? nblk : nblk_u;
        TIMED(HELM3D_PROF_REDUCE, launch_reduce(c->part, nu, 1, c->nrm + j, s, hrow + 2 * j, skip));
        if (record) HIPCHK(hipEventRecord(c->ring_ev[j], s));
        return HELM3D_OK;
    };

    // a whole cycle of steps ahead of the host: replayed from a captured
    // graph on small grids
    auto enqueue_cycle = [&](bool use_graph) -> int {
        if (use_graph)
            return cycle_graph<R>(c, variant, m, V, ld, nblk, skip, [&](hipStream_t s) -> int {
                for (int j = 1; j <= m; ++j) CHK(enqueue_on(j, s, false));
                return HELM3D_OK;
            });
        for (int j = 1; j <= m; ++j) CHK(enqueue_on(j, c->stream, false));
        return HELM3D_OK;
    };

    if (device_driven) {
        // Every cycle the host loop below could run (floor(it/m) < maxit with
        // m steps per cycle) is enqueued; the device takes its decisions: the
        // least squares and the quasi-residual test after the cycle's steps
        // (steps past the first one at or below tol are dropped from the
        // combine, as the host loop would not have run them), the true
        // residual test at a restart, and a zero initial residual.  The
        // kernels of a cycle after the stop return at once.
        HIPCHK(launch_gmres_start(c->nrm, st, c->stream));
        const bool use_graph = N <= kDeepLookahead && c->use_graphs && !c->prof;
        for (long cyc = 0; (1 + cyc * long(m)) / m < maxit; ++cyc) {
            if (cyc > 0) {
                HIPCHK(launch_gmres_gate(st, c->stream));
                CHK(stencil<R>(c, variant, op, x, slot(0), R(-1), b, R(1), c->part, skip));
                TIMED(HELM3D_PROF_REDUCE, launch_reduce(c->part, c->grid.blocks, 1, c->nrm, c->stream, nullptr, skip));
                HIPCHK(launch_gmres_restart(c->nrm, tol, st, c->stream));
            }
            CHK(enqueue_cycle(use_graph));
            HIPCHK(launch_gmres_lsq(c->gs, kGsRow, m, tol, st, c->stream));
            TIMED(HELM3D_PROF_COMBINE, launch_combine_dev<R>(V, ld, m, st->coef, &st->kept, x, N, nblk_u, c->stream, skip));
        }
        // no host copy of the outcome: the caller uses x only
        res->iters = -1;
        res->cycles = -1;
        res->relres = res->quasi = std::nan("");
        res->converged = false;
        return HELM3D_OK;
    }

    // Without P and with no result wanted, ||r0|| is read with the first
    // step's row (one host round trip less): the steps need it only on the
    // device, and with r0 = 0 they are discarded.
    bool defer_norm = !P && !need_result && maxit >= 1 && tol < 1.0;
    double normr0 = 0.0, beta = 0.0;
    double hlast = 1.0;
    if (hst) hst->push_back(1.0);
    auto take_norm = [&]() {
        normr0 = std::sqrt(c->small_host[0]);
        beta = normr0;
    };
    auto zero_residual = [&]() {
        // zero residual: the solution is x (opAbstractMatrix.m:118 returns zeros for b = 0)
        res->relres = 0.0;
        res->quasi = 0.0;
        res->converged = true;
        return HELM3D_OK;
    };
    if (!defer_norm) {
        HIPCHK(hipStreamSynchronize(c->stream));
        if (c->prof) CHK(prof_fold(c));
        take_norm();
        if (normr0 == 0.0) return zero_residual();
    }
    long it = 1, inner = 0;
    std::vector<C> H(size_t(m + 1) * m), y;
    std::vector<double> sig(m + 1, 0.0);
    bool capped = false;

    while (hlast > tol && (it / m) < maxit) {
        std::fill(H.begin(), H.end(), C(0));
        sig[0] = defer_norm ? 0.0 : 1.0 / beta;
        int k = 1;
        if (!P) {
            auto enqueue = [&](int j) { return enqueue_on(j, c->stream, true); };
            // may a step follow once `done` steps have been taken (iteration cap)
            auto may_continue = [&](long done) { return iter_cap <= 0 || done < iter_cap; };
            // steps in flight: the one the host waits for and those after it
            const int depth = N <= kDeepLookahead ? m : 2;
            int next = 1;  // next step of the cycle to enqueue
            // a whole cycle ahead: replay it from a captured graph, one event
            // after its last step
            const bool graph = depth == m && c->use_graphs && !c->prof && may_continue(inner + m - 1);
            if (graph) {
                CHK(enqueue_cycle(true));
                HIPCHK(hipEventRecord(c->ring_ev[m], c->stream));
                next = m + 1;
            }
            while (k <= m && hlast > tol) {
                while (next <= m && next < k + depth && may_continue(inner + (next - k))) CHK(enqueue(next++));
                HIPCHK(hipEventSynchronize(c->ring_ev[graph ? m : k]));
                if (defer_norm) {
                    take_norm();
                    if (normr0 == 0.0) return zero_residual();
                    sig[0] = 1.0 / beta;
                    defer_norm = false;
                }
                const double* h = c->ring_host + size_t(k) * kGsRow;
                for (int q = 0; q < k; ++q) H[q + (k - 1) * (m + 1)] = sig[q] * C(h[2 * q], h[2 * q + 1]);
                const double hn = std::sqrt(h[2 * k]);
                H[k + (k - 1) * (m + 1)] = hn;
                sig[k] = 1.0 / hn;
                hlast = hess_lsq(k, m + 1, H, beta, y) / normr0;
                if (hst) hst->push_back(hlast);
                ++k;
                ++it;
                ++inner;
                if (!may_continue(inner)) {
                    capped = true;
                    break;
                }
            }
        }
        while (P && k <= m && hlast > tol) {
            // w = A z_k into slot k
            {
                CHK(P->apply(slot(k - 1), sig[k - 1], Z + size_t(k - 1) * ld));
                CHK(stencil<R>(c, variant, op, Z + size_t(k - 1) * ld, slot(k), R(1), nullptr, R(0), nullptr));
            }
            // classical Gram-Schmidt, FGMRES.m:100-103; d (2k values) and
            // ||w||^2 reach c->small and its pinned mirror
            CoefPack scl{};
            for (int q = 0; q < k; ++q) scl.v[q] = sig[q] * sig[q];
            TIMED(HELM3D_PROF_DOTS, launch_dots<R>(V, ld, k, slot(k), N, c->part, nblk_d, c->stream));
            TIMED(HELM3D_PROF_REDUCE, launch_reduce(c->part, nblk_d, 2 * k, c->small, c->stream, c->small_host_dev));
            TIMED(HELM3D_PROF_UPDATE,
                  launch_update<R>(V, ld, k, c->small, scl, slot(k), N, c->part, nblk_u, c->stream));
            TIMED(HELM3D_PROF_REDUCE,
                  launch_reduce(c->part, nblk_u, 1, c->small + 2 * k, c->stream, c->small_host_dev + 2 * k));
            HIPCHK(hipStreamSynchronize(c->stream));
            if (c->prof) CHK(prof_fold(c));
            for (int q = 0; q < k; ++q)
                H[q + (k - 1) * (m + 1)] = sig[q] * C(c->small_host[2 * q], c->small_host[2 * q + 1]);
            const double hn = std::sqrt(c->small_host[2 * k]);
            H[k + (k - 1) * (m + 1)] = hn;
            sig[k] = 1.0 / hn;
            hlast = hess_lsq(k, m + 1, H, beta, y) / normr0;
            if (hst) hst->push_back(hlast);
            ++k;
            ++it;
            ++inner;
            if (iter_cap > 0 && inner >= iter_cap) {
                capped = true;
                break;
            }
        }
        // x += V y (or Z y), FGMRES.m:115-119
        CoefPack a{};
        for (int q = 0; q < k - 1; ++q) {
            const C v = P ? y[q] : y[q] * sig[q];
            a.v[2 * q] = v.real();
            a.v[2 * q + 1] = v.imag();
        }
        if (P) {
            TIMED(HELM3D_PROF_COMBINE, launch_combine<R>(Z, ld, k - 1, a, x, N, nblk_u, c->stream));
        } else {
            TIMED(HELM3D_PROF_COMBINE, launch_combine<R>(V, ld, k - 1, a, x, N, nblk_u, c->stream));
        }
        // no further cycle (quasi-residual below tol, maxit or the iteration
        // cap): the true residual decides nothing, so a caller that uses only
        // x skips it
        if (!need_result && (hlast <= tol || (it / m) >= maxit || capped)) {
            res->relres = hlast;
            ++res->cycles;
            break;
        }
        // true residual r = b - A x into slot 0, FGMRES.m:120-124
        CHK(stencil<R>(c, variant, op, x, slot(0), R(-1), b, R(1), c->part));
        CHK(reduce_to_small(c, c->grid.blocks, 1, 0, c->nrm));
        CHK(fetch_small(c, 1));
        const double rn = std::sqrt(c->small_host[0]);
        res->relres = rn / normr0;
        ++res->cycles;
        if (res->relres < tol) break;
        beta = rn;
        if (capped) break;
    }
    res->iters = inner;
    res->quasi = hlast;
    res->converged = res->relres < tol;
    return HELM3D_OK;
}

template int gmres_engine<double>(helm3d_ctx*, int, const cpx<double>*, cpx<double>*, bool, double, int, long, long,
                                  PrecondT<double>*, GmresResult*, std::vector<double>*, bool);
template int gmres_engine<float>(helm3d_ctx*, int, const cpx<float>*, cpx<float>*, bool, double, int, long, long,
                                 PrecondT<float>*, GmresResult*, std::vector<double>*, bool);

int model_variant(const helm3d_ctx* c, int mode_adj)
{
    return (mode_adj ? VAR_ADJ : VAR_FORW) | (c->wnc ? VAR_WNC : 0);
}

}  // namespace helm3d

namespace {

template <typename R>
Vecs<R>& bicg_of(helm3d_ctx* c);
template <>
Vecs<double>& bicg_of<double>(helm3d_ctx* c) { return c->BC64; }
template <>
Vecs<float>& bicg_of<float>(helm3d_ctx* c) { return c->BC32; }

// BiCGStab of BICGSTAB.m:37-74 (identity preconditioner) from z = 0, all
// scalars on the device; the host reads ||r|| every check_every iterations.
// Stops when the recursive residual ||r|| / ||b|| < tol (BICGSTAB.m:63) or
// after maxit iterations (two matvecs each).
template <typename R>
int bicgstab_engine(helm3d_ctx* c, int variant, const cpx<R>* b, cpx<R>* z, double tol, long maxit,
                    int check_every, long* iters_out, double* relres_out)
{
    const long N = c->N;
    const long ld = slot_ld(N);
    Vecs<R>& W = bicg_of<R>(c);
    CHK(ensure_vecs(W, N, 6));
    CHK(ensure_work(c));
    CHK(ensure_t(&c->bsc, &c->bsc_cap, BS_SLOTS));
    cpx<R>* r = W.p;
    cpx<R>* rh = W.p + ld;
    cpx<R>* p = W.p + 2 * ld;
    cpx<R>* v = W.p + 3 * ld;
    cpx<R>* sv = W.p + 4 * ld;
    cpx<R>* t = W.p + 5 * ld;
    const OpDesc<R>& op = op_of<R>(c);
    const int nblk = krylov_blocks(N);
    double* sc = c->bsc;
    const cpx<R>* none = nullptr;
    HIPCHK((launch_axpby<R, R>(1.0, 0.0, b, 0.0, 0.0, r, N, false, c->stream)));
    HIPCHK((launch_axpby<R, R>(1.0, 0.0, b, 0.0, 0.0, rh, N, false, c->stream)));
    HIPCHK((launch_axpby<R, R>(1.0, 0.0, b, 0.0, 0.0, p, N, false, c->stream)));
    HIPCHK((launch_axpby<R, R>(0.0, 0.0, none, 0.0, 0.0, z, N, false, c->stream)));
    PairArgs<R> P0{{rh, nullptr}, {r, nullptr}};
    HIPCHK(launch_dots_pairs<R>(P0, 1, true, N, c->part, nblk, c->stream));
    TIMED(HELM3D_PROF_REDUCE, launch_reduce(c->part, nblk, 3, sc + BS_RHO0, c->stream));
    HIPCHK(hipMemcpyAsync(c->small_host, sc + BS_RHO0 + 2, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const double nb = std::sqrt(c->small_host[0]);
    *iters_out = 0;
    *relres_out = 0.0;
    if (nb == 0.0) return HELM3D_OK;
    double rel = 1.0;
    long it = 0;
    for (; it < maxit; ++it) {
        const int cur = (it & 1) ? BS_RHO1 : BS_RHO0;
        const int nxt = (it & 1) ? BS_RHO0 : BS_RHO1;
        CHK(stencil<R>(c, variant, op, p, v, R(1), nullptr, R(0), nullptr));
        PairArgs<R> P1{{rh, nullptr}, {v, nullptr}};
        TIMED(HELM3D_PROF_DOTS, launch_dots_pairs<R>(P1, 1, false, N, c->part, nblk, c->stream));
        TIMED(HELM3D_PROF_REDUCE, launch_reduce(c->part, nblk, 2, sc + BS_RV, c->stream));
        TIMED(HELM3D_PROF_UPDATE, launch_bicg_s<R>(r, v, sv, N, sc, cur, c->stream));
        CHK(stencil<R>(c, variant, op, sv, t, R(1), nullptr, R(0), nullptr));
        PairArgs<R> P2{{t, nullptr}, {sv, nullptr}};
        TIMED(HELM3D_PROF_DOTS, launch_dots_pairs<R>(P2, 1, true, N, c->part, nblk, c->stream));
        TIMED(HELM3D_PROF_REDUCE, launch_reduce(c->part, nblk, 3, sc + BS_TS, c->stream));
        TIMED(HELM3D_PROF_UPDATE, launch_bicg_xr<R>(p, sv, t, rh, z, r, N, sc, c->part, nblk, c->stream));
        TIMED(HELM3D_PROF_REDUCE, launch_reduce(c->part, nblk, 3, sc + nxt, c->stream));
        if ((it + 1) % check_every == 0 || it + 1 == maxit) {
            HIPCHK(hipMemcpyAsync(c->small_host, sc + nxt + 2, sizeof(double), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            if (c->prof) CHK(prof_fold(c));
            rel = std::sqrt(c->small_host[0]) / nb;
            if (!(rel >= tol)) {  // converged (or NaN: stop)
                ++it;
                break;
            }
        }
        TIMED(HELM3D_PROF_COMBINE, launch_bicg_p<R>(r, v, p, N, sc, nxt, cur, c->stream));
    }
    *iters_out = it;
    *relres_out = rel;
    return HELM3D_OK;
}

// Fixed-length BiCGStab (exactly `iters` iterations, no host reads, no
// profiling events) on stream st: the body captured into a HIP graph.  The
// breakdown guards of the scalar kernels (zero denominators give zero steps)
// make a zero right-hand side return z = 0.
template <typename R>
int bicgstab_fixed(helm3d_ctx* c, hipStream_t st, int variant, const cpx<R>* b, cpx<R>* z, long iters)
{
    const long N = c->N;
    const long ld = slot_ld(N);
    cpx<R>* W = bicg_of<R>(c).p;
    cpx<R>* r = W;
    cpx<R>* rh = W + ld;
    cpx<R>* p = W + 2 * ld;
    cpx<R>* v = W + 3 * ld;
    cpx<R>* sv = W + 4 * ld;
    cpx<R>* t = W + 5 * ld;
    const OpDesc<R>& op = op_of<R>(c);
    const int nblk = krylov_blocks(N);
    double* sc = c->bsc;
    const cpx<R>* none = nullptr;
    auto sten = [&](const cpx<R>* x, cpx<R>* y) -> int {
        StencilLaunch<R> L{};
        L.op = op;
        L.x = x;
        L.y = y;
        L.alpha = R(1);
        L.sink = static_cast<cpx<R>*>(c->sink);
        HIPCHK(launch_stencil<R>(variant, L, c->grid, st));
        return HELM3D_OK;
    };
    HIPCHK((launch_axpby<R, R>(1.0, 0.0, b, 0.0, 0.0, r, N, false, st)));
    HIPCHK((launch_axpby<R, R>(1.0, 0.0, b, 0.0, 0.0, rh, N, false, st)));
    HIPCHK((launch_axpby<R, R>(1.0, 0.0, b, 0.0, 0.0, p, N, false, st)));
    HIPCHK((launch_axpby<R, R>(0.0, 0.0, none, 0.0, 0.0, z, N, false, st)));
    PairArgs<R> P0{{rh, nullptr}, {r, nullptr}};
    HIPCHK(launch_dots_pairs<R>(P0, 1, true, N, c->part, nblk, st));
    HIPCHK(launch_reduce(c->part, nblk, 3, sc + BS_RHO0, st));
    for (long it = 0; it < iters; ++it) {
        const int cur = (it & 1) ? BS_RHO1 : BS_RHO0;
        const int nxt = (it & 1) ? BS_RHO0 : BS_RHO1;
        CHK(sten(p, v));
        PairArgs<R> P1{{rh, nullptr}, {v, nullptr}};
        HIPCHK(launch_dots_pairs<R>(P1, 1, false, N, c->part, nblk, st));
        HIPCHK(launch_reduce(c->part, nblk, 2, sc + BS_RV, st));
        HIPCHK(launch_bicg_s<R>(r, v, sv, N, sc, cur, st));
        CHK(sten(sv, t));
        PairArgs<R> P2{{t, nullptr}, {sv, nullptr}};
        HIPCHK(launch_dots_pairs<R>(P2, 1, true, N, c->part, nblk, st));
        HIPCHK(launch_reduce(c->part, nblk, 3, sc + BS_TS, st));
        HIPCHK(launch_bicg_xr<R>(p, sv, t, rh, z, r, N, sc, c->part, nblk, st));
        HIPCHK(launch_reduce(c->part, nblk, 3, sc + nxt, st));
        if (it + 1 < iters) HIPCHK(launch_bicg_p<R>(r, v, p, N, sc, nxt, cur, st));
    }
    return HELM3D_OK;
}

// Run the fixed-length complex64 inner solve T32b = BiCGStab(T32a) on the
// context stream, replaying a captured graph (captured on the first call for
// this variant and length).
int inner_fixed_c64(helm3d_ctx* c, int variant, long iters)
{
    CHK(ensure_vecs(c->BC32, c->N, 6));
    CHK(ensure_t(&c->bsc, &c->bsc_cap, BS_SLOTS));
    if (!c->use_graphs) return bicgstab_fixed<float>(c, c->stream, variant, c->T32a.p, c->T32b.p, iters);
    if (!c->inner_graph || c->inner_graph_variant != variant || c->inner_graph_iters != iters) {
        if (c->inner_graph) {
            HIPCHK(hipGraphExecDestroy(c->inner_graph));
            c->inner_graph = nullptr;
        }
        if (!c->cap_stream) HIPCHK(hipStreamCreateWithFlags(&c->cap_stream, hipStreamNonBlocking));
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(c->cap_stream, hipStreamCaptureModeThreadLocal));
        const int st = bicgstab_fixed<float>(c, c->cap_stream, variant, c->T32a.p, c->T32b.p, iters);
        const hipError_t e = hipStreamEndCapture(c->cap_stream, &g);
        if (st != HELM3D_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        HIPCHK(e);
        const hipError_t ei = hipGraphInstantiate(&c->inner_graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(ei);
        c->inner_graph_variant = variant;
        c->inner_graph_iters = iters;
    }
    HIPCHK(hipGraphLaunch(c->inner_graph, c->stream));
    return HELM3D_OK;
}

// complex64 inner solve as the flexible preconditioner of the FP64 outer loop
struct MixedPrecond : Precond {
    helm3d_ctx* c;
    int variant;
    double inner_tol;
    int inner_m;
    long inner_cap;
    int inner_solver = HELM3D_INNER_BICGSTAB;
    long inner_total = 0;
    int apply(const cpx<double>* v_raw, double s, cpx<double>* z) override
    {
        const long N = c->N;
        // T32a = c64(s * v), solve A t = T32a in complex64, z = c128(t)
        HIPCHK((launch_axpby<double, float>(s, 0.0, v_raw, 0.0, 0.0, c->T32a.p, N, false, c->stream)));
        if (inner_solver == HELM3D_INNER_BICGSTAB && inner_tol <= 0.0) {
            // fixed number of iterations: one captured graph per outer step
            CHK(inner_fixed_c64(c, variant, inner_cap));
            inner_total += 2 * inner_cap;
        } else if (inner_solver == HELM3D_INNER_BICGSTAB) {
            long its = 0;
            double rel = 0;
            CHK(bicgstab_engine<float>(c, variant, c->T32a.p, c->T32b.p, inner_tol, inner_cap, 4, &its, &rel));
            inner_total += 2 * its;  // matvecs
        } else {
            GmresResult r;
            CHK(gmres_engine<float>(c, variant, c->T32a.p, c->T32b.p, false, inner_tol, inner_m, 1L << 40,
                                    inner_cap, nullptr, &r, nullptr));
            inner_total += r.iters;
        }
        HIPCHK((launch_axpby<float, double>(1.0, 0.0, c->T32b.p, 0.0, 0.0, z, N, false, c->stream)));
        return HELM3D_OK;
    }
};

int check_model(helm3d_ctx* c)
{
    CHK(bind(c));
    if (!c->has_model) return fail(HELM3D_ERR_NOMODEL, "no model set");
    return HELM3D_OK;
}

std::mutex g_dropin_mu;
helm3d_ctx* g_dropin[64] = {nullptr};

}  // namespace

extern "C" {

const char* helm3d_version(void) { return "helm3d-mi355x 0.1"; }

int helm3d_profile(helm3d_ctx* c, int enable)
{
    CHK(bind(c));
    HIPCHK(hipStreamSynchronize(c->stream));
    CHK(prof_fold(c));
    for (int k = 0; k < HELM3D_PROF_KINDS; ++k) {
        c->prof_ms[k] = 0;
        c->prof_n[k] = 0;
    }
    c->prof = enable != 0;
    return HELM3D_OK;
}

int helm3d_profile_read(helm3d_ctx* c, int kind, double* total_ms, long long* launches)
{
    CHK(bind(c));
    if (kind < 0 || kind >= HELM3D_PROF_KINDS || !total_ms || !launches) return fail(HELM3D_ERR_ARG, "bad kind");
    HIPCHK(hipStreamSynchronize(c->stream));
    CHK(prof_fold(c));
    *total_ms = c->prof_ms[kind];
    *launches = c->prof_n[kind];
    return HELM3D_OK;
}

const char* helm3d_last_error(void) { return g_err.c_str(); }

int helm3d_create(int device, void* stream, helm3d_ctx** out)
{
    if (!out) return fail(HELM3D_ERR_ARG, "null out");
    *out = nullptr;
    int dev = device;
    if (dev < 0) HIPCHK(hipGetDevice(&dev));
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (dev >= ndev) return fail(HELM3D_ERR_ARG, "device index out of range");
    HIPCHK(hipSetDevice(dev));
    helm3d_ctx* c = new helm3d_ctx();
    c->device = dev;
    c->stream = static_cast<hipStream_t>(stream);
    *out = c;
    return HELM3D_OK;
}

int helm3d_destroy(helm3d_ctx* ctx)
{
    if (!ctx) return HELM3D_OK;
    delete ctx;
    return HELM3D_OK;
}

int helm3d_set_stream(helm3d_ctx* ctx, void* stream)
{
    if (!ctx) return fail(HELM3D_ERR_ARG, "null context");
    ctx->stream = static_cast<hipStream_t>(stream);
    return HELM3D_OK;
}

long long helm3d_numel(const helm3d_ctx* ctx) { return (ctx && ctx->has_model) ? ctx->N : 0; }

int helm3d_set_model(helm3d_ctx* c, const int n[3], const double h[3], const int npml[6], const void* wn_dev,
                     int wn_complex)
{
    CHK(bind(c));
    if (!n || !h || !npml || !wn_dev) return fail(HELM3D_ERR_ARG, "null argument");
    for (int a = 0; a < 3; ++a)
        if (n[a] < 1) return fail(HELM3D_ERR_NSIZE, "n must hold 3 positive integers");
    for (int a = 0; a < 6; ++a)
        if (npml[a] < 0) return fail(HELM3D_ERR_ARG, "npml must be non-negative");
    const long N = long(n[0]) * n[1] * n[2];
    if (N >= (1L << 40)) return fail(HELM3D_ERR_ARG, "grid too large");
    // a new model on the same grid keeps the launch shape (tuned or chosen)
    const bool same_grid = c->has_model && c->n[0] == n[0] && c->n[1] == n[1] && c->n[2] == n[2];
    std::memcpy(c->n, n, sizeof(c->n));
    std::memcpy(c->npml, npml, sizeof(c->npml));
    std::memcpy(c->h, h, sizeof(c->h));
    c->N = N;
    c->wnc = wn_complex != 0;
    c->hc = make_hconst(h);

    // PML tables: 6 arrays of n_a + 2 complex values
    const size_t len = size_t(n[0] + 2) * 2 + size_t(n[1] + 2) * 2 + size_t(n[2] + 2) * 2;
    std::vector<cpx<double>> t64(len);
    std::vector<cpx<float>> t32(len);
    size_t off[6], pos = 0;
    for (int a = 0; a < 3; ++a) {
        const PmlAxis ax = make_pml_axis(n[a], npml[2 * a], npml[2 * a + 1]);
        off[2 * a] = pos;
        for (int t = 0; t < n[a] + 2; ++t, ++pos) t64[pos] = cpx<double>{ax.lo[t].real(), ax.lo[t].imag()};
        off[2 * a + 1] = pos;
        for (int t = 0; t < n[a] + 2; ++t, ++pos) t64[pos] = cpx<double>{ax.hi[t].real(), ax.hi[t].imag()};
    }
    for (size_t q = 0; q < len; ++q) t32[q] = cpx<float>{float(t64[q].re), float(t64[q].im)};
    // interior index range per axis: slots i..i+2 all exactly (1, 1)
    int ilo[3], ihi[3];
    for (int a = 0; a < 3; ++a) pml_interior(n[a], npml[2 * a], npml[2 * a + 1], &ilo[a], &ihi[a]);
    CHK(ensure_t(&c->tab64, &c->tab64_cap, len));
    CHK(ensure_t(&c->tab32, &c->tab32_cap, len));
    HIPCHK(hipMemcpyAsync(c->tab64, t64.data(), len * sizeof(cpx<double>), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->tab32, t32.data(), len * sizeof(cpx<float>), hipMemcpyHostToDevice, c->stream));

    // wn copies (FP64 as given, FP32 for the complex64 path)
    const size_t esz = c->wnc ? 2 : 1;
    CHK(ensure(&c->wn64, &c->wn64_cap, size_t(N) * esz * sizeof(double)));
    CHK(ensure(&c->wn32, &c->wn32_cap, size_t(N) * esz * sizeof(float)));
    HIPCHK(hipMemcpyAsync(c->wn64, wn_dev, size_t(N) * esz * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    if (c->wnc)
        HIPCHK((launch_axpby<double, float>(1.0, 0.0, static_cast<const cpx<double>*>(c->wn64), 0.0, 0.0,
                                            static_cast<cpx<float>*>(c->wn32), N, false, c->stream)));
    else
        HIPCHK(launch_cast_real(static_cast<const double*>(c->wn64), static_cast<float*>(c->wn32), N, c->stream));

    fill_opdesc(c->op64, n, c->hc, c->tab64, off, c->wn64, c->wnc, ilo, ihi);
    fill_opdesc(c->op32, n, c->hc, c->tab32, off, c->wn32, c->wnc, ilo, ihi);
    if (!same_grid) c->grid = stencil_grid<double>(n[0], n[1], n[2], 0);
    CHK(ensure_work(c));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->has_model = true;
    c->inner_graph_iters = -1;  // captured graphs refer to the old model's buffers
    return HELM3D_OK;
}

int helm3d_set_launch(helm3d_ctx* c, int kchunk, int tile_rows, int prefetch, int rows_per_thread)
{
    CHK(check_model(c));
    if (!(tile_rows == 4 || tile_rows == 8) || prefetch < 1 || prefetch > 4 || kchunk < 0 ||
        !(rows_per_thread == 1 || (rows_per_thread == 2 && tile_rows == 8)))
        return fail(HELM3D_ERR_ARG,
                    "tile_rows in {4,8}, prefetch in 1..4, kchunk >= 0, rows_per_thread 1 (or 2 with 8 rows)");
    c->grid = stencil_grid<double>(c->n[0], c->n[1], c->n[2], kchunk, tile_rows, prefetch, rows_per_thread);
    CHK(ensure_work(c));
    c->inner_graph_iters = -1;
    return HELM3D_OK;
}

int helm3d_set_option(helm3d_ctx* c, int option, long long value)
{
    CHK(bind(c));
    switch (option) {
    case HELM3D_OPT_GRAPHS:
        c->use_graphs = value != 0;
        return HELM3D_OK;
    default:
        return fail(HELM3D_ERR_ARG, "unknown option");
    }
}

int helm3d_get_launch(helm3d_ctx* c, int* kchunk, int* tile_rows, int* prefetch, int* rows_per_thread, int* blocks)
{
    CHK(check_model(c));
    if (kchunk) *kchunk = c->grid.kchunk;
    if (tile_rows) *tile_rows = c->grid.ty;
    if (prefetch) *prefetch = c->grid.pf;
    if (rows_per_thread) *rows_per_thread = c->grid.ry;
    if (blocks) *blocks = c->grid.blocks;
    return HELM3D_OK;
}

int helm3d_autotune(helm3d_ctx* c, int mode, int precision, int reps, int* kchunk, int* tile_rows, int* prefetch,
                    int* rows_per_thread, double* best_ms)
{
    CHK(check_model(c));
    if (precision != HELM3D_C128 && precision != HELM3D_C64) return fail(HELM3D_ERR_ARG, "bad precision");
    reps = std::max(1, std::min(reps, 1000));
    const int variant = model_variant(c, mode != 1);
    const long N = c->N;
    const long ld = slot_ld(N);
    // scratch: two vectors of the context's staging buffer (complex128
    // slots hold the complex64 pair as well)
    CHK(ensure_vecs(c->T64, N, 2));
    HIPCHK(hipMemsetAsync(c->T64.p, 0, size_t(2 * ld) * sizeof(cpx<double>), c->stream));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const StencilGrid keep = c->grid;
    StencilGrid best = keep;
    double best_t = 1e300;
    int rc = HELM3D_OK;
    // candidate z chunks: whole column, halves, ... down to 8 planes.  A grid
    // of few column tiles (the coarse multigrid levels) also tries chunks of
    // 4, 2 and 1 planes: its launches are short marches on a few compute
    // units, and shorter marches on more blocks cut their latency even though
    // each chunk reads its two halo planes again.
    std::vector<int> chunks;
    for (int d : {1, 2, 3, 4, 6, 8, 12}) {
        const int kc = (c->n[2] + d - 1) / d;
        if ((kc >= 8 || d == 1) && std::find(chunks.begin(), chunks.end(), kc) == chunks.end()) chunks.push_back(kc);
    }
    const long tiles = long((c->n[0] + 63) / 64) * ((c->n[1] + 3) / 4);
    if (tiles < 128) {
        for (int kc : {4, 2, 1}) {
            if (kc < c->n[2] && std::find(chunks.begin(), chunks.end(), kc) == chunks.end()) chunks.push_back(kc);
        }
    }
    // (tile rows, rows per thread): 64 x 4 and 64 x 8 tiles, the latter also
    // with two rows per thread
    const int shapes[3][2] = {{4, 1}, {8, 1}, {8, 2}};
    for (const auto& sh : shapes) {
        const int ty = sh[0], ry = sh[1];
        for (int pf : {1, 2, 3, 4}) {
            for (int kc : chunks) {
                c->grid = stencil_grid<double>(c->n[0], c->n[1], c->n[2], kc, ty, pf, ry);
                CHK(ensure_work(c));
                float ms = 0.f;
                auto run = [&]() -> int {
                    if (precision == HELM3D_C128) {
                        const cpx<double>* x = c->T64.p;
                        cpx<double>* y = c->T64.p + ld;
                        return stencil<double>(c, variant, c->op64, x, y, 1.0, nullptr, 0.0, nullptr);
                    }
                    const cpx<float>* x = reinterpret_cast<const cpx<float>*>(c->T64.p);
                    cpx<float>* y = reinterpret_cast<cpx<float>*>(c->T64.p + ld);
                    return stencil<float>(c, variant, c->op32, x, y, 1.0f, nullptr, 0.0f, nullptr);
                };
                for (int w = 0; w < 2 && rc == HELM3D_OK; ++w) rc = run();
                if (rc == HELM3D_OK && hipEventRecord(e0, c->stream) != hipSuccess) rc = fail(HELM3D_ERR_HIP, "event");
                for (int r = 0; r < reps && rc == HELM3D_OK; ++r) rc = run();
                if (rc == HELM3D_OK && hipEventRecord(e1, c->stream) != hipSuccess) rc = fail(HELM3D_ERR_HIP, "event");
                if (rc == HELM3D_OK && hipEventSynchronize(e1) != hipSuccess) rc = fail(HELM3D_ERR_HIP, "event sync");
                if (rc == HELM3D_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = fail(HELM3D_ERR_HIP, "event time");
                if (rc != HELM3D_OK) break;
                const double t = double(ms) / reps;
                if (t < best_t) {
                    best_t = t;
                    best = c->grid;
                }
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->grid = rc == HELM3D_OK ? best : keep;
    CHK(ensure_work(c));
    c->inner_graph_iters = -1;  // captured graphs hold the old launch shape
    if (rc != HELM3D_OK) return rc;
    if (kchunk) *kchunk = best.kchunk;
    if (tile_rows) *tile_rows = best.ty;
    if (prefetch) *prefetch = best.pf;
    if (rows_per_thread) *rows_per_thread = best.ry;
    if (best_ms) *best_ms = best_t;
    return HELM3D_OK;
}

int helm3d_apply(helm3d_ctx* c, int variant, int precision, const void* x, void* y)
{
    CHK(check_model(c));
    if (!x || !y) return fail(HELM3D_ERR_ARG, "null vector");
    const int v = model_variant(c, variant & HELM3D_ADJ);
    if (precision == HELM3D_C128)
        return stencil<double>(c, v, c->op64, static_cast<const cpx<double>*>(x), static_cast<cpx<double>*>(y), 1.0,
                               nullptr, 0.0, nullptr);
    if (precision == HELM3D_C64)
        return stencil<float>(c, v, c->op32, static_cast<const cpx<float>*>(x), static_cast<cpx<float>*>(y), 1.0f,
                              nullptr, 0.0f, nullptr);
    return fail(HELM3D_ERR_ARG, "bad precision");
}

int helm3d_apply_deriv(helm3d_ctx* c, int adjoint, int precision, const void* dwn, const void* x, void* y)
{
    CHK(check_model(c));
    if (!x || !y || !dwn) return fail(HELM3D_ERR_ARG, "null argument");
    const int v = (adjoint ? VAR_ADJ : VAR_FORW) | VAR_DERIV;
    if (precision == HELM3D_C128) {
        OpDesc<double> op = c->op64;
        op.wn = static_cast<const double*>(dwn);
        op.wnc = nullptr;
        return stencil<double>(c, v, op, static_cast<const cpx<double>*>(x), static_cast<cpx<double>*>(y), 1.0,
                               nullptr, 0.0, nullptr);
    }
    if (precision == HELM3D_C64) {
        OpDesc<float> op = c->op32;
        op.wn = static_cast<const float*>(dwn);
        op.wnc = nullptr;
        return stencil<float>(c, v, op, static_cast<const cpx<float>*>(x), static_cast<cpx<float>*>(y), 1.0f,
                              nullptr, 0.0f, nullptr);
    }
    return fail(HELM3D_ERR_ARG, "bad precision");
}

void helm3d_solve_opts_default(helm3d_solve_opts* o)
{
    if (!o) return;
    o->tol = 1e-6;
    o->maxinnerit = 20;
    o->maxit = 10000;
    o->use_x0 = 0;
    o->inner_tol = 1e-2;
    o->inner_m = 20;
    o->inner_maxit = 200;
    o->inner_solver = HELM3D_INNER_BICGSTAB;
}

int helm3d_solve_bicgstab(helm3d_ctx* c, int mode, const void* b, void* x, const helm3d_solve_opts* opts,
                          helm3d_solve_info* info)
{
    CHK(check_model(c));
    if (!b || !x) return fail(HELM3D_ERR_ARG, "null vector");
    helm3d_solve_opts o;
    helm3d_solve_opts_default(&o);
    if (opts) o = *opts;
    const int variant = model_variant(c, mode != 1);
    long its = 0;
    double rel = 0;
    auto* bb = static_cast<const cpx<double>*>(b);
    auto* xx = static_cast<cpx<double>*>(x);
    CHK(bicgstab_engine<double>(c, variant, bb, xx, o.tol, long(o.maxit), 1, &its, &rel));
    // true residual ||b - A x|| / ||b|| in FP64
    CHK(ensure_vecs(c->T64, c->N, 1));
    CHK(stencil<double>(c, variant, c->op64, xx, c->T64.p, -1.0, bb, 1.0, c->part));
    CHK(reduce_to_small(c, c->grid.blocks, 1, 0));
    HIPCHK(launch_norm2<double>(bb, c->N, c->part, krylov_blocks(c->N), c->stream));
    CHK(reduce_to_small(c, krylov_blocks(c->N), 1, 1));
    CHK(fetch_small(c, 2));
    const double relres = c->small_host[1] > 0 ? std::sqrt(c->small_host[0] / c->small_host[1]) : 0.0;
    if (info) {
        info->iters = 2 * its;
        info->inner_iters = 0;
        info->cycles = its;
        info->relres = relres;
        info->quasi_relres = rel;
        info->converged = rel < o.tol;
    }
    return rel < o.tol ? HELM3D_OK : fail(HELM3D_ERR_NOTCONV, "maxit reached above tolerance");
}

int helm3d_solve(helm3d_ctx* c, int mode, const void* b, void* x, const helm3d_solve_opts* opts,
                 helm3d_solve_info* info, double* hst, long long hst_cap, long long* n_hst)
{
    CHK(check_model(c));
    if (!b || !x) return fail(HELM3D_ERR_ARG, "null vector");
    helm3d_solve_opts o;
    helm3d_solve_opts_default(&o);
    if (opts) o = *opts;
    std::vector<double> h;
    GmresResult r;
    CHK(gmres_engine<double>(c, model_variant(c, mode != 1), static_cast<const cpx<double>*>(b),
                             static_cast<cpx<double>*>(x), o.use_x0 != 0, o.tol, o.maxinnerit, long(o.maxit), 0,
                             nullptr, &r, hst ? &h : nullptr));
    if (hst) {
        const long long nh = std::min<long long>(hst_cap, (long long)h.size());
        std::memcpy(hst, h.data(), size_t(nh) * sizeof(double));
        if (n_hst) *n_hst = nh;
    }
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

int helm3d_solve_mixed(helm3d_ctx* c, int mode, const void* b, void* x, const helm3d_solve_opts* opts,
                       helm3d_solve_info* info)
{
    CHK(check_model(c));
    if (!b || !x) return fail(HELM3D_ERR_ARG, "null vector");
    helm3d_solve_opts o;
    helm3d_solve_opts_default(&o);
    if (opts) o = *opts;
    CHK(ensure_vecs(c->T32a, c->N, 1));
    CHK(ensure_vecs(c->T32b, c->N, 1));
    MixedPrecond P;
    P.c = c;
    P.variant = model_variant(c, mode != 1);
    P.inner_tol = o.inner_tol;
    P.inner_m = o.inner_m;
    P.inner_cap = o.inner_maxit;
    P.inner_solver = o.inner_solver;
    GmresResult r;
    CHK(gmres_engine<double>(c, P.variant, static_cast<const cpx<double>*>(b), static_cast<cpx<double>*>(x),
                             o.use_x0 != 0, o.tol, o.maxinnerit, long(o.maxit), 0, &P, &r, nullptr));
    if (info) {
        info->iters = r.iters;
        info->inner_iters = P.inner_total;
        info->cycles = r.cycles;
        info->relres = r.relres;
        info->quasi_relres = r.quasi;
        info->converged = r.converged;
    }
    return r.converged ? HELM3D_OK : fail(HELM3D_ERR_NOTCONV, "maxit reached above tolerance");
}

int helm3d_fwi_grad_accum(helm3d_ctx* c, const double* dwn, const void* U, const void* V, double scale, int grid,
                          double* g)
{
    CHK(check_model(c));
    if (!U || !V || !g) return fail(HELM3D_ERR_ARG, "null argument");
    if (grid != HELM3D_GRID_PHYS && grid != HELM3D_GRID_PML) return fail(HELM3D_ERR_ARG, "bad grid");
    StencilLaunch<double> L{};
    L.op = c->op64;
    L.op.wn = dwn;
    L.op.wnc = nullptr;
    L.op.wn_const = 1.0;
    L.x = static_cast<const cpx<double>*>(V);
    L.y = nullptr;
    L.alpha = 1.0;
    L.gu = static_cast<const cpx<double>*>(U);
    L.gg = g;
    L.sink = static_cast<cpx<double>*>(c->sink);
    L.gscale = scale;
    for (int a = 0; a < 3; ++a) {
        const bool phys = grid == HELM3D_GRID_PHYS;
        L.p0[a] = phys ? c->npml[2 * a] : 0;
        L.pn[a] = phys ? c->n[a] - c->npml[2 * a] - c->npml[2 * a + 1] : c->n[a];
        if (L.pn[a] < 1) return fail(HELM3D_ERR_ARG, "PML covers the whole axis");
    }
    HIPCHK(launch_stencil<double>(VAR_ADJ | VAR_DERIV, L, c->grid, c->stream));
    return HELM3D_OK;
}

namespace {

int ext_grid(const int n_phys[3], const int npml[6], ExtGrid* G)
{
    if (!n_phys || !npml) return fail(HELM3D_ERR_ARG, "null argument");
    for (int a = 0; a < 3; ++a) {
        if (n_phys[a] < 1 || npml[2 * a] < 0 || npml[2 * a + 1] < 0)
            return fail(HELM3D_ERR_ARG, "bad grid sizes");
        G->np[a] = n_phys[a];
        G->lo[a] = npml[2 * a];
        const long nt = long(n_phys[a]) + npml[2 * a] + npml[2 * a + 1];
        if (nt > 1000000) return fail(HELM3D_ERR_ARG, "grid too large");
        G->nt[a] = int(nt);
    }
    return HELM3D_OK;
}

}  // namespace

int helm3d_param2wavenum(helm3d_ctx* c, const int n_phys[3], const int npml[6], const double* v, int unit,
                         double freq, double* wn, double* dwn, double* ddwn)
{
    if (!c || !v || !wn) return fail(HELM3D_ERR_ARG, "null argument");
    if (unit != HELM3D_UNIT_MS && unit != HELM3D_UNIT_S2M2 && unit != HELM3D_UNIT_S2KM2)
        return fail(HELM3D_ERR_ARG, "Unknown unit");
    ExtGrid G;
    CHK(ext_grid(n_phys, npml, &G));
    CHK(bind(c));
    const double w = 2.0 * M_PI * freq;
    double cc = w * w;
    if (unit == HELM3D_UNIT_S2KM2) cc *= 1e-6;
    HIPCHK(launch_param2wavenum(G, v, unit == HELM3D_UNIT_MS ? 0 : 1, cc, wn, dwn, ddwn, c->stream));
    return HELM3D_OK;
}

int helm3d_extension_adjoint(helm3d_ctx* c, const int n_phys[3], const int npml[6], const double* gpml,
                             double* g, int accumulate)
{
    if (!c || !gpml || !g) return fail(HELM3D_ERR_ARG, "null argument");
    ExtGrid G;
    CHK(ext_grid(n_phys, npml, &G));
    CHK(bind(c));
    HIPCHK(launch_extension_adjoint(G, gpml, g, accumulate != 0, c->stream));
    return HELM3D_OK;
}

int helm3d_csr_apply(helm3d_ctx* c, int nrows, const long long* rowptr, const long long* col, const double* w,
                     const long long* oidx, double alpha, const void* in, void* out, int accumulate)
{
    if (!c || nrows < 0) return fail(HELM3D_ERR_ARG, "bad argument");
    if (nrows == 0) return HELM3D_OK;
    if (!rowptr || !col || !w || !in || !out) return fail(HELM3D_ERR_ARG, "null argument");
    CHK(bind(c));
    HIPCHK(launch_csr_apply(nrows, rowptr, col, w, oidx, alpha, static_cast<const cpx<double>*>(in),
                            static_cast<cpx<double>*>(out), accumulate != 0, c->stream));
    return HELM3D_OK;
}

// ---------------------------------------------------------------------------
// Drop-in matvec at the MEX boundary
// ---------------------------------------------------------------------------

int helm3d_mvp(int variant, const double* wnr, const double* wni, const double* h, const double* n,
               long long n_len, const double* npml, double* yr, double* yi, const double* xr, const double* xi,
               long long x_rows, int nthreads)
{
    (void)nthreads;
    if (n_len != 3 || !n) return fail(HELM3D_ERR_NSIZE, "n must be a 3-length vector");
    int ni[3];
    for (int a = 0; a < 3; ++a) {
        if (!(n[a] >= 1) || n[a] != std::floor(n[a]) || n[a] > 2147483647.0)
            return fail(HELM3D_ERR_NSIZE, "n must be a 3-length vector");
        ni[a] = int(n[a]);
    }
    const long long N = (long long)ni[0] * ni[1] * ni[2];
    if (x_rows != N) return fail(HELM3D_ERR_XSIZE, "x must have nx*ny*nz elements");
    if (variant < 0 || variant > 7) return fail(HELM3D_ERR_ARG, "bad variant");
    const bool wnc = variant & HELM3D_WNC;
    if (!wnr || !h || !npml || !yr || !yi || !xr || (wnc && !wni)) return fail(HELM3D_ERR_ARG, "null argument");
    int np[6];
    for (int a = 0; a < 6; ++a) {
        if (!(npml[a] >= 0)) return fail(HELM3D_ERR_ARG, "npml must be non-negative");
        np[a] = int(npml[a]);
    }

    std::lock_guard<std::mutex> lock(g_dropin_mu);
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(HELM3D_ERR_ARG, "device index out of range");
    if (!g_dropin[dev]) CHK(helm3d_create(dev, nullptr, &g_dropin[dev]));
    helm3d_ctx* c = g_dropin[dev];
    // staging: wn (interleaved when complex), x, y on device
    Vecs<double>& st = c->T64;
    CHK(ensure_vecs(st, N, 3));
    const long ld = slot_ld(N);
    double* d_planar = reinterpret_cast<double*>(st.p);           // 2N doubles of slot 0
    cpx<double>* d_x = st.p + ld;                                 // slot 1
    cpx<double>* d_y = st.p + 2 * ld;                             // slot 2
    // wn -> slot 0 (real: first N doubles; complex: interleaved)
    if (wnc) {
        HIPCHK(hipMemcpyAsync(d_planar, wnr, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(d_planar + N, wni, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(launch_interleave(d_planar, d_planar + N, d_y, N, c->stream));  // y slot as scratch
        CHK(helm3d_set_model(c, ni, h, np, d_y, 1));
    } else {
        HIPCHK(hipMemcpyAsync(d_planar, wnr, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
        CHK(helm3d_set_model(c, ni, h, np, d_planar, 0));
    }
    // x planar -> interleaved (imaginary zeros for a real x, Helm3dmvp_mex.c:87-93)
    HIPCHK(hipMemcpyAsync(d_planar, xr, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (xi) HIPCHK(hipMemcpyAsync(d_planar + N, xi, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(launch_interleave(d_planar, xi ? d_planar + N : nullptr, d_x, N, c->stream));
    const int v = (variant & (HELM3D_ADJ | HELM3D_DERIV)) | (wnc ? VAR_WNC : 0);
    CHK(stencil<double>(c, v, c->op64, d_x, d_y, 1.0, nullptr, 0.0, nullptr));
    HIPCHK(launch_deinterleave(d_y, d_planar, d_planar + N, N, c->stream));
    HIPCHK(hipMemcpyAsync(yr, d_planar, N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(yi, d_planar + N, N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return HELM3D_OK;
}

#define HELM3D_DO_HMVP(name, var)                                                                          \
    int name(const double* wnr, const double* wni, const double* h, const double* n, const double* npml,   \
             double* yr, double* yi, const double* xr, const double* xi, int nthreads)                     \
    {                                                                                                      \
        if (!n) return fail(HELM3D_ERR_NSIZE, "n must be a 3-length vector");                             \
        const long long N = (long long)n[0] * (long long)n[1] * (long long)n[2];                           \
        return helm3d_mvp(var, wnr, wni, h, n, 3, npml, yr, yi, xr, xi, N, nthreads);                      \
    }
HELM3D_DO_HMVP(helm3d_do_Hmvp_forw, HELM3D_FORW)
HELM3D_DO_HMVP(helm3d_do_Hmvp_adj, HELM3D_ADJ)
HELM3D_DO_HMVP(helm3d_do_Hmvp_forw_deriv, HELM3D_FORW | HELM3D_DERIV)
HELM3D_DO_HMVP(helm3d_do_Hmvp_adj_deriv, HELM3D_ADJ | HELM3D_DERIV)
HELM3D_DO_HMVP(helm3d_do_Hmvp_forw_wnc, HELM3D_FORW | HELM3D_WNC)
HELM3D_DO_HMVP(helm3d_do_Hmvp_adj_wnc, HELM3D_ADJ | HELM3D_WNC)

}  // extern "C"
