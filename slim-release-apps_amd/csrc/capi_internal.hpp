// Internal host-side declarations shared by the C-ABI sources: error
// reporting, device allocation helpers, the Hessenberg least squares and the
// operator context.
#pragma once

#include <cmath>
#include <complex>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "bicgstab.hpp"
#include "fwi.hpp"
#include "helm27.hpp"
#include "helm3d.h"
#include "krylov.hpp"

namespace helm3d {

// message of the calling thread's last failure (defined in helm3d_capi.cpp)
extern thread_local std::string g_err;

inline int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                                \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return fail(e_ == hipErrorOutOfMemory ? HELM3D_ERR_NOMEM : HELM3D_ERR_HIP,              \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                         \
    } while (0)

#define CHK(expr)                      \
    do {                               \
        int s_ = (expr);               \
        if (s_ != HELM3D_OK) return s_; \
    } while (0)

// Grow-only device allocation.
inline int ensure(void** p, size_t* cap, size_t bytes)
{
    if (*cap >= bytes && *p) return HELM3D_OK;
    if (*p) {
        HIPCHK(hipFree(*p));
        *p = nullptr;
        *cap = 0;
    }
    if (bytes == 0) return HELM3D_OK;
    HIPCHK(hipMalloc(p, bytes));
    *cap = bytes;
    return HELM3D_OK;
}

template <typename T>
inline int ensure_t(T** p, size_t* cap, size_t count)
{
    void* q = *p;
    int s = ensure(&q, cap, count * sizeof(T));
    *p = static_cast<T*>(q);
    return s;
}

// Least squares min ||beta e1 - H y|| for the leading (k+1) x k block of the
// column-major (ldh x *) Hessenberg matrix H, by Givens rotations; returns the
// explicit residual ||beta e1 - H y|| as FGMRES.m:109-111 forms it.
inline double hess_lsq(int k, int ldh, const std::vector<std::complex<double>>& H, double beta,
                std::vector<std::complex<double>>& y)
{
    using C = std::complex<double>;
    std::vector<C> R((k + 1) * k), g(k + 1, C(0));
    for (int c = 0; c < k; ++c)
        for (int r = 0; r <= k; ++r) R[r + c * (k + 1)] = (r <= c + 1) ? H[r + c * ldh] : C(0);
    g[0] = beta;
    for (int c = 0; c < k; ++c) {
        const C a = R[c + c * (k + 1)], b = R[c + 1 + c * (k + 1)];
        const double den = std::sqrt(std::norm(a) + std::norm(b));
        C cs(1), sn(0);
        if (den != 0) {
            cs = a / den;
            sn = b / den;
        }
        for (int cc = c; cc < k; ++cc) {
            const C u = R[c + cc * (k + 1)], v = R[c + 1 + cc * (k + 1)];
            R[c + cc * (k + 1)] = std::conj(cs) * u + std::conj(sn) * v;
            R[c + 1 + cc * (k + 1)] = -sn * u + cs * v;
        }
        const C u = g[c], v = g[c + 1];
        g[c] = std::conj(cs) * u + std::conj(sn) * v;
        g[c + 1] = -sn * u + cs * v;
    }
    y.assign(k, C(0));
    for (int r = k - 1; r >= 0; --r) {
        C s = g[r];
        for (int c = r + 1; c < k; ++c) s -= R[r + c * (k + 1)] * y[c];
        y[r] = s / R[r + r * (k + 1)];
    }
    double rs = 0;
    for (int r = 0; r <= k; ++r) {
        C s = (r == 0) ? C(beta) : C(0);
        for (int c = 0; c < k; ++c)
            if (r <= c + 1) s -= H[r + c * ldh] * y[c];
        rs += std::norm(s);
    }
    return std::sqrt(rs);
}

template <typename R>
struct Vecs {
    cpx<R>* p = nullptr;
    size_t cap = 0;
};

}  // namespace helm3d

using namespace helm3d;

struct helm3d_ctx {
    int device = 0;
    hipStream_t stream = nullptr;

    bool has_model = false;
    int n[3] = {0, 0, 0};
    int npml[6] = {0, 0, 0, 0, 0, 0};
    double h[3] = {0, 0, 0};
    bool wnc = false;
    long N = 0;
    HConst hc{};

    cpx<double>* tab64 = nullptr;
    size_t tab64_cap = 0;
    cpx<float>* tab32 = nullptr;
    size_t tab32_cap = 0;
    void* wn64 = nullptr;
    size_t wn64_cap = 0;
    void* wn32 = nullptr;
    size_t wn32_cap = 0;
    OpDesc<double> op64{};
    OpDesc<float> op32{};
    StencilGrid grid{};

    double* part = nullptr;
    size_t part_cap = 0;
    void* sink = nullptr;  // stencil sink (kSinkSlots complex doubles)
    size_t sink_cap = 0;
    double* small = nullptr;  // device: 2*kMaxBasis dots + norm
    size_t small_cap = 0;
    double* small_host = nullptr;  // pinned mirror
    // Arnoldi steps enqueued ahead of the host: per step k a row of
    // kGsRow doubles (dots d_0..d_{k-1}, then ||w||^2), the squared norms of
    // the basis slots (scales of the update and of the next matvec), and
    // pinned rows the reductions write into (one row and event per step of a
    // cycle)
    double* gs = nullptr;  // device: (kMaxBasis + 1) rows
    size_t gs_cap = 0;
    double* nrm = nullptr;  // device: kMaxBasis + 1 values
    size_t nrm_cap = 0;
    GmresDevState* gst = nullptr;  // device state of the result-free solves
    size_t gst_cap = 0;
    double* ring_host = nullptr;  // pinned: one row per step of a cycle
    hipEvent_t ring_ev[kMaxBasis + 1] = {};
    // device addresses of the pinned mirrors: the reduction kernels write
    // reduced values straight into them (no copy command per step)
    double* small_host_dev = nullptr;
    double* ring_dev = nullptr;

    Vecs<double> V64, Z64, T64, BC64;
    Vecs<float> V32, Z32, T32a, T32b, BC32;
    double* bsc = nullptr;  // BiCGStab device scalars (BS_SLOTS doubles)
    size_t bsc_cap = 0;

    // captured Arnoldi cycle of the unpreconditioned GMRES on small grids,
    // one per precision (0: complex128, 1: complex64), with the bytes of the
    // launch arguments it was captured with (a different key recaptures)
    hipGraphExec_t cyc_graph[2] = {nullptr, nullptr};
    std::vector<unsigned char> cyc_key[2];

    // captured fixed-length complex64 inner solve (mixed precision)
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t inner_graph = nullptr;
    int inner_graph_variant = -1;
    long inner_graph_iters = -1;
    bool use_graphs = true;

    // optional per-kernel-class timing with HIP events on `stream`
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<int, size_t>> ev_pending;  // (kind, index of start event)
    double prof_ms[HELM3D_PROF_KINDS] = {0};
    long long prof_n[HELM3D_PROF_KINDS] = {0};

    ~helm3d_ctx()
    {
        (void)hipSetDevice(device);
        void* bufs[] = {tab64, tab32, wn64, wn32, part, small, sink, V64.p, Z64.p, T64.p, BC64.p,
                        V32.p, Z32.p, T32a.p, T32b.p, BC32.p, bsc, gs, nrm, gst};
        for (void* b : bufs)
            if (b) (void)hipFree(b);
        if (small_host) (void)hipHostFree(small_host);
        if (ring_host) (void)hipHostFree(ring_host);
        for (hipEvent_t e : ring_ev)
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
        if (inner_graph) (void)hipGraphExecDestroy(inner_graph);
        for (hipGraphExec_t g : cyc_graph)
            if (g) (void)hipGraphExecDestroy(g);
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
    }
};

namespace helm3d {

// Operator descriptor of H(wn) for the kernels: grid, PML tables (tab + off[q]
// for xlo, xhi, ylo, yhi, zlo, zhi), wavenumber (real or interleaved complex,
// device), the h-dependent constants and the interior index ranges.
template <typename R>
void fill_opdesc(OpDesc<R>& op, const int n[3], const HConst& k, const cpx<R>* tab, const size_t off[6],
                 const void* wn, bool wnc, const int ilo[3], const int ihi[3])
{
    op.nx = n[0];
    op.ny = n[1];
    op.nz = n[2];
    op.xlo = tab + off[0];
    op.xhi = tab + off[1];
    op.ylo = tab + off[2];
    op.yhi = tab + off[3];
    op.zlo = tab + off[4];
    op.zhi = tab + off[5];
    if (wnc) {
        op.wn = nullptr;
        op.wnc = static_cast<const cpx<R>*>(wn);
    } else {
        op.wn = static_cast<const R*>(wn);
        op.wnc = nullptr;
    }
    op.wn_const = R(0);
    for (int a = 0; a < 3; ++a) op.c[a] = R(k.c[a]);
    op.xy = R(k.xy);
    op.xz = R(k.xz);
    op.yz = R(k.yz);
    op.w3a2 = R(k.w3a2);
    op.wn_coef = R(k.wn_coef);
    op.wm2 = R(kWM2);
    op.wm3 = R(kWM3);
    op.wm4 = R(kWM4);
    // interior coefficients: xi_x = xi_y = 2, side values 1
    op.fi[0] = R(-2 * k.c[0] - 2 * k.c[1]);
    op.fi[1] = R(k.c[0] + 2 * k.xy);
    op.fi[2] = R(k.c[1] + 2 * k.xy);
    op.fi[3] = R(-2 * k.xy);
    op.gi[0] = R(2 * k.xz + 2 * k.yz);
    op.gi[1] = R(-k.xz + 2 * k.w3a2);
    op.gi[2] = R(-k.yz + 2 * k.w3a2);
    op.gi[3] = R(-2 * k.w3a2);
    for (int a = 0; a < 3; ++a) {
        op.int_lo[a] = ilo[a];
        op.int_hi[a] = ihi[a];
    }
}

inline int bind(helm3d_ctx* c)
{
    if (!c) return fail(HELM3D_ERR_ARG, "null context");
    HIPCHK(hipSetDevice(c->device));
    return HELM3D_OK;
}

// Leading dimension of a vector slot: N rounded up to 256 complex values so
// every slot starts on a 4 KiB boundary.
inline long slot_ld(long N) { return (N + 255) / 256 * 256; }

// Flexible preconditioner hook of the GMRES engine: z = M^{-1} (s * v_raw),
// in the engine's precision.
template <typename R>
struct PrecondT {
    virtual ~PrecondT() = default;
    virtual int apply(const cpx<R>* v_raw, double s, cpx<R>* z) = 0;
};
using Precond = PrecondT<double>;

struct GmresResult {
    long iters = 0;
    long cycles = 0;
    double relres = 1.0;
    double quasi = 1.0;
    bool converged = false;
};

// Restarted GMRES of FGMRES.m:33-128 on the context's operator (variant) and
// workspace (helm3d_capi.cpp).  P makes it flexible.  need_result = false is
// for callers that use only x (multigrid smoothers and coarse solves): the
// true residual after the last cycle (which decides nothing) is not formed and
// res->relres holds the quasi-residual; x is the same.
template <typename R>
int gmres_engine(helm3d_ctx* c, int variant, const cpx<R>* b, cpx<R>* x, bool use_x0, double tol, int m,
                 long maxit, long iter_cap, PrecondT<R>* P, GmresResult* res, std::vector<double>* hst,
                 bool need_result = true);

int model_variant(const helm3d_ctx* c, int mode_adj);

constexpr int kGsRow = 2 * kMaxBasis + 2;  // doubles per Arnoldi-step row
// grids of at most this many points enqueue a whole Arnoldi cycle ahead of
// the host (gmres_engine without a preconditioner)
constexpr long kDeepLookahead = 2L << 20;
// result-free calls without a preconditioner and with at most this many
// restart cycles run all their decisions on the device (gmres_engine)
constexpr long kMaxDeviceCycles = 16;

// Context work buffers: reduction partials, reduced values (+ pinned host
// mirror) and the stencil's write-only sink.
inline int ensure_work(helm3d_ctx* c)
{
    const size_t nblk = std::max<size_t>(c->grid.blocks, kMaxPartials);
    CHK(ensure_t(&c->part, &c->part_cap, nblk * 2 * kMaxBasis));
    CHK(ensure_t(&c->small, &c->small_cap, 2 * kMaxBasis + 8));
    CHK(ensure(&c->sink, &c->sink_cap, size_t(kSinkSlots) * sizeof(cpx<double>)));
    if (!c->small_host) {
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->small_host), (2 * kMaxBasis + 8) * sizeof(double)));
        HIPCHK(hipHostGetDevicePointer(reinterpret_cast<void**>(&c->small_host_dev), c->small_host, 0));
    }
    CHK(ensure_t(&c->gs, &c->gs_cap, size_t(kMaxBasis + 1) * kGsRow));
    CHK(ensure_t(&c->nrm, &c->nrm_cap, size_t(kMaxBasis + 1)));
    CHK(ensure_t(&c->gst, &c->gst_cap, 1));
    if (!c->ring_host) {
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->ring_host), (kMaxBasis + 1) * kGsRow * sizeof(double)));
        HIPCHK(hipHostGetDevicePointer(reinterpret_cast<void**>(&c->ring_dev), c->ring_host, 0));
    }

    for (hipEvent_t& e : c->ring_ev)
        if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return HELM3D_OK;
}

template <typename R>
inline int ensure_vecs(Vecs<R>& v, long N, int slots)
{
    return ensure_t(&v.p, &v.cap, size_t(slot_ld(N)) * slots);
}

}  // namespace helm3d
