// y = H x / H^H x in complex128 without any device allocation: the variant of
// helm3d_apply for callers that cannot call hipMalloc (code under HIP graph
// capture, tight serving loops).  The operator is given by its grid, spacing,
// PML counts and the caller's device wn; the kernels' scratch -- the six 1-D
// PML tables and the stencil's write-only sink -- lives in the file-scope
// device array `scratch` below, one copy per GPU, held until the process ends.
//
// Every call on one GPU shares that GPU's array: calls that overlap on one
// GPU (different streams, or host threads without ordering) race on the PML
// tables.  Each call rewrites the tables it reads before the stencil runs, on
// the same stream; the sink holds no values anyone reads, so nothing needs
// clearing.
#include <mutex>

#include "capi_internal.hpp"

namespace {

constexpr int kMaxAxis = HELM3D_NOALLOC_MAX_AXIS;           // points per axis, PML included
constexpr size_t kAxisSlots = size_t(kMaxAxis) + 2;          // slots of one table
constexpr size_t kTabBytes = 6 * kAxisSlots * sizeof(cpx<double>);
constexpr size_t kSinkBytes = size_t(kSinkSlots) * sizeof(cpx<double>);
constexpr size_t kScratchBytes = kTabBytes + kSinkBytes;

}  // namespace

// The kernels' scratch on every GPU the library runs on (1.4 MiB).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {nullptr};  // device address of `scratch` per device

int scratch_of(int dev, void** out)
{
    if (dev < 0 || dev >= 64) return fail(HELM3D_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(HELM3D_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return HELM3D_OK;
}

// The PML profile of Helm3d_27pt.h:326-390 evaluated on the device (device
// cos): the same branches as the host tables of helm3d_pml.cpp, including the
// lower-count interior test (Q1).
__device__ double gamma_lower_d(double x, double n, double nlo)
{
    return cos(M_PI * ((x - 1) * n / (2 * (n + 1) * nlo)));
}
__device__ double gamma_upper_d(double x, double n, double nhi)
{
    return cos(M_PI * ((1 - (x - 1) / (n + 1)) * n / (2 * nhi)));
}
// 2 / [(2 - ga (ga + gb)) + i (3 ga + gb)]
__device__ cpx<double> xi_from_d(double ga, double gb)
{
    const double re = 2.0 - ga * (ga + gb), im = 3.0 * ga + gb;
    const double d = re * re + im * im;
    return cpx<double>{2.0 * re / d, -2.0 * im / d};
}

// One thread per (axis, side, slot): slot t of axis a, table lo (side 0) or
// hi (side 1), at tab[(2 a + side) * kAxisSlots + t].
__global__ void pml_tables_kernel(cpx<double>* tab, int nx, int ny, int nz, int l0, int h0, int l1, int h1, int l2,
                                  int h2)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;  // 2 * axis + side
    const int a = q >> 1;
    const int nn = a == 0 ? nx : (a == 1 ? ny : nz);
    if (t >= nn + 2) return;
    const double n = nn;
    const double nlo = a == 0 ? l0 : (a == 1 ? l1 : l2);
    const double nhi = a == 0 ? h0 : (a == 1 ? h1 : h2);
    const double x = t;  // 1-based position of 0-based index t - 1
    double ga, gb;
    if ((q & 1) == 0) {
        // xi_lo: Helm3d_27pt.h:346-367
        if (x <= nlo) {
            ga = (x == nlo) ? 0 : gamma_lower_d(x + 1, n, nlo);
            gb = gamma_lower_d(x, n, nlo);
        } else if (x <= n + 2 - nlo) {
            gb = 0;
            ga = (x == n + 2 - nlo) ? gamma_upper_d(x + 1, n, nhi) : 0;
        } else {
            gb = gamma_upper_d(x, n, nhi);
            ga = gamma_upper_d(x + 1, n, nhi);
        }
    } else {
        // xi_hi: Helm3d_27pt.h:369-390
        if (x <= nlo) {
            ga = (x == nlo) ? 0 : gamma_lower_d(x + 1, n, nlo);
            gb = (x >= nlo - 1) ? 0 : gamma_lower_d(x + 2, n, nlo);
        } else if (x <= n + 2 - nlo) {
            gb = (x >= n + 1 - nlo) ? gamma_upper_d(x + 2, n, nhi) : 0;
            ga = (x == n + 2 - nlo) ? gamma_upper_d(x + 1, n, nhi) : 0;
        } else {
            gb = gamma_upper_d(x + 2, n, nhi);
            ga = gamma_upper_d(x + 1, n, nhi);
        }
    }
    tab[size_t(q) * kAxisSlots + t] = xi_from_d(ga, gb);
}

}  // namespace

extern "C" {

int helm3d_apply_noalloc(int variant, const int n[3], const double h[3], const int npml[6], const void* wn_dev,
                         int wn_complex, const void* x_dev, void* y_dev, void* stream)
{
    if (!n || !h || !npml || !wn_dev || !x_dev || !y_dev) return fail(HELM3D_ERR_ARG, "null argument");
    if (variant != HELM3D_FORW && variant != HELM3D_ADJ) return fail(HELM3D_ERR_ARG, "variant must be FORW or ADJ");
    for (int a = 0; a < 3; ++a) {
        if (n[a] < 1) return fail(HELM3D_ERR_NSIZE, "n must hold 3 positive integers");
        if (n[a] > kMaxAxis)
            return fail(HELM3D_ERR_ARG, "helm3d_apply_noalloc: a grid axis exceeds HELM3D_NOALLOC_MAX_AXIS points");
    }
    for (int a = 0; a < 6; ++a)
        if (npml[a] < 0) return fail(HELM3D_ERR_ARG, "npml must be non-negative");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    void* base = nullptr;
    CHK(scratch_of(dev, &base));
    cpx<double>* tab = static_cast<cpx<double>*>(base);
    cpx<double>* sink = reinterpret_cast<cpx<double>*>(static_cast<unsigned char*>(base) + kTabBytes);
    const hipStream_t s = static_cast<hipStream_t>(stream);

    const int mx = std::max(n[0], std::max(n[1], n[2])) + 2;
    hipLaunchKernelGGL(pml_tables_kernel, dim3((mx + 255) / 256, 6), dim3(256), 0, s, tab, n[0], n[1], n[2], npml[0],
                       npml[1], npml[2], npml[3], npml[4], npml[5]);
    HIPCHK(hipGetLastError());

    int ilo[3], ihi[3];
    for (int a = 0; a < 3; ++a) pml_interior(n[a], npml[2 * a], npml[2 * a + 1], &ilo[a], &ihi[a]);
    const size_t off[6] = {0, kAxisSlots, 2 * kAxisSlots, 3 * kAxisSlots, 4 * kAxisSlots, 5 * kAxisSlots};
    StencilLaunch<double> L{};
    fill_opdesc(L.op, n, make_hconst(h), tab, off, wn_dev, wn_complex != 0, ilo, ihi);
    L.x = static_cast<const cpx<double>*>(x_dev);
    L.y = static_cast<cpx<double>*>(y_dev);
    L.alpha = 1.0;
    L.beta = 0.0;
    L.sink = sink;
    const StencilGrid g = stencil_grid<double>(n[0], n[1], n[2], 0);
    const int v = variant | (wn_complex ? VAR_WNC : 0);
    HIPCHK(launch_stencil<double>(v, L, g, s));
    return HELM3D_OK;
}

}  // extern "C"
