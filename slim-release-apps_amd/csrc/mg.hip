// Grid transfers of the multigrid V-cycle (construct_multigrid.m:30-40):
// prolongation = trilinear interpolation between the PML-extended grids of two
// levels, both spanning linspace(0,1,n) per axis (fine2coarse.m:40-57 with
// LinInterp1D.m), restriction = 1/8 times its transpose.  Per axis the host
// supplies, for every fine index, the coarse bracket I and the weights of I
// and I+1, and for every coarse index the list of fine indices that use it
// (row-compressed), so both directions are gathers.  Sums are formed in FP64
// for complex128 and complex64 vectors alike.
#include "mg.hpp"

namespace helm3d {

namespace {

constexpr int KB = 256;

// Both transfers run one thread per output point on a 3-D grid of 64 x 4
// thread blocks (x along the lanes, one y row per wave, one z plane per grid
// layer): no index division, and the y and z table entries are uniform over
// a wave.

// xf += P xc at fine points: the 2 x 2 x 2 coarse neighbours of the bracket.
template <typename R>
__global__ __launch_bounds__(KB) void prolong_add_kernel(MgAxes T, const cpx<R>* __restrict__ xc,
                                                         cpx<R>* __restrict__ xf)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = __builtin_amdgcn_readfirstlane(int(blockIdx.y) * 4 + int(threadIdx.x >> 6));
    const int k = blockIdx.z;
    if (j >= T.nf[1]) return;
    const int ic = min(i, T.nf[0] - 1);
    const long ncx = T.nc[0], ncxy = long(T.nc[0]) * T.nc[1];
    const int Ix = T.ax[0].I[ic], Iy = T.ax[1].I[j], Iz = T.ax[2].I[k];
    const double wx[2] = {T.ax[0].wa[ic], T.ax[0].wb[ic]};
    const double wy[2] = {T.ax[1].wa[j], T.ax[1].wb[j]};
    const double wz[2] = {T.ax[2].wa[k], T.ax[2].wb[k]};
    const int jx[2] = {Ix, min(Ix + 1, T.nc[0] - 1)};
    const long jy[2] = {Iy * ncx, min(Iy + 1, T.nc[1] - 1) * ncx};
    const long jz[2] = {Iz * ncxy, min(Iz + 1, T.nc[2] - 1) * ncxy};
    double re = 0.0, im = 0.0;
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const double wyz = wy[b] * wz[c];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const double w = wx[a] * wyz;
                const cpx<R> v = xc[jx[a] + jy[b] + jz[c]];
                re = fma(w, double(v.re), re);
                im = fma(w, double(v.im), im);
            }
        }
    if (i < T.nf[0]) {
        const long p = i + long(j) * T.nf[0] + long(k) * T.nf[0] * T.nf[1];
        cpx<R> o = xf[p];
        o.re = R(double(o.re) + re);
        o.im = R(double(o.im) + im);
        xf[p] = o;
    }
}

// yc = scale * P^T rf at coarse points: per (z, y) entry of the transposed
// lists (uniform over the wave), the x entries of the lane's coarse index
// from the padded list (WX entries, zero weights past the real ones), so the
// loads of a coarse point are independent and issued together.  WX = 0: the
// padded width is read at run time.
template <typename R, int WX>
__global__ __launch_bounds__(KB) void restrict_kernel(MgAxes T, double scale, const cpx<R>* __restrict__ rf,
                                                      cpx<R>* __restrict__ yc)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = __builtin_amdgcn_readfirstlane(int(blockIdx.y) * 4 + int(threadIdx.x >> 6));
    const int k = blockIdx.z;
    if (j >= T.nc[1]) return;
    const int ic = min(i, T.nc[0] - 1);
    const long nfx = T.nf[0], nfxy = long(T.nf[0]) * T.nf[1];
    const int pw = WX > 0 ? WX : T.ax[0].pw;
    constexpr int KX = WX > 0 ? WX : 8;
    int fx[KX];
    double wxv[KX];
#pragma unroll
    for (int e = 0; e < KX; ++e) {
        const bool real = WX > 0 || e < pw;
        fx[e] = real ? T.ax[0].pfi[long(ic) * pw + e] : 0;
        wxv[e] = real ? T.ax[0].pfw[long(ic) * pw + e] : 0.0;
    }
    double re = 0.0, im = 0.0;
    const int z0 = T.ax[2].rp[k], z1 = T.ax[2].rp[k + 1];
    const int y0 = T.ax[1].rp[j], y1 = T.ax[1].rp[j + 1];
    for (int ez = z0; ez < z1; ++ez) {
        const long fz = T.ax[2].fi[ez] * nfxy;
        const double wz = T.ax[2].fw[ez];
        for (int ey = y0; ey < y1; ++ey) {
            const long base = T.ax[1].fi[ey] * nfx + fz;
            const double wyz = wz * T.ax[1].fw[ey];
            double sr = 0.0, si = 0.0;
#pragma unroll
            for (int e = 0; e < KX; ++e) {
                if (WX > 0 || e < pw) {
                    const cpx<R> v = rf[base + fx[e]];
                    sr = fma(wxv[e], double(v.re), sr);
                    si = fma(wxv[e], double(v.im), si);
                }
            }
            re = fma(wyz, sr, re);
            im = fma(wyz, si, im);
        }
    }
    if (i < T.nc[0]) yc[i + long(j) * T.nc[0] + long(k) * T.nc[0] * T.nc[1]] = cpx<R>{R(scale * re), R(scale * im)};
}

dim3 transfer_grid(const int n[3]) { return dim3((n[0] + 63) / 64, (n[1] + 3) / 4, n[2]); }

}  // namespace

template <typename R>
hipError_t launch_prolong_add(const MgAxes& T, const cpx<R>* xc, cpx<R>* xf, hipStream_t s)
{
    hipLaunchKernelGGL(prolong_add_kernel<R>, transfer_grid(T.nf), dim3(KB), 0, s, T, xc, xf);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_restrict(const MgAxes& T, double scale, const cpx<R>* rf, cpx<R>* yc, hipStream_t s)
{
    const dim3 g = transfer_grid(T.nc);
    switch (T.ax[0].pw) {
    case 2: hipLaunchKernelGGL((restrict_kernel<R, 2>), g, dim3(KB), 0, s, T, scale, rf, yc); break;
    case 3: hipLaunchKernelGGL((restrict_kernel<R, 3>), g, dim3(KB), 0, s, T, scale, rf, yc); break;
    case 4: hipLaunchKernelGGL((restrict_kernel<R, 4>), g, dim3(KB), 0, s, T, scale, rf, yc); break;
    case 5: hipLaunchKernelGGL((restrict_kernel<R, 5>), g, dim3(KB), 0, s, T, scale, rf, yc); break;
    default:
        if (T.ax[0].pw < 1 || T.ax[0].pw > 8) return hipErrorInvalidValue;
        hipLaunchKernelGGL((restrict_kernel<R, 0>), g, dim3(KB), 0, s, T, scale, rf, yc);
        break;
    }
    return hipGetLastError();
}

template hipError_t launch_prolong_add<double>(const MgAxes&, const cpx<double>*, cpx<double>*, hipStream_t);
template hipError_t launch_prolong_add<float>(const MgAxes&, const cpx<float>*, cpx<float>*, hipStream_t);
template hipError_t launch_restrict<double>(const MgAxes&, double, const cpx<double>*, cpx<double>*, hipStream_t);
template hipError_t launch_restrict<float>(const MgAxes&, double, const cpx<float>*, cpx<float>*, hipStream_t);

}  // namespace helm3d
