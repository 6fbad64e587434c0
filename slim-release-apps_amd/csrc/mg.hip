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

template <typename R>
__global__ __launch_bounds__(KB) void prolong_add_kernel(MgAxes T, const cpx<R>* __restrict__ xc,
                                                         cpx<R>* __restrict__ xf)
{
    const long nfx = T.nf[0], nfy = T.nf[1], nfz = T.nf[2];
    const long N = nfx * nfy * nfz;
    const long ncx = T.nc[0], ncxy = long(T.nc[0]) * T.nc[1];
    for (long p = long(blockIdx.x) * KB + threadIdx.x; p < N; p += long(gridDim.x) * KB) {
        const int i = int(p % nfx), j = int((p / nfx) % nfy), k = int(p / (nfx * nfy));
        const int Ix = T.ax[0].I[i], Iy = T.ax[1].I[j], Iz = T.ax[2].I[k];
        const double wx[2] = {T.ax[0].wa[i], T.ax[0].wb[i]};
        const double wy[2] = {T.ax[1].wa[j], T.ax[1].wb[j]};
        const double wz[2] = {T.ax[2].wa[k], T.ax[2].wb[k]};
        const int jx[2] = {Ix, min(Ix + 1, T.nc[0] - 1)};
        const int jy[2] = {Iy, min(Iy + 1, T.nc[1] - 1)};
        const int jz[2] = {Iz, min(Iz + 1, T.nc[2] - 1)};
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int a = 0; a < 2; ++a) {
                    const double w = wx[a] * wy[b] * wz[c];
                    const cpx<R> v = xc[jx[a] + jy[b] * ncx + jz[c] * ncxy];
                    re += w * double(v.re);
                    im += w * double(v.im);
                }
        cpx<R> o = xf[p];
        o.re = R(double(o.re) + re);
        o.im = R(double(o.im) + im);
        xf[p] = o;
    }
}

template <typename R>
__global__ __launch_bounds__(KB) void restrict_kernel(MgAxes T, double scale, const cpx<R>* __restrict__ rf,
                                                      cpx<R>* __restrict__ yc)
{
    const long ncx = T.nc[0], ncy = T.nc[1], ncz = T.nc[2];
    const long N = ncx * ncy * ncz;
    const long nfx = T.nf[0], nfxy = long(T.nf[0]) * T.nf[1];
    for (long q = long(blockIdx.x) * KB + threadIdx.x; q < N; q += long(gridDim.x) * KB) {
        const int i = int(q % ncx), j = int((q / ncx) % ncy), k = int(q / (ncx * ncy));
        double re = 0.0, im = 0.0;
        for (int ez = T.ax[2].rp[k]; ez < T.ax[2].rp[k + 1]; ++ez) {
            const long fz = T.ax[2].fi[ez];
            const double wz = T.ax[2].fw[ez];
            for (int ey = T.ax[1].rp[j]; ey < T.ax[1].rp[j + 1]; ++ey) {
                const long fy = T.ax[1].fi[ey];
                const double wyz = wz * T.ax[1].fw[ey];
                for (int ex = T.ax[0].rp[i]; ex < T.ax[0].rp[i + 1]; ++ex) {
                    const double w = wyz * T.ax[0].fw[ex];
                    const cpx<R> v = rf[T.ax[0].fi[ex] + fy * nfx + fz * nfxy];
                    re += w * double(v.re);
                    im += w * double(v.im);
                }
            }
        }
        yc[q] = cpx<R>{R(scale * re), R(scale * im)};
    }
}

int blocks_for(long n) { return int(std::min<long>(std::max<long>((n + KB - 1) / KB, 1), 8192)); }

}  // namespace

template <typename R>
hipError_t launch_prolong_add(const MgAxes& T, const cpx<R>* xc, cpx<R>* xf, hipStream_t s)
{
    const long N = long(T.nf[0]) * T.nf[1] * T.nf[2];
    hipLaunchKernelGGL(prolong_add_kernel<R>, dim3(blocks_for(N)), dim3(KB), 0, s, T, xc, xf);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_restrict(const MgAxes& T, double scale, const cpx<R>* rf, cpx<R>* yc, hipStream_t s)
{
    const long N = long(T.nc[0]) * T.nc[1] * T.nc[2];
    hipLaunchKernelGGL(restrict_kernel<R>, dim3(blocks_for(N)), dim3(KB), 0, s, T, scale, rf, yc);
    return hipGetLastError();
}

template hipError_t launch_prolong_add<double>(const MgAxes&, const cpx<double>*, cpx<double>*, hipStream_t);
template hipError_t launch_prolong_add<float>(const MgAxes&, const cpx<float>*, cpx<float>*, hipStream_t);
template hipError_t launch_restrict<double>(const MgAxes&, double, const cpx<double>*, cpx<double>*, hipStream_t);
template hipError_t launch_restrict<float>(const MgAxes&, double, const cpx<float>*, cpx<float>*, hipStream_t);

}  // namespace helm3d
