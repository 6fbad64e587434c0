// Shared definitions of the MI355X Helmholtz library: complex storage type,
// stencil constants, PML tables and the operator descriptor handed to kernels.
//
// Stencil semantics follow SLIM-release-apps
// tools/algorithms/3DFreqModeling/Helm3d_27pt.h (27-point mixed-grid stencil of
// Operto et al. 2007 with mass lumping and PML); each item cites the lines it
// follows.
#pragma once

#include <hip/hip_runtime.h>
#include <complex>
#include <cstdint>
#include <vector>

namespace helm3d {

// Interleaved complex value (re, im) with natural 2*sizeof(R) alignment so a
// double complex moves as one 16-byte access.
template <typename R>
struct alignas(2 * sizeof(R)) cpx {
    R re, im;
};

template <typename R>
__host__ __device__ inline cpx<R> mk(R a, R b) { return cpx<R>{a, b}; }
template <typename R>
__host__ __device__ inline cpx<R> operator+(cpx<R> a, cpx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R>
__host__ __device__ inline cpx<R> operator-(cpx<R> a, cpx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <typename R>
__host__ __device__ inline cpx<R> operator-(cpx<R> a) { return {-a.re, -a.im}; }
template <typename R>
__host__ __device__ inline cpx<R> operator*(R s, cpx<R> a) { return {s * a.re, s * a.im}; }
// The fused forms are written out so that every kernel instantiation rounds
// the same way (a * b - c * d left to the compiler may be contracted with
// either product, differently in different instantiations).
template <typename R>
__host__ __device__ inline cpx<R> cmul(cpx<R> a, cpx<R> b)
{
    return {R(fma(a.re, b.re, -(a.im * b.im))), R(fma(a.re, b.im, a.im * b.re))};
}
// acc + a*b
template <typename R>
__host__ __device__ inline cpx<R> cfma(cpx<R> acc, cpx<R> a, cpx<R> b)
{
    acc.re = R(fma(a.re, b.re, acc.re));
    acc.re = R(fma(-a.im, b.im, acc.re));
    acc.im = R(fma(a.re, b.im, acc.im));
    acc.im = R(fma(a.im, b.re, acc.im));
    return acc;
}
// acc + s*a, s real
template <typename R>
__host__ __device__ inline cpx<R> rfma(cpx<R> acc, R s, cpx<R> a)
{
    return {R(fma(s, a.re, acc.re)), R(fma(s, a.im, acc.im))};
}
template <typename R>
__host__ __device__ inline cpx<R> conjc(cpx<R> a) { return {a.re, -a.im}; }

// Stencil weights: the C literals of Helm3d_27pt.h:194-200 (not the exact
// fractions of the MATLAB files; the MEX C path is the parity target).
constexpr double kW1 = 1.8395262e-5;
constexpr double kW2 = 0.296692333333333;
constexpr double kW3 = 0.027476150000000;
constexpr double kWM1 = 0.49649658;
constexpr double kWM2 = 0.075168750000000;
constexpr double kWM3 = 0.004373916666667;
constexpr double kWM4 = 5.690375e-07;

// h-dependent scalars, Helm3d_27pt.h:639-671 (same evaluation order).
struct HConst {
    double wn_coef;      // centre mass weight
    double c[3];         // face weights pml{x,y,z}_coef (negative)
    double xy, xz, yz;   // cross weights W2 / (2 (h_a^2 + h_b^2))
    double w3a2;         // 2 * 3 W3 / (4 |h|^2)
};
HConst make_hconst(const double h[3]);

// One axis of the PML: xi_lo and xi_hi for 0-based indices -1 .. n, stored at
// slot index+1 (n+2 values).  Evaluated on the host in FP64 with the profile
// of Helm3d_27pt.h:326-390 at the 1-based position index+1
// (Helm3d_27pt.h:406-408, 425-427, 467-469); the adjoint reads the
// neighbour's slot (Helm3d_27pt.h:397-402, 437-463).
struct PmlAxis {
    std::vector<std::complex<double>> lo, hi;
};
PmlAxis make_pml_axis(int n, int nlo, int nhi);
// Slot t of both tables of one axis (no allocation).
void pml_slot(int n, int nlo, int nhi, int t, std::complex<double>* lo, std::complex<double>* hi);
// Interior index range of one axis: indices i whose slots i, i+1, i+2 all
// hold xi_lo = xi_hi = 1 exactly, as one contiguous run [ilo, ihi] (ilo > ihi
// when there is none).  No allocation.
void pml_interior(int n, int nlo, int nhi, int* ilo, int* ihi);

// Device-side description of one operator H(wn) on an nx*ny*nz grid
// (x fastest, idx = i + j*nx + k*nx*ny as Helm3d_27pt.h:21-23).
template <typename R>
struct OpDesc {
    int nx, ny, nz;
    // PML tables (device, n+2 complex entries per axis, slot = index+1)
    const cpx<R>* xlo;
    const cpx<R>* xhi;
    const cpx<R>* ylo;
    const cpx<R>* yhi;
    const cpx<R>* zlo;
    const cpx<R>* zhi;
    // wavenumber: real (wn) or complex (wnc); when both are null the real
    // constant wn_const is used (a constant dwn, e.g. (2 pi f)^2 for s^2/m^2)
    const R* wn;
    const cpx<R>* wnc;
    R wn_const;
    // constants
    R c[3], xy, xz, yz, w3a2, wn_coef, wm2, wm3, wm4;
    // interior (no x/y PML) column coefficients on (centre, x sum, y sum,
    // diagonal sum): F for the own plane, G for planes k+-1
    R fi[4], gi[4];
    // per axis, indices i whose PML slots i, i+1, i+2 (indices i-1..i+1) all
    // hold xi_lo = xi_hi = 1 exactly; empty when lo > hi
    int int_lo[3], int_hi[3];
};

// Status codes returned by every C-ABI entry point.
enum Status : int {
    OK = 0,
    ERR_NSIZE = 1,      // n is not a 3-vector of positive integers
    ERR_XSIZE = 2,      // x does not have prod(n) rows
    ERR_ARG = 3,        // invalid argument (null pointer, bad variant, bad npml)
    ERR_HIP = 4,        // HIP runtime failure
    ERR_NOMODEL = 5,    // context has no model
    ERR_NOMEM = 6,      // device allocation failed
    ERR_NOTCONV = 7,    // solver stopped at maxit above tolerance (x still returned)
};

}  // namespace helm3d
