// Host-side setup of the stencil: h-dependent constants and 1-D PML tables.
//
// Both are evaluated once per operator in FP64 on the host (glibc cos), so the
// device kernels read exact reference values instead of recomputing up to four
// cosines and two complex divisions per point as the CPU reference does.
#include <cmath>

#include "helm3d_common.hpp"

namespace helm3d {

// Helm3d_27pt.h:639-671.
HConst make_hconst(const double h[3])
{
    HConst k;
    const double hx2 = h[0] * h[0], hy2 = h[1] * h[1], hz2 = h[2] * h[2];
    const double hxy = hx2 + hy2, hxz = hx2 + hz2, hyz = hy2 + hz2;
    const double hxyz = hx2 + hy2 + hz2;
    const double w3a = (kW3) * 3 / (4 * hxyz);
    k.w3a2 = 2 * w3a;
    k.wn_coef = -(kW1 + 3 * kW2 + 16 * w3a * hxyz / 3 + kWM1 - 1);
    k.c[0] = -(kW1 / hx2 + kW2 / hx2 + kW2 / hxz + kW2 / hxy + 8 * w3a);
    k.c[1] = -(kW1 / hy2 + kW2 / hy2 + kW2 / hyz + kW2 / hxy + 8 * w3a);
    k.c[2] = -(kW1 / hz2 + kW2 / hz2 + kW2 / hxz + kW2 / hyz + 8 * w3a);
    k.xz = kW2 / (2 * hxz);
    k.xy = kW2 / (2 * hxy);
    k.yz = kW2 / (2 * hyz);
    return k;
}

namespace {

// Damping profile at 1-based position x, Helm3d_27pt.h:326-334.
double gamma_lower(double x, double n, double nlo)
{
    return std::cos(M_PI * ((x - 1) * n / (2 * (n + 1) * nlo)));
}
double gamma_upper(double x, double n, double nhi)
{
    return std::cos(M_PI * ((1 - (x - 1) / (n + 1)) * n / (2 * nhi)));
}

std::complex<double> xi_from(double ga, double gb)
{
    // 2 / [(2 - ga (ga + gb)) + i (3 ga + gb)]
    return std::complex<double>(2.0, 0.0) /
           std::complex<double>(2.0 - ga * (ga + gb), 3.0 * ga + gb);
}

// Helm3d_27pt.h:346-367.  The interior test compares with n+2-nlo at both ends
// (the reference uses the lower PML count there; kept for parity).
std::complex<double> xi_lower(double x, double n, double nlo, double nhi)
{
    double g0, g1;
    if (x <= nlo) {
        g1 = (x == nlo) ? 0 : gamma_lower(x + 1, n, nlo);
        g0 = gamma_lower(x, n, nlo);
    } else if (x <= n + 2 - nlo) {
        g0 = 0;
        g1 = (x == n + 2 - nlo) ? gamma_upper(x + 1, n, nhi) : 0;
    } else {
        g0 = gamma_upper(x, n, nhi);
        g1 = gamma_upper(x + 1, n, nhi);
    }
    return xi_from(g1, g0);
}

// Helm3d_27pt.h:369-390.
std::complex<double> xi_upper(double x, double n, double nlo, double nhi)
{
    double g1, g2;
    if (x <= nlo) {
        g1 = (x == nlo) ? 0 : gamma_lower(x + 1, n, nlo);
        g2 = (x >= nlo - 1) ? 0 : gamma_lower(x + 2, n, nlo);
    } else if (x <= n + 2 - nlo) {
        g2 = (x >= n + 1 - nlo) ? gamma_upper(x + 2, n, nhi) : 0;
        g1 = (x == n + 2 - nlo) ? gamma_upper(x + 1, n, nhi) : 0;
    } else {
        g2 = gamma_upper(x + 2, n, nhi);
        g1 = gamma_upper(x + 1, n, nhi);
    }
    return xi_from(g1, g2);
}

}  // namespace

void pml_slot(int n, int nlo, int nhi, int t, std::complex<double>* lo, std::complex<double>* hi)
{
    // slot t holds 0-based index t-1, whose 1-based position is t
    *lo = xi_lower(double(t), double(n), double(nlo), double(nhi));
    *hi = xi_upper(double(t), double(n), double(nlo), double(nhi));
}

PmlAxis make_pml_axis(int n, int nlo, int nhi)
{
    PmlAxis a;
    a.lo.resize(n + 2);
    a.hi.resize(n + 2);
    for (int t = 0; t < n + 2; ++t) pml_slot(n, nlo, nhi, t, &a.lo[t], &a.hi[t]);
    return a;
}

void pml_interior(int n, int nlo, int nhi, int* ilo, int* ihi)
{
    auto one = [&](int t) {
        std::complex<double> lo, hi;
        pml_slot(n, nlo, nhi, t, &lo, &hi);
        return lo == std::complex<double>(1.0, 0.0) && hi == std::complex<double>(1.0, 0.0);
    };
    // i qualifies when slots i, i+1, i+2 all hold exactly (1, 1); the
    // qualifying indices must form one contiguous run, else none qualify
    *ilo = 1;
    *ihi = 0;
    int first = -1, last = -1, run = 0;
    bool prev_ok = false, contiguous = true;
    for (int i = 0; i < n; ++i) {
        bool ok;
        if (i == 0) {
            run = int(one(0)) + int(one(1)) + int(one(2));
            ok = run == 3;
        } else {
            // slide the window i-1..i+1 -> i..i+2
            run += int(one(i + 2)) - int(one(i - 1));
            ok = run == 3;
        }
        if (ok) {
            if (first < 0) first = i;
            else if (!prev_ok) contiguous = false;
            last = i;
        }
        prev_ok = ok;
    }
    if (first >= 0 && contiguous) {
        *ilo = first;
        *ihi = last;
    }
}

}  // namespace helm3d
