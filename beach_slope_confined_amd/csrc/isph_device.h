// isph_device.h -- device-side physics of the ISPH step (gfx950).
//
// Pairwise formulas of the reference with its evaluation order kept
// (Fortran left-to-right within a precedence level, x**(-1) as 1/x,
// x**2 as x*x, x**3 as (x*x)*x, x**4 as (x*x)*(x*x)) and the float32
// values of its default-real literals.  The library is compiled with
// -ffp-contract=off, so each product and sum rounds as written; the FP64
// div and sqrt of the AMDGPU backend are correctly rounded.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/isph_capi.h"

namespace isph {

constexpr double C_HYB1 = 87.92000579833984375;   // float32(28.0*3.14), kernel.f90:61
constexpr double C_HYB2 = 9.4200000762939453125;  // float32(3.0*3.14),  kernel.f90:62
constexpr double ALP = 0.1;                       // kernel.f90:30
constexpr double TINY_DET = 9.999999747378751635551453e-06;  // float32(1e-5), functions.f90:75
constexpr double TINY6 = 9.999999974752427078783512e-07;     // float32(1e-6), integrator.f90:497
constexpr double DM_SALT = 9.9999997171806853657471947e-10;  // float32(1e-9), scalar.f90:118

// Kernel constants derived once per launch from the parameters.
struct KernelConsts {
    double h, h2, h3, twoh;
    double n1, n2;     // hybrid normalisers 10/(c1 h^2), 1/(c2 h^2)
    double wden;       // Wendland gradient denominator (4*3.14) h^4
    double wnorm;      // Wendland 7/((4*3.14) h^2)
    int wendland;
};

__host__ __device__ inline KernelConsts make_kc(const isph_params& p) {
    KernelConsts k;
    k.h = p.h;
    k.h2 = p.h * p.h;
    k.h3 = k.h2 * p.h;
    k.twoh = 2.0 * p.h;
    k.n1 = 10.0 / (C_HYB1 * k.h2);
    k.n2 = 1.0 / (C_HYB2 * k.h2);
    k.wden = (4.0 * 3.140) * (k.h2 * k.h2);
    k.wnorm = 7.0 / ((4.0 * 3.140) * k.h2);
    k.wendland = p.kernel == ISPH_KERNEL_WENDLAND_C2;
    return k;
}

// Wab, kernel.f90:34-66 (hybrid) / :42-43 (Wendland C2).
__device__ inline double kern_w(const KernelConsts& k, double r) {
    if (k.wendland) {
        double w = 1.0 - r / (2.0 * k.h);
        double w2 = w * w;
        return (k.wnorm * (1.0 + (2.0 * r) / k.h)) * (w2 * w2);
    }
    double t1, t2;
    if (r < k.h) {
        double q = r / k.h;
        t1 = (4.0 - (6.0 * (r * r)) / k.h2) + (3.0 * ((r * r) * r)) / (k.h2 * k.h);
        t2 = (((q * q) * q) - (6.0 * r) / k.h) + 6.0;
    } else {
        double u = 2.0 - r / k.h;
        t1 = (u * u) * u;
        t2 = t1;
    }
    return (ALP * t1) * k.n1 + ((1.0 - ALP) * t2) * k.n2;
}

// Wo, kernel.f90:68-92.
__device__ inline double kern_w0(const KernelConsts& k) {
    if (k.wendland) return k.wnorm;
    return (ALP * 4.0) * k.n1 + ((1.0 - ALP) * 6.0) * k.n2;
}

// Wabx/Waby(p_j, p_i, r, h), kernel.f90:94-162: gradient with respect to
// x_i; dx = x_i - x_j, dy = y_i - y_j.
__device__ inline void kern_gradw(const KernelConsts& k, double r, double dx, double dy,
                                  double& wx, double& wy) {
    const double mdx = -dx, mdy = -dy;  // p1%x - p2%x = x_j - x_i
    if (k.wendland) {
        double w = 1.0 - r / (2.0 * k.h);
        double w3 = (w * w) * w;
        wx = ((35.0 * mdx) * w3) / k.wden;
        wy = ((35.0 * mdy) * w3) / k.wden;
        return;
    }
    double tx1, tx2, ty1, ty2;
    if (r < k.h) {
        double ax1 = (9.0 * r) / k.h3 - 12.0 / k.h2;
        double ax2 = (3.0 * r) / k.h3 - 6.0 / (r * k.h);
        tx1 = ax1 * dx; tx2 = ax2 * dx;
        ty1 = ax1 * dy; ty2 = ax2 * dy;
    } else {
        double u = 2.0 - r / k.h;
        double a = (3.0 / (r * k.h)) * (u * u);
        tx1 = a * mdx; tx2 = tx1;
        ty1 = a * mdy; ty2 = ty1;
    }
    wx = (ALP * tx1) * k.n1 + ((1.0 - ALP) * tx2) * k.n2;
    wy = (ALP * ty1) * k.n1 + ((1.0 - ALP) * ty2) * k.n2;
}

// dist (functions.f90:28-42) + gradient, from positions.
__device__ inline void pair_geom(const KernelConsts& k, double2 pi, double2 pj, double& r,
                                 double& wx, double& wy) {
    double dx = pi.x - pj.x, dy = pi.y - pj.y;
    r = sqrt(dx * dx + dy * dy);
    kern_gradw(k, r, dx, dy, wx, wy);
}

// invertmat2D, functions.f90:62-96.
__device__ inline void invert2(double& a0, double& a1, double& a2, double& a3) {
    double det = a0 * a3 - a1 * a2;
    if (fabs(det) >= TINY_DET && isfinite(det)) {
        double r0 = a3 / det, r1 = (-a1) / det, r2 = (-a2) / det, r3 = a0 / det;
        a0 = r0; a1 = r1; a2 = r2; a3 = r3;
    } else {
        a0 = 1.0; a1 = 0.0; a2 = 0.0; a3 = 1.0;
    }
}

// vis(oden), porous.f90:12-30 (visc_law 0) or mu = nu*oden (visc_law 1).
__device__ inline double visc(int visc_law, double nu, double oden) {
    if (visc_law == 1) return nu * oden;
    double sp = oden - 1000.0;
    return ((1.002 + 0.001652 * sp) + 0.0000083 * (sp * sp)) * 0.001;
}

}  // namespace isph
