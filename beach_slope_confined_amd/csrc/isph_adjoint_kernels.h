// isph_adjoint_kernels.h -- kernels of the discrete adjoint of the ISPH step.
//
// Every stage of the forward step (isph_kernels.hip) is a gather over a
// particle's neighbour list: out_i = F(u_i, sum_j g(u_i, u_j)).  Its adjoint
// runs in two passes, both gathers over the same symmetric lists, so no two
// threads write one address and the sums have a fixed order:
//   self pass (thread = row i): reads the adjoint of out_i, forms the
//     adjoint of the row's accumulators (stored in `w` for the second pass)
//     and of u_i, evaluating g with u_i seeded as dual numbers;
//   neighbour pass (thread = particle j): for every i in the list of j,
//     evaluates g(u_i, u_j) with u_j seeded and contracts it with row i's
//     accumulator adjoint.
// The templated formulas below restate the forward kernels' arithmetic in
// the same order; with T = double they reproduce the forward values.
#pragma once
#include "isph_dual.h"
#include "isph_kernels.h"

namespace isph {
namespace adj {

// ---------------------------------------------------------------------------
// templated physics (kernel.f90, functions.f90; same order as isph_device.h)
// ---------------------------------------------------------------------------
template <class T>
__device__ inline T kw(const KernelConsts& k, const T& r) {
    if (k.wendland) {
        T w = 1.0 - r / (2.0 * k.h);
        T w2 = w * w;
        return (k.wnorm * (1.0 + (2.0 * r) / k.h)) * (w2 * w2);
    }
    T t1, t2;
    if (r < k.h) {
        T q = r / k.h;
        t1 = (4.0 - (6.0 * (r * r)) / k.h2) + (3.0 * ((r * r) * r)) / (k.h2 * k.h);
        t2 = (((q * q) * q) - (6.0 * r) / k.h) + 6.0;
    } else {
        T u = 2.0 - r / k.h;
        t1 = (u * u) * u;
        t2 = t1;
    }
    return (ALP * t1) * k.n1 + ((1.0 - ALP) * t2) * k.n2;
}

template <class T>
__device__ inline void kgradw(const KernelConsts& k, const T& r, const T& dx, const T& dy, T& wx,
                              T& wy) {
    const T mdx = -dx, mdy = -dy;
    if (k.wendland) {
        T w = 1.0 - r / (2.0 * k.h);
        T w3 = (w * w) * w;
        wx = ((35.0 * mdx) * w3) / k.wden;
        wy = ((35.0 * mdy) * w3) / k.wden;
        return;
    }
    T tx1, tx2, ty1, ty2;
    if (r < k.h) {
        T ax1 = (9.0 * r) / k.h3 - 12.0 / k.h2;
        T ax2 = (3.0 * r) / k.h3 - 6.0 / (r * k.h);
        tx1 = ax1 * dx; tx2 = ax2 * dx;
        ty1 = ax1 * dy; ty2 = ax2 * dy;
    } else {
        T u = 2.0 - r / k.h;
        T a = (3.0 / (r * k.h)) * (u * u);
        tx1 = a * mdx; tx2 = tx1;
        ty1 = a * mdy; ty2 = ty1;
    }
    wx = (ALP * tx1) * k.n1 + ((1.0 - ALP) * tx2) * k.n2;
    wy = (ALP * ty1) * k.n1 + ((1.0 - ALP) * ty2) * k.n2;
}

template <class T>
__device__ inline void pgeom(const KernelConsts& k, const T& pix, const T& piy, const T& pjx,
                             const T& pjy, T& r, T& wx, T& wy) {
    T dx = pix - pjx, dy = piy - pjy;
    r = sqrt(dx * dx + dy * dy);
    kgradw(k, r, dx, dy, wx, wy);
}

template <class T>
__device__ inline void inv2(T& a0, T& a1, T& a2, T& a3) {
    T det = a0 * a3 - a1 * a2;
    if (fabs(det) >= TINY_DET && isfinite(val(det))) {
        T r0 = a3 / det, r1 = (-a1) / det, r2 = (-a2) / det, r3 = a0 / det;
        a0 = r0; a1 = r1; a2 = r2; a3 = r3;
    } else {
        a0 = T(1.0); a1 = T(0.0); a2 = T(0.0); a3 = T(1.0);
    }
}

// adjoint arrays, one per differentiable field of StateView (and L); `th`
// holds per-slot contributions to dJ/d(por, alpha, beta) (3 x cap), `aum`
// per-slot contributions to dJ/d umax, `w` per-row stage weights (6 x cap)
struct AdjView {
    double2 *pos, *vel, *sp, *sv;
    double *por, *rho, *R, *nut, *tke, *varts, *p, *velm;
    double* L;
    double* th;
    double* aum;
    double* w;
    size_t cap;
};

__device__ __forceinline__ void add2(double2* a, int i, double x, double y) {
    double2 t = a[i];
    t.x = t.x + x;
    t.y = t.y + y;
    a[i] = t;
}

__device__ __forceinline__ void ld_L(const double* L, int i, double* l) {
    const double2* L2 = reinterpret_cast<const double2*>(L);
    double2 a = L2[2 * (size_t)i], b = L2[2 * (size_t)i + 1];
    l[0] = a.x; l[1] = a.y; l[2] = b.x; l[3] = b.y;
}

__device__ __forceinline__ void add_L(double* L, int i, const double* g) {
    double2* L2 = reinterpret_cast<double2*>(L);
    double2 a = L2[2 * (size_t)i], b = L2[2 * (size_t)i + 1];
    a.x = a.x + g[0]; a.y = a.y + g[1]; b.x = b.x + g[2]; b.y = b.y + g[3];
    L2[2 * (size_t)i] = a;
    L2[2 * (size_t)i + 1] = b;
}

__device__ __forceinline__ void set_L(double* L, int i, double a0, double a1, double a2, double a3) {
    double2* L2 = reinterpret_cast<double2*>(L);
    L2[2 * (size_t)i] = make_double2(a0, a1);
    L2[2 * (size_t)i + 1] = make_double2(a2, a3);
}

__device__ __forceinline__ bool is_active_fluid(int tid, int fl) {
    return tid == 3 && !(fl & ISPH_FLAG_BUFFER) && !(fl & ISPH_FLAG_DEAD);
}

// ===========================================================================
// projection (isph.f90:16-47): sp = x; x += dt v / n        (fluid, not dead)
// ===========================================================================
__global__ void k_adj_projection(int n, StateView s, const SimScalars* ss, AdjView a) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    if (s.tid[i] == 3 && !(s.flags[i] & ISPH_FLAG_DEAD)) {
        const double dt = ss->dt;
        const double2 v = s.vel[i];
        const double nn = s.por[i];
        const double2 ap = a.pos[i], as = a.sp[i];
        // x_out = x + (dt*v)/n
        const double gx = ap.x * (dt / nn), gy = ap.y * (dt / nn);
        add2(a.vel, i, gx, gy);
        a.por[i] = a.por[i] - (ap.x * ((dt * v.x) / (nn * nn)) + ap.y * ((dt * v.y) / (nn * nn)));
        a.pos[i] = make_double2(ap.x + as.x, ap.y + as.y);
    }
    a.sp[i] = make_double2(0.0, 0.0);
}

// ===========================================================================
// comp_pos (integrator.f90:867-906): x = ((x + xs) + dt v / n) / 2
// ===========================================================================
__global__ void k_adj_comp_pos(int n, StateView s, const SimScalars* ss, AdjView a) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    if (s.tid[i] == 3 && !(s.flags[i] & ISPH_FLAG_DEAD)) {
        const double dt = ss->dt;
        const double2 v = s.vel[i];
        const double nn = s.por[i];
        const double2 ap = a.pos[i];
        const double hx = ap.x / 2.0, hy = ap.y / 2.0;
        add2(a.sp, i, hx, hy);
        add2(a.vel, i, hx * (dt / nn), hy * (dt / nn));
        a.por[i] = a.por[i] - (hx * ((dt * v.x) / (nn * nn)) + hy * ((dt * v.y) / (nn * nn)));
        a.pos[i] = make_double2(hx, hy);
    }
}

// ===========================================================================
// neighbour / KGC (search.f90:77-85, 276-278): A_i = sum_j grad W (x) V_j dx,
// L_i = inv(A_i) for fluid rows, identity otherwise.
// row seeds: x_i, y_i; neighbour seeds: x_j, y_j, rho_j
// ===========================================================================
template <class T>
__device__ inline void kgc_pair(const KernelConsts& kc, const T& pix, const T& piy, const T& pjx,
                                const T& pjy, const T& rhoj, double mj, T* acc) {
    T dx = pix - pjx, dy = piy - pjy;
    T r = sqrt(dx * dx + dy * dy);
    T wx, wy;
    kgradw(kc, r, dx, dy, wx, wy);
    T res1 = (mj * (pjx - pix)) / rhoj;
    T res3 = (mj * (pjy - piy)) / rhoj;
    acc[0] = wx * res1;
    acc[1] = wy * res1;
    acc[2] = wx * res3;
    acc[3] = wy * res3;
}

__global__ void __launch_bounds__(BLOCK) k_adj_kgc_self(int n, KernelConsts kc, size_t ld,
                                                        StateView s, const int* __restrict__ nbr,
                                                        const int* __restrict__ cnt, AdjView a) {
    int i = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (i >= n) return;
    double lb[4];
    ld_L(a.L, i, lb);
    set_L(a.L, i, 0.0, 0.0, 0.0, 0.0);  // L is overwritten for every row
    double* W = a.w;
    const size_t cp = a.cap;
    if (s.tid[i] != 3) {
        for (int q = 0; q < 4; ++q) W[q * cp + i] = 0.0;
        return;
    }
    const double2 pi = s.pos[i];
    const int c = cnt[i];
    double A[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < c; ++k) {
        const int j = nbr[(size_t)k * ld + i];
        const double2 pj = s.pos[j];
        double t[4];
        kgc_pair<double>(kc, pi.x, pi.y, pj.x, pj.y, s.rho[j], s.mass[j], t);
        for (int q = 0; q < 4; ++q) A[q] = A[q] + t[q];
    }
    // adjoint of the 2x2 inverse
    Dual<4> d0 = seed<4>(A[0], 0), d1 = seed<4>(A[1], 1), d2 = seed<4>(A[2], 2),
            d3 = seed<4>(A[3], 3);
    inv2(d0, d1, d2, d3);
    double Ab[4];
    for (int q = 0; q < 4; ++q)
        Ab[q] = ((lb[0] * d0.d[q] + lb[1] * d1.d[q]) + lb[2] * d2.d[q]) + lb[3] * d3.d[q];
    for (int q = 0; q < 4; ++q) W[q * cp + i] = Ab[q];
    Dual<2> px = seed<2>(pi.x, 0), py = seed<2>(pi.y, 1);
    Dual<2> sum(0.0);
    for (int k = 0; k < c; ++k) {
        const int j = nbr[(size_t)k * ld + i];
        const double2 pj = s.pos[j];
        Dual<2> t[4];
        kgc_pair<Dual<2>>(kc, px, py, Dual<2>(pj.x), Dual<2>(pj.y), Dual<2>(s.rho[j]), s.mass[j], t);
        for (int q = 0; q < 4; ++q) sum += t[q] * Ab[q];
    }
    add2(a.pos, i, sum.d[0], sum.d[1]);
}

__global__ void __launch_bounds__(BLOCK) k_adj_kgc_nbr(int n, KernelConsts kc, size_t ld,
                                                       StateView s, const int* __restrict__ nbr,
                                                       const int* __restrict__ cnt, AdjView a) {
    int j = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (j >= n) return;
    const double2 pj = s.pos[j];
    Dual<3> px = seed<3>(pj.x, 0), py = seed<3>(pj.y, 1), rj = seed<3>(s.rho[j], 2);
    const double mj = s.mass[j];
    const size_t cp = a.cap;
    Dual<3> sum(0.0);
    const int c = cnt[j];
    for (int k = 0; k < c; ++k) {
        const int i = nbr[(size_t)k * ld + j];
        if (s.tid[i] != 3) continue;
        const double2 pi = s.pos[i];
        Dual<3> t[4];
        kgc_pair<Dual<3>>(kc, Dual<3>(pi.x), Dual<3>(pi.y), px, py, rj, mj, t);
        for (int q = 0; q < 4; ++q) sum += t[q] * a.w[q * cp + i];
    }
    add2(a.pos, j, sum.d[0], sum.d[1]);
    a.rho[j] = a.rho[j] + sum.d[2];
}

// ===========================================================================
// effpor (porous.f90:107-224): n, rho = n oden, R from x_i, v_i and the
// porous parameters (por, alpha, beta).  Row-local: the soil is fixed.
// seeds: x, y, vx, vy, por, alpha, beta
// ===========================================================================
template <class T>
__device__ inline void effpor_row(const GridInfo& g, const KernelConsts& kc, const PorousParams& pp,
                                  const T& px, const T& py, const T& vx, const T& vy,
                                  const T& por, const T& alpha, const T& beta, double oden,
                                  const double2* spos, const double* sm, const double* sr,
                                  const int* sstart, T& nn, T& rho, T& R) {
    bool oob = false;
    const int cell = cell_of(g, val(px), val(py), oob);
    const int cx = cell / g.ny, cy = cell - cx * g.ny;
    T sf(0.0);
    const T cpor = 1.0 - por * 1.0;
    for (int dc = -1; dc <= 1; ++dc) {
        const int col = (cx + dc) * g.ny + cy;
        const int lo = sstart[col - 1], hi = sstart[col + 2];
        for (int q = lo; q < hi; ++q) {
            const double2 ps = spos[q];
            T dx = px - ps.x, dy = py - ps.y;
            T r = sqrt(dx * dx + dy * dy);
            if (r < kc.twoh) sf = sf + ((cpor * kw(kc, r)) * sm[q]) / sr[q];
        }
    }
    if (sf > (1.0 - por)) sf = 1.0 - por;
    nn = 1.0 - sf;
    rho = nn * oden;
    const T n3 = (nn * nn) * nn;
    const T vm = sqrt(vx * vx + vy * vy);
    const T omn = 1.0 - nn;
    if (pp.drag_law == ISPH_DRAG_REFERENCE) {
        T fch = 1.750 / sqrt(150.0 * n3);
        T t1 = (((visc(pp.visc_law, pp.nu, oden) / oden) * alpha) * (omn * omn)) /
               (n3 * (pp.d50 * pp.d50));
        T t2 = (((fch * vm) * sqrt(alpha)) * omn) / (sqrt(n3) * pp.d50);
        R = (-t1) - t2;
    } else {
        T t1 = ((alpha * pp.nu) * (omn * omn)) / (n3 * (pp.d50 * pp.d50));
        T t2 = ((beta * vm) * omn) / (n3 * pp.d50);
        R = (-t1) - t2;
    }
}

__global__ void k_adj_effpor(int n, GridInfo g, KernelConsts kc, PorousParams pp, StateView s,
                             const double2* __restrict__ spos, const double* __restrict__ sm,
                             const double* __restrict__ sr, const int* __restrict__ sstart,
                             AdjView a) {
    int i = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (i >= n) return;
    const double bR = a.R[i];
    a.R[i] = 0.0;  // R is overwritten for every particle
    if (s.tid[i] != 3) return;
    const double bn = a.por[i], br = a.rho[i];
    a.por[i] = 0.0;
    a.rho[i] = 0.0;
    const double2 p = s.pos[i], v = s.vel[i];
    Dual<7> nn, rho, R;
    effpor_row<Dual<7>>(g, kc, pp, seed<7>(p.x, 0), seed<7>(p.y, 1), seed<7>(v.x, 2),
                        seed<7>(v.y, 3), seed<7>(pp.por, 4), seed<7>(pp.alpha, 5),
                        seed<7>(pp.beta, 6), s.oden[i], spos, sm, sr, sstart, nn, rho, R);
    double gr[7];
    for (int q = 0; q < 7; ++q) gr[q] = (bn * nn.d[q] + br * rho.d[q]) + bR * R.d[q];
    add2(a.pos, i, gr[0], gr[1]);
    add2(a.vel, i, gr[2], gr[3]);
    const size_t cp = a.cap;
    a.th[i] = a.th[i] + gr[4];
    a.th[cp + i] = a.th[cp + i] + gr[5];
    a.th[2 * cp + i] = a.th[2 * cp + i] + gr[6];
}

// ===========================================================================
// eddyvis (turbulence.f90:12-110) and varts (integrator.f90:467-506)
// row seeds: x, y, vx, vy, n, L0..L3; neighbour seeds: x, y, vx, vy, n, rho
// ===========================================================================
template <class T>
__device__ inline void strain_pair(const KernelConsts& kc, const T* ri, const T* nj, double mj,
                                   bool eddy, T* acc) {
    // ri: x y vx vy n l0 l1 l2 l3; nj: x y vx vy n rho
    T r, wx, wy;
    pgeom(kc, ri[0], ri[1], nj[0], nj[1], r, wx, wy);
    T gx = ri[5] * wx + ri[6] * wy;
    T gy = ri[7] * wx + ri[8] * wy;
    const T ui = ri[2] / ri[4], wi = ri[3] / ri[4];
    if (eddy) {
        T du = nj[2] / nj[4] - ui;
        T dv = nj[3] / nj[4] - wi;
        acc[0] = ((mj * du) * gx) / nj[5];
        acc[1] = ((mj * dv) * gy) / nj[5];
        acc[2] = ((0.50 * mj) * (du * gy + dv * gx)) / nj[5];
    } else {
        T vol = mj / nj[5];
        acc[0] = ((nj[2] / nj[4] - ui) * gx) * vol;
        acc[1] = ((nj[3] / nj[4] - wi) * gy) * vol;
        acc[2] = T(0.0);
    }
}

template <int NR>
__device__ __forceinline__ void load_strain_row(const StateView& s, const double* L, int i, Dual<NR>* ri,
                                                bool seeded) {
    const double2 p = s.pos[i], v = s.vel[i];
    double l[4];
    ld_L(L, i, l);
    const double x[9] = {p.x, p.y, v.x, v.y, s.por[i], l[0], l[1], l[2], l[3]};
    for (int q = 0; q < 9; ++q) ri[q] = seeded ? seed<NR>(x[q], q) : Dual<NR>(x[q]);
}

// eddy = true: eddyvis (fluid rows), false: varts (free rows)
__global__ void __launch_bounds__(BLOCK) k_adj_strain_self(int n, KernelConsts kc, double dl,
                                                           int eddy, size_t ld, StateView s,
                                                           const int* __restrict__ nbr,
                                                           const int* __restrict__ cnt,
                                                           const double* __restrict__ L,
                                                           const SimScalars* ss, AdjView a) {
    int i = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (i >= n) return;
    const size_t cp = a.cap;
    const bool act = eddy ? (s.tid[i] == 3) : ((s.flags[i] & ISPH_FLAG_FREE) != 0);
    if (!act) {
        for (int q = 0; q < 3; ++q) a.w[q * cp + i] = 0.0;
        return;
    }
    double ob0, ob1;
    if (eddy) {
        ob0 = a.nut[i];
        ob1 = a.tke[i];
        a.nut[i] = 0.0;
        a.tke[i] = 0.0;
    } else {
        ob0 = a.varts[i];
        ob1 = 0.0;
        a.varts[i] = 0.0;
    }
    const int c = cnt[i];
    double acc[3] = {0.0, 0.0, 0.0};
    {
        double ri[9];
        const double2 p = s.pos[i], v = s.vel[i];
        double l[4];
        ld_L(L, i, l);
        ri[0] = p.x; ri[1] = p.y; ri[2] = v.x; ri[3] = v.y; ri[4] = s.por[i];
        ri[5] = l[0]; ri[6] = l[1]; ri[7] = l[2]; ri[8] = l[3];
        for (int k = 0; k < c; ++k) {
            const int j = nbr[(size_t)k * ld + i];
            const double2 pj = s.pos[j], vj = s.vel[j];
            const double nj[6] = {pj.x, pj.y, vj.x, vj.y, s.por[j], s.rho[j]};
            double t[3];
            strain_pair<double>(kc, ri, nj, s.mass[j], eddy, t);
            if (eddy) {
                acc[0] = acc[0] + t[0];
                acc[1] = acc[1] + t[1];
                acc[2] = acc[2] + t[2];
            } else {
                acc[0] = acc[0] + t[0];
                acc[1] = acc[1] + t[1];
            }
        }
    }
    // finish: seeds t1, t2, t3 (eddy) or t1, t2, vx, vy, n, umax (varts)
    double Ab[3] = {0.0, 0.0, 0.0};
    double dvx = 0.0, dvy = 0.0, dn = 0.0, dum = 0.0;
    if (eddy) {
        Dual<3> t1 = seed<3>(acc[0], 0), t2 = seed<3>(acc[1], 1), t3 = seed<3>(acc[2], 2);
        const double cs = 0.150, cv = 0.080, ce = 1.0;
        Dual<3> t4 = sqrt(2.0 * (t1 * t1 + t2 * t2) + (2.0 * t3) * (2.0 * t3));
        Dual<3> nut = ((cs * dl) * (cs * dl)) * t4;
        Dual<3> tke = (cv / ce) * ((dl * t4) * (dl * t4));
        for (int q = 0; q < 3; ++q) Ab[q] = ob0 * nut.d[q] + ob1 * tke.d[q];
    } else {
        const double2 v = s.vel[i];
        Dual<6> t1 = seed<6>(acc[0], 0), t2 = seed<6>(acc[1], 1), vx = seed<6>(v.x, 2),
                vy = seed<6>(v.y, 3), ni = seed<6>(s.por[i], 4), um = seed<6>(ss->umax, 5);
        const double ps = 1.0;
        Dual<6> vt = (((ps * (dl * dl)) * sqrt(t1 * t1 + t2 * t2)) * sqrt(vx * vx + vy * vy)) /
                     ((um + TINY6) * ni);
        Ab[0] = ob0 * vt.d[0];
        Ab[1] = ob0 * vt.d[1];
        dvx = ob0 * vt.d[2];
        dvy = ob0 * vt.d[3];
        dn = ob0 * vt.d[4];
        dum = ob0 * vt.d[5];
    }
    for (int q = 0; q < 3; ++q) a.w[q * cp + i] = Ab[q];
    Dual<9> ri[9];
    load_strain_row<9>(s, L, i, ri, true);
    Dual<9> sum(0.0);
    for (int k = 0; k < c; ++k) {
        const int j = nbr[(size_t)k * ld + i];
        const double2 pj = s.pos[j], vj = s.vel[j];
        const Dual<9> nj[6] = {Dual<9>(pj.x), Dual<9>(pj.y), Dual<9>(vj.x),
                               Dual<9>(vj.y), Dual<9>(s.por[j]), Dual<9>(s.rho[j])};
        Dual<9> t[3];
        strain_pair<Dual<9>>(kc, ri, nj, s.mass[j], eddy, t);
        sum += t[0] * Ab[0];
        sum += t[1] * Ab[1];
        if (eddy) sum += t[2] * Ab[2];
    }
    add2(a.pos, i, sum.d[0], sum.d[1]);
    add2(a.vel, i, sum.d[2] + dvx, sum.d[3] + dvy);
    a.por[i] = a.por[i] + (sum.d[4] + dn);
    const double lg[4] = {sum.d[5], sum.d[6], sum.d[7], sum.d[8]};
    add_L(a.L, i, lg);
    if (!eddy) a.aum[i] = a.aum[i] + dum;
}

__global__ void __launch_bounds__(BLOCK) k_adj_strain_nbr(int n, KernelConsts kc, int eddy,
                                                          size_t ld, StateView s,
                                                          const int* __restrict__ nbr,
                                                          const int* __restrict__ cnt,
                                                          const double* __restrict__ L, AdjView a) {
    int j = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (j >= n) return;
    const size_t cp = a.cap;
    const double2 pj = s.pos[j], vj = s.vel[j];
    const Dual<6> nj[6] = {seed<6>(pj.x, 0), seed<6>(pj.y, 1), seed<6>(vj.x, 2),
                           seed<6>(vj.y, 3), seed<6>(s.por[j], 4), seed<6>(s.rho[j], 5)};
    const double mj = s.mass[j];
    Dual<6> sum(0.0);
    const int c = cnt[j];
    for (int k = 0; k < c; ++k) {
        const int i = nbr[(size_t)k * ld + j];
        const bool act = eddy ? (s.tid[i] == 3) : ((s.flags[i] & ISPH_FLAG_FREE) != 0);
        if (!act) continue;
        Dual<6> ri[9];
        load_strain_row<6>(s, L, i, ri, false);
        Dual<6> t[3];
        strain_pair<Dual<6>>(kc, ri, nj, mj, eddy, t);
        sum += t[0] * a.w[i];
        sum += t[1] * a.w[cp + i];
        if (eddy) sum += t[2] * a.w[2 * cp + i];
    }
    add2(a.pos, j, sum.d[0], sum.d[1]);
    add2(a.vel, j, sum.d[2], sum.d[3]);
    a.por[j] = a.por[j] + sum.d[4];
    a.rho[j] = a.rho[j] + sum.d[5];
}

// ===========================================================================
// int_vel v* (integrator.f90:509-698)
// row seeds (12): x y vx vy n nut tke varts L0..L3
// neighbour seeds (8): x y vx vy rho nut tke varts
// ===========================================================================
struct VstarConst {
    KernelConsts kc;
    PorousParams pp;
    double lam, g, w0;
};

template <class T>
__device__ inline void vstar_pair(const VstarConst& vc, const T* ri, int fi, double oi, const T* nj,
                                  int fj, int tj, double oj, double mj, T* acc) {
    const KernelConsts& kc = vc.kc;
    T r, wx, wy;
    pgeom(kc, ri[0], ri[1], nj[0], nj[1], r, wx, wy);
    T gx = ri[8] * wx + ri[9] * wy;
    T gy = ri[10] * wx + ri[11] * wy;
    const double mui = visc(vc.pp.visc_law, vc.pp.nu, oi);
    double cmu;
    if (tj == 3) {
        const double muj = visc(vc.pp.visc_law, vc.pp.nu, oj);
        cmu = (((2.0 * mui) * muj) * ((1.0 / oi) * (1.0 / oj))) / ((mui / oi) + (muj / oj));
    } else {
        cmu = mui * (1.0 / oi);
    }
    T vart(0.0);
    if (fi & ISPH_FLAG_FREE) vart = ri[7];
    else if ((fj & ISPH_FLAG_FREE) && !(fi & ISPH_FLAG_BUFFER)) vart = (kw(kc, r) * nj[7]) / vc.w0;
    const T pvol = mj / nj[4];
    const T pdist = 1.0 / (r * r + vc.lam);
    const T dxp = ri[0] - nj[0], dyp = ri[1] - nj[1];
    const T nuti = ri[5], nutj = nj[5];
    T sx(0.0), sy(0.0);
    T dvx = ri[2] - nj[2];
    T t1 = ((((2.0 * pvol) * dvx) * dxp) * gx) * pdist;
    T t2 = ((((2.0 * pvol) * dvx) * dyp) * gy) * pdist;
    sx = (t1 + t2) * (cmu + vart);
    bool small = false;
    if (tj > 2) {
        small = nuti < TINY6 && nutj < TINY6;
        if (small) sx = sx + ((t1 + t2) * (nuti + nutj)) * 0.50;
        else sx = sx + (t1 + t2) * (((2.0 * nuti) * nutj) / (nuti + nutj));
    }
    T dvy = ri[3] - nj[3];
    t1 = ((((2.0 * pvol) * dvy) * dxp) * gx) * pdist;
    t2 = ((((2.0 * pvol) * dvy) * dyp) * gy) * pdist;
    sy = (t1 + t2) * (cmu + vart);
    if (tj > 2) {
        if (small) sy = sy + ((t1 + t2) * (nuti + nutj)) * 0.50;
        else sy = sy + (t1 + t2) * (((2.0 * nuti) * nutj) / (nuti + nutj));
    }
    const T dtk = nj[6] - ri[6];
    t1 = (pvol * dtk) * gx;
    sx = sx - ((t1 * ri[4]) * 2.0) / 3.0;
    t1 = (pvol * dtk) * gy;
    sy = sy - ((t1 * ri[4]) * 2.0) / 3.0;
    acc[0] = sx;
    acc[1] = sy;
}

template <int N>
__device__ __forceinline__ void load_vstar_row(const StateView& s, const double* L, int i, Dual<N>* ri,
                                               bool seeded) {
    const double2 p = s.pos[i], v = s.vel[i];
    double l[4];
    ld_L(L, i, l);
    const double x[12] = {p.x, p.y, v.x, v.y, s.por[i], s.nut[i], s.tke[i], s.varts[i],
                          l[0], l[1], l[2], l[3]};
    for (int q = 0; q < 12; ++q) ri[q] = seeded ? seed<N>(x[q], q) : Dual<N>(x[q]);
}

template <int N>
__device__ __forceinline__ void load_vstar_nbr(const StateView& s, int j, Dual<N>* nj, bool seeded) {
    const double2 p = s.pos[j], v = s.vel[j];
    const double x[8] = {p.x, p.y, v.x, v.y, s.rho[j], s.nut[j], s.tke[j], s.varts[j]};
    for (int q = 0; q < 8; ++q) nj[q] = seeded ? seed<N>(x[q], q) : Dual<N>(x[q]);
}

__global__ void __launch_bounds__(BLOCK) k_adj_vstar_self(int n, VstarConst vc, size_t ld,
                                                          StateView s, const int* __restrict__ nbr,
                                                          const int* __restrict__ cnt,
                                                          const double* __restrict__ L,
                                                          const SimScalars* ss, AdjView a) {
    int i = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (i >= n) return;
    const size_t cp = a.cap;
    const int fi = s.flags[i];
    const double dt = ss->dt;
    const double2 ob = a.sv[i];
    if (fi & ISPH_FLAG_BUFFER) {
        // sv = (vx, g n dt)
        a.sv[i] = make_double2(0.0, 0.0);
        add2(a.vel, i, ob.x, 0.0);
        a.por[i] = a.por[i] + ob.y * (vc.g * dt);
    }
    if (!is_active_fluid(s.tid[i], fi)) {
        a.w[i] = 0.0;
        a.w[cp + i] = 0.0;
        return;
    }
    a.sv[i] = make_double2(0.0, 0.0);
    const double oi = s.oden[i];
    const int c = cnt[i];
    double acc[2] = {0.0, 0.0};
    {
        Dual<1> ri[12];
        load_vstar_row<1>(s, L, i, ri, false);
        for (int k = 0; k < c; ++k) {
            const int j = nbr[(size_t)k * ld + i];
            Dual<1> nj[8];
            load_vstar_nbr<1>(s, j, nj, false);
            Dual<1> t[2];
            vstar_pair<Dual<1>>(vc, ri, fi, oi, nj, s.flags[j], s.tid[j], s.oden[j], s.mass[j], t);
            acc[0] = acc[0] + t[0].v;
            acc[1] = acc[1] + t[1].v;
        }
    }
    // finish: seeds sx, sy, vx, vy, R, n
    const double2 v = s.vel[i];
    Dual<6> sx = seed<6>(acc[0], 0), sy = seed<6>(acc[1], 1), vx = seed<6>(v.x, 2),
            vy = seed<6>(v.y, 3), Ri = seed<6>(s.R[i], 4), ni = seed<6>(s.por[i], 5);
    Dual<6> den = 1.0 - (dt * Ri) * ni;
    Dual<6> vxs = (vx + dt * sx) / den;
    Dual<6> vys = (vy + dt * sy) / den;
    vys = vys + (vc.g * ni) * dt;
    double f[6];
    for (int q = 0; q < 6; ++q) f[q] = ob.x * vxs.d[q] + ob.y * vys.d[q];
    a.w[i] = f[0];
    a.w[cp + i] = f[1];
    Dual<12> ri[12];
    load_vstar_row<12>(s, L, i, ri, true);
    Dual<12> sum(0.0);
    for (int k = 0; k < c; ++k) {
        const int j = nbr[(size_t)k * ld + i];
        Dual<12> nj[8];
        load_vstar_nbr<12>(s, j, nj, false);
        Dual<12> t[2];
        vstar_pair<Dual<12>>(vc, ri, fi, oi, nj, s.flags[j], s.tid[j], s.oden[j], s.mass[j], t);
        sum += t[0] * f[0];
        sum += t[1] * f[1];
    }
    add2(a.pos, i, sum.d[0], sum.d[1]);
    add2(a.vel, i, sum.d[2] + f[2], sum.d[3] + f[3]);
    a.R[i] = a.R[i] + f[4];
    a.por[i] = a.por[i] + (sum.d[4] + f[5]);
    a.nut[i] = a.nut[i] + sum.d[5];
    a.tke[i] = a.tke[i] + sum.d[6];
    a.varts[i] = a.varts[i] + sum.d[7];
    const double lg[4] = {sum.d[8], sum.d[9], sum.d[10], sum.d[11]};
    add_L(a.L, i, lg);
}

__global__ void __launch_bounds__(BLOCK) k_adj_vstar_nbr(int n, VstarConst vc, size_t ld,
                                                         StateView s, const int* __restrict__ nbr,
                                                         const int* __restrict__ cnt,
                                                         const double* __restrict__ L, AdjView a) {
    int j = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (j >= n) return;
    const size_t cp = a.cap;
    Dual<8> nj[8];
    load_vstar_nbr<8>(s, j, nj, true);
    const int fj = s.flags[j], tj = s.tid[j];
    const double oj = s.oden[j], mj = s.mass[j];
    Dual<8> sum(0.0);
    const int c = cnt[j];
    for (int k = 0; k < c; ++k) {
        const int i = nbr[(size_t)k * ld + j];
        const int fi = s.flags[i];
        if (!is_active_fluid(s.tid[i], fi)) continue;
        Dual<8> ri[12];
        load_vstar_row<8>(s, L, i, ri, false);
        Dual<8> t[2];
        vstar_pair<Dual<8>>(vc, ri, fi, s.oden[i], nj, fj, tj, oj, mj, t);
        sum += t[0] * a.w[i];
        sum += t[1] * a.w[cp + i];
    }
    add2(a.pos, j, sum.d[0], sum.d[1]);
    add2(a.vel, j, sum.d[2], sum.d[3]);
    a.rho[j] = a.rho[j] + sum.d[4];
    a.nut[j] = a.nut[j] + sum.d[5];
    a.tke[j] = a.tke[j] + sum.d[6];
    a.varts[j] = a.varts[j] + sum.d[7];
}

// ===========================================================================
// PPE assembly (isph.f90:123-456).  Row i before Jacobi scaling: off-diagonal
// v_ij, diagonal dia_i = sum_j dterm_ij, rhs bs_i = sum_j bterm_ij; scaled:
// A_ij = v_ij / dia_i, b_i = bs_i / dia_i.  With lambda = A^-T pbar:
// bbar_i = lambda_i, Abar_ij = -lambda_i p_j, so
//   vbar_ij = -lambda_i p_j / dia_i,  bsbar_i = lambda_i / dia_i,
//   diabar_i = lambda_i (sum_j v_ij p_j - bs_i) / dia_i^2.
// row seeds (10): x y svx svy n rho L0..L3; neighbour seeds (10): the same
// ===========================================================================
struct AsmConst {
    KernelConsts kc;
    double lam, lampd;
};

template <class T>
__device__ inline void asm_pair(const AsmConst& ac, const T* ri, int ti, bool freei, double mi,
                                const T* nj, int tj, double odenj, double mj, double dt, T& v,
                                T& dterm, T& bterm) {
    const KernelConsts& kc = ac.kc;
    v = T(0.0);
    dterm = T(0.0);
    bterm = T(0.0);
    const double lamp = (freei && tj == 3) ? (1.0 + odenj / ac.lampd) : 1.0;
    const T ni = ri[4], rhoi = ri[5];
    const T nj_ = nj[4], rhoj = nj[5];
    const T rho_j = rhoj * (1.0 / nj_);
    const T pvol = mj * (1.0 / rhoj);
    T ddx = nj[0] - ri[0], ddy = nj[1] - ri[1];
    T rr = sqrt(ddx * ddx + ddy * ddy);
    T pdist = 1.0 / (rr * rr + ac.lam);
    T r, wx, wy;
    pgeom(kc, ri[0], ri[1], nj[0], nj[1], r, wx, wy);
    const T dxp = ri[0] - nj[0], dyp = ri[1] - nj[1];
    const T rho_i = rhoi * (1.0 / ni);
    T t1, t2;
    if (ti == 3) {
        const T rho_ij = (((2.0 * rhoj) * rhoi) * (1.0 / (nj_ * ni))) *
                         (1.0 / (rhoj * (1.0 / nj_) + rhoi * (1.0 / ni)));
        T gx = ri[6] * wx + ri[7] * wy;
        T gy = ri[8] * wx + ri[9] * wy;
        t1 = (((2.0 * pvol) * gx) * dxp) * pdist;
        t2 = (((2.0 * pvol) * gy) * dyp) * pdist;
        v = -(t1 + t2);
        dterm = (t1 + t2) * lamp;
        if (tj == 3) {
            t1 = (pvol * ((nj[2] * rho_ij) / nj_ - (ri[2] * rho_ij) / ni)) * gx;
            t2 = (pvol * ((nj[3] * rho_ij) / nj_ - (ri[3] * rho_ij) / ni)) * gy;
        } else {
            t1 = (pvol * ((nj[2] * rho_i) / ni - (ri[2] * rho_i) / ni)) * gx;
            t2 = (pvol * ((nj[3] * rho_i) / ni - (ri[3] * rho_i) / ni)) * gy;
        }
        bterm = (t1 + t2) / dt;
    } else if (ti == 2) {
        t1 = (((2.0 * pvol) * wx) * dxp) * pdist;
        t2 = (((2.0 * pvol) * wy) * dyp) * pdist;
        if (tj <= 2) {
            v = -(t1 + t2);
            dterm = t1 + t2;
        }
    } else if (ti == 1) {
        if (tj == 3) {
            T gjx = nj[6] * wx + nj[7] * wy;
            T gjy = nj[8] * wx + nj[9] * wy;
            t1 = ((((2.0 * mi) * gjx) * dxp) * pdist) / rhoi;
            t2 = ((((2.0 * mi) * gjy) * dyp) * pdist) / rhoi;
            v = (-(t1 + t2)) * 1.0;
            dterm = t1 + t2;
            t1 = ((mi * ((nj[2] * rho_j) / nj_ - (ri[2] * rho_j) / ni)) * gjx) / rhoi;
            t2 = ((mi * ((nj[3] * rho_j) / nj_ - (ri[3] * rho_j) / ni)) * gjy) / rhoi;
            bterm = (t1 + t2) / dt;
        } else if (tj <= 2) {
            t1 = (((2.0 * pvol) * wx) * dxp) * pdist;
            t2 = (((2.0 * pvol) * wy) * dyp) * pdist;
            v = -(t1 + t2);
            dterm = t1 + t2;
        }
    }
}

template <int N>
__device__ __forceinline__ void load_asm(const StateView& s, const double* L, int i, Dual<N>* u,
                                         bool seeded) {
    const double2 p = s.pos[i], sv = s.sv[i];
    double l[4];
    ld_L(L, i, l);
    const double x[10] = {p.x, p.y, sv.x, sv.y, s.por[i], s.rho[i], l[0], l[1], l[2], l[3]};
    for (int q = 0; q < 10; ++q) u[q] = seeded ? seed<N>(x[q], q) : Dual<N>(x[q]);
}

__device__ __forceinline__ void add_u10(const StateView& s, AdjView& a, int i, const double* d) {
    add2(a.pos, i, d[0], d[1]);
    add2(a.sv, i, d[2], d[3]);
    a.por[i] = a.por[i] + d[4];
    a.rho[i] = a.rho[i] + d[5];
    const double lg[4] = {d[6], d[7], d[8], d[9]};
    add_L(a.L, i, lg);
}

// lam: the adjoint PPE solution, p: the pressure the forward solve returned
__global__ void __launch_bounds__(BLOCK) k_adj_asm_self(int n, AsmConst ac, size_t ld, StateView s,
                                                        const int* __restrict__ nbr,
                                                        const int* __restrict__ cnt,
                                                        const double* __restrict__ L,
                                                        const double* __restrict__ lam,
                                                        const double* __restrict__ p,
                                                        const SimScalars* ss, AdjView a) {
    int i = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (i >= n) return;
    const size_t cp = a.cap;
    const double dt = ss->dt;
    const int ti = s.tid[i];
    const bool freei = (s.flags[i] & ISPH_FLAG_FREE) != 0;
    const double mi = s.mass[i];
    const int c = cnt[i];
    double dia = 0.0, bs = 0.0, vp = 0.0;
    {
        Dual<1> ri[10];
        load_asm<1>(s, L, i, ri, false);
        for (int k = 0; k < c; ++k) {
            const int j = nbr[(size_t)k * ld + i];
            Dual<1> nj[10];
            load_asm<1>(s, L, j, nj, false);
            Dual<1> v, dt_, bt;
            asm_pair<Dual<1>>(ac, ri, ti, freei, mi, nj, s.tid[j], s.oden[j], s.mass[j], dt, v, dt_, bt);
            dia = dia + dt_.v;
            bs = bs + bt.v;
            vp = vp + v.v * p[j];
        }
    }
    const double li = lam[i];
    const double w0 = -li / dia;                      // vbar_ij = w0 * p_j
    const double w1 = (li * (vp - bs)) / (dia * dia);  // diabar
    const double w2 = li / dia;                        // bsbar
    a.w[i] = w0;
    a.w[cp + i] = w1;
    a.w[2 * cp + i] = w2;
    Dual<10> ri[10];
    load_asm<10>(s, L, i, ri, true);
    Dual<10> sum(0.0);
    for (int k = 0; k < c; ++k) {
        const int j = nbr[(size_t)k * ld + i];
        Dual<10> nj[10];
        load_asm<10>(s, L, j, nj, false);
        Dual<10> v, dt_, bt;
        asm_pair<Dual<10>>(ac, ri, ti, freei, mi, nj, s.tid[j], s.oden[j], s.mass[j], dt, v, dt_, bt);
        sum += v * (w0 * p[j]);
        sum += dt_ * w1;
        sum += bt * w2;
    }
    add_u10(s, a, i, sum.d);
}

__global__ void __launch_bounds__(BLOCK) k_adj_asm_nbr(int n, AsmConst ac, size_t ld, StateView s,
                                                       const int* __restrict__ nbr,
                                                       const int* __restrict__ cnt,
                                                       const double* __restrict__ L,
                                                       const double* __restrict__ p,
                                                       const SimScalars* ss, AdjView a) {
    int j = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (j >= n) return;
    const size_t cp = a.cap;
    const double dt = ss->dt;
    Dual<10> nj[10];
    load_asm<10>(s, L, j, nj, true);
    const int tj = s.tid[j];
    const double oj = s.oden[j], mj = s.mass[j], pj = p[j];
    Dual<10> sum(0.0);
    const int c = cnt[j];
    for (int k = 0; k < c; ++k) {
        const int i = nbr[(size_t)k * ld + j];
        Dual<10> ri[10];
        load_asm<10>(s, L, i, ri, false);
        Dual<10> v, dt_, bt;
        asm_pair<Dual<10>>(ac, ri, s.tid[i], (s.flags[i] & ISPH_FLAG_FREE) != 0, s.mass[i], nj, tj,
                           oj, mj, dt, v, dt_, bt);
        sum += v * (a.w[i] * pj);
        sum += dt_ * a.w[cp + i];
        sum += bt * a.w[2 * cp + i];
    }
    add_u10(s, a, j, sum.d);
}

// ===========================================================================
// comp_vel (integrator.f90:715-865): v = v* + sum_j (pressure gradient) dt,
// vel = |v| / n.  row seeds (9): x y n rho p L0..L3;
// neighbour seeds (5): x y n rho p
// ===========================================================================
template <class T>
__device__ inline void cv_pair(const KernelConsts& kc, const T* ri, const T* nj, int tj, double mj,
                               double dt, T* acc) {
    T r, wx, wy;
    pgeom(kc, ri[0], ri[1], nj[0], nj[1], r, wx, wy);
    T gx = ri[5] * wx + ri[6] * wy;
    T gy = ri[7] * wx + ri[8] * wy;
    const T vj = mj / nj[3];
    const T ni = ri[2], rhoi = ri[3], p_i = ri[4];
    T t1;
    if (tj == 3) {
        const T nj_ = nj[2], rhoj = nj[3], pj = nj[4];
        T q = ((pj * rhoi) / ni) + ((p_i * rhoj) / nj_);
        T den = rhoi * ((rhoi / ni) + (rhoj / nj_));
        t1 = (2.0 * (((vj * (ni * ni)) * q) * gx)) / den;
        acc[0] = -(t1 * dt);
        t1 = (2.0 * (((vj * (ni * ni)) * q) * gy)) / den;
        acc[1] = -(t1 * dt);
    } else {
        T q = nj[4] + p_i;
        t1 = (((vj * (ni * ni)) * q) * gx) / rhoi;
        acc[0] = -(t1 * dt);
        t1 = (((vj * (ni * ni)) * q) * gy) / rhoi;
        acc[1] = -(t1 * dt);
    }
}

template <int N>
__device__ __forceinline__ void load_cv_row(const StateView& s, const double* L, const double* p, int i,
                                            Dual<N>* u, bool seeded) {
    const double2 q = s.pos[i];
    double l[4];
    ld_L(L, i, l);
    const double x[9] = {q.x, q.y, s.por[i], s.rho[i], p[i], l[0], l[1], l[2], l[3]};
    for (int k = 0; k < 9; ++k) u[k] = seeded ? seed<N>(x[k], k) : Dual<N>(x[k]);
}

template <int N>
__device__ __forceinline__ void load_cv_nbr(const StateView& s, const double* p, int j, Dual<N>* u,
                                            bool seeded) {
    const double2 q = s.pos[j];
    const double x[5] = {q.x, q.y, s.por[j], s.rho[j], p[j]};
    for (int k = 0; k < 5; ++k) u[k] = seeded ? seed<N>(x[k], k) : Dual<N>(x[k]);
}

__global__ void __launch_bounds__(BLOCK) k_adj_cv_self(int n, KernelConsts kc, size_t ld, StateView s,
                                                       const int* __restrict__ nbr,
                                                       const int* __restrict__ cnt,
                                                       const double* __restrict__ L,
                                                       const double* __restrict__ p,
                                                       const SimScalars* ss, AdjView a) {
    int i = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (i >= n) return;
    const size_t cp = a.cap;
    const int fi = s.flags[i];
    const double dt = ss->dt;
    const double2 ov = a.vel[i];
    const double om = a.velm[i];
    if (fi & ISPH_FLAG_BUFFER) {  // vel = (v*x, 0)
        a.vel[i] = make_double2(0.0, 0.0);
        add2(a.sv, i, ov.x, 0.0);
    }
    if ((fi & ISPH_FLAG_DEAD) && !(fi & ISPH_FLAG_BUFFER)) {
        a.vel[i] = make_double2(0.0, 0.0);
        a.sv[i] = make_double2(0.0, 0.0);
    }
    if (!is_active_fluid(s.tid[i], fi)) {
        a.w[i] = 0.0;
        a.w[cp + i] = 0.0;
        return;
    }
    a.vel[i] = make_double2(0.0, 0.0);
    a.velm[i] = 0.0;
    const int c = cnt[i];
    double acc[2] = {0.0, 0.0};
    {
        Dual<1> ri[9];
        load_cv_row<1>(s, L, p, i, ri, false);
        for (int k = 0; k < c; ++k) {
            const int j = nbr[(size_t)k * ld + i];
            Dual<1> nj[5];
            load_cv_nbr<1>(s, p, j, nj, false);
            Dual<1> t[2];
            cv_pair<Dual<1>>(kc, ri, nj, s.tid[j], s.mass[j], dt, t);
            acc[0] = acc[0] + t[0].v;
            acc[1] = acc[1] + t[1].v;
        }
    }
    // finish: vel = sv + acc, vel_m = |vel| / n; seeds ax, ay, svx, svy, n
    const double2 sv = s.sv[i];
    Dual<5> ax = seed<5>(acc[0], 0), ay = seed<5>(acc[1], 1), svx = seed<5>(sv.x, 2),
            svy = seed<5>(sv.y, 3), ni = seed<5>(s.por[i], 4);
    Dual<5> vx = svx + ax, vy = svy + ay;
    Dual<5> vm = sqrt(vx * vx + vy * vy) / ni;
    double f[5];
    for (int q = 0; q < 5; ++q) f[q] = (ov.x * vx.d[q] + ov.y * vy.d[q]) + om * vm.d[q];
    a.w[i] = f[0];
    a.w[cp + i] = f[1];
    add2(a.sv, i, f[2], f[3]);
    Dual<9> ri[9];
    load_cv_row<9>(s, L, p, i, ri, true);
    Dual<9> sum(0.0);
    for (int k = 0; k < c; ++k) {
        const int j = nbr[(size_t)k * ld + i];
        Dual<9> nj[5];
        load_cv_nbr<9>(s, p, j, nj, false);
        Dual<9> t[2];
        cv_pair<Dual<9>>(kc, ri, nj, s.tid[j], s.mass[j], dt, t);
        sum += t[0] * f[0];
        sum += t[1] * f[1];
    }
    add2(a.pos, i, sum.d[0], sum.d[1]);
    a.por[i] = a.por[i] + (sum.d[2] + f[4]);
    a.rho[i] = a.rho[i] + sum.d[3];
    a.p[i] = a.p[i] + sum.d[4];
    const double lg[4] = {sum.d[5], sum.d[6], sum.d[7], sum.d[8]};
    add_L(a.L, i, lg);
}

__global__ void __launch_bounds__(BLOCK) k_adj_cv_nbr(int n, KernelConsts kc, size_t ld, StateView s,
                                                      const int* __restrict__ nbr,
                                                      const int* __restrict__ cnt,
                                                      const double* __restrict__ L,
                                                      const double* __restrict__ p,
                                                      const SimScalars* ss, AdjView a) {
    int j = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (j >= n) return;
    const size_t cp = a.cap;
    const double dt = ss->dt;
    Dual<5> nj[5];
    load_cv_nbr<5>(s, p, j, nj, true);
    const int tj = s.tid[j];
    const double mj = s.mass[j];
    Dual<5> sum(0.0);
    const int c = cnt[j];
    for (int k = 0; k < c; ++k) {
        const int i = nbr[(size_t)k * ld + j];
        if (!is_active_fluid(s.tid[i], s.flags[i])) continue;
        Dual<5> ri[9];
        load_cv_row<5>(s, L, p, i, ri, false);
        Dual<5> t[2];
        cv_pair<Dual<5>>(kc, ri, nj, tj, mj, dt, t);
        sum += t[0] * a.w[i];
        sum += t[1] * a.w[cp + i];
    }
    add2(a.pos, j, sum.d[0], sum.d[1]);
    a.por[j] = a.por[j] + sum.d[2];
    a.rho[j] = a.rho[j] + sum.d[3];
    a.p[j] = a.p[j] + sum.d[4];
}

// ===========================================================================
// opt2_shift + apply (part_shift.f90:12-192): x += clamp(dx), v += dv with
// dx = -cf sum_j V_j grad W, dv = (t1 dx + t2 dy, t3 dx + t4 dy).
// row seeds (8): x y vx vy L0..L3; neighbour seeds (5): x y vx vy rho
// ===========================================================================
template <class T>
__device__ inline void shift_pair(const KernelConsts& kc, double cf, const T* ri, const T* nj,
                                  double mj, T* acc) {
    T r, wx, wy;
    pgeom(kc, ri[0], ri[1], nj[0], nj[1], r, wx, wy);
    const T vj = mj / nj[4];
    acc[0] = -((cf * wx) * vj);
    acc[1] = -((cf * wy) * vj);
    T gx = ri[4] * wx + ri[5] * wy;
    T gy = ri[6] * wx + ri[7] * wy;
    T du = nj[2] - ri[2], dv = nj[3] - ri[3];
    acc[2] = (gx * du) * vj;
    acc[3] = (gy * du) * vj;
    acc[4] = (gx * dv) * vj;
    acc[5] = (gy * dv) * vj;
}

template <int N>
__device__ __forceinline__ void load_shift_row(const StateView& s, const double* L, int i, Dual<N>* u,
                                               bool seeded) {
    const double2 p = s.pos[i], v = s.vel[i];
    double l[4];
    ld_L(L, i, l);
    const double x[8] = {p.x, p.y, v.x, v.y, l[0], l[1], l[2], l[3]};
    for (int k = 0; k < 8; ++k) u[k] = seeded ? seed<N>(x[k], k) : Dual<N>(x[k]);
}

template <int N>
__device__ __forceinline__ void load_shift_nbr(const StateView& s, int j, Dual<N>* u, bool seeded) {
    const double2 p = s.pos[j], v = s.vel[j];
    const double x[5] = {p.x, p.y, v.x, v.y, s.rho[j]};
    for (int k = 0; k < 5; ++k) u[k] = seeded ? seed<N>(x[k], k) : Dual<N>(x[k]);
}

__device__ __forceinline__ bool shift_row(const StateView& s, const int* cnt, int i) {
    const int fi = s.flags[i];
    return s.tid[i] == 3 && !(fi & ISPH_FLAG_BUFFER) && !(fi & ISPH_FLAG_DEAD) && cnt[i] > 0;
}

__device__ __forceinline__ double shift_cf(const KernelConsts& kc, double csh, int fi) {
    const double frac = (fi & ISPH_FLAG_FREE) ? 0.01 : 1.0;
    return (csh * (kc.h * kc.h)) * frac;
}

__global__ void __launch_bounds__(BLOCK) k_adj_shift_self(int n, KernelConsts kc, double csh,
                                                          double cap, size_t ld, StateView s,
                                                          const int* __restrict__ nbr,
                                                          const int* __restrict__ cnt,
                                                          const double* __restrict__ L, AdjView a) {
    int i = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (i >= n) return;
    const size_t cp = a.cap;
    if (!shift_row(s, cnt, i)) {
        for (int q = 0; q < 6; ++q) a.w[q * cp + i] = 0.0;
        return;
    }
    const double cf = shift_cf(kc, csh, s.flags[i]);
    const int c = cnt[i];
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    {
        Dual<1> ri[8];
        load_shift_row<1>(s, L, i, ri, false);
        for (int k = 0; k < c; ++k) {
            const int j = nbr[(size_t)k * ld + i];
            Dual<1> nj[5];
            load_shift_nbr<1>(s, j, nj, false);
            Dual<1> t[6];
            shift_pair<Dual<1>>(kc, cf, ri, nj, s.mass[j], t);
            for (int q = 0; q < 6; ++q) acc[q] = acc[q] + t[q].v;
        }
    }
    const double xs = acc[0], ys = acc[1];
    const double2 ap = a.pos[i], av = a.vel[i];
    // clamp(x) = sign(x) min(|x|, cap): slope 1 below the cap, 0 above
    const double sx = fabs(xs) < cap ? 1.0 : 0.0, sy = fabs(ys) < cap ? 1.0 : 0.0;
    double f[6];
    f[0] = (ap.x * sx + av.x * acc[2]) + av.y * acc[4];
    f[1] = (ap.y * sy + av.x * acc[3]) + av.y * acc[5];
    f[2] = av.x * xs;
    f[3] = av.x * ys;
    f[4] = av.y * xs;
    f[5] = av.y * ys;
    for (int q = 0; q < 6; ++q) a.w[q * cp + i] = f[q];
    Dual<8> ri[8];
    load_shift_row<8>(s, L, i, ri, true);
    Dual<8> sum(0.0);
    for (int k = 0; k < c; ++k) {
        const int j = nbr[(size_t)k * ld + i];
        Dual<8> nj[5];
        load_shift_nbr<8>(s, j, nj, false);
        Dual<8> t[6];
        shift_pair<Dual<8>>(kc, cf, ri, nj, s.mass[j], t);
        for (int q = 0; q < 6; ++q) sum += t[q] * f[q];
    }
    add2(a.pos, i, sum.d[0], sum.d[1]);
    add2(a.vel, i, sum.d[2], sum.d[3]);
    const double lg[4] = {sum.d[4], sum.d[5], sum.d[6], sum.d[7]};
    add_L(a.L, i, lg);
}

__global__ void __launch_bounds__(BLOCK) k_adj_shift_nbr(int n, KernelConsts kc, double csh,
                                                         size_t ld, StateView s,
                                                         const int* __restrict__ nbr,
                                                         const int* __restrict__ cnt,
                                                         const double* __restrict__ L, AdjView a) {
    int j = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (j >= n) return;
    const size_t cp = a.cap;
    Dual<5> nj[5];
    load_shift_nbr<5>(s, j, nj, true);
    const double mj = s.mass[j];
    Dual<5> sum(0.0);
    const int c = cnt[j];
    for (int k = 0; k < c; ++k) {
        const int i = nbr[(size_t)k * ld + j];
        if (!shift_row(s, cnt, i)) continue;
        Dual<5> ri[8];
        load_shift_row<5>(s, L, i, ri, false);
        Dual<5> t[6];
        shift_pair<Dual<5>>(kc, shift_cf(kc, csh, s.flags[i]), ri, nj, mj, t);
        for (int q = 0; q < 6; ++q) sum += t[q] * a.w[q * cp + i];
    }
    add2(a.pos, j, sum.d[0], sum.d[1]);
    add2(a.vel, j, sum.d[2], sum.d[3]);
    a.rho[j] = a.rho[j] + sum.d[4];
}

// ===========================================================================
// timestep (integrator.f90:937-947): umax = max_i vel_i; the adjoint goes
// to the particle(s) attaining the maximum
// ===========================================================================
__global__ void k_adj_umax(int n, const double* __restrict__ velm, const SimScalars* ss_end,
                           const double* __restrict__ aumax, AdjView a) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const double um = ss_end->umax;
    if (um > 0.0 && fabs(velm[i]) == um) a.velm[i] = a.velm[i] + aumax[0];
}

// ===========================================================================
// objective: probe pressure (Shepard, probe.f90:99-120) and water depth at
// gauges (a 1-D Wendland C2 average of the fluid volume, support 2h)
// ===========================================================================
constexpr int OBJ_MAXP = 8;

struct ObjConst {
    KernelConsts kc;
    int np, ng;
    double px[OBJ_MAXP], py[OBJ_MAXP], gx[OBJ_MAXP];
};

template <class T>
__device__ inline T depth_w(const KernelConsts& kc, const T& dx) {
    T q = fabs(dx) / kc.h;
    if (q >= 2.0) return T(0.0);
    T u = 1.0 - q * 0.5;
    return ((5.0 / (8.0 * kc.h)) * ((u * u) * u)) * (1.5 * q + 1.0);
}

// per particle: for every probe (num, den) and every gauge the volume term
template <class T>
__device__ inline void obj_terms(const ObjConst& oc, const T& x, const T& y, const T& rho,
                                 const T& p, double m, T* out) {
    for (int k = 0; k < oc.np; ++k) {
        T dx = x - oc.px[k], dy = y - oc.py[k];
        T r = sqrt(dx * dx + dy * dy);
        T num(0.0), den(0.0);
        if (r < oc.kc.twoh) {
            T w = (m * kw(oc.kc, r)) / rho;
            num = p * w;
            den = w;
        }
        out[2 * k] = num;
        out[2 * k + 1] = den;
    }
    for (int k = 0; k < oc.ng; ++k) out[2 * oc.np + k] = (m / rho) * depth_w(oc.kc, x - oc.gx[k]);
}

// block partials of the objective's sums: part[m * nb + block]
__global__ void k_obj_sums(int n, ObjConst oc, StateView s, double* __restrict__ part) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    const int M = 2 * oc.np + oc.ng;
    double t[3 * OBJ_MAXP];
    for (int m = 0; m < M; ++m) t[m] = 0.0;
    if (i < n && s.tid[i] == 3 && !(s.flags[i] & ISPH_FLAG_DEAD)) {
        const double2 q = s.pos[i];
        obj_terms<double>(oc, q.x, q.y, s.rho[i], s.p[i], s.mass[i], t);
    }
    for (int m = 0; m < M; ++m) {
        double v = block_sum(t[m]);
        if (threadIdx.x == 0) part[(size_t)m * gridDim.x + blockIdx.x] = v;
    }
}

// sums[m] = sum over nb partials, fixed order (one block per m)
__global__ void k_sum_parts(int nb, const double* __restrict__ part, double* __restrict__ sums) {
    const int m = blockIdx.x;
    double v = 0.0;
    for (int k = threadIdx.x; k < nb; k += BLOCK) v = v + part[(size_t)m * nb + k];
    v = block_sum(v);
    if (threadIdx.x == 0) sums[m] = v;
}

// J_step = sum_k num_k / den_k + sum_g depth_g, accumulated times weight
__global__ void k_obj_accum(ObjConst oc, const double* __restrict__ sums, double weight,
                            double* __restrict__ J) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double j = 0.0;
    for (int k = 0; k < oc.np; ++k) {
        const double den = sums[2 * k + 1];
        j = j + (den > 0.0 ? sums[2 * k] / den : 0.0);
    }
    for (int k = 0; k < oc.ng; ++k) j = j + sums[2 * oc.np + k];
    J[0] = J[0] + weight * j;
}

// adjoint of weight * J_step at the state s (seeds x, y, rho, p)
__global__ void k_adj_obj(int n, ObjConst oc, StateView s, const double* __restrict__ sums,
                          double weight, AdjView a) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    if (s.tid[i] != 3 || (s.flags[i] & ISPH_FLAG_DEAD)) return;
    double cw[3 * OBJ_MAXP];
    for (int k = 0; k < oc.np; ++k) {
        const double num = sums[2 * k], den = sums[2 * k + 1];
        cw[2 * k] = den > 0.0 ? weight / den : 0.0;
        cw[2 * k + 1] = den > 0.0 ? -(weight * num) / (den * den) : 0.0;
    }
    for (int k = 0; k < oc.ng; ++k) cw[2 * oc.np + k] = weight;
    const double2 q = s.pos[i];
    Dual<4> t[3 * OBJ_MAXP];
    obj_terms<Dual<4>>(oc, seed<4>(q.x, 0), seed<4>(q.y, 1), seed<4>(s.rho[i], 2),
                       seed<4>(s.p[i], 3), s.mass[i], t);
    double g[4] = {0.0, 0.0, 0.0, 0.0};
    const int M = 2 * oc.np + oc.ng;
    for (int m = 0; m < M; ++m)
        for (int q4 = 0; q4 < 4; ++q4) g[q4] = g[q4] + cw[m] * t[m].d[q4];
    add2(a.pos, i, g[0], g[1]);
    a.rho[i] = a.rho[i] + g[2];
    a.p[i] = a.p[i] + g[3];
}

// ===========================================================================
// re-binning (integrator.f90:32-391): dst[i] = src[perm[i]]; adjoint
// scatters back, src_bar[perm[i]] = dst_bar[i]
// ===========================================================================
__global__ void k_adj_unperm(int n, const int* __restrict__ perm, AdjView d, AdjView s) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const int o = perm[i];
    s.pos[o] = d.pos[i];
    s.vel[o] = d.vel[i];
    s.sp[o] = d.sp[i];
    s.sv[o] = d.sv[i];
    s.por[o] = d.por[i];
    s.rho[o] = d.rho[i];
    s.R[o] = d.R[i];
    s.nut[o] = d.nut[i];
    s.tke[o] = d.tke[i];
    s.varts[o] = d.varts[i];
    s.p[o] = d.p[i];
    s.velm[o] = d.velm[i];
    const double2* dl = reinterpret_cast<const double2*>(d.L);
    double2* sl = reinterpret_cast<double2*>(s.L);
    sl[2 * (size_t)o] = dl[2 * (size_t)i];
    sl[2 * (size_t)o + 1] = dl[2 * (size_t)i + 1];
}

// ===========================================================================
// transposed PPE values: valT[k ld + i] = A_{j i} with j = nbr[k ld + i]
// (the lists are symmetric and ascend in particle index)
// ===========================================================================
__global__ void k_transpose_vals(int n, size_t ld, const int* __restrict__ nbr,
                                 const int* __restrict__ cnt, const double* __restrict__ val,
                                 double* __restrict__ valT, uint32_t* err) {
    int i = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (i >= n) return;
    const int c = cnt[i];
    for (int k = 0; k < c; ++k) {
        const int j = nbr[(size_t)k * ld + i];
        int lo = 0, hi = cnt[j] - 1, m = -1;
        while (lo <= hi) {
            const int mid = (lo + hi) >> 1;
            const int v = nbr[(size_t)mid * ld + j];
            if (v == i) { m = mid; break; }
            if (v < i) lo = mid + 1; else hi = mid - 1;
        }
        double t = 0.0;
        if (m >= 0) t = val[(size_t)m * ld + j];
        else atomicOr(err, ERR_CAPACITY);
        valT[(size_t)k * ld + i] = t;
    }
}

// sum of `count` doubles in block partials (one value per block)
__global__ void k_sum_blocks(int64_t count, const double* __restrict__ x, double* __restrict__ part) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    double v = i < count ? x[i] : 0.0;
    v = block_sum(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ void k_zero(int64_t count, double* __restrict__ x) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < count) x[i] = 0.0;
}

// lambda by particle id <-> lambda by row (ids are unique; an id outside
// [0, npid) has no stored value and starts from zero)
__global__ void k_lam_gather(int n, const int* __restrict__ pid, const double* __restrict__ by_pid,
                             int npid, double* __restrict__ lam) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const int q = pid[i];
    lam[i] = (q >= 0 && q < npid) ? by_pid[q] : 0.0;
}

__global__ void k_lam_scatter(int n, const int* __restrict__ pid, const double* __restrict__ lam,
                              double* __restrict__ by_pid, int npid) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const int q = pid[i];
    if (q >= 0 && q < npid) by_pid[q] = lam[i];
}

}  // namespace adj
}  // namespace isph
