// isph_scalar.hip -- salinity transport and variable density (scalar.f90),
// the sea-water part of the reference's step (main.f90:199-263).
//
// con in [-0.5, 0.5] is the salinity of a fluid particle.  scalart computes
// its rate of change cdiff by a pairwise anisotropic diffusion over the
// stored neighbour lists; scalarupdate integrates it over one substep;
// densityupdate maps con linearly to the fluid density.
//
// Layout: con and the particle volume ovol travel with the particle (they
// are members of StateView and follow every re-binning); cdiff and the
// per-particle coefficients (Dx, Dy, Dm + nut/ts, nut) are scratch in the
// current order.  The coefficients depend only on the particle's own
// velocity, porosity, eddy viscosity and dt, none of which change during
// the transport substeps, so k_scalar_coef evaluates them once per step, in
// one pass over N, instead of once per pair as scalar.f90:160-192 does; the
// values are the same expressions and the same bits.
//
// scalart reads the lists as the reference stores them: pair distances r
// from the positions of the last neighbour search (pos_nb) and position
// differences from the current positions (scalar.f90:224-229 mixes the
// stored dist(m) with the current x, which the shift has moved).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "isph_host.h"
#include "isph_kernels.h"

namespace isph {

// scalar.f90:118 (al, at, tor are double literals; Dm is DM_SALT)
constexpr double AL_SALT = 0.0025;
constexpr double AT_SALT = 0.00025;
constexpr double TOR_SALT = 1.0;

__device__ __forceinline__ bool salt_live(const StateView& s, int i) {
    return s.tid[i] == 3 && !(s.flags[i] & ISPH_FLAG_DEAD);
}

// Dx, Dy of scalar.f90:138-154: Dm + (al vx^2 + at vy^2 + (al-at) vx vy)/(|v| n)
// and its mirror; Dm where that is not positive (including 0/0 at v = 0),
// at most 0.12 dl^2/dt.
__device__ inline double2 salt_dispersion(double vx, double vy, double n, double cap) {
    const double v = sqrt(vx * vx + vy * vy);
    const double cross = ((AL_SALT - AT_SALT) * vx) * vy;
    double dx = DM_SALT + ((AL_SALT * (vx * vx) + AT_SALT * (vy * vy)) + cross) / (v * n);
    double dy = DM_SALT + ((AT_SALT * (vx * vx) + AL_SALT * (vy * vy)) + cross) / (v * n);
    dx = (dx > 0.0) ? dx : DM_SALT;
    dx = (dx <= cap) ? dx : cap;
    dy = (dy > 0.0) ? dy : DM_SALT;
    dy = (dy <= cap) ? dy : cap;
    return make_double2(dx, dy);
}

__global__ void k_scalar_coef(int n, SaltConsts sc, StateView s, const SimScalars* ss,
                              double4* __restrict__ coef) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n || !salt_live(s, i)) return;
    const double cap = (0.12 * (sc.dl * sc.dl)) / ss->dt;
    const double2 v = s.vel[i];
    const double2 d = salt_dispersion(v.x, v.y, s.por[i], cap);
    const double nut = s.nut[i];
    coef[i] = make_double4(d.x, d.y, DM_SALT + nut / sc.tschmidt, nut);
}

// scalart, scalar.f90:114-236: one thread per fluid particle, pull over its
// list; only live fluid neighbours contribute (:158).
__global__ void k_scalart(int n, KernelConsts kh, SaltConsts sc, size_t ld, StateView s,
                          const double2* __restrict__ pos_nb, const int* __restrict__ nbr,
                          const int* __restrict__ cnt, const double4* __restrict__ coef,
                          double* __restrict__ cdiff) {
    int i = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (i >= n || !salt_live(s, i)) return;
    const double con1 = s.con[i];
    const double4 di = coef[i];
    const double ni = s.por[i], oi = s.ovol[i];
    const double ai = di.z, nuti = di.w;
    const double2 pi = s.pos[i], qi = pos_nb[i];
    double acc = 0.0;
    const int c = cnt[i];
    for (int k = 0; k < c; ++k) {
        const int j = nbr[(size_t)k * ld + i];
        if (!salt_live(s, j)) continue;
        const double nj = s.por[j];
        const double4 dj = coef[j];
        double t1x, t1y;
        if (ni >= 0.8 || nj >= 0.8) {  // open water: molecular + turbulent (:184-196)
            if (nuti < TINY6 && dj.w < TINY6) {
                t1x = DM_SALT + ((nuti + dj.w) * 0.50) / sc.tschmidt;
            } else {
                t1x = 2.0 * ((ai * dj.z) / (ai + dj.z));
            }
            t1y = t1x;
        } else {  // in the pores: harmonic mean of the dispersion (:200-201)
            t1x = (((2.0 * TOR_SALT) * di.x) * dj.x) / (di.x + dj.x);
            t1y = (((2.0 * TOR_SALT) * di.y) * dj.y) / (di.y + dj.y);
        }
        const double t3 = (oi * (ni + nj)) / (ni * nj);
        const double2 qj = pos_nb[j];
        const double ex = qi.x - qj.x, ey = qi.y - qj.y;
        const double r = sqrt(ex * ex + ey * ey);
        const double2 pj = s.pos[j];
        const double dx = pi.x - pj.x, dy = pi.y - pj.y;
        double hx, hy;
        kern_gradw(kh, r, dx, dy, hx, hy);  // hWabx/hWaby, kernel.f90:204-248
        const double dc = con1 - s.con[j];
        const double den = r * r + sc.lam;
        const double tx = ((((t1x * t3) * dc) * dx) * hx) / den;  // t2 = 1 (:206)
        const double ty = ((((t1y * t3) * dc) * dy) * hy) / den;
        acc = (acc + tx) + ty;
    }
    cdiff[i] = acc;
}

// scalarupdate, scalar.f90:238-270 (stime = ss->dtsol when ss is given)
__global__ void k_scalarupdate(int n, SaltConsts sc, StateView s, const double* __restrict__ cdiff,
                               double stime, const SimScalars* ss) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n || !salt_live(s, i)) return;
    const double st = ss ? ss->dtsol : stime;
    double c = s.con[i] + cdiff[i] * st;
    if ((s.flags[i] & ISPH_FLAG_BUFFER) && s.pos[i].x < sc.xbuf) c = 0.5;
    s.con[i] = c;
}

// densityupdate, scalar.f90:272-305
__global__ void k_densityupdate(int n, SaltConsts sc, StateView s) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n || !salt_live(s, i)) return;
    const double od = 0.5 * (sc.rhomax + sc.rhomin) + s.con[i] * (sc.rhomax - sc.rhomin);
    s.oden[i] = od;
    s.mass[i] = od * s.ovol[i];
    s.rho[i] = od * s.por[i];
}

// salt-water start, main.f90:160-195 (every fluid particle, dead or not)
__global__ void k_salt_init(int n, StateView s, double xl, double yl, double line_grad,
                            double rel_den) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n || s.tid[i] != 3) return;
    const double2 p = s.pos[i];
    double c = -0.50;
    if (((p.y - yl) - line_grad * (p.x - xl)) > 0.0) {
        s.mass[i] = s.mass[i] * rel_den;
        s.oden[i] = s.oden[i] * rel_den;
        s.rho[i] = s.rho[i] * rel_den;
        c = 0.50;
    }
    s.con[i] = c;
}

namespace host {

SaltConsts salt_consts(const isph_ctx* c) {
    const isph_params& p = c->prm;
    SaltConsts sc;
    sc.lam = p.lam;
    sc.dl = p.dl;
    sc.tschmidt = p.tschmidt;
    sc.rhomax = p.rho0 * p.rel_den;  // setup.f90:19
    sc.rhomin = p.rhomin;
    sc.xbuf = p.salt_xbuf;
    return sc;
}

// hWabx/hWaby are always the hybrid kernel, whatever the run's kernel
static KernelConsts hybrid_kc(const isph_ctx* c) {
    KernelConsts k = make_kc(c->prm);
    k.wendland = 0;
    return k;
}

int run_scalar_coef(isph_ctx* c) {
    if (c->n)
        k_scalar_coef<<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, salt_consts(c), c->st[c->cur],
                                                            c->ss, c->salt_coef);
    HIPCHK(c, hipGetLastError());
    return ISPH_OK;
}

// cdiff from the coefficients of the last run_scalar_coef
int run_scalart_only(isph_ctx* c) {
    if (!c->have_nbr) return fail(c, ISPH_E_STATE, "scalart before neighbour");
    if (c->n) {
        prof_record(c, PROF_SCALART, true);
        k_scalart<<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, hybrid_kc(c), salt_consts(c),
                                                        (size_t)c->cap, c->st[c->cur], c->pos_nb,
                                                        c->nbr, c->cnt, c->salt_coef, c->cdiff);
        prof_record(c, PROF_SCALART, false);
    }
    HIPCHK(c, hipGetLastError());
    return ISPH_OK;
}

int run_scalart(isph_ctx* c) {
    if (!c->have_nbr) return fail(c, ISPH_E_STATE, "scalart before neighbour");
    int rc = run_scalar_coef(c);
    if (rc) return rc;
    return run_scalart_only(c);
}

int run_scalarupdate(isph_ctx* c, double stime, bool from_ss) {
    if (c->n)
        k_scalarupdate<<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, salt_consts(c), c->st[c->cur],
                                                             c->cdiff, stime,
                                                             from_ss ? c->ss : nullptr);
    HIPCHK(c, hipGetLastError());
    return ISPH_OK;
}

int run_densityupdate(isph_ctx* c) {
    if (c->n)
        k_densityupdate<<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, salt_consts(c), c->st[c->cur]);
    HIPCHK(c, hipGetLastError());
    return ISPH_OK;
}

// The transport substeps at the start of a salt-water step (main.f90:204-215):
// solsteps x (scalart, scalarupdate(dtsol)) as the last timestep set them.
// Without lists from a previous step there is nothing to transport with.
int run_salt_substeps(isph_ctx* c) {
    if (!c->have_nbr || !c->n) return ISPH_OK;
    HIPCHK(c, hipMemcpyAsync(c->ss_host, c->ss, sizeof(SimScalars), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int steps = c->ss_host->solsteps;
    if (steps < 0 || steps > MAX_SUBSTEPS)
        return fail(c, ISPH_E_CAPACITY, "salinity transport needs too many substeps");
    if (steps == 0) return ISPH_OK;
    int rc = run_scalar_coef(c);
    if (rc) return rc;
    for (int k = 0; k < steps; ++k) {
        if ((rc = run_scalart_only(c))) return rc;
        if ((rc = run_scalarupdate(c, 0.0, true))) return rc;
    }
    return ISPH_OK;
}

}  // namespace host
}  // namespace isph

using namespace isph;
using namespace isph::host;

extern "C" {

int isph_scalart(isph_ctx* c) {
    if (!c) return ISPH_E_ARG;
    int rc = run_scalart(c);
    prof_collect(c);
    if (rc) return rc;
    return sync_check(c);
}

int isph_scalarupdate(isph_ctx* c, double stime) {
    if (!c) return ISPH_E_ARG;
    int rc = run_scalarupdate(c, stime, false);
    if (rc) return rc;
    return sync_check(c);
}

int isph_densityupdate(isph_ctx* c) {
    if (!c) return ISPH_E_ARG;
    int rc = run_densityupdate(c);
    if (rc) return rc;
    return sync_check(c);
}

int isph_salt_init(isph_ctx* c, double xl, double yl, double line_grad) {
    if (!c) return ISPH_E_ARG;
    if (c->n)
        k_salt_init<<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, c->st[c->cur], xl, yl, line_grad,
                                                          c->prm.rel_den);
    HIPCHK(c, hipGetLastError());
    return sync_check(c);
}

int isph_get_substeps(isph_ctx* c, int32_t* solsteps, double* dtsol) {
    if (!c) return ISPH_E_ARG;
    HIPCHK(c, hipMemcpyAsync(c->ss_host, c->ss, sizeof(SimScalars), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (solsteps) *solsteps = c->ss_host->solsteps;
    if (dtsol) *dtsol = c->ss_host->dtsol;
    return ISPH_OK;
}

}  // extern "C"
