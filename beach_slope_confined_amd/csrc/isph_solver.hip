// isph_solver.hip -- BiCGSTAB for the Jacobi-scaled PPE on gfx950.
//
// Recurrences of the reference's bicgstab (solver.f90:162-213, van der
// Vorst 1992).  One iteration is three row kernels, each followed by a
// one-block reduction kernel (k_finish) that turns the block partials into
// the iteration's scalars and stop flags:
//   k_krylov_1  p = r + beta q (or p = r after a (re)start), v = A p, (v, rhat)
//   k_krylov_2  s = r - alpha v, t = A s, (t, s), (t, t)   (rows in reverse order)
//   k_krylov_3  x += alpha p + omega s, r = s - omega t, q = p - omega v,
//               (r, rhat), (r, r)
// q carries p - omega v from one iteration to the next so that the p update
// needs (r_j, q_j) instead of three vectors, with the same rounding as
// r + beta (p - omega v); (r, q) and (r, v) are stored as pairs so that each
// neighbour is one gather.  Partials are plain stores; the kernel boundary
// makes them visible, so no fences or global atomics are needed.
//
// Flavours (template TV matrix value, TX vector, TI column index):
//   <double, double, short>  FP64 drop-in (int when an offset does not fit 16 bits)
//   <float,  float,  short>  mixed: FP32 values and FP32 Krylov vectors
//   <float,  double, short>  mixed: FP32 values, FP64 Krylov vectors,
//                            conservation-form SpMV with FP64 row sums
// short indices are offsets j - i (built only when every offset fits).  The
// mixed flavours add reliable updates: when the recurrence residual has
// dropped by `delta`, x is flushed into the FP64 solution and the FP64 true
// residual b - A64 x replaces the recurrence residual (search direction
// kept).  Dot products always accumulate in FP64.
#include <type_traits>

#include "isph_kernels.h"
#include "isph_spmv.h"

namespace isph {

// Conservation form for FP32 values with FP64 vectors (MIXED_V64):
// (A x)_i = s_i x_i + sum_j a_ij (x_j - x_i) with the FP64 row sum s_i.  The
// FP32 rounding of a_ij then multiplies x_j - x_i, which is small for the
// smooth modes whose eigenvalues are small, instead of x_j itself.
template <class TV, class TX, class TI, class F>
__device__ __forceinline__ TX spmv_row_diff(int i, int c, double si, const TI* __restrict__ nbr,
                                            const TV* __restrict__ val, size_t ld, TX xi, F&& xj) {
    TX acc = (TX)0;
    int k = 0;
    for (; k + 4 <= c; k += 4) {
        const size_t o = (size_t)k * ld + i;
        const int j0 = col_of(nbr[o], i), j1 = col_of(nbr[o + ld], i);
        const int j2 = col_of(nbr[o + 2 * ld], i), j3 = col_of(nbr[o + 3 * ld], i);
        const TX v0 = (TX)val[o], v1 = (TX)val[o + ld], v2 = (TX)val[o + 2 * ld],
                 v3 = (TX)val[o + 3 * ld];
        const TX x0 = xj(j0), x1 = xj(j1), x2 = xj(j2), x3 = xj(j3);
        acc = acc + v0 * (x0 - xi);
        acc = acc + v1 * (x1 - xi);
        acc = acc + v2 * (x2 - xi);
        acc = acc + v3 * (x3 - xi);
    }
    for (; k < c; ++k) {
        const size_t o = (size_t)k * ld + i;
        const int j = col_of(nbr[o], i);
        const TX v = (TX)val[o];
        acc = acc + v * (xj(j) - xi);
    }
    return (TX)si * xi + acc;
}

template <class TV, class TX, class TI, class F>
__device__ __forceinline__ TX spmv_any(const KrylovArgs<TV, TX, TI>& a, int i, TX xi, F&& xj) {
    if constexpr (std::is_same<TV, float>::value && std::is_same<TX, double>::value)
        return spmv_row_diff<TV, TX, TI>(i, a.cnt[i], a.rsum[i], a.nbr, a.val, a.ld, xi, xj);
    else
        return spmv_row<TV, TX, TI>(i, a.cnt[i], a.dpos[i], a.nbr, a.val, a.ld, xi, xj);
}

template <class TV, class TX, class TI, bool FIRST>
__global__ void __launch_bounds__(BLOCK) k_krylov_1t(KrylovArgs<TV, TX, TI> a,
                                                     const SolverScalars* sc) {
    using P2 = typename Pair2<TX>::type;
    // the done flag and beta are loaded together, and the row's own pair is
    // requested before the done test waits on them
    const int done = sc->done;
    const TX beta = (TX)sc->betas;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int i = lb * BLOCK + threadIdx.x;
    const P2* __restrict__ rq = a.rq;
    P2 ui{};
    if (i < a.n) ui = rq[i];
    if (done) return;
    auto pn = [&](int j) -> TX {
        const P2 u = rq[j];
        return FIRST ? u.x : u.x + beta * u.y;
    };
    double v = 0.0;
    if (i < a.n) {
        const TX pi = FIRST ? ui.x : ui.x + beta * ui.y;
        const TX s = spmv_any(a, i, pi, pn);
        a.p[i] = pi;
        P2 w;
        w.x = ui.x;
        w.y = s;
        a.rv[i] = w;
        v = (double)s * (double)a.rhat[i];
    }
    v = block_sum(v);
    if (threadIdx.x == 0) a.part[lb] = v;
}

template <class TV, class TX, class TI>
__global__ void __launch_bounds__(BLOCK) k_krylov_2(KrylovArgs<TV, TX, TI> a,
                                                    const SolverScalars* sc) {
    using P2 = typename Pair2<TX>::type;
    // loads ahead of the done test, as in k_krylov_1t
    const int done = sc->done;
    const TX alpha = (TX)sc->alpha;
    const int lb = xcd_block_rev(blockIdx.x, gridDim.x);  // opposite order to k_krylov_1t
    const int i = lb * BLOCK + threadIdx.x;
    const P2* __restrict__ rv = a.rv;
    P2 ui{};
    if (i < a.n) ui = rv[i];
    if (done) return;
    auto sf = [&](int j) -> TX {
        const P2 u = rv[j];
        return u.x - alpha * u.y;
    };
    double v0 = 0.0, v1 = 0.0;
    if (i < a.n) {
        const TX si = ui.x - alpha * ui.y;
        const TX s = spmv_any(a, i, si, sf);
        a.si[i] = si;
        a.mm2[i] = s;
        v0 = (double)s * (double)si;
        v1 = (double)s * (double)s;
    }
    v0 = block_sum(v0);
    v1 = block_sum(v1);
    if (threadIdx.x == 0) {
        a.part[lb] = v0;
        a.part[gridDim.x + lb] = v1;
    }
}

template <class TV, class TX, class TI>
__global__ void __launch_bounds__(BLOCK) k_krylov_3(KrylovArgs<TV, TX, TI> a,
                                                    const SolverScalars* sc) {
    using P2 = typename Pair2<TX>::type;
    // loads ahead of the done test, as in k_krylov_1t
    const int done = sc->done;
    const TX alpha = (TX)sc->alpha, w = (TX)sc->wi;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int i = lb * BLOCK + threadIdx.x;
    TX si = (TX)0, p = (TX)0;
    if (i < a.n) {
        si = a.si[i];
        p = a.p[i];
    }
    if (done) return;
    double v0 = 0.0, v1 = 0.0;
    if (i < a.n) {
        const TX mm1 = a.rv[i].y;
        const TX x = (a.xs[i] + alpha * p) + w * si;
        const TX r = si - w * a.mm2[i];
        const TX qq = p - w * mm1;
        a.xs[i] = x;
        P2 u;
        u.x = r;
        u.y = qq;
        a.rq[i] = u;
        v0 = (double)r * (double)a.rhat[i];
        v1 = (double)r * (double)r;
    }
    v0 = block_sum(v0);
    v1 = block_sum(v1);
    if (threadIdx.x == 0) {
        a.part[lb] = v0;
        a.part[gridDim.x + lb] = v1;
    }
}

// One block of FIN_BLOCK threads: finish the reductions of the preceding row
// kernel and update the solver scalars.  Thread t sums partials t,
// t+FIN_BLOCK, ... in order, then the block tree.
__global__ void __launch_bounds__(FIN_BLOCK) k_finish(int kind, int nbp,
                                                      const double* __restrict__ part,
                                                      SolverScalars* sc, SolveCtl ctl) {
    const bool two = kind != FIN_ALPHA && kind != FIN_START && kind != FIN_REPLACE;
    // The partials and the solver state are loaded before anything waits on
    // either, so neither the done test nor the update below adds its load
    // latency to the chain; the update works on this copy and stores it.
    double p0[8], p1[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int k = threadIdx.x + u * FIN_BLOCK;
        p0[u] = k < nbp ? part[k] : 0.0;
        p1[u] = (two && k < nbp) ? part[nbp + k] : 0.0;
    }
    SolverScalars s = *sc;
    if ((kind == FIN_ALPHA || kind == FIN_OMEGA || kind == FIN_BETA) && s.done) return;
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        if (threadIdx.x + u * FIN_BLOCK < nbp) {
            s0 = s0 + p0[u];
            s1 = s1 + p1[u];
        }
    }
    for (int k = threadIdx.x + 8 * FIN_BLOCK; k < nbp; k += FIN_BLOCK) {
        s0 = s0 + part[k];
        if (two) s1 = s1 + part[nbp + k];
    }
    const double t0 = block_sum_fin(s0);
    const double t1 = two ? block_sum_fin(s1) : 0.0;
    if (threadIdx.x != 0) return;
    switch (kind) {
        case FIN_INIT: {
            s.bb = t0;
            s.nb = sqrt(t0);
            s.rr = t1;
            s.prevr = t1;
            s.iters = 0;
            s.total = 0;
            s.first = 1;
            s.status = 0;
            s.nupd = 0;
            s.rh2 = t1;
            s.brk = 0;
            if (t0 == 0.0) {
                s.done = DONE_ZERO_RHS;
                s.relres = 0.0;
            } else if (sqrt(t1) <= ctl.tol * s.nb) {
                s.done = DONE_CONVERGED;
                s.relres = sqrt(t1) / s.nb;
            } else {
                s.done = 0;
            }
            break;
        }
        // breakdown ((rhat, v) = 0, or (t, t) = 0 / (t, s) = 0): the step
        // degenerates to a minimal-residual step (alpha = 0) or to the half
        // step x + alpha p (omega = 0) and ends in a restart
        case FIN_ALPHA: {
            double a = s.prevr / t0;
            if (!isfinite(a)) {
                a = 0.0;
                s.brk = 1;
            }
            s.alpha = a;
            break;
        }
        case FIN_OMEGA: {
            double w = t0 / t1;
            if (!isfinite(w) || w == 0.0) {
                w = 0.0;
                s.brk = 1;
            }
            s.wi = w;
            break;
        }
        case FIN_BETA: {
            const double betas = (t0 / s.prevr) * (s.alpha / s.wi);
            s.betas = betas;
            s.prevr_prev = s.prevr;
            s.prevr = t0;
            s.rr = t1;
            s.iters = s.iters + 1;
            s.total = s.total + 1;
            s.first = 0;
            const double nr = sqrt(t1);
            if (!(fabs(t0) > BRK_COS * sqrt(t1 * s.rh2))) s.brk = 1;
            if (s.brk) {
                // FP64: true residual + restart; mixed: flush + full restart
                // (the flag stays set for the host until FIN_START)
                if (!ctl.mixed) {
                    s.brk = 0;
                    s.done = DONE_RECURRENCE;
                } else {
                    s.done = DONE_UPDATE;
                }
            } else if (!isfinite(betas) || !isfinite(t1)) {
                s.done = DONE_FAILED;
                s.status = ERR_NOCONV;
                s.relres = nr / s.nb;
            } else if (!ctl.mixed && nr <= ctl.tol * s.nb) {
                s.done = DONE_RECURRENCE;
            } else if (ctl.mixed && (nr <= ctl.tol * s.nb || nr <= ctl.delta * s.rupd)) {
                s.done = DONE_UPDATE;
            } else if (s.total >= ctl.maxit) {
                s.done = DONE_FAILED;
                s.status = ERR_NOCONV;
                s.relres = nr / s.nb;
            }
            break;
        }
        case FIN_TRUE: {
            const double nr = sqrt(t1);
            if (nr <= ctl.tol * s.nb) {
                s.done = DONE_CONVERGED;
                s.relres = nr / s.nb;
            } else if (!isfinite(nr)) {
                s.done = DONE_FAILED;
                s.status = ERR_NOCONV;
                s.relres = nr;
            } else {
                s.rr = t1;
                s.prevr = t1;
                s.rh2 = t1;
                s.first = 1;
                s.done = 0;
                if (s.total >= ctl.maxit) {
                    s.done = DONE_FAILED;
                    s.status = ERR_NOCONV;
                    s.relres = nr / s.nb;
                }
            }
            break;
        }
        case FIN_RESID: {
            // mixed outer residual; nupd counts the flushes (ctl.mixed == 2 after a flush)
            s.bb = t0;
            s.nb = sqrt(t0);
            s.nr = sqrt(t1);
            if (ctl.mixed == 2) s.nupd = s.nupd + 1;
            if (t0 == 0.0) {
                s.done = DONE_ZERO_RHS;
                s.relres = 0.0;
                break;
            }
            s.relres = s.nr / s.nb;
            if (s.nr <= ctl.tol * s.nb) {
                s.done = DONE_CONVERGED;
            } else if (ctl.mixed == 2 &&
                       (s.nupd >= ctl.maxupd || s.total >= ctl.maxit || !isfinite(s.nr))) {
                s.done = DONE_FAILED;
                s.status = ERR_NOCONV;
            } else {
                s.done = 0;
            }
            break;
        }
        case FIN_START:
            s.prevr = t0;
            s.prevr_prev = t0;
            s.rh2 = t0;
            s.brk = 0;
            s.rupd = s.nr;
            s.first = 1;
            s.iters = 0;
            s.done = 0;
            break;
        case FIN_REPLACE: {
            s.betas = (t0 / s.prevr_prev) * (s.alpha / s.wi);
            s.prevr = t0;
            s.rupd = s.nr;
            s.done = 0;
            break;
        }
        default:
            break;
    }
    *sc = s;
}

// r = b - A64 x in FP64; partials (b,b), (r,r).  mode 1 also restarts the
// FP64 recurrence from r (rs = rhat = r).
__global__ void __launch_bounds__(BLOCK) k_resid(ResidArgs a, int mode, double2* __restrict__ rq,
                                                 double* __restrict__ rhat) {
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int i = lb * BLOCK + threadIdx.x;
    double vb = 0.0, vr = 0.0;
    if (i < a.n) {
        const double* __restrict__ x = a.x;
        auto xj = [&](int j) { return x[j]; };
        // the same sum in the same order with either index form
        const double s =
            a.nbr16 ? spmv_row<double, double, short>(i, a.cnt[i], a.dpos[i], a.nbr16, a.val, a.ld,
                                                      x[i], xj)
                    : spmv_row<double, double, int>(i, a.cnt[i], a.dpos[i], a.nbr, a.val, a.ld,
                                                    x[i], xj);
        const double bi = a.b[i];
        const double r = bi - s;
        a.r[i] = r;
        if (mode == 1) {
            rq[i] = make_double2(r, 0.0);
            rhat[i] = r;
        }
        vb = bi * bi;
        vr = r * r;
    }
    vb = block_sum(vb);
    vr = block_sum(vr);
    if (threadIdx.x == 0) {
        a.part[lb] = vb;
        a.part[gridDim.x + lb] = vr;
    }
}

template <class TX>
__global__ void __launch_bounds__(BLOCK) k_mx_start(int n, const double* __restrict__ r64,
                                                    typename Pair2<TX>::type* __restrict__ rq,
                                                    TX* __restrict__ rhat, TX* __restrict__ xs,
                                                    double* __restrict__ part,
                                                    const SolverScalars* sc) {
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int i = lb * BLOCK + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        const TX r = (TX)r64[i];
        typename Pair2<TX>::type u;
        u.x = r;
        u.y = (TX)0;
        rq[i] = u;
        rhat[i] = r;
        xs[i] = (TX)0;
        v = (double)r * (double)r;
    }
    v = block_sum(v);
    if (threadIdx.x == 0) part[lb] = v;
}

template <class TX>
__global__ void __launch_bounds__(BLOCK) k_mx_replace(int n, const double* __restrict__ r64,
                                                      typename Pair2<TX>::type* __restrict__ rq,
                                                      const TX* __restrict__ rhat,
                                                      double* __restrict__ part,
                                                      const SolverScalars* sc) {
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int i = lb * BLOCK + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        const TX r = (TX)r64[i];
        rq[i].x = r;
        v = (double)r * (double)rhat[i];
    }
    v = block_sum(v);
    if (threadIdx.x == 0) part[lb] = v;
}

template <class TX>
__global__ void k_mx_flush(int n, double* __restrict__ x64, TX* __restrict__ xs) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    x64[i] = x64[i] + (double)xs[i];
    xs[i] = (TX)0;
}

// 16-bit column offsets j - i for the mixed-precision SpMV
__global__ void k_to_short(int n, size_t ld, const int* __restrict__ nbr,
                           const int* __restrict__ cnt, short* __restrict__ out, uint32_t* err) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const int c = cnt[i];
    bool bad = false;
    for (int k = 0; k < c; ++k) {
        const int d = nbr[(size_t)k * ld + i] - i;
        if (d < -32768 || d > 32767) bad = true;
        out[(size_t)k * ld + i] = (short)d;
    }
    if (bad) atomicOr(err, ERR_INDEX16);
}

// instantiations
template __global__ void k_krylov_1t<double, double, int, true>(KrylovArgs<double, double, int>,
                                                               const SolverScalars*);
template __global__ void k_krylov_1t<double, double, int, false>(KrylovArgs<double, double, int>,
                                                                const SolverScalars*);
template __global__ void k_krylov_2<double, double, int>(KrylovArgs<double, double, int>,
                                                         const SolverScalars*);
template __global__ void k_krylov_3<double, double, int>(KrylovArgs<double, double, int>,
                                                         const SolverScalars*);
template __global__ void k_krylov_1t<double, double, short, true>(KrylovArgs<double, double, short>,
                                                                 const SolverScalars*);
template __global__ void k_krylov_1t<double, double, short, false>(KrylovArgs<double, double, short>,
                                                                  const SolverScalars*);
template __global__ void k_krylov_2<double, double, short>(KrylovArgs<double, double, short>,
                                                           const SolverScalars*);
template __global__ void k_krylov_3<double, double, short>(KrylovArgs<double, double, short>,
                                                           const SolverScalars*);
#define ISPH_MIXED_FLAVOUR(TX, TI)                                                              \
    template __global__ void k_krylov_1t<float, TX, TI, true>(KrylovArgs<float, TX, TI>,        \
                                                              const SolverScalars*);           \
    template __global__ void k_krylov_1t<float, TX, TI, false>(KrylovArgs<float, TX, TI>,       \
                                                               const SolverScalars*);          \
    template __global__ void k_krylov_2<float, TX, TI>(KrylovArgs<float, TX, TI>,               \
                                                       const SolverScalars*);                  \
    template __global__ void k_krylov_3<float, TX, TI>(KrylovArgs<float, TX, TI>,               \
                                                       const SolverScalars*);
ISPH_MIXED_FLAVOUR(float, short)
ISPH_MIXED_FLAVOUR(float, int)
ISPH_MIXED_FLAVOUR(double, short)
ISPH_MIXED_FLAVOUR(double, int)
template __global__ void k_mx_start<float>(int, const double*, float2*, float*, float*, double*,
                                           const SolverScalars*);
template __global__ void k_mx_start<double>(int, const double*, double2*, double*, double*,
                                            double*, const SolverScalars*);
template __global__ void k_mx_replace<float>(int, const double*, float2*, const float*, double*,
                                             const SolverScalars*);
template __global__ void k_mx_replace<double>(int, const double*, double2*, const double*, double*,
                                              const SolverScalars*);
template __global__ void k_mx_flush<float>(int, double*, float*);
template __global__ void k_mx_flush<double>(int, double*, double*);

}  // namespace isph
