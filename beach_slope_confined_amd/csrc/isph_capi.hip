// isph_capi.hip -- host side of the C ABI declared in include/isph_capi.h.
//
// A context owns every device buffer (allocated once at isph_create, sized
// for cap_particles, like memory.f90:10-147 sizes the reference's cells)
// and one HIP stream of default priority.  Each entry point enqueues its
// kernels on that stream and synchronises before returning; isph_step()
// runs whole steps and synchronises only inside the PPE solve (to poll the
// device-side convergence flag) and at the end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/isph_capi.h"
#include "isph_kernels.h"

using namespace isph;

#include "isph_host.h"

namespace isph {
namespace host {

int fail(isph_ctx* c, int code, const char* m) {
    c->msg = m;
    return code;
}

GridInfo make_grid(const isph_params& p) {
    GridInfo g;
    g.xmin = p.xmin;
    g.ymin = p.ymin;
    g.cs = p.cell_size;
    g.nx = (int)std::floor((p.xmax - p.xmin) / p.cell_size) + 3;
    g.ny = (int)std::floor((p.ymax - p.ymin) / p.cell_size) + 3;
    g.order = p.cell_order;
    g.per = p.cell_order ? 10 : 2;
    g.nbins = (long long)g.per * g.nx * g.ny;
    return g;
}

bool prec_mixed(int precision) {
    return precision == ISPH_PREC_MIXED || precision == ISPH_PREC_MIXED_V64 ||
           precision == ISPH_PREC_MIXED_MG || precision == ISPH_PREC_IR ||
           precision == ISPH_PREC_FP64_MG;
}

bool prec_mg(int precision) {
    return precision == ISPH_PREC_MIXED_MG || precision == ISPH_PREC_FP64_MG;
}

int check_params(isph_ctx* c, const isph_params* p) {
    if (!(p->h > 0.0) || !(p->cell_size >= 2.0 * p->h) || !(p->xmax > p->xmin) ||
        !(p->ymax > p->ymin) || p->max_nbr <= 0)
        return fail(c, ISPH_E_ARG, "invalid geometry parameters");
    if (p->cell_order != ISPH_ORDER_CELL && p->cell_order != ISPH_ORDER_REFERENCE)
        return fail(c, ISPH_E_ARG, "unknown cell_order");
    if (p->precision < ISPH_PREC_FP64 || p->precision > ISPH_PREC_FP64_MG)
        return fail(c, ISPH_E_ARG, "unknown precision");
    if (!(p->inner_tol >= 0.0 && p->inner_tol < 1.0))
        return fail(c, ISPH_E_ARG, "inner_tol must lie in [0, 1)");
    if (p->cell_order == ISPH_ORDER_REFERENCE && !(p->cell_size >= 4.0 * p->h))
        return fail(c, ISPH_E_ARG, "the reference order needs cell_size >= 4h (2x2-cell blocks)");
    GridInfo g = make_grid(*p);
    if ((double)g.nx * (double)g.ny * g.per > 2.0e9) return fail(c, ISPH_E_ARG, "cell grid too large");
    return ISPH_OK;
}

PorousParams porous_params(const isph_params& p) {
    PorousParams pp;
    pp.por = p.por;
    pp.alpha = p.alpha;
    pp.beta = p.beta;
    pp.d50 = p.d50;
    pp.nu = p.nu;
    pp.drag_law = p.drag_law;
    pp.visc_law = p.visc_law;
    return pp;
}

int alloc_state(isph_ctx* c, StateView& s, size_t cap) {
    int rc = 0;
    rc |= dalloc(c, &s.pos, cap);
    rc |= dalloc(c, &s.vel, cap);
    rc |= dalloc(c, &s.sp, cap);
    rc |= dalloc(c, &s.sv, cap);
    double** d[] = {&s.mass, &s.rho, &s.oden, &s.p, &s.por, &s.R, &s.nut, &s.tke, &s.varts, &s.velm,
                    &s.con, &s.ovol};
    for (auto pp : d) rc |= dalloc(c, pp, cap);
    rc |= dalloc(c, &s.tid, cap);
    rc |= dalloc(c, &s.flags, cap);
    rc |= dalloc(c, &s.pid, cap);
    return rc ? ISPH_E_HIP : ISPH_OK;
}

void free_state(StateView& s) {
    void* p[] = {s.pos, s.vel, s.sp, s.sv, s.mass, s.rho, s.oden, s.p, s.por, s.R, s.nut, s.tke,
                 s.varts, s.velm, s.con, s.ovol, s.tid, s.flags, s.pid};
    for (auto q : p)
        if (q) (void)hipFree(q);
    s = StateView{};
}

// read and clear the device error word; map to a status
int take_errors(isph_ctx* c) {
    uint32_t e = 0;
    HIPCHK(c, hipMemcpyAsync(&e, c->err, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (!e) return ISPH_OK;
    HIPCHK(c, hipMemsetAsync(c->err, 0, sizeof(uint32_t), c->stream));
    if (e & ERR_NAN) return fail(c, ISPH_E_NAN, "non-finite velocity");
    if (e & ERR_SINGULAR) return fail(c, ISPH_E_SINGULAR, "PPE row with zero or non-finite diagonal");
    if (e & ERR_CAPACITY) return fail(c, ISPH_E_CAPACITY, "neighbour list capacity exceeded");
    if (e & ERR_DOMAIN) return fail(c, ISPH_E_DOMAIN, "particle outside the cell grid");
    if (e & ERR_NOCONV) return fail(c, ISPH_E_NOCONV, "PPE solve did not converge");
    if (e & ERR_SUBSTEPS)
        return fail(c, ISPH_E_CAPACITY, "salinity transport needs more than 2^20 substeps");
    return fail(c, ISPH_E_HIP, "unknown device error");
}

int sync_check(isph_ctx* c) {
    HIPCHK(c, hipGetLastError());
    return take_errors(c);
}

// exclusive scan of `count` ints in place
int scan(isph_ctx* c, int* data, int64_t count) {
    int nb = nblk(count, SCAN_BLOCK);
    k_scan_blocks<<<nb, SCAN_BLOCK, 0, c->stream>>>((int)count, data, c->scanbuf);
    k_scan_sums<<<1, SCAN_BLOCK, 0, c->stream>>>(nb, c->scanbuf);
    k_scan_add<<<nb, SCAN_BLOCK, 0, c->stream>>>((int)count, data, c->scanbuf);
    HIPCHK(c, hipGetLastError());
    return ISPH_OK;
}

// Stable LSD radix sort of the n keys c->key (values in [0, maxkey]): the
// sorted keys into c->key_s, the permutation (sorted position -> input
// position) into c->perm.  Passes alternate between those and c->rank /
// c->rs_val so that the last one lands in key_s / perm.
int radix_sort(isph_ctx* c, int n, int64_t maxkey) {
    int bits = 1;
    while (bits < 31 && (int64_t(1) << bits) <= maxkey) ++bits;
    const int passes = (bits + RS_BITS - 1) / RS_BITS;
    const int ntiles = (n + RS_TILE - 1) / RS_TILE;
    const int* kin = c->key;
    const int* vin = nullptr;
    for (int p = 0; p < passes; ++p) {
        const bool last_parity = ((passes - 1 - p) & 1) == 0;
        int* kout = last_parity ? c->key_s : c->rank;
        int* vout = last_parity ? c->perm : c->rs_val;
        k_rs_hist<<<ntiles, BLOCK, 0, c->stream>>>(n, kin, p * RS_BITS, ntiles, c->rs_hist);
        int rc = scan(c, c->rs_hist, (int64_t)RS_DIGITS * ntiles);
        if (rc) return rc;
        k_rs_scatter<<<ntiles, BLOCK, 0, c->stream>>>(n, kin, vin, p * RS_BITS, ntiles, c->rs_hist,
                                                      kout, vout);
        kin = kout;
        vin = vout;
    }
    HIPCHK(c, hipGetLastError());
    return ISPH_OK;
}

// stable sort by bin key (deleted: key nbins); `first`: no previous binning
// (isph_set_particles)
int sort_particles(isph_ctx* c, bool with_sink, bool first) {
    const int n = c->n;
    const int64_t nb = c->grid.nbins;
    StateView& s = c->st[c->cur];
    StateView& d = c->st[c->cur ^ 1];
    HIPCHK(c, hipMemsetAsync(c->bincnt, 0, sizeof(int) * (nb + 2), c->stream));
    if (n > 0)
        k_bin_keys<<<nblk(n), BLOCK, 0, c->stream>>>(
            n, c->grid, s.pos, s.tid, s.flags, with_sink && c->prm.sink, c->prm.xlcutoff,
            c->prm.ytcutoff, first ? nullptr : c->key_s, c->key, c->bincnt, c->err);
    int rc = scan(c, c->bincnt, nb + 2);
    if (rc) return rc;
    if (n > 0 && (rc = radix_sort(c, n, nb))) return rc;
    int nnew = n;
    if (with_sink && c->prm.sink) {
        HIPCHK(c, hipMemcpyAsync(&nnew, c->bincnt + nb, sizeof(int), hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    if (nnew > 0)
        k_gather_state<<<nblk(nnew), BLOCK, 0, c->stream>>>(nnew, c->perm, s, d, nullptr, nullptr);
    HIPCHK(c, hipGetLastError());
    c->cur ^= 1;
    c->n = nnew;
    c->have_nbr = false;
    c->assembled = false;
    return ISPH_OK;
}

KernelConsts kconsts(const isph_ctx* c) { return make_kc(c->prm); }

__global__ void k_fill(int n, double* x, double v) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) x[i] = v;
}

// ---- stage launchers (enqueue only) ----------------------------------------
int run_projection(isph_ctx* c) {
    if (c->n) k_projection<<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, c->st[c->cur], c->ss);
    HIPCHK(c, hipGetLastError());
    return ISPH_OK;
}

int run_neighbour(isph_ctx* c) {
    const isph_params& p = c->prm;
    if (c->n)
        k_neighbour<<<nblk(c->n), BLOCK, 0, c->stream>>>(
            c->n, c->grid, kconsts(c), p.lamfs * p.maxdivr, p.max_nbr, (size_t)c->cap,
            c->st[c->cur], c->key_s, c->bincnt, c->nbr, c->cnt, c->L, c->divr, c->err);
    HIPCHK(c, hipGetLastError());
    // the lists keep these positions' distances (the reference stores dist and
    // klt per pair, search.f90:65-74): the stages of the salt-water step that
    // run after the shift has moved the particles (scalart, the last eddyvis)
    // take their pair distances from here
    if (c->n)
        HIPCHK(c, hipMemcpyAsync(c->pos_nb, c->st[c->cur].pos, sizeof(double2) * c->n,
                                 hipMemcpyDeviceToDevice, c->stream));
    c->have_nbr = true;
    c->assembled = false;
    return ISPH_OK;
}

int run_effpor(isph_ctx* c) {
    if (c->n)
        k_effpor<<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, c->grid, kconsts(c),
                                                       porous_params(c->prm), c->st[c->cur],
                                                       c->key_s, c->spos, c->smass, c->srho,
                                                       c->sstart, c->err);
    HIPCHK(c, hipGetLastError());
    return ISPH_OK;
}

// Pair geometry from the current positions (the list stages of this library
// take it from there; in the reference's step order that is where the last
// search left the particles) or, with `stored`, from the positions of the
// last search, as the reference's stored klt (turbulence.f90:36) for the
// eddyvis that main.f90:260 calls after the shift has moved them.
static int eddyvis_on(isph_ctx* c, bool stored) {
    if (!c->have_nbr) return fail(c, ISPH_E_STATE, "eddyvis before neighbour");
    StateView v = c->st[c->cur];
    if (stored) v.pos = c->pos_nb;
    if (c->n)
        k_eddyvis<<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, kconsts(c), c->prm.dl, (size_t)c->cap,
                                                        v, c->nbr, c->cnt, c->L);
    HIPCHK(c, hipGetLastError());
    return ISPH_OK;
}

int run_eddyvis(isph_ctx* c) { return eddyvis_on(c, false); }
int run_eddyvis_lagged(isph_ctx* c) { return eddyvis_on(c, true); }

int run_int_vel(isph_ctx* c) {
    if (!c->have_nbr) return fail(c, ISPH_E_STATE, "int_vel before neighbour");
    if (c->n) {
        k_varts<<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, kconsts(c), c->prm.dl, (size_t)c->cap,
                                                      c->st[c->cur], c->nbr, c->cnt, c->L, c->ss);
        k_vstar<<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, kconsts(c), porous_params(c->prm),
                                                      c->prm.lam, c->prm.g, (size_t)c->cap,
                                                      c->st[c->cur], c->nbr, c->cnt, c->L, c->ss);
    }
    HIPCHK(c, hipGetLastError());
    return ISPH_OK;
}

int run_assemble(isph_ctx* c) {
    if (!c->have_nbr) return fail(c, ISPH_E_STATE, "assembly before neighbour");
    const isph_params& p = c->prm;
    if (c->n)
        k_assemble<<<nblk(c->n), BLOCK, 0, c->stream>>>(
            c->n, kconsts(c), p.lam, p.rho0 * p.rel_den, (size_t)c->cap, c->st[c->cur], c->nbr,
            c->cnt, c->L, c->val64, prec_mixed(p.precision) ? c->val32 : nullptr,
            (p.precision == ISPH_PREC_MIXED_V64 || prec_mg(p.precision)) ? c->rsum : nullptr,
            c->dpos, c->rhs, c->ss, c->err);
    HIPCHK(c, hipGetLastError());
    c->assembled = true;
    c->assembled32 = prec_mixed(p.precision);
    return ISPH_OK;
}

int run_comp_vel(isph_ctx* c) {
    if (!c->have_nbr) return fail(c, ISPH_E_STATE, "comp_vel before neighbour");
    if (c->n)
        k_comp_vel<<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, kconsts(c), (size_t)c->cap,
                                                         c->st[c->cur], c->nbr, c->cnt, c->L,
                                                         c->ss, c->err);
    HIPCHK(c, hipGetLastError());
    return ISPH_OK;
}

int run_comp_pos(isph_ctx* c) {
    if (c->n)
        k_comp_pos<<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, c->st[c->cur], c->ss,
                                                        c->prm.xrelease_lo, c->prm.xrelease_hi);
    HIPCHK(c, hipGetLastError());
    return ISPH_OK;
}

int run_freesurf(isph_ctx* c) {
    if (!c->have_nbr) return fail(c, ISPH_E_STATE, "freesurf before neighbour");
    if (c->n)
        k_freesurf<<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, c->prm.lamfs * c->prm.maxdivr,
                                                         (size_t)c->cap, c->st[c->cur], c->nbr,
                                                         c->cnt, c->divr);
    HIPCHK(c, hipGetLastError());
    return ISPH_OK;
}

int run_shift(isph_ctx* c) {
    if (!c->have_nbr) return fail(c, ISPH_E_STATE, "opt2_shift before neighbour");
    if (!c->prm.shifting || !c->n) return ISPH_OK;
    k_shift<<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, kconsts(c), c->prm.csh,
                                                  c->prm.maxshift * c->prm.dl, (size_t)c->cap,
                                                  c->st[c->cur], c->nbr, c->cnt, c->L);
    k_shift_apply<<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, c->st[c->cur]);
    HIPCHK(c, hipGetLastError());
    return ISPH_OK;
}

int run_timestep(isph_ctx* c) {
    const int nb = std::max(nblk(c->n), 1);
    if (c->n)
        k_maxes<<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, c->st[c->cur].velm, c->st[c->cur].nut,
                                                      c->part);
    else
        k_fill<<<1, BLOCK, 0, c->stream>>>(2, c->part, 0.0);
    const isph_params& p = c->prm;
    k_timestep<<<1, BLOCK, 0, c->stream>>>(nb, c->part, c->ss, p.dl, p.sig1, p.mu, p.rho0,
                                           p.adaptive_dt, p.tschmidt, p.salinity, c->err);
    HIPCHK(c, hipGetLastError());
    return ISPH_OK;
}

// ---- PPE solve -------------------------------------------------------------

int read_sc(isph_ctx* c) {
    HIPCHK(c, hipMemcpyAsync(c->sc_host, c->sc, sizeof(SolverScalars), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ISPH_OK;
}

void prof_record(isph_ctx* c, int k, bool start) {
    Profiler& pr = c->prof;
    if (!pr.on) return;
    if (start) {
        pr.cur[k] = (pr.seen[k]++ % pr.stride) == 0;
        if (!pr.cur[k]) return;
        if (pr.used[k] == pr.pool[k].size()) {
            EventPair e;
            (void)hipEventCreate(&e.a);
            (void)hipEventCreate(&e.b);
            pr.pool[k].push_back(e);
            pr.real[k].push_back(0);
        }
        (void)hipEventRecord(pr.pool[k][pr.used[k]].a, c->stream);
    } else {
        if (!pr.cur[k]) return;
        (void)hipEventRecord(pr.pool[k][pr.used[k]].b, c->stream);
        pr.real[k][pr.used[k]] = 1;
        pr.batch_recorded[k]++;
        pr.used[k]++;
    }
}

// Of the launches of the last batch only the first `real_iters` did work;
// the rest returned at once because the device had set `done`.  Recorded
// launches are the ones whose index in the slot's sequence is a multiple of
// the stride, so the recorded launches of the batch are matched to their
// launch index before the no-op tail is dropped.
void prof_mark(isph_ctx* c, int k0, int real_iters, int issued) {
    Profiler& pr = c->prof;
    if (!pr.on) return;
    for (int k = k0; k < k0 + 2; ++k) {
        const int rec = pr.batch_recorded[k];
        const int64_t first_launch = pr.seen[k] - issued;  // index of the batch's first launch
        size_t q = pr.used[k] - (size_t)rec;
        for (int64_t l = first_launch; l < pr.seen[k]; ++l) {
            if (l % pr.stride != 0) continue;
            if (l - first_launch >= real_iters) pr.real[k][q] = 0;
            ++q;
        }
        pr.batch_recorded[k] = 0;
    }
}

void prof_collect(isph_ctx* c) {
    Profiler& pr = c->prof;
    if (!pr.on) return;
    (void)hipStreamSynchronize(c->stream);
    for (int k = 0; k < PROF_SLOTS; ++k) {
        for (size_t q = 0; q < pr.used[k]; ++q) {
            if (!pr.real[k][q]) continue;
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, pr.pool[k][q].a, pr.pool[k][q].b) == hipSuccess) {
                pr.total_ms[k] += ms;
                pr.launches[k] += 1;
            }
        }
        pr.used[k] = 0;
    }
}

SolveCtl solve_ctl(const isph_ctx* c, int mixed) {
    SolveCtl t;
    t.tol = c->prm.tol;
    // inner_tol 0: the precision's default reliable-update factor
    t.delta = c->prm.inner_tol > 0.0 ? c->prm.inner_tol : prec_mg(c->prm.precision) ? 1e-3 : 0.1;
    t.maxit = c->prm.maxit;
    t.maxupd = c->prm.max_outer;
    t.mixed = mixed;
    return t;
}

ResidArgs resid_args(isph_ctx* c, double* x, double* r, const double* val, const double* rhs) {
    ResidArgs a;
    a.n = c->n;
    a.ld = (size_t)c->cap;
    a.nbr = c->nbr;
    a.nbr16 = nullptr;
    a.cnt = c->cnt;
    a.dpos = c->dpos;
    a.val = val ? val : c->val64;
    a.b = rhs ? rhs : c->rhs;
    a.x = x;
    a.r = r;
    a.part = c->part;
    return a;
}

// vector slots: v[0] rhat, v[1] p, v[2] s, v[3] t, v[4] x (or its segment)
template <class TV, class TX, class TI>
KrylovArgs<TV, TX, TI> krylov_args(isph_ctx* c, const TI* nbr, const TV* val, TX* const* v,
                                   typename Pair2<TX>::type* rq, typename Pair2<TX>::type* rv) {
    KrylovArgs<TV, TX, TI> a;
    a.n = c->n;
    a.ld = (size_t)c->cap;
    a.nbr = nbr;
    a.cnt = c->cnt;
    a.dpos = c->dpos;
    a.val = val;
    a.rsum = c->rsum;
    a.rq = rq;
    a.rv = rv;
    a.rhat = v[0];
    a.p = v[1];
    a.si = v[2];
    a.mm2 = v[3];
    a.xs = v[4];
    a.part = c->part;
    return a;
}

// BiCGSTAB iterations in batches until the device sets sc->done; the host
// polls after each batch.  `first` selects p = r for the first iteration
// after a (re)start.  prof slots k0, k0+1: the two SpMV kernels.
template <class TV, class TX, class TI>
int iterate(isph_ctx* c, const KrylovArgs<TV, TX, TI>& a, const SolveCtl& ctl, bool& first,
            int k0) {
    const int grid = nblk8(c->n);
    for (;;) {
        const int before = c->sc_host->total;
        int batch = 8;
        const SolverScalars& h = *c->sc_host;
        if (h.nb > 0 && h.rr > 0) {
            double target = ctl.tol * h.nb;
            if (ctl.mixed) target = std::max(target, ctl.delta * h.rupd);
            const double ratio = std::sqrt(h.rr) / target;
            batch = ratio > 100.0 ? 32 : (ratio > 3.0 ? 16 : 4);
        }
        for (int b = 0; b < batch; ++b) {
            prof_record(c, k0, true);
            if (first)
                k_krylov_1t<TV, TX, TI, true><<<grid, BLOCK, 0, c->stream>>>(a, c->sc);
            else
                k_krylov_1t<TV, TX, TI, false><<<grid, BLOCK, 0, c->stream>>>(a, c->sc);
            prof_record(c, k0, false);
            k_finish<<<1, FIN_BLOCK, 0, c->stream>>>(FIN_ALPHA, grid, c->part, c->sc, ctl);
            prof_record(c, k0 + 1, true);
            k_krylov_2<TV, TX, TI><<<grid, BLOCK, 0, c->stream>>>(a, c->sc);
            prof_record(c, k0 + 1, false);
            k_finish<<<1, FIN_BLOCK, 0, c->stream>>>(FIN_OMEGA, grid, c->part, c->sc, ctl);
            k_krylov_3<TV, TX, TI><<<grid, BLOCK, 0, c->stream>>>(a, c->sc);
            k_finish<<<1, FIN_BLOCK, 0, c->stream>>>(FIN_BETA, grid, c->part, c->sc, ctl);
            first = false;
        }
        HIPCHK(c, hipGetLastError());
        int rc = read_sc(c);
        if (rc) return rc;
        prof_mark(c, k0, c->sc_host->total - before, batch);
        if (c->sc_host->done) return ISPH_OK;
    }
}

template <class TI>
int solve_fp64_t(isph_ctx* c, const TI* nbr, const double* val, const double* rhs,
                 double* x) {
    const int grid = nblk8(c->n);
    // the FP64 solve builds x in place
    double* vv[5] = {c->v64[0], c->v64[1], c->v64[2], c->v64[3], x};
    auto ka = krylov_args<double, double, TI>(c, nbr, val, vv, c->rq64, c->rv64);
    const SolveCtl ctl = solve_ctl(c, 0);
    ResidArgs ra = resid_args(c, x, c->v64[9], val, rhs);
    if constexpr (std::is_same<TI, short>::value) ra.nbr16 = nbr;
    k_resid<<<grid, BLOCK, 0, c->stream>>>(ra, 1, c->rq64, c->v64[0]);
    k_finish<<<1, FIN_BLOCK, 0, c->stream>>>(FIN_INIT, grid, c->part, c->sc, ctl);
    HIPCHK(c, hipGetLastError());
    int rc = read_sc(c);
    if (rc) return rc;
    bool first = true;
    for (;;) {
        const int d = c->sc_host->done;
        if (d == DONE_ZERO_RHS) {
            k_fill<<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, x, 0.0);
            return ISPH_OK;
        }
        if (d == DONE_CONVERGED) return ISPH_OK;
        if (d == DONE_FAILED) return fail(c, ISPH_E_NOCONV, "PPE solve did not converge");
        if (d == DONE_RECURRENCE) {
            k_resid<<<grid, BLOCK, 0, c->stream>>>(ra, 1, c->rq64, c->v64[0]);
            k_finish<<<1, FIN_BLOCK, 0, c->stream>>>(FIN_TRUE, grid, c->part, c->sc, ctl);
            HIPCHK(c, hipGetLastError());
            if ((rc = read_sc(c))) return rc;
            first = true;
            continue;
        }
        if ((rc = iterate(c, ka, ctl, first, 0))) return rc;
    }
}

// 16-bit column offsets j - i when every offset of the lists fits (the index
// stream of the two SpMVs per iteration drops from 4 to 2 bytes per entry),
// else the int32 columns; the arithmetic is the same either way.
int solve_fp64_on(isph_ctx* c, const double* val, const double* rhs, double* x) {
    k_to_short<<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, (size_t)c->cap, c->nbr, c->cnt, c->nbr16,
                                                     c->err);
    uint32_t e = 0;
    HIPCHK(c, hipMemcpyAsync(&e, c->err, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (e & ERR_INDEX16) {
        HIPCHK(c, hipMemsetAsync(c->err, 0, sizeof(uint32_t), c->stream));
        c->narrow_idx = false;
        return solve_fp64_t<int>(c, c->nbr, val, rhs, x);
    }
    c->narrow_idx = true;
    return solve_fp64_t<short>(c, c->nbr16, val, rhs, x);
}

int solve_fp64(isph_ctx* c) { return solve_fp64_on(c, c->val64, c->rhs, c->v64[8]); }

template <class TX, class TI>
int solve_mixed_t(isph_ctx* c, const TI* nbr, TX* const* v, typename Pair2<TX>::type* rq,
                  typename Pair2<TX>::type* rv, int k0) {
    const int grid = nblk8(c->n);
    double* x = c->v64[8];
    double* r64 = c->v64[9];
    auto ka = krylov_args<float, TX, TI>(c, nbr, c->val32, v, rq, rv);
    SolveCtl ctl = solve_ctl(c, 1);
    SolveCtl ctl_flush = solve_ctl(c, 2);
    ResidArgs ra = resid_args(c, x, r64);
    if constexpr (std::is_same<TI, short>::value) ra.nbr16 = nbr;
    k_resid<<<grid, BLOCK, 0, c->stream>>>(ra, 0, nullptr, nullptr);
    k_finish<<<1, FIN_BLOCK, 0, c->stream>>>(FIN_INIT, grid, c->part, c->sc, solve_ctl(c, 0));
    k_finish<<<1, FIN_BLOCK, 0, c->stream>>>(FIN_RESID, grid, c->part, c->sc, ctl);
    HIPCHK(c, hipGetLastError());
    int rc = read_sc(c);
    if (rc) return rc;
    if (c->sc_host->done == DONE_ZERO_RHS) {
        k_fill<<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, x, 0.0);
        return ISPH_OK;
    }
    if (c->sc_host->done == DONE_CONVERGED) return ISPH_OK;
    k_mx_start<TX><<<grid, BLOCK, 0, c->stream>>>(c->n, r64, rq, v[0], v[4], c->part, c->sc);
    k_finish<<<1, FIN_BLOCK, 0, c->stream>>>(FIN_START, grid, c->part, c->sc, ctl);
    HIPCHK(c, hipGetLastError());
    if ((rc = read_sc(c))) return rc;
    bool first = true;
    for (;;) {
        if ((rc = iterate(c, ka, ctl, first, k0))) return rc;
        const int d = c->sc_host->done;
        if (d == DONE_FAILED) return fail(c, ISPH_E_NOCONV, "mixed-precision PPE solve did not converge");
        if (d != DONE_UPDATE) return fail(c, ISPH_E_STATE, "unexpected solver state");
        // reliable update: flush x, FP64 true residual, replace r
        k_mx_flush<TX><<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, x, v[4]);
        k_resid<<<grid, BLOCK, 0, c->stream>>>(ra, 0, nullptr, nullptr);
        k_finish<<<1, FIN_BLOCK, 0, c->stream>>>(FIN_RESID, grid, c->part, c->sc, ctl_flush);
        HIPCHK(c, hipGetLastError());
        if ((rc = read_sc(c))) return rc;
        if (c->sc_host->done == DONE_CONVERGED) return ISPH_OK;
        if (c->sc_host->done == DONE_FAILED)
            return fail(c, ISPH_E_NOCONV, "mixed-precision PPE solve did not converge");
        if (c->sc_host->brk || c->prm.precision == ISPH_PREC_IR) {
            // breakdown, or iterative refinement: the next correction starts
            // from zero on the true residual, rhat = r
            k_mx_start<TX><<<grid, BLOCK, 0, c->stream>>>(c->n, r64, rq, v[0], v[4], c->part, c->sc);
            k_finish<<<1, FIN_BLOCK, 0, c->stream>>>(FIN_START, grid, c->part, c->sc, ctl);
            first = true;
        } else {
            k_mx_replace<TX><<<grid, BLOCK, 0, c->stream>>>(c->n, r64, rq, v[0], c->part, c->sc);
            k_finish<<<1, FIN_BLOCK, 0, c->stream>>>(FIN_REPLACE, grid, c->part, c->sc, ctl);
        }
        HIPCHK(c, hipGetLastError());
    }
}

int solve_mixed(isph_ctx* c) {
    // 16-bit column offsets when every offset fits, else the int32 columns
    k_to_short<<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, (size_t)c->cap, c->nbr, c->cnt, c->nbr16,
                                                     c->err);
    uint32_t e = 0;
    HIPCHK(c, hipMemcpyAsync(&e, c->err, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const bool narrow = !(e & ERR_INDEX16);
    if (!narrow) HIPCHK(c, hipMemsetAsync(c->err, 0, sizeof(uint32_t), c->stream));
    c->narrow_idx = narrow;
    if (c->prm.precision == ISPH_PREC_MIXED_V64) {
        double* vv[5] = {c->v64[0], c->v64[1], c->v64[2], c->v64[3], c->v64[4]};
        return narrow ? solve_mixed_t<double, short>(c, c->nbr16, vv, c->rq64, c->rv64, 2)
                      : solve_mixed_t<double, int>(c, c->nbr, vv, c->rq64, c->rv64, 2);
    }
    float* vv[5] = {c->v32[0], c->v32[1], c->v32[2], c->v32[3], c->v32[4]};
    return narrow ? solve_mixed_t<float, short>(c, c->nbr16, vv, c->rq32, c->rv32, 2)
                  : solve_mixed_t<float, int>(c, c->nbr, vv, c->rq32, c->rv32, 2);
}

int run_solve(isph_ctx* c) {
    if (!c->assembled) return fail(c, ISPH_E_STATE, "solve before assembly");
    const bool mixed = prec_mixed(c->prm.precision);
    if (mixed && !c->assembled32)
        return fail(c, ISPH_E_STATE, "mixed precision needs an assembly with a mixed precision mode");
    StateView& s = c->st[c->cur];
    double* x = c->v64[8];
    if (c->n == 0) return ISPH_OK;
    if (c->prm.precision == ISPH_PREC_FGMRES) {
        // warm start fsol in row order (solver.f90:319); the solution is
        // fetched into fsol only on convergence (solver.f90:330-333), and
        // the pressure is fsol in either case (isph.f90:465-483)
        const size_t nd = sizeof(double) * c->n;
        HIPCHK(c, hipMemcpyAsync(x, c->fsol, nd, hipMemcpyDeviceToDevice, c->stream));
        bool conv = false;
        int rc = solve_gmres(c, x, conv);
        c->outer = 0;
        if (rc) return rc;
        if (conv) HIPCHK(c, hipMemcpyAsync(c->fsol, x, nd, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(s.p, c->fsol, nd, hipMemcpyDeviceToDevice, c->stream));
        return ISPH_OK;
    }
    const double* start = c->solve_replay ? c->solve_replay : s.p;
    HIPCHK(c, hipMemcpyAsync(x, start, sizeof(double) * c->n, hipMemcpyDeviceToDevice, c->stream));
    int rc = prec_mg(c->prm.precision) ? solve_mg(c)
             : mixed                                ? solve_mixed(c)
                                                    : solve_fp64(c);
    // every solve returns on success right after reading the solver state
    // (a second read would cost a synchronisation); after a failure the state
    // is read again for the counts below
    if (rc != ISPH_OK) (void)read_sc(c);
    c->iters = c->sc_host->total;
    c->relres = c->sc_host->relres;
    c->outer = mixed ? c->sc_host->nupd : 0;
    if (rc != ISPH_OK) return rc;
    // pressure scatter (isph.f90:465-483): rows are particles
    HIPCHK(c, hipMemcpyAsync(s.p, x, sizeof(double) * c->n, hipMemcpyDeviceToDevice, c->stream));
    if (c->solve_record)
        HIPCHK(c, hipMemcpyAsync(c->solve_record, x, sizeof(double) * c->n,
                                 hipMemcpyDeviceToDevice, c->stream));
    return ISPH_OK;
}

// One step: main.f90:41-66, or with salinity main.f90:199-263 (transport
// substeps first, densityupdate after the shift, eddyvis after timestep).
int run_step(isph_ctx* c) {
    int rc;
    const bool salt = c->prm.salinity != 0;
    double told = 0.0, tnew = 0.0;
    if (c->nzones) {  // the buffer schedule compares the times around the advance
        HIPCHK(c, hipMemcpyAsync(c->ss_host, c->ss, sizeof(SimScalars), hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        told = c->ss_host->t;
        tnew = told + c->ss_host->dt;  // k_advance_time's sum
    }
    k_advance_time<<<1, 64, 0, c->stream>>>(c->ss);
    if (salt && (rc = run_salt_substeps(c))) return rc;
    if (c->nzones && (rc = run_buffer_schedule(c, told, tnew))) return rc;
    if ((rc = run_projection(c))) return rc;
    if ((rc = sort_particles(c, false, false))) return rc;
    if ((rc = run_neighbour(c))) return rc;
    if ((rc = run_effpor(c))) return rc;
    if ((rc = run_eddyvis(c))) return rc;
    if ((rc = run_int_vel(c))) return rc;
    if ((rc = run_assemble(c))) return rc;
    if ((rc = run_solve(c))) return rc;
    if ((rc = take_errors(c))) return rc;
    if ((rc = run_comp_vel(c))) return rc;
    if ((rc = run_comp_pos(c))) return rc;
    if ((rc = sort_particles(c, true, false))) return rc;
    if ((rc = run_neighbour(c))) return rc;
    if ((rc = run_freesurf(c))) return rc;
    if ((rc = run_effpor(c))) return rc;
    if ((rc = run_shift(c))) return rc;
    if (salt && (rc = run_densityupdate(c))) return rc;
    if ((rc = run_timestep(c))) return rc;
    if (salt && (rc = run_eddyvis_lagged(c))) return rc;
    return ISPH_OK;
}

int copy_state(isph_ctx* c, const StateView& a, const StateView& b, int n) {
    auto D2D = hipMemcpyDeviceToDevice;
    const size_t n2 = sizeof(double2) * n, nd = sizeof(double) * n, ni = sizeof(int) * n;
    HIPCHK(c, hipMemcpyAsync(b.pos, a.pos, n2, D2D, c->stream));
    HIPCHK(c, hipMemcpyAsync(b.vel, a.vel, n2, D2D, c->stream));
    HIPCHK(c, hipMemcpyAsync(b.sp, a.sp, n2, D2D, c->stream));
    HIPCHK(c, hipMemcpyAsync(b.sv, a.sv, n2, D2D, c->stream));
    const double* da[] = {a.mass, a.rho, a.oden, a.p,     a.por, a.R,
                          a.nut,  a.tke, a.varts, a.velm, a.con, a.ovol};
    double* db[] = {b.mass, b.rho, b.oden, b.p, b.por, b.R, b.nut, b.tke, b.varts, b.velm, b.con, b.ovol};
    for (int k = 0; k < 12; ++k) HIPCHK(c, hipMemcpyAsync(db[k], da[k], nd, D2D, c->stream));
    HIPCHK(c, hipMemcpyAsync(b.tid, a.tid, ni, D2D, c->stream));
    HIPCHK(c, hipMemcpyAsync(b.flags, a.flags, ni, D2D, c->stream));
    HIPCHK(c, hipMemcpyAsync(b.pid, a.pid, ni, D2D, c->stream));
    return ISPH_OK;
}

}  // namespace host
}  // namespace isph

using namespace isph::host;

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

int isph_create(const isph_params* prm, int32_t cap, int32_t cap_soil, isph_ctx** out) {
    if (!prm || !out || cap <= 0 || cap_soil < 0) return ISPH_E_ARG;
    isph_ctx* c = new isph_ctx();
    *out = nullptr;
    int rc = check_params(c, prm);
    if (rc) {
        delete c;
        return rc;
    }
    c->prm = *prm;
    c->cap = cap;
    c->cap_soil = cap_soil;
    c->grid = make_grid(*prm);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return ISPH_E_HIP;
    }
    // the device address of this GPU's isph_probe_noalloc scratch array
    if ((rc = probe_scratch_attach(c))) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    const size_t N = (size_t)cap, M = (size_t)prm->max_nbr;
    const int64_t nb = c->grid.nbins;
    rc |= alloc_state(c, c->st[0], N);
    rc |= alloc_state(c, c->st[1], N);
    rc |= dalloc(c, &c->L, 4 * N);
    rc |= dalloc(c, &c->divr, N);
    rc |= dalloc(c, &c->key, N);
    rc |= dalloc(c, &c->key_s, N);
    rc |= dalloc(c, &c->rank, N);
    rc |= dalloc(c, &c->perm, N);
    rc |= dalloc(c, &c->rs_val, N);
    rc |= dalloc(c, &c->rs_hist, (size_t)RS_DIGITS * ((N + RS_TILE - 1) / RS_TILE));
    rc |= dalloc(c, &c->bincnt, (size_t)nb + 2);
    // block sums of the scans: over the bins of a re-binning, over the tile
    // digit counts of a radix pass and over the particles of a buffer removal
    const int64_t scan_max = std::max<int64_t>(
        std::max<int64_t>(nb + 2, (int64_t)N), (int64_t)RS_DIGITS * ((N + RS_TILE - 1) / RS_TILE));
    rc |= dalloc(c, &c->scanbuf, (size_t)nblk(scan_max, SCAN_BLOCK) + 1);
    rc |= dalloc(c, &c->nbr, N * M);
    rc |= dalloc(c, &c->cnt, N);
    rc |= dalloc(c, &c->val64, N * M);
    rc |= dalloc(c, &c->val32, N * M);
    rc |= dalloc(c, &c->dpos, N);
    rc |= dalloc(c, &c->nbr16, N * M);
    rc |= dalloc(c, &c->rhs, N);
    rc |= dalloc(c, &c->rsum, N);
    for (int k = 0; k < 10; ++k) rc |= dalloc(c, &c->v64[k], N);
    for (int k = 0; k < 9; ++k) rc |= dalloc(c, &c->v32[k], N);
    rc |= dalloc(c, &c->rq64, N);
    rc |= dalloc(c, &c->rv64, N);
    rc |= dalloc(c, &c->rq32, N);
    rc |= dalloc(c, &c->rv32, N);
    rc |= dalloc(c, &c->part, 4 * (size_t)nblk8(cap));
    rc |= dalloc(c, &c->counters, 16);
    rc |= dalloc(c, &c->sc, 1);
    rc |= dalloc(c, &c->ss, 1);
    rc |= dalloc(c, &c->err, 1);
    rc |= dalloc(c, &c->spos, (size_t)std::max(cap_soil, 1));
    rc |= dalloc(c, &c->smass, (size_t)std::max(cap_soil, 1));
    rc |= dalloc(c, &c->srho, (size_t)std::max(cap_soil, 1));
    rc |= dalloc(c, &c->sstart, (size_t)c->grid.nx * c->grid.ny + 1);
    rc |= dalloc(c, &c->pos_nb, N);
    rc |= dalloc(c, &c->cdiff, N);
    rc |= dalloc(c, &c->salt_coef, N);
    rc |= dalloc(c, &c->fsol, N);
    if (hipHostMalloc((void**)&c->sc_host, sizeof(SolverScalars), hipHostMallocDefault) != hipSuccess)
        rc = ISPH_E_HIP;
    if (hipHostMalloc((void**)&c->ss_host, sizeof(SimScalars), hipHostMallocDefault) != hipSuccess)
        rc = ISPH_E_HIP;
    if (rc) {
        isph_destroy(c);
        return ISPH_E_HIP;
    }
    memset(c->sc_host, 0, sizeof(SolverScalars));
    SimScalars ss0 = {0.0, 0.0, prm->dt, 0.0, prm->dt, 0, 0};
    if (hipMemcpyAsync(c->ss, &ss0, sizeof(ss0), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        isph_destroy(c);
        return ISPH_E_HIP;
    }
    *out = c;
    return ISPH_OK;
}

int isph_destroy(isph_ctx* c) {
    if (!c) return ISPH_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_state(c->st[0]);
    free_state(c->st[1]);
    void* p[] = {c->L, c->divr, c->key, c->key_s, c->rank, c->perm, c->rs_val, c->rs_hist,
                 c->bincnt, c->scanbuf, c->nbr16,
                 c->rq64, c->rv64, c->rq32, c->rv32, c->rsum,
                 c->nbr, c->cnt, c->val64, c->val32, c->dpos, c->rhs, c->part, c->counters,
                 c->sc, c->ss, c->err, c->spos, c->smass, c->srho, c->sstart,
                 c->pos_nb, c->cdiff, c->salt_coef, c->fsol};
    for (auto q : p)
        if (q) (void)hipFree(q);
    for (int k = 0; k < 12; ++k)
        if (c->v64[k]) (void)hipFree(c->v64[k]);
    for (int k = 0; k < 10; ++k)
        if (c->v32[k]) (void)hipFree(c->v32[k]);
    if (c->sc_host) (void)hipHostFree(c->sc_host);
    if (c->ss_host) (void)hipHostFree(c->ss_host);
    adj_free(c);
    free_zones(c);
    mg_free(c);
    gm_free(c);
    for (auto& sl : c->slots) {
        if (!sl.v.pos) continue;
        free_state(sl.v);
        if (sl.key) (void)hipFree(sl.key);
        if (sl.fsol) (void)hipFree(sl.fsol);
        if (sl.bins) (void)hipFree(sl.bins);
    }
    for (int k = 0; k < PROF_SLOTS; ++k)
        for (auto& e : c->prof.pool[k]) {
            (void)hipEventDestroy(e.a);
            (void)hipEventDestroy(e.b);
        }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return ISPH_OK;
}

const char* isph_last_error(const isph_ctx* c) { return c ? c->msg.c_str() : "null context"; }

int isph_set_params(isph_ctx* c, const isph_params* prm) {
    if (!c || !prm) return ISPH_E_ARG;
    int rc = check_params(c, prm);
    if (rc) return rc;
    GridInfo g = make_grid(*prm);
    if (g.nx != c->grid.nx || g.ny != c->grid.ny || prm->max_nbr != c->prm.max_nbr)
        return fail(c, ISPH_E_ARG, "grid and max_nbr are fixed at create time");
    c->prm = *prm;
    c->grid = g;
    c->assembled = false;
    return ISPH_OK;
}

int isph_get_params(const isph_ctx* c, isph_params* prm) {
    if (!c || !prm) return ISPH_E_ARG;
    *prm = c->prm;
    return ISPH_OK;
}

int isph_set_particles(isph_ctx* c, int32_t n, const double* x, const double* y,
                       const double* vx, const double* vy, const double* mass,
                       const double* density, const double* oden, const double* pressure,
                       const int32_t* tid, const int32_t* flags, const int32_t* pid) {
    if (!c) return ISPH_E_ARG;
    if (n < 0 || n > c->cap) return fail(c, ISPH_E_ARG, "n exceeds capacity");
    if (n > 0 && (!x || !y || !vx || !vy || !mass || !density || !oden || !pressure || !tid || !pid))
        return fail(c, ISPH_E_ARG, "null particle array");
    StateView& s = c->st[c->cur];
    std::vector<double2> a(n), b(n), z(n, make_double2(0.0, 0.0));
    std::vector<double> velm(n), one(n, 1.0), zero(n, 0.0), ovol(n);
    std::vector<int> fl(n, 0);
    for (int i = 0; i < n; ++i) {
        a[i] = make_double2(x[i], y[i]);
        b[i] = make_double2(vx[i], vy[i]);
        velm[i] = std::sqrt(vx[i] * vx[i] + vy[i] * vy[i]);
        ovol[i] = mass[i] / oden[i];  // fmass/rho at generation (internal.f90:103)
        if (flags) fl[i] = flags[i];
    }
    const size_t nd = sizeof(double) * n, ni = sizeof(int) * n, n2 = sizeof(double2) * n;
    auto H2D = hipMemcpyHostToDevice;
    HIPCHK(c, hipMemcpyAsync(s.pos, a.data(), n2, H2D, c->stream));
    HIPCHK(c, hipMemcpyAsync(s.vel, b.data(), n2, H2D, c->stream));
    HIPCHK(c, hipMemcpyAsync(s.sp, z.data(), n2, H2D, c->stream));
    HIPCHK(c, hipMemcpyAsync(s.sv, z.data(), n2, H2D, c->stream));
    HIPCHK(c, hipMemcpyAsync(s.mass, mass, nd, H2D, c->stream));
    HIPCHK(c, hipMemcpyAsync(s.rho, density, nd, H2D, c->stream));
    HIPCHK(c, hipMemcpyAsync(s.oden, oden, nd, H2D, c->stream));
    HIPCHK(c, hipMemcpyAsync(s.p, pressure, nd, H2D, c->stream));
    HIPCHK(c, hipMemcpyAsync(s.por, one.data(), nd, H2D, c->stream));
    double* zs[] = {s.R, s.nut, s.tke, s.varts, s.con};
    HIPCHK(c, hipMemcpyAsync(s.ovol, ovol.data(), nd, H2D, c->stream));
    for (auto q : zs) HIPCHK(c, hipMemcpyAsync(q, zero.data(), nd, H2D, c->stream));
    HIPCHK(c, hipMemcpyAsync(s.velm, velm.data(), nd, H2D, c->stream));
    HIPCHK(c, hipMemcpyAsync(s.tid, tid, ni, H2D, c->stream));
    HIPCHK(c, hipMemcpyAsync(s.flags, fl.data(), ni, H2D, c->stream));
    HIPCHK(c, hipMemcpyAsync(s.pid, pid, ni, H2D, c->stream));
    std::vector<double> Lh(4 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        Lh[4 * i] = 1.0; Lh[4 * i + 1] = 0.0; Lh[4 * i + 2] = 0.0; Lh[4 * i + 3] = 1.0;
    }
    HIPCHK(c, hipMemcpyAsync(c->L, Lh.data(), 4 * nd, H2D, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->divr, zero.data(), nd, H2D, c->stream));
    SimScalars ss0 = {0.0, 0.0, c->prm.dt, 0.0, c->prm.dt, 0, 0};
    HIPCHK(c, hipMemcpyAsync(c->ss, &ss0, sizeof(ss0), H2D, c->stream));
    c->n = n;
    c->next_pid = 0;
    for (int i = 0; i < n; ++i) c->next_pid = std::max(c->next_pid, (int64_t)pid[i] + 1);
    int rc = sort_particles(c, false, true);
    if (rc) return rc;
    // fsol = fguess: the uploaded pressures in the first row order
    // (memory.f90:113-144)
    if (n > 0)
        HIPCHK(c, hipMemcpyAsync(c->fsol, c->st[c->cur].p, sizeof(double) * n,
                                 hipMemcpyDeviceToDevice, c->stream));
    return sync_check(c);
}

int isph_get_particles(isph_ctx* c, int32_t* n, double* x, double* y, double* vx, double* vy,
                       double* mass, double* density, double* oden, double* pressure,
                       int32_t* tid, int32_t* flags, int32_t* pid) {
    if (!c) return ISPH_E_ARG;
    if (n) *n = c->n;
    const int m = c->n;
    StateView& s = c->st[c->cur];
    const size_t nd = sizeof(double) * m, ni = sizeof(int) * m;
    auto D2H = hipMemcpyDeviceToHost;
    std::vector<double2> a;
    if (x || y) {
        a.resize(m);
        HIPCHK(c, hipMemcpyAsync(a.data(), s.pos, sizeof(double2) * m, D2H, c->stream));
    }
    std::vector<double2> b;
    if (vx || vy) {
        b.resize(m);
        HIPCHK(c, hipMemcpyAsync(b.data(), s.vel, sizeof(double2) * m, D2H, c->stream));
    }
    if (mass) HIPCHK(c, hipMemcpyAsync(mass, s.mass, nd, D2H, c->stream));
    if (density) HIPCHK(c, hipMemcpyAsync(density, s.rho, nd, D2H, c->stream));
    if (oden) HIPCHK(c, hipMemcpyAsync(oden, s.oden, nd, D2H, c->stream));
    if (pressure) HIPCHK(c, hipMemcpyAsync(pressure, s.p, nd, D2H, c->stream));
    if (tid) HIPCHK(c, hipMemcpyAsync(tid, s.tid, ni, D2H, c->stream));
    if (flags) HIPCHK(c, hipMemcpyAsync(flags, s.flags, ni, D2H, c->stream));
    if (pid) HIPCHK(c, hipMemcpyAsync(pid, s.pid, ni, D2H, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < m; ++i) {
        if (x) x[i] = a[i].x;
        if (y) y[i] = a[i].y;
        if (vx) vx[i] = b[i].x;
        if (vy) vy[i] = b[i].y;
    }
    return ISPH_OK;
}

int isph_set_soil(isph_ctx* c, int32_t ns, const double* x, const double* y, const double* mass,
                  const double* density) {
    if (!c) return ISPH_E_ARG;
    if (ns < 0 || ns > c->cap_soil) return fail(c, ISPH_E_ARG, "ns exceeds soil capacity");
    const GridInfo& g = c->grid;
    const int64_t nc = (int64_t)g.nx * g.ny;
    std::vector<int> cnt(nc + 1, 0), k(ns);
    bool oob = false;
    for (int i = 0; i < ns; ++i) {
        double fx = std::floor((x[i] - g.xmin) / g.cs), fy = std::floor((y[i] - g.ymin) / g.cs);
        int cx, cy;
        if (!(fx >= 0.0)) { cx = 1; oob = true; }
        else if (fx > (double)(g.nx - 3)) { cx = g.nx - 2; oob = true; }
        else cx = (int)fx + 1;
        if (!(fy >= 0.0)) { cy = 1; oob = true; }
        else if (fy > (double)(g.ny - 3)) { cy = g.ny - 2; oob = true; }
        else cy = (int)fy + 1;
        k[i] = cx * g.ny + cy;
        cnt[k[i] + 1]++;
    }
    for (int64_t b = 0; b < nc; ++b) cnt[b + 1] += cnt[b];
    std::vector<int> pos(cnt.begin(), cnt.end() - 1);
    std::vector<double2> sp(ns);
    std::vector<double> sm(ns), sr(ns);
    for (int i = 0; i < ns; ++i) {
        int d = pos[k[i]]++;
        sp[d] = make_double2(x[i], y[i]);
        sm[d] = mass[i];
        sr[d] = density[i];
    }
    auto H2D = hipMemcpyHostToDevice;
    if (ns) {
        HIPCHK(c, hipMemcpyAsync(c->spos, sp.data(), sizeof(double2) * ns, H2D, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->smass, sm.data(), sizeof(double) * ns, H2D, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->srho, sr.data(), sizeof(double) * ns, H2D, c->stream));
    }
    HIPCHK(c, hipMemcpyAsync(c->sstart, cnt.data(), sizeof(int) * (nc + 1), H2D, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->ns = ns;
    if (oob) return fail(c, ISPH_E_DOMAIN, "soil particle outside the cell grid");
    return ISPH_OK;
}

static double* field_ptr(isph_ctx* c, int32_t f, int* width, bool* pair, int* comp) {
    StateView& s = c->st[c->cur];
    *width = 1;
    *pair = false;
    *comp = 0;
    switch (f) {
        case ISPH_FIELD_XS: *pair = true; *comp = 0; return (double*)s.sp;
        case ISPH_FIELD_YS: *pair = true; *comp = 1; return (double*)s.sp;
        case ISPH_FIELD_VXS: *pair = true; *comp = 0; return (double*)s.sv;
        case ISPH_FIELD_VYS: *pair = true; *comp = 1; return (double*)s.sv;
        case ISPH_FIELD_POROSITY: return s.por;
        case ISPH_FIELD_RESIST: return s.R;
        case ISPH_FIELD_COFF: *width = 4; return c->L;
        case ISPH_FIELD_DIVR: return c->divr;
        case ISPH_FIELD_NUT: return s.nut;
        case ISPH_FIELD_TKE: return s.tke;
        case ISPH_FIELD_VARTS: return s.varts;
        case ISPH_FIELD_VEL: return s.velm;
        case ISPH_FIELD_RHS: return c->rhs;
        case ISPH_FIELD_X: *pair = true; *comp = 0; return (double*)s.pos;
        case ISPH_FIELD_Y: *pair = true; *comp = 1; return (double*)s.pos;
        case ISPH_FIELD_VX: *pair = true; *comp = 0; return (double*)s.vel;
        case ISPH_FIELD_VY: *pair = true; *comp = 1; return (double*)s.vel;
        case ISPH_FIELD_DENSITY: return s.rho;
        case ISPH_FIELD_PRESSURE: return s.p;
        case ISPH_FIELD_CON: return s.con;
        case ISPH_FIELD_OVOL: return s.ovol;
        case ISPH_FIELD_CDIFF: return c->cdiff;
        case ISPH_FIELD_MASS: return s.mass;
        case ISPH_FIELD_ODEN: return s.oden;
        default: return nullptr;
    }
}

int isph_get_field(isph_ctx* c, int32_t f, double* out) {
    if (!c || !out) return ISPH_E_ARG;
    int w, comp;
    bool pair;
    double* p = field_ptr(c, f, &w, &pair, &comp);
    if (!p) return fail(c, ISPH_E_ARG, "unknown field");
    const int m = c->n;
    if (pair) {
        std::vector<double> t(2 * (size_t)m);
        HIPCHK(c, hipMemcpyAsync(t.data(), p, sizeof(double) * 2 * m, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int i = 0; i < m; ++i) out[i] = t[2 * (size_t)i + comp];
        return ISPH_OK;
    }
    HIPCHK(c, hipMemcpyAsync(out, p, sizeof(double) * w * m, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ISPH_OK;
}

int isph_set_field(isph_ctx* c, int32_t f, const double* in) {
    if (!c || !in) return ISPH_E_ARG;
    int w, comp;
    bool pair;
    double* p = field_ptr(c, f, &w, &pair, &comp);
    if (!p) return fail(c, ISPH_E_ARG, "unknown field");
    const int m = c->n;
    if (pair) {
        std::vector<double> t(2 * (size_t)m);
        HIPCHK(c, hipMemcpyAsync(t.data(), p, sizeof(double) * 2 * m, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int i = 0; i < m; ++i) t[2 * (size_t)i + comp] = in[i];
        HIPCHK(c, hipMemcpyAsync(p, t.data(), sizeof(double) * 2 * m, hipMemcpyHostToDevice, c->stream));
    } else {
        HIPCHK(c, hipMemcpyAsync(p, in, sizeof(double) * w * m, hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ISPH_OK;
}

int isph_get_neighbours(isph_ctx* c, int32_t* ptr, int32_t* idx) {
    if (!c || !ptr) return ISPH_E_ARG;
    if (!c->have_nbr) return fail(c, ISPH_E_STATE, "no neighbour list");
    const int m = c->n;
    std::vector<int> cnt(m);
    HIPCHK(c, hipMemcpyAsync(cnt.data(), c->cnt, sizeof(int) * m, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    ptr[0] = 0;
    int maxc = 0;
    for (int i = 0; i < m; ++i) {
        ptr[i + 1] = ptr[i] + cnt[i];
        maxc = std::max(maxc, cnt[i]);
    }
    if (!idx) return ISPH_OK;
    std::vector<int> ell((size_t)maxc * c->cap);
    for (int k = 0; k < maxc; ++k)
        HIPCHK(c, hipMemcpyAsync(ell.data() + (size_t)k * c->cap, c->nbr + (size_t)k * c->cap,
                                 sizeof(int) * m, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < m; ++i)
        for (int k = 0; k < cnt[i]; ++k) idx[ptr[i] + k] = ell[(size_t)k * c->cap + i];
    return ISPH_OK;
}

int isph_get_csr(isph_ctx* c, int32_t* rowptr, int32_t* col, double* val, double* rhs) {
    if (!c || !rowptr) return ISPH_E_ARG;
    if (!c->assembled) return fail(c, ISPH_E_STATE, "no assembled matrix");
    const int m = c->n;
    std::vector<int> cnt(m), dp(m);
    HIPCHK(c, hipMemcpyAsync(cnt.data(), c->cnt, sizeof(int) * m, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(dp.data(), c->dpos, sizeof(int) * m, hipMemcpyDeviceToHost, c->stream));
    if (rhs) HIPCHK(c, hipMemcpyAsync(rhs, c->rhs, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    rowptr[0] = 0;
    int maxc = 0;
    for (int i = 0; i < m; ++i) {
        rowptr[i + 1] = rowptr[i] + cnt[i] + 1;
        maxc = std::max(maxc, cnt[i]);
    }
    if (!col && !val) return ISPH_OK;
    std::vector<int> ell((size_t)maxc * c->cap);
    std::vector<double> ev((size_t)maxc * c->cap);
    for (int k = 0; k < maxc; ++k) {
        HIPCHK(c, hipMemcpyAsync(ell.data() + (size_t)k * c->cap, c->nbr + (size_t)k * c->cap,
                                 sizeof(int) * m, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(ev.data() + (size_t)k * c->cap, c->val64 + (size_t)k * c->cap,
                                 sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < m; ++i) {
        int base = rowptr[i];
        for (int k = 0; k < cnt[i]; ++k) {
            int slot = base + (k < dp[i] ? k : k + 1);
            if (col) col[slot] = ell[(size_t)k * c->cap + i];
            if (val) val[slot] = ev[(size_t)k * c->cap + i];
        }
        if (col) col[base + dp[i]] = i;
        if (val) val[base + dp[i]] = 1.0;
    }
    return ISPH_OK;
}

int isph_get_scalars(isph_ctx* c, double* umax, double* dt, double* t, int32_t* iters,
                     int32_t* outer, double* relres) {
    if (!c) return ISPH_E_ARG;
    SimScalars s;
    HIPCHK(c, hipMemcpyAsync(&s, c->ss, sizeof(s), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (umax) *umax = s.umax;
    if (dt) *dt = s.dt;
    if (t) *t = s.t;
    if (iters) *iters = c->iters;
    if (outer) *outer = c->outer;
    if (relres) *relres = c->relres;
    return ISPH_OK;
}

int isph_set_scalars(isph_ctx* c, double umax, double dt, double t) {
    if (!c) return ISPH_E_ARG;
    SimScalars s;
    HIPCHK(c, hipMemcpyAsync(&s, c->ss, sizeof(s), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    s.umax = umax;
    s.dt = dt;
    s.t = t;
    HIPCHK(c, hipMemcpyAsync(c->ss, &s, sizeof(s), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ISPH_OK;
}

#define STAGE(name, fn)                     \
    int name(isph_ctx* c) {                 \
        if (!c) return ISPH_E_ARG;          \
        int rc = fn(c);                     \
        if (rc) return rc;                  \
        return sync_check(c);               \
    }

static int cellshift_(isph_ctx* c) { return sort_particles(c, false, false); }
static int cellshiftalt_(isph_ctx* c) { return sort_particles(c, true, false); }
static int resetid_(isph_ctx* c) { (void)c; return ISPH_OK; }

STAGE(isph_projection, run_projection)
STAGE(isph_cellshift, cellshift_)
STAGE(isph_resetid, resetid_)
STAGE(isph_neighbour, run_neighbour)
STAGE(isph_effpor, run_effpor)
STAGE(isph_eddyvis, run_eddyvis)
STAGE(isph_eddyvis_lagged, run_eddyvis_lagged)
STAGE(isph_int_vel, run_int_vel)
STAGE(isph_ppe_assemble, run_assemble)
STAGE(isph_comp_vel, run_comp_vel)
STAGE(isph_comp_pos, run_comp_pos)
STAGE(isph_cellshiftalt, cellshiftalt_)
STAGE(isph_freesurf, run_freesurf)
STAGE(isph_opt2_shift, run_shift)
STAGE(isph_timestep, run_timestep)

int isph_ppe_solve(isph_ctx* c, int32_t* iters, double* relres) {
    if (!c) return ISPH_E_ARG;
    int rc = run_solve(c);
    if (iters) *iters = c->iters;
    if (relres) *relres = c->relres;
    prof_collect(c);
    if (rc) return rc;
    return sync_check(c);
}

int isph_ppesolve(isph_ctx* c, int32_t* iters, double* relres) {
    if (!c) return ISPH_E_ARG;
    int rc = run_assemble(c);
    if (rc) return rc;
    if ((rc = take_errors(c))) return rc;
    return isph_ppe_solve(c, iters, relres);
}

int isph_step(isph_ctx* c, int32_t nsteps) {
    if (!c || nsteps < 0) return ISPH_E_ARG;
    for (int s = 0; s < nsteps; ++s) {
        int rc = run_step(c);
        if (rc) {
            prof_collect(c);
            return rc;
        }
    }
    prof_collect(c);
    return sync_check(c);
}

int isph_checkpoint(isph_ctx* c, int32_t slot) {
    if (!c || slot < 0 || slot >= 64) return ISPH_E_ARG;
    auto& sl = c->slots[slot];
    if (!sl.v.pos) {
        // all or nothing: a slot is either complete or left empty
        int rc = alloc_state(c, sl.v, (size_t)c->cap);
        rc |= dalloc(c, &sl.key, (size_t)c->cap);
        rc |= dalloc(c, &sl.fsol, (size_t)c->cap);
        rc |= dalloc(c, &sl.bins, (size_t)c->grid.nbins + 2);
        if (rc) {
            free_state(sl.v);
            void* q[] = {sl.key, sl.fsol, sl.bins};
            for (auto x : q)
                if (x) (void)hipFree(x);
            sl.key = nullptr;
            sl.fsol = nullptr;
            sl.bins = nullptr;
            return fail(c, ISPH_E_HIP, "checkpoint allocation failed");
        }
    }
    int rc = copy_state(c, c->st[c->cur], sl.v, c->n);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(sl.key, c->key_s, sizeof(int) * c->n, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(sl.fsol, c->fsol, sizeof(double) * c->cap, hipMemcpyDeviceToDevice,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(sl.bins, c->bincnt, sizeof(int) * (c->grid.nbins + 2),
                             hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(&sl.ss, c->ss, sizeof(SimScalars), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    sl.n = c->n;
    sl.nzones = c->nzones;
    for (int k = 0; k < c->nzones; ++k) sl.zone_counter[k] = c->zones[k].counter;
    sl.zone_shift = c->zone_shift;
    sl.next_pid = c->next_pid;
    sl.used = true;
    return ISPH_OK;
}

int isph_restore(isph_ctx* c, int32_t slot) {
    if (!c || slot < 0 || slot >= 64) return ISPH_E_ARG;
    auto& sl = c->slots[slot];
    if (!sl.used) return fail(c, ISPH_E_STATE, "restore from an empty checkpoint slot");
    if (sl.nzones != c->nzones)
        return fail(c, ISPH_E_STATE, "buffer zones were installed or removed since this checkpoint");
    int rc = copy_state(c, sl.v, c->st[c->cur], sl.n);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(c->key_s, sl.key, sizeof(int) * sl.n, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->fsol, sl.fsol, sizeof(double) * c->cap, hipMemcpyDeviceToDevice,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(c->bincnt, sl.bins, sizeof(int) * (c->grid.nbins + 2),
                             hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->ss, &sl.ss, sizeof(SimScalars), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->n = sl.n;
    for (int k = 0; k < c->nzones; ++k) c->zones[k].counter = sl.zone_counter[k];
    c->zone_shift = sl.zone_shift;
    c->next_pid = sl.next_pid;
    c->have_nbr = false;
    c->assembled = false;
    return ISPH_OK;
}

int isph_set_progress(isph_ctx* c, isph_progress_fn fn, void* user) {
    if (!c) return ISPH_E_ARG;
    c->progress = fn;
    c->progress_user = user;
    return ISPH_OK;
}

int isph_profile_enable(isph_ctx* c, int32_t enable) {
    if (!c || enable < 0) return ISPH_E_ARG;
    c->prof.on = enable != 0;
    c->prof.stride = enable > 0 ? enable : 1;
    for (int k = 0; k < PROF_SLOTS; ++k) {
        c->prof.launches[k] = 0;
        c->prof.total_ms[k] = 0.0;
        c->prof.used[k] = 0;
        c->prof.seen[k] = 0;
        c->prof.batch_recorded[k] = 0;
    }
    return ISPH_OK;
}

int isph_get_kernel_time(isph_ctx* c, int32_t kernel, int64_t* launches, double* total_ms) {
    if (!c || kernel < 0 || kernel >= PROF_SLOTS) return ISPH_E_ARG;
    if (launches) *launches = c->prof.launches[kernel];
    if (total_ms) *total_ms = c->prof.total_ms[kernel];
    return ISPH_OK;
}

int isph_copy_bench(isph_ctx* c, int64_t bytes, int32_t reps, double* ms) {
    if (!c || bytes <= 0 || reps <= 0 || !ms) return ISPH_E_ARG;
    size_t n2 = (size_t)bytes / sizeof(double2);
    double2 *a = nullptr, *b = nullptr;
    HIPCHK(c, hipMalloc((void**)&a, n2 * sizeof(double2)));
    if (hipMalloc((void**)&b, n2 * sizeof(double2)) != hipSuccess) {
        (void)hipFree(a);
        return fail(c, ISPH_E_HIP, "copy_bench allocation failed");
    }
    (void)hipMemsetAsync(a, 0, n2 * sizeof(double2), c->stream);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    int grid = 256 * 8 * 4;
    k_copy<<<grid, BLOCK, 0, c->stream>>>(a, b, n2);
    (void)hipEventRecord(e0, c->stream);
    for (int r = 0; r < reps; ++r) k_copy<<<grid, BLOCK, 0, c->stream>>>(a, b, n2);
    (void)hipEventRecord(e1, c->stream);
    (void)hipEventSynchronize(e1);
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e0, e1);
    *ms = (double)t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return ISPH_OK;
}

}  // extern "C"
