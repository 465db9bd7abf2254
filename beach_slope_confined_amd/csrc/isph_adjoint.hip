// isph_adjoint.hip -- discrete adjoint of the ISPH step (host side).
//
// Forward (isph_adjoint_forward): the porous closure and timestep at the
// initial state, then the steps of main.f90:41-66; the state at the start
// of every segment of `every` steps and the final state stay in HBM.
//
// Reverse (isph_adjoint_backward), segment by segment from the last: the
// segment's steps are recomputed from its checkpoint with a tape of each
// step's intermediate states (three state copies, both neighbour lists and
// KGC matrices, the two re-binning permutations, the pressure); then the
// steps are reversed, each stage of a step in the opposite order:
//   objective term -> timestep -> shift -> effpor -> KGC -> re-binning ->
//   comp_pos -> comp_vel -> PPE (transposed solve, assembly) -> v* ->
//   varts -> eddyvis -> effpor -> KGC -> re-binning -> projection
// and at the end the initial porous closure.  The PPE adjoint is
// lambda = A^-T pbar, solved by the forward's BiCGSTAB (or its multigrid-
// preconditioned form) on the transposed values (the lists are symmetric,
// so A^T has A's sparsity).  Parameter
// gradients accumulate per particle slot and are summed in a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "isph_adjoint_kernels.h"
#include "isph_host.h"

using namespace isph;
using namespace isph::adj;

namespace isph {
namespace host {

struct Ckpt {
    StateView v{};
    int* key = nullptr;   // the sorted bin keys and bin starts of the state
    int* bins = nullptr;
    SimScalars ss{};
    int n = 0;
};

struct StepTape {
    StateView Sa{}, Sc{}, Sf{};  // step start, before the solve, before the shift
    int *perm1 = nullptr, *perm2 = nullptr;
    int *nbr1 = nullptr, *cnt1 = nullptr, *nbr2 = nullptr, *cnt2 = nullptr;
    double *L1 = nullptr, *L2 = nullptr;
    double *rho0 = nullptr, *rho1p = nullptr, *p = nullptr;
    double2* vel1 = nullptr;
    SimScalars *ssc = nullptr, *ssend = nullptr;
    bool ok = false;
};

struct AdjointState {
    // forward record
    bool ready = false;
    int nsteps = 0, every = 0;
    isph_params prm{};
    ObjConst oc{};
    double J = 0.0;
    std::vector<Ckpt> seg;
    Ckpt fin;
    // reverse buffers
    std::vector<StepTape> tape;
    bool have_adj = false;
    AdjView A[2]{};
    int ca = 0;
    double* th = nullptr;
    double* aum = nullptr;
    double* w = nullptr;
    double* lam = nullptr;
    double* valT = nullptr;
    float* val32T = nullptr;  // FP32 values and FP64 row sums of A^T for the
    double* rsumT = nullptr;  // multigrid-preconditioned transposed solve
    double* dsc = nullptr;   // [0] aumax, [1] J, [2..25] objective sums, [26..29] gradients
    double* part = nullptr;  // 3 * OBJ_MAXP * nblk(cap)
    // lambda of the step reversed last, by particle id: the warm start of
    // the next transposed solve (the adjacent step, rows in another order)
    double* lam_pid = nullptr;
    int npid = 0;       // entries allocated in lam_pid
    // every forward step's PPE solution (nsteps x rec_n doubles, row order):
    // the backward's recomputation starts each solve from it, so the solve
    // stops at its first true-residual test with the same bits; null when
    // it did not fit in memory (the recomputation then solves in full)
    double* rec = nullptr;
    int64_t rec_cap = 0;  // doubles allocated in rec
    int64_t rec_n = 0;    // particles of the recorded run (constant: no zones, no sink)
    int warm_npid = 0;  // entries the transposed solves read and write: npid in
                        // the backward sweep, 0 (zero start) otherwise
};

namespace {

int alloc_ckpt(isph_ctx* c, Ckpt& k) {
    if (k.v.pos) return ISPH_OK;
    int rc = alloc_state(c, k.v, (size_t)c->cap);
    rc |= dalloc(c, &k.key, (size_t)c->cap);
    rc |= dalloc(c, &k.bins, (size_t)c->grid.nbins + 2);
    return rc;
}

int save_ckpt(isph_ctx* c, Ckpt& k) {
    int rc = alloc_ckpt(c, k);
    if (rc) return fail(c, ISPH_E_HIP, "adjoint checkpoint allocation failed");
    if ((rc = copy_state(c, c->st[c->cur], k.v, c->n))) return rc;
    HIPCHK(c, hipMemcpyAsync(k.key, c->key_s, sizeof(int) * c->n, hipMemcpyDeviceToDevice,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(k.bins, c->bincnt, sizeof(int) * (c->grid.nbins + 2),
                             hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(&k.ss, c->ss, sizeof(SimScalars), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    k.n = c->n;
    return ISPH_OK;
}

int load_ckpt(isph_ctx* c, const Ckpt& k) {
    int rc = copy_state(c, k.v, c->st[c->cur], k.n);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(c->key_s, k.key, sizeof(int) * k.n, hipMemcpyDeviceToDevice,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(c->bincnt, k.bins, sizeof(int) * (c->grid.nbins + 2),
                             hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->ss, &k.ss, sizeof(SimScalars), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->n = k.n;
    c->have_nbr = false;
    c->assembled = false;
    return ISPH_OK;
}

void free_ckpt(Ckpt& k) {
    if (k.v.pos) free_state(k.v);
    if (k.key) (void)hipFree(k.key);
    if (k.bins) (void)hipFree(k.bins);
    k = Ckpt{};
}

int alloc_tape(isph_ctx* c, StepTape& t) {
    if (t.ok) return ISPH_OK;
    const size_t N = (size_t)c->cap, M = (size_t)c->prm.max_nbr;
    int rc = 0;
    rc |= alloc_state(c, t.Sa, N);
    rc |= alloc_state(c, t.Sc, N);
    rc |= alloc_state(c, t.Sf, N);
    rc |= dalloc(c, &t.perm1, N);
    rc |= dalloc(c, &t.perm2, N);
    rc |= dalloc(c, &t.nbr1, N * M);
    rc |= dalloc(c, &t.nbr2, N * M);
    rc |= dalloc(c, &t.cnt1, N);
    rc |= dalloc(c, &t.cnt2, N);
    rc |= dalloc(c, &t.L1, 4 * N);
    rc |= dalloc(c, &t.L2, 4 * N);
    rc |= dalloc(c, &t.rho0, N);
    rc |= dalloc(c, &t.rho1p, N);
    rc |= dalloc(c, &t.p, N);
    rc |= dalloc(c, &t.vel1, N);
    rc |= dalloc(c, &t.ssc, 1);
    rc |= dalloc(c, &t.ssend, 1);
    if (rc) return fail(c, ISPH_E_HIP, "adjoint tape allocation failed");
    t.ok = true;
    return ISPH_OK;
}

void free_tape(StepTape& t) {
    if (!t.ok) return;
    free_state(t.Sa);
    free_state(t.Sc);
    free_state(t.Sf);
    void* p[] = {t.perm1, t.perm2, t.nbr1, t.nbr2, t.cnt1, t.cnt2, t.L1, t.L2,
                 t.rho0, t.rho1p, t.p, t.vel1, t.ssc, t.ssend};
    for (auto q : p)
        if (q) (void)hipFree(q);
    t = StepTape{};
}

int alloc_adjview(isph_ctx* c, AdjView& a) {
    const size_t N = (size_t)c->cap;
    int rc = 0;
    rc |= dalloc(c, &a.pos, N);
    rc |= dalloc(c, &a.vel, N);
    rc |= dalloc(c, &a.sp, N);
    rc |= dalloc(c, &a.sv, N);
    double** d[] = {&a.por, &a.rho, &a.R, &a.nut, &a.tke, &a.varts, &a.p, &a.velm};
    for (auto q : d) rc |= dalloc(c, q, N);
    rc |= dalloc(c, &a.L, 4 * N);
    a.cap = N;
    return rc;
}

void free_adjview(AdjView& a) {
    void* p[] = {a.pos, a.vel, a.sp, a.sv, a.por, a.rho, a.R, a.nut, a.tke, a.varts, a.p, a.velm, a.L};
    for (auto q : p)
        if (q) (void)hipFree(q);
    a = AdjView{};
}

int zero_adjview(isph_ctx* c, const AdjView& a) {
    const size_t N = a.cap;
    HIPCHK(c, hipMemsetAsync(a.pos, 0, sizeof(double2) * N, c->stream));
    HIPCHK(c, hipMemsetAsync(a.vel, 0, sizeof(double2) * N, c->stream));
    HIPCHK(c, hipMemsetAsync(a.sp, 0, sizeof(double2) * N, c->stream));
    HIPCHK(c, hipMemsetAsync(a.sv, 0, sizeof(double2) * N, c->stream));
    double* d[] = {a.por, a.rho, a.R, a.nut, a.tke, a.varts, a.p, a.velm};
    for (auto q : d) HIPCHK(c, hipMemsetAsync(q, 0, sizeof(double) * N, c->stream));
    HIPCHK(c, hipMemsetAsync(a.L, 0, sizeof(double) * 4 * N, c->stream));
    return ISPH_OK;
}

AdjointState* state(isph_ctx* c) {
    if (!c->adj) c->adj = new AdjointState();
    return c->adj;
}

// adjoint arrays, parameter and umax accumulators, solve vectors
int ensure_adj(isph_ctx* c) {
    AdjointState* S = state(c);
    if (S->have_adj) return ISPH_OK;
    const size_t N = (size_t)c->cap, M = (size_t)c->prm.max_nbr;
    int rc = 0;
    rc |= alloc_adjview(c, S->A[0]);
    rc |= alloc_adjview(c, S->A[1]);
    rc |= dalloc(c, &S->th, 3 * N);
    rc |= dalloc(c, &S->aum, N);
    rc |= dalloc(c, &S->w, 6 * N);
    rc |= dalloc(c, &S->lam, N);
    rc |= dalloc(c, &S->valT, N * M);
    rc |= dalloc(c, &S->dsc, 8 + 3 * OBJ_MAXP);
    rc |= dalloc(c, &S->part, (size_t)3 * OBJ_MAXP * nblk(c->cap));
    if (rc) return fail(c, ISPH_E_HIP, "adjoint buffer allocation failed");
    for (int k = 0; k < 2; ++k) {
        S->A[k].th = S->th;
        S->A[k].aum = S->aum;
        S->A[k].w = S->w;
    }
    S->have_adj = true;
    return ISPH_OK;
}

int zero_all(isph_ctx* c) {
    AdjointState* S = state(c);
    const size_t N = (size_t)c->cap;
    int rc = zero_adjview(c, S->A[0]);
    if (!rc) rc = zero_adjview(c, S->A[1]);
    if (rc) return rc;
    HIPCHK(c, hipMemsetAsync(S->th, 0, sizeof(double) * 3 * N, c->stream));
    HIPCHK(c, hipMemsetAsync(S->aum, 0, sizeof(double) * N, c->stream));
    HIPCHK(c, hipMemsetAsync(S->dsc, 0, sizeof(double) * (8 + 3 * OBJ_MAXP), c->stream));
    S->ca = 0;
    return ISPH_OK;
}

int make_obj(isph_ctx* c, const isph_objective* o, ObjConst& oc) {
    if (!o || o->nprobe < 0 || o->ngauge < 0 || o->nprobe > ISPH_OBJ_MAX ||
        o->ngauge > ISPH_OBJ_MAX || o->nprobe + o->ngauge == 0)
        return fail(c, ISPH_E_ARG, "objective needs 1..8 probes and/or 1..8 gauges");
    memset(&oc, 0, sizeof(oc));
    oc.kc = kconsts(c);
    oc.np = o->nprobe;
    oc.ng = o->ngauge;
    for (int k = 0; k < oc.np; ++k) {
        oc.px[k] = o->probe_x[k];
        oc.py[k] = o->probe_y[k];
    }
    for (int k = 0; k < oc.ng; ++k) oc.gx[k] = o->gauge_x[k];
    return ISPH_OK;
}

// objective sums at state s into dsc[2..]; weight*J_step added to dsc[1]
// when accum
int obj_sums(isph_ctx* c, const ObjConst& oc, const StateView& s, double weight, bool accum) {
    AdjointState* S = state(c);
    const int nb = std::max(nblk(c->n), 1);
    const int M = 2 * oc.np + oc.ng;
    k_obj_sums<<<nb, BLOCK, 0, c->stream>>>(c->n, oc, s, S->part);
    k_sum_parts<<<M, BLOCK, 0, c->stream>>>(nb, S->part, S->dsc + 2);
    if (accum) k_obj_accum<<<1, 64, 0, c->stream>>>(oc, S->dsc + 2, weight, S->dsc + 1);
    HIPCHK(c, hipGetLastError());
    return ISPH_OK;
}

// sum of count doubles -> *out (device), fixed order
int dsum(isph_ctx* c, int64_t count, const double* x, double* out) {
    AdjointState* S = state(c);
    const int nb = std::max(nblk(count), 1);
    k_sum_blocks<<<nb, BLOCK, 0, c->stream>>>(count, x, S->part);
    k_sum_parts<<<1, BLOCK, 0, c->stream>>>(nb, S->part, out);
    HIPCHK(c, hipGetLastError());
    return ISPH_OK;
}

int copy_dev(isph_ctx* c, void* dst, const void* src, size_t bytes) {
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    return ISPH_OK;
}

// one step of main.f90:41-66 (the run_step order) recording the tape
int tape_step(isph_ctx* c, StepTape& t) {
    const size_t N = (size_t)c->cap, M = (size_t)c->prm.max_nbr;
    int rc;
    if ((rc = copy_state(c, c->st[c->cur], t.Sa, c->n))) return rc;
    k_advance_time<<<1, 64, 0, c->stream>>>(c->ss);
    if ((rc = run_projection(c))) return rc;
    if ((rc = sort_particles(c, false, false))) return rc;
    if ((rc = copy_dev(c, t.perm1, c->perm, sizeof(int) * c->n))) return rc;
    if ((rc = copy_dev(c, t.rho0, c->st[c->cur].rho, sizeof(double) * c->n))) return rc;
    if ((rc = run_neighbour(c))) return rc;
    if ((rc = copy_dev(c, t.nbr1, c->nbr, sizeof(int) * N * M))) return rc;
    if ((rc = copy_dev(c, t.cnt1, c->cnt, sizeof(int) * c->n))) return rc;
    if ((rc = copy_dev(c, t.L1, c->L, sizeof(double) * 4 * c->n))) return rc;
    if ((rc = run_effpor(c))) return rc;
    if ((rc = run_eddyvis(c))) return rc;
    if ((rc = run_int_vel(c))) return rc;
    if ((rc = run_assemble(c))) return rc;
    if ((rc = copy_state(c, c->st[c->cur], t.Sc, c->n))) return rc;
    if ((rc = copy_dev(c, t.ssc, c->ss, sizeof(SimScalars)))) return rc;
    if ((rc = run_solve(c))) return rc;
    if ((rc = take_errors(c))) return rc;
    if ((rc = copy_dev(c, t.p, c->st[c->cur].p, sizeof(double) * c->n))) return rc;
    if ((rc = run_comp_vel(c))) return rc;
    if ((rc = copy_dev(c, t.vel1, c->st[c->cur].vel, sizeof(double2) * c->n))) return rc;
    if ((rc = run_comp_pos(c))) return rc;
    if ((rc = sort_particles(c, true, false))) return rc;
    if ((rc = copy_dev(c, t.perm2, c->perm, sizeof(int) * c->n))) return rc;
    if ((rc = copy_dev(c, t.rho1p, c->st[c->cur].rho, sizeof(double) * c->n))) return rc;
    if ((rc = run_neighbour(c))) return rc;
    if ((rc = copy_dev(c, t.nbr2, c->nbr, sizeof(int) * N * M))) return rc;
    if ((rc = copy_dev(c, t.cnt2, c->cnt, sizeof(int) * c->n))) return rc;
    if ((rc = copy_dev(c, t.L2, c->L, sizeof(double) * 4 * c->n))) return rc;
    if ((rc = run_freesurf(c))) return rc;
    if ((rc = run_effpor(c))) return rc;
    if ((rc = copy_state(c, c->st[c->cur], t.Sf, c->n))) return rc;
    if ((rc = run_shift(c))) return rc;
    if ((rc = run_timestep(c))) return rc;
    if ((rc = copy_dev(c, t.ssend, c->ss, sizeof(SimScalars)))) return rc;
    return sync_check(c);
}

VstarConst vstar_const(const isph_ctx* c) {
    VstarConst v;
    v.kc = kconsts(c);
    v.pp = porous_params(c->prm);
    v.lam = c->prm.lam;
    v.g = c->prm.g;
    v.w0 = c->prm.kernel == ISPH_KERNEL_WENDLAND_C2 ? v.kc.wnorm
                                                    : (ALP * 4.0) * v.kc.n1 + ((1.0 - ALP) * 6.0) * v.kc.n2;
    return v;
}

AsmConst asm_const(const isph_ctx* c) {
    AsmConst a;
    a.kc = kconsts(c);
    a.lam = c->prm.lam;
    a.lampd = c->prm.rho0 * c->prm.rel_den;
    return a;
}

// ---- stage adjoints (enqueue only) ------------------------------------------
int adj_kgc(isph_ctx* c, const AdjView& a, const StateView& s, const int* nbr, const int* cnt) {
    const int n = c->n;
    const size_t ld = (size_t)c->cap;
    k_adj_kgc_self<<<nblk(n), BLOCK, 0, c->stream>>>(n, kconsts(c), ld, s, nbr, cnt, a);
    k_adj_kgc_nbr<<<nblk(n), BLOCK, 0, c->stream>>>(n, kconsts(c), ld, s, nbr, cnt, a);
    HIPCHK(c, hipGetLastError());
    return ISPH_OK;
}

int adj_effpor(isph_ctx* c, const AdjView& a, const StateView& s) {
    k_adj_effpor<<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, c->grid, kconsts(c),
                                                      porous_params(c->prm), s, c->spos, c->smass,
                                                      c->srho, c->sstart, a);
    HIPCHK(c, hipGetLastError());
    return ISPH_OK;
}

int adj_strain(isph_ctx* c, const AdjView& a, const StateView& s, const int* nbr, const int* cnt,
               const double* L, const SimScalars* ss, int eddy) {
    const int n = c->n;
    const size_t ld = (size_t)c->cap;
    k_adj_strain_self<<<nblk(n), BLOCK, 0, c->stream>>>(n, kconsts(c), c->prm.dl, eddy, ld, s, nbr,
                                                        cnt, L, ss, a);
    k_adj_strain_nbr<<<nblk(n), BLOCK, 0, c->stream>>>(n, kconsts(c), eddy, ld, s, nbr, cnt, L, a);
    HIPCHK(c, hipGetLastError());
    return ISPH_OK;
}

int adj_vstar(isph_ctx* c, const AdjView& a, const StateView& s, const int* nbr, const int* cnt,
              const double* L, const SimScalars* ss) {
    const int n = c->n;
    const size_t ld = (size_t)c->cap;
    const VstarConst vc = vstar_const(c);
    k_adj_vstar_self<<<nblk(n), BLOCK, 0, c->stream>>>(n, vc, ld, s, nbr, cnt, L, ss, a);
    k_adj_vstar_nbr<<<nblk(n), BLOCK, 0, c->stream>>>(n, vc, ld, s, nbr, cnt, L, a);
    HIPCHK(c, hipGetLastError());
    return ISPH_OK;
}

// PPE: assemble A from s (the forward's values), lambda = A^-T abar.p,
// then the assembly adjoint.  p = the forward solution.
int adj_ppe(isph_ctx* c, const AdjView& a, const StateView& s, int* nbr, int* cnt, const double* L,
            const double* p, SimScalars* ss, int64_t* iters) {
    AdjointState* S = state(c);
    const int n = c->n;
    const size_t ld = (size_t)c->cap;
    const isph_params& pr = c->prm;
    int* nbr0 = c->nbr;
    int* cnt0 = c->cnt;
    c->nbr = nbr;
    c->cnt = cnt;
    k_assemble<<<nblk(n), BLOCK, 0, c->stream>>>(n, kconsts(c), pr.lam, pr.rho0 * pr.rel_den, ld, s,
                                                  nbr, cnt, L, c->val64, nullptr, nullptr, c->dpos,
                                                  c->rhs, ss, c->err);
    k_transpose_vals<<<nblk(n), BLOCK, 0, c->stream>>>(n, ld, nbr, cnt, c->val64, S->valT, c->err);
    // warm start: each particle's lambda of the step reversed before this
    // one (zero for a particle that step did not have)
    k_lam_gather<<<nblk(n), BLOCK, 0, c->stream>>>(n, s.pid, S->lam_pid, S->warm_npid, S->lam);
    int rc = take_errors(c);
    if (!rc && prec_mg(pr.precision)) {
        // the forward solves with the multigrid preconditioner; so does the
        // transposed solve, on A^T's own Galerkin hierarchy
        if (!S->val32T) {
            const size_t N = (size_t)c->cap, M = (size_t)pr.max_nbr;
            if (dalloc(c, &S->val32T, N * M) || dalloc(c, &S->rsumT, N))
                rc = fail(c, ISPH_E_HIP, "adjoint buffer allocation failed");
        }
        if (!rc) {
            // the hierarchy reads the bin of every row: the taped state is
            // sorted by the keys of its positions, which have not moved
            // since that step's re-binning, so the keys and bin starts are
            // recomputed from them (the forward recomputation re-sorts)
            const int64_t nb = c->grid.nbins;
            HIPCHK(c, hipMemsetAsync(c->bincnt, 0, sizeof(int) * (nb + 2), c->stream));
            k_bin_keys<<<nblk(n), BLOCK, 0, c->stream>>>(n, c->grid, s.pos, s.tid, s.flags, 0, 0.0,
                                                          0.0, nullptr, c->key_s, c->bincnt,
                                                          c->err);
            rc = scan(c, c->bincnt, nb + 2);
        }
        if (!rc) rc = take_errors(c);
        if (!rc) {
            k_mg_lower<<<nblk(n), BLOCK, 0, c->stream>>>(n, ld, cnt, c->dpos, S->valT, S->val32T,
                                                          S->rsumT);
            rc = solve_mg_on(c, MGSystem{S->valT, S->val32T, S->rsumT, a.p, S->lam});
        }
    } else if (!rc) {
        rc = solve_fp64_on(c, S->valT, a.p, S->lam);
    }
    if (!rc) rc = read_sc(c);
    if (!rc && iters) *iters += c->sc_host->total;
    if (!rc && S->warm_npid > 0)
        k_lam_scatter<<<nblk(n), BLOCK, 0, c->stream>>>(n, s.pid, S->lam, S->lam_pid, S->warm_npid);
    c->nbr = nbr0;
    c->cnt = cnt0;
    c->assembled = false;
    if (rc) return rc;
    HIPCHK(c, hipMemsetAsync(a.p, 0, sizeof(double) * n, c->stream));
    const AsmConst ac = asm_const(c);
    k_adj_asm_self<<<nblk(n), BLOCK, 0, c->stream>>>(n, ac, ld, s, nbr, cnt, L, S->lam, p, ss, a);
    k_adj_asm_nbr<<<nblk(n), BLOCK, 0, c->stream>>>(n, ac, ld, s, nbr, cnt, L, p, ss, a);
    HIPCHK(c, hipGetLastError());
    return ISPH_OK;
}

int adj_comp_vel(isph_ctx* c, const AdjView& a, const StateView& s, const int* nbr, const int* cnt,
                 const double* L, const double* p, const SimScalars* ss) {
    const int n = c->n;
    const size_t ld = (size_t)c->cap;
    k_adj_cv_self<<<nblk(n), BLOCK, 0, c->stream>>>(n, kconsts(c), ld, s, nbr, cnt, L, p, ss, a);
    k_adj_cv_nbr<<<nblk(n), BLOCK, 0, c->stream>>>(n, kconsts(c), ld, s, nbr, cnt, L, p, ss, a);
    HIPCHK(c, hipGetLastError());
    return ISPH_OK;
}

int adj_shift(isph_ctx* c, const AdjView& a, const StateView& s, const int* nbr, const int* cnt,
              const double* L) {
    if (!c->prm.shifting) return ISPH_OK;
    const int n = c->n;
    const size_t ld = (size_t)c->cap;
    const double cap = c->prm.maxshift * c->prm.dl;
    k_adj_shift_self<<<nblk(n), BLOCK, 0, c->stream>>>(n, kconsts(c), c->prm.csh, cap, ld, s, nbr,
                                                       cnt, L, a);
    k_adj_shift_nbr<<<nblk(n), BLOCK, 0, c->stream>>>(n, kconsts(c), c->prm.csh, ld, s, nbr, cnt,
                                                      L, a);
    HIPCHK(c, hipGetLastError());
    return ISPH_OK;
}

// adjoint arrays: A[ca] current; the re-binning scatters into the other one
int adj_unperm(isph_ctx* c, const int* perm) {
    AdjointState* S = state(c);
    k_adj_unperm<<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, perm, S->A[S->ca], S->A[S->ca ^ 1]);
    HIPCHK(c, hipGetLastError());
    S->ca ^= 1;
    return ISPH_OK;
}

// reverse of one taped step; `end` is the state after the step
int reverse_step(isph_ctx* c, StepTape& t, const StateView& end, double weight, int64_t* iters) {
    AdjointState* S = state(c);
    const int n = c->n;
    int rc;
    // objective term at the end of the step
    if ((rc = obj_sums(c, S->oc, end, weight, false))) return rc;
    k_adj_obj<<<nblk(n), BLOCK, 0, c->stream>>>(n, S->oc, end, S->dsc + 2, weight, S->A[S->ca]);
    // timestep: umax of this step feeds the next step's varts
    k_adj_umax<<<nblk(n), BLOCK, 0, c->stream>>>(n, t.Sf.velm, t.ssend, S->dsc, S->A[S->ca]);
    HIPCHK(c, hipGetLastError());
    if ((rc = adj_shift(c, S->A[S->ca], t.Sf, t.nbr2, t.cnt2, t.L2))) return rc;
    if ((rc = adj_effpor(c, S->A[S->ca], t.Sf))) return rc;
    StateView se = t.Sf;
    se.rho = t.rho1p;
    if ((rc = adj_kgc(c, S->A[S->ca], se, t.nbr2, t.cnt2))) return rc;
    if ((rc = adj_unperm(c, t.perm2))) return rc;
    StateView sd = t.Sc;
    sd.vel = t.vel1;
    k_adj_comp_pos<<<nblk(n), BLOCK, 0, c->stream>>>(n, sd, t.ssc, S->A[S->ca]);
    HIPCHK(c, hipGetLastError());
    if ((rc = adj_comp_vel(c, S->A[S->ca], t.Sc, t.nbr1, t.cnt1, t.L1, t.p, t.ssc))) return rc;
    if ((rc = adj_ppe(c, S->A[S->ca], t.Sc, t.nbr1, t.cnt1, t.L1, t.p, t.ssc, iters))) return rc;
    if ((rc = adj_vstar(c, S->A[S->ca], t.Sc, t.nbr1, t.cnt1, t.L1, t.ssc))) return rc;
    if ((rc = adj_strain(c, S->A[S->ca], t.Sc, t.nbr1, t.cnt1, t.L1, t.ssc, 0))) return rc;
    if ((rc = adj_strain(c, S->A[S->ca], t.Sc, t.nbr1, t.cnt1, t.L1, t.ssc, 1))) return rc;
    if ((rc = adj_effpor(c, S->A[S->ca], t.Sc))) return rc;
    StateView sb = t.Sc;
    sb.rho = t.rho0;
    if ((rc = adj_kgc(c, S->A[S->ca], sb, t.nbr1, t.cnt1))) return rc;
    if ((rc = adj_unperm(c, t.perm1))) return rc;
    k_adj_projection<<<nblk(n), BLOCK, 0, c->stream>>>(n, t.Sa, t.ssc, S->A[S->ca]);
    HIPCHK(c, hipGetLastError());
    // dJ/d umax of the previous step's timestep
    if ((rc = dsum(c, n, S->aum, S->dsc))) return rc;
    HIPCHK(c, hipMemsetAsync(S->aum, 0, sizeof(double) * n, c->stream));
    return sync_check(c);
}

int check_adjoint_params(isph_ctx* c) {
    if (c->prm.adaptive_dt) return fail(c, ISPH_E_ARG, "the adjoint needs a fixed dt (adaptive_dt = 0)");
    if (c->prm.sink) return fail(c, ISPH_E_ARG, "the adjoint does not support the outflow sink");
    if (c->prm.salinity)
        return fail(c, ISPH_E_ARG, "the adjoint does not support salinity transport (salinity = 0)");
    if (c->prm.cell_order != ISPH_ORDER_CELL)
        return fail(c, ISPH_E_ARG, "the adjoint runs with cell_order = ISPH_ORDER_CELL");
    if (c->prm.precision == ISPH_PREC_FGMRES)
        return fail(c, ISPH_E_ARG, "the adjoint needs a BiCGSTAB forward solve (not FGMRES)");
    // the backward's recomputation does not replay buffer removal/insertion
    if (c->nzones > 0)
        return fail(c, ISPH_E_ARG, "the adjoint does not support buffer zones (none may be installed)");
    return ISPH_OK;
}

}  // namespace

void adj_free(isph_ctx* c) {
    AdjointState* S = c->adj;
    if (!S) return;
    for (auto& k : S->seg) free_ckpt(k);
    free_ckpt(S->fin);
    for (auto& t : S->tape) free_tape(t);
    if (S->have_adj) {
        free_adjview(S->A[0]);
        free_adjview(S->A[1]);
        void* p[] = {S->th, S->aum, S->w, S->lam, S->valT, S->val32T, S->rsumT, S->dsc, S->part};
        for (auto q : p)
            if (q) (void)hipFree(q);
    }
    if (S->lam_pid) (void)hipFree(S->lam_pid);
    if (S->rec) (void)hipFree(S->rec);
    delete S;
    c->adj = nullptr;
}

}  // namespace host
}  // namespace isph

using namespace isph::host;

extern "C" {

int isph_adjoint_forward(isph_ctx* c, const isph_objective* obj, int32_t nsteps, int32_t every,
                         double* J) {
    if (!c || nsteps <= 0 || every <= 0) return ISPH_E_ARG;
    int rc = check_adjoint_params(c);
    if (rc) return rc;
    AdjointState* S = state(c);
    S->ready = false;
    if ((rc = make_obj(c, obj, S->oc))) return rc;
    if ((rc = ensure_adj(c))) return rc;
    const int nseg = (nsteps + every - 1) / every;
    if ((int)S->seg.size() < nseg) S->seg.resize(nseg);
    HIPCHK(c, hipMemsetAsync(S->dsc, 0, sizeof(double) * (8 + 3 * OBJ_MAXP), c->stream));
    // porous closure and umax of the initial state (main.f90:35-37)
    if ((rc = run_effpor(c)) || (rc = run_timestep(c)) || (rc = sync_check(c))) return rc;
    // room for every step's PPE solution; without it the backward solves
    // its recomputation in full
    const int64_t need = (int64_t)nsteps * std::max(c->n, 1);
    if (S->rec_cap < need) {
        if (S->rec) (void)hipFree(S->rec);
        S->rec = nullptr;
        S->rec_cap = 0;
        if (hipMalloc((void**)&S->rec, sizeof(double) * (size_t)need) == hipSuccess)
            S->rec_cap = need;
        else
            (void)hipGetLastError();  // out of memory is not an error here
    }
    S->rec_n = c->n;
    const bool record = S->rec != nullptr;
    const double weight = 1.0 / nsteps;
    for (int k = 0; k < nsteps; ++k) {
        if (k % every == 0 && (rc = save_ckpt(c, S->seg[k / every]))) return rc;
        c->solve_record = record ? S->rec + (size_t)k * c->n : nullptr;
        rc = run_step(c);
        c->solve_record = nullptr;
        if (rc) {
            prof_collect(c);
            return rc;
        }
        if ((rc = obj_sums(c, S->oc, c->st[c->cur], weight, true))) return rc;
        if (c->progress && ((k + 1) % every == 0 || k + 1 == nsteps)) {
            if ((rc = sync_check(c))) return rc;
            c->progress(c->progress_user, "adjoint_forward", k / every + 1, nseg);
        }
    }
    prof_collect(c);
    if ((rc = sync_check(c))) return rc;
    if ((rc = save_ckpt(c, S->fin))) return rc;
    HIPCHK(c, hipMemcpyAsync(&S->J, S->dsc + 1, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    S->nsteps = nsteps;
    S->every = every;
    S->prm = c->prm;
    S->ready = true;
    if (J) *J = S->J;
    return ISPH_OK;
}

int isph_adjoint_backward(isph_ctx* c, double* grad3, double* J, int64_t* adj_iters) {
    if (!c || !grad3) return ISPH_E_ARG;
    AdjointState* S = c->adj;
    if (!S || !S->ready) return fail(c, ISPH_E_STATE, "isph_adjoint_backward needs isph_adjoint_forward first");
    if (memcmp(&S->prm, &c->prm, sizeof(isph_params)) != 0)
        return fail(c, ISPH_E_STATE, "parameters changed since isph_adjoint_forward");
    if (c->nzones > 0)
        return fail(c, ISPH_E_ARG, "the adjoint does not support buffer zones (none may be installed)");
    int rc;
    if ((rc = zero_all(c))) return rc;
    // the transposed solves' warm starts by particle id, zero at the start.
    // A sparse id space (max pid far above the particle count) would need a
    // table far larger than the state: those solves start from zero instead.
    const int64_t npid = std::max<int64_t>(c->next_pid, 1);
    S->warm_npid = 0;
    if (npid <= 4 * (int64_t)c->cap + 1024) {
        if (S->npid < npid) {
            if (S->lam_pid) (void)hipFree(S->lam_pid);
            S->lam_pid = nullptr;
            S->npid = 0;
            if (dalloc(c, &S->lam_pid, (size_t)npid))
                return fail(c, ISPH_E_HIP, "adjoint buffer allocation failed");
            S->npid = (int)npid;
        }
        k_zero<<<nblk(npid), BLOCK, 0, c->stream>>>((int)npid, S->lam_pid);
        S->warm_npid = (int)npid;
    }
    const int every = S->every, nsteps = S->nsteps;
    const int nseg = (nsteps + every - 1) / every;
    const double weight = 1.0 / nsteps;
    int64_t iters = 0;
    if ((int)S->tape.size() < every) S->tape.resize(every);
    for (int g = nseg - 1; g >= 0; --g) {
        const int k0 = g * every, m = std::min(every, nsteps - k0);
        if ((rc = load_ckpt(c, S->seg[g]))) return rc;
        for (int s = 0; s < m; ++s) {
            if ((rc = alloc_tape(c, S->tape[s]))) return rc;
            const bool replay = S->rec && c->n == S->rec_n;
            c->solve_replay = replay ? S->rec + (size_t)(k0 + s) * c->n : nullptr;
            rc = tape_step(c, S->tape[s]);
            c->solve_replay = nullptr;
            if (rc) return rc;
        }
        for (int s = m - 1; s >= 0; --s) {
            const StateView& end = s + 1 < m ? S->tape[s + 1].Sa
                                             : (g + 1 < nseg ? S->seg[g + 1].v : S->fin.v);
            if ((rc = reverse_step(c, S->tape[s], end, weight, &iters))) return rc;
        }
        if (c->progress) c->progress(c->progress_user, "adjoint_backward", nseg - g, nseg);
    }
    // initial porous closure
    if ((rc = adj_effpor(c, S->A[S->ca], S->seg[0].v))) return rc;
    const int n = c->n;
    for (int q = 0; q < 3; ++q)
        if ((rc = dsum(c, n, S->th + (size_t)q * c->cap, S->dsc + 26 + q))) return rc;
    double g3[3];
    HIPCHK(c, hipMemcpyAsync(g3, S->dsc + 26, sizeof(g3), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((rc = sync_check(c))) return rc;
    if ((rc = load_ckpt(c, S->fin))) return rc;
    for (int q = 0; q < 3; ++q) grad3[q] = g3[q];
    if (J) *J = S->J;
    if (adj_iters) *adj_iters = iters;
    return ISPH_OK;
}

// ---- stage adjoints on the current state ------------------------------------
int isph_adjoint_zero(isph_ctx* c) {
    if (!c) return ISPH_E_ARG;
    int rc = ensure_adj(c);
    if (!rc) rc = zero_all(c);
    if (!rc) rc = sync_check(c);
    return rc;
}

static int adj_field(isph_ctx* c, int32_t f, double** p, int* width, int* comp) {
    AdjointState* S = state(c);
    const AdjView& a = S->A[S->ca];
    *width = 1;
    *comp = -1;
    switch (f) {
        case ISPH_FIELD_XS: *p = (double*)a.sp; *comp = 0; return ISPH_OK;
        case ISPH_FIELD_YS: *p = (double*)a.sp; *comp = 1; return ISPH_OK;
        case ISPH_FIELD_VXS: *p = (double*)a.sv; *comp = 0; return ISPH_OK;
        case ISPH_FIELD_VYS: *p = (double*)a.sv; *comp = 1; return ISPH_OK;
        case ISPH_FIELD_POROSITY: *p = a.por; return ISPH_OK;
        case ISPH_FIELD_RESIST: *p = a.R; return ISPH_OK;
        case ISPH_FIELD_COFF: *p = a.L; *width = 4; return ISPH_OK;
        case ISPH_FIELD_NUT: *p = a.nut; return ISPH_OK;
        case ISPH_FIELD_TKE: *p = a.tke; return ISPH_OK;
        case ISPH_FIELD_VARTS: *p = a.varts; return ISPH_OK;
        case ISPH_FIELD_VEL: *p = a.velm; return ISPH_OK;
        case ISPH_FIELD_X: *p = (double*)a.pos; *comp = 0; return ISPH_OK;
        case ISPH_FIELD_Y: *p = (double*)a.pos; *comp = 1; return ISPH_OK;
        case ISPH_FIELD_VX: *p = (double*)a.vel; *comp = 0; return ISPH_OK;
        case ISPH_FIELD_VY: *p = (double*)a.vel; *comp = 1; return ISPH_OK;
        case ISPH_FIELD_DENSITY: *p = a.rho; return ISPH_OK;
        case ISPH_FIELD_PRESSURE: *p = a.p; return ISPH_OK;
        default: return fail(c, ISPH_E_ARG, "field has no adjoint");
    }
}

int isph_adjoint_get(isph_ctx* c, int32_t f, double* out) {
    if (!c || !out) return ISPH_E_ARG;
    int rc = ensure_adj(c);
    if (rc) return rc;
    double* p;
    int w, comp;
    if ((rc = adj_field(c, f, &p, &w, &comp))) return rc;
    const int n = c->n;
    if (comp >= 0) {
        std::vector<double> t(2 * (size_t)n);
        HIPCHK(c, hipMemcpyAsync(t.data(), p, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int i = 0; i < n; ++i) out[i] = t[2 * (size_t)i + comp];
        return ISPH_OK;
    }
    HIPCHK(c, hipMemcpyAsync(out, p, sizeof(double) * w * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ISPH_OK;
}

int isph_adjoint_set(isph_ctx* c, int32_t f, const double* in) {
    if (!c || !in) return ISPH_E_ARG;
    int rc = ensure_adj(c);
    if (rc) return rc;
    double* p;
    int w, comp;
    if ((rc = adj_field(c, f, &p, &w, &comp))) return rc;
    const int n = c->n;
    if (comp >= 0) {
        std::vector<double> t(2 * (size_t)n);
        HIPCHK(c, hipMemcpyAsync(t.data(), p, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int i = 0; i < n; ++i) t[2 * (size_t)i + comp] = in[i];
        HIPCHK(c, hipMemcpyAsync(p, t.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice, c->stream));
    } else {
        HIPCHK(c, hipMemcpyAsync(p, in, sizeof(double) * w * n, hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ISPH_OK;
}

int isph_adjoint_params(isph_ctx* c, double* grad3, double* umax_bar) {
    if (!c) return ISPH_E_ARG;
    int rc = ensure_adj(c);
    if (rc) return rc;
    AdjointState* S = state(c);
    const int n = c->n;
    for (int q = 0; q < 3; ++q)
        if ((rc = dsum(c, n, S->th + (size_t)q * c->cap, S->dsc + 26 + q))) return rc;
    if ((rc = dsum(c, n, S->aum, S->dsc + 29))) return rc;
    double g[4];
    HIPCHK(c, hipMemcpyAsync(g, S->dsc + 26, sizeof(g), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (grad3)
        for (int q = 0; q < 3; ++q) grad3[q] = g[q];
    if (umax_bar) *umax_bar = g[3];
    return sync_check(c);
}

int isph_objective_eval(isph_ctx* c, const isph_objective* obj, double* J) {
    if (!c || !J) return ISPH_E_ARG;
    int rc = ensure_adj(c);
    if (rc) return rc;
    AdjointState* S = state(c);
    ObjConst oc;
    if ((rc = make_obj(c, obj, oc))) return rc;
    HIPCHK(c, hipMemsetAsync(S->dsc + 1, 0, sizeof(double), c->stream));
    if ((rc = obj_sums(c, oc, c->st[c->cur], 1.0, true))) return rc;
    HIPCHK(c, hipMemcpyAsync(J, S->dsc + 1, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return sync_check(c);
}

int isph_adjoint_stage(isph_ctx* c, int32_t stage, const isph_objective* obj) {
    if (!c) return ISPH_E_ARG;
    int rc = ensure_adj(c);
    if (rc) return rc;
    AdjointState* S = state(c);
    const AdjView& a = S->A[S->ca];
    StateView& s = c->st[c->cur];
    const int n = c->n;
    const bool lists = stage == ISPH_ADJ_NEIGHBOUR || stage == ISPH_ADJ_EDDYVIS ||
                       stage == ISPH_ADJ_VARTS || stage == ISPH_ADJ_VSTAR ||
                       stage == ISPH_ADJ_PPESOLVE || stage == ISPH_ADJ_COMP_VEL ||
                       stage == ISPH_ADJ_SHIFT;
    if (lists && !c->have_nbr) return fail(c, ISPH_E_STATE, "stage adjoint needs neighbour lists");
    if (n == 0) return ISPH_OK;
    switch (stage) {
        case ISPH_ADJ_PROJECTION:
            k_adj_projection<<<nblk(n), BLOCK, 0, c->stream>>>(n, s, c->ss, a);
            break;
        case ISPH_ADJ_NEIGHBOUR: rc = adj_kgc(c, a, s, c->nbr, c->cnt); break;
        case ISPH_ADJ_EFFPOR: rc = adj_effpor(c, a, s); break;
        case ISPH_ADJ_EDDYVIS: rc = adj_strain(c, a, s, c->nbr, c->cnt, c->L, c->ss, 1); break;
        case ISPH_ADJ_VARTS: rc = adj_strain(c, a, s, c->nbr, c->cnt, c->L, c->ss, 0); break;
        case ISPH_ADJ_VSTAR: rc = adj_vstar(c, a, s, c->nbr, c->cnt, c->L, c->ss); break;
        case ISPH_ADJ_PPESOLVE: {
            // p: the pressure of the forward solve (the adjoint leaves s.p alone).
            // The transposed solve starts from zero: the by-pid warm start
            // belongs to the backward sweep, not to a lone stage.
            int64_t it = 0;
            S->warm_npid = 0;
            rc = adj_ppe(c, a, s, c->nbr, c->cnt, c->L, s.p, c->ss, &it);
            break;
        }
        case ISPH_ADJ_COMP_VEL: rc = adj_comp_vel(c, a, s, c->nbr, c->cnt, c->L, s.p, c->ss); break;
        case ISPH_ADJ_COMP_POS:
            k_adj_comp_pos<<<nblk(n), BLOCK, 0, c->stream>>>(n, s, c->ss, a);
            break;
        case ISPH_ADJ_SHIFT: rc = adj_shift(c, a, s, c->nbr, c->cnt, c->L); break;
        case ISPH_ADJ_OBJECTIVE: {
            ObjConst oc;
            if ((rc = make_obj(c, obj, oc))) break;
            if ((rc = obj_sums(c, oc, s, 1.0, false))) break;
            k_adj_obj<<<nblk(n), BLOCK, 0, c->stream>>>(n, oc, s, S->dsc + 2, 1.0, a);
            break;
        }
        default: return fail(c, ISPH_E_ARG, "unknown adjoint stage");
    }
    if (rc) return rc;
    HIPCHK(c, hipGetLastError());
    return sync_check(c);
}

}  // extern "C"
