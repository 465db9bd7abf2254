/*
 * isph_oracle.c -- CPU restatement of the reference ISPH step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, smoke() and the
 * cpu_baseline leg of bench.py load this library; nothing in the product
 * links or calls it.
 *
 * Each routine restates one reference subroutine; the citation in front of
 * it names the lines it follows.  Expressions keep the reference's
 * evaluation order as Fortran parses it (left to right within one
 * precedence level, `x**(-1)` as 1/x, `x**2` as x*x, `x**3` as (x*x)*x,
 * `x**4` as (x*x)*(x*x)), and the default-real literals keep their float32
 * values (28.0*3.14, 3.0*3.14, 1e-5, 1e-6).  The file is compiled with
 * -ffp-contract=off so no multiply-add is fused.
 *
 * Where the reference depends on an order the GPU chooses differently the
 * restatement follows the GPU order, which is fixed and documented:
 *   - particles are kept sorted by (cell, fluid bit, previous position),
 *     which plays the role of the reference's cellshift+resetid order;
 *   - neighbour lists are ascending in that order (the reference's lists
 *     ascend in matid after resetid, isph.f90:167-176);
 *   - dot products use the fixed tree of ora_dot_tree() (the reference's
 *     OpenMP reduction order depends on the thread count, solver.f90:17).
 * Reference bugs that are not reproduced (documented in the build notes):
 *   - the shared `num2` loop bound race (isph.f90:159-162): each row uses
 *     its own neighbour count;
 *   - 0/0 in the shift clamp (part_shift.f90:90-91,113-114): a zero shift
 *     stays zero;
 *   - fluid particles skipped by opt2_shift keep xs = old position and jump
 *     by it (part_shift.f90:40,180): their shift is zero.
 * Parity pinning: tests/test_oracle_pinning.py and its neighbours check
 * every routine against closed forms, brute force, scipy or the reference's
 * known answers; no routine is left "parity unpinned".
 */
#include "isph_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* float32 values of the reference's default-real literals */
#define C_HYB1 87.92000579833984375       /* 28.0*3.14, kernel.f90:61 */
#define C_HYB2 9.4200000762939453125      /* 3.0*3.14,  kernel.f90:62 */
#define ALP 0.1                           /* kernel.f90:30 */
#define TINY_DET 9.999999747378751635551453e-06  /* 1e-5, functions.f90:75 */
#define TINY6 9.999999974752427078783512e-07     /* 1e-6, integrator.f90:497,582 */
#define DM_SALT 9.9999997171806853657471947e-10   /* 1e-9, scalar.f90:118, integrator.f90:914 */
#define MAX_SUBSTEPS (1 << 20)             /* cap on transport substeps per step */
#define RBLOCK 256                        /* reduction block of the GPU kernels */

struct ora_ctx {
    isph_params prm;
    int32_t cap, n, cap_soil, ns;
    double *x, *y, *vx, *vy, *xs, *ys, *vxs, *vys, *mass, *rho, *oden, *p;
    double *por, *R, *nut, *tke, *varts, *vel, *L, *divr;
    double *con, *ovol, *cdiff;   /* salinity, particle volume, d con/dt */
    double *xnb, *ynb;            /* positions at the last neighbour search */
    int32_t *tid, *flags, *pid;
    /* cell grid */
    int32_t nx, ny;
    int32_t per;            /* bins per cell: 2 (cell order) or 10 (reference order) */
    int64_t nbins;          /* per * nx * ny */
    int32_t *key;           /* sorted key per particle */
    int32_t *bstart;        /* nbins + 2 */
    int32_t *perm;
    double *tmpd;
    int32_t *tmpi;
    /* neighbours */
    int32_t *nptr, *nidx;
    int have_nbr;
    /* soil, sorted by cell */
    double *sx, *sy, *sm, *sr;
    int32_t *sstart;        /* ncells + 1 */
    /* PPE */
    int32_t *rowptr, *col;
    double *val, *rhs;
    int assembled;
    double *fsol;           /* GMRES warm start and result in row order */
    double *w[10];
    float *f[8];
    double umax, numax, dt, t, relres, dtsol;
    int32_t iters, outer, solsteps;
    /* buffer zones (internal.f90:475-614): the template particles inside the
     * entry cells in insertion order, a per-cell entry mask, the schedule */
    struct {
        int32_t nins, remove_right;
        double *x, *y, *vx, *vy, *p;
        int32_t *flags, *key;
        unsigned char* mask;
        double mass, density, ovol, con, period, xcut;
        int64_t counter;
    } zone[ISPH_MAX_ZONES];
    int32_t nzones, next_pid;
    double zone_shift;
    char err[256];
};

static int set_err(ora_ctx* c, int code, const char* msg) {
    snprintf(c->err, sizeof(c->err), "%s", msg);
    return code;
}

const char* ora_last_error(const ora_ctx* c) { return c ? c->err : "null context"; }

int ora_set_threads(int32_t nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
    return omp_get_max_threads();
#else
    (void)nthreads;
    return 1;
#endif
}

/* ---------------------------------------------------------------------- */
/* kernel.f90                                                              */
/* ---------------------------------------------------------------------- */

/* Wab (kernel.f90:34-66) hybrid cubic-spline/hyperbolic kernel; Wendland C2
 * from the commented form kernel.f90:42-43. */
double ora_kernel_w(const isph_params* prm, double r) {
    const double h = prm->h;
    const double h2 = h * h;
    if (prm->kernel == ISPH_KERNEL_WENDLAND_C2) {
        double w = 1.0 - r / (2.0 * h);
        double w2 = w * w;
        return ((7.0 / ((4.0 * 3.140) * h2)) * (1.0 + (2.0 * r) / h)) * (w2 * w2);
    }
    double t1, t2;
    if (r < h) {
        double q = r / h;
        t1 = (4.0 - (6.0 * (r * r)) / h2) + (3.0 * ((r * r) * r)) / (h2 * h);
        t2 = (((q * q) * q) - (6.0 * r) / h) + 6.0;
    } else {
        double u = 2.0 - r / h;
        t1 = (u * u) * u;
        t2 = (u * u) * u;
    }
    return (ALP * t1) * (10.0 / (C_HYB1 * h2)) + ((1.0 - ALP) * t2) * (1.0 / (C_HYB2 * h2));
}

/* Wo (kernel.f90:68-92): W(0). */
double ora_kernel_w0(const isph_params* prm) {
    const double h = prm->h;
    const double h2 = h * h;
    if (prm->kernel == ISPH_KERNEL_WENDLAND_C2) return 7.0 / ((4.0 * 3.140) * h2);
    return (ALP * 4.0) * (10.0 / (C_HYB1 * h2)) + ((1.0 - ALP) * 6.0) * (1.0 / (C_HYB2 * h2));
}

/* Wabx/Waby (kernel.f90:94-162) called as Wabx(p_j, p_i, r, h): the gradient
 * with respect to x_i.  dx = x_i - x_j, dy = y_i - y_j.  Wendland C2 from
 * kernel.f90:102-103,137-138. */
void ora_kernel_gradw(const isph_params* prm, double r, double dx, double dy,
                      double* wx, double* wy) {
    const double h = prm->h;
    const double h2 = h * h;
    const double mdx = -dx, mdy = -dy; /* p1%x - p2%x = x_j - x_i */
    if (prm->kernel == ISPH_KERNEL_WENDLAND_C2) {
        double w = 1.0 - r / (2.0 * h);
        double w3 = (w * w) * w;
        double den = (4.0 * 3.140) * (h2 * h2);
        *wx = ((35.0 * mdx) * w3) / den;
        *wy = ((35.0 * mdy) * w3) / den;
        return;
    }
    const double h3 = h2 * h;
    double ax1, ax2;
    double tx1, tx2, ty1, ty2;
    if (r < h) {
        ax1 = (9.0 * r) / h3 - 12.0 / h2;
        ax2 = (3.0 * r) / h3 - 6.0 / (r * h);
        tx1 = ax1 * dx; tx2 = ax2 * dx;
        ty1 = ax1 * dy; ty2 = ax2 * dy;
    } else {
        double u = 2.0 - r / h;
        double a = (3.0 / (r * h)) * (u * u);
        tx1 = a * mdx; tx2 = a * mdx;
        ty1 = a * mdy; ty2 = a * mdy;
    }
    const double n1 = 10.0 / (C_HYB1 * h2), n2 = 1.0 / (C_HYB2 * h2);
    *wx = (ALP * tx1) * n1 + ((1.0 - ALP) * tx2) * n2;
    *wy = (ALP * ty1) * n1 + ((1.0 - ALP) * ty2) * n2;
}

/* ---------------------------------------------------------------------- */
/* functions.f90                                                           */
/* ---------------------------------------------------------------------- */

/* invertmat2D (functions.f90:62-96): a = [a1 a2; a3 a4] stored row-wise. */
void ora_invert2(double* a) {
    double det = a[0] * a[3] - a[1] * a[2];
    if (fabs(det) >= TINY_DET && isfinite(det)) {
        double r0 = a[3] / det, r1 = (-a[1]) / det, r2 = (-a[2]) / det, r3 = a[0] / det;
        a[0] = r0; a[1] = r1; a[2] = r2; a[3] = r3;
    } else {
        a[0] = 1.0; a[1] = 0.0; a[2] = 0.0; a[3] = 1.0;
    }
}

/* vis (porous.f90:12-30, integrator.f90:12-30) dynamic viscosity of
 * saline water; visc_law 1 uses mu = nu*oden. */
static inline double visc(const isph_params* prm, double oden) {
    if (prm->visc_law == 1) return prm->nu * oden;
    double sp = oden - 1000.0;
    return ((1.002 + 0.001652 * sp) + 0.0000083 * (sp * sp)) * 0.001;
}

/* ---------------------------------------------------------------------- */
/* fixed reduction tree shared with the GPU kernels                        */
/* ---------------------------------------------------------------------- */

/* One 64-lane wavefront butterfly (xor 1,2,...,32) then the 4 waves of a
 * 256-thread block as (w0+w1)+(w2+w3). */
static double block_tree(double* v /* RBLOCK, destroyed */) {
    double w[4];
    for (int k = 0; k < 4; ++k) {
        double* a = v + 64 * k;
        for (int s = 1; s < 64; s <<= 1)
            for (int l = 0; l < 64; l += 2 * s) a[l] = a[l] + a[l + s];
        w[k] = a[0];
    }
    return (w[0] + w[1]) + (w[2] + w[3]);
}

/* The block partials are finished by one block of FBLOCK threads: thread t
 * sums partials t, t+FBLOCK, ... in order, then each wave's butterfly, then
 * the 16 wave sums as a balanced pairwise tree. */
#define FBLOCK 1024
static double tree_finish(const double* part, int64_t nb) {
    static double v[FBLOCK];
    double w[FBLOCK / 64];
    for (int t = 0; t < FBLOCK; ++t) {
        double s = 0.0;
        for (int64_t k = t; k < nb; k += FBLOCK) s += part[k];
        v[t] = s;
    }
    for (int k = 0; k < FBLOCK / 64; ++k) {
        double* a = v + 64 * k;
        for (int st = 1; st < 64; st <<= 1)
            for (int l = 0; l < 64; l += 2 * st) a[l] = a[l] + a[l + st];
        w[k] = a[0];
    }
    for (int st = 1; st < FBLOCK / 64; st <<= 1)
        for (int l = 0; l < FBLOCK / 64; l += 2 * st) w[l] = w[l] + w[l + st];
    return w[0];
}

/* the GPU grid of a reduction is a multiple of 8 blocks (one share per
 * XCD); the extra blocks contribute zero partials */
static int64_t nblocks8(int64_t n) { return ((n + RBLOCK - 1) / RBLOCK + 7) / 8 * 8; }

double ora_dot_tree(const double* a, const double* b, int64_t n) {
    int64_t nb = nblocks8(n);
    if (nb == 0) return 0.0;
    double* part = (double*)malloc(sizeof(double) * nb);
#pragma omp parallel for schedule(static)
    for (int64_t blk = 0; blk < nb; ++blk) {
        double v[RBLOCK];
        for (int t = 0; t < RBLOCK; ++t) {
            int64_t i = blk * RBLOCK + t;
            v[t] = (i < n) ? a[i] * b[i] : 0.0;
        }
        part[blk] = block_tree(v);
    }
    double r = tree_finish(part, nb);
    free(part);
    return r;
}

/* same tree over float vectors, products formed in double (exact) */
static double dot_tree_f(const float* a, const float* b, int64_t n) {
    int64_t nb = nblocks8(n);
    if (nb == 0) return 0.0;
    double* part = (double*)malloc(sizeof(double) * nb);
#pragma omp parallel for schedule(static)
    for (int64_t blk = 0; blk < nb; ++blk) {
        double v[RBLOCK];
        for (int t = 0; t < RBLOCK; ++t) {
            int64_t i = blk * RBLOCK + t;
            v[t] = (i < n) ? (double)a[i] * (double)b[i] : 0.0;
        }
        part[blk] = block_tree(v);
    }
    double r = tree_finish(part, nb);
    free(part);
    return r;
}

/* ---------------------------------------------------------------------- */
/* context                                                                 */
/* ---------------------------------------------------------------------- */

static void grid_dims(const isph_params* prm, int32_t* nx, int32_t* ny) {
    *nx = (int32_t)floor((prm->xmax - prm->xmin) / prm->cell_size) + 3;
    *ny = (int32_t)floor((prm->ymax - prm->ymin) / prm->cell_size) + 3;
}

static int check_params(ora_ctx* c, const isph_params* prm) {
    if (!(prm->h > 0.0) || !(prm->cell_size >= 2.0 * prm->h) ||
        !(prm->xmax > prm->xmin) || !(prm->ymax > prm->ymin) || prm->max_nbr <= 0)
        return set_err(c, ISPH_E_ARG, "invalid geometry parameters");
    if (prm->cell_order != ISPH_ORDER_CELL && prm->cell_order != ISPH_ORDER_REFERENCE)
        return set_err(c, ISPH_E_ARG, "unknown cell_order");
    if (prm->cell_order == ISPH_ORDER_REFERENCE && !(prm->cell_size >= 4.0 * prm->h))
        return set_err(c, ISPH_E_ARG, "the reference order needs cell_size >= 4h");
    if (prm->precision < ISPH_PREC_FP64 || prm->precision > ISPH_PREC_FP64_MG)
        return set_err(c, ISPH_E_ARG, "unknown precision");
    int32_t nx, ny;
    grid_dims(prm, &nx, &ny);
    if ((double)nx * (double)ny * 2.0 > 2.0e9) return set_err(c, ISPH_E_ARG, "cell grid too large");
    return ISPH_OK;
}

static void* xmalloc(size_t n) { return calloc(n ? n : 1, 1); }

int ora_create(const isph_params* prm, int32_t cap, int32_t cap_soil, ora_ctx** out) {
    if (!prm || !out || cap <= 0 || cap_soil < 0) return ISPH_E_ARG;
    ora_ctx* c = (ora_ctx*)xmalloc(sizeof(ora_ctx));
    int rc = check_params(c, prm);
    if (rc) { free(c); return rc; }
    c->prm = *prm;
    c->cap = cap;
    c->cap_soil = cap_soil;
    size_t nd = sizeof(double) * (size_t)cap;
    double** dp[] = {&c->x, &c->y, &c->vx, &c->vy, &c->xs, &c->ys, &c->vxs, &c->vys,
                     &c->mass, &c->rho, &c->oden, &c->p, &c->por, &c->R, &c->nut,
                     &c->tke, &c->varts, &c->vel, &c->divr, &c->tmpd, &c->rhs,
                     &c->con, &c->ovol, &c->cdiff, &c->xnb, &c->ynb, &c->fsol};
    for (size_t k = 0; k < sizeof(dp) / sizeof(dp[0]); ++k) *dp[k] = (double*)xmalloc(nd);
    c->L = (double*)xmalloc(4 * nd);
    for (int k = 0; k < 10; ++k) c->w[k] = (double*)xmalloc(nd);
    for (int k = 0; k < 8; ++k) c->f[k] = (float*)xmalloc(sizeof(float) * (size_t)cap);
    size_t ni = sizeof(int32_t) * (size_t)cap;
    c->tid = (int32_t*)xmalloc(ni);
    c->flags = (int32_t*)xmalloc(ni);
    c->pid = (int32_t*)xmalloc(ni);
    c->key = (int32_t*)xmalloc(ni);
    c->perm = (int32_t*)xmalloc(ni);
    c->tmpi = (int32_t*)xmalloc(ni);
    c->nptr = (int32_t*)xmalloc(ni + sizeof(int32_t));
    c->nidx = (int32_t*)xmalloc(ni * (size_t)prm->max_nbr);
    c->rowptr = (int32_t*)xmalloc(ni + sizeof(int32_t));
    c->col = (int32_t*)xmalloc(ni * (size_t)(prm->max_nbr + 1));
    c->val = (double*)xmalloc(nd * (size_t)(prm->max_nbr + 1));
    grid_dims(prm, &c->nx, &c->ny);
    c->per = prm->cell_order ? 10 : 2;
    c->nbins = c->per * (int64_t)c->nx * c->ny;
    c->bstart = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)(c->nbins + 2));
    c->sstart = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)((int64_t)c->nx * c->ny + 1));
    size_t ns = sizeof(double) * (size_t)(cap_soil ? cap_soil : 1);
    c->sx = (double*)xmalloc(ns);
    c->sy = (double*)xmalloc(ns);
    c->sm = (double*)xmalloc(ns);
    c->sr = (double*)xmalloc(ns);
    c->dt = prm->dt;
    c->dtsol = prm->dt;
    *out = c;
    return ISPH_OK;
}

static void free_zones(ora_ctx* c) {
    for (int k = 0; k < ISPH_MAX_ZONES; ++k) {
        free(c->zone[k].x); free(c->zone[k].y); free(c->zone[k].vx); free(c->zone[k].vy);
        free(c->zone[k].p); free(c->zone[k].flags); free(c->zone[k].key); free(c->zone[k].mask);
        memset(&c->zone[k], 0, sizeof(c->zone[k]));
    }
    c->nzones = 0;
}

int ora_destroy(ora_ctx* c) {
    if (!c) return ISPH_OK;
    free_zones(c);
    double* dp[] = {c->x, c->y, c->vx, c->vy, c->xs, c->ys, c->vxs, c->vys, c->mass,
                    c->rho, c->oden, c->p, c->por, c->R, c->nut, c->tke, c->varts,
                    c->vel, c->divr, c->tmpd, c->rhs, c->L, c->val, c->sx, c->sy,
                    c->sm, c->sr, c->con, c->ovol, c->cdiff, c->xnb, c->ynb, c->fsol};
    for (size_t k = 0; k < sizeof(dp) / sizeof(dp[0]); ++k) free(dp[k]);
    for (int k = 0; k < 10; ++k) free(c->w[k]);
    for (int k = 0; k < 8; ++k) free(c->f[k]);
    int32_t* ip[] = {c->tid, c->flags, c->pid, c->key, c->perm, c->tmpi, c->nptr,
                     c->nidx, c->rowptr, c->col, c->bstart, c->sstart};
    for (size_t k = 0; k < sizeof(ip) / sizeof(ip[0]); ++k) free(ip[k]);
    free(c);
    return ISPH_OK;
}

int ora_set_params(ora_ctx* c, const isph_params* prm) {
    int32_t nx, ny;
    int rc = check_params(c, prm);
    if (rc) return rc;
    grid_dims(prm, &nx, &ny);
    if (nx != c->nx || ny != c->ny || prm->max_nbr != c->prm.max_nbr ||
        prm->cell_order != c->prm.cell_order)
        return set_err(c, ISPH_E_ARG, "grid, cell order and max_nbr are fixed at create time");
    c->prm = *prm;
    return ISPH_OK;
}

/* cell of a point: 1-cell halo around [xmin,xmax]x[ymin,ymax]; points outside
 * are clamped into the edge cells and reported. */
static inline int32_t cell_of(const ora_ctx* c, double x, double y, int* oob) {
    const isph_params* prm = &c->prm;
    double fx = floor((x - prm->xmin) / prm->cell_size);
    double fy = floor((y - prm->ymin) / prm->cell_size);
    int32_t cx, cy;
    if (!(fx >= 0.0)) { cx = 1; *oob = 1; }
    else if (fx > (double)(c->nx - 3)) { cx = c->nx - 2; *oob = 1; }
    else cx = (int32_t)fx + 1;
    if (!(fy >= 0.0)) { cy = 1; *oob = 1; }
    else if (fy > (double)(c->ny - 3)) { cy = c->ny - 2; *oob = 1; }
    else cy = (int32_t)fy + 1;
    return cx * c->ny + cy;
}

/* Bin layout.  Cell order: bin = 2 (cx ny + cy) + fluid bit.  Reference
 * order: the reference's resetid numbering (isph.f90:96-109: cells column by
 * column, row i = 1 at the top, so y descending) times 10 plus a class:
 * 0 walls, 1 fluid that stayed, 2..9 the arrivals in the gather order of
 * integrator.f90:318-389 (from below, above, right, left, upper left, lower
 * left, upper right, lower right). */
static inline int32_t cell_rank(const ora_ctx* c, int32_t cx, int32_t cy) {
    return c->prm.cell_order ? cx * c->ny + (c->ny - 1 - cy) : cx * c->ny + cy;
}

static inline void key_cell(const ora_ctx* c, int32_t key, int32_t* cx, int32_t* cy) {
    if (c->prm.cell_order) {
        int32_t r = key / 10;
        *cx = r / c->ny;
        *cy = c->ny - 1 - (r - *cx * c->ny);
    } else {
        *cx = (key >> 1) / c->ny;
        *cy = (key >> 1) - *cx * c->ny;
    }
}

/* first and one-past-last bin of the cells ylo..yhi of column cx */
static inline void column_bins(const ora_ctx* c, int32_t cx, int32_t ylo, int32_t yhi,
                               int32_t* lo, int32_t* hi) {
    if (c->prm.cell_order) {
        *lo = 10 * (cx * c->ny + (c->ny - 1 - yhi));
        *hi = 10 * (cx * c->ny + (c->ny - 1 - ylo)) + 10;
    } else {
        *lo = 2 * (cx * c->ny + ylo);
        *hi = 2 * (cx * c->ny + yhi) + 2;
    }
}

/* cellshift's transfer list of a fluid particle that crossed (dx, dy)
 * cells (integrator.f90:271-310), as the class of its arrival */
static inline int32_t move_class(int32_t dx, int32_t dy) {
    static const int32_t t[3][3] = {{8, 3, 6}, {4, 1, 5}, {9, 2, 7}};
    return t[dy + 1][dx + 1];
}

/* permute every per-particle array: new[i] = old[perm[i]] */
static void apply_perm(ora_ctx* c, int32_t nnew) {
    double* dp[] = {c->x, c->y, c->vx, c->vy, c->xs, c->ys, c->vxs, c->vys, c->mass,
                    c->rho, c->oden, c->p, c->por, c->R, c->nut, c->tke, c->varts,
                    c->vel, c->con, c->ovol};
    for (size_t k = 0; k < sizeof(dp) / sizeof(dp[0]); ++k) {
        double* a = dp[k];
        for (int32_t i = 0; i < nnew; ++i) c->tmpd[i] = a[c->perm[i]];
        memcpy(a, c->tmpd, sizeof(double) * (size_t)nnew);
    }
    int32_t* ip[] = {c->tid, c->flags, c->pid};
    for (size_t k = 0; k < sizeof(ip) / sizeof(ip[0]); ++k) {
        int32_t* a = ip[k];
        for (int32_t i = 0; i < nnew; ++i) c->tmpi[i] = a[c->perm[i]];
        memcpy(a, c->tmpi, sizeof(int32_t) * (size_t)nnew);
    }
}

/* Stable counting sort by bin key; deleted particles get key nbins and are
 * dropped.  Plays the role of cellshift/cellshiftalt's re-binning
 * (integrator.f90:32-391) plus resetid (isph.f90:88-113).  In the reference
 * order a fluid particle goes to the cell next to its previous one in the
 * direction it crossed (one cell at most, :271-310); `first`: no previous
 * binning. */
static int sort_particles(ora_ctx* c, int with_sink, int first) {
    const isph_params* prm = &c->prm;
    int oob = 0;
    int64_t nb = c->nbins;
    int32_t* cnt = c->bstart;
    memset(cnt, 0, sizeof(int32_t) * (size_t)(nb + 2));
    for (int32_t i = 0; i < c->n; ++i) {
        int32_t k;
        int fl = (c->tid[i] == 3);
        if (with_sink && prm->sink && fl && !(c->flags[i] & ISPH_FLAG_BUFFER) &&
            c->x[i] <= prm->xlcutoff && c->y[i] > prm->ytcutoff) {
            k = (int32_t)nb;
        } else {
            int32_t cell = cell_of(c, c->x[i], c->y[i], &oob);
            if (!c->prm.cell_order) {
                k = 2 * cell + fl;
            } else {
                int32_t cx = cell / c->ny, cy = cell % c->ny, cls = fl;
                if (fl && !first) {
                    int32_t ox, oy;
                    key_cell(c, c->key[i], &ox, &oy);
                    int32_t dx = cx - ox, dy = cy - oy;
                    dx = dx < -1 ? -1 : (dx > 1 ? 1 : dx);
                    dy = dy < -1 ? -1 : (dy > 1 ? 1 : dy);
                    cx = ox + dx;
                    cy = oy + dy;
                    cls = move_class(dx, dy);
                }
                k = 10 * cell_rank(c, cx, cy) + cls;
            }
        }
        c->tmpi[i] = k;
        cnt[k + 1] += 1;
    }
    for (int64_t b = 0; b < nb + 1; ++b) cnt[b + 1] += cnt[b];
    /* cnt[b] is now the start of bin b; place in order (stable) */
    int32_t* pos = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nb + 1));
    memcpy(pos, cnt, sizeof(int32_t) * (size_t)(nb + 1));
    for (int32_t i = 0; i < c->n; ++i) {
        int32_t k = c->tmpi[i];
        int32_t d = pos[k]++;
        c->perm[d] = i;
        c->key[d] = k;
    }
    free(pos);
    int32_t nnew = cnt[nb];
    apply_perm(c, nnew);
    c->n = nnew;
    c->have_nbr = 0;
    c->assembled = 0;
    if (oob) return set_err(c, ISPH_E_DOMAIN, "particle outside the cell grid");
    return ISPH_OK;
}

int ora_set_particles(ora_ctx* c, int32_t n, const double* x, const double* y,
                      const double* vx, const double* vy, const double* mass,
                      const double* density, const double* oden, const double* pressure,
                      const int32_t* tid, const int32_t* flags, const int32_t* pid) {
    if (n < 0 || n > c->cap) return set_err(c, ISPH_E_ARG, "n exceeds capacity");
    c->n = n;
    size_t nd = sizeof(double) * (size_t)n;
    memcpy(c->x, x, nd); memcpy(c->y, y, nd);
    memcpy(c->vx, vx, nd); memcpy(c->vy, vy, nd);
    memcpy(c->mass, mass, nd); memcpy(c->rho, density, nd);
    memcpy(c->oden, oden, nd); memcpy(c->p, pressure, nd);
    memcpy(c->tid, tid, sizeof(int32_t) * (size_t)n);
    memcpy(c->pid, pid, sizeof(int32_t) * (size_t)n);
    for (int32_t i = 0; i < n; ++i) {
        c->flags[i] = flags ? flags[i] : 0;
        c->xs[i] = 0.0; c->ys[i] = 0.0; c->vxs[i] = 0.0; c->vys[i] = 0.0;
        c->por[i] = 1.0; c->R[i] = 0.0; c->nut[i] = 0.0; c->tke[i] = 0.0;
        c->varts[i] = 0.0;
        c->con[i] = 0.0;
        c->ovol[i] = mass[i] / oden[i];  /* fmass/rho, internal.f90:103 */
        c->vel[i] = sqrt(vx[i] * vx[i] + vy[i] * vy[i]);
        c->divr[i] = 0.0;
        c->L[4 * i] = 1.0; c->L[4 * i + 1] = 0.0; c->L[4 * i + 2] = 0.0; c->L[4 * i + 3] = 1.0;
    }
    c->umax = 0.0; c->numax = 0.0; c->t = 0.0; c->dt = c->prm.dt;
    c->dtsol = c->prm.dt; c->solsteps = 0;  /* solsteps = 0, initialize.f90:56 */
    c->next_pid = 0;
    for (int32_t i = 0; i < n; ++i)
        if (pid[i] + 1 > c->next_pid) c->next_pid = pid[i] + 1;
    int rc = sort_particles(c, 0, 1);
    if (rc) return rc;
    /* fsol = fguess: the pressures in the first row order (memory.f90:113-144) */
    memcpy(c->fsol, c->p, sizeof(double) * (size_t)n);
    return ISPH_OK;
}

int ora_get_particles(ora_ctx* c, int32_t* n, double* x, double* y, double* vx,
                      double* vy, double* mass, double* density, double* oden,
                      double* pressure, int32_t* tid, int32_t* flags, int32_t* pid) {
    if (n) *n = c->n;
    size_t nd = sizeof(double) * (size_t)c->n, ni = sizeof(int32_t) * (size_t)c->n;
    if (x) memcpy(x, c->x, nd);
    if (y) memcpy(y, c->y, nd);
    if (vx) memcpy(vx, c->vx, nd);
    if (vy) memcpy(vy, c->vy, nd);
    if (mass) memcpy(mass, c->mass, nd);
    if (density) memcpy(density, c->rho, nd);
    if (oden) memcpy(oden, c->oden, nd);
    if (pressure) memcpy(pressure, c->p, nd);
    if (tid) memcpy(tid, c->tid, ni);
    if (flags) memcpy(flags, c->flags, ni);
    if (pid) memcpy(pid, c->pid, ni);
    return ISPH_OK;
}

/* soil particles sorted by cell, stable (porlist, boundary.f90:569-588) */
int ora_set_soil(ora_ctx* c, int32_t ns, const double* x, const double* y,
                 const double* mass, const double* density) {
    if (ns < 0 || ns > c->cap_soil) return set_err(c, ISPH_E_ARG, "ns exceeds soil capacity");
    int64_t nc = (int64_t)c->nx * c->ny;
    int32_t* cnt = c->sstart;
    memset(cnt, 0, sizeof(int32_t) * (size_t)(nc + 1));
    int32_t* k = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ns ? ns : 1));
    int oob = 0;
    for (int32_t i = 0; i < ns; ++i) {
        k[i] = cell_of(c, x[i], y[i], &oob);
        cnt[k[i] + 1] += 1;
    }
    for (int64_t b = 0; b < nc; ++b) cnt[b + 1] += cnt[b];
    int32_t* pos = (int32_t*)malloc(sizeof(int32_t) * (size_t)nc);
    memcpy(pos, cnt, sizeof(int32_t) * (size_t)nc);
    for (int32_t i = 0; i < ns; ++i) {
        int32_t d = pos[k[i]]++;
        c->sx[d] = x[i]; c->sy[d] = y[i]; c->sm[d] = mass[i]; c->sr[d] = density[i];
    }
    free(pos);
    free(k);
    c->ns = ns;
    if (oob) return set_err(c, ISPH_E_DOMAIN, "soil particle outside the cell grid");
    return ISPH_OK;
}

static double* field_ptr(ora_ctx* c, int32_t f, int* width) {
    *width = 1;
    switch (f) {
        case ISPH_FIELD_XS: return c->xs;
        case ISPH_FIELD_YS: return c->ys;
        case ISPH_FIELD_VXS: return c->vxs;
        case ISPH_FIELD_VYS: return c->vys;
        case ISPH_FIELD_POROSITY: return c->por;
        case ISPH_FIELD_RESIST: return c->R;
        case ISPH_FIELD_COFF: *width = 4; return c->L;
        case ISPH_FIELD_DIVR: return c->divr;
        case ISPH_FIELD_NUT: return c->nut;
        case ISPH_FIELD_TKE: return c->tke;
        case ISPH_FIELD_VARTS: return c->varts;
        case ISPH_FIELD_VEL: return c->vel;
        case ISPH_FIELD_RHS: return c->rhs;
        case ISPH_FIELD_X: return c->x;
        case ISPH_FIELD_Y: return c->y;
        case ISPH_FIELD_VX: return c->vx;
        case ISPH_FIELD_VY: return c->vy;
        case ISPH_FIELD_DENSITY: return c->rho;
        case ISPH_FIELD_PRESSURE: return c->p;
        case ISPH_FIELD_CON: return c->con;
        case ISPH_FIELD_OVOL: return c->ovol;
        case ISPH_FIELD_CDIFF: return c->cdiff;
        case ISPH_FIELD_MASS: return c->mass;
        case ISPH_FIELD_ODEN: return c->oden;
        default: return NULL;
    }
}

int ora_get_field(ora_ctx* c, int32_t f, double* out) {
    int w;
    double* a = field_ptr(c, f, &w);
    if (!a) return set_err(c, ISPH_E_ARG, "unknown field");
    memcpy(out, a, sizeof(double) * (size_t)c->n * (size_t)w);
    return ISPH_OK;
}

int ora_set_field(ora_ctx* c, int32_t f, const double* in) {
    int w;
    double* a = field_ptr(c, f, &w);
    if (!a) return set_err(c, ISPH_E_ARG, "unknown field");
    memcpy(a, in, sizeof(double) * (size_t)c->n * (size_t)w);
    return ISPH_OK;
}

int ora_get_neighbours(ora_ctx* c, int32_t* ptr, int32_t* idx) {
    if (!c->have_nbr) return set_err(c, ISPH_E_STATE, "no neighbour list");
    int32_t M = c->prm.max_nbr;
    ptr[0] = 0;
    for (int32_t i = 0; i < c->n; ++i) ptr[i + 1] = ptr[i] + c->nptr[i];
    if (idx)
        for (int32_t i = 0; i < c->n; ++i)
            memcpy(idx + ptr[i], c->nidx + (size_t)i * M, sizeof(int32_t) * (size_t)c->nptr[i]);
    return ISPH_OK;
}

int ora_get_csr(ora_ctx* c, int32_t* rowptr, int32_t* col, double* val, double* rhs) {
    if (!c->assembled) return set_err(c, ISPH_E_STATE, "no assembled matrix");
    memcpy(rowptr, c->rowptr, sizeof(int32_t) * (size_t)(c->n + 1));
    int32_t nnz = c->rowptr[c->n];
    if (col) memcpy(col, c->col, sizeof(int32_t) * (size_t)nnz);
    if (val) memcpy(val, c->val, sizeof(double) * (size_t)nnz);
    if (rhs) memcpy(rhs, c->rhs, sizeof(double) * (size_t)c->n);
    return ISPH_OK;
}

int ora_get_scalars(ora_ctx* c, double* umax, double* dt, double* t, int32_t* iters,
                    int32_t* outer, double* relres) {
    if (umax) *umax = c->umax;
    if (dt) *dt = c->dt;
    if (t) *t = c->t;
    if (iters) *iters = c->iters;
    if (outer) *outer = c->outer;
    if (relres) *relres = c->relres;
    return ISPH_OK;
}

int ora_set_scalars(ora_ctx* c, double umax, double dt, double t) {
    c->umax = umax; c->dt = dt; c->t = t;
    return ISPH_OK;
}

/* ---------------------------------------------------------------------- */
/* isph.f90:16-47  projection                                              */
/* ---------------------------------------------------------------------- */
int ora_projection(ora_ctx* c) {
    const double dt = c->dt;
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < c->n; ++i) {
        if (c->tid[i] == 3 && !(c->flags[i] & ISPH_FLAG_DEAD)) {
            c->xs[i] = c->x[i];
            c->x[i] = c->x[i] + (dt * c->vx[i]) / c->por[i];
            c->ys[i] = c->y[i];
            c->y[i] = c->y[i] + (dt * c->vy[i]) / c->por[i];
        }
    }
    return ISPH_OK;
}

int ora_cellshift(ora_ctx* c) { return sort_particles(c, 0, 0); }
int ora_cellshiftalt(ora_ctx* c) { return sort_particles(c, 1, 0); }
/* resetid (isph.f90:88-113): the sorted position is the matid already. */
int ora_resetid(ora_ctx* c) { (void)c; return ISPH_OK; }

/* ---------------------------------------------------------------------- */
/* search.f90:13-399  neighbour + nsearch2 (list, KGC, divr, free flag)    */
/* ---------------------------------------------------------------------- */
int ora_neighbour(ora_ctx* c) {
    const isph_params* prm = &c->prm;
    const double h2x = 2.0 * prm->h;
    const double thr = prm->lamfs * prm->maxdivr;
    const int32_t M = prm->max_nbr;
    int over = 0;
#pragma omp parallel for schedule(static) reduction(| : over)
    for (int32_t i = 0; i < c->n; ++i) {
        int32_t cx, cy;
        key_cell(c, c->key[i], &cx, &cy);
        int32_t cnt = 0;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        const double xi = c->x[i], yi = c->y[i];
        for (int32_t dc = -1; dc <= 1; ++dc) {
            int32_t blo, bhi;
            column_bins(c, cx + dc, cy - 1, cy + 1, &blo, &bhi);
            int32_t lo = c->bstart[blo], hi = c->bstart[bhi];
            for (int32_t j = lo; j < hi; ++j) {
                if (j == i) continue;
                double dx = xi - c->x[j], dy = yi - c->y[j];
                double r = sqrt(dx * dx + dy * dy);
                if (!(r < h2x)) continue;
                if (cnt < M) c->nidx[(size_t)i * M + cnt] = j;
                cnt++;
                double wx, wy;
                ora_kernel_gradw(prm, r, dx, dy, &wx, &wy);
                double res1 = (c->mass[j] * (c->x[j] - xi)) / c->rho[j];
                double res3 = (c->mass[j] * (c->y[j] - yi)) / c->rho[j];
                a0 = a0 + wx * res1;
                a1 = a1 + wy * res1;
                a2 = a2 + wx * res3;
                a3 = a3 + wy * res3;
            }
        }
        if (cnt > M) { over = 1; cnt = M; }
        c->nptr[i] = cnt;
        int32_t fl = c->flags[i] & ~ISPH_FLAG_FREE;
        c->divr[i] = 0.0;
        if (c->tid[i] == 3) {
            fl &= ~ISPH_FLAG_VICINITY;
            double dv = a0 + a3;
            c->divr[i] = dv;
            double a[4] = {a0, a1, a2, a3};
            ora_invert2(a);
            memcpy(c->L + 4 * (size_t)i, a, sizeof(a));
            if (dv <= thr && !(c->flags[i] & ISPH_FLAG_BUFFER)) fl |= ISPH_FLAG_FREE;
        } else {
            c->L[4 * (size_t)i] = 1.0; c->L[4 * (size_t)i + 1] = 0.0;
            c->L[4 * (size_t)i + 2] = 0.0; c->L[4 * (size_t)i + 3] = 1.0;
        }
        c->flags[i] = fl;
    }
    /* the reference stores dist and klt per pair (search.f90:65-74); keep
     * the positions they were computed at for stages that run after the
     * shift has moved the particles */
    memcpy(c->xnb, c->x, sizeof(double) * (size_t)c->n);
    memcpy(c->ynb, c->y, sizeof(double) * (size_t)c->n);
    c->have_nbr = 1;
    c->assembled = 0;
    if (over) return set_err(c, ISPH_E_CAPACITY, "neighbour list capacity exceeded");
    return ISPH_OK;
}

/* ---------------------------------------------------------------------- */
/* porous.f90:107-224 effpor (+ porsearch :32-105, vis :12-30)             */
/* ---------------------------------------------------------------------- */
int ora_effpor(ora_ctx* c) {
    const isph_params* prm = &c->prm;
    const double h2x = 2.0 * prm->h;
    const int32_t ny = c->ny;
    int oob = 0;
#pragma omp parallel for schedule(static) reduction(| : oob)
    for (int32_t i = 0; i < c->n; ++i) {
        c->R[i] = 0.0;
        if (c->tid[i] != 3) continue;
        const double xi = c->x[i], yi = c->y[i];
        int o = 0;
        int32_t cell = cell_of(c, xi, yi, &o);
        oob |= o;
        int32_t cx = cell / ny, cy = cell % ny;
        double sf = 0.0;
        /* soil cells visited: the 3x3 block column by column (cell order), or
         * porous.f90:127-176 in the reference order: the particle's own cell
         * (its plist cell), then cell2..cell4 of the 2x2 block on its side of
         * the cell midlines (x <= mid: left; y <= mid: lower) */
        int32_t cells[9], nc = 0;
        if (!prm->cell_order) {
            for (int32_t dc = -1; dc <= 1; ++dc)
                for (int32_t dr = -1; dr <= 1; ++dr) cells[nc++] = (cx + dc) * ny + cy + dr;
        } else {
            key_cell(c, c->key[i], &cx, &cy);
            double xleft = prm->xmin + (double)(cx - 1) * prm->cell_size;
            double xright = prm->xmin + (double)cx * prm->cell_size;
            double ybot = prm->ymin + (double)(cy - 1) * prm->cell_size;
            double ytop = prm->ymin + (double)cy * prm->cell_size;
            double xmid = xleft + fabs((xleft - xright) / 2.0);
            double ymid = ybot + fabs((ytop - ybot) / 2.0);
            int32_t d[3][2];
            if (xi <= xmid && yi <= ymid) {          /* (i,j-1), (i+1,j), (i+1,j-1) */
                d[0][0] = -1; d[0][1] = 0; d[1][0] = 0; d[1][1] = -1; d[2][0] = -1; d[2][1] = -1;
            } else if (xi > xmid && yi <= ymid) {    /* (i,j+1), (i+1,j), (i+1,j+1) */
                d[0][0] = 1; d[0][1] = 0; d[1][0] = 0; d[1][1] = -1; d[2][0] = 1; d[2][1] = -1;
            } else if (xi > xmid && yi > ymid) {     /* (i-1,j+1), (i-1,j), (i,j+1) */
                d[0][0] = 1; d[0][1] = 1; d[1][0] = 0; d[1][1] = 1; d[2][0] = 1; d[2][1] = 0;
            } else {                                 /* (i-1,j), (i-1,j-1), (i,j-1) */
                d[0][0] = 0; d[0][1] = 1; d[1][0] = -1; d[1][1] = 1; d[2][0] = -1; d[2][1] = 0;
            }
            cells[nc++] = cx * ny + cy;
            for (int q = 0; q < 3; ++q) cells[nc++] = (cx + d[q][0]) * ny + cy + d[q][1];
        }
        for (int32_t q = 0; q < nc; ++q) {
            for (int32_t s = c->sstart[cells[q]]; s < c->sstart[cells[q] + 1]; ++s) {
                double dx = xi - c->sx[s], dy = yi - c->sy[s];
                double r = sqrt(dx * dx + dy * dy);
                if (r < h2x)
                    sf = sf + (((1.0 - prm->por * 1.0) * ora_kernel_w(prm, r)) * c->sm[s]) / c->sr[s];
            }
        }
        if (sf > (1.0 - prm->por)) sf = 1.0 - prm->por;
        double n = 1.0 - sf;
        c->por[i] = n;
        c->rho[i] = n * c->oden[i];
        double n3 = (n * n) * n;
        double vm = sqrt(c->vx[i] * c->vx[i] + c->vy[i] * c->vy[i]);
        double omn = 1.0 - n;
        double R;
        if (prm->drag_law == ISPH_DRAG_REFERENCE) {
            double fch = 1.750 / sqrt(150.0 * n3);
            double t1 = (((visc(prm, c->oden[i]) / c->oden[i]) * prm->alpha) * (omn * omn)) /
                        (n3 * (prm->d50 * prm->d50));
            double t2 = (((fch * vm) * sqrt(prm->alpha)) * omn) / (sqrt(n3) * prm->d50);
            R = (-t1) - t2;
        } else {
            double t1 = ((prm->alpha * prm->nu) * (omn * omn)) / (n3 * (prm->d50 * prm->d50));
            double t2 = ((prm->beta * vm) * omn) / (n3 * prm->d50);
            R = (-t1) - t2;
        }
        c->R[i] = R;
    }
    if (oob) return set_err(c, ISPH_E_DOMAIN, "particle outside the cell grid");
    return ISPH_OK;
}

/* recompute r and grad W for pair (i, j) exactly as the search did */
static inline void pair_geom(const ora_ctx* c, int32_t i, int32_t j, double* r,
                             double* wx, double* wy) {
    double dx = c->x[i] - c->x[j], dy = c->y[i] - c->y[j];
    *r = sqrt(dx * dx + dy * dy);
    ora_kernel_gradw(&c->prm, *r, dx, dy, wx, wy);
}

/* r and grad W of pair (i, j) as stored by the last neighbour search */
static inline void pair_geom_nb(const ora_ctx* c, int32_t i, int32_t j, double* r,
                                double* wx, double* wy) {
    double dx = c->xnb[i] - c->xnb[j], dy = c->ynb[i] - c->ynb[j];
    *r = sqrt(dx * dx + dy * dy);
    ora_kernel_gradw(&c->prm, *r, dx, dy, wx, wy);
}

/* ---------------------------------------------------------------------- */
/* turbulence.f90:12-110 eddyvis                                           */
/* ---------------------------------------------------------------------- */
static int eddyvis_core(ora_ctx* c, int stored) {
    const isph_params* prm = &c->prm;
    const int32_t M = prm->max_nbr;
    const double cs = 0.150, cv = 0.080, ce = 1.0;
    const double dl = prm->dl;
    if (!c->have_nbr) return set_err(c, ISPH_E_STATE, "eddyvis before neighbour");
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < c->n; ++i) {
        if (c->tid[i] != 3) continue;
        const double* L = c->L + 4 * (size_t)i;
        double t1 = 0.0, t2 = 0.0, t3 = 0.0;
        const double ui = c->vx[i] / c->por[i], vi = c->vy[i] / c->por[i];
        for (int32_t k = 0; k < c->nptr[i]; ++k) {
            int32_t j = c->nidx[(size_t)i * M + k];
            double r, wx, wy;
            if (stored) pair_geom_nb(c, i, j, &r, &wx, &wy);
            else pair_geom(c, i, j, &r, &wx, &wy);
            double gx = L[0] * wx + L[1] * wy;
            double gy = L[2] * wx + L[3] * wy;
            double du = c->vx[j] / c->por[j] - ui;
            double dv = c->vy[j] / c->por[j] - vi;
            t1 = t1 + ((c->mass[j] * du) * gx) / c->rho[j];
            t2 = t2 + ((c->mass[j] * dv) * gy) / c->rho[j];
            t3 = t3 + ((0.50 * c->mass[j]) * (du * gy + dv * gx)) / c->rho[j];
        }
        double t4 = sqrt(2.0 * (t1 * t1 + t2 * t2) + (2.0 * t3) * (2.0 * t3));
        c->nut[i] = ((cs * dl) * (cs * dl)) * t4;
        c->tke[i] = (cv / ce) * ((dl * t4) * (dl * t4));
    }
    return ISPH_OK;
}

/* ---------------------------------------------------------------------- */
/* integrator.f90:458-713 int_vel                                          */
/* ---------------------------------------------------------------------- */
int ora_eddyvis(ora_ctx* c) { return eddyvis_core(c, 0); }
/* eddyvis on the stored pair geometry of the last search (the reference's
 * klt), as main.f90:260 calls it after the shift */
int ora_eddyvis_lagged(ora_ctx* c) { return eddyvis_core(c, 1); }

int ora_int_vel(ora_ctx* c) {
    const isph_params* prm = &c->prm;
    const int32_t M = prm->max_nbr;
    const double dt = c->dt, lam = prm->lam, dl = prm->dl, ps = 1.0;
    const double w0 = ora_kernel_w0(prm);
    if (!c->have_nbr) return set_err(c, ISPH_E_STATE, "int_vel before neighbour");
    /* (i) artificial viscosity of free-surface particles (:467-506) */
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < c->n; ++i) {
        if (!(c->flags[i] & ISPH_FLAG_FREE)) continue;
        const double* L = c->L + 4 * (size_t)i;
        double t1 = 0.0, t2 = 0.0;
        const double ui = c->vx[i] / c->por[i], vi = c->vy[i] / c->por[i];
        for (int32_t k = 0; k < c->nptr[i]; ++k) {
            int32_t j = c->nidx[(size_t)i * M + k];
            double r, wx, wy;
            pair_geom(c, i, j, &r, &wx, &wy);
            double gx = L[0] * wx + L[1] * wy;
            double gy = L[2] * wx + L[3] * wy;
            t1 = t1 + ((c->vx[j] / c->por[j] - ui) * gx) * (c->mass[j] / c->rho[j]);
            t2 = t2 + ((c->vy[j] / c->por[j] - vi) * gy) * (c->mass[j] / c->rho[j]);
        }
        c->varts[i] = (((ps * (dl * dl)) * sqrt(t1 * t1 + t2 * t2)) *
                       sqrt(c->vx[i] * c->vx[i] + c->vy[i] * c->vy[i])) /
                      ((c->umax + TINY6) * c->por[i]);
    }
    /* (ii) non-pressure velocity v* (:509-698) */
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < c->n; ++i) {
        const int32_t fi = c->flags[i];
        if (c->tid[i] == 3 && !(fi & ISPH_FLAG_BUFFER) && !(fi & ISPH_FLAG_DEAD)) {
            const double* L = c->L + 4 * (size_t)i;
            const double ni = c->por[i], oi = c->oden[i];
            const double mui = visc(prm, oi);
            const double nuti = c->nut[i], tkei = c->tke[i];
            double sx = 0.0, sy = 0.0;
            for (int32_t k = 0; k < c->nptr[i]; ++k) {
                int32_t j = c->nidx[(size_t)i * M + k];
                double r, wx, wy;
                pair_geom(c, i, j, &r, &wx, &wy);
                double gx = L[0] * wx + L[1] * wy;
                double gy = L[2] * wx + L[3] * wy;
                double cmu;
                if (c->tid[j] == 3) {
                    double oj = c->oden[j], muj = visc(prm, oj);
                    cmu = (((2.0 * mui) * muj) * ((1.0 / oi) * (1.0 / oj))) / ((mui / oi) + (muj / oj));
                } else {
                    cmu = mui * (1.0 / oi);
                }
                double vart;
                if (fi & ISPH_FLAG_FREE) vart = c->varts[i];
                else if ((c->flags[j] & ISPH_FLAG_FREE) && !(fi & ISPH_FLAG_BUFFER))
                    vart = (ora_kernel_w(prm, r) * c->varts[j]) / w0;
                else vart = 0.0;
                double pvol = c->mass[j] / c->rho[j];
                double pdist = 1.0 / (r * r + lam);
                double dxp = c->x[i] - c->x[j], dyp = c->y[i] - c->y[j];
                /* x component */
                double dvx = c->vx[i] - c->vx[j];
                double t1 = ((((2.0 * pvol) * dvx) * dxp) * gx) * pdist;
                double t2 = ((((2.0 * pvol) * dvx) * dyp) * gy) * pdist;
                sx = sx + (t1 + t2) * (cmu + vart);
                if (c->tid[j] > 2) {
                    double nutj = c->nut[j];
                    if (nuti < TINY6 && nutj < TINY6) sx = sx + ((t1 + t2) * (nuti + nutj)) * 0.50;
                    else sx = sx + (t1 + t2) * (((2.0 * nuti) * nutj) / (nuti + nutj));
                }
                /* y component */
                double dvy = c->vy[i] - c->vy[j];
                t1 = ((((2.0 * pvol) * dvy) * dxp) * gx) * pdist;
                t2 = ((((2.0 * pvol) * dvy) * dyp) * gy) * pdist;
                sy = sy + (t1 + t2) * (cmu + vart);
                if (c->tid[j] > 2) {
                    double nutj = c->nut[j];
                    if (nuti < TINY6 && nutj < TINY6) sy = sy + ((t1 + t2) * (nuti + nutj)) * 0.50;
                    else sy = sy + (t1 + t2) * (((2.0 * nuti) * nutj) / (nuti + nutj));
                }
                /* SGS kinetic energy gradient (tid_i == 3 always here) */
                double dtk = c->tke[j] - tkei;
                t1 = (pvol * dtk) * gx;
                sx = sx - ((t1 * ni) * 2.0) / 3.0;
                t1 = (pvol * dtk) * gy;
                sy = sy - ((t1 * ni) * 2.0) / 3.0;
            }
            double den = 1.0 - (dt * c->R[i]) * ni;
            double vxs = (c->vx[i] + dt * sx) / den;
            double vys = (c->vy[i] + dt * sy) / den;
            vys = vys + (prm->g * ni) * dt;
            c->vxs[i] = vxs;
            c->vys[i] = vys;
        }
        if (fi & ISPH_FLAG_BUFFER) {
            c->vxs[i] = c->vx[i];
            c->vys[i] = (prm->g * c->por[i]) * dt;
        }
    }
    return ISPH_OK;
}

/* ---------------------------------------------------------------------- */
/* isph.f90:115-456 PPE assembly (rows init :123-131, rules :136-421,      */
/* diagonal placement :425-426, Jacobi scaling :434-442)                   */
/* ---------------------------------------------------------------------- */
int ora_ppe_assemble(ora_ctx* c) {
    const isph_params* prm = &c->prm;
    const int32_t M = prm->max_nbr;
    const double dt = c->dt, lam = prm->lam;
    const double lampd = prm->rho0 * prm->rel_den;
    if (!c->have_nbr) return set_err(c, ISPH_E_STATE, "assembly before neighbour");
    c->rowptr[0] = 0;
    for (int32_t i = 0; i < c->n; ++i) c->rowptr[i + 1] = c->rowptr[i] + c->nptr[i] + 1;
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int32_t i = 0; i < c->n; ++i) {
        const int32_t ti = c->tid[i];
        const int freei = (c->flags[i] & ISPH_FLAG_FREE) != 0;
        const double* Li = c->L + 4 * (size_t)i;
        const double xi = c->x[i], yi = c->y[i];
        const double ni = c->por[i], rhoi = c->rho[i];
        const double rho_i = rhoi * (1.0 / ni);
        const int32_t base = c->rowptr[i];
        const int32_t cnt = c->nptr[i];
        int32_t dpos = 0;
        for (int32_t k = 0; k < cnt; ++k)
            if (c->nidx[(size_t)i * M + k] < i) dpos = k + 1;
        double dia = 0.0, b = 0.0;
        for (int32_t k = 0; k < cnt; ++k) {
            int32_t j = c->nidx[(size_t)i * M + k];
            int32_t slot = base + (k < dpos ? k : k + 1);
            c->col[slot] = j;
            double v = 0.0;
            const int32_t tj = c->tid[j];
            double lamp = (freei && tj == 3) ? (1.0 + c->oden[j] / lampd) : 1.0;
            double nj = c->por[j], rhoj = c->rho[j];
            double rho_j = rhoj * (1.0 / nj);
            double pvol = c->mass[j] * (1.0 / rhoj);
            double ddx = c->x[j] - xi, ddy = c->y[j] - yi;
            double rr = sqrt(ddx * ddx + ddy * ddy);
            double pdist = 1.0 / (rr * rr + lam);
            double rho_ij = (((2.0 * rhoj) * rhoi) * (1.0 / (nj * ni))) *
                            (1.0 / (rhoj * (1.0 / nj) + rhoi * (1.0 / ni)));
            double r, wx, wy;
            pair_geom(c, i, j, &r, &wx, &wy);
            double dxp = xi - c->x[j], dyp = yi - c->y[j];
            double t1, t2;
            if (ti == 3) {
                double gx = Li[0] * wx + Li[1] * wy;
                double gy = Li[2] * wx + Li[3] * wy;
                t1 = (((2.0 * pvol) * gx) * dxp) * pdist;
                t2 = (((2.0 * pvol) * gy) * dyp) * pdist;
                v = -(t1 + t2);
                dia = dia + (t1 + t2) * lamp;
                if (tj == 3) {
                    t1 = (pvol * ((c->vxs[j] * rho_ij) / nj - (c->vxs[i] * rho_ij) / ni)) * gx;
                    t2 = (pvol * ((c->vys[j] * rho_ij) / nj - (c->vys[i] * rho_ij) / ni)) * gy;
                } else {
                    t1 = (pvol * ((c->vxs[j] * rho_i) / ni - (c->vxs[i] * rho_i) / ni)) * gx;
                    t2 = (pvol * ((c->vys[j] * rho_i) / ni - (c->vys[i] * rho_i) / ni)) * gy;
                }
                b = b + (t1 + t2) / dt;
            } else if (ti == 2) {
                t1 = (((2.0 * pvol) * wx) * dxp) * pdist;
                t2 = (((2.0 * pvol) * wy) * dyp) * pdist;
                if (tj <= 2) {
                    v = -(t1 + t2);
                    dia = dia + (t1 + t2);
                }
            } else if (ti == 1) {
                if (tj == 3) {
                    const double* Lj = c->L + 4 * (size_t)j;
                    double gjx = Lj[0] * wx + Lj[1] * wy;
                    double gjy = Lj[2] * wx + Lj[3] * wy;
                    const double mi = c->mass[i];
                    t1 = ((((2.0 * mi) * gjx) * dxp) * pdist) / rhoi;
                    t2 = ((((2.0 * mi) * gjy) * dyp) * pdist) / rhoi;
                    v = (-(t1 + t2)) * 1.0;
                    dia = dia + (t1 + t2);
                    t1 = ((mi * ((c->vxs[j] * rho_j) / nj - (c->vxs[i] * rho_j) / ni)) * gjx) / rhoi;
                    t2 = ((mi * ((c->vys[j] * rho_j) / nj - (c->vys[i] * rho_j) / ni)) * gjy) / rhoi;
                    b = b + (t1 + t2) / dt;
                } else if (tj <= 2) {
                    t1 = (((2.0 * pvol) * wx) * dxp) * pdist;
                    t2 = (((2.0 * pvol) * wy) * dyp) * pdist;
                    v = -(t1 + t2);
                    dia = dia + (t1 + t2);
                }
            }
            c->val[slot] = v;
        }
        c->col[base + dpos] = i;
        c->val[base + dpos] = dia;
        if (!(dia != 0.0) || !isfinite(dia)) bad = 1;
        c->rhs[i] = b / dia;
        for (int32_t k = 0; k < cnt + 1; ++k) c->val[base + k] = c->val[base + k] / dia;
    }
    c->assembled = 1;
    if (bad) return set_err(c, ISPH_E_SINGULAR, "PPE row with zero or non-finite diagonal");
    return ISPH_OK;
}

/* CSR matvec, each row summed in column order (spmv2, solver.f90:84-127) */
static void spmv_csr(int32_t n, const int32_t* rp, const int32_t* col, const double* val,
                     const double* x, double* y) {
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) s = s + val[k] * x[col[k]];
        y[i] = s;
    }
}

static void spmv_csr_f(int32_t n, const int32_t* rp, const int32_t* col, const float* val,
                       const float* x, float* y) {
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) {
        float s = 0.0f;
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) s = s + val[k] * x[col[k]];
        y[i] = s;
    }
}

/* Restart threshold on the BiCGSTAB shadow inner product: rho = (r, rhat)
 * carries a rounding error of about eps ||r|| ||rhat||, so once
 * |rho| <= BRK_COS ||r|| ||rhat|| it has lost its digits and the recurrence
 * restarts from the true residual with rhat = r. */
#define BRK_COS 1e-15

/* BiCGSTAB with the recurrences of solver.f90:162-213 (van der Vorst 1992),
 * stop when ||r||_2 <= tol ||b||_2 on the recurrence residual and then on the
 * true residual b - Ax; a failed true-residual check restarts from it. */
static int bicgstab_core(int32_t n, const int32_t* rp, const int32_t* col, const double* val,
                         const double* b, double* sl, double tol, int32_t maxit,
                         double** w, int32_t* iters, double* relres) {
    double *rs = w[0], *p = w[1], *res = w[2], *mm1 = w[3], *mm2 = w[4], *si = w[5];
    double bb = ora_dot_tree(b, b, n);
    *iters = 0;
    if (bb == 0.0) {
        for (int32_t i = 0; i < n; ++i) sl[i] = 0.0;
        *relres = 0.0;
        return ISPH_OK;
    }
    const double nb = sqrt(bb);
    spmv_csr(n, rp, col, val, sl, res);
    for (int32_t i = 0; i < n; ++i) {
        rs[i] = b[i] - res[i];
        p[i] = rs[i];
        res[i] = rs[i];
    }
    double rr = ora_dot_tree(rs, rs, n);
    double prevr = rr;
    double rh2 = rr; /* ||rhat||^2 */
    if (sqrt(rr) <= tol * nb) { *relres = sqrt(rr) / nb; return ISPH_OK; }
    int it = 0;
    for (;;) {
        if (it >= maxit) { *iters = it; *relres = sqrt(rr) / nb; return ISPH_E_NOCONV; }
        it++;
        spmv_csr(n, rp, col, val, p, mm1);
        double d1 = ora_dot_tree(mm1, res, n);
        double alpha = prevr / d1;
        /* breakdown ((rhat, v) = 0, or (t, t) = 0 / (t, s) = 0 below): the
         * step degenerates to a minimal-residual step (alpha = 0) or to the
         * half step x + alpha p (omega = 0), then restarts from the true
         * residual with rhat = r */
        int brk = 0;
        if (!isfinite(alpha)) { alpha = 0.0; brk = 1; }
        for (int32_t i = 0; i < n; ++i) si[i] = rs[i] - alpha * mm1[i];
        spmv_csr(n, rp, col, val, si, mm2);
        double d2 = ora_dot_tree(mm2, si, n);
        double d3 = ora_dot_tree(mm2, mm2, n);
        double wi = d2 / d3;
        if (!isfinite(wi) || wi == 0.0) { wi = 0.0; brk = 1; }
        for (int32_t i = 0; i < n; ++i) {
            sl[i] = (sl[i] + alpha * p[i]) + wi * si[i];
            rs[i] = si[i] - wi * mm2[i];
        }
        double d4 = ora_dot_tree(rs, res, n);
        rr = ora_dot_tree(rs, rs, n);
        double betas = (d4 / prevr) * (alpha / wi);
        prevr = d4;
        if (!(fabs(d4) > BRK_COS * sqrt(rr * rh2))) brk = 1;
        if (!brk && (!isfinite(betas) || !isfinite(rr))) {
            *iters = it; *relres = NAN;
            return ISPH_E_NOCONV;
        }
        if (brk || sqrt(rr) <= tol * nb) {
            /* true residual */
            spmv_csr(n, rp, col, val, sl, mm1);
            for (int32_t i = 0; i < n; ++i) rs[i] = b[i] - mm1[i];
            double tr = ora_dot_tree(rs, rs, n);
            if (sqrt(tr) <= tol * nb) { *iters = it; *relres = sqrt(tr) / nb; return ISPH_OK; }
            if (!isfinite(tr)) { *iters = it; *relres = NAN; return ISPH_E_NOCONV; }
            for (int32_t i = 0; i < n; ++i) { p[i] = rs[i]; res[i] = rs[i]; }
            rr = tr;
            prevr = tr;
            rh2 = tr;
            continue;
        }
        for (int32_t i = 0; i < n; ++i) p[i] = rs[i] + betas * (p[i] - wi * mm1[i]);
    }
}

int ora_bicgstab_csr(int32_t n, const int32_t* rowptr, const int32_t* col, const double* val,
                     const double* b, double* x, double tol, int32_t maxit, int32_t* iters,
                     double* relres) {
    double* w[6];
    for (int k = 0; k < 6; ++k) w[k] = (double*)xmalloc(sizeof(double) * (size_t)(n ? n : 1));
    int rc = bicgstab_core(n, rowptr, col, val, b, x, tol, maxit, w, iters, relres);
    for (int k = 0; k < 6; ++k) free(w[k]);
    return rc;
}

/* FP32 values, FP64 vectors, conservation form: y_i = s_i x_i +
 * sum_{j != i} a_ij (x_j - x_i) in column order, s_i the FP64 row sum of
 * the scaled row (columns in order, diagonal included). */
static const double* g_rsum;
static void spmv_csr_fd(int32_t n, const int32_t* rp, const int32_t* col, const float* val,
                        const double* x, double* y) {
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) {
        double acc = 0.0;
        const double xi = x[i];
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
            if (col[k] == i) continue;
            acc = acc + (double)val[k] * (x[col[k]] - xi);
        }
        y[i] = g_rsum[i] * xi + acc;
    }
}

/* Mixed precision: BiCGSTAB (same recurrences) with FP32 matrix values,
 * Krylov vectors of type TX (float: ISPH_PREC_MIXED, double:
 * ISPH_PREC_MIXED_V64), dot products accumulated in FP64, and reliable
 * updates (Sleijpen & van der Vorst 1996; van der Vorst & Ye 2000): when
 * the recurrence residual has dropped by delta = inner_tol (0: 0.1) since the last
 * update (or below tol), the iterate is added to the FP64 solution, the
 * FP64 true residual b - A64 x replaces the recurrence residual and the
 * iteration continues with its search direction kept.  The stop test is the
 * FP64 true residual; max_outer caps the number of updates. */
#define MIXED_SOLVE(NAME, TX, SPMV, DOT, VEC)                                                \
static int NAME(ora_ctx* c, const float* val32, double* x, int32_t* iters, int32_t* outer,  \
                double* relres) {                                                           \
    const isph_params* prm = &c->prm;                                                       \
    /* inner_tol 0: the default factor 0.1 (isph_capi.h) */                                 \
    const double delta = prm->inner_tol > 0.0 ? prm->inner_tol : 0.1;                       \
    const int32_t n = c->n;                                                                 \
    double* r64 = c->w[6];                                                                  \
    double* ax = c->w[7];                                                                   \
    TX *rs = VEC(0), *p = VEC(1), *res = VEC(2), *mm1 = VEC(3), *mm2 = VEC(4), *si = VEC(5), \
       *sl = VEC(6);                                                                        \
    const double* b = c->rhs;                                                               \
    *iters = 0;                                                                             \
    *outer = 0;                                                                             \
    int rc = ISPH_OK;                                                                       \
    spmv_csr(n, c->rowptr, c->col, c->val, x, ax);                                          \
    for (int32_t i = 0; i < n; ++i) r64[i] = b[i] - ax[i];                                 \
    double bb = ora_dot_tree(b, b, n);                                                      \
    double rr = ora_dot_tree(r64, r64, n);                                                  \
    if (bb == 0.0) {                                                                        \
        for (int32_t i = 0; i < n; ++i) x[i] = 0.0;                                         \
        *relres = 0.0;                                                                      \
        return ISPH_OK;                                                                     \
    }                                                                                       \
    const double nb = sqrt(bb);                                                             \
    *relres = sqrt(rr) / nb;                                                                \
    if (sqrt(rr) <= prm->tol * nb) return ISPH_OK;                                          \
    double rupd = sqrt(rr);                                                                 \
    for (int32_t i = 0; i < n; ++i) {                                                       \
        rs[i] = (TX)r64[i];                                                                 \
        res[i] = rs[i];                                                                     \
        sl[i] = (TX)0;                                                                      \
    }                                                                                       \
    double prevr = DOT(res, rs, n), prevr_prev = prevr, rh2 = prevr;                        \
    double alpha = 0.0, wi = 0.0, betas = 0.0;                                              \
    int first = 1;                                                                          \
    int32_t total = 0, nupd = 0;                                                            \
    for (;;) {                                                                              \
        if (total >= prm->maxit) { rc = ISPH_E_NOCONV; break; }                             \
        total++;                                                                            \
        if (first) {                                                                        \
            for (int32_t i = 0; i < n; ++i) p[i] = rs[i];                                   \
        } else {                                                                            \
            TX bt = (TX)betas, wt = (TX)wi;                                                 \
            for (int32_t i = 0; i < n; ++i) p[i] = rs[i] + bt * (p[i] - wt * mm1[i]);       \
        }                                                                                   \
        first = 0;                                                                          \
        SPMV(n, c->rowptr, c->col, val32, p, mm1);                                          \
        double d1 = DOT(mm1, res, n);                                                       \
        alpha = prevr / d1;                                                                 \
        int brk = 0;                                                                        \
        if (!isfinite(alpha)) { alpha = 0.0; brk = 1; }                                     \
        TX at = (TX)alpha;                                                                  \
        for (int32_t i = 0; i < n; ++i) si[i] = rs[i] - at * mm1[i];                        \
        SPMV(n, c->rowptr, c->col, val32, si, mm2);                                         \
        double d2 = DOT(mm2, si, n);                                                        \
        double d3 = DOT(mm2, mm2, n);                                                       \
        wi = d2 / d3;                                                                       \
        if (!isfinite(wi) || wi == 0.0) { wi = 0.0; brk = 1; }                              \
        TX wt = (TX)wi;                                                                     \
        for (int32_t i = 0; i < n; ++i) {                                                   \
            sl[i] = (sl[i] + at * p[i]) + wt * si[i];                                       \
            rs[i] = si[i] - wt * mm2[i];                                                    \
        }                                                                                   \
        double d4 = DOT(rs, res, n);                                                        \
        double rrk = DOT(rs, rs, n);                                                        \
        betas = (d4 / prevr) * (alpha / wi);                                                \
        prevr_prev = prevr;                                                                 \
        prevr = d4;                                                                         \
        if (!(fabs(d4) > BRK_COS * sqrt(rrk * rh2))) brk = 1;                               \
        if (!brk && (!isfinite(betas) || !isfinite(rrk))) { rc = ISPH_E_NOCONV; break; }    \
        if (brk || sqrt(rrk) <= prm->tol * nb || sqrt(rrk) <= delta * rupd) {               \
            for (int32_t i = 0; i < n; ++i) {                                               \
                x[i] = x[i] + (double)sl[i];                                                \
                sl[i] = (TX)0;                                                              \
            }                                                                               \
            spmv_csr(n, c->rowptr, c->col, c->val, x, ax);                                  \
            for (int32_t i = 0; i < n; ++i) r64[i] = b[i] - ax[i];                         \
            double tr = ora_dot_tree(r64, r64, n);                                          \
            nupd++;                                                                         \
            *relres = sqrt(tr) / nb;                                                        \
            if (sqrt(tr) <= prm->tol * nb) break;                                           \
            if (nupd >= prm->max_outer || total >= prm->maxit || !isfinite(tr)) {          \
                rc = ISPH_E_NOCONV;                                                         \
                break;                                                                      \
            }                                                                               \
            for (int32_t i = 0; i < n; ++i) rs[i] = (TX)r64[i];                             \
            rupd = sqrt(tr);                                                                \
            if (brk || prm->precision == ISPH_PREC_IR) { /* restart from zero: rhat = r */  \
                for (int32_t i = 0; i < n; ++i) res[i] = rs[i];                             \
                prevr = DOT(res, rs, n);                                                    \
                prevr_prev = prevr;                                                         \
                rh2 = prevr;                                                                \
                first = 1;                                                                  \
            } else {                                                                        \
                d4 = DOT(rs, res, n);                                                       \
                betas = (d4 / prevr_prev) * (alpha / wi);                                   \
                prevr = d4;                                                                 \
            }                                                                               \
        }                                                                                   \
    }                                                                                       \
    *iters = total;                                                                         \
    *outer = nupd;                                                                          \
    return rc;                                                                              \
}

/* work vectors: w[6] = r64, w[7] = A x, w[8] = x are taken */
#define VEC_F(k) (c->f[k])
#define VEC_D(k) (c->w[(k) == 6 ? 9 : (k)])
MIXED_SOLVE(mixed_solve_f32, float, spmv_csr_f, dot_tree_f, VEC_F)
MIXED_SOLVE(mixed_solve_v64, double, spmv_csr_fd, ora_dot_tree, VEC_D)

static int mixed_solve(ora_ctx* c, double* x, int32_t* iters, int32_t* outer, double* relres) {
    const int32_t nnz = c->rowptr[c->n];
    float* val32 = (float*)xmalloc(sizeof(float) * (size_t)(nnz ? nnz : 1));
    for (int32_t k = 0; k < nnz; ++k) val32[k] = (float)c->val[k];
    double* rsum = (double*)xmalloc(sizeof(double) * (size_t)(c->n ? c->n : 1));
    for (int32_t i = 0; i < c->n; ++i) {
        double s = 0.0;
        for (int32_t k = c->rowptr[i]; k < c->rowptr[i + 1]; ++k) s = s + c->val[k];
        rsum[i] = s;
    }
    g_rsum = rsum;
    int rc = c->prm.precision == ISPH_PREC_MIXED_V64
                 ? mixed_solve_v64(c, val32, x, iters, outer, relres)
                 : mixed_solve_f32(c, val32, x, iters, outer, relres);
    free(val32);
    free(rsum);
    g_rsum = NULL;
    return rc;
}

/* ---------------------------------------------------------------------- */
/* solver.f90:306-347 fgmres: MKL's RCI dfgmres as the reference sets it   */
/* ---------------------------------------------------------------------- */
/* No preconditioner (ipar(11) = 0), so FGMRES (Saad 1993) is restarted
 * GMRES with z_j = v_j.  Stop tests after every iteration, in this order:
 * the residual estimate |g_(j+1)| <= tol * ||r0|| (ipar(9) = 1, dpar(1) =
 * tol, dpar(3) = ||r0|| of the warm start), the zero norm of the new Arnoldi
 * vector <= 1e-12 (ipar(12) = 1, dpar(8)), the iteration limit `maxit`
 * (ipar(5)), the end of the restart cycle of m = min(restart or 150, n)
 * columns (ipar(15)).  The Arnoldi vectors are orthogonalised by classical
 * Gram-Schmidt applied twice; every dot product is ora_dot_tree, so the GPU
 * kernels' trees give the same bits.  On convergence (or a zero norm) and at
 * the end of a cycle x += V y with y = R^-1 g; at the iteration limit x is
 * left as it was at the start of the cycle.  Returns 1 when converged, 0 at
 * the limit, -1 on a breakdown. */
#define GM_DEFAULT_M 150
#define GM_ZERO_NORM 1e-12
static int fgmres_core(int32_t n, const int32_t* rp, const int32_t* col, const double* val,
                       const double* b, double* x, double tol, int32_t maxit, int32_t restart,
                       int32_t* iters, double* relres) {
    int32_t m = restart > 0 ? restart : GM_DEFAULT_M;
    if (m > n) m = n;
    *iters = 0;
    *relres = 0.0;
    if (n == 0) return 1;
    double* V = (double*)xmalloc(sizeof(double) * (size_t)(m + 1) * (size_t)n);
    double* w = (double*)xmalloc(sizeof(double) * (size_t)n);
    double* u = (double*)xmalloc(sizeof(double) * (size_t)n);
    double* R = (double*)xmalloc(sizeof(double) * (size_t)(m + 1) * (size_t)m);
    double* cs = (double*)xmalloc(sizeof(double) * (size_t)m);
    double* sn = (double*)xmalloc(sizeof(double) * (size_t)m);
    double* g = (double*)xmalloc(sizeof(double) * (size_t)(m + 1));
    double* y = (double*)xmalloc(sizeof(double) * (size_t)m);
    double* h1 = (double*)xmalloc(sizeof(double) * (size_t)(m + 1));
    double* h2 = (double*)xmalloc(sizeof(double) * (size_t)(m + 1));
    enum { RUN, CONV, HAPPY, CYCLE, MAXIT, FAIL } st;
    /* warm-start residual */
    spmv_csr(n, rp, col, val, x, w);
    for (int32_t r = 0; r < n; ++r) u[r] = b[r] - w[r];
    double rn = sqrt(ora_dot_tree(u, u, n));
    const double r0 = rn, tolabs = tol * rn;
    double res = rn;
    int32_t total = 0;
    st = !isfinite(rn) ? FAIL : (rn == 0.0 || rn <= tolabs) ? CONV : RUN;
    while (st == RUN) {
        double beta = rn;
        g[0] = rn;
        int32_t j = 0;
        while (st == RUN) {
            double* vj = V + (size_t)j * n;
            for (int32_t r = 0; r < n; ++r) vj[r] = u[r] / beta;
            spmv_csr(n, rp, col, val, vj, w);
            for (int32_t i = 0; i <= j; ++i) h1[i] = ora_dot_tree(V + (size_t)i * n, w, n);
            for (int32_t r = 0; r < n; ++r) {
                double t = w[r];
                for (int32_t i = 0; i <= j; ++i) t = t - h1[i] * V[(size_t)i * n + r];
                u[r] = t;
            }
            for (int32_t i = 0; i <= j; ++i) h2[i] = ora_dot_tree(V + (size_t)i * n, u, n);
            for (int32_t r = 0; r < n; ++r) {
                double t = u[r];
                for (int32_t i = 0; i <= j; ++i) t = t - h2[i] * V[(size_t)i * n + r];
                u[r] = t;
            }
            beta = sqrt(ora_dot_tree(u, u, n));
            double* cj = R + (size_t)j * (m + 1);
            for (int32_t i = 0; i <= j; ++i) cj[i] = h1[i] + h2[i];
            for (int32_t i = 0; i < j; ++i) {
                const double a = cj[i], bb = cj[i + 1];
                cj[i] = cs[i] * a + sn[i] * bb;
                cj[i + 1] = cs[i] * bb - sn[i] * a;
            }
            const double a = cj[j];
            const double d = sqrt(a * a + beta * beta);
            total++;
            if (!(d > 0.0) || !isfinite(d)) { st = FAIL; break; }
            cs[j] = a / d;
            sn[j] = beta / d;
            cj[j] = d;
            const double gj = g[j];
            g[j + 1] = -(sn[j] * gj);
            g[j] = cs[j] * gj;
            j++;
            res = fabs(g[j]);
            if (!isfinite(res)) st = FAIL;
            else if (res <= tolabs) st = CONV;
            else if (beta <= GM_ZERO_NORM) st = HAPPY;
            else if (total >= maxit) st = MAXIT;
            else if (j >= m) st = CYCLE;
        }
        if (st == CONV || st == HAPPY || st == CYCLE) {
            for (int32_t i = j - 1; i >= 0; --i) {
                double sacc = g[i];
                for (int32_t l = i + 1; l < j; ++l) sacc = sacc - R[(size_t)l * (m + 1) + i] * y[l];
                y[i] = sacc / R[(size_t)i * (m + 1) + i];
            }
            for (int32_t r = 0; r < n; ++r) {
                double t = x[r];
                for (int32_t i = 0; i < j; ++i) t = t + y[i] * V[(size_t)i * n + r];
                x[r] = t;
            }
        }
        if (st == CYCLE) {
            spmv_csr(n, rp, col, val, x, w);
            for (int32_t r = 0; r < n; ++r) u[r] = b[r] - w[r];
            rn = sqrt(ora_dot_tree(u, u, n));
            res = rn;
            st = !isfinite(rn) ? FAIL : (rn == 0.0 || rn <= tolabs) ? CONV : RUN;
        }
    }
    *iters = total;
    *relres = r0 > 0.0 ? res / r0 : 0.0;
    free(V); free(w); free(u); free(R); free(cs); free(sn); free(g); free(y); free(h1); free(h2);
    return st == FAIL ? -1 : (st == MAXIT ? 0 : 1);
}

int ora_fgmres_csr(int32_t n, const int32_t* rowptr, const int32_t* col, const double* val,
                   const double* b, double* x, double tol, int32_t maxit, int32_t restart,
                   int32_t* iters, double* relres, int32_t* converged) {
    int rc = fgmres_core(n, rowptr, col, val, b, x, tol, maxit, restart, iters, relres);
    if (converged) *converged = rc == 1;
    return rc < 0 ? ISPH_E_NOCONV : ISPH_OK;
}

/* solve (isph.f90:460) and pressure scatter (isph.f90:465-483); the warm
 * start is the particles' current pressure (the reference warm-starts
 * FGMRES from fsol).  A failed solve leaves p unchanged, as the reference's
 * RCI loop does when dfgmres never returns 0 (solver.f90:330-343). */
int ora_ppe_solve(ora_ctx* c, int32_t* iters, double* relres) {
    const isph_params* prm = &c->prm;
    if (!c->assembled) return set_err(c, ISPH_E_STATE, "solve before assembly");
    double* x = c->w[8];
    memcpy(x, c->p, sizeof(double) * (size_t)c->n);
    int32_t it = 0, outer = 0;
    double rel = 0.0;
    int rc;
    if (prm->precision == ISPH_PREC_MIXED_MG || prm->precision == ISPH_PREC_FP64_MG)
        return set_err(c, ISPH_E_ARG, "the multigrid solve is checked against the FP64 solve, "
                                      "it has no restatement here");
    if (prm->precision == ISPH_PREC_FGMRES) {
        /* warm start fsol in row order (solver.f90:319); fsol takes the
         * solution only on convergence (solver.f90:330-333); p = fsol */
        memcpy(x, c->fsol, sizeof(double) * (size_t)c->n);
        int32_t conv = 0;
        rc = ora_fgmres_csr(c->n, c->rowptr, c->col, c->val, c->rhs, x, prm->tol, prm->maxit,
                            prm->inner_maxit, &it, &rel, &conv);
        c->iters = it;
        c->outer = 0;
        c->relres = rel;
        if (iters) *iters = it;
        if (relres) *relres = rel;
        if (rc != ISPH_OK) return set_err(c, rc, "GMRES breakdown (non-finite or singular)");
        if (conv) memcpy(c->fsol, x, sizeof(double) * (size_t)c->n);
        memcpy(c->p, c->fsol, sizeof(double) * (size_t)c->n);
        return ISPH_OK;
    }
    if (prm->precision == ISPH_PREC_MIXED || prm->precision == ISPH_PREC_MIXED_V64 ||
        prm->precision == ISPH_PREC_IR) {
        rc = mixed_solve(c, x, &it, &outer, &rel);
    } else {
        rc = bicgstab_core(c->n, c->rowptr, c->col, c->val, c->rhs, x, prm->tol, prm->maxit,
                           c->w, &it, &rel);
    }
    c->iters = it;
    c->outer = outer;
    c->relres = rel;
    if (iters) *iters = it;
    if (relres) *relres = rel;
    if (rc != ISPH_OK) return set_err(c, rc, "PPE solve did not converge");
    memcpy(c->p, x, sizeof(double) * (size_t)c->n);
    return ISPH_OK;
}

int ora_ppesolve(ora_ctx* c, int32_t* iters, double* relres) {
    int rc = ora_ppe_assemble(c);
    if (rc) return rc;
    return ora_ppe_solve(c, iters, relres);
}

/* ---------------------------------------------------------------------- */
/* integrator.f90:715-865 comp_vel                                         */
/* ---------------------------------------------------------------------- */
int ora_comp_vel(ora_ctx* c) {
    const isph_params* prm = &c->prm;
    const int32_t M = prm->max_nbr;
    const double dt = c->dt;
    if (!c->have_nbr) return set_err(c, ISPH_E_STATE, "comp_vel before neighbour");
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < c->n; ++i) {
        const int32_t fi = c->flags[i];
        if (c->tid[i] == 3 && !(fi & ISPH_FLAG_BUFFER) && !(fi & ISPH_FLAG_DEAD)) {
            const double* L = c->L + 4 * (size_t)i;
            const double ni = c->por[i], rhoi = c->rho[i], pi = c->p[i];
            double vx = 0.0, vy = 0.0;
            for (int32_t k = 0; k < c->nptr[i]; ++k) {
                int32_t j = c->nidx[(size_t)i * M + k];
                double r, wx, wy;
                pair_geom(c, i, j, &r, &wx, &wy);
                double gx = L[0] * wx + L[1] * wy;
                double gy = L[2] * wx + L[3] * wy;
                double vj = c->mass[j] / c->rho[j];
                double t1;
                if (c->tid[j] == 3) {
                    double nj = c->por[j], rhoj = c->rho[j], pj = c->p[j];
                    double q = ((pj * rhoi) / ni) + ((pi * rhoj) / nj);
                    double den = rhoi * ((rhoi / ni) + (rhoj / nj));
                    t1 = (2.0 * (((vj * (ni * ni)) * q) * gx)) / den;
                    vx = vx - t1 * dt;
                    t1 = (2.0 * (((vj * (ni * ni)) * q) * gy)) / den;
                    vy = vy - t1 * dt;
                } else {
                    double q = c->p[j] + pi;
                    t1 = (((vj * (ni * ni)) * q) * gx) / rhoi;
                    vx = vx - t1 * dt;
                    t1 = (((vj * (ni * ni)) * q) * gy) / rhoi;
                    vy = vy - t1 * dt;
                }
            }
            vx = c->vxs[i] + vx;
            vy = c->vys[i] + vy;
            c->vx[i] = vx;
            c->vy[i] = vy;
            c->vel[i] = sqrt(vx * vx + vy * vy) / ni;
        }
        if (fi & ISPH_FLAG_BUFFER) {
            c->vx[i] = c->vxs[i];
            c->vy[i] = 0.0;
        }
        if ((fi & ISPH_FLAG_DEAD) && !(fi & ISPH_FLAG_BUFFER)) {
            c->vx[i] = 0.0; c->vy[i] = 0.0; c->vxs[i] = 0.0; c->vys[i] = 0.0;
        }
    }
    return ISPH_OK;
}

/* integrator.f90:867-906 comp_pos, with the buffer release of :888-897: a
 * buffer particle with xrelease_lo < x < xrelease_hi after the update stops
 * being a buffer particle */
int ora_comp_pos(ora_ctx* c) {
    const double dt = c->dt;
    const double lo = c->prm.xrelease_lo, hi = c->prm.xrelease_hi;
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < c->n; ++i) {
        if (c->tid[i] == 3 && !(c->flags[i] & ISPH_FLAG_DEAD)) {
            c->x[i] = ((c->x[i] + c->xs[i]) + (dt * c->vx[i]) / c->por[i]) / 2.0;
            c->y[i] = ((c->y[i] + c->ys[i]) + (dt * c->vy[i]) / c->por[i]) / 2.0;
            if ((c->flags[i] & ISPH_FLAG_BUFFER) && c->x[i] > lo && c->x[i] < hi)
                c->flags[i] &= ~ISPH_FLAG_BUFFER;
        }
    }
    return ISPH_OK;
}

/* isph.f90:49-86 freesurf, in pull form over the symmetric lists: a particle
 * is `vicinity` when a neighbouring non-buffer fluid particle has
 * divr <= lamfs*maxdivr.  Nothing downstream reads it
 * (part_shift.f90:68-69 zeroes the normals either way). */
int ora_freesurf(ora_ctx* c) {
    const isph_params* prm = &c->prm;
    const int32_t M = prm->max_nbr;
    const double thr = prm->lamfs * prm->maxdivr;
    if (!c->have_nbr) return set_err(c, ISPH_E_STATE, "freesurf before neighbour");
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < c->n; ++i) {
        int v = 0;
        for (int32_t k = 0; k < c->nptr[i]; ++k) {
            int32_t j = c->nidx[(size_t)i * M + k];
            if (c->tid[j] == 3 && !(c->flags[j] & ISPH_FLAG_BUFFER) && c->divr[j] <= thr) v = 1;
        }
        if (v) c->flags[i] |= ISPH_FLAG_VICINITY;
        else c->flags[i] &= ~ISPH_FLAG_VICINITY;
    }
    return ISPH_OK;
}

/* ---------------------------------------------------------------------- */
/* part_shift.f90:12-192 opt2_shift                                        */
/* ---------------------------------------------------------------------- */
int ora_opt2_shift(ora_ctx* c) {
    const isph_params* prm = &c->prm;
    const int32_t M = prm->max_nbr;
    const double heff = prm->h;
    const double cap = prm->maxshift * prm->dl;
    if (!c->have_nbr) return set_err(c, ISPH_E_STATE, "opt2_shift before neighbour");
    if (!prm->shifting) return ISPH_OK;
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < c->n; ++i) {
        const int32_t fi = c->flags[i];
        if (c->tid[i] != 3 || (fi & ISPH_FLAG_BUFFER)) continue;
        c->xs[i] = 0.0; c->ys[i] = 0.0; c->vxs[i] = 0.0; c->vys[i] = 0.0;
        if ((fi & ISPH_FLAG_DEAD) || c->nptr[i] == 0) continue;
        const double* L = c->L + 4 * (size_t)i;
        /* normals (:46-66) are overwritten with 0 at :68-69; pint = 0 */
        const double frac = (fi & ISPH_FLAG_FREE) ? 0.01 : 1.0;
        const double cf = (prm->csh * (heff * heff)) * frac;
        double xs = 0.0, ys = 0.0;
        for (int32_t k = 0; k < c->nptr[i]; ++k) {
            int32_t j = c->nidx[(size_t)i * M + k];
            double r, wx, wy;
            pair_geom(c, i, j, &r, &wx, &wy);
            double vj = c->mass[j] / c->rho[j];
            xs = xs - (cf * wx) * vj;
        }
        for (int32_t k = 0; k < c->nptr[i]; ++k) {
            int32_t j = c->nidx[(size_t)i * M + k];
            double r, wx, wy;
            pair_geom(c, i, j, &r, &wx, &wy);
            double vj = c->mass[j] / c->rho[j];
            ys = ys - (cf * wy) * vj;
        }
        double axs = fabs(xs), ays = fabs(ys);
        c->xs[i] = (xs == 0.0) ? 0.0 : ((axs < cap ? axs : cap) * xs) / axs;
        c->ys[i] = (ys == 0.0) ? 0.0 : ((ays < cap ? ays : cap) * ys) / ays;
        double t1 = 0.0, t2 = 0.0, t3 = 0.0, t4 = 0.0;
        for (int32_t k = 0; k < c->nptr[i]; ++k) {
            int32_t j = c->nidx[(size_t)i * M + k];
            double r, wx, wy;
            pair_geom(c, i, j, &r, &wx, &wy);
            double vj = c->mass[j] / c->rho[j];
            double gx = L[0] * wx + L[1] * wy;
            double gy = L[2] * wx + L[3] * wy;
            double du = c->vx[j] - c->vx[i], dv = c->vy[j] - c->vy[i];
            t1 = t1 + (gx * du) * vj;
            t2 = t2 + (gy * du) * vj;
            t3 = t3 + (gx * dv) * vj;
            t4 = t4 + (gy * dv) * vj;
        }
        c->vxs[i] = t1 * xs + t2 * ys;
        c->vys[i] = t3 * xs + t4 * ys;
    }
    /* apply (:172-190) */
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < c->n; ++i) {
        if (c->tid[i] == 3 && !(c->flags[i] & ISPH_FLAG_BUFFER)) {
            c->x[i] = c->x[i] + c->xs[i];
            c->y[i] = c->y[i] + c->ys[i];
            c->vx[i] = c->vx[i] + c->vxs[i];
            c->vy[i] = c->vy[i] + c->vys[i];
        }
    }
    return ISPH_OK;
}

/* integrator.f90:908-976 timestep: umax/numax (:937-947) and, when enabled,
 * the adaptive dt (:950-960) with its float32 conversions. */
int ora_timestep(ora_ctx* c) {
    const isph_params* prm = &c->prm;
    double umax = 0.0, numax = 0.0;
    for (int32_t i = 0; i < c->n; ++i) {
        double v = fabs(c->vel[i]);
        double nt = fabs(c->nut[i]);
        if (v > umax) umax = v;
        if (nt > numax) numax = nt;
    }
    c->umax = umax;
    c->numax = numax;
    if (prm->adaptive_dt) {
        const double dl = prm->dl, s1 = prm->sig1, mu = prm->mu, rho = prm->rho0;
        double t = c->t, dt;
        if (t <= 1.0) {
            dt = fmin(fmin(s1 * dl / umax, s1 * (dl * dl) / ((mu / rho) + numax)), 0.0010);
        } else if (t > 1.0 && t <= 2.50) {
            float a = (float)(s1 * dl / umax), b = (float)(s1 * (dl * dl) * rho / mu);
            float m = a < b ? a : b;
            m = m < 0.002f ? m : 0.002f;
            dt = (double)m;
        } else if (t > 3.0 && t <= 4.0) {
            float a = (float)(s1 * dl / umax), b = (float)(s1 * (dl * dl) * rho / mu);
            float m = a < b ? a : b;
            m = m < 0.004f ? m : 0.004f;
            dt = (double)m;
        } else {
            dt = fmin(fmin(s1 * dl / umax, s1 * (dl * dl) / ((mu / rho) + numax)), 0.0050);
        }
        c->dt = dt;
    }
    /* transport substeps, integrator.f90:961-972 */
    if (!isfinite(numax)) c->numax = 0.0;
    {
        const double dl = prm->dl, dt = c->dt;
        double dts = (prm->sig1 * (dl * dl)) / (DM_SALT + c->numax / prm->tschmidt);
        int32_t steps = 1;
        if (dts < dt) {
            double q = ceil(dt / dts);
            if (!(q <= (double)MAX_SUBSTEPS)) {
                steps = 0;
                if (prm->salinity) {
                    c->dtsol = (dts < dt) ? dts : dt;
                    c->solsteps = 0;
                    return set_err(c, ISPH_E_CAPACITY, "salinity transport needs more than 2^20 substeps");
                }
            } else {
                steps = (int32_t)q;
                dts = dt / (double)steps;
            }
        }
        c->dtsol = (dts < dt) ? dts : dt;
        c->solsteps = steps;
    }
    return ISPH_OK;
}

/* ---------------------------------------------------------------------- */
/* scalar.f90 salinity transport and variable density                      */
/* ---------------------------------------------------------------------- */
static int salt_live(const ora_ctx* c, int32_t i) {
    return c->tid[i] == 3 && !(c->flags[i] & ISPH_FLAG_DEAD);
}

/* scalar.f90:114-236 scalart: cdiff of every live fluid particle, pairwise
 * over its stored list.  As in the reference, the dispersion of both
 * particles is evaluated per pair (:138-154, :160-170), the pair distance is
 * the stored dist(m) (positions of the last neighbour search) and the
 * position differences are current (:214-229). */
int ora_scalart(ora_ctx* c) {
    const isph_params* prm = &c->prm;
    const int32_t M = prm->max_nbr;
    const double al = 0.0025, at = 0.00025, Dm = DM_SALT, tor = 1.0;
    const double dl = prm->dl, dt = c->dt, lam = prm->lam, ts = prm->tschmidt;
    const double cap = (0.12 * (dl * dl)) / dt;
    isph_params hyb = *prm;
    hyb.kernel = ISPH_KERNEL_HYBRID;   /* hWabx/hWaby, kernel.f90:204-248 */
    if (!c->have_nbr) return set_err(c, ISPH_E_STATE, "scalart before neighbour");
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < c->n; ++i) {
        if (!salt_live(c, i)) continue;
        double cdiff = 0.0;
        const double con1 = c->con[i];
        const double vxi = c->vx[i], vyi = c->vy[i], pori = c->por[i];
        const double vi = sqrt(vxi * vxi + vyi * vyi);
        double Dxi = Dm + ((al * (vxi * vxi) + at * (vyi * vyi)) + ((al - at) * vxi) * vyi) / (vi * pori);
        Dxi = (Dxi > 0.0) ? Dxi : Dm;
        Dxi = (Dxi <= cap) ? Dxi : cap;
        double Dyi = Dm + ((at * (vxi * vxi) + al * (vyi * vyi)) + ((al - at) * vxi) * vyi) / (vi * pori);
        Dyi = (Dyi > 0.0) ? Dyi : Dm;
        Dyi = (Dyi <= cap) ? Dyi : cap;
        for (int32_t k = 0; k < c->nptr[i]; ++k) {
            int32_t j = c->nidx[(size_t)i * M + k];
            if (!salt_live(c, j)) continue;
            const double con2 = c->con[j];
            const double vxj = c->vx[j], vyj = c->vy[j], porj = c->por[j];
            const double vj = sqrt(vxj * vxj + vyj * vyj);
            double Dxj = Dm + ((al * (vxj * vxj) + at * (vyj * vyj)) + ((al - at) * vxj) * vyj) / (vj * porj);
            double Dyj = Dm + ((at * (vxj * vxj) + al * (vyj * vyj)) + ((al - at) * vxj) * vyj) / (vj * porj);
            Dxj = (Dxj > 0.0) ? Dxj : Dm;
            Dxj = (Dxj <= cap) ? Dxj : cap;
            Dyj = (Dyj > 0.0) ? Dyj : Dm;
            Dyj = (Dyj <= cap) ? Dyj : cap;
            double t1x, t1y;
            if (pori >= 0.8 || porj >= 0.8) {
                if (c->nut[i] < TINY6 && c->nut[j] < TINY6) {
                    t1x = Dm + ((c->nut[i] + c->nut[j]) * 0.50) / ts;
                } else {
                    t1x = 2.0 * (((Dm + c->nut[i] / ts) * (Dm + c->nut[j] / ts)) /
                                 ((Dm + c->nut[i] / ts) + (Dm + c->nut[j] / ts)));
                }
                t1y = t1x;
            } else {
                t1x = (((2.0 * tor) * Dxi) * Dxj) / (Dxi + Dxj);
                t1y = (((2.0 * tor) * Dyi) * Dyj) / (Dyi + Dyj);
            }
            const double t2 = 1.0;
            const double t3 = (c->ovol[i] * (pori + porj)) / (pori * porj);
            const double ex = c->xnb[i] - c->xnb[j], ey = c->ynb[i] - c->ynb[j];
            const double r = sqrt(ex * ex + ey * ey);
            const double dx = c->x[i] - c->x[j], dy = c->y[i] - c->y[j];
            double hx, hy;
            ora_kernel_gradw(&hyb, r, dx, dy, &hx, &hy);
            cdiff = (cdiff + (((((t1x * t2) * t3) * (con1 - con2)) * dx) * hx) / (r * r + lam)) +
                    (((((t1y * t2) * t3) * (con1 - con2)) * dy) * hy) / (r * r + lam);
        }
        c->cdiff[i] = cdiff;
    }
    return ISPH_OK;
}

/* scalar.f90:238-270 scalarupdate */
int ora_scalarupdate(ora_ctx* c, double stime) {
    const double xbuf = c->prm.salt_xbuf;
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < c->n; ++i) {
        if (!salt_live(c, i)) continue;
        c->con[i] = c->con[i] + c->cdiff[i] * stime;
        if ((c->flags[i] & ISPH_FLAG_BUFFER) && c->x[i] < xbuf) c->con[i] = 0.5;
    }
    return ISPH_OK;
}

/* scalar.f90:272-305 densityupdate (rhomax = rho rel_den, setup.f90:19) */
int ora_densityupdate(ora_ctx* c) {
    const double rhomax = c->prm.rho0 * c->prm.rel_den, rhomin = c->prm.rhomin;
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < c->n; ++i) {
        if (!salt_live(c, i)) continue;
        c->oden[i] = 0.5 * (rhomax + rhomin) + c->con[i] * (rhomax - rhomin);
        c->mass[i] = c->oden[i] * c->ovol[i];
        c->rho[i] = c->oden[i] * c->por[i];
    }
    return ISPH_OK;
}

/* main.f90:160-195: salt-water start */
int ora_salt_init(ora_ctx* c, double xl, double yl, double line_grad) {
    const double rel = c->prm.rel_den;
    for (int32_t i = 0; i < c->n; ++i) {
        if (c->tid[i] != 3) continue;
        c->con[i] = -0.50;
        if (((c->y[i] - yl) - line_grad * (c->x[i] - xl)) > 0.0) {
            c->mass[i] = c->mass[i] * rel;
            c->oden[i] = c->oden[i] * rel;
            c->rho[i] = c->rho[i] * rel;
            c->con[i] = 0.50;
        }
    }
    return ISPH_OK;
}

/* probe.f90:86-130 probevalue: Shepard pressure at each point over fluid
 * (tid > 2) within 2 blen h (blen = 1) in the point's cell and its eight
 * neighbours, visited in the order of probe.f90:92-94 ((i,j), (i,j-1),
 * (i,j+1), (i+1,j), (i-1,j), (i+1,j+1), (i+1,j-1), (i-1,j-1), (i-1,j+1)
 * with i the row (y) and j the column (x) of the cell). */
int ora_probe(ora_ctx* c, int32_t np, const double* px, const double* py, double* out) {
    static const int oy[9] = {0, 0, 0, 1, -1, 1, 1, -1, -1};
    static const int ox[9] = {0, -1, 1, 0, 0, 1, -1, -1, 1};
    const double h2x = 2.0 * c->prm.h;
    int oob = 0;
    for (int32_t k = 0; k < np; ++k) {
        int32_t cell = cell_of(c, px[k], py[k], &oob);
        int32_t cx = cell / c->ny, cy = cell % c->ny;
        double term1 = 0.0, pr = 0.0;
        int32_t count = 0;
        for (int q = 0; q < 9; ++q) {
            int32_t blo, bhi;
            column_bins(c, cx + ox[q], cy + oy[q], cy + oy[q], &blo, &bhi);
            for (int32_t j = c->bstart[blo]; j < c->bstart[bhi]; ++j) {
                if (c->tid[j] <= 2) continue;
                double dx = px[k] - c->x[j], dy = py[k] - c->y[j];
                double r = sqrt(dx * dx + dy * dy);
                if (!(r < h2x)) continue;
                double w = ora_kernel_w(&c->prm, r);
                term1 = term1 + (c->mass[j] * w) / c->rho[j];
                pr = pr + ((c->p[j] * c->mass[j]) * w) / c->rho[j];
                count++;
            }
        }
        out[k] = count ? pr / term1 : 0.0;
    }
    if (oob) return set_err(c, ISPH_E_DOMAIN, "probe outside the cell grid");
    return ISPH_OK;
}

int ora_get_substeps(ora_ctx* c, int32_t* solsteps, double* dtsol) {
    if (solsteps) *solsteps = c->solsteps;
    if (dtsol) *dtsol = c->dtsol;
    return ISPH_OK;
}

/* ---------------------------------------------------------------------- */
/* internal.f90:475-614 buffer zones, main.f90:103-121,218-236 schedule    */
/* ---------------------------------------------------------------------- */
int ora_set_buffer_zones(ora_ctx* c, int32_t nz, const isph_buffer_zone* zs, double shift) {
    if (nz < 0 || nz > ISPH_MAX_ZONES || (nz && !zs)) return set_err(c, ISPH_E_ARG, "bad zones");
    free_zones(c);
    const int64_t ncell = (int64_t)c->nx * c->ny;
    for (int k = 0; k < nz; ++k) {
        const isph_buffer_zone* q = &zs[k];
        if (!(q->period > 0.0)) { free_zones(c); return set_err(c, ISPH_E_ARG, "invalid buffer zone"); }
        unsigned char* mask = (unsigned char*)xmalloc((size_t)ncell);
        int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)ncell);
        for (int64_t e = 0; e < ncell; ++e) order[e] = -1;
        for (int32_t e = 0; e < q->ncells; ++e) {
            int32_t cx = q->cells[2 * e], cy = q->cells[2 * e + 1];
            if (cx < 1 || cx > c->nx - 2 || cy < 1 || cy > c->ny - 2) {
                free(mask); free(order); free_zones(c);
                return set_err(c, ISPH_E_ARG, "buffer entry cell outside the cell grid");
            }
            mask[cx * c->ny + cy] = 1;
            if (order[cx * c->ny + cy] < 0) order[cx * c->ny + cy] = e;
        }
        /* template particles in the entry cells, entry cell by entry cell,
         * each in template order (internal.f90:482-504) */
        int32_t* sel = (int32_t*)malloc(sizeof(int32_t) * (size_t)(q->ntemplate + 1));
        int32_t m = 0;
        for (int32_t e = 0; e < q->ncells; ++e) {
            int32_t cell = q->cells[2 * e] * c->ny + q->cells[2 * e + 1];
            if (order[cell] != e) continue;
            for (int32_t t = 0; t < q->ntemplate; ++t) {
                double fx = floor((q->x[t] - c->prm.xmin) / c->prm.cell_size);
                double fy = floor((q->y[t] - c->prm.ymin) / c->prm.cell_size);
                if (!(fx >= 0.0) || !(fy >= 0.0) || fx > c->nx - 3 || fy > c->ny - 3) continue;
                if (((int32_t)fx + 1) * c->ny + ((int32_t)fy + 1) == cell) sel[m++] = t;
            }
        }
        c->zone[k].x = (double*)xmalloc(sizeof(double) * (size_t)m);
        c->zone[k].y = (double*)xmalloc(sizeof(double) * (size_t)m);
        c->zone[k].vx = (double*)xmalloc(sizeof(double) * (size_t)m);
        c->zone[k].vy = (double*)xmalloc(sizeof(double) * (size_t)m);
        c->zone[k].p = (double*)xmalloc(sizeof(double) * (size_t)m);
        c->zone[k].flags = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)m);
        c->zone[k].key = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)m);
        for (int32_t u = 0; u < m; ++u) {
            int32_t t = sel[u];
            c->zone[k].x[u] = q->x[t]; c->zone[k].y[u] = q->y[t];
            c->zone[k].vx[u] = q->vx[t]; c->zone[k].vy[u] = q->vy[t];
            c->zone[k].p[u] = q->pressure[t]; c->zone[k].flags[u] = q->flags[t];
            int32_t cx = (int32_t)floor((q->x[t] - c->prm.xmin) / c->prm.cell_size) + 1;
            int32_t cy = (int32_t)floor((q->y[t] - c->prm.ymin) / c->prm.cell_size) + 1;
            c->zone[k].key[u] = c->prm.cell_order ? 10 * cell_rank(c, cx, cy) + 1
                                                  : 2 * (cx * c->ny + cy) + 1;
        }
        free(sel);
        free(order);
        c->zone[k].mask = mask;
        c->zone[k].nins = m;
        c->zone[k].mass = q->mass; c->zone[k].density = q->density;
        c->zone[k].ovol = q->ovol; c->zone[k].con = q->con;
        c->zone[k].period = q->period; c->zone[k].xcut = q->xcut;
        c->zone[k].remove_right = q->remove_right;
        c->zone[k].counter = 1;
        c->nzones = k + 1;
    }
    c->zone_shift = shift;
    return ISPH_OK;
}

/* remove_buffer1/2: buffer particles of the entry cells past the line leave,
 * the rest keep their order */
int ora_remove_buffer(ora_ctx* c, int32_t zi) {
    if (zi < 0 || zi >= c->nzones) return set_err(c, ISPH_E_ARG, "no such zone");
    int32_t m = 0;
    for (int32_t i = 0; i < c->n; ++i) {
        int32_t cx, cy;
        key_cell(c, c->key[i], &cx, &cy);
        int keep = 1;
        if (c->zone[zi].mask[cx * c->ny + cy] && (c->flags[i] & ISPH_FLAG_BUFFER))
            keep = c->zone[zi].remove_right ? (c->x[i] < c->zone[zi].xcut)
                                            : (c->x[i] > c->zone[zi].xcut);
        if (keep) c->perm[m++] = i;
    }
    if (m == c->n) return ISPH_OK;
    for (int32_t u = 0; u < m; ++u) c->tmpi[u] = c->key[c->perm[u]];
    memcpy(c->key, c->tmpi, sizeof(int32_t) * (size_t)m);
    apply_perm(c, m);
    c->n = m;
    c->have_nbr = 0;
    c->assembled = 0;
    return ISPH_OK;
}

/* insert_buffer1/2: the zone's particles appended as fluid */
int ora_insert_buffer(ora_ctx* c, int32_t zi) {
    if (zi < 0 || zi >= c->nzones) return set_err(c, ISPH_E_ARG, "no such zone");
    const int32_t m = c->zone[zi].nins, n0 = c->n;
    if ((int64_t)n0 + m > c->cap) return set_err(c, ISPH_E_CAPACITY, "buffer insertion exceeds the particle capacity");
    for (int32_t u = 0; u < m; ++u) {
        int32_t i = n0 + u;
        double vx = c->zone[zi].vx[u], vy = c->zone[zi].vy[u];
        c->x[i] = c->zone[zi].x[u]; c->y[i] = c->zone[zi].y[u];
        c->vx[i] = vx; c->vy[i] = vy;
        c->xs[i] = 0.0; c->ys[i] = 0.0; c->vxs[i] = 0.0; c->vys[i] = 0.0;
        c->mass[i] = c->zone[zi].mass; c->rho[i] = c->zone[zi].density;
        c->oden[i] = c->zone[zi].density; c->p[i] = c->zone[zi].p[u];
        c->por[i] = 1.0; c->R[i] = 0.0; c->nut[i] = 0.0; c->tke[i] = 0.0; c->varts[i] = 0.0;
        c->vel[i] = sqrt(vx * vx + vy * vy);
        c->con[i] = c->zone[zi].con; c->ovol[i] = c->zone[zi].ovol;
        c->tid[i] = 3; c->flags[i] = c->zone[zi].flags[u]; c->pid[i] = c->next_pid + u;
        c->key[i] = c->zone[zi].key[u];
    }
    c->n = n0 + m;
    c->next_pid += m;
    c->have_nbr = 0;
    c->assembled = 0;
    return ISPH_OK;
}

static int buffer_schedule(ora_ctx* c, double told, double tnew) {
    for (int k = 0; k < c->nzones; ++k) {
        double at = (double)c->zone[k].counter * c->zone[k].period;
        if ((told + c->zone_shift) < at && (tnew + c->zone_shift) >= at) {
            int rc = ora_remove_buffer(c, k);
            if (rc) return rc;
            if ((rc = ora_insert_buffer(c, k))) return rc;
            c->zone[k].counter += 1;
        }
    }
    return ISPH_OK;
}

/* One step in the order of main.f90:41-66 (t advances first, :42-43); with
 * salinity in the order of main.f90:199-263 (buffer zones after the
 * transport, :217-235):
 * the transport substeps on the previous step's lists first (:204-215),
 * densityupdate after the shift (:258), eddyvis after timestep (:260). */
int ora_step(ora_ctx* c, int32_t nsteps) {
    int rc;
    const int salt = c->prm.salinity != 0;
    for (int32_t s = 0; s < nsteps; ++s) {
        const double told = c->t;
        c->t = c->t + c->dt;
        if (salt && c->have_nbr) {
            for (int32_t k = 0; k < c->solsteps; ++k) {
                if ((rc = ora_scalart(c))) return rc;
                if ((rc = ora_scalarupdate(c, c->dtsol))) return rc;
            }
        }
        if (c->nzones && (rc = buffer_schedule(c, told, c->t))) return rc;
        if ((rc = ora_projection(c))) return rc;
        if ((rc = ora_cellshift(c))) return rc;
        if ((rc = ora_resetid(c))) return rc;
        if ((rc = ora_neighbour(c))) return rc;
        if ((rc = ora_effpor(c))) return rc;
        if ((rc = ora_eddyvis(c))) return rc;
        if ((rc = ora_int_vel(c))) return rc;
        if ((rc = ora_ppesolve(c, NULL, NULL))) return rc;
        if ((rc = ora_comp_vel(c))) return rc;
        if ((rc = ora_comp_pos(c))) return rc;
        if ((rc = ora_cellshiftalt(c))) return rc;
        if ((rc = ora_neighbour(c))) return rc;
        if ((rc = ora_freesurf(c))) return rc;
        if ((rc = ora_effpor(c))) return rc;
        if ((rc = ora_opt2_shift(c))) return rc;
        if (salt && (rc = ora_densityupdate(c))) return rc;
        if ((rc = ora_timestep(c))) return rc;
        if (salt && (rc = ora_eddyvis_lagged(c))) return rc;
    }
    return ISPH_OK;
}
