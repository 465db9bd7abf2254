/*
 * isph_capi.h -- C ABI of the gfx950 ISPH step (libisph.so).
 *
 * The reference solver (a Fortran 2008 / OpenMP 2-D ISPH code) keeps its
 * particle state in module globals and exposes one zero-argument subroutine
 * per stage of a time step; `main.f90:41-66` calls them in a fixed order.
 * This header replaces each of those subroutines by one `extern "C"` entry
 * point that takes an opaque context.  All particle state lives in HBM
 * inside the context; the Fortran driver (iso_c_binding, see
 * beach_slope_confined_amd/fortran/isph_capi.f90) and the Python harness
 * (ctypes) only pass plain pointers and sizes.
 *
 * Conventions
 *   - every function returns an int status: ISPH_OK (0) or a negative
 *     ISPH_E_* code; isph_last_error() returns a message for the last one.
 *     Nothing aborts or exits across the ABI.
 *   - one host thread per context; all work is ordered on the context's own
 *     HIP stream (default priority) and every entry point is synchronous
 *     when it returns (isph_step() synchronises once per call).
 *   - host pointers are borrowed for the duration of the call.
 *   - particle arrays are structure-of-arrays, 0-based, in the context's
 *     current order.  After isph_cellshift()/isph_cellshiftalt() that order
 *     is the cell-sorted order, which is also the PPE row/column order (the
 *     reference's `matid`, `isph.f90:88-113`).  `pid` travels with the
 *     particle and identifies it across re-orderings.  pids must be unique
 *     and non-negative (not checked): buffer insertion hands out max pid + 1
 *     upward, and the adjoint's warm starts are kept by pid (with repeated
 *     pids they, and so the gradient within solver tolerance, vary from run
 *     to run).  A pid space much larger than the particle count costs the
 *     adjoint its warm starts, not memory.
 *   - particle types follow the reference (`particle.f90:92`,
 *     `boundary.f90:50-83`): tid 1 = fixed wall (innermost layer),
 *     tid 2 = dummy wall, tid 3 = fluid.  Soil particles (the reference's
 *     tid 4 `porlist`, `boundary.f90:523-593`) are passed separately with
 *     isph_set_soil().
 */
#ifndef ISPH_CAPI_H
#define ISPH_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define ISPH_OK            0
#define ISPH_E_ARG        -1   /* bad argument or size                         */
#define ISPH_E_HIP        -2   /* HIP runtime error                            */
#define ISPH_E_CAPACITY   -3   /* neighbour list or particle capacity exceeded */
#define ISPH_E_NAN        -4   /* non-finite value produced                    */
#define ISPH_E_NOCONV     -5   /* PPE solve did not converge / broke down      */
#define ISPH_E_DOMAIN     -6   /* particle outside the cell grid               */
#define ISPH_E_SINGULAR   -7   /* PPE row with zero or non-finite diagonal     */
#define ISPH_E_STATE      -8   /* call out of order (e.g. solve before assembly) */

/* ---- enumerations ------------------------------------------------------ */
#define ISPH_KERNEL_HYBRID       0   /* kernel.f90:52-62,113-124 (alp = 0.1)    */
#define ISPH_KERNEL_WENDLAND_C2  1   /* kernel.f90:42-43,75,102-103             */

#define ISPH_DRAG_REFERENCE 0  /* porous.f90:192,204-212: lamk = alpha, vis(oden) */
#define ISPH_DRAG_ERGUN     1  /* R = -alpha nu (1-n)^2/(n^3 d50^2) - beta |v| (1-n)/(n^3 d50) */

#define ISPH_PREC_FP64      0  /* FP64 matrix, FP64 BiCGSTAB                        */
#define ISPH_PREC_MIXED     1  /* FP32 matrix values and FP32 Krylov vectors, FP64
                                  dot products, FP64 residual + reliable updates     */
#define ISPH_PREC_MIXED_V64 2  /* FP32 matrix values, FP64 Krylov vectors, FP64
                                  residual + reliable updates                        */
#define ISPH_PREC_MIXED_MG  3  /* as MIXED_V64, right-preconditioned by one FP32
                                  V-cycle of an aggregation multigrid on the cell
                                  grid (cells, then 2x2 blocks of cells)             */
#define ISPH_PREC_IR        5  /* iterative refinement: FP64 outer residual
                                  b - A64 x and update x += d; each correction d
                                  solves A d = r with FP32 matrix values and an
                                  FP32-vector BiCGSTAB (FP64 dot products) started
                                  from zero (rhat = r, p = r) and stopped when its
                                  recurrence residual fell by inner_tol             */
#define ISPH_PREC_FP64_MG   6  /* MIXED_MG's algorithm with every value FP64: the
                                  FP64 matrix in the outer products and an FP64
                                  V-cycle on FP64 Galerkin operators (the baseline
                                  that isolates what FP32 buys in MIXED_MG)         */
#define ISPH_PREC_FGMRES    4  /* FP64 restarted GMRES, the reference's live solve
                                  (solver.f90:306-347, MKL dfgmres without a
                                  preconditioner): relative tolerance `tol` on the
                                  residual estimate against the initial residual
                                  (1e-3 there), zero-norm test 1e-12, restart length
                                  inner_maxit (0: min(150, n)), at most `maxit`
                                  iterations (150 there); warm start from the last
                                  solution in row order (fsol); without convergence
                                  the pressure keeps that warm start               */

/* Particle order kept by cellshift/cellshiftalt (isph_params.cell_order).
 * ISPH_ORDER_CELL: by cell (column-major, y ascending), walls first, then
 *   fluid, each in its previous order.
 * ISPH_ORDER_REFERENCE: the reference's plist order after cellshift and
 *   resetid (integrator.f90:250-391, isph.f90:88-113): cells numbered column
 *   by column with y descending; in a cell the walls, then the fluid that
 *   stayed, then arrivals from the cells below, above, right, left, upper
 *   left, lower left, upper right and lower right, each in its previous
 *   order; a fluid particle moves at most one cell per re-binning.  effpor
 *   then visits the soil in porous.f90:127-176's 2x2-cell order.  Needs
 *   cell_size >= 4h, and cell boundaries at the reference's: xmin and ymin
 *   multiples of cell_size (its cells start at x = 0 and y = 0,
 *   setup.f90:120-123). */
#define ISPH_ORDER_CELL      0
#define ISPH_ORDER_REFERENCE 1

/* particle flag bits (particle.f90:123-124) */
#define ISPH_FLAG_DEAD      1
#define ISPH_FLAG_BUFFER    2
#define ISPH_FLAG_FREE      4
#define ISPH_FLAG_VICINITY  8

/* Run parameters.  The reference keeps these as module globals
 * (initialize.f90:19-79) and scattered constants (kernel.f90:30,
 * porous.f90:111, part_shift.f90:18,71, turbulence.f90:16); setup.f90:47-82
 * derives h1, dl and lam from them.  All doubles come first, then int32s,
 * so that C, Fortran bind(C) and ctypes agree on the layout. */
typedef struct isph_params {
    double h;          /* smoothing length (reference h1)                      */
    double cell_size;  /* cell edge, >= 2h (reference uses 4h + quadrants)    */
    double xmin, ymin; /* lower-left corner of the cell grid                   */
    double xmax, ymax; /* upper-right corner of the cell grid                  */
    double dl;         /* particle spacing (setup.f90:53)                      */
    double lam;        /* eta in 1/(r^2+eta), 1e-3 h^2 (setup.f90:82)          */
    double dt;         /* time step                                            */
    double rho0;       /* reference density `rho` (initialize.f90:21)          */
    double rel_den;    /* salt/fresh density ratio (initialize.f90:25)         */
    double g;          /* signed gravity along y (initialize.f90:21)           */
    double nu;         /* kinematic viscosity for ISPH_DRAG_ERGUN              */
    double csh;        /* shifting coefficient (initialize.f90:22)             */
    double maxshift;   /* shift cap as a fraction of dl (initialize.f90:36)    */
    double lamfs;      /* free-surface factor (initialize.f90:42)              */
    double maxdivr;    /* free-surface divergence scale (initialize.f90:41)    */
    double por;        /* porosity of the soil skeleton (initialize.f90:24)    */
    double alpha;      /* Darcy coefficient (reference lamk, porous.f90:111)   */
    double beta;       /* Forchheimer coefficient (ISPH_DRAG_ERGUN only)       */
    double d50;        /* grain diameter (porous.f90:111)                      */
    double tol;        /* PPE relative residual tolerance ||b-Ax||/||b||        */
    double inner_tol;  /* mixed precision: reliable-update factor delta (the FP64
                          residual replaces the FP32 one each time it fell by
                          delta); ISPH_PREC_IR: the corrections' reduction.
                          0 selects the precision's default: 0.1, and 1e-3 for
                          MIXED_MG and FP64_MG (whose reliable updates cost a
                          FP64 residual pass against about 8 iterations'
                          reduction).  Range [0, 1).                         */
    double xlcutoff;   /* outflow sink (integrator.f90:60-61), active if sink  */
    double ytcutoff;
    double sig1;       /* adaptive dt factor (initialize.f90:35)               */
    double mu;         /* dynamic viscosity used by adaptive dt (initialize.f90:35) */
    double rhomin;     /* fresh-water density of the salinity law (initialize.f90:22);
                          rhomax = rho0 * rel_den (setup.f90:19)               */
    double tschmidt;   /* turbulent Schmidt number (initialize.f90:23)         */
    double salt_xbuf;  /* buffer particles with x < salt_xbuf are reset to
                          con = 0.5 after each transport substep (scalar.f90:252) */
    double xrelease_lo; /* comp_pos releases a buffer particle (it becomes fluid)
                           once xrelease_lo < x < xrelease_hi (integrator.f90:888-897:
                           8 prrealx + brrealx + 0.001 and xrcutoff, setup.f90:76);
                           the defaults (both 1e300) release nothing */
    double xrelease_hi;
    int32_t kernel;      /* ISPH_KERNEL_*                                       */
    int32_t drag_law;    /* ISPH_DRAG_*                                         */
    int32_t precision;   /* ISPH_PREC_*                                         */
    int32_t maxit;       /* BiCGSTAB iteration cap                              */
    int32_t inner_maxit; /* FGMRES restart length (0: min(150, n))              */
    int32_t max_outer;   /* mixed precision: cap on reliable updates            */
    int32_t max_nbr;     /* neighbour-list capacity per particle                */
    int32_t shifting;    /* 1: opt2_shift moves particles                       */
    int32_t adaptive_dt; /* 1: isph_timestep() updates dt (integrator.f90:950-960) */
    int32_t sink;        /* 1: isph_cellshiftalt() deletes outflow particles    */
    int32_t visc_law;    /* 0: mu = vis(oden) (porous.f90:12-30); 1: mu = nu*oden */
    int32_t salinity;    /* 1: isph_step() runs the salt-water step of main.f90:199-263
                            (transport substeps first, densityupdate after the
                            shift, eddyvis after timestep); 0: main.f90:41-66    */
    int32_t cell_order;  /* ISPH_ORDER_*: particle order kept by the re-binnings */
    int32_t reserved;    /* keeps the struct a multiple of 8 bytes; set to 0    */
} isph_params;

typedef struct isph_ctx isph_ctx;

/* Fields readable with isph_get_field() (double arrays of n, or 4n for L). */
#define ISPH_FIELD_XS       0   /* projection's old x, then shift dx          */
#define ISPH_FIELD_YS       1
#define ISPH_FIELD_VXS      2   /* intermediate velocity v*, then shift dv     */
#define ISPH_FIELD_VYS      3
#define ISPH_FIELD_POROSITY 4   /* n  (porous.f90:188)                          */
#define ISPH_FIELD_RESIST   5   /* R  (porous.f90:204-212, resistx = resisty)   */
#define ISPH_FIELD_COFF     6   /* KGC matrix L, 4 per particle (search.f90:276-278) */
#define ISPH_FIELD_DIVR     7   /* gradvx = A11 + A22 (search.f90:272)          */
#define ISPH_FIELD_NUT      8   /* eddy viscosity (turbulence.f90:80)           */
#define ISPH_FIELD_TKE      9   /* turbulent kinetic energy (turbulence.f90:83) */
#define ISPH_FIELD_VARTS   10   /* free-surface artificial viscosity (integrator.f90:495) */
#define ISPH_FIELD_VEL     11   /* |v|/n (integrator.f90:825)                   */
#define ISPH_FIELD_RHS     12   /* Jacobi-scaled PPE right-hand side            */
#define ISPH_FIELD_X       13   /* position x (particle.f90:33)                 */
#define ISPH_FIELD_Y       14   /* position y                                   */
#define ISPH_FIELD_VX      15   /* velocity x (particle.f90:34)                 */
#define ISPH_FIELD_VY      16   /* velocity y                                   */
#define ISPH_FIELD_DENSITY 17   /* density rho = n * oden (porous.f90:190)       */
#define ISPH_FIELD_PRESSURE 18  /* pressure (isph.f90:465-483)                  */
#define ISPH_FIELD_CON     19   /* salinity con in [-0.5, 0.5] (particle.f90:37)  */
#define ISPH_FIELD_OVOL    20   /* particle volume fmass/rho (internal.f90:103)   */
#define ISPH_FIELD_CDIFF   21   /* d con/dt of the last scalart (scalar.f90:129) */
#define ISPH_FIELD_MASS    22   /* mass                                          */
#define ISPH_FIELD_ODEN    23   /* fluid density oden (scalar.f90:293)           */

/* ---- context ------------------------------------------------------------ */
/* Allocates every device buffer for up to cap_particles particles and
 * cap_soil soil particles (replaces memory.f90:10-147). */
int isph_create(const isph_params* prm, int32_t cap_particles, int32_t cap_soil,
                isph_ctx** out);
int isph_destroy(isph_ctx* ctx);
const char* isph_last_error(const isph_ctx* ctx);
int isph_set_params(isph_ctx* ctx, const isph_params* prm);
int isph_get_params(const isph_ctx* ctx, isph_params* prm);

/* ---- state transfer ----------------------------------------------------- */
/* Uploads particles (replaces gen_boundary/gen_fluid's placement into
 * dpcell, boundary.f90:12, internal.f90:10) and bins them into cells.
 * flags may be NULL (all zero).  Sets vel = |v|, umax = 0, t = 0. */
int isph_set_particles(isph_ctx* ctx, int32_t n,
                       const double* x, const double* y,
                       const double* vx, const double* vy,
                       const double* mass, const double* density,
                       const double* oden, const double* pressure,
                       const int32_t* tid, const int32_t* flags,
                       const int32_t* pid);
/* Downloads particles in the current order; any pointer may be NULL. */
int isph_get_particles(isph_ctx* ctx, int32_t* n,
                       double* x, double* y, double* vx, double* vy,
                       double* mass, double* density, double* oden,
                       double* pressure, int32_t* tid, int32_t* flags,
                       int32_t* pid);
/* Soil skeleton particles with mass and density (porlist,
 * boundary.f90:569-588); porsearch uses m/rho (porous.f90:54-55). */
int isph_set_soil(isph_ctx* ctx, int32_t ns, const double* x, const double* y,
                  const double* mass, const double* density);
int isph_get_field(isph_ctx* ctx, int32_t field, double* out);
int isph_set_field(isph_ctx* ctx, int32_t field, const double* in);
/* Neighbour lists of the last isph_neighbour() as CSR over the current
 * order: ptr[n+1], idx[ptr[n]] ascending per row (the verlet list,
 * interactions.f90:26-35).  Pass idx = NULL to read ptr only. */
int isph_get_neighbours(isph_ctx* ctx, int32_t* ptr, int32_t* idx);
/* PPE matrix of the last assembly as CSR with the diagonal in place
 * (fmatrix/fvec after Jacobi scaling, isph.f90:423-442): rowptr[n+1],
 * col[nnz] strictly ascending per row, val[nnz], rhs[n].  Any of col, val,
 * rhs may be NULL. */
int isph_get_csr(isph_ctx* ctx, int32_t* rowptr, int32_t* col, double* val,
                 double* rhs);
/* Scalars: umax, dt, t, and the last solve's statistics. */
int isph_get_scalars(isph_ctx* ctx, double* umax, double* dt, double* t,
                     int32_t* iters, int32_t* outer, double* relres);
int isph_set_scalars(isph_ctx* ctx, double umax, double dt, double t);

/* ---- one entry per reference subroutine (main.f90:41-66 order) --------- */
int isph_projection(isph_ctx* ctx);    /* isph.f90:16-47                     */
int isph_cellshift(isph_ctx* ctx);     /* integrator.f90:250-391             */
int isph_resetid(isph_ctx* ctx);       /* isph.f90:88-113                    */
int isph_neighbour(isph_ctx* ctx);     /* search.f90:310-399 + nsearch2 :13-308 */
int isph_effpor(isph_ctx* ctx);        /* porous.f90:107-224 (+ porsearch :32) */
int isph_eddyvis(isph_ctx* ctx);       /* turbulence.f90:12-110              */
/* eddyvis with the pair distances and kernel gradients of the last
 * isph_neighbour() (the reference's stored klt, turbulence.f90:36) instead
 * of the current positions: the eddyvis of the salt-water step, which
 * main.f90:260 calls after opt2_shift has moved the particles. */
int isph_eddyvis_lagged(isph_ctx* ctx);
int isph_int_vel(isph_ctx* ctx);       /* integrator.f90:458-713             */
/* ppesolve = assembly (isph.f90:123-456) + solve (:460) + scatter (:465-483) */
int isph_ppe_assemble(isph_ctx* ctx);
int isph_ppe_solve(isph_ctx* ctx, int32_t* iters, double* relres);
int isph_ppesolve(isph_ctx* ctx, int32_t* iters, double* relres);
int isph_comp_vel(isph_ctx* ctx);      /* integrator.f90:715-865             */
int isph_comp_pos(isph_ctx* ctx);      /* integrator.f90:867-906             */
int isph_cellshiftalt(isph_ctx* ctx);  /* integrator.f90:32-248              */
int isph_freesurf(isph_ctx* ctx);      /* isph.f90:49-86                     */
int isph_opt2_shift(isph_ctx* ctx);    /* part_shift.f90:12-192              */
int isph_timestep(isph_ctx* ctx);      /* integrator.f90:908-976             */
/* nsteps whole steps in the main.f90:41-66 order (no host round trip
 * between stages); with params.salinity = 1 in the order of main.f90:199-263
 * (without the buffer insertion of :217-235). */
int isph_step(isph_ctx* ctx, int32_t nsteps);

/* ---- salinity transport and variable density (scalar.f90) --------------
 * con is the salinity of a fluid particle, -0.5 (fresh) to 0.5 (salt); the
 * fluid density follows it linearly between rhomin and rhomax.  The
 * transport uses the neighbour lists, pair distances and porosity of the
 * last isph_neighbour()/isph_effpor() with the current positions and
 * velocities, as the reference's stored lists do when scalart runs at the
 * start of a step.  isph_set_particles() sets con = 0 and ovol = mass/oden. */
int isph_scalart(isph_ctx* ctx);                  /* scalar.f90:114-236: cdiff */
int isph_scalarupdate(isph_ctx* ctx, double stime); /* scalar.f90:238-270        */
int isph_densityupdate(isph_ctx* ctx);            /* scalar.f90:272-305        */
/* Salt-water start of main.f90:160-195: every fluid particle gets
 * con = -0.5; those above the line (y - yl) - line_grad (x - xl) > 0 get
 * mass, oden and density times rel_den and con = 0.5. */
int isph_salt_init(isph_ctx* ctx, double xl, double yl, double line_grad);
/* Transport substeps of the next step and their length, as the last
 * isph_timestep() set them (integrator.f90:961-972); 0 substeps before the
 * first timestep, as the reference's solsteps = 0 (initialize.f90:56). */
int isph_get_substeps(isph_ctx* ctx, int32_t* solsteps, double* dtsol);

/* ---- inflow/outflow buffers (internal.f90:475-614, main.f90:103-121,218-236) */
/* A buffer zone: its entry cells, the template lattice inserted into them
 * every `period` seconds of simulated time, and the line past which its
 * buffer particles leave.  insert_buffer1/2 append the template particles
 * that lie in the entry cells as fluid (tid 3) with the zone's mass,
 * density = oden, ovol and con; remove_buffer1 drops the buffer particles of
 * the entry cells with x >= xcut (remove_right = 1), remove_buffer2 those
 * with x <= xcut (remove_right = 0). */
#define ISPH_MAX_ZONES 4
typedef struct isph_buffer_zone {
    int32_t ntemplate;     /* template particles                               */
    int32_t ncells;        /* entry cells                                      */
    const double *x, *y, *vx, *vy, *pressure;   /* template state (ntemplate)  */
    const int32_t* flags;  /* template flags: ISPH_FLAG_BUFFER, ISPH_FLAG_DEAD */
    const int32_t* cells;  /* ncells (cx, cy) pairs on the context's cell grid:
                              cx = floor((x - xmin)/cell_size) + 1, likewise cy */
    double mass, density, ovol, con;  /* of every inserted particle            */
    double period;         /* insertion period (the reference's ins_1, ins_2)  */
    double xcut;           /* removal line (rv_buf_r, rv_buf_l)                */
    int32_t remove_right;  /* 1: x >= xcut leaves; 0: x <= xcut leaves         */
    int32_t reserved;      /* set to 0                                         */
} isph_buffer_zone;
/* Installs up to ISPH_MAX_ZONES zones (0 removes them) and restarts their
 * schedules: at each isph_step, with told and t the times before and after
 * its advance, zone k runs remove then insert when
 * told + time_shift < m_k period_k <= t + time_shift, and m_k then counts on
 * (m_k = 1 after this call, main.f90:93-94,194-196).  In a salt-water step
 * this happens after the transport substeps, as main.f90:210-236 orders it. */
int isph_set_buffer_zones(isph_ctx* ctx, int32_t nzones, const isph_buffer_zone* zones,
                          double time_shift);
int isph_remove_buffer(isph_ctx* ctx, int32_t zone);   /* remove_buffer1/2 */
int isph_insert_buffer(isph_ctx* ctx, int32_t zone);   /* insert_buffer1/2 */

/* ---- text output and probes (output.f90, probe.f90) --------------------- */
/* Fortran edit descriptors Fw.d and ESw.d as a Fortran run-time library
 * prints them (correct rounding, ties to even; w asterisks when the value
 * does not fit; ESw.d exponents E+dd, or +ddd without the E from 100 to 999;
 * Inf, -Inf, NaN right-justified).  out receives w characters and a NUL. */
int isph_format_f(double v, int32_t w, int32_t d, char* out);
int isph_format_es(double v, int32_t w, int32_t d, char* out);
/* One snapshot record, output.f90:206-209:
 * (F10.3,1X,F10.3,1X,F10.3,1X,ES10.3,1X,ES10.3,1X,F10.5) of
 * x, y, p, vx, vy, con + 0.5; out receives 65 characters and a NUL. */
int isph_snapshot_record(double x, double y, double p, double vx, double vy,
                         double con, char* out);
/* combined (output.f90:188-234): writes <dir>/fluid/fluid_<index>.txt (live
 * fluid), <dir>/buffer/buffer_<index>.txt (live buffer fluid) and
 * <dir>/fluid/dead_<index>.txt (dead non-buffer fluid), one record per
 * particle in the context's cell order.  The folders must exist. */
int isph_combined(isph_ctx* ctx, const char* dir, int32_t index);
/* probevalue (probe.f90:86-130): Shepard pressure at np points,
 * sum_j p_j m_j W_j / rho_j over sum_j m_j W_j / rho_j for fluid j within
 * 2h in the 3x3 cells around the point's cell; 0 where no fluid is in range. */
int isph_probe(isph_ctx* ctx, int32_t np, const double* px, const double* py,
               double* out);
/* isph_probe without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving or sampling loops, harnesses that
 * forbid hipMalloc).  Same arguments, results (bit for bit) and errors as
 * isph_probe, except that np may not exceed ISPH_PROBE_NOALLOC_MAX
 * (ISPH_E_ARG beyond it).  The points and values are staged in a fixed
 * device array of 3 * ISPH_PROBE_NOALLOC_MAX doubles that every context on
 * a GPU shares and that holds the last call's contents; calls that overlap
 * in time on one GPU (from several threads or contexts) race.  Synchronous
 * when it returns. */
#define ISPH_PROBE_NOALLOC_MAX 4096
int isph_probe_noalloc(isph_ctx* ctx, int32_t np, const double* px, const double* py,
                       double* out);
/* The probe series file of output.f90:150-166: n records (F10.3,1X,F10.3)
 * of a[i], b[i] (the reference writes t sqrt(|g|/wc) and
 * p / (rho |g| wc), probe.f90:126-127). */
int isph_write_probe_series(const char* path, int32_t n, const double* a,
                            const double* b);

/* ---- in-memory checkpoints (no reference counterpart: the reference has no
 * restart, output.f90 only writes text snapshots) ---------------------------- */
/* Saves the particle state, the cell order and the scalars (umax, dt, t) to
 * device-memory slot `slot` (0 <= slot < 64); the slot's memory is
 * allocated on first use and kept until isph_destroy.  isph_restore puts
 * the state back; neighbour lists and the PPE matrix are rebuilt by the
 * next step. */
int isph_checkpoint(isph_ctx* ctx, int32_t slot);
int isph_restore(isph_ctx* ctx, int32_t slot);

/* ---- instrumentation (no reference counterpart) -------------------------- */
/* enable = s > 0: HIP events bracket every s-th launch of each BiCGSTAB SpMV
 * kernel and of the scalart kernel on the context stream (0 turns it off);
 * isph_get_kernel_time() returns the timed launches and their total
 * milliseconds since the last enable.  kernel: 0 = FP64 step-1 SpMV
 * (p update, v = A p), 1 = FP64 step-2 SpMV (s, t = A s), 2 and 3 = the
 * same for the mixed-precision flavours, 4 = scalart. */
int isph_profile_enable(isph_ctx* ctx, int32_t enable);
/* Progress of the long calls: fn(user, what, done, total) is called on the
 * calling thread after each checkpoint segment of isph_adjoint_forward
 * (what = "adjoint_forward") and of isph_adjoint_backward
 * ("adjoint_backward").  NULL turns it off.  fn must not call into the
 * same context. */
typedef void (*isph_progress_fn)(void* user, const char* what, int32_t done, int32_t total);
int isph_set_progress(isph_ctx* ctx, isph_progress_fn fn, void* user);
int isph_get_kernel_time(isph_ctx* ctx, int32_t kernel, int64_t* launches,
                         double* total_ms);
/* Copies `bytes` bytes device-to-device with the library's copy kernel
 * `reps` times on the context stream; returns the mean ms per copy.  Used to
 * calibrate attainable HBM bandwidth and the FETCH_SIZE/WRITE_SIZE counters. */
int isph_copy_bench(isph_ctx* ctx, int64_t bytes, int32_t reps, double* ms);

/* ---- discrete adjoint (no reference counterpart: the reference has no
 * derivatives; this is the calibration path for the porous parameters) ---- */
/* Objective J = (1/nsteps) sum over steps of [ sum_k P_k + sum_g D_g ]
 * evaluated after every step:
 *   P_k  Shepard pressure at probe k, sum_j p_j m_j W_j / rho_j over
 *        sum_j m_j W_j / rho_j, fluid j within 2h (the reference's probe,
 *        probe.f90:86-130); 0 when no fluid particle is within 2h;
 *   D_g  water depth at x = gauge_x[g]: sum_j (m_j / rho_j) w(x_j - x_g) over
 *        fluid j with the normalised 1-D Wendland C2 weight
 *        w(d) = 5/(8h) (1 - q/2)^3 (1.5 q + 1), q = |d|/h < 2.
 * At most 8 probes and 8 gauges. */
#define ISPH_OBJ_MAX 8
typedef struct isph_objective {
    int32_t nprobe, ngauge;
    double probe_x[ISPH_OBJ_MAX], probe_y[ISPH_OBJ_MAX];
    double gauge_x[ISPH_OBJ_MAX];
} isph_objective;

/* Forward run for the adjoint: porous closure + timestep at the current
 * state (as before the reference's time loop, main.f90:35-37), then nsteps
 * steps; the state at the start of every `ckpt_every`-th step and the final
 * state are kept in device memory.  Returns J.  Needs a fixed dt
 * (adaptive_dt = 0) and no outflow sink. */
int isph_adjoint_forward(isph_ctx* ctx, const isph_objective* obj, int32_t nsteps,
                         int32_t ckpt_every, double* J);
/* Reverse sweep of the last isph_adjoint_forward (same parameters): for each
 * checkpoint segment, last first, recomputes its steps with a per-step tape
 * and runs the adjoint of every stage in reverse order; the PPE adjoint is
 * one BiCGSTAB solve with the transposed matrix per step.  Returns
 * grad = dJ/d(por, alpha, beta), J, and the total BiCGSTAB iterations of the
 * adjoint solves.  Leaves the context at the final state of the forward run. */
int isph_adjoint_backward(isph_ctx* ctx, double* grad3, double* J, int64_t* adj_iters);

/* Stage adjoints on the current state, for testing the reverse sweep one
 * stage at a time.  The adjoint arrays (same field ids as isph_get_field;
 * ISPH_FIELD_COFF is the 4-per-particle L) hold the adjoint of the stage's
 * outputs on entry and are updated in place to the adjoint of its inputs.
 * The state must be the stage's input state (for ISPH_ADJ_PPESOLVE: after
 * the forward solve, whose pressure it uses); list stages use the lists and
 * L of the last isph_neighbour. */
#define ISPH_ADJ_PROJECTION 1
#define ISPH_ADJ_NEIGHBOUR  3   /* KGC matrices L = inv(A)             */
#define ISPH_ADJ_EFFPOR     4
#define ISPH_ADJ_EDDYVIS    5
#define ISPH_ADJ_VARTS      6
#define ISPH_ADJ_VSTAR      7
#define ISPH_ADJ_PPESOLVE   8   /* assembly + transposed solve         */
#define ISPH_ADJ_COMP_VEL  10
#define ISPH_ADJ_COMP_POS  11
#define ISPH_ADJ_SHIFT     16   /* opt2_shift + its position/velocity update */
#define ISPH_ADJ_OBJECTIVE 20   /* one step's objective term, weight 1 */
int isph_adjoint_stage(isph_ctx* ctx, int32_t stage, const isph_objective* obj);
int isph_adjoint_zero(isph_ctx* ctx);
int isph_adjoint_get(isph_ctx* ctx, int32_t field, double* out);
int isph_adjoint_set(isph_ctx* ctx, int32_t field, const double* in);
/* dJ/d(por, alpha, beta) accumulated by the stage adjoints, and dJ/d umax */
int isph_adjoint_params(isph_ctx* ctx, double* grad3, double* umax_bar);
/* one step's objective term at the current state */
int isph_objective_eval(isph_ctx* ctx, const isph_objective* obj, double* J);

#ifdef __cplusplus
}
#endif
#endif /* ISPH_CAPI_H */
