/*
 * isph_oracle.h -- CPU restatement of the reference ISPH step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the gfx950
 * library (libisph.so): only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never calls it.
 *
 * It restates the reference Fortran routines (file:line cited per function
 * in isph_oracle.c) with the same expression order, in the same particle
 * order and with the same reduction trees as the GPU kernels, so that the
 * two can be compared bit for bit.  The API mirrors include/isph_capi.h
 * with an `ora_` prefix and host arrays.
 */
#ifndef ISPH_ORACLE_H
#define ISPH_ORACLE_H

#include <stdint.h>
#include "../include/isph_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_ctx ora_ctx;

int ora_create(const isph_params* prm, int32_t cap_particles, int32_t cap_soil,
               ora_ctx** out);
int ora_destroy(ora_ctx* c);
const char* ora_last_error(const ora_ctx* c);
int ora_set_threads(int32_t nthreads);
int ora_set_params(ora_ctx* c, const isph_params* prm);
int ora_set_particles(ora_ctx* c, int32_t n, const double* x, const double* y,
                      const double* vx, const double* vy, const double* mass,
                      const double* density, const double* oden,
                      const double* pressure, const int32_t* tid,
                      const int32_t* flags, const int32_t* pid);
int ora_get_particles(ora_ctx* c, int32_t* n, double* x, double* y, double* vx,
                      double* vy, double* mass, double* density, double* oden,
                      double* pressure, int32_t* tid, int32_t* flags,
                      int32_t* pid);
int ora_set_soil(ora_ctx* c, int32_t ns, const double* x, const double* y,
                 const double* mass, const double* density);
int ora_get_field(ora_ctx* c, int32_t field, double* out);
int ora_set_field(ora_ctx* c, int32_t field, const double* in);
int ora_get_neighbours(ora_ctx* c, int32_t* ptr, int32_t* idx);
int ora_get_csr(ora_ctx* c, int32_t* rowptr, int32_t* col, double* val,
                double* rhs);
int ora_get_scalars(ora_ctx* c, double* umax, double* dt, double* t,
                    int32_t* iters, int32_t* outer, double* relres);
int ora_set_scalars(ora_ctx* c, double umax, double dt, double t);

int ora_projection(ora_ctx* c);
int ora_cellshift(ora_ctx* c);
int ora_resetid(ora_ctx* c);
int ora_neighbour(ora_ctx* c);
int ora_effpor(ora_ctx* c);
int ora_eddyvis(ora_ctx* c);
int ora_int_vel(ora_ctx* c);
int ora_ppe_assemble(ora_ctx* c);
int ora_ppe_solve(ora_ctx* c, int32_t* iters, double* relres);
int ora_ppesolve(ora_ctx* c, int32_t* iters, double* relres);
int ora_comp_vel(ora_ctx* c);
int ora_comp_pos(ora_ctx* c);
int ora_cellshiftalt(ora_ctx* c);
int ora_freesurf(ora_ctx* c);
int ora_opt2_shift(ora_ctx* c);
int ora_timestep(ora_ctx* c);
int ora_step(ora_ctx* c, int32_t nsteps);
int ora_eddyvis_lagged(ora_ctx* c);
int ora_scalart(ora_ctx* c);
int ora_scalarupdate(ora_ctx* c, double stime);
int ora_densityupdate(ora_ctx* c);
int ora_salt_init(ora_ctx* c, double xl, double yl, double line_grad);
int ora_get_substeps(ora_ctx* c, int32_t* solsteps, double* dtsol);
int ora_probe(ora_ctx* c, int32_t np, const double* px, const double* py, double* out);
int ora_set_buffer_zones(ora_ctx* c, int32_t nzones, const isph_buffer_zone* zones,
                         double time_shift);
int ora_remove_buffer(ora_ctx* c, int32_t zone);
int ora_insert_buffer(ora_ctx* c, int32_t zone);

/* Building blocks exposed for the oracle's own pinning tests. */
double ora_kernel_w(const isph_params* prm, double r);
double ora_kernel_w0(const isph_params* prm);
/* gradient of W with respect to particle i; dx = x_i - x_j, dy = y_i - y_j */
void ora_kernel_gradw(const isph_params* prm, double r, double dx, double dy,
                      double* wx, double* wy);
/* invertmat2D (functions.f90:62-96) in place on a[4] */
void ora_invert2(double* a);
/* dot product in the fixed reduction tree the GPU kernels use */
double ora_dot_tree(const double* a, const double* b, int64_t n);
/* Unpreconditioned BiCGSTAB (solver.f90:129-254 recurrences) on a CSR
 * system with the true-residual stop; x holds the guess on entry. */
int ora_bicgstab_csr(int32_t n, const int32_t* rowptr, const int32_t* col,
                     const double* val, const double* b, double* x, double tol,
                     int32_t maxit, int32_t* iters, double* relres);
/* restarted GMRES of the reference's live solve (solver.f90:306-347):
 * restart 0 = min(150, n); x is the warm start and, when converged (or at
 * the end of a restart cycle), the solution; converged = 0 at maxit */
int ora_fgmres_csr(int32_t n, const int32_t* rowptr, const int32_t* col,
                   const double* val, const double* b, double* x, double tol,
                   int32_t maxit, int32_t restart, int32_t* iters, double* relres,
                   int32_t* converged);

#ifdef __cplusplus
}
#endif
#endif
