// isph_kernels.h -- structures shared by the gfx950 kernels and the host API.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "isph_device.h"

namespace isph {

constexpr int BLOCK = 256;       // threads per block (4 wave64)
constexpr int SCAN_BLOCK = 1024;

// device error word bits
constexpr uint32_t ERR_DOMAIN = 1u;
constexpr uint32_t ERR_CAPACITY = 2u;
constexpr uint32_t ERR_NAN = 4u;
constexpr uint32_t ERR_SINGULAR = 8u;
constexpr uint32_t ERR_NOCONV = 16u;

// solver `done` states
constexpr int DONE_RUNNING = 0;
constexpr int DONE_RECURRENCE = 1;  // recurrence residual passed, true residual pending
constexpr int DONE_CONVERGED = 2;
constexpr int DONE_ZERO_RHS = 3;
constexpr int DONE_FAILED = 4;
constexpr int DONE_INNER = 5;       // mixed precision inner solve finished

struct GridInfo {
    double xmin, ymin, cs;
    int nx, ny;
    long long nbins;  // 2 * nx * ny
};

struct PorousParams {
    double por, alpha, beta, d50, nu;
    int drag_law, visc_law;
};

// one copy of the per-particle state (the arrays that travel with a particle
// through a re-binning)
struct StateView {
    double2 *pos, *vel, *sp, *sv;  // (x,y), (vx,vy), (xs,ys), (vxs,vys)
    double *mass, *rho, *oden, *p, *por, *R, *nut, *tke, *varts, *velm;
    int *tid, *flags, *pid;
};

struct SimScalars {
    double umax, numax, dt, t;
};

struct SolverScalars {
    double bb, nb, prevr, alpha, wi, betas, rr, rr0, nr, relres;
    int iters, total, first, done, status, pad;
};

struct SolveArgs {
    int n;
    size_t ld;
    const int* nbr;
    const int* cnt;
    const int* dpos;
    const double* val;
    const double* b;
    double* x;
    double* rs;
    double* rhat;
    double* part;
    uint32_t* counters;
    double tol;
    int maxit;
};

template <class T>
struct SolveArgsT {
    int n;
    size_t ld;
    const int* nbr;
    const int* cnt;
    const int* dpos;
    const T* val;
    T* rs;
    T* rhat;
    T* si;
    T* mm2;
    T* xs;  // the solution being built
    double* part;
    uint32_t* counters;
    double tol;
    int maxit;
    int inner_maxit;
};

__global__ void k_bin_keys(int n, GridInfo g, const double2* pos, const int* tid, const int* flags,
                           int sink, double xlcut, double ytcut, int* key, int* bincnt, int* rank,
                           uint32_t* err);
__global__ void k_scan_blocks(int n, int* data, int* bsum);
__global__ void k_scan_sums(int nb, int* bsum);
__global__ void k_scan_add(int n, int* data, const int* bsum);
__global__ void k_bin_scatter(int n, const int* key, const int* rank, const int* bstart, int* perm);
__global__ void k_bin_fixup(int nbins, const int* bstart, int* perm);
__global__ void k_gather_state(int n, const int* perm, StateView src, StateView dst, const int* key,
                               int* key_sorted);
__global__ void k_projection(int n, StateView s, const SimScalars* ss);
__global__ void k_neighbour(int n, GridInfo g, KernelConsts kc, double thr, int maxnbr, size_t ld,
                            StateView s, const int* key, const int* bstart, int* nbr, int* cnt,
                            double* L, double* divr, uint32_t* err);
__global__ void k_effpor(int n, GridInfo g, KernelConsts kc, PorousParams pp, StateView s,
                         const double2* spos, const double* sm, const double* sr,
                         const int* sstart, uint32_t* err);
__global__ void k_eddyvis(int n, KernelConsts kc, double dl, size_t ld, StateView s, const int* nbr,
                          const int* cnt, const double* L);
__global__ void k_varts(int n, KernelConsts kc, double dl, size_t ld, StateView s, const int* nbr,
                        const int* cnt, const double* L, const SimScalars* ss);
__global__ void k_vstar(int n, KernelConsts kc, PorousParams pp, double lam, double g, size_t ld,
                        StateView s, const int* nbr, const int* cnt, const double* L,
                        const SimScalars* ss);
__global__ void k_assemble(int n, KernelConsts kc, double lam, double lampd, size_t ld,
                           StateView s, const int* nbr, const int* cnt, const double* L,
                           double* val, float* val32, int* dpos, double* rhs,
                           const SimScalars* ss, uint32_t* err);
__global__ void k_bicg_init(SolveArgs a, SolverScalars* sc);
template <class T>
__global__ void k_bicg_1(SolveArgsT<T> a, SolverScalars* sc, const T* p_old, const T* mm1_old,
                         T* p_new, T* mm1_new);
template <class T>
__global__ void k_bicg_2(SolveArgsT<T> a, SolverScalars* sc, const T* mm1);
template <class T>
__global__ void k_bicg_3(SolveArgsT<T> a, SolverScalars* sc, const T* p, int inner);
__global__ void k_bicg_true(SolveArgs a, SolverScalars* sc);
__global__ void k_mixed_resid(SolveArgs a, SolverScalars* sc, double* r64);
__global__ void k_mixed_start(SolveArgsT<float> f, SolverScalars* sc, const double* r64);
__global__ void k_mixed_update(int n, double* x, const float* d, const SolverScalars* sc);
__global__ void k_comp_vel(int n, KernelConsts kc, size_t ld, StateView s, const int* nbr,
                           const int* cnt, const double* L, const SimScalars* ss, uint32_t* err);
__global__ void k_comp_pos(int n, StateView s, const SimScalars* ss);
__global__ void k_freesurf(int n, double thr, size_t ld, StateView s, const int* nbr,
                           const int* cnt, const double* divr);
__global__ void k_shift(int n, KernelConsts kc, double csh, double cap, size_t ld, StateView s,
                        const int* nbr, const int* cnt, const double* L);
__global__ void k_shift_apply(int n, StateView s);
__global__ void k_maxes(int n, const double* velm, const double* nut, unsigned long long* out);
__global__ void k_timestep(const unsigned long long* mx, SimScalars* ss, double dl, double s1,
                           double mu, double rho, int adaptive);
__global__ void k_advance_time(SimScalars* ss);
__global__ void k_copy(const double2* a, double2* b, size_t n);

}  // namespace isph
