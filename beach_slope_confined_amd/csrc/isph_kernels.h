// isph_kernels.h -- structures shared by the gfx950 kernels and the host API.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "isph_device.h"

namespace isph {

constexpr int BLOCK = 256;       // threads per block (4 wave64)
constexpr int SCAN_BLOCK = 1024;

// LSD radix sort of the bin keys: 8-bit digits, tiles of RS_IPT keys per
// thread
constexpr int RS_BITS = 8;
constexpr int RS_DIGITS = 1 << RS_BITS;  // == BLOCK: one digit per thread
constexpr int RS_IPT = 8;
constexpr int RS_TILE = BLOCK * RS_IPT;

// device error word bits
constexpr uint32_t ERR_DOMAIN = 1u;
constexpr uint32_t ERR_CAPACITY = 2u;
constexpr uint32_t ERR_NAN = 4u;
constexpr uint32_t ERR_SINGULAR = 8u;
constexpr uint32_t ERR_NOCONV = 16u;
constexpr uint32_t ERR_INDEX16 = 32u;  // a column offset does not fit 16 bits
constexpr uint32_t ERR_SUBSTEPS = 64u; // salinity transport needs > MAX_SUBSTEPS substeps

// cap on the transport substeps of one step (integrator.f90:966-970 has none;
// its ceiling() would overflow a default integer long before this)
constexpr int MAX_SUBSTEPS = 1 << 20;

// solver `done` states
constexpr int DONE_RUNNING = 0;
constexpr int DONE_RECURRENCE = 1;  // recurrence residual passed, true residual pending
constexpr int DONE_CONVERGED = 2;
constexpr int DONE_ZERO_RHS = 3;
constexpr int DONE_FAILED = 4;
constexpr int DONE_UPDATE = 5;      // mixed precision: reliable update requested

struct GridInfo {
    double xmin, ymin, cs;
    int nx, ny;
    long long nbins;  // per * nx * ny
    int order;        // ISPH_ORDER_*
    int per;          // bins per cell: 2 (walls, fluid) or 10 (walls, stayers, 8 arrival sides)
};

// Bin layout.  ISPH_ORDER_CELL: bin = 2 * (cx*ny + cy) + fluid bit.
// ISPH_ORDER_REFERENCE: cells ranked column by column with y descending (the
// reference's resetid numbering, isph.f90:96-109, row i = 1 at the top), bin
// = 10 * rank + class, class 0 walls, 1 fluid that stayed, 2..9 arrivals in
// the gather order of integrator.f90:318-389.
__host__ __device__ inline int cell_rank(const GridInfo& g, int cx, int cy) {
    return g.order ? cx * g.ny + (g.ny - 1 - cy) : cx * g.ny + cy;
}

// cell of a bin key
__host__ __device__ inline void key_cell(const GridInfo& g, int key, int& cx, int& cy) {
    if (g.order) {
        const int r = key / 10;
        cx = r / g.ny;
        cy = g.ny - 1 - (r - cx * g.ny);
    } else {
        const int c = key >> 1;
        cx = c / g.ny;
        cy = c - cx * g.ny;
    }
}

// bins [lo, hi) of the cells ylo..yhi of column cx (contiguous in both layouts)
__host__ __device__ inline void column_bins(const GridInfo& g, int cx, int ylo, int yhi, int& lo,
                                            int& hi) {
    if (g.order) {
        lo = 10 * (cx * g.ny + (g.ny - 1 - yhi));
        hi = 10 * (cx * g.ny + (g.ny - 1 - ylo)) + 10;
    } else {
        lo = 2 * (cx * g.ny + ylo);
        hi = 2 * (cx * g.ny + yhi) + 2;
    }
}

// class of a fluid particle that moved by (dcx, dcy) in {-1,0,1}^2 cells
// (y up): stay 1; from below 2 (tn), above 3 (ts), right 4 (tw), left 5 (te),
// upper left 6 (tse), lower left 7 (tne), upper right 8 (tsw), lower right 9 (tnw)
__host__ __device__ inline int move_class(int dcx, int dcy) {
    // [dcy + 1][dcx + 1]
    const int t[3][3] = {{8, 3, 6}, {4, 1, 5}, {9, 2, 7}};
    return t[dcy + 1][dcx + 1];
}

struct PorousParams {
    double por, alpha, beta, d50, nu;
    int drag_law, visc_law;
};

// one copy of the per-particle state (the arrays that travel with a particle
// through a re-binning)
struct StateView {
    double2 *pos, *vel, *sp, *sv;  // (x,y), (vx,vy), (xs,ys), (vxs,vys)
    double *mass, *rho, *oden, *p, *por, *R, *nut, *tke, *varts, *velm;
    double *con, *ovol;            // salinity and particle volume (scalar.f90)
    int *tid, *flags, *pid;
};

struct SimScalars {
    double umax, numax, dt, t;
    double dtsol;   // transport substep (integrator.f90:964-972)
    int solsteps;   // transport substeps per step
    int pad;
};

struct SolverScalars {
    double bb, nb, prevr, prevr_prev, alpha, wi, betas, rr, rupd, nr, relres;
    double rh2;  // ||rhat||^2 since the last (re)start
    int iters, total, first, done, status, nupd;
    int brk;     // breakdown in this iteration: restart from the true residual
};

// Restart threshold on the shadow inner product rho = (r, rhat): its
// rounding error is about eps ||r|| ||rhat||, so once |rho| <= BRK_COS
// ||r|| ||rhat|| it has lost its digits and the recurrence restarts.
constexpr double BRK_COS = 1e-15;

// ---------------------------------------------------------------------------
// fixed-tree reductions: wave64 butterfly (xor 1,2,..,32), then the four
// waves of a 256-thread block as (w0+w1)+(w2+w3)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) v = v + __shfl_xor(v, s, 64);
    return v;
}

// block sum in the fixed tree; the result is valid in every thread
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double red[4];
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return (red[0] + red[1]) + (red[2] + red[3]);
}

__device__ __forceinline__ double block_max(double v) {
    __shared__ double red[4];
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) v = fmax(v, __shfl_xor(v, s, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// Logical block of a solver launch: physical blocks are dealt round-robin
// to the 8 XCDs, so XCD x gets the contiguous logical range
// [x*per, (x+1)*per) and its L2 holds the vector entries its rows gather.
__device__ __forceinline__ int xcd_block(int bid, int nblocks) {
    int per = nblocks >> 3;  // nblocks is a multiple of 8
    return (bid & 7) * per + (bid >> 3);
}

// Row tile of physical block bid for a grid of any size nb: the hardware
// deals workgroups to the 8 XCDs round-robin (bid & 7), so XCD x runs the
// blocks x, x + 8, ...; this gives XCD x the contiguous logical range
// [x q + min(x, r), ...) of q + (x < r) blocks (q = nb / 8, r = nb % 8).
// A particle's neighbours lie within a few cell columns of it in the sorted
// order, so with contiguous ranges an XCD's L2 holds the neighbour data its
// rows gather instead of every XCD fetching the same rows.
__device__ __forceinline__ int xcd_tile(int bid, int nb) {
    const int q = nb >> 3, r = nb & 7, x = bid & 7;
    return x * q + min(x, r) + (bid >> 3);
}

// The same partition with each XCD walking its range from the end.  The two
// SpMV kernels of a BiCGSTAB iteration read the same matrix (~250 MB at 1M
// rows, about the size of the 256 MB Infinity Cache); running the second
// one in the opposite order starts it on the rows the first one read last,
// which are still cached.  Partials stay indexed by logical block, so the
// reductions are unchanged.
__device__ __forceinline__ int xcd_block_rev(int bid, int nblocks) {
    int per = nblocks >> 3;
    return (bid & 7) * per + (per - 1 - (bid >> 3));
}

template <class T> struct Pair2;
template <> struct Pair2<double> { using type = double2; };
template <> struct Pair2<float> { using type = float2; };

// Operands of one BiCGSTAB flavour: TV matrix values, TX vectors, TI column
// index (int: absolute, short: offset from the row).  The two gathered
// operands are stored as pairs so that a neighbour costs one load:
// rq = (r, q) for the p update of step 1, rv = (r, v) for s = r - alpha v
// of step 2.
template <class TV, class TX, class TI>
struct KrylovArgs {
    using P2 = typename Pair2<TX>::type;
    int n;
    size_t ld;
    const TI* nbr;
    const int* cnt;
    const int* dpos;
    const TV* val;
    const double* rsum;  // FP64 row sums (conservation-form SpMV of MIXED_V64)
    P2 *rq, *rv;
    TX *rhat, *p, *si, *mm2, *xs;
    double* part;  // 2 * nbp block partials
};

constexpr int FIN_BLOCK = 1024;  // threads of the one-block reduction kernel

// cell (column-major, 1-based interior) of a position; out-of-grid
// positions are clamped to the border cells and flagged
__device__ inline int cell_of(const GridInfo& g, double x, double y, bool& oob) {
    double fx = floor((x - g.xmin) / g.cs);
    double fy = floor((y - g.ymin) / g.cs);
    int cx, cy;
    if (!(fx >= 0.0)) { cx = 1; oob = true; }
    else if (fx > (double)(g.nx - 3)) { cx = g.nx - 2; oob = true; }
    else cx = (int)fx + 1;
    if (!(fy >= 0.0)) { cy = 1; oob = true; }
    else if (fy > (double)(g.ny - 3)) { cy = g.ny - 2; oob = true; }
    else cy = (int)fy + 1;
    return cx * g.ny + cy;
}

// block sum over FIN_BLOCK threads: wave butterflies, then the 16 wave sums
// as a balanced pairwise tree; valid in every thread
__device__ __forceinline__ double block_sum_fin(double v) {
    __shared__ double red[FIN_BLOCK / 64];
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double a[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = red[2 * k] + red[2 * k + 1];
#pragma unroll
    for (int k = 0; k < 4; ++k) a[k] = a[2 * k] + a[2 * k + 1];
    return ((a[0] + a[1]) + (a[2] + a[3]));
}

// k_finish's reduction of one column of nbp block partials in a block of
// FIN_BLOCK threads: thread t sums partials t, t + FIN_BLOCK, ... in order,
// then block_sum_fin; valid in every thread
__device__ __forceinline__ double fin_sum(const double* __restrict__ part, int nbp) {
    double p[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int k = threadIdx.x + u * FIN_BLOCK;
        p[u] = k < nbp ? part[k] : 0.0;
    }
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u)
        if (threadIdx.x + u * FIN_BLOCK < nbp) s = s + p[u];
    for (int k = threadIdx.x + 8 * FIN_BLOCK; k < nbp; k += FIN_BLOCK) s = s + part[k];
    return block_sum_fin(s);
}

// Operands of an FP64 residual b - A64 x.
struct ResidArgs {
    int n;
    size_t ld;
    const int* nbr;
    const short* nbr16;  // offsets j - i of the same lists when every one fits, else null
    const int* cnt;
    const int* dpos;
    const double* val;
    const double* b;
    const double* x;
    double* r;
    double* part;
};

struct SolveCtl {
    double tol;    // ||b - A x|| <= tol ||b||
    double delta;  // reliable update when the residual fell by delta
    int maxit;     // iteration cap (total)
    int maxupd;    // reliable update cap
    int mixed;     // 1: reliable-update logic, 0: plain FP64 BiCGSTAB
};

// reductions finished by one block (k_finish)
constexpr int FIN_INIT = 0;     // (b,b), (r,r) of the FP64 start
constexpr int FIN_ALPHA = 1;    // (v,rhat)
constexpr int FIN_OMEGA = 2;    // (t,s), (t,t)
constexpr int FIN_BETA = 3;     // (r,rhat), (r,r) + stop tests
constexpr int FIN_TRUE = 4;     // (r,r) of the FP64 true residual
constexpr int FIN_RESID = 5;    // (b,b), (r,r) of the mixed outer residual
constexpr int FIN_START = 6;    // (rhat,r) of the mixed start
constexpr int FIN_REPLACE = 7;  // (r,rhat) after a reliable update

__global__ void k_bin_keys(int n, GridInfo g, const double2* pos, const int* tid, const int* flags,
                           int sink, double xlcut, double ytcut, const int* prev_key, int* key,
                           int* bincnt, uint32_t* err);
__global__ void k_scan_blocks(int n, int* data, int* bsum);
__global__ void k_scan_sums(int nb, int* bsum);
__global__ void k_scan_add(int n, int* data, const int* bsum);
__global__ void k_rs_hist(int n, const int* keys, int shift, int ntiles, int* hist);
__global__ void k_rs_scatter(int n, const int* kin, const int* vin, int shift, int ntiles,
                             const int* hist, int* kout, int* vout);
__global__ void k_gather_state(int n, const int* perm, StateView src, StateView dst, const int* key,
                               int* key_sorted);
__global__ void k_projection(int n, StateView s, const SimScalars* ss);
__global__ void k_neighbour(int n, GridInfo g, KernelConsts kc, double thr, int maxnbr, size_t ld,
                            StateView s, const int* key, const int* bstart, int* nbr, int* cnt,
                            double* L, double* divr, uint32_t* err);
__global__ void k_effpor(int n, GridInfo g, KernelConsts kc, PorousParams pp, StateView s,
                         const int* key, const double2* spos, const double* sm, const double* sr,
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
                           double* val, float* val32, double* rsum, int* dpos, double* rhs,
                           const SimScalars* ss, uint32_t* err);
template <class TV, class TX, class TI, bool FIRST>
__global__ void k_krylov_1t(KrylovArgs<TV, TX, TI> a, const SolverScalars* sc);
template <class TV, class TX, class TI>
__global__ void k_krylov_2(KrylovArgs<TV, TX, TI> a, const SolverScalars* sc);
template <class TV, class TX, class TI>
__global__ void k_krylov_3(KrylovArgs<TV, TX, TI> a, const SolverScalars* sc);
__global__ void k_finish(int kind, int nbp, const double* part, SolverScalars* sc, SolveCtl ctl);
__global__ void k_resid(ResidArgs a, int mode, double2* rq, double* rhat);
template <class TX>
__global__ void k_mx_start(int n, const double* r64, typename Pair2<TX>::type* rq, TX* rhat,
                           TX* xs, double* part, const SolverScalars* sc);
template <class TX>
__global__ void k_mx_replace(int n, const double* r64, typename Pair2<TX>::type* rq,
                             const TX* rhat, double* part, const SolverScalars* sc);
template <class TX>
__global__ void k_mx_flush(int n, double* x64, TX* xs);
__global__ void k_to_short(int n, size_t ld, const int* nbr, const int* cnt, short* out,
                           uint32_t* err);
__global__ void k_mg_lower(int n, size_t ld, const int* cnt, const int* dpos, const double* val,
                           float* val32, double* rsum);
__global__ void k_comp_vel(int n, KernelConsts kc, size_t ld, StateView s, const int* nbr,
                           const int* cnt, const double* L, const SimScalars* ss, uint32_t* err);
__global__ void k_comp_pos(int n, StateView s, const SimScalars* ss, double xrel_lo,
                           double xrel_hi);
__global__ void k_freesurf(int n, double thr, size_t ld, StateView s, const int* nbr,
                           const int* cnt, const double* divr);
__global__ void k_shift(int n, KernelConsts kc, double csh, double cap, size_t ld, StateView s,
                        const int* nbr, const int* cnt, const double* L);
__global__ void k_shift_apply(int n, StateView s);
__global__ void k_maxes(int n, const double* velm, const double* nut, double* part);
__global__ void k_timestep(int nb, const double* part, SimScalars* ss, double dl, double s1,
                           double mu, double rho, int adaptive, double tschmidt, int salinity,
                           uint32_t* err);
__global__ void k_advance_time(SimScalars* ss);
__global__ void k_copy(const double2* a, double2* b, size_t n);

// salinity transport and variable density (isph_scalar.hip)
struct SaltConsts {
    double lam, dl, tschmidt;
    double rhomax, rhomin, xbuf;
};
__global__ void k_scalar_coef(int n, SaltConsts sc, StateView s, const SimScalars* ss,
                              double4* coef);
__global__ void k_scalart(int n, KernelConsts kh, SaltConsts sc, size_t ld, StateView s,
                          const double2* pos_nb, const int* nbr, const int* cnt,
                          const double4* coef, double* cdiff);
__global__ void k_scalarupdate(int n, SaltConsts sc, StateView s, const double* cdiff,
                               double stime, const SimScalars* ss);
__global__ void k_densityupdate(int n, SaltConsts sc, StateView s);
__global__ void k_salt_init(int n, StateView s, double xl, double yl, double line_grad,
                            double rel_den);

}  // namespace isph
