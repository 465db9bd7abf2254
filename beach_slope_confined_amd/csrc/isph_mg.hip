// isph_mg.hip -- multigrid-preconditioned PPE solve: mixed precision
// (ISPH_PREC_MIXED_MG, the V-cycle and the outer products in FP32 values)
// and the same algorithm with every value FP64 (ISPH_PREC_FP64_MG), the
// baseline that isolates the effect of precision.
//
// BiCGSTAB with right preconditioning by one V-cycle of an aggregation
// multigrid built on the particle cell grid:
//   level 0  the PPE rows (particles), FP32 values of the assembled matrix;
//   level 1  the cells: aggregate = the particles of a cell, piecewise
//            constant prolongation, Galerkin operator A1 = P^T A P, which is
//            a 9-point stencil on the cell grid because every neighbour of a
//            particle lies in the 3x3 cells around its own;
//   level l+1  2x2 blocks of level-l cells, again Galerkin and 9-point;
//   block     the levels from the first with at most MG_CMAX cells down to
//            a single cell: V-cycles inside one block (k_mg_block).
// Smoothing: the particle level takes one damped-Jacobi sweep after the
// coarse correction and none before it (x = 0, the residual is b), so its
// iterate is constant on cells and the sweep reads per-row sums by cell
// offset instead of the neighbour lists (k_mg_fine_up); the grid
// levels one sweep before (from x = 0) and one after; the block levels
// MG_BLOCK_NU sweeps each way, with max(2, ceil(K / 3)) V-cycles for K block
// levels.  Every coarse correction is scaled by MG_OVER.  The V-cycle is a fixed linear
// operator, so the preconditioned recurrences are the ordinary ones:
//   p = r + beta (p - omega v),  p^ = M p,  v = A p^,  alpha = rho / (rhat, v)
//   s = r - alpha v,             s^ = M s,  t = A s^,  omega = (t, s)/(t, t)
//   x += alpha p^ + omega s^,    r = s - omega t
// The outer vectors and all dot products are FP64; A p^ and A s^ use the
// FP32 values in the conservation form with FP64 row sums (as the
// MIXED_V64 solve); x accumulates in a segment flushed into the FP64
// solution by reliable updates against the FP64 true residual, which is
// also the stop test (||b - A64 x|| <= tol ||b||).  The scalars follow the
// same protocol as the unpreconditioned solves and are finished by
// k_finish.  Only the preconditioner and the matrix values are FP32; in
// ISPH_PREC_FP64_MG they are FP64 (kernels templated on the value type T).
//
// The hierarchy is rebuilt for every solve (the matrix changes every step):
// each row's entries summed by neighbour cell offset (k_mg_rows9), the cell
// operator from those sums (k_mg_galerkin1), then one launch per coarser
// level (k_mg_galerkin), each sum in a fixed order, so the set-up is
// deterministic.  The block levels' operators and the exact level's
// inverse are formed on a side stream (mg_setup, mg_join).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "isph_host.h"
#include "isph_kernels.h"

namespace isph {

constexpr int MG_CMAX = 1024;  // top block level: at most this many cells (one thread each)
constexpr int MG_CLOW = 768;   // the block's lower levels together: at most this many cells
constexpr int MG_CLEV = 16;    // at most this many block levels
constexpr int MG_MAXLEV = 32;  // grid levels in all
constexpr int MG_BLOCK_NU = 3;      // block levels: sweeps before and after the correction
// block V-cycles per coarse solve: a deeper block hierarchy compounds the
// under-correction of its aggregates, so it gets more cycles
inline int mg_block_cycles(int K) { return std::max(2, (K + 2) / 3); }
// factor on every coarse-grid correction: the piecewise-constant
// aggregates under-correct, and 1.5 cuts the BiCGSTAB iterations of the
// 1024x1024 tank from 94 to 30 (1.3: 34; 1.7 in the block: 50)
constexpr float MG_OVER = 1.5f;
// Jacobi damping 0.8, rounded to the V-cycle's value type T
template <class T>
__device__ __forceinline__ T mg_omega() { return (T)0.8; }
template <>
__device__ __forceinline__ float mg_omega<float>() { return 0.8f; }

__device__ __forceinline__ int mg_col(int e, int) { return e; }
__device__ __forceinline__ int mg_col(short e, int i) { return i + (int)e; }

// cell of every particle on the level-1 grid (cx * ny + cy of its bin)
__global__ void k_mg_cells(int n, GridInfo g, const int* __restrict__ key, int* __restrict__ cell) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    int cx, cy;
    key_cell(g, key[i], cx, cy);
    cell[i] = cx * g.ny + cy;
}

// A1 = P^T A P in two passes.  k_mg_rows9: every row's entries summed by the
// neighbour's cell offset (9 sums, the unit diagonal first), one thread per
// row; an entry outside the 3x3 cells raises ERR_DOMAIN (cannot happen for
// lists built from the cell stencil).  k_mg_galerkin1: per cell, its rows'
// sums in row order.  A thread per cell walking its rows' entries serially
// (about 140 at 2h cells) left the GPU nearly idle: 606 blocks at 1M.
template <class TI>
__global__ void __launch_bounds__(BLOCK) k_mg_rows9(int n, int ny, size_t ld,
                                                    const TI* __restrict__ nbr,
                                                    const int* __restrict__ cnt,
                                                    const double* __restrict__ val,
                                                    const int* __restrict__ cell,
                                                    double* __restrict__ part9,
                                                    float* __restrict__ p9f, uint32_t* err) {
    const int i = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (i >= n) return;
    const int ci = cell[i];
    const int cx = ci / ny, cy = ci - cx * ny;
    double acc[9] = {0, 0, 0, 0, 1.0, 0, 0, 0, 0};
    bool bad = false;
    auto add = [&](int cj, double v) {
        const int jx = cj / ny, jy = cj - jx * ny;
        const int dx = jx - cx, dy = jy - cy;
        if (dx < -1 || dx > 1 || dy < -1 || dy > 1) {
            bad = true;
            return;
        }
        const int dd = (dx + 1) * 3 + (dy + 1);
#pragma unroll
        for (int q = 0; q < 9; ++q)  // constant indices: acc stays in registers
            if (q == dd) acc[q] += v;
    };
    const int c = cnt[i];
    int k = 0;
    // the loads of four entries are in flight together; the sums keep the
    // entry order
    for (; k + 4 <= c; k += 4) {
        const size_t o = (size_t)k * ld + i;
        const int j0 = mg_col(nbr[o], i), j1 = mg_col(nbr[o + ld], i);
        const int j2 = mg_col(nbr[o + 2 * ld], i), j3 = mg_col(nbr[o + 3 * ld], i);
        const double v0 = val[o], v1 = val[o + ld], v2 = val[o + 2 * ld], v3 = val[o + 3 * ld];
        const int c0 = cell[j0], c1 = cell[j1], c2 = cell[j2], c3 = cell[j3];
        add(c0, v0);
        add(c1, v1);
        add(c2, v2);
        add(c3, v3);
    }
    for (; k < c; ++k) {
        const size_t o = (size_t)k * ld + i;
        add(cell[mg_col(nbr[o], i)], val[o]);
    }
    if (bad) atomicOr(err, ERR_DOMAIN);
#pragma unroll
    for (int d = 0; d < 9; ++d) part9[(size_t)d * ld + i] = acc[d];
    if (p9f)
#pragma unroll
        for (int d = 0; d < 9; ++d) p9f[(size_t)d * ld + i] = (float)acc[d];
}

template <class T>
__global__ void k_mg_galerkin1(GridInfo g, const int* __restrict__ bstart, size_t ld,
                               const double* __restrict__ part9, T* __restrict__ A1) {
    const int nc = g.nx * g.ny;
    int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= nc) return;
    const int cx = c / g.ny, cy = c - cx * g.ny;
    int lo, hi;
    column_bins(g, cx, cy, cy, lo, hi);
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int i1 = bstart[hi];
    int i = bstart[lo];
    // two rows' sums loaded before they are added, in row order per plane
    for (; i + 2 <= i1; i += 2) {
        double r0[9], r1[9];
#pragma unroll
        for (int d = 0; d < 9; ++d) {
            r0[d] = part9[(size_t)d * ld + i];
            r1[d] = part9[(size_t)d * ld + i + 1];
        }
#pragma unroll
        for (int d = 0; d < 9; ++d) acc[d] = (acc[d] + r0[d]) + r1[d];
    }
    if (i < i1)
#pragma unroll
        for (int d = 0; d < 9; ++d) acc[d] += part9[(size_t)d * ld + i];
#pragma unroll
    for (int d = 0; d < 9; ++d) A1[(size_t)d * nc + c] = (T)acc[d];
}

// A_{l+1} from A_l: coarse cell C = (X, Y) owns the fine cells (2X+a, 2Y+b)
template <class T>
__global__ void k_mg_galerkin(int nx, int ny, const T* __restrict__ A, int nx2, int ny2,
                              T* __restrict__ A2) {
    const int nc = nx * ny, nc2 = nx2 * ny2;
    int C = blockIdx.x * BLOCK + threadIdx.x;
    if (C >= nc2) return;
    const int X = C / ny2, Y = C - X * ny2;
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            const int cx = 2 * X + a, cy = 2 * Y + b;
            if (cx >= nx || cy >= ny) continue;
            const int c = cx * ny + cy;
            for (int d = 0; d < 9; ++d) {
                const int ex = cx + d / 3 - 1, ey = cy + d % 3 - 1;
                if (ex < 0 || ex >= nx || ey < 0 || ey >= ny) continue;
                const int D = ((ex >> 1) - X + 1) * 3 + ((ey >> 1) - Y + 1);
                acc[D] += (double)A[(size_t)d * nc + c];
            }
        }
    for (int d = 0; d < 9; ++d) A2[(size_t)d * nc2 + C] = (T)acc[d];
}

// ---------------------------------------------------------------------------
// V-cycle kernels (value type T: float for MIXED_MG, double for FP64_MG)
// ---------------------------------------------------------------------------
// Level 0 down.  The right-hand side b is formed from the outer pairs:
// MODE 0: b = p = r + beta q (r on a restart), from rq = (r, q);
// MODE 1: b = s = r - alpha v, from rv = (r, v).  b goes to bout (FP64, the
// outer recurrences need it).  Without pre-smoothing the particle iterate is
// 0 and its residual is b itself, r0 = (T) b.
template <class T, int MODE, bool FIRST>
__global__ void k_mg_bpoint(int n, const double2* __restrict__ pairs, const SolverScalars* sc,
                            double* __restrict__ bout, T* __restrict__ r0) {
    // the flag, the coefficient and the pair are requested together
    const int done = sc->done;
    const double coef = MODE == 0 ? sc->betas : sc->alpha;
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const double2 u = pairs[i];
    if (done) return;
    const double bi = MODE == 0 ? (FIRST ? u.x : u.x + coef * u.y) : u.x - coef * u.y;
    bout[i] = bi;
    r0[i] = (T)bi;
}
// restriction to the cells: b1[c] = sum of r0 over the cell's particles
template <class T>
__global__ void k_mg_restrict1(GridInfo g, const int* __restrict__ bstart,
                               const T* __restrict__ r0, T* __restrict__ b1,
                               const SolverScalars* sc) {
    const int done = sc->done;  // tested once the cell's bin range is requested
    const int nc = g.nx * g.ny;
    int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= nc) return;
    const int cx = c / g.ny, cy = c - cx * g.ny;
    int lo, hi;
    column_bins(g, cx, cy, cy, lo, hi);
    const int i0 = bstart[lo], i1 = bstart[hi];
    if (done) return;
    T s = (T)0;
    for (int i = i0; i < i1; ++i) s += r0[i];
    b1[c] = s;
}

// neighbour index clamped into [0, n): the stencil entries that point
// outside the grid are zero on every level (the Galerkin sums never reach
// them), so a clamped neighbour adds 0 and the sweeps need no branches
__device__ __forceinline__ int mg_clamp(int e, int n) { return min(max(e, 0), n - 1); }

// c / ny for a block cell, c < MG_CMAX and ny <= MG_CMAX, from rny = RN(1/ny)
// in FP32: with c = k ny + r, (c + 1/2) / ny lies at least 1/(2 ny) >=
// 2^-11 inside (k, k + 1), and the two FP32 roundings move it by at most
// 2^-22 (c + 1/2) / ny < 2^-12, so the truncation is k.
__device__ __forceinline__ int mg_div(int c, float rny) {
    return (int)(((float)c + 0.5f) * rny);
}

template <class T>
__device__ __forceinline__ T mg_jac0(const T* __restrict__ A, int nc, int c,
                                     const T* __restrict__ b) {
    const T dg = A[(size_t)4 * nc + c];
    return dg != (T)0 ? mg_omega<T>() * b[c] / dg : (T)0;
}

// grid level down and restriction: four consecutive lanes take the four
// children k of one cell C = (X, Y) of the next level, child k at
// (2X + k/2, 2Y + k%2); each smooths its child from zero (x = omega D^-1 b)
// and forms its residual r = b - A x; b2[C] = (r0 + r1) + (r2 + r3), the
// same sum in every lane of the quad (missing children add 0)
template <class T>
__global__ void k_mg_down(int nx, int ny, const T* __restrict__ A,
                          const T* __restrict__ b, T* __restrict__ x, int nx2, int ny2,
                          T* __restrict__ b2, const SolverScalars* sc) {
    const int done = sc->done;
    const int nc = nx * ny;
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    const int C = t >> 2, k = t & 3;
    T r = (T)0, xc = (T)0;
    int cw = -1;
    if (C < nx2 * ny2) {
        const int X = C / ny2, Y = C - X * ny2;
        const int cx = 2 * X + (k >> 1), cy = 2 * Y + (k & 1);
        if (cx < nx && cy < ny) {
            const int c = cx * ny + cy;
            T acc = (T)0;
#pragma unroll
            for (int d = 0; d < 9; ++d) {
                const int ex = mg_clamp(cx + d / 3 - 1, nx), ey = mg_clamp(cy + d % 3 - 1, ny);
                acc += A[(size_t)d * nc + c] * mg_jac0(A, nc, ex * ny + ey, b);
            }
            xc = mg_jac0(A, nc, c, b);
            r = b[c] - acc;
            cw = c;
        }
    }
    // the flag was loaded with the first operands; uniform over the grid
    if (done) return;
    if (cw >= 0) x[cw] = xc;
    const T s = r + __shfl_xor(r, 1);
    const T q = s + __shfl_xor(s, 2);
    if (k == 0 && C < nx2 * ny2) b2[C] = q;
}

// The coarsest levels, from the first with at most MG_CMAX cells down to a
// single cell, are solved inside one block of MG_CMAX threads by V-cycles:
// thread c owns cell c of every block level.  The top block level keeps its
// stencil in registers; the lower levels' stencils, all iterates and
// right-hand sides sit in LDS, so a sweep costs one barrier.
struct MGBlock {
    int K;                        // block levels (1..MG_CLEV); 0 is the top, K-1 one cell
    int nu;                       // damped-Jacobi sweeps before and after the correction
    int cycles;                   // V-cycles per coarse solve
    float over;                   // factor on the coarse correction
    int nx[MG_CLEV], ny[MG_CLEV];
    int off[MG_CLEV];             // offset of level l >= 1 in the lower-level arrays
    float rny[MG_CLEV];           // 1 / ny in FP32 (mg_div)
    const void* A[MG_CLEV];       // stencils, of the V-cycle's value type
    int D;                        // level solved exactly by its inverse (0: none)
};

// The exact coarse solve.  The first block level below the top with at most
// MG_DMAX cells (and more than one) is solved by its inverse, formed once per
// hierarchy by k_mg_coarse_inv: Gauss-Jordan with partial pivoting in FP64 on
// the dense matrix of the level's 9-point stencil.  A cell with a zero
// diagonal (an empty aggregate, which no other cell couples to) gets an
// identity row, so its value stays 0 as in the sweeps.  If a pivot falls
// below 1e-12 of the largest entry, `ok` is 0 and the V-cycles recurse to
// the single cell instead.  The inverse is stored transposed, Ainv[j n + c].
constexpr int MG_DMAX = 64;

template <class T>
__global__ void __launch_bounds__(MG_CMAX) k_mg_coarse_inv(int nx, int ny, const T* __restrict__ A,
                                                           T* __restrict__ Ainv,
                                                           int* __restrict__ ok) {
    constexpr int W = 2 * MG_DMAX;  // [M | I]
    __shared__ double M[MG_DMAX][W + 1];
    __shared__ double fac[MG_DMAX];
    __shared__ double amax;
    __shared__ int piv, good;
    const int n = nx * ny, t = threadIdx.x;
    for (int e = t; e < MG_DMAX * W; e += MG_CMAX) M[e / W][e % W] = 0.0;
    __syncthreads();
    if (t < n) {
        const int rx = t / ny, ry = t - rx * ny;
        for (int d = 0; d < 9; ++d) {
            const int ex = rx + d / 3 - 1, ey = ry + d % 3 - 1;
            if (ex >= 0 && ex < nx && ey >= 0 && ey < ny)
                M[t][ex * ny + ey] = (double)A[(size_t)d * n + t];
        }
        if (M[t][t] == 0.0) M[t][t] = 1.0;  // empty aggregate: x = b = 0
        M[t][MG_DMAX + t] = 1.0;
    }
    __syncthreads();
    if (t < 64) {  // largest |entry|, the scale of the pivot test
        double v = 0.0;
        if (t < n)
            for (int j = 0; j < n; ++j) v = fmax(v, fabs(M[t][j]));
        for (int s = 1; s < 64; s <<= 1) v = fmax(v, __shfl_xor(v, s, 64));
        if (t == 0) {
            amax = v;
            good = 1;
        }
    }
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        if (t < 64) {  // pivot: largest |M[r][k]| over r >= k, lowest r on ties
            double v = (t >= k && t < n) ? fabs(M[t][k]) : -1.0;
            int r = t;
            for (int s = 1; s < 64; s <<= 1) {
                const double v2 = __shfl_xor(v, s, 64);
                const int r2 = __shfl_xor(r, s, 64);
                if (v2 > v || (v2 == v && r2 < r)) {
                    v = v2;
                    r = r2;
                }
            }
            if (t == 0) {
                piv = r;
                if (!(v > 1e-12 * amax)) good = 0;
            }
        }
        __syncthreads();
        if (!good) break;
        const int p = piv;
        if (p != k && t < W) {
            const double u = M[k][t];
            M[k][t] = M[p][t];
            M[p][t] = u;
        }
        __syncthreads();
        const double pv = M[k][k];
        if (t < n) fac[t] = M[t][k];
        __syncthreads();
        if (t < W) M[k][t] = M[k][t] / pv;
        __syncthreads();
        for (int e = t; e < n * W; e += MG_CMAX) {
            const int r = e / W, j = e - r * W;
            if (r != k) M[r][j] = M[r][j] - fac[r] * M[k][j];
        }
        __syncthreads();
    }
    for (int e = t; e < n * n; e += MG_CMAX) {
        const int j = e / n, c = e - j * n;
        Ainv[e] = (T)M[c][MG_DMAX + j];
    }
    if (t == 0) *ok = good;
}

template <class T>
__global__ void __launch_bounds__(MG_CMAX) k_mg_block(MGBlock B, const T* __restrict__ b,
                                                      T* __restrict__ x,
                                                      const T* __restrict__ Ainv,
                                                      const int* __restrict__ Aok,
                                                      const SolverScalars* sc) {
    const int done = sc->done;  // tested once the prologue's loads are requested
    __shared__ T xt[2][MG_CMAX];   // top level: iterate (two parities), residual
    __shared__ T rt[MG_CMAX];
    __shared__ T As[9 * MG_CLOW];  // lower levels: stencil d of cell c at 9 off + d n + c
    __shared__ T xl[2][MG_CLOW];
    __shared__ T bl[MG_CLOW];
    __shared__ T Ai[MG_DMAX * MG_DMAX];  // the exact level's inverse, transposed
    const T omega = mg_omega<T>();
    const int c = threadIdx.x;
    const int K = B.K;
    const int ny0 = B.ny[0], n0 = B.nx[0] * ny0;
    const bool own = c < n0;
    const int cx0 = own ? c / ny0 : 0, cy0 = own ? c - cx0 * ny0 : 0;
    T a[9], w = (T)0, bc = (T)0;
    int nb[9], pa[9];   // neighbour cells and their parents on level 1
    int pc = 0;
    bool live = false;
    if (own) {
        const int ny1 = K > 1 ? B.ny[1] : 1;
        for (int d = 0; d < 9; ++d) {
            const int ex = cx0 + d / 3 - 1, ey = cy0 + d % 3 - 1;
            const bool in = ex >= 0 && ex < B.nx[0] && ey >= 0 && ey < ny0;
            a[d] = in ? static_cast<const T*>(B.A[0])[(size_t)d * n0 + c] : (T)0;
            nb[d] = in ? ex * ny0 + ey : c;
            pa[d] = in ? (ex >> 1) * ny1 + (ey >> 1) : 0;
        }
        pc = (cx0 >> 1) * ny1 + (cy0 >> 1);
        live = a[4] != (T)0;
        w = live ? omega / a[4] : (T)0;
        bc = b[c];
    }
    // the lower levels' stencils, concatenated as in As (level l from
    // 9 off[l]; the last level is one cell): every load is issued before the
    // first store, so the copy costs one memory latency, not one per level
    constexpr int NST = (9 * MG_CLOW + MG_CMAX - 1) / MG_CMAX;
    const int tot9 = K > 1 ? 9 * (B.off[K - 1] + 1) : 0;
    T st[NST];
#pragma unroll
    for (int r = 0; r < NST; ++r) {
        const int e = c + r * MG_CMAX;
        st[r] = (T)0;
        if (e < tot9) {
            const void* src = B.A[1];
            int base = 0;
#pragma unroll
            for (int q = 2; q < MG_CLEV; ++q)  // constant indices into B
                if (q < K && e >= 9 * B.off[q]) {
                    src = B.A[q];
                    base = 9 * B.off[q];
                }
            st[r] = static_cast<const T*>(src)[e - base];
        }
    }
#pragma unroll
    for (int r = 0; r < NST; ++r)
        if (c + r * MG_CMAX < tot9) As[c + r * MG_CMAX] = st[r];
    // the exact level (uniform: every thread reads the same flag)
    const int D = (B.D > 0 && *Aok) ? B.D : 0;
    const int nD = D ? B.nx[D] * B.ny[D] : 0;
    for (int e = c; e < nD * nD; e += MG_CMAX) Ai[e] = Ainv[e];
    // LDS keeps what the last kernel on this CU left there: every iterate
    // starts from zero, and a sweep reads the coarser level's iterate only
    // when it applies its correction
    xt[0][c] = (T)0;
    for (int e = c; e < MG_CLOW; e += MG_CMAX) {
        xl[0][e] = (T)0;
        xl[1][e] = (T)0;
    }
    if (done) return;  // uniform over the block
    __syncthreads();

    // top sweep from parity p (z = x + e(parent) on level 1 at parity q if corr)
    auto top_sweep = [&](int p, bool corr, int q) {
        if (own) {
            const T* xc = xt[p];
            T acc = (T)0;
            T zi;
            if (corr) {
                const T* e = xl[q] + B.off[1];
                const T f = (T)B.over;
                for (int d = 0; d < 9; ++d) acc += a[d] * (xc[nb[d]] + f * e[pa[d]]);
                zi = xc[c] + f * e[pc];
            } else {
                for (int d = 0; d < 9; ++d) acc += a[d] * xc[nb[d]];
                zi = xc[c];
            }
            xt[p ^ 1][c] = live ? zi + w * (bc - acc) : (T)0;
        }
        __syncthreads();
    };
    // lower-level sweep: from zero (x1 = omega D^-1 b into parity 1), or from
    // parity p, with the correction from level l+1 at parity q if corr
    // Every LDS load of a sweep is issued before its first use: the diagonal
    // selects the result at the end instead of guarding the loads, and the
    // cell's column comes from B.rny (see mg_div) instead of an integer
    // division; the arithmetic is unchanged.
    auto low_sweep = [&](int l, int p, bool zero, bool corr, int q) {
        const int nx = B.nx[l], ny = B.ny[l], n = nx * ny, o = B.off[l];
        if (c < n) {
            const T* A = As + 9 * o;
            const T dg = A[4 * n + c];
            const T bc = bl[o + c];
            T nv;
            if (zero) {
                nv = omega * bc / dg;
            } else {
                const int cx = mg_div(c, B.rny[l]), cy = c - cx * ny;
                const T* xc = xl[p] + o;
                T acc = (T)0;
                T zi;
                if (corr) {
                    const int ny2 = B.ny[l + 1];
                    const T* e = xl[q] + B.off[l + 1];
                    const T f = (T)B.over;
#pragma unroll
                    for (int d = 0; d < 9; ++d) {
                        const int ex = mg_clamp(cx + d / 3 - 1, nx);
                        const int ey = mg_clamp(cy + d % 3 - 1, ny);
                        acc += A[d * n + c] * (xc[ex * ny + ey] + f * e[(ex >> 1) * ny2 + (ey >> 1)]);
                    }
                    zi = xc[c] + f * e[(cx >> 1) * ny2 + (cy >> 1)];
                } else {
#pragma unroll
                    for (int d = 0; d < 9; ++d) {
                        const int ex = mg_clamp(cx + d / 3 - 1, nx);
                        const int ey = mg_clamp(cy + d % 3 - 1, ny);
                        acc += A[d * n + c] * xc[ex * ny + ey];
                    }
                    zi = xc[c];
                }
                nv = zi + omega * (bc - acc) / dg;
            }
            xl[zero ? 1 : p ^ 1][o + c] = dg != (T)0 ? nv : (T)0;
        }
        __syncthreads();
    };
    // The sweep from zero of level l (x = omega D^-1 b into parity 1) is
    // pointwise, so the thread that restricts b into cell c of level l also
    // sweeps it, the same expression as low_sweep's, and the level's first
    // barrier phase goes away.  Not on the level solved exactly.
    const int Lx = D ? D : K - 1;  // the level solved exactly
    auto zero_sweep = [&](int l, int cc, T bcc) {
        const int n = B.nx[l] * B.ny[l], o = B.off[l];
        const T dg = As[9 * o + 4 * n + cc];
        const T nv = omega * bcc / dg;
        xl[1][o + cc] = dg != (T)0 ? nv : (T)0;
    };
    // residual of level l at parity p, summed over the children of each cell
    // of level l+1 into its right-hand side (and the sweep from zero of
    // level l+1 unless it is solved exactly)
    auto low_restrict = [&](int l, int p) {
        const int nx = B.nx[l], ny = B.ny[l], n = nx * ny, o = B.off[l];
        const int ny2 = B.ny[l + 1];
        if (c < B.nx[l + 1] * ny2) {
            const int X = mg_div(c, B.rny[l + 1]), Y = c - X * ny2;
            const T* A = As + 9 * o;
            const T* xc = xl[p] + o;
            T s = (T)0;
            for (int u = 0; u < 2; ++u)
                for (int v = 0; v < 2; ++v) {
                    const int cx = 2 * X + u, cy = 2 * Y + v;
                    if (cx >= nx || cy >= ny) continue;
                    const int cc = cx * ny + cy;
                    T acc = (T)0;
#pragma unroll
                    for (int d = 0; d < 9; ++d) {
                        const int ex = mg_clamp(cx + d / 3 - 1, nx), ey = mg_clamp(cy + d % 3 - 1, ny);
                        acc += A[d * n + cc] * xc[ex * ny + ey];
                    }
                    s += bl[o + cc] - acc;
                }
            bl[B.off[l + 1] + c] = s;
            if (l + 1 < Lx) zero_sweep(l + 1, c, s);
        }
        __syncthreads();
    };

    if (K == 1) {  // a single cell: exact
        if (own) x[c] = live ? bc / a[4] : (T)0;
        return;
    }
    const int nu = B.nu;
    int pt = 0;  // parity of the top iterate
    for (int cyc = 0; cyc < B.cycles; ++cyc) {
        for (int s = 0; s < nu; ++s) {
            top_sweep(pt, false, 0);
            pt ^= 1;
        }
        if (own) {
            T acc = (T)0;
            for (int d = 0; d < 9; ++d) acc += a[d] * xt[pt][nb[d]];
            rt[c] = bc - acc;
        }
        __syncthreads();
        {
            const int ny1 = B.ny[1];
            if (c < B.nx[1] * ny1) {
                const int X = c / ny1, Y = c - X * ny1;
                T s = (T)0;
                for (int u = 0; u < 2; ++u)
                    for (int v = 0; v < 2; ++v) {
                        const int cx = 2 * X + u, cy = 2 * Y + v;
                        if (cx < B.nx[0] && cy < ny0) s += rt[cx * ny0 + cy];
                    }
                bl[B.off[1] + c] = s;
                if (1 < Lx) zero_sweep(1, c, s);
            }
            __syncthreads();
        }
        // down the lower levels: smoothing (the sweep from zero was made with
        // the restriction into the level), restriction
        for (int l = 1; l < Lx; ++l) {
            int p = 1;
            for (int s = 1; s < nu; ++s) {
                low_sweep(l, p, false, false, 0);
                p ^= 1;
            }
            low_restrict(l, p);
        }
        if (D) {  // x = Ainv b on level D
            const int o = B.off[D];
            if (c < nD) {
                T acc = (T)0;
                for (int j = 0; j < nD; ++j) acc += Ai[j * nD + c] * bl[o + j];
                xl[1][o + c] = acc;
            }
            __syncthreads();
        } else {  // the single cell of the last level
            const int o = B.off[K - 1];
            if (c == 0) {
                const T dg = As[9 * o + 4];
                xl[1][o] = dg != (T)0 ? bl[o] / dg : (T)0;
            }
            __syncthreads();
        }
        // up: correction folded into the first post-smoothing sweep
        int q = 1;  // parity of level l+1's iterate
        for (int l = Lx - 1; l >= 1; --l) {
            int p = nu & 1;
            low_sweep(l, p, false, true, q);
            p ^= 1;
            for (int s = 1; s < nu; ++s) {
                low_sweep(l, p, false, false, 0);
                p ^= 1;
            }
            q = p;
        }
        top_sweep(pt, true, q);
        pt ^= 1;
        for (int s = 1; s < nu; ++s) {
            top_sweep(pt, false, 0);
            pt ^= 1;
        }
    }
    if (own) x[c] = xt[pt][c];
}

// grid level up: z = x + over e(parent), out = z + omega D^-1 (b - A z)
template <class T>
__global__ void k_mg_grid_up(int nx, int ny, const T* __restrict__ A,
                             const T* __restrict__ b, const T* __restrict__ x, int ny2,
                             const T* __restrict__ e, float over, T* __restrict__ out,
                             const SolverScalars* sc) {
    const int done = sc->done;  // tested before the store
    const int nc = nx * ny;
    int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= nc) return;
    const int cx = c / ny, cy = c - cx * ny;
    const T f = (T)over;
    auto z = [&](int ex, int ey) { return x[ex * ny + ey] + f * e[(ex >> 1) * ny2 + (ey >> 1)]; };
    T acc = (T)0;
#pragma unroll
    for (int d = 0; d < 9; ++d) {
        const int ex = mg_clamp(cx + d / 3 - 1, nx), ey = mg_clamp(cy + d % 3 - 1, ny);
        acc += A[(size_t)d * nc + c] * z(ex, ey);
    }
    const T dg = A[(size_t)4 * nc + c];
    const T zc = z(cx, cy);
    const T v = dg != (T)0 ? zc + mg_omega<T>() * (b[c] - acc) / dg : (T)0;
    if (done) return;
    out[c] = v;
}

// Two grid levels in one launch.  Each dependent launch of a cell level
// costs a few microseconds whatever its size, so consecutive levels are
// paired: one block takes a tile of MG_TX x MG_TY cells of the coarser of
// the pair's outputs and recomputes the one-cell halos it needs on the
// finer levels.  Every value is formed by the same expression, in the same
// order, with the same clamped neighbours as k_mg_down / k_mg_grid_up, so
// the results are theirs bit for bit; each cell is written by the one tile
// that owns it.
constexpr int MG_TX = 2, MG_TY = 8;

// k_mg_down for levels 0 and 1 of the pair: level 0 (nx0 x ny0) smoothed
// from zero and restricted into level 1, level 1 likewise into level 2.
// The tile owns the level-2 cells [X0, X0+TX) x [Y0, Y0+TY), their level-1
// children and level-0 grandchildren.
template <class T>
__global__ void __launch_bounds__(BLOCK) k_mg_down2(int nx0, int ny0, const T* __restrict__ A0,
                                                    const T* __restrict__ b0, T* __restrict__ x0,
                                                    int nx1, int ny1, const T* __restrict__ A1,
                                                    T* __restrict__ b1, T* __restrict__ x1,
                                                    int nx2, int ny2, T* __restrict__ b2,
                                                    const SolverScalars* sc) {
    const int done = sc->done;  // tested once the first loads are requested
    constexpr int TX = MG_TX, TY = MG_TY;
    constexpr int MW = 2 * TX + 2, MH = 2 * TY + 2;  // level 1 under the tile, halo 1
    constexpr int RW = 2 * MW, RH = 2 * MH;          // their level-0 children
    constexpr int FW = RW + 2, FH = RH + 2;          // ... with a halo of 1
    __shared__ T xf[FW * FH];
    __shared__ T rf[RW * RH];
    __shared__ T bm[MW * MH];
    __shared__ T xm[MW * MH];
    __shared__ T rm[4 * TX * TY];
    const int nty = (ny2 + TY - 1) / TY;
    const int X0 = (blockIdx.x / nty) * TX, Y0 = (blockIdx.x % nty) * TY;
    const int mx0 = 2 * X0 - 1, my0 = 2 * Y0 - 1;  // origin of the level-1 region
    const int rx0 = 2 * mx0, ry0 = 2 * my0;        // origin of its children
    const int fx0 = rx0 - 1, fy0 = ry0 - 1;        // origin of the level-0 region
    const int nc0 = nx0 * ny0, nc1 = nx1 * ny1;
    const T omega = mg_omega<T>();
    // level 0 from zero, x = omega D^-1 b (each position holds its clamped
    // cell's value, as the sweeps read clamped neighbours)
    constexpr int NF = (FW * FH + BLOCK - 1) / BLOCK;
    T vf[NF];
#pragma unroll
    for (int q = 0; q < NF; ++q) {
        const int e = threadIdx.x + q * BLOCK;
        vf[q] = (T)0;
        if (e < FW * FH) {
            const int i = e / FH, j = e - i * FH;
            vf[q] = mg_jac0(A0, nc0, mg_clamp(fx0 + i, nx0) * ny0 + mg_clamp(fy0 + j, ny0), b0);
        }
    }
    if (done) return;  // uniform over the grid
#pragma unroll
    for (int q = 0; q < NF; ++q) {
        const int e = threadIdx.x + q * BLOCK;
        if (e < FW * FH) {
            const int i = e / FH, j = e - i * FH;
            const int ox = fx0 + i, oy = fy0 + j;
            xf[e] = vf[q];
            if (ox >= 4 * X0 && ox < 4 * (X0 + TX) && ox < nx0 && oy >= 4 * Y0 &&
                oy < 4 * (Y0 + TY) && oy < ny0)
                x0[mg_clamp(ox, nx0) * ny0 + mg_clamp(oy, ny0)] = vf[q];
        }
    }
    __syncthreads();
    // level-0 residuals of the children of the level-1 region (0 outside)
    for (int e = threadIdx.x; e < RW * RH; e += BLOCK) {
        const int i = e / RH, j = e - i * RH;
        const int gx = rx0 + i, gy = ry0 + j;
        T r = (T)0;
        if (gx >= 0 && gx < nx0 && gy >= 0 && gy < ny0) {
            const int c = gx * ny0 + gy;
            T acc = (T)0;
#pragma unroll
            for (int d = 0; d < 9; ++d) {
                const int ex = mg_clamp(gx + d / 3 - 1, nx0), ey = mg_clamp(gy + d % 3 - 1, ny0);
                acc += A0[(size_t)d * nc0 + c] * xf[(ex - fx0) * FH + (ey - fy0)];
            }
            r = b0[c] - acc;
        }
        rf[e] = r;
    }
    __syncthreads();
    // level-1 right-hand sides: the children's quad sum (r0 + r1) + (r2 + r3)
    for (int e = threadIdx.x; e < MW * MH; e += BLOCK) {
        const int p = e / MH, q = e - p * MH;
        const int li = 2 * p, lj = 2 * q;
        const T b = (rf[li * RH + lj] + rf[li * RH + lj + 1]) +
                    (rf[(li + 1) * RH + lj] + rf[(li + 1) * RH + lj + 1]);
        bm[e] = b;
        const int gx = mx0 + p, gy = my0 + q;
        if (gx >= 2 * X0 && gx < 2 * (X0 + TX) && gx < nx1 && gy >= 2 * Y0 &&
            gy < 2 * (Y0 + TY) && gy < ny1)
            b1[gx * ny1 + gy] = b;
    }
    __syncthreads();
    // level 1 from zero, at clamped cells
    for (int e = threadIdx.x; e < MW * MH; e += BLOCK) {
        const int p = e / MH, q = e - p * MH;
        const int ox = mx0 + p, oy = my0 + q;
        const int gx = mg_clamp(ox, nx1), gy = mg_clamp(oy, ny1);
        const T dg = A1[(size_t)4 * nc1 + gx * ny1 + gy];
        const T v = dg != (T)0 ? omega * bm[(gx - mx0) * MH + (gy - my0)] / dg : (T)0;
        xm[e] = v;
        if (ox >= 2 * X0 && ox < 2 * (X0 + TX) && ox < nx1 && oy >= 2 * Y0 &&
            oy < 2 * (Y0 + TY) && oy < ny1)
            x1[gx * ny1 + gy] = v;
    }
    __syncthreads();
    // level-1 residuals of the tile's level-1 cells (0 outside)
    for (int e = threadIdx.x; e < 4 * TX * TY; e += BLOCK) {
        const int p = e / (2 * TY), q = e - p * (2 * TY);
        const int gx = 2 * X0 + p, gy = 2 * Y0 + q;
        T r = (T)0;
        if (gx < nx1 && gy < ny1) {
            const int c = gx * ny1 + gy;
            T acc = (T)0;
#pragma unroll
            for (int d = 0; d < 9; ++d) {
                const int ex = mg_clamp(gx + d / 3 - 1, nx1), ey = mg_clamp(gy + d % 3 - 1, ny1);
                acc += A1[(size_t)d * nc1 + c] * xm[(ex - mx0) * MH + (ey - my0)];
            }
            r = bm[(gx - mx0) * MH + (gy - my0)] - acc;
        }
        rm[e] = r;
    }
    __syncthreads();
    // level-2 right-hand sides
    if (threadIdx.x < TX * TY) {
        const int P = threadIdx.x / TY, Q = threadIdx.x - P * TY;
        const int X = X0 + P, Y = Y0 + Q;
        if (X < nx2 && Y < ny2) {
            const int li = 2 * P, lj = 2 * Q;
            b2[X * ny2 + Y] = (rm[li * 2 * TY + lj] + rm[li * 2 * TY + lj + 1]) +
                              (rm[(li + 1) * 2 * TY + lj] + rm[(li + 1) * 2 * TY + lj + 1]);
        }
    }
}

// k_mg_grid_up for levels 1 and 0 of the pair: level 1 corrected from
// e2 (level 2's result) and swept into out1, then level 0 corrected from
// out1 and swept into out0.  The tile owns the level-1 cells
// [2 X0, 2 (X0+TX)) x [2 Y0, 2 (Y0+TY)) and their level-0 children.
template <class T>
__global__ void __launch_bounds__(BLOCK) k_mg_up2(int nx0, int ny0, const T* __restrict__ A0,
                                                  const T* __restrict__ b0,
                                                  const T* __restrict__ x0, int nx1, int ny1,
                                                  const T* __restrict__ A1,
                                                  const T* __restrict__ b1,
                                                  const T* __restrict__ x1, int nx2, int ny2,
                                                  const T* __restrict__ e2, float over,
                                                  T* __restrict__ out1, T* __restrict__ out0,
                                                  const SolverScalars* sc) {
    const int done = sc->done;  // tested once the first loads are requested
    constexpr int TX = MG_TX, TY = MG_TY;
    constexpr int FW = 4 * TX + 2, FH = 4 * TY + 2;  // level 0 under the tile, halo 1
    constexpr int MW = 2 * TX + 2, MH = 2 * TY + 2;  // their level-1 parents
    constexpr int ZW = MW + 2, ZH = MH + 2;          // ... with a halo of 1
    __shared__ T z1[ZW * ZH];
    __shared__ T o1[MW * MH];
    __shared__ T z0[FW * FH];
    const int nty = (ny2 + TY - 1) / TY;
    const int X0 = (blockIdx.x / nty) * TX, Y0 = (blockIdx.x % nty) * TY;
    const int fx0 = 4 * X0 - 1, fy0 = 4 * Y0 - 1;  // origin of the level-0 region
    const int mx0 = 2 * X0 - 1, my0 = 2 * Y0 - 1;  // origin of the level-1 region
    const int zx0 = mx0 - 1, zy0 = my0 - 1;        // ... with its halo
    const int nc0 = nx0 * ny0, nc1 = nx1 * ny1;
    const T f = (T)over;
    const T omega = mg_omega<T>();
    // level 1 corrected, z = x + over e(parent), at clamped cells
    for (int e = threadIdx.x; e < ZW * ZH; e += BLOCK) {
        const int i = e / ZH, j = e - i * ZH;
        const int ex = mg_clamp(zx0 + i, nx1), ey = mg_clamp(zy0 + j, ny1);
        z1[e] = x1[ex * ny1 + ey] + f * e2[(ex >> 1) * ny2 + (ey >> 1)];
    }
    if (done) return;  // uniform over the grid
    __syncthreads();
    // level-1 sweep over its region; the tile's own cells are written
    for (int e = threadIdx.x; e < MW * MH; e += BLOCK) {
        const int p = e / MH, q = e - p * MH;
        const int gx = mx0 + p, gy = my0 + q;
        T v = (T)0;
        if (gx >= 0 && gx < nx1 && gy >= 0 && gy < ny1) {
            const int c = gx * ny1 + gy;
            T acc = (T)0;
#pragma unroll
            for (int d = 0; d < 9; ++d) {
                const int ex = mg_clamp(gx + d / 3 - 1, nx1), ey = mg_clamp(gy + d % 3 - 1, ny1);
                acc += A1[(size_t)d * nc1 + c] * z1[(ex - zx0) * ZH + (ey - zy0)];
            }
            const T dg = A1[(size_t)4 * nc1 + c];
            const T zc = z1[(gx - zx0) * ZH + (gy - zy0)];
            v = dg != (T)0 ? zc + omega * (b1[c] - acc) / dg : (T)0;
            if (gx >= 2 * X0 && gx < 2 * (X0 + TX) && gy >= 2 * Y0 && gy < 2 * (Y0 + TY))
                out1[c] = v;
        }
        o1[e] = v;
    }
    __syncthreads();
    // level 0 corrected from the new level-1 values, at clamped cells
    for (int e = threadIdx.x; e < FW * FH; e += BLOCK) {
        const int i = e / FH, j = e - i * FH;
        const int ex = mg_clamp(fx0 + i, nx0), ey = mg_clamp(fy0 + j, ny0);
        z0[e] = x0[ex * ny0 + ey] + f * o1[((ex >> 1) - mx0) * MH + ((ey >> 1) - my0)];
    }
    __syncthreads();
    // level-0 sweep of the tile's cells
    for (int e = threadIdx.x; e < 16 * TX * TY; e += BLOCK) {
        const int p = e / (4 * TY), q = e - p * (4 * TY);
        const int gx = 4 * X0 + p, gy = 4 * Y0 + q;
        if (gx >= nx0 || gy >= ny0) continue;
        const int c = gx * ny0 + gy;
        T acc = (T)0;
#pragma unroll
        for (int d = 0; d < 9; ++d) {
            const int ex = mg_clamp(gx + d / 3 - 1, nx0), ey = mg_clamp(gy + d % 3 - 1, ny0);
            acc += A0[(size_t)d * nc0 + c] * z0[(ex - fx0) * FH + (ey - fy0)];
        }
        const T dg = A0[(size_t)4 * nc0 + c];
        const T zc = z0[(gx - fx0) * FH + (gy - fy0)];
        out0[c] = dg != (T)0 ? zc + omega * (b0[c] - acc) / dg : (T)0;
    }
}

// Level 0 up.  The particle iterate before the sweep is the coarse
// correction alone (the iterate was 0, with no sweep on the way down), so
// z_j = over e1(cell of j) is constant on every cell and
//   (A z)_i = sum_d p9[d][i] z(cell_i + d),
// p9 being row i's entries summed by the neighbour's cell offset d (the
// 9 planes of k_mg_rows9, unit diagonal in the centre plane).  The sweep
// out = z + omega (b - A z) reads 9 values a row in place of the row's
// neighbour list and gathers, and z is never stored.  A particle's cell is
// interior (binning clamps to cells 1 .. n-2 in both directions), so the
// 9 cells around it are in the grid.
template <class T>
__global__ void __launch_bounds__(BLOCK) k_mg_fine_up(int n, size_t ld, int ny,
                                                      const int* __restrict__ cell,
                                                      const T* __restrict__ p9,
                                                      const double* __restrict__ b,
                                                      const T* __restrict__ e1, float over,
                                                      double* __restrict__ out,
                                                      const SolverScalars* sc) {
    const int done = sc->done;  // tested once the row's cell is requested
    int i = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (i >= n) return;
    const int ci = cell[i];
    if (done) return;
    const T f = (T)over;
    T pv[9], zv[9];
#pragma unroll
    for (int d = 0; d < 9; ++d) {
        pv[d] = p9[(size_t)d * ld + i];
        zv[d] = f * e1[ci + (d / 3 - 1) * ny + (d % 3 - 1)];
    }
    T acc = (T)0;
#pragma unroll
    for (int d = 0; d < 9; ++d) acc += pv[d] * zv[d];
    out[i] = (double)(zv[4] + mg_omega<T>() * ((T)b[i] - acc));
}

// ---------------------------------------------------------------------------
// outer BiCGSTAB kernels (FP64 vectors, partials per XCD-mapped block)
// ---------------------------------------------------------------------------
template <class TI, class TV>
__device__ __forceinline__ double mg_spmv(int i, int c, double si, const TI* __restrict__ nbr,
                                          const TV* __restrict__ val, size_t ld,
                                          const double* __restrict__ x) {
    const double xi = x[i];
    double acc = 0.0;
    int k = 0;
    // loads of four entries ahead of their gathers, sum in entry order
    for (; k + 4 <= c; k += 4) {
        const size_t o = (size_t)k * ld + i;
        const int j0 = mg_col(nbr[o], i), j1 = mg_col(nbr[o + ld], i);
        const int j2 = mg_col(nbr[o + 2 * ld], i), j3 = mg_col(nbr[o + 3 * ld], i);
        const double v0 = (double)val[o], v1 = (double)val[o + ld];
        const double v2 = (double)val[o + 2 * ld], v3 = (double)val[o + 3 * ld];
        const double x0 = x[j0], x1 = x[j1], x2 = x[j2], x3 = x[j3];
        acc = acc + v0 * (x0 - xi);
        acc = acc + v1 * (x1 - xi);
        acc = acc + v2 * (x2 - xi);
        acc = acc + v3 * (x3 - xi);
    }
    for (; k < c; ++k) {
        const size_t o = (size_t)k * ld + i;
        acc = acc + (double)val[o] * (x[mg_col(nbr[o], i)] - xi);
    }
    return si * xi + acc;
}

// v = A p^; rv = (r, v); partial (v, rhat)
template <class TI, class TV>
__global__ void __launch_bounds__(BLOCK) k_mg_v(int n, size_t ld, const TI* __restrict__ nbr,
                                                const int* __restrict__ cnt,
                                                const TV* __restrict__ val,
                                                const double* __restrict__ rsum,
                                                const double* __restrict__ ph,
                                                const double2* __restrict__ rq,
                                                const double* __restrict__ rhat,
                                                double2* __restrict__ rv, double* __restrict__ part,
                                                const SolverScalars* sc) {
    // the row's count and row sum are requested before the done test waits
    const int done = sc->done;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int i = lb * BLOCK + threadIdx.x;
    int ci = 0;
    double si = 0.0;
    if (i < n) {
        ci = cnt[i];
        si = rsum[i];
    }
    if (done) return;
    double v = 0.0;
    if (i < n) {
        const double s = mg_spmv<TI, TV>(i, ci, si, nbr, val, ld, ph);
        rv[i] = make_double2(rq[i].x, s);
        v = s * rhat[i];
    }
    v = block_sum(v);
    if (threadIdx.x == 0) part[lb] = v;
}

// t = A s^; partials (t, s), (t, t)
template <class TI, class TV>
__global__ void __launch_bounds__(BLOCK) k_mg_t(int n, size_t ld, const TI* __restrict__ nbr,
                                                const int* __restrict__ cnt,
                                                const TV* __restrict__ val,
                                                const double* __restrict__ rsum,
                                                const double* __restrict__ sh,
                                                const double* __restrict__ s,
                                                double* __restrict__ t, double* __restrict__ part,
                                                const SolverScalars* sc) {
    const int done = sc->done;  // tested once the row's first loads are requested
    // rows in the opposite order to k_mg_v: this pass starts on the matrix
    // rows that k_mg_v read last, and the next k_mg_v on those this one read
    // last, while they are still in the Infinity Cache
    const int lb = xcd_block_rev(blockIdx.x, gridDim.x);
    const int i = lb * BLOCK + threadIdx.x;
    int ci = 0;
    double si = 0.0;
    if (i < n) {
        ci = cnt[i];
        si = rsum[i];
    }
    if (done) return;
    double v0 = 0.0, v1 = 0.0;
    if (i < n) {
        const double ti = mg_spmv<TI, TV>(i, ci, si, nbr, val, ld, sh);
        t[i] = ti;
        v0 = ti * s[i];
        v1 = ti * ti;
    }
    v0 = block_sum(v0);
    v1 = block_sum(v1);
    if (threadIdx.x == 0) {
        part[lb] = v0;
        part[gridDim.x + lb] = v1;
    }
}

// xs += alpha p^ + omega s^; r = s - omega t; q = p - omega v; partials (r, rhat), (r, r)
__global__ void __launch_bounds__(BLOCK) k_mg_x(int n, const double* __restrict__ p,
                                                const double* __restrict__ ph,
                                                const double* __restrict__ s,
                                                const double* __restrict__ sh,
                                                const double* __restrict__ t,
                                                const double2* __restrict__ rv,
                                                const double* __restrict__ rhat,
                                                double* __restrict__ xs, double2* __restrict__ rq,
                                                double* __restrict__ part, const SolverScalars* sc) {
    // the scalars and the row's first operands are requested before the
    // done test waits
    const int done = sc->done;
    const double al = sc->alpha, w = sc->wi;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int i = lb * BLOCK + threadIdx.x;
    double xi = 0.0, phi = 0.0;
    if (i < n) {
        xi = xs[i];
        phi = ph[i];
    }
    if (done) return;
    double v0 = 0.0, v1 = 0.0;
    if (i < n) {
        xs[i] = (xi + al * phi) + w * sh[i];
        const double r = s[i] - w * t[i];
        const double q = p[i] - w * rv[i].y;
        rq[i] = make_double2(r, q);
        v0 = r * rhat[i];
        v1 = r * r;
    }
    v0 = block_sum(v0);
    v1 = block_sum(v1);
    if (threadIdx.x == 0) {
        part[lb] = v0;
        part[gridDim.x + lb] = v1;
    }
}

// FP32 values and FP64 row sums of a Jacobi-scaled FP64 matrix in the
// assembly's layout (unit diagonal at dpos, row sum in column order as
// k_assemble forms it): the mixed-precision operands of a matrix that was
// not assembled, such as the adjoint's transpose
__global__ void k_mg_lower(int n, size_t ld, const int* __restrict__ cnt,
                           const int* __restrict__ dpos, const double* __restrict__ val,
                           float* __restrict__ val32, double* __restrict__ rsum) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const int c = cnt[i], dp = dpos[i];
    double rs = 0.0;
    for (int k = 0; k < c; ++k) {
        const double v = val[(size_t)k * ld + i];
        val32[(size_t)k * ld + i] = (float)v;
        if (k == dp) rs = rs + 1.0;
        rs = rs + v;
    }
    if (dp == c) rs = rs + 1.0;
    rsum[i] = rs;
}

namespace host {

// Level arrays are allocated with 8-byte elements and read as the solve's
// value type (float or double).
struct MGLevel {
    int nx = 0, ny = 0;
    double *A = nullptr, *b = nullptr, *x = nullptr, *x2 = nullptr;
};

struct MGData {
    int nlev = 0;  // grid levels (level 1 .. nlev); the last is a single cell
    int lb = 0;    // first level solved inside one block (k_mg_block)
    MGBlock blk;
    MGLevel lv[MG_MAXLEV];
    double* z0 = nullptr;  // particle residual r0, restricted to the cells
    int* cell = nullptr;   // level-1 cell of every particle
    double* part9 = nullptr;  // per row, its entries summed by neighbour cell offset (9 x cap)
    float* p9f = nullptr;     // the same sums rounded to FP32, for the FP32 V-cycle
    double* ainv = nullptr;   // inverse of the exact block level (MG_DMAX^2 values)
    int* aok = nullptr;       // 1 when ainv holds it
    // The block levels' operators and the inverse are read only by k_mg_block,
    // so they are formed on a side stream while the main stream computes the
    // first residual; the main stream waits for `formed` before the first
    // V-cycle (mg_join).
    hipStream_t side = nullptr;
    hipEvent_t ready = nullptr, formed = nullptr;
    bool pending = false;  // side-stream work not yet joined
};

template <class T>
static inline T* as(double* p) { return reinterpret_cast<T*>(p); }

// the row sums by cell offset in the V-cycle's value type
template <class T>
static inline const T* mg_p9(const MGData* m);
template <>
inline const float* mg_p9<float>(const MGData* m) { return m->p9f; }
template <>
inline const double* mg_p9<double>(const MGData* m) { return m->part9; }

// the matrix values the V-cycle and the outer products read
template <class T>
static inline const T* mg_vals(const MGSystem& sys);
template <>
inline const float* mg_vals<float>(const MGSystem& sys) { return sys.val32; }
template <>
inline const double* mg_vals<double>(const MGSystem& sys) { return sys.val64; }

void mg_free(isph_ctx* c) {
    MGData* m = c->mg;
    if (!m) return;
    if (m->side) (void)hipStreamSynchronize(m->side);
    for (int l = 0; l < m->nlev; ++l) {
        void* p[] = {m->lv[l].A, m->lv[l].b, m->lv[l].x, m->lv[l].x2};
        for (auto q : p)
            if (q) (void)hipFree(q);
    }
    void* p[] = {m->z0, m->cell, m->part9, m->p9f, m->ainv, m->aok};
    for (auto q : p)
        if (q) (void)hipFree(q);
    if (m->ready) (void)hipEventDestroy(m->ready);
    if (m->formed) (void)hipEventDestroy(m->formed);
    if (m->side) (void)hipStreamDestroy(m->side);
    delete m;
    c->mg = nullptr;
}

static int mg_alloc(isph_ctx* c) {
    if (c->mg) return ISPH_OK;
    MGData* m = new MGData();
    c->mg = m;
    int nx = c->grid.nx, ny = c->grid.ny;
    int rc = 0;
    for (;;) {
        if (m->nlev >= MG_MAXLEV) {
            mg_free(c);
            return fail(c, ISPH_E_ARG, "cell grid too large for the multigrid hierarchy");
        }
        MGLevel& L = m->lv[m->nlev++];
        L.nx = nx;
        L.ny = ny;
        const size_t nc = (size_t)nx * ny;
        rc |= dalloc(c, &L.A, 9 * nc);
        rc |= dalloc(c, &L.b, nc);
        rc |= dalloc(c, &L.x, nc);
        rc |= dalloc(c, &L.x2, nc);
        if (nx == 1 && ny == 1) break;
        nx = (nx + 1) / 2;
        ny = (ny + 1) / 2;
    }
    const size_t N = (size_t)c->cap;
    rc |= dalloc(c, &m->z0, N);
    rc |= dalloc(c, &m->cell, N);
    rc |= dalloc(c, &m->part9, 9 * N);
    rc |= dalloc(c, &m->p9f, 9 * N);
    rc |= dalloc(c, &m->ainv, (size_t)MG_DMAX * MG_DMAX);
    rc |= dalloc(c, &m->aok, 1);
    if (!rc && (hipStreamCreateWithFlags(&m->side, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&m->ready, hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&m->formed, hipEventDisableTiming) != hipSuccess))
        rc = ISPH_E_HIP;
    if (rc) {
        mg_free(c);
        return fail(c, ISPH_E_HIP, "multigrid hierarchy allocation failed");
    }
    // the block's top level: the first with at most MG_CMAX cells whose
    // lower levels fit the block's LDS (the last level, one cell, always does)
    for (int l = 0; l < m->nlev; ++l) {
        int64_t low = 0;
        for (int k = l + 1; k < m->nlev; ++k) low += (int64_t)m->lv[k].nx * m->lv[k].ny;
        if ((int64_t)m->lv[l].nx * m->lv[l].ny <= MG_CMAX && low <= MG_CLOW &&
            m->nlev - l <= MG_CLEV) {
            m->lb = l;
            break;
        }
    }
    MGBlock& B = m->blk;
    B.K = m->nlev - m->lb;
    B.nu = MG_BLOCK_NU;
    B.cycles = mg_block_cycles(B.K);
    B.over = MG_OVER;
    int off = 0;
    for (int k = 0; k < B.K; ++k) {
        const MGLevel& L = m->lv[m->lb + k];
        B.nx[k] = L.nx;
        B.ny[k] = L.ny;
        B.A[k] = L.A;
        B.rny[k] = 1.0f / (float)L.ny;
        B.off[k] = k == 0 ? 0 : off;
        if (k > 0) off += L.nx * L.ny;
    }
    // the exact level: the first below the top with 2..MG_DMAX cells
    B.D = 0;
    for (int k = 1; k + 1 < B.K; ++k)
        if (B.nx[k] * B.ny[k] <= MG_DMAX) {
            B.D = k;
            break;
        }
    return ISPH_OK;
}

// Galerkin hierarchy of the system's matrix (FP64 sums, stored as T)
template <class T, class TI>
static int mg_setup(isph_ctx* c, const MGSystem& sys, const TI* nbr) {
    int rc = mg_alloc(c);
    if (rc) return rc;
    MGData* m = c->mg;
    const GridInfo& g = c->grid;
    k_mg_cells<<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, g, c->key_s, m->cell);
    if (c->n)
        k_mg_rows9<TI><<<nblk(c->n), BLOCK, 0, c->stream>>>(
            c->n, g.ny, (size_t)c->cap, nbr, c->cnt, sys.val64, m->cell, m->part9,
            std::is_same<T, float>::value ? m->p9f : nullptr, c->err);
    k_mg_galerkin1<T><<<nblk((int64_t)g.nx * g.ny), BLOCK, 0, c->stream>>>(
        g, c->bincnt, (size_t)c->cap, m->part9, as<T>(m->lv[0].A));
    // the grid levels' operators on the main stream, the block levels'
    // (lb and below) and the inverse on the side stream
    auto galerkin = [&](int l, hipStream_t st) {
        const MGLevel &a = m->lv[l], &b = m->lv[l + 1];
        k_mg_galerkin<T><<<nblk((int64_t)b.nx * b.ny), BLOCK, 0, st>>>(a.nx, a.ny, as<T>(a.A),
                                                                      b.nx, b.ny, as<T>(b.A));
    };
    int l = 0;
    for (; l + 1 < m->nlev && l + 1 < m->lb; ++l) galerkin(l, c->stream);
    HIPCHK(c, hipEventRecord(m->ready, c->stream));
    HIPCHK(c, hipStreamWaitEvent(m->side, m->ready, 0));
    for (; l + 1 < m->nlev; ++l) galerkin(l, m->side);
    if (m->blk.D) {
        const MGLevel& L = m->lv[m->lb + m->blk.D];
        k_mg_coarse_inv<T><<<1, MG_CMAX, 0, m->side>>>(L.nx, L.ny, as<T>(L.A), as<T>(m->ainv),
                                                        m->aok);
    }
    HIPCHK(c, hipEventRecord(m->formed, m->side));
    m->pending = true;
    HIPCHK(c, hipGetLastError());
    return ISPH_OK;
}

// the main stream waits for the side stream's operators and inverse: before
// the first V-cycle, and before the solve returns (the next set-up rewrites
// what the side stream reads)
static int mg_join(isph_ctx* c) {
    MGData* m = c->mg;
    if (!m || !m->pending) return ISPH_OK;
    m->pending = false;
    HIPCHK(c, hipStreamWaitEvent(c->stream, m->formed, 0));
    return ISPH_OK;
}

// one V-cycle: b formed from `pairs` (MODE 0: p from rq, MODE 1: s from rv)
// into bout, preconditioned vector into out
template <class T, class TI, int MODE, bool FIRST>
static void mg_vcycle(isph_ctx* c, const MGSystem& sys, const TI* nbr, const double2* pairs,
                      double* bout, double* out) {
    MGData* m = c->mg;
    const GridInfo& g = c->grid;
    const int n = c->n;
    const size_t ld = (size_t)c->cap;
    k_mg_bpoint<T, MODE, FIRST><<<nblk(n), BLOCK, 0, c->stream>>>(n, pairs, c->sc, bout,
                                                                   as<T>(m->z0));
    k_mg_restrict1<T><<<nblk((int64_t)g.nx * g.ny), BLOCK, 0, c->stream>>>(
        g, c->bincnt, as<T>(m->z0), as<T>(m->lv[0].b), c->sc);
    const int lb = m->lb;
    // the grid levels down, two per launch (k_mg_down2), the last one alone
    // when their number is odd
    int l = 0;
    for (; l + 1 < lb; l += 2) {
        MGLevel &a = m->lv[l], &b = m->lv[l + 1], &d = m->lv[l + 2];
        const int tiles = ((d.nx + MG_TX - 1) / MG_TX) * ((d.ny + MG_TY - 1) / MG_TY);
        k_mg_down2<T><<<tiles, BLOCK, 0, c->stream>>>(
            a.nx, a.ny, as<T>(a.A), as<T>(a.b), as<T>(a.x), b.nx, b.ny, as<T>(b.A), as<T>(b.b),
            as<T>(b.x), d.nx, d.ny, as<T>(d.b), c->sc);
    }
    if (l < lb) {
        MGLevel &a = m->lv[l], &b = m->lv[l + 1];
        k_mg_down<T><<<nblk(4 * (int64_t)b.nx * b.ny), BLOCK, 0, c->stream>>>(
            a.nx, a.ny, as<T>(a.A), as<T>(a.b), as<T>(a.x), b.nx, b.ny, as<T>(b.b), c->sc);
    }
    k_mg_block<T><<<1, MG_CMAX, 0, c->stream>>>(m->blk, as<T>(m->lv[lb].b), as<T>(m->lv[lb].x),
                                                 as<T>(m->ainv), m->aok, c->sc);
    // up: the odd level alone first, then pairs (k_mg_up2)
    l = lb - 1;
    if (lb & 1) {
        MGLevel &a = m->lv[l], &b = m->lv[l + 1];
        k_mg_grid_up<T><<<nblk((int64_t)a.nx * a.ny), BLOCK, 0, c->stream>>>(
            a.nx, a.ny, as<T>(a.A), as<T>(a.b), as<T>(a.x), b.ny, as<T>(b.x), MG_OVER, as<T>(a.x2),
            c->sc);
        std::swap(a.x, a.x2);
        --l;
    }
    for (; l >= 1; l -= 2) {
        MGLevel &a = m->lv[l - 1], &b = m->lv[l], &d = m->lv[l + 1];
        const int tiles = ((d.nx + MG_TX - 1) / MG_TX) * ((d.ny + MG_TY - 1) / MG_TY);
        k_mg_up2<T><<<tiles, BLOCK, 0, c->stream>>>(
            a.nx, a.ny, as<T>(a.A), as<T>(a.b), as<T>(a.x), b.nx, b.ny, as<T>(b.A), as<T>(b.b),
            as<T>(b.x), d.nx, d.ny, as<T>(d.x), MG_OVER, as<T>(b.x2), as<T>(a.x2), c->sc);
        std::swap(b.x, b.x2);
        std::swap(a.x, a.x2);
    }
    k_mg_fine_up<T><<<nblk(n), BLOCK, 0, c->stream>>>(n, ld, g.ny, m->cell, mg_p9<T>(m), bout,
                                                       as<T>(m->lv[0].x), MG_OVER, out, c->sc);
}

template <class T, class TI>
static int mg_iterate(isph_ctx* c, const MGSystem& sys, const TI* nbr, const SolveCtl& ctl,
                      bool& first) {
    const T* val = mg_vals<T>(sys);
    const int grid = nblk8(c->n);
    const int n = c->n;
    const size_t ld = (size_t)c->cap;
    double *rhat = c->v64[0], *p = c->v64[1], *s = c->v64[2], *t = c->v64[3], *xs = c->v64[4];
    double *ph = c->v64[5], *sh = c->v64[6];
    if (int rc = mg_join(c)) return rc;
    for (;;) {
        int batch = 4;
        const SolverScalars& h = *c->sc_host;
        if (h.nb > 0 && h.rr > 0) {
            double target = std::max(ctl.tol * h.nb, ctl.delta * h.rupd);
            const double ratio = std::sqrt(h.rr) / target;
            batch = ratio > 100.0 ? 8 : (ratio > 3.0 ? 4 : 2);
        }
        for (int b = 0; b < batch; ++b) {
            if (first)
                mg_vcycle<T, TI, 0, true>(c, sys, nbr, c->rq64, p, ph);
            else
                mg_vcycle<T, TI, 0, false>(c, sys, nbr, c->rq64, p, ph);
            k_mg_v<TI, T><<<grid, BLOCK, 0, c->stream>>>(n, ld, nbr, c->cnt, val, sys.rsum, ph,
                                                          c->rq64, rhat, c->rv64, c->part, c->sc);
            k_finish<<<1, FIN_BLOCK, 0, c->stream>>>(FIN_ALPHA, grid, c->part, c->sc, ctl);
            mg_vcycle<T, TI, 1, false>(c, sys, nbr, c->rv64, s, sh);
            k_mg_t<TI, T><<<grid, BLOCK, 0, c->stream>>>(n, ld, nbr, c->cnt, val, sys.rsum, sh, s,
                                                          t, c->part, c->sc);
            k_finish<<<1, FIN_BLOCK, 0, c->stream>>>(FIN_OMEGA, grid, c->part, c->sc, ctl);
            k_mg_x<<<grid, BLOCK, 0, c->stream>>>(n, p, ph, s, sh, t, c->rv64, rhat, xs, c->rq64,
                                                   c->part, c->sc);
            k_finish<<<1, FIN_BLOCK, 0, c->stream>>>(FIN_BETA, grid, c->part, c->sc, ctl);
            first = false;
        }
        HIPCHK(c, hipGetLastError());
        int rc = read_sc(c);
        if (rc) return rc;
        if (c->sc_host->done) return ISPH_OK;
    }
}

template <class T, class TI>
static int solve_mg_body(isph_ctx* c, const MGSystem& sys, const TI* nbr) {
    const int grid = nblk8(c->n);
    double* x = sys.x;
    double* r64 = c->v64[9];
    int rc = mg_setup<T, TI>(c, sys, nbr);
    if (rc) return rc;
    SolveCtl ctl = solve_ctl(c, 1);
    SolveCtl ctl_flush = solve_ctl(c, 2);
    ResidArgs ra = resid_args(c, x, r64, sys.val64, sys.rhs);
    if constexpr (std::is_same<TI, short>::value) ra.nbr16 = nbr;
    k_resid<<<grid, BLOCK, 0, c->stream>>>(ra, 0, nullptr, nullptr);
    k_finish<<<1, FIN_BLOCK, 0, c->stream>>>(FIN_INIT, grid, c->part, c->sc, solve_ctl(c, 0));
    k_finish<<<1, FIN_BLOCK, 0, c->stream>>>(FIN_RESID, grid, c->part, c->sc, ctl);
    HIPCHK(c, hipGetLastError());
    if ((rc = read_sc(c))) return rc;
    if (c->sc_host->done == DONE_ZERO_RHS) {
        k_fill<<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, x, 0.0);
        return ISPH_OK;
    }
    if (c->sc_host->done == DONE_CONVERGED) return ISPH_OK;
    k_mx_start<double><<<grid, BLOCK, 0, c->stream>>>(c->n, r64, c->rq64, c->v64[0], c->v64[4],
                                                      c->part, c->sc);
    k_finish<<<1, FIN_BLOCK, 0, c->stream>>>(FIN_START, grid, c->part, c->sc, ctl);
    HIPCHK(c, hipGetLastError());
    if ((rc = read_sc(c))) return rc;
    bool first = true;
    for (;;) {
        if ((rc = mg_iterate<T, TI>(c, sys, nbr, ctl, first))) return rc;
        const int d = c->sc_host->done;
        if (d == DONE_FAILED) return fail(c, ISPH_E_NOCONV, "multigrid PPE solve did not converge");
        if (d != DONE_UPDATE) return fail(c, ISPH_E_STATE, "unexpected solver state");
        k_mx_flush<double><<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, x, c->v64[4]);
        k_resid<<<grid, BLOCK, 0, c->stream>>>(ra, 0, nullptr, nullptr);
        k_finish<<<1, FIN_BLOCK, 0, c->stream>>>(FIN_RESID, grid, c->part, c->sc, ctl_flush);
        HIPCHK(c, hipGetLastError());
        if ((rc = read_sc(c))) return rc;
        if (c->sc_host->done == DONE_CONVERGED) return ISPH_OK;
        if (c->sc_host->done == DONE_FAILED)
            return fail(c, ISPH_E_NOCONV, "multigrid PPE solve did not converge");
        if (c->sc_host->brk) {
            k_mx_start<double><<<grid, BLOCK, 0, c->stream>>>(c->n, r64, c->rq64, c->v64[0],
                                                              c->v64[4], c->part, c->sc);
            k_finish<<<1, FIN_BLOCK, 0, c->stream>>>(FIN_START, grid, c->part, c->sc, ctl);
            first = true;
        } else {
            k_mx_replace<double><<<grid, BLOCK, 0, c->stream>>>(c->n, r64, c->rq64, c->v64[0],
                                                                c->part, c->sc);
            k_finish<<<1, FIN_BLOCK, 0, c->stream>>>(FIN_REPLACE, grid, c->part, c->sc, ctl);
        }
        HIPCHK(c, hipGetLastError());
        if ((rc = read_sc(c))) return rc;
    }
}

template <class T, class TI>
static int solve_mg_t(isph_ctx* c, const MGSystem& sys, const TI* nbr) {
    const int rc = solve_mg_body<T, TI>(c, sys, nbr);
    const int rj = mg_join(c);
    return rc ? rc : rj;
}

int solve_mg(isph_ctx* c) {
    return solve_mg_on(c, MGSystem{c->val64, c->val32, c->rsum, c->rhs, c->v64[8]});
}

int solve_mg_on(isph_ctx* c, const MGSystem& sys) {
    k_to_short<<<nblk(c->n), BLOCK, 0, c->stream>>>(c->n, (size_t)c->cap, c->nbr, c->cnt, c->nbr16,
                                                     c->err);
    uint32_t e = 0;
    HIPCHK(c, hipMemcpyAsync(&e, c->err, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const bool narrow = !(e & ERR_INDEX16);
    if (!narrow) HIPCHK(c, hipMemsetAsync(c->err, 0, sizeof(uint32_t), c->stream));
    c->narrow_idx = narrow;
    if (c->prm.precision == ISPH_PREC_FP64_MG)
        return narrow ? solve_mg_t<double, short>(c, sys, c->nbr16)
                      : solve_mg_t<double, int>(c, sys, c->nbr);
    return narrow ? solve_mg_t<float, short>(c, sys, c->nbr16)
                  : solve_mg_t<float, int>(c, sys, c->nbr);
}

}  // namespace host
}  // namespace isph
