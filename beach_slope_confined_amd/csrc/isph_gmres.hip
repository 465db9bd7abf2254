// isph_gmres.hip -- the reference's live PPE solve (ISPH_PREC_FGMRES):
// restarted GMRES on the Jacobi-scaled matrix, as solver.f90:306-347 drives
// MKL's RCI dfgmres.  No preconditioner is set (ipar(11) = 0), so FGMRES
// (Saad 1993) reduces to GMRES with z_j = v_j.  The reference's settings:
//   relative tolerance dpar(1) = tol (1e-3 there) on the residual estimate
//   |g_(j+1)| of the Givens rotations against the initial residual
//   (ipar(9) = 1: dpar(5) <= dpar(1) * dpar(3));
//   zero-norm test of every new Arnoldi vector (ipar(12) = 1,
//   dpar(8) = 1e-12);
//   restart and iteration limit min(150, n) (the defaults of ipar(15) and
//   ipar(5); `inner_maxit` sets another restart length, `maxit` the limit);
//   warm start from the previous solution vector in row order (fsol);
//   the solution is fetched only on convergence (dfgmres_get runs for
//   rci_request = 0 alone), so on reaching the limit fsol and the pressure
//   keep the warm start.
// The tests are taken in that order after every iteration: residual, zero
// norm, iteration limit, end of the restart cycle.
//
// One Arnoldi step j (classical Gram-Schmidt applied twice, CGS2):
//   k_gm_arnoldi  v_j = u / beta, w = A v_j, partials (v_i, w), i <= j
//   k_gm_fin      h1_i, one block per i, k_finish's tree
//   k_gm_orth     u = w - sum_i h1_i v_i (i ascending), partials (v_i, u)
//   k_gm_fin      h2_i
//   k_gm_orth2    u = u - sum_i h2_i v_i, partial (u, u)
//   k_gm_givens   one thread: column h1 + h2, the earlier rotations, a new
//                 rotation, g, the stop tests
// Every dot product uses the block tree and k_finish's tree of the BiCGSTAB
// kernels, the row product is spmv_row, and the scalar work runs in one
// thread, so the oracle repeats each operation in the same order.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "isph_host.h"
#include "isph_kernels.h"
#include "isph_spmv.h"

namespace isph {

constexpr int GM_DEFAULT_M = 150;       // MKL's default restart length and limit
constexpr int GM_MAXM = 512;            // longest restart cycle accepted
constexpr double GM_ZERO_NORM = 1e-12;  // dpar(8)
constexpr int GM_CHUNK = 8;             // basis vectors per block reduction round
constexpr int GM_SUB = 8;               // basis vectors loaded ahead in a subtraction
constexpr int GM_BATCH = 8;             // Arnoldi steps enqueued between host reads

enum { GM_RUN = 0, GM_CONV = 1, GM_HAPPY = 2, GM_CYCLE = 3, GM_MAXIT = 4, GM_FAIL = 5 };

struct GmState {
    double beta;    // norm of the newest Arnoldi vector before normalisation
    double r0;      // ||b - A x0|| of the warm start
    double tolabs;  // tol * r0 (dpar(4))
    double res;     // residual estimate |g_(j+1)| (dpar(5))
    int j;          // Arnoldi columns of this cycle
    int total;      // iterations of the solve
    int done;       // GM_*
    int pad;
};

template <class TI>
struct GmMat {
    int n;
    size_t ld;
    const TI* nbr;
    const int* cnt;
    const int* dpos;
    const double* val;
};

// K block sums in the tree of block_sum (wave butterfly, (w0+w1)+(w2+w3))
template <int K>
__device__ __forceinline__ void block_sum_k(double (&v)[K]) {
    __shared__ double red[K][4];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) red[k][threadIdx.x >> 6] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = (red[k][0] + red[k][1]) + (red[k][2] + red[k][3]);
}

// partials (v_i, w) for i = 0..j into part[i * nbp + lb]; rows past n add 0
__device__ __forceinline__ void gm_dots(bool valid, int r, int j, const double* __restrict__ V,
                                        size_t cap, double w, double* __restrict__ part,
                                        int nbp, int lb) {
    for (int i0 = 0; i0 <= j; i0 += GM_CHUNK) {
        double v[GM_CHUNK];
#pragma unroll
        for (int k = 0; k < GM_CHUNK; ++k) {
            const int i = i0 + k;
            v[k] = (valid && i <= j) ? V[(size_t)i * cap + r] * w : 0.0;
        }
        block_sum_k<GM_CHUNK>(v);
        if (threadIdx.x == 0)
#pragma unroll
            for (int k = 0; k < GM_CHUNK; ++k)
                if (i0 + k <= j) part[(size_t)(i0 + k) * nbp + lb] = v[k];
    }
}

// x - h_0 v_0 - h_1 v_1 - ... - h_j v_j at row r, subtracted in order; the
// loads of GM_SUB basis vectors are issued ahead of their subtractions
__device__ __forceinline__ double gm_subtract(double x, int r, int j, const double* __restrict__ V,
                                              size_t cap, const double* __restrict__ h) {
    for (int i0 = 0; i0 <= j; i0 += GM_SUB) {
        double v[GM_SUB];
#pragma unroll
        for (int k = 0; k < GM_SUB; ++k) v[k] = i0 + k <= j ? V[(size_t)(i0 + k) * cap + r] : 0.0;
#pragma unroll
        for (int k = 0; k < GM_SUB; ++k)
            if (i0 + k <= j) x = x - h[i0 + k] * v[k];
    }
    return x;
}

// u = b - A x, partial (u, u)
template <class TI>
__global__ void __launch_bounds__(BLOCK) k_gm_resid(GmMat<TI> A, const double* __restrict__ b,
                                                    const double* __restrict__ x,
                                                    double* __restrict__ u,
                                                    double* __restrict__ part) {
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int r = lb * BLOCK + threadIdx.x;
    double q = 0.0;
    if (r < A.n) {
        auto xj = [&](int c) { return x[c]; };
        const double s = spmv_row<double, double, TI>(r, A.cnt[r], A.dpos[r], A.nbr, A.val, A.ld,
                                                      x[r], xj);
        const double ur = b[r] - s;
        u[r] = ur;
        q = ur * ur;
    }
    q = block_sum(q);
    if (threadIdx.x == 0) part[lb] = q;
}

// start of a cycle from the residual in u: beta = g_0 = ||u||; on the first
// cycle also r0 and the absolute tolerance
__global__ void __launch_bounds__(FIN_BLOCK) k_gm_start(int nbp, const double* __restrict__ part,
                                                        double* __restrict__ g, GmState* st,
                                                        double tol, int first) {
    const double t = fin_sum(part, nbp);
    if (threadIdx.x != 0) return;
    const double r = sqrt(t);
    if (first) {
        st->r0 = r;
        st->tolabs = tol * r;
        st->total = 0;
    }
    st->beta = r;
    st->res = r;
    st->j = 0;
    g[0] = r;
    if (!isfinite(r))
        st->done = GM_FAIL;
    else if (r == 0.0 || r <= st->tolabs)
        st->done = GM_CONV;
    else
        st->done = GM_RUN;
}

// v_j = u / beta (stored), w = A v_j, partials (v_i, w) for i <= j
template <class TI>
__global__ void __launch_bounds__(BLOCK) k_gm_arnoldi(GmMat<TI> A, int j, double* __restrict__ V,
                                                      size_t cap, const double* __restrict__ u,
                                                      double* __restrict__ w, const GmState* st,
                                                      double* __restrict__ part) {
    if (st->done) return;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int r = lb * BLOCK + threadIdx.x;
    const bool valid = r < A.n;
    const double beta = st->beta;
    double wr = 0.0;
    if (valid) {
        const double vr = u[r] / beta;
        V[(size_t)j * cap + r] = vr;
        auto vj = [&](int c) { return u[c] / beta; };
        wr = spmv_row<double, double, TI>(r, A.cnt[r], A.dpos[r], A.nbr, A.val, A.ld, vr, vj);
        w[r] = wr;
    }
    gm_dots(valid, r, j, V, cap, wr, part, gridDim.x, lb);
}

// out[i] = the finished sum of partial column i (one block per column)
__global__ void __launch_bounds__(FIN_BLOCK) k_gm_fin(int nbp, const double* __restrict__ part,
                                                      double* __restrict__ out, const GmState* st) {
    if (st->done) return;
    const double t = fin_sum(part + (size_t)blockIdx.x * nbp, nbp);
    if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// u = w - sum_i h_i v_i, partials (v_i, u) for the second pass
__global__ void __launch_bounds__(BLOCK) k_gm_orth(int n, int j, const double* __restrict__ V,
                                                   size_t cap, const double* __restrict__ w,
                                                   const double* __restrict__ h,
                                                   double* __restrict__ u, const GmState* st,
                                                   double* __restrict__ part) {
    if (st->done) return;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int r = lb * BLOCK + threadIdx.x;
    const bool valid = r < n;
    double x = 0.0;
    if (valid) {
        x = gm_subtract(w[r], r, j, V, cap, h);
        u[r] = x;
    }
    gm_dots(valid, r, j, V, cap, x, part, gridDim.x, lb);
}

// u = u - sum_i h2_i v_i, partial (u, u)
__global__ void __launch_bounds__(BLOCK) k_gm_orth2(int n, int j, const double* __restrict__ V,
                                                    size_t cap, const double* __restrict__ h,
                                                    double* __restrict__ u, const GmState* st,
                                                    double* __restrict__ part) {
    if (st->done) return;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int r = lb * BLOCK + threadIdx.x;
    double q = 0.0;
    if (r < n) {
        const double x = gm_subtract(u[r], r, j, V, cap, h);
        u[r] = x;
        q = x * x;
    }
    q = block_sum(q);
    if (threadIdx.x == 0) part[lb] = q;
}

// column j of the Hessenberg matrix (h1 + h2, beta), rotated by the earlier
// Givens rotations and a new one that zeroes beta; R holds the rotated
// columns (column j at R + j (m + 1)), g the rotated right-hand side
__global__ void __launch_bounds__(FIN_BLOCK) k_gm_givens(int j, int m, int maxit, int nbp,
                                                         const double* __restrict__ part,
                                                         const double* __restrict__ h1,
                                                         const double* __restrict__ h2,
                                                         double* __restrict__ R,
                                                         double* __restrict__ cs,
                                                         double* __restrict__ sn,
                                                         double* __restrict__ g, GmState* st) {
    if (st->done) return;
    // the column and the earlier rotations staged in LDS for thread 0
    __shared__ double scol[GM_MAXM + 1], scs[GM_MAXM], ssn[GM_MAXM];
    for (int i = threadIdx.x; i <= j; i += FIN_BLOCK) {
        scol[i] = h1[i] + h2[i];
        if (i < j) {
            scs[i] = cs[i];
            ssn[i] = sn[i];
        }
    }
    const double t = fin_sum(part, nbp);  // its barriers also order the staging
    double* col = R + (size_t)j * (m + 1);
    if (threadIdx.x == 0) {
        for (int i = 0; i < j; ++i) {
            const double a = scol[i], b = scol[i + 1];
            scol[i] = scs[i] * a + ssn[i] * b;
            scol[i + 1] = scs[i] * b - ssn[i] * a;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < j; i += FIN_BLOCK) col[i] = scol[i];
    if (threadIdx.x != 0) return;
    const double beta = sqrt(t);
    const double a = scol[j];
    const double d = sqrt(a * a + beta * beta);
    st->beta = beta;
    st->total = st->total + 1;
    if (!(d > 0.0) || !isfinite(d)) {
        st->done = GM_FAIL;
        return;
    }
    cs[j] = a / d;
    sn[j] = beta / d;
    col[j] = d;
    const double gj = g[j];
    g[j + 1] = -(sn[j] * gj);
    g[j] = cs[j] * gj;
    st->j = j + 1;
    const double res = fabs(g[j + 1]);
    st->res = res;
    if (!isfinite(res))
        st->done = GM_FAIL;
    else if (res <= st->tolabs)
        st->done = GM_CONV;
    else if (beta <= GM_ZERO_NORM)
        st->done = GM_HAPPY;
    else if (st->total >= maxit)
        st->done = GM_MAXIT;
    else if (j + 1 >= m)
        st->done = GM_CYCLE;
}

// y = R^-1 g on the first k columns (back substitution, one thread)
__global__ void k_gm_lsq(int k, int m, const double* __restrict__ R, const double* __restrict__ g,
                         double* __restrict__ y) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int i = k - 1; i >= 0; --i) {
        double s = g[i];
        for (int l = i + 1; l < k; ++l) s = s - R[(size_t)l * (m + 1) + i] * y[l];
        y[i] = s / R[(size_t)i * (m + 1) + i];
    }
}

// x = x + sum_i y_i v_i (i ascending)
__global__ void __launch_bounds__(BLOCK) k_gm_update(int n, int k, const double* __restrict__ V,
                                                     size_t cap, const double* __restrict__ y,
                                                     double* __restrict__ x) {
    const int r = blockIdx.x * BLOCK + threadIdx.x;
    if (r >= n) return;
    double t = x[r];
    for (int i = 0; i < k; ++i) t = t + y[i] * V[(size_t)i * cap + r];
    x[r] = t;
}

namespace host {

struct GmData {
    int m = 0;
    int64_t cap = 0;
    double *V = nullptr, *w = nullptr, *u = nullptr, *part = nullptr;
    double *h1 = nullptr, *h2 = nullptr, *R = nullptr, *cs = nullptr, *sn = nullptr;
    double *g = nullptr, *y = nullptr;
    GmState* st = nullptr;
    GmState* st_host = nullptr;  // pinned
};

void gm_free(isph_ctx* c) {
    GmData* d = c->gm;
    if (!d) return;
    void* p[] = {d->V, d->w, d->u, d->part, d->h1, d->h2, d->R, d->cs, d->sn, d->g, d->y, d->st};
    for (auto q : p)
        if (q) (void)hipFree(q);
    if (d->st_host) (void)hipHostFree(d->st_host);
    delete d;
    c->gm = nullptr;
}

static int gm_alloc(isph_ctx* c, int m) {
    if (c->gm && c->gm->m >= m && c->gm->cap >= c->cap) return ISPH_OK;
    gm_free(c);
    GmData* d = new GmData();
    c->gm = d;
    d->m = m;
    d->cap = c->cap;
    const size_t cap = (size_t)c->cap, nbp = (size_t)nblk8(c->cap);
    int rc = 0;
    rc |= dalloc(c, &d->V, (size_t)(m + 1) * cap);
    rc |= dalloc(c, &d->w, cap);
    rc |= dalloc(c, &d->u, cap);
    rc |= dalloc(c, &d->part, (size_t)(m + 1) * nbp);
    rc |= dalloc(c, &d->h1, (size_t)m + 1);
    rc |= dalloc(c, &d->h2, (size_t)m + 1);
    rc |= dalloc(c, &d->R, (size_t)(m + 1) * m);
    rc |= dalloc(c, &d->cs, (size_t)m);
    rc |= dalloc(c, &d->sn, (size_t)m);
    rc |= dalloc(c, &d->g, (size_t)m + 1);
    rc |= dalloc(c, &d->y, (size_t)m);
    if (hipMalloc((void**)&d->st, sizeof(GmState)) != hipSuccess) rc = 1;
    if (hipHostMalloc((void**)&d->st_host, sizeof(GmState)) != hipSuccess) rc = 1;
    if (rc) {
        gm_free(c);
        return fail(c, ISPH_E_HIP, "GMRES workspace allocation failed");
    }
    return ISPH_OK;
}

static int gm_read(isph_ctx* c) {
    GmData* d = c->gm;
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(d->st_host, d->st, sizeof(GmState), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ISPH_OK;
}

template <class TI>
static int gm_run(isph_ctx* c, const GmMat<TI>& A, double* x, int m, bool& converged) {
    GmData* d = c->gm;
    const int n = A.n;
    const int grid = nblk8(n);
    const size_t cap = (size_t)c->cap;
    const int maxit = c->prm.maxit;
    k_gm_resid<TI><<<grid, BLOCK, 0, c->stream>>>(A, c->rhs, x, d->u, d->part);
    k_gm_start<<<1, FIN_BLOCK, 0, c->stream>>>(grid, d->part, d->g, d->st, c->prm.tol, 1);
    int rc;
    for (;;) {
        if ((rc = gm_read(c))) return rc;
        const GmState h = *d->st_host;
        if (h.done == GM_RUN) {
            const int nb = std::min(GM_BATCH, m - h.j);
            for (int b = 0; b < nb; ++b) {
                const int j = h.j + b;
                k_gm_arnoldi<TI><<<grid, BLOCK, 0, c->stream>>>(A, j, d->V, cap, d->u, d->w, d->st,
                                                                d->part);
                k_gm_fin<<<j + 1, FIN_BLOCK, 0, c->stream>>>(grid, d->part, d->h1, d->st);
                k_gm_orth<<<grid, BLOCK, 0, c->stream>>>(n, j, d->V, cap, d->w, d->h1, d->u, d->st,
                                                         d->part);
                k_gm_fin<<<j + 1, FIN_BLOCK, 0, c->stream>>>(grid, d->part, d->h2, d->st);
                k_gm_orth2<<<grid, BLOCK, 0, c->stream>>>(n, j, d->V, cap, d->h2, d->u, d->st,
                                                          d->part);
                k_gm_givens<<<1, FIN_BLOCK, 0, c->stream>>>(j, m, maxit, grid, d->part, d->h1,
                                                            d->h2, d->R, d->cs, d->sn, d->g, d->st);
            }
            continue;
        }
        c->iters = h.total;
        c->relres = h.r0 > 0.0 ? h.res / h.r0 : 0.0;
        if (h.done == GM_FAIL) return fail(c, ISPH_E_NOCONV, "GMRES breakdown (non-finite or singular)");
        if (h.done == GM_MAXIT) {
            converged = false;
            return ISPH_OK;
        }
        // GM_CONV, GM_HAPPY, GM_CYCLE: x += V y over the cycle's columns
        if (h.j > 0) {
            k_gm_lsq<<<1, 64, 0, c->stream>>>(h.j, m, d->R, d->g, d->y);
            k_gm_update<<<nblk(n), BLOCK, 0, c->stream>>>(n, h.j, d->V, cap, d->y, x);
        }
        if (h.done != GM_CYCLE) {
            converged = true;
            HIPCHK(c, hipGetLastError());
            return ISPH_OK;
        }
        k_gm_resid<TI><<<grid, BLOCK, 0, c->stream>>>(A, c->rhs, x, d->u, d->part);
        k_gm_start<<<1, FIN_BLOCK, 0, c->stream>>>(grid, d->part, d->g, d->st, c->prm.tol, 0);
    }
}

int solve_gmres(isph_ctx* c, double* x, bool& converged) {
    converged = false;
    const int n = c->n;
    if (n == 0) {
        converged = true;
        c->iters = 0;
        c->relres = 0.0;
        return ISPH_OK;
    }
    const int want = c->prm.inner_maxit > 0 ? c->prm.inner_maxit : GM_DEFAULT_M;
    if (want > GM_MAXM) return fail(c, ISPH_E_ARG, "GMRES restart length (inner_maxit) above 512");
    if (c->prm.maxit <= 0) return fail(c, ISPH_E_ARG, "GMRES needs maxit > 0");
    const int m = std::min(want, n);
    int rc = gm_alloc(c, std::max(m, std::min(GM_DEFAULT_M, (int)c->cap)));
    if (rc) return rc;
    HIPCHK(c, hipMemsetAsync(c->gm->st, 0, sizeof(GmState), c->stream));
    GmMat<int> A{n, (size_t)c->cap, c->nbr, c->cnt, c->dpos, c->val64};
    return gm_run<int>(c, A, x, m, converged);
}

}  // namespace host
}  // namespace isph
