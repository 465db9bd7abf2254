// isph_kernels.hip -- gfx950 kernels of the ISPH step.
//
// One thread per particle / PPE row, 256-thread blocks (4 wave64).  Particle
// state is structure-of-arrays in HBM, positions and velocities as double2
// so that a neighbour gather is one 16-byte load.  Neighbour lists and the
// PPE matrix are column-major ELL with leading dimension `ld` (the particle
// capacity): entry k of row i sits at k*ld + i, so the 64 lanes of a wave
// read 64 consecutive entries.  Entries of a row ascend in particle index
// (the reference's matid order); the PPE diagonal is 1 after Jacobi
// scaling and is not stored, only its position `dpos` in the row.
#include "isph_kernels.h"

namespace isph {

__device__ __forceinline__ void raise_err(uint32_t* err, uint32_t bit) { atomicOr(err, bit); }

// ---------------------------------------------------------------------------
// cell binning: stable sort by bin key (cellshift/cellshiftalt re-binning,
// integrator.f90:32-391, + resetid, isph.f90:88-113).  k_bin_keys computes
// each particle's key and the bin counts (their exclusive scan is bstart);
// an LSD radix sort of (key, index) pairs gives the permutation, ties in the
// previous order.  With the reference order a fluid particle moves at most
// one cell from its previous bin (prev_key; null on the first binning), as
// cellshift only hands a particle to the adjacent cell it crossed into
// (integrator.f90:271-310).
// ---------------------------------------------------------------------------
__global__ void k_bin_keys(int n, GridInfo g, const double2* __restrict__ pos,
                           const int* __restrict__ tid, const int* __restrict__ flags, int sink,
                           double xlcut, double ytcut, const int* __restrict__ prev_key,
                           int* __restrict__ key, int* __restrict__ bincnt, uint32_t* err) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    double2 p = pos[i];
    int fl = tid[i] == 3;
    int k;
    if (sink && fl && !(flags[i] & ISPH_FLAG_BUFFER) && p.x <= xlcut && p.y > ytcut) {
        k = (int)g.nbins;
    } else {
        bool oob = false;
        const int c = cell_of(g, p.x, p.y, oob);
        if (oob) raise_err(err, ERR_DOMAIN);
        if (!g.order) {
            k = 2 * c + fl;
        } else {
            int cx = c / g.ny, cy = c - cx * g.ny, cls = fl;
            if (fl && prev_key) {
                int ox, oy;
                key_cell(g, prev_key[i], ox, oy);
                const int dx = max(-1, min(1, cx - ox)), dy = max(-1, min(1, cy - oy));
                cx = ox + dx;
                cy = oy + dy;
                cls = move_class(dx, dy);
            }
            k = 10 * cell_rank(g, cx, cy) + cls;
        }
    }
    key[i] = k;
    // one atomic per distinct key of the wave: consecutive particles of a
    // sorted state mostly share their bin, and atomics on one address
    // serialise
    unsigned long long todo = __ballot(1);
    const int lane = threadIdx.x & 63;
    while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const int lk = __shfl(k, leader, 64);
        const unsigned long long same = __ballot(k == lk);
        if (lane == leader) atomicAdd(&bincnt[lk], (int)__popcll(same));
        todo &= ~same;
    }
}

// exclusive scan of int32 in place: per-block scan, scan of block sums, add
__global__ void k_scan_blocks(int n, int* __restrict__ data, int* __restrict__ bsum) {
    __shared__ int s[SCAN_BLOCK];
    int i = blockIdx.x * SCAN_BLOCK + threadIdx.x;
    int v = i < n ? data[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_BLOCK; off <<= 1) {
        int t = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    if (i < n) data[i] = s[threadIdx.x] - v;
    if (threadIdx.x == SCAN_BLOCK - 1) bsum[blockIdx.x] = s[threadIdx.x];
}

__global__ void k_scan_sums(int nb, int* __restrict__ bsum) {
    __shared__ int s[SCAN_BLOCK];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += SCAN_BLOCK) {
        int i = base + threadIdx.x;
        int v = i < nb ? bsum[i] : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < SCAN_BLOCK; off <<= 1) {
            int t = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) bsum[i] = carry + s[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == SCAN_BLOCK - 1) carry += s[threadIdx.x];
        __syncthreads();
    }
}

__global__ void k_scan_add(int n, int* __restrict__ data, const int* __restrict__ bsum) {
    int i = blockIdx.x * SCAN_BLOCK + threadIdx.x;
    if (i < n) data[i] += bsum[blockIdx.x];
}

// LSD radix sort, one RS_BITS-bit digit per pass.  A tile of RS_TILE keys
// is RS_IPT rounds of BLOCK consecutive keys; per pass, k_rs_hist counts
// each tile's digits into hist[digit * ntiles + tile], whose exclusive scan
// is the output offset of every (digit, tile), and k_rs_scatter writes each
// key to its offset plus its rank among the equal digits before it in the
// tile.  Every pass is stable, so keys end sorted with ties in input order.
// Both kernels load all RS_IPT rounds of their tile before the first use, so
// a tile costs one memory latency, not one per round.

// lanes of the wave whose digit equals this lane's (among the `ok` lanes)
__device__ __forceinline__ unsigned long long digit_peers(bool ok, int d) {
    unsigned long long peers = __ballot(ok);
#pragma unroll
    for (int b = 0; b < RS_BITS; ++b) {
        const unsigned long long m = __ballot((d >> b) & 1);
        peers &= ((d >> b) & 1) ? m : ~m;
    }
    return peers;
}

static_assert(RS_DIGITS == BLOCK, "one digit per thread");

__global__ void __launch_bounds__(BLOCK) k_rs_hist(int n, const int* __restrict__ keys, int shift,
                                                   int ntiles, int* __restrict__ hist) {
    __shared__ int h[RS_DIGITS];
    const int base = blockIdx.x * RS_TILE + threadIdx.x;
    int kk[RS_IPT];
#pragma unroll
    for (int e = 0; e < RS_IPT; ++e) {
        const int idx = base + e * BLOCK;
        kk[e] = idx < n ? keys[idx] : 0;
    }
    h[threadIdx.x] = 0;
    __syncthreads();
    const unsigned long long below = (1ull << (threadIdx.x & 63)) - 1;
#pragma unroll
    for (int e = 0; e < RS_IPT; ++e) {
        const bool ok = base + e * BLOCK < n;
        const int d = (kk[e] >> shift) & (RS_DIGITS - 1);
        const unsigned long long peers = digit_peers(ok, d);
        if (ok && !(peers & below)) atomicAdd(&h[d], (int)__popcll(peers));
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// vin null: the values are the input positions
__global__ void __launch_bounds__(BLOCK) k_rs_scatter(int n, const int* __restrict__ kin,
                                                      const int* __restrict__ vin, int shift,
                                                      int ntiles, const int* __restrict__ hist,
                                                      int* __restrict__ kout,
                                                      int* __restrict__ vout) {
    constexpr int NW = BLOCK / 64;
    __shared__ int base[RS_DIGITS];         // next offset of each digit in this tile
    __shared__ int wcnt[2][NW][RS_DIGITS];  // per round (double-buffered): count per wave
    const int wv = threadIdx.x >> 6;
    const unsigned long long lower = (1ull << (threadIdx.x & 63)) - 1;
    const int t0 = blockIdx.x * RS_TILE + threadIdx.x;
    int kk[RS_IPT], vv[RS_IPT];
#pragma unroll
    for (int e = 0; e < RS_IPT; ++e) {
        const int idx = t0 + e * BLOCK;
        kk[e] = idx < n ? kin[idx] : 0;
        vv[e] = idx < n ? (vin ? vin[idx] : idx) : 0;
    }
    base[threadIdx.x] = hist[(size_t)threadIdx.x * ntiles + blockIdx.x];
#pragma unroll
    for (int e = 0; e < RS_IPT; ++e) {
        int(*wc)[RS_DIGITS] = wcnt[e & 1];
        const bool ok = t0 + e * BLOCK < n;
        const int k = kk[e];
        const int d = (k >> shift) & (RS_DIGITS - 1);
        for (int q = 0; q < NW; ++q) wc[q][threadIdx.x] = 0;
        __syncthreads();  // zeroed counts; the previous round's base update done
        const unsigned long long peers = digit_peers(ok, d);
        const int rank = (int)__popcll(peers & lower);
        if (ok && rank == 0) wc[wv][d] = (int)__popcll(peers);
        __syncthreads();
        if (ok) {
            int pos = base[d] + rank;
            for (int q = 0; q < wv; ++q) pos += wc[q][d];
            kout[pos] = k;
            vout[pos] = vv[e];
        }
        __syncthreads();  // every read of base done before its update
        int tot = 0;
        for (int q = 0; q < NW; ++q) tot += wc[q][threadIdx.x];
        base[threadIdx.x] += tot;
        // wc is zeroed again two rounds on, after the next round's first
        // barrier has ordered this update's reads
    }
}

// dst[i] = src[perm[i]]; key_sorted[i] = key[perm[i]] unless key_sorted is
// null (the radix sort leaves the sorted keys in place)
__global__ void k_gather_state(int n, const int* __restrict__ perm, StateView src, StateView dst,
                               const int* __restrict__ key, int* __restrict__ key_sorted) {
    int i = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (i >= n) return;
    int o = perm[i];
    dst.pos[i] = src.pos[o];
    dst.vel[i] = src.vel[o];
    dst.sp[i] = src.sp[o];
    dst.sv[i] = src.sv[o];
    dst.mass[i] = src.mass[o];
    dst.rho[i] = src.rho[o];
    dst.oden[i] = src.oden[o];
    dst.p[i] = src.p[o];
    dst.por[i] = src.por[o];
    dst.R[i] = src.R[o];
    dst.nut[i] = src.nut[o];
    dst.tke[i] = src.tke[o];
    dst.varts[i] = src.varts[o];
    dst.velm[i] = src.velm[o];
    dst.con[i] = src.con[o];
    dst.ovol[i] = src.ovol[o];
    dst.tid[i] = src.tid[o];
    dst.flags[i] = src.flags[o];
    dst.pid[i] = src.pid[o];
    if (key_sorted) key_sorted[i] = key[o];
}

// ---------------------------------------------------------------------------
// projection, isph.f90:16-47
// ---------------------------------------------------------------------------
__global__ void k_projection(int n, StateView s, const SimScalars* ss) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    if (s.tid[i] == 3 && !(s.flags[i] & ISPH_FLAG_DEAD)) {
        const double dt = ss->dt;
        double2 p = s.pos[i], v = s.vel[i];
        double n_ = s.por[i];
        s.sp[i] = p;
        p.x = p.x + (dt * v.x) / n_;
        p.y = p.y + (dt * v.y) / n_;
        s.pos[i] = p;
    }
}

// ---------------------------------------------------------------------------
// neighbour + nsearch2, search.f90:13-399: list, KGC matrix A (:77-85),
// divr (:272), L = inv(A) (:276-278), free flag (:302-304)
// ---------------------------------------------------------------------------
// Candidates of a particle in cell R (R = key / per, the cell's place in bin
// order) are the particles of cells R + dc*ny + {-1, 0, 1}, dc in {-1, 0, 1}:
// three contiguous index ranges, one per column.  A block's particles are
// consecutive in bin order, so for each dc the ranges of all its particles
// lie inside one window, [bstart[per*(R0+dc*ny-1)], bstart[per*(R1+dc*ny+2)))
// for the block's first and last cells R0, R1.  The positions of a window
// of up to NB_WIN particles are staged in LDS once for the block (2h cells
// hold about 7 fluid particles, so a block's window is about 280).
//
// The scan runs in two phases.  Phase 1 keeps a candidate when its squared
// distance d2 does not exceed twoh^2 (1 + 2^-30): every rejected candidate
// has RN(sqrt(d2)) >= twoh (d2 > twoh^2 (1 + 2^-31) gives sqrt(d2) >
// twoh (1 + 2^-33), and rounding is monotonic), so it fails the reference's
// test r < 2h exactly as before.  Phase 2 walks the kept candidates in scan
// order, applies the exact test to r = sqrt(d2) and accumulates A.  The lanes
// of a wave reach the expensive part (sqrt, kernel gradient, two divisions)
// together on every trip, where a one-phase scan pays it on nearly every
// candidate trip (about one candidate in three falls inside 2h: the circle
// covers 35% of the 3 x 3 cells of side 2h) because some lane of the wave is
// nearly always on a hit.  List order, list
// contents and the order of the sums are those of the one-phase scan.
__global__ void __launch_bounds__(BLOCK) k_neighbour(int n, GridInfo g, KernelConsts kc,
                                                     double thr, int maxnbr, size_t ld,
                                                     StateView s, const int* __restrict__ key,
                                                     const int* __restrict__ bstart,
                                                     int* __restrict__ nbr, int* __restrict__ cnt,
                                                     double* __restrict__ L,
                                                     double* __restrict__ divr, uint32_t* err) {
    // the first NB_STAGE entries of each list are kept in LDS and stored after
    // the scan, slot by slot, so that a wave's stores of slot k are one
    // coalesced row segment (stored as found, the lanes' k-th entries arrive
    // on different trips and each store is a scattered part of a line)
    constexpr int NB_STAGE = 24;
    constexpr int NB_WIN = 320;  // staged candidate positions per column offset
    __shared__ int snb[NB_STAGE][BLOCK];
    __shared__ double2 swin[3][NB_WIN];
    __shared__ int wlo[3], whi[3];
    const int i0 = xcd_tile(blockIdx.x, gridDim.x) * BLOCK;
    if (i0 >= n) return;  // whole block past the end: no barrier is pending
    const int i = i0 + threadIdx.x;
    if (threadIdx.x < 3) {
        const int dc = (int)threadIdx.x - 1;
        const int r0 = key[i0] / g.per, r1 = key[min(n, i0 + BLOCK) - 1] / g.per;
        wlo[threadIdx.x] = bstart[g.per * (r0 + dc * g.ny - 1)];
        whi[threadIdx.x] = bstart[g.per * (r1 + dc * g.ny + 2)];
    }
    __syncthreads();
    for (int w = 0; w < 3; ++w) {
        const int lo = wlo[w], len = whi[w] - lo;
        if (len <= NB_WIN)
            for (int k = threadIdx.x; k < len; k += BLOCK) swin[w][k] = s.pos[lo + k];
    }
    __syncthreads();
    if (i >= n) return;
    const int R = key[i] / g.per;
    const double2 pi = s.pos[i];
    const double t2 = (kc.twoh * kc.twoh) * (1.0 + 0x1p-30);
    int lo[3], hi[3];
    for (int w = 0; w < 3; ++w) {
        const int cr = R + (w - 1) * g.ny;
        lo[w] = bstart[g.per * (cr - 1)];
        hi[w] = bstart[g.per * (cr + 2)];
    }
    // phase 1: candidates within the widened radius, in scan order.  A range
    // outside its window (keys out of order, as after a buffer insertion
    // that no re-binning has followed yet) is read from global memory.
    int c1 = 0;
    for (int w = 0; w < 3; ++w) {
        const int wl = wlo[w];
        const bool staged = whi[w] - wl <= NB_WIN && lo[w] >= wl && hi[w] <= whi[w];
        for (int j = lo[w]; j < hi[w]; ++j) {
            if (j == i) continue;
            const double2 pj = staged ? swin[w][j - wl] : s.pos[j];
            const double dx = pi.x - pj.x, dy = pi.y - pj.y;
            if (dx * dx + dy * dy > t2) continue;
            if (c1 < NB_STAGE) snb[c1][threadIdx.x] = j;
            else if (c1 < maxnbr) nbr[(size_t)c1 * ld + i] = j;
            c1++;
        }
    }
    int c = 0;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int staged_n;
    if (c1 <= maxnbr) {
        // phase 2: exact test and the sums of A, compacting the list in
        // place (slot c <= slot k: no kept entry is overwritten unread)
        for (int k = 0; k < c1; ++k) {
            const int j = k < NB_STAGE ? snb[k][threadIdx.x] : nbr[(size_t)k * ld + i];
            const double2 pj = s.pos[j];
            double dx = pi.x - pj.x, dy = pi.y - pj.y;
            double r = sqrt(dx * dx + dy * dy);
            if (!(r < kc.twoh)) continue;
            if (c < NB_STAGE) snb[c][threadIdx.x] = j;
            else nbr[(size_t)c * ld + i] = j;
            c++;
            double wx, wy;
            kern_gradw(kc, r, dx, dy, wx, wy);
            const double mj = s.mass[j], rj = s.rho[j];
            double res1 = (mj * (pj.x - pi.x)) / rj;
            double res3 = (mj * (pj.y - pi.y)) / rj;
            a0 = a0 + wx * res1;
            a1 = a1 + wy * res1;
            a2 = a2 + wx * res3;
            a3 = a3 + wy * res3;
        }
        staged_n = c < NB_STAGE ? c : NB_STAGE;
    } else {
        // more candidates than list slots (phase 1 could not keep them all):
        // the one-phase scan, entries straight to the list, every hit in the
        // sums
        for (int w = 0; w < 3; ++w) {
            for (int j = lo[w]; j < hi[w]; ++j) {
                if (j == i) continue;
                const double2 pj = s.pos[j];
                double dx = pi.x - pj.x, dy = pi.y - pj.y;
                double r = sqrt(dx * dx + dy * dy);
                if (!(r < kc.twoh)) continue;
                if (c < maxnbr) nbr[(size_t)c * ld + i] = j;
                c++;
                double wx, wy;
                kern_gradw(kc, r, dx, dy, wx, wy);
                const double mj = s.mass[j], rj = s.rho[j];
                double res1 = (mj * (pj.x - pi.x)) / rj;
                double res3 = (mj * (pj.y - pi.y)) / rj;
                a0 = a0 + wx * res1;
                a1 = a1 + wy * res1;
                a2 = a2 + wx * res3;
                a3 = a3 + wy * res3;
            }
        }
        if (c > maxnbr) {
            raise_err(err, ERR_CAPACITY);
            c = maxnbr;
        }
        staged_n = 0;
    }
    for (int k = 0; k < staged_n; ++k) nbr[(size_t)k * ld + i] = snb[k][threadIdx.x];
    cnt[i] = c;
    int fl = s.flags[i] & ~ISPH_FLAG_FREE;
    double dv = 0.0;
    double l0 = 1.0, l1 = 0.0, l2 = 0.0, l3 = 1.0;
    if (s.tid[i] == 3) {
        fl &= ~ISPH_FLAG_VICINITY;
        dv = a0 + a3;
        l0 = a0; l1 = a1; l2 = a2; l3 = a3;
        invert2(l0, l1, l2, l3);
        if (dv <= thr && !(s.flags[i] & ISPH_FLAG_BUFFER)) fl |= ISPH_FLAG_FREE;
    }
    divr[i] = dv;
    double2* L2 = reinterpret_cast<double2*>(L);
    L2[2 * (size_t)i] = make_double2(l0, l1);
    L2[2 * (size_t)i + 1] = make_double2(l2, l3);
    s.flags[i] = fl;
}

// ---------------------------------------------------------------------------
// effpor, porous.f90:107-224 (+ porsearch :32-105)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double soil_cells(const KernelConsts& kc, double cpor, double2 pi,
                                             int lo, int hi, const double2* __restrict__ spos,
                                             const double* __restrict__ sm,
                                             const double* __restrict__ sr, double sf) {
    for (int q = lo; q < hi; ++q) {
        const double2 ps = spos[q];
        double dx = pi.x - ps.x, dy = pi.y - ps.y;
        double r = sqrt(dx * dx + dy * dy);
        if (r < kc.twoh) sf = sf + ((cpor * kern_w(kc, r)) * sm[q]) / sr[q];
    }
    return sf;
}

__global__ void k_effpor(int n, GridInfo g, KernelConsts kc, PorousParams pp, StateView s,
                         const int* __restrict__ key, const double2* __restrict__ spos,
                         const double* __restrict__ sm, const double* __restrict__ sr,
                         const int* __restrict__ sstart, uint32_t* err) {
    int i = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (i >= n) return;
    s.R[i] = 0.0;
    if (s.tid[i] != 3) return;
    const double2 pi = s.pos[i];
    bool oob = false;
    const int cell = cell_of(g, pi.x, pi.y, oob);
    if (oob) raise_err(err, ERR_DOMAIN);
    double sf = 0.0;
    const double cpor = 1.0 - pp.por * 1.0;
    if (!g.order) {
        const int cx = cell / g.ny, cy = cell - cx * g.ny;
        for (int dc = -1; dc <= 1; ++dc) {
            const int col = (cx + dc) * g.ny + cy;
            sf = soil_cells(kc, cpor, pi, sstart[col - 1], sstart[col + 2], spos, sm, sr, sf);
        }
    } else {
        // porous.f90:127-176: the particle's own cell (its bin), then three
        // cells of the 2x2 block on the side of the cell's midlines it lies on
        int cx, cy;
        key_cell(g, key[i], cx, cy);
        const double xl = g.xmin + (double)(cx - 1) * g.cs, xr = g.xmin + (double)cx * g.cs;
        const double yb = g.ymin + (double)(cy - 1) * g.cs, yt = g.ymin + (double)cy * g.cs;
        const bool left = pi.x <= xl + fabs((xl - xr) / 2.0);
        const bool low = pi.y <= yb + fabs((yt - yb) / 2.0);
        int ox[4], oy[4];
        ox[0] = 0; oy[0] = 0;
        if (left && low) { ox[1] = -1; oy[1] = 0; ox[2] = 0; oy[2] = -1; ox[3] = -1; oy[3] = -1; }
        else if (!left && low) { ox[1] = 1; oy[1] = 0; ox[2] = 0; oy[2] = -1; ox[3] = 1; oy[3] = -1; }
        else if (!left) { ox[1] = 1; oy[1] = 1; ox[2] = 0; oy[2] = 1; ox[3] = 1; oy[3] = 0; }
        else { ox[1] = 0; oy[1] = 1; ox[2] = -1; oy[2] = 1; ox[3] = -1; oy[3] = 0; }
        for (int q = 0; q < 4; ++q) {
            const int c = (cx + ox[q]) * g.ny + (cy + oy[q]);
            sf = soil_cells(kc, cpor, pi, sstart[c], sstart[c + 1], spos, sm, sr, sf);
        }
    }
    if (sf > (1.0 - pp.por)) sf = 1.0 - pp.por;
    const double nn = 1.0 - sf;
    s.por[i] = nn;
    const double oden = s.oden[i];
    s.rho[i] = nn * oden;
    const double n3 = (nn * nn) * nn;
    const double2 v = s.vel[i];
    const double vm = sqrt(v.x * v.x + v.y * v.y);
    const double omn = 1.0 - nn;
    double R;
    if (pp.drag_law == ISPH_DRAG_REFERENCE) {
        double fch = 1.750 / sqrt(150.0 * n3);
        double t1 = (((visc(pp.visc_law, pp.nu, oden) / oden) * pp.alpha) * (omn * omn)) /
                    (n3 * (pp.d50 * pp.d50));
        double t2 = (((fch * vm) * sqrt(pp.alpha)) * omn) / (sqrt(n3) * pp.d50);
        R = (-t1) - t2;
    } else {
        double t1 = ((pp.alpha * pp.nu) * (omn * omn)) / (n3 * (pp.d50 * pp.d50));
        double t2 = ((pp.beta * vm) * omn) / (n3 * pp.d50);
        R = (-t1) - t2;
    }
    s.R[i] = R;
}

__device__ __forceinline__ void load_L(const double* L, int i, double& l0, double& l1, double& l2,
                                       double& l3) {
    const double2* L2 = reinterpret_cast<const double2*>(L);
    double2 a = L2[2 * (size_t)i], b = L2[2 * (size_t)i + 1];
    l0 = a.x; l1 = a.y; l2 = b.x; l3 = b.y;
}

// ---------------------------------------------------------------------------
// eddyvis, turbulence.f90:12-110
// ---------------------------------------------------------------------------
__global__ void k_eddyvis(int n, KernelConsts kc, double dl, size_t ld, StateView s,
                          const int* __restrict__ nbr, const int* __restrict__ cnt,
                          const double* __restrict__ L) {
    int i = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (i >= n || s.tid[i] != 3) return;
    double l0, l1, l2, l3;
    load_L(L, i, l0, l1, l2, l3);
    const double2 pi = s.pos[i], vi = s.vel[i];
    const double ni = s.por[i];
    const double ui = vi.x / ni, wi = vi.y / ni;
    double t1 = 0.0, t2 = 0.0, t3 = 0.0;
    const int c = cnt[i];
    for (int k = 0; k < c; ++k) {
        const int j = nbr[(size_t)k * ld + i];
        double r, wx, wy;
        pair_geom(kc, pi, s.pos[j], r, wx, wy);
        double gx = l0 * wx + l1 * wy;
        double gy = l2 * wx + l3 * wy;
        const double2 vj = s.vel[j];
        const double nj = s.por[j], mj = s.mass[j], rj = s.rho[j];
        double du = vj.x / nj - ui;
        double dv = vj.y / nj - wi;
        t1 = t1 + ((mj * du) * gx) / rj;
        t2 = t2 + ((mj * dv) * gy) / rj;
        t3 = t3 + ((0.50 * mj) * (du * gy + dv * gx)) / rj;
    }
    const double cs = 0.150, cv = 0.080, ce = 1.0;
    double t4 = sqrt(2.0 * (t1 * t1 + t2 * t2) + (2.0 * t3) * (2.0 * t3));
    s.nut[i] = ((cs * dl) * (cs * dl)) * t4;
    s.tke[i] = (cv / ce) * ((dl * t4) * (dl * t4));
}

// ---------------------------------------------------------------------------
// int_vel, integrator.f90:458-713: (i) varts of free particles (:467-506)
// ---------------------------------------------------------------------------
__global__ void k_varts(int n, KernelConsts kc, double dl, size_t ld, StateView s,
                        const int* __restrict__ nbr, const int* __restrict__ cnt,
                        const double* __restrict__ L, const SimScalars* ss) {
    int i = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (i >= n || !(s.flags[i] & ISPH_FLAG_FREE)) return;
    double l0, l1, l2, l3;
    load_L(L, i, l0, l1, l2, l3);
    const double2 pi = s.pos[i], vi = s.vel[i];
    const double ni = s.por[i];
    const double ui = vi.x / ni, wi = vi.y / ni;
    double t1 = 0.0, t2 = 0.0;
    const int c = cnt[i];
    for (int k = 0; k < c; ++k) {
        const int j = nbr[(size_t)k * ld + i];
        double r, wx, wy;
        pair_geom(kc, pi, s.pos[j], r, wx, wy);
        double gx = l0 * wx + l1 * wy;
        double gy = l2 * wx + l3 * wy;
        const double2 vj = s.vel[j];
        const double nj = s.por[j];
        const double vol = s.mass[j] / s.rho[j];
        t1 = t1 + ((vj.x / nj - ui) * gx) * vol;
        t2 = t2 + ((vj.y / nj - wi) * gy) * vol;
    }
    const double ps = 1.0;
    s.varts[i] = (((ps * (dl * dl)) * sqrt(t1 * t1 + t2 * t2)) * sqrt(vi.x * vi.x + vi.y * vi.y)) /
                 ((ss->umax + TINY6) * ni);
}

// (ii) v* of fluid particles (:509-698)
__global__ void __launch_bounds__(BLOCK) k_vstar(int n, KernelConsts kc, PorousParams pp,
                                                 double lam, double g, size_t ld, StateView s,
                                                 const int* __restrict__ nbr,
                                                 const int* __restrict__ cnt,
                                                 const double* __restrict__ L,
                                                 const SimScalars* ss) {
    int i = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (i >= n) return;
    const int fi = s.flags[i];
    const double dt = ss->dt;
    if (s.tid[i] == 3 && !(fi & ISPH_FLAG_BUFFER) && !(fi & ISPH_FLAG_DEAD)) {
        double l0, l1, l2, l3;
        load_L(L, i, l0, l1, l2, l3);
        const double2 pi = s.pos[i], vi = s.vel[i];
        const double ni = s.por[i], oi = s.oden[i];
        const double mui = visc(pp.visc_law, pp.nu, oi);
        const double nuti = s.nut[i], tkei = s.tke[i];
        const double w0 = kern_w0(kc);
        double sx = 0.0, sy = 0.0;
        const int c = cnt[i];
        for (int k = 0; k < c; ++k) {
            const int j = nbr[(size_t)k * ld + i];
            const double2 pj = s.pos[j];
            double r, wx, wy;
            pair_geom(kc, pi, pj, r, wx, wy);
            double gx = l0 * wx + l1 * wy;
            double gy = l2 * wx + l3 * wy;
            const int tj = s.tid[j];
            double cmu;
            if (tj == 3) {
                const double oj = s.oden[j];
                const double muj = visc(pp.visc_law, pp.nu, oj);
                cmu = (((2.0 * mui) * muj) * ((1.0 / oi) * (1.0 / oj))) / ((mui / oi) + (muj / oj));
            } else {
                cmu = mui * (1.0 / oi);
            }
            double vart;
            if (fi & ISPH_FLAG_FREE) vart = s.varts[i];
            else if ((s.flags[j] & ISPH_FLAG_FREE) && !(fi & ISPH_FLAG_BUFFER))
                vart = (kern_w(kc, r) * s.varts[j]) / w0;
            else vart = 0.0;
            const double pvol = s.mass[j] / s.rho[j];
            const double pdist = 1.0 / (r * r + lam);
            const double dxp = pi.x - pj.x, dyp = pi.y - pj.y;
            const double2 vj = s.vel[j];
            double dvx = vi.x - vj.x;
            double t1 = ((((2.0 * pvol) * dvx) * dxp) * gx) * pdist;
            double t2 = ((((2.0 * pvol) * dvx) * dyp) * gy) * pdist;
            sx = sx + (t1 + t2) * (cmu + vart);
            double nutj = 0.0;
            if (tj > 2) {
                nutj = s.nut[j];
                if (nuti < TINY6 && nutj < TINY6) sx = sx + ((t1 + t2) * (nuti + nutj)) * 0.50;
                else sx = sx + (t1 + t2) * (((2.0 * nuti) * nutj) / (nuti + nutj));
            }
            double dvy = vi.y - vj.y;
            t1 = ((((2.0 * pvol) * dvy) * dxp) * gx) * pdist;
            t2 = ((((2.0 * pvol) * dvy) * dyp) * gy) * pdist;
            sy = sy + (t1 + t2) * (cmu + vart);
            if (tj > 2) {
                if (nuti < TINY6 && nutj < TINY6) sy = sy + ((t1 + t2) * (nuti + nutj)) * 0.50;
                else sy = sy + (t1 + t2) * (((2.0 * nuti) * nutj) / (nuti + nutj));
            }
            const double dtk = s.tke[j] - tkei;
            t1 = (pvol * dtk) * gx;
            sx = sx - ((t1 * ni) * 2.0) / 3.0;
            t1 = (pvol * dtk) * gy;
            sy = sy - ((t1 * ni) * 2.0) / 3.0;
        }
        const double den = 1.0 - (dt * s.R[i]) * ni;
        double vxs = (vi.x + dt * sx) / den;
        double vys = (vi.y + dt * sy) / den;
        vys = vys + (g * ni) * dt;
        s.sv[i] = make_double2(vxs, vys);
    }
    if (fi & ISPH_FLAG_BUFFER) s.sv[i] = make_double2(s.vel[i].x, (g * s.por[i]) * dt);
}

// ---------------------------------------------------------------------------
// PPE assembly, isph.f90:123-456 (row rules :145-421, diagonal position
// :167-176,425-426, Jacobi scaling :434-442)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BLOCK) k_assemble(int n, KernelConsts kc, double lam,
                                                    double lampd, size_t ld, StateView s,
                                                    const int* __restrict__ nbr,
                                                    const int* __restrict__ cnt,
                                                    const double* __restrict__ L,
                                                    double* __restrict__ val,
                                                    float* __restrict__ val32,
                                                    double* __restrict__ rsum,
                                                    int* __restrict__ dpos, double* __restrict__ rhs,
                                                    const SimScalars* ss, uint32_t* err) {
    int i = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (i >= n) return;
    const double dt = ss->dt;
    const int ti = s.tid[i];
    const bool freei = (s.flags[i] & ISPH_FLAG_FREE) != 0;
    double li0, li1, li2, li3;
    load_L(L, i, li0, li1, li2, li3);
    const double2 pi = s.pos[i], svi = s.sv[i];
    const double ni = s.por[i], rhoi = s.rho[i], mi = s.mass[i];
    const double rho_i = rhoi * (1.0 / ni);
    const int c = cnt[i];
    int dp = 0;
    double dia = 0.0, b = 0.0;
    for (int k = 0; k < c; ++k) {
        const int j = nbr[(size_t)k * ld + i];
        if (j < i) dp = k + 1;
        double v = 0.0;
        const int tj = s.tid[j];
        const double lamp = (freei && tj == 3) ? (1.0 + s.oden[j] / lampd) : 1.0;
        const double nj = s.por[j], rhoj = s.rho[j];
        const double rho_j = rhoj * (1.0 / nj);
        const double pvol = s.mass[j] * (1.0 / rhoj);
        const double2 pj = s.pos[j];
        double ddx = pj.x - pi.x, ddy = pj.y - pi.y;
        double rr = sqrt(ddx * ddx + ddy * ddy);
        double pdist = 1.0 / (rr * rr + lam);
        // pair_geom's r is rr: (pi - pj)^2 and (pj - pi)^2 round alike
        const double dxp = pi.x - pj.x, dyp = pi.y - pj.y;
        double wx, wy;
        kern_gradw(kc, rr, dxp, dyp, wx, wy);
        double t1, t2;
        if (ti == 3) {
            const double rho_ij = (((2.0 * rhoj) * rhoi) * (1.0 / (nj * ni))) *
                                  (1.0 / (rhoj * (1.0 / nj) + rhoi * (1.0 / ni)));
            double gx = li0 * wx + li1 * wy;
            double gy = li2 * wx + li3 * wy;
            t1 = (((2.0 * pvol) * gx) * dxp) * pdist;
            t2 = (((2.0 * pvol) * gy) * dyp) * pdist;
            v = -(t1 + t2);
            dia = dia + (t1 + t2) * lamp;
            const double2 svj = s.sv[j];
            if (tj == 3) {
                t1 = (pvol * ((svj.x * rho_ij) / nj - (svi.x * rho_ij) / ni)) * gx;
                t2 = (pvol * ((svj.y * rho_ij) / nj - (svi.y * rho_ij) / ni)) * gy;
            } else {
                t1 = (pvol * ((svj.x * rho_i) / ni - (svi.x * rho_i) / ni)) * gx;
                t2 = (pvol * ((svj.y * rho_i) / ni - (svi.y * rho_i) / ni)) * gy;
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
                double lj0, lj1, lj2, lj3;
                load_L(L, j, lj0, lj1, lj2, lj3);
                double gjx = lj0 * wx + lj1 * wy;
                double gjy = lj2 * wx + lj3 * wy;
                t1 = ((((2.0 * mi) * gjx) * dxp) * pdist) / rhoi;
                t2 = ((((2.0 * mi) * gjy) * dyp) * pdist) / rhoi;
                v = (-(t1 + t2)) * 1.0;
                dia = dia + (t1 + t2);
                const double2 svj = s.sv[j];
                t1 = ((mi * ((svj.x * rho_j) / nj - (svi.x * rho_j) / ni)) * gjx) / rhoi;
                t2 = ((mi * ((svj.y * rho_j) / nj - (svi.y * rho_j) / ni)) * gjy) / rhoi;
                b = b + (t1 + t2) / dt;
            } else if (tj <= 2) {
                t1 = (((2.0 * pvol) * wx) * dxp) * pdist;
                t2 = (((2.0 * pvol) * wy) * dyp) * pdist;
                v = -(t1 + t2);
                dia = dia + (t1 + t2);
            }
        }
        val[(size_t)k * ld + i] = v;
    }
    if (!(dia != 0.0) || !isfinite(dia)) raise_err(err, ERR_SINGULAR);
    double rs = 0.0;  // row sum of the scaled row, columns in order, diagonal 1
    for (int k = 0; k < c; ++k) {
        double v = val[(size_t)k * ld + i] / dia;
        val[(size_t)k * ld + i] = v;
        if (val32) val32[(size_t)k * ld + i] = (float)v;
        if (k == dp) rs = rs + 1.0;
        rs = rs + v;
    }
    if (dp == c) rs = rs + 1.0;
    if (rsum) rsum[i] = rs;
    dpos[i] = dp;
    rhs[i] = b / dia;
}

// ---------------------------------------------------------------------------
// comp_vel, integrator.f90:715-865
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BLOCK) k_comp_vel(int n, KernelConsts kc, size_t ld,
                                                    StateView s, const int* __restrict__ nbr,
                                                    const int* __restrict__ cnt,
                                                    const double* __restrict__ L,
                                                    const SimScalars* ss, uint32_t* err) {
    int i = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (i >= n) return;
    const int fi = s.flags[i];
    const double dt = ss->dt;
    if (s.tid[i] == 3 && !(fi & ISPH_FLAG_BUFFER) && !(fi & ISPH_FLAG_DEAD)) {
        double l0, l1, l2, l3;
        load_L(L, i, l0, l1, l2, l3);
        const double2 pi = s.pos[i];
        const double ni = s.por[i], rhoi = s.rho[i], p_i = s.p[i];
        double vx = 0.0, vy = 0.0;
        const int c = cnt[i];
        for (int k = 0; k < c; ++k) {
            const int j = nbr[(size_t)k * ld + i];
            double r, wx, wy;
            pair_geom(kc, pi, s.pos[j], r, wx, wy);
            double gx = l0 * wx + l1 * wy;
            double gy = l2 * wx + l3 * wy;
            const double vj = s.mass[j] / s.rho[j];
            double t1;
            if (s.tid[j] == 3) {
                const double nj = s.por[j], rhoj = s.rho[j], pj = s.p[j];
                double q = ((pj * rhoi) / ni) + ((p_i * rhoj) / nj);
                double den = rhoi * ((rhoi / ni) + (rhoj / nj));
                t1 = (2.0 * (((vj * (ni * ni)) * q) * gx)) / den;
                vx = vx - t1 * dt;
                t1 = (2.0 * (((vj * (ni * ni)) * q) * gy)) / den;
                vy = vy - t1 * dt;
            } else {
                double q = s.p[j] + p_i;
                t1 = (((vj * (ni * ni)) * q) * gx) / rhoi;
                vx = vx - t1 * dt;
                t1 = (((vj * (ni * ni)) * q) * gy) / rhoi;
                vy = vy - t1 * dt;
            }
        }
        const double2 sv = s.sv[i];
        vx = sv.x + vx;
        vy = sv.y + vy;
        s.vel[i] = make_double2(vx, vy);
        const double vm = sqrt(vx * vx + vy * vy) / ni;
        s.velm[i] = vm;
        if (!isfinite(vm)) raise_err(err, ERR_NAN);
    }
    if (fi & ISPH_FLAG_BUFFER) s.vel[i] = make_double2(s.sv[i].x, 0.0);
    if ((fi & ISPH_FLAG_DEAD) && !(fi & ISPH_FLAG_BUFFER)) {
        s.vel[i] = make_double2(0.0, 0.0);
        s.sv[i] = make_double2(0.0, 0.0);
    }
}

// comp_pos, integrator.f90:867-906 (buffer release :888-897 not modelled)
// comp_pos, integrator.f90:867-906: the midpoint position update, then the
// release of a buffer particle that has moved into (xrel_lo, xrel_hi)
// (:888-897): it stops being a buffer particle.
__global__ void k_comp_pos(int n, StateView s, const SimScalars* ss, double xrel_lo,
                           double xrel_hi) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const int f = s.flags[i];
    if (s.tid[i] == 3 && !(f & ISPH_FLAG_DEAD)) {
        const double dt = ss->dt;
        double2 p = s.pos[i], q = s.sp[i], v = s.vel[i];
        const double nn = s.por[i];
        p.x = ((p.x + q.x) + (dt * v.x) / nn) / 2.0;
        p.y = ((p.y + q.y) + (dt * v.y) / nn) / 2.0;
        s.pos[i] = p;
        if ((f & ISPH_FLAG_BUFFER) && p.x > xrel_lo && p.x < xrel_hi)
            s.flags[i] = f & ~ISPH_FLAG_BUFFER;
    }
}

// freesurf, isph.f90:49-86, pull form over the symmetric lists
__global__ void k_freesurf(int n, double thr, size_t ld, StateView s, const int* __restrict__ nbr,
                           const int* __restrict__ cnt, const double* __restrict__ divr) {
    int i = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (i >= n) return;
    bool v = false;
    const int c = cnt[i];
    for (int k = 0; k < c; ++k) {
        const int j = nbr[(size_t)k * ld + i];
        if (s.tid[j] == 3 && !(s.flags[j] & ISPH_FLAG_BUFFER) && divr[j] <= thr) v = true;
    }
    int fl = s.flags[i];
    s.flags[i] = v ? (fl | ISPH_FLAG_VICINITY) : (fl & ~ISPH_FLAG_VICINITY);
}

// ---------------------------------------------------------------------------
// opt2_shift, part_shift.f90:12-192
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BLOCK) k_shift(int n, KernelConsts kc, double csh, double cap,
                                                 size_t ld, StateView s,
                                                 const int* __restrict__ nbr,
                                                 const int* __restrict__ cnt,
                                                 const double* __restrict__ L) {
    int i = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (i >= n) return;
    const int fi = s.flags[i];
    if (s.tid[i] != 3 || (fi & ISPH_FLAG_BUFFER)) return;
    const int c = cnt[i];
    if ((fi & ISPH_FLAG_DEAD) || c == 0) {
        s.sp[i] = make_double2(0.0, 0.0);
        s.sv[i] = make_double2(0.0, 0.0);
        return;
    }
    double l0, l1, l2, l3;
    load_L(L, i, l0, l1, l2, l3);
    const double2 pi = s.pos[i], vi = s.vel[i];
    const double frac = (fi & ISPH_FLAG_FREE) ? 0.01 : 1.0;
    const double cf = (csh * (kc.h * kc.h)) * frac;
    double xs = 0.0, ys = 0.0;
    double t1 = 0.0, t2 = 0.0, t3 = 0.0, t4 = 0.0;
    for (int k = 0; k < c; ++k) {
        const int j = nbr[(size_t)k * ld + i];
        double r, wx, wy;
        pair_geom(kc, pi, s.pos[j], r, wx, wy);
        const double vj = s.mass[j] / s.rho[j];
        xs = xs - (cf * wx) * vj;
        ys = ys - (cf * wy) * vj;
        double gx = l0 * wx + l1 * wy;
        double gy = l2 * wx + l3 * wy;
        const double2 vel = s.vel[j];
        double du = vel.x - vi.x, dv = vel.y - vi.y;
        t1 = t1 + (gx * du) * vj;
        t2 = t2 + (gy * du) * vj;
        t3 = t3 + (gx * dv) * vj;
        t4 = t4 + (gy * dv) * vj;
    }
    double axs = fabs(xs), ays = fabs(ys);
    double cxs = (xs == 0.0) ? 0.0 : ((axs < cap ? axs : cap) * xs) / axs;
    double cys = (ys == 0.0) ? 0.0 : ((ays < cap ? ays : cap) * ys) / ays;
    s.sp[i] = make_double2(cxs, cys);
    s.sv[i] = make_double2(t1 * xs + t2 * ys, t3 * xs + t4 * ys);
}

__global__ void k_shift_apply(int n, StateView s) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    if (s.tid[i] == 3 && !(s.flags[i] & ISPH_FLAG_BUFFER)) {
        double2 p = s.pos[i], q = s.sp[i], v = s.vel[i], w = s.sv[i];
        s.pos[i] = make_double2(p.x + q.x, p.y + q.y);
        s.vel[i] = make_double2(v.x + w.x, v.y + w.y);
    }
}

// ---------------------------------------------------------------------------
// timestep, integrator.f90:908-976: umax = max vel, numax = max |nut|
// (max is order-independent; non-negative doubles compare as uint64)
// ---------------------------------------------------------------------------
__global__ void k_maxes(int n, const double* __restrict__ velm, const double* __restrict__ nut,
                        double* __restrict__ part) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    double a = 0.0, b = 0.0;
    if (i < n) {
        a = fabs(velm[i]);
        b = fabs(nut[i]);
        if (!(a == a)) a = 0.0;
        if (!(b == b)) b = 0.0;
    }
    a = block_max(a);
    b = block_max(b);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = a;
        part[gridDim.x + blockIdx.x] = b;
    }
}

__global__ void k_timestep(int nb, const double* __restrict__ part, SimScalars* ss, double dl,
                           double s1, double mu, double rho, int adaptive, double tschmidt,
                           int salinity, uint32_t* err) {
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < nb; k += BLOCK) {
        a = fmax(a, part[k]);
        b = fmax(b, part[nb + k]);
    }
    a = block_max(a);
    b = block_max(b);
    if (threadIdx.x != 0) return;
    const double umax = a;
    const double numax = b;
    ss->umax = umax;
    ss->numax = numax;
    if (adaptive) {
        const double t = ss->t;
        double dt;
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
        ss->dt = dt;
    }
    // transport substeps (integrator.f90:961-972); Dm is the float32 literal 1e-9
    double nm = numax;
    if (!isfinite(nm)) nm = 0.0;
    ss->numax = nm;
    const double dt = ss->dt;
    double dts = (s1 * (dl * dl)) / (DM_SALT + nm / tschmidt);
    int steps = 1;
    if (dts < dt) {
        const double q = ceil(dt / dts);
        if (!(q <= (double)MAX_SUBSTEPS)) {
            steps = 0;
            if (salinity) raise_err(err, ERR_SUBSTEPS);
        } else {
            steps = (int)q;
            dts = dt / (double)steps;
        }
    }
    ss->dtsol = (dts < dt) ? dts : dt;
    ss->solsteps = steps;
}

__global__ void k_advance_time(SimScalars* ss) {
    if (threadIdx.x == 0 && blockIdx.x == 0) ss->t = ss->t + ss->dt;
}

// plain streaming copy (bandwidth and counter calibration): four 16-byte
// loads in flight per lane before their stores, then a one-deep tail
__global__ void k_copy(const double2* __restrict__ a, double2* __restrict__ b, size_t n) {
    size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * BLOCK;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const double2 v0 = a[i], v1 = a[i + stride], v2 = a[i + 2 * stride],
                      v3 = a[i + 3 * stride];
        b[i] = v0;
        b[i + stride] = v1;
        b[i + 2 * stride] = v2;
        b[i + 3 * stride] = v3;
    }
    for (; i < n; i += stride) b[i] = a[i];
}

}  // namespace isph
