// isph_buffers.hip -- inflow/outflow buffer zones (internal.f90:475-614) and
// their schedule (main.f90:103-121, 218-236).
//
// Removal keeps the particle order: a keep flag per particle, an exclusive
// scan of the flags and a gather of the state (the bin keys travel along),
// as remove_buffer1/2 compact each entry cell's list in place.  Insertion
// appends the zone's particles after the last one; the next re-binning puts
// them after the fluid already in their cell, where insert_buffer1/2 append
// them to the cell's list.  Their bin key is their entry cell's (the cell
// they are placed in), so the reference order classifies their next move
// from there.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "isph_host.h"
#include "isph_kernels.h"

namespace isph {

// remove_buffer1 keeps (.not. buffer) .or. (buffer .and. x < xcut)
// (internal.f90:555-556), remove_buffer2 keeps x > xcut (:591-592); only
// the particles listed in the entry cells are tested
__global__ void k_buffer_keep(int n, GridInfo g, StateView s, const int* __restrict__ key,
                              const unsigned char* __restrict__ cellmask, double xcut, int right,
                              int* __restrict__ keep) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    int cx, cy;
    key_cell(g, key[i], cx, cy);
    int k = 1;
    if (cellmask[cx * g.ny + cy] && (s.flags[i] & ISPH_FLAG_BUFFER)) {
        const double x = s.pos[i].x;
        k = right ? (x < xcut) : (x > xcut);
    }
    keep[i] = k;
}

__global__ void k_buffer_perm(int n, const int* __restrict__ keep, const int* __restrict__ at,
                              int* __restrict__ perm) {
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n && keep[i]) perm[at[i]] = i;
}

// insert_buffer1/2 (internal.f90:475-542): template particle -> fluid
__global__ void k_buffer_insert(int m, int n0, const double2* __restrict__ pos,
                                const double2* __restrict__ vel, const double* __restrict__ p,
                                const int* __restrict__ flags, const int* __restrict__ zkey,
                                double mass, double density, double ovol, double con, int pid0,
                                StateView s, int* __restrict__ key) {
    int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= m) return;
    const int i = n0 + k;
    const double2 v = vel[k];
    s.pos[i] = pos[k];
    s.vel[i] = v;
    s.sp[i] = make_double2(0.0, 0.0);
    s.sv[i] = make_double2(0.0, 0.0);
    s.mass[i] = mass;
    s.rho[i] = density;
    s.oden[i] = density;
    s.p[i] = p[k];
    s.por[i] = 1.0;
    s.R[i] = 0.0;
    s.nut[i] = 0.0;
    s.tke[i] = 0.0;
    s.varts[i] = 0.0;
    s.velm[i] = sqrt(v.x * v.x + v.y * v.y);
    s.con[i] = con;
    s.ovol[i] = ovol;
    s.tid[i] = 3;
    s.flags[i] = flags[k];
    s.pid[i] = pid0 + k;
    key[i] = zkey[k];
}

namespace host {

static void free_zone(BufferZone& z) {
    void* p[] = {z.pos, z.vel, z.p, z.flags, z.key, z.cellmask};
    for (auto q : p)
        if (q) (void)hipFree(q);
    z = BufferZone();
}

void free_zones(isph_ctx* c) {
    for (auto& z : c->zones) free_zone(z);
    c->nzones = 0;
}

static int remove_buffer(isph_ctx* c, int zi) {
    BufferZone& z = c->zones[zi];
    const int n = c->n;
    if (n == 0) return ISPH_OK;
    StateView& s = c->st[c->cur];
    // scratch outside a re-binning: keep in c->rank, its exclusive scan in
    // c->perm, the compaction permutation in c->key (all cap ints)
    int* keep = c->rank;
    int* at = c->perm;
    k_buffer_keep<<<nblk(n), BLOCK, 0, c->stream>>>(n, c->grid, s, c->key_s, z.cellmask, z.xcut,
                                                     z.remove_right, keep);
    HIPCHK(c, hipMemcpyAsync(at, keep, sizeof(int) * n, hipMemcpyDeviceToDevice, c->stream));
    int rc = scan(c, at, (int64_t)n);
    if (rc) return rc;
    int last[2] = {0, 0};
    HIPCHK(c, hipMemcpyAsync(&last[0], at + (n - 1), sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&last[1], keep + (n - 1), sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int nnew = last[0] + last[1];
    if (nnew == n) return ISPH_OK;
    int* perm = c->key;
    k_buffer_perm<<<nblk(n), BLOCK, 0, c->stream>>>(n, keep, at, perm);
    StateView& d = c->st[c->cur ^ 1];
    if (nnew > 0)
        k_gather_state<<<nblk(nnew), BLOCK, 0, c->stream>>>(nnew, perm, s, d, c->key_s, c->rank);
    HIPCHK(c, hipGetLastError());
    // the gathered keys are in c->rank; make them the current keys
    HIPCHK(c, hipMemcpyAsync(c->key_s, c->rank, sizeof(int) * std::max(nnew, 1),
                             hipMemcpyDeviceToDevice, c->stream));
    c->cur ^= 1;
    c->n = nnew;
    c->have_nbr = false;
    c->assembled = false;
    return ISPH_OK;
}

static int insert_buffer(isph_ctx* c, int zi) {
    BufferZone& z = c->zones[zi];
    if (z.nins == 0) return ISPH_OK;
    if ((int64_t)c->n + z.nins > c->cap)
        return fail(c, ISPH_E_CAPACITY, "buffer insertion exceeds the particle capacity");
    if (c->next_pid + z.nins > (int64_t)INT32_MAX + 1)
        return fail(c, ISPH_E_CAPACITY, "particle ids exhausted: inserted pids would exceed INT32_MAX");
    k_buffer_insert<<<nblk(z.nins), BLOCK, 0, c->stream>>>(
        z.nins, c->n, z.pos, z.vel, z.p, z.flags, z.key, z.mass, z.density, z.ovol, z.con,
        (int)c->next_pid, c->st[c->cur], c->key_s);
    HIPCHK(c, hipGetLastError());
    c->n += z.nins;
    c->next_pid += z.nins;
    c->have_nbr = false;
    c->assembled = false;
    return ISPH_OK;
}

// main.f90:103-121 / 218-236 for every installed zone, in zone order
int run_buffer_schedule(isph_ctx* c, double told, double tnew) {
    for (int k = 0; k < c->nzones; ++k) {
        BufferZone& z = c->zones[k];
        const double at = (double)z.counter * z.period;
        if ((told + c->zone_shift) < at && (tnew + c->zone_shift) >= at) {
            int rc = remove_buffer(c, k);
            if (rc) return rc;
            if ((rc = insert_buffer(c, k))) return rc;
            z.counter += 1;
        }
    }
    return ISPH_OK;
}

}  // namespace host
}  // namespace isph

using namespace isph;
using namespace isph::host;

extern "C" {

int isph_set_buffer_zones(isph_ctx* c, int32_t nz, const isph_buffer_zone* zs, double shift) {
    if (!c || nz < 0 || nz > ISPH_MAX_ZONES || (nz && !zs)) return ISPH_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_zones(c);
    const GridInfo& g = c->grid;
    const int64_t ncell = (int64_t)g.nx * g.ny;
    for (int k = 0; k < nz; ++k) {
        const isph_buffer_zone& q = zs[k];
        if (q.ntemplate < 0 || q.ncells < 0 || (q.ntemplate && (!q.x || !q.y || !q.vx || !q.vy ||
                                                                 !q.pressure || !q.flags)) ||
            (q.ncells && !q.cells) || !(q.period > 0.0)) {
            free_zones(c);
            return fail(c, ISPH_E_ARG, "invalid buffer zone");
        }
        std::vector<unsigned char> mask(ncell, 0);
        std::vector<int> order(ncell, -1);
        for (int e = 0; e < q.ncells; ++e) {
            const int cx = q.cells[2 * e], cy = q.cells[2 * e + 1];
            if (cx < 1 || cx > g.nx - 2 || cy < 1 || cy > g.ny - 2) {
                free_zones(c);
                return fail(c, ISPH_E_ARG, "buffer entry cell outside the cell grid");
            }
            mask[(int64_t)cx * g.ny + cy] = 1;
            if (order[(int64_t)cx * g.ny + cy] < 0) order[(int64_t)cx * g.ny + cy] = e;
        }
        // the template particles inside the entry cells, entry cell by entry
        // cell, each in template order (internal.f90:482-504)
        std::vector<std::pair<int, int>> sel;
        for (int t = 0; t < q.ntemplate; ++t) {
            const double fx = std::floor((q.x[t] - g.xmin) / g.cs), fy = std::floor((q.y[t] - g.ymin) / g.cs);
            if (!(fx >= 0.0) || !(fy >= 0.0) || fx > g.nx - 3 || fy > g.ny - 3) continue;
            const int64_t cell = ((int64_t)fx + 1) * g.ny + ((int64_t)fy + 1);
            if (mask[cell]) sel.push_back({order[cell], t});
        }
        std::stable_sort(sel.begin(), sel.end(),
                         [](const std::pair<int, int>& a, const std::pair<int, int>& b) {
                             return a.first < b.first;
                         });
        const int m = (int)sel.size();
        std::vector<double2> pos(m), vel(m);
        std::vector<double> p(m);
        std::vector<int> fl(m), key(m);
        for (int u = 0; u < m; ++u) {
            const int t = sel[u].second;
            pos[u] = make_double2(q.x[t], q.y[t]);
            vel[u] = make_double2(q.vx[t], q.vy[t]);
            p[u] = q.pressure[t];
            fl[u] = q.flags[t];
            const int cx = (int)std::floor((q.x[t] - g.xmin) / g.cs) + 1;
            const int cy = (int)std::floor((q.y[t] - g.ymin) / g.cs) + 1;
            key[u] = g.order ? 10 * cell_rank(g, cx, cy) + 1 : 2 * (cx * g.ny + cy) + 1;
        }
        BufferZone& z = c->zones[k];
        int rc = 0;
        rc |= dalloc(c, &z.pos, m);
        rc |= dalloc(c, &z.vel, m);
        rc |= dalloc(c, &z.p, m);
        rc |= dalloc(c, &z.flags, m);
        rc |= dalloc(c, &z.key, m);
        rc |= dalloc(c, &z.cellmask, ncell);
        if (rc) {
            free_zones(c);
            return fail(c, ISPH_E_HIP, "buffer zone allocation failed");
        }
        auto H2D = hipMemcpyHostToDevice;
        if (m) {
            HIPCHK(c, hipMemcpyAsync(z.pos, pos.data(), sizeof(double2) * m, H2D, c->stream));
            HIPCHK(c, hipMemcpyAsync(z.vel, vel.data(), sizeof(double2) * m, H2D, c->stream));
            HIPCHK(c, hipMemcpyAsync(z.p, p.data(), sizeof(double) * m, H2D, c->stream));
            HIPCHK(c, hipMemcpyAsync(z.flags, fl.data(), sizeof(int) * m, H2D, c->stream));
            HIPCHK(c, hipMemcpyAsync(z.key, key.data(), sizeof(int) * m, H2D, c->stream));
        }
        HIPCHK(c, hipMemcpyAsync(z.cellmask, mask.data(), ncell, H2D, c->stream));
        z.nins = m;
        z.mass = q.mass;
        z.density = q.density;
        z.ovol = q.ovol;
        z.con = q.con;
        z.period = q.period;
        z.xcut = q.xcut;
        z.remove_right = q.remove_right;
        z.counter = 1;
        c->nzones = k + 1;
    }
    c->zone_shift = shift;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ISPH_OK;
}

int isph_remove_buffer(isph_ctx* c, int32_t zone) {
    if (!c || zone < 0 || zone >= c->nzones) return ISPH_E_ARG;
    int rc = remove_buffer(c, zone);
    if (rc) return rc;
    return sync_check(c);
}

int isph_insert_buffer(isph_ctx* c, int32_t zone) {
    if (!c || zone < 0 || zone >= c->nzones) return ISPH_E_ARG;
    int rc = insert_buffer(c, zone);
    if (rc) return rc;
    return sync_check(c);
}

}  // extern "C"
