// isph_output.hip -- the reference's text output and probes.
//
//   - Fortran edit descriptors Fw.d and ESw.d (isph_format_f/_es) with the
//     output of a Fortran run-time library: correctly rounded from the
//     binary value (ties to even), "-0.000" for negative values that round
//     to zero, the leading zero dropped when only it does not fit, w
//     asterisks when the field is too narrow, ESw.d exponents as E+dd up to
//     99 and +ddd (no E) from 100 to 999, Inf/-Inf/NaN right-justified;
//   - the snapshot writer of output.f90:188-234 (isph_combined): fluid,
//     buffer and dead fluid particles to three files, one record
//     (x, y, p: F10.3; vx, vy: ES10.3; con + 0.5: F10.5) per particle, in
//     the context's cell order;
//   - the Shepard pressure probes of probe.f90:86-130 (isph_probe), one
//     thread per probe summing over the 3x3 cells around the probe's cell;
//   - the probe time series file of output.f90:150-166 (F10.3, 1X, F10.3).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "isph_host.h"
#include "isph_kernels.h"

// Device memory of isph_probe_noalloc: the probe abscissae, ordinates and
// values of one call, 3 * ISPH_PROBE_NOALLOC_MAX doubles (96 KiB).  A
// file-scope array, so the library never allocates it: each GPU holds its
// own copy from load until the process ends.  In host code `scratch` only
// names the symbol; its device address comes from hipGetSymbolAddress.
static __device__ __align__(256) double scratch[3 * ISPH_PROBE_NOALLOC_MAX];

namespace isph {

// probevalue, probe.f90:86-130: fluid within 2 blen h (blen = 1,
// initialize.f90:34) of the probe, searched in the probe's cell and its 8
// neighbours in the order of probe.f90:92-94; Shepard-normalised pressure,
// 0 when no fluid particle is in range.
__global__ void k_probe(int np, const double* __restrict__ px, const double* __restrict__ py,
                        GridInfo g, KernelConsts kc, StateView s, const int* __restrict__ bstart,
                        double* __restrict__ out, uint32_t* err) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    bool oob = false;
    const int cell = cell_of(g, px[k], py[k], oob);
    if (oob) atomicOr(err, ERR_DOMAIN);
    const int cx = cell / g.ny, cy = cell - cx * g.ny;
    // (y, x) offsets: (i,j),(i,j-1),(i,j+1),(i+1,j),(i-1,j),(i+1,j+1),(i+1,j-1),(i-1,j-1),(i-1,j+1)
    const int oy[9] = {0, 0, 0, 1, -1, 1, 1, -1, -1};
    const int ox[9] = {0, -1, 1, 0, 0, 1, -1, -1, 1};
    double term1 = 0.0, pr = 0.0;
    int count = 0;
    for (int q = 0; q < 9; ++q) {
        int blo, bhi;
        column_bins(g, cx + ox[q], cy + oy[q], cy + oy[q], blo, bhi);
        const int lo = bstart[blo], hi = bstart[bhi];
        for (int j = lo; j < hi; ++j) {
            if (s.tid[j] <= 2) continue;
            const double2 pj = s.pos[j];
            const double dx = px[k] - pj.x, dy = py[k] - pj.y;
            const double r = sqrt(dx * dx + dy * dy);
            if (!(r < 2.0 * kc.h)) continue;
            const double w = kern_w(kc, r);
            const double m = s.mass[j], rho = s.rho[j];
            term1 = term1 + (m * w) / rho;
            pr = pr + ((s.p[j] * m) * w) / rho;
            ++count;
        }
    }
    out[k] = count ? pr / term1 : 0.0;
}

__global__ void k_probe_zero(int np, double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < np) out[k] = 0.0;
}

namespace host {

namespace {
std::mutex scratch_mu;
double* scratch_dev[ISPH_MAX_DEVICES] = {};  // device address of `scratch`, per GPU
}  // namespace

// Takes the device address of this GPU's `scratch` array (once per device,
// with that device current) and keeps it in the context.  Called by
// isph_create on the context's device.
int probe_scratch_attach(isph_ctx* c) {
    int dev = -1;
    HIPCHK(c, hipGetDevice(&dev));
    if (dev < 0 || dev >= ISPH_MAX_DEVICES) return fail(c, ISPH_E_HIP, "device index out of range");
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            c->msg = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return ISPH_E_HIP;
        }
        scratch_dev[dev] = static_cast<double*>(p);
    }
    c->probe_scratch = scratch_dev[dev];
    return ISPH_OK;
}

static void right_justify(const char* s, int len, int w, char* out) {
    if (len > w) {
        memset(out, '*', (size_t)w);
    } else {
        memset(out, ' ', (size_t)(w - len));
        memcpy(out + (w - len), s, (size_t)len);
    }
    out[w] = '\0';
}

static bool special(double v, int w, char* out) {
    if (std::isnan(v)) {
        right_justify("NaN", 3, w, out);
        return true;
    }
    if (std::isinf(v)) {
        if (v > 0) right_justify("Inf", 3, w, out);
        else right_justify("-Inf", 4, w, out);
        return true;
    }
    return false;
}

// Fw.d
void format_f(double v, int w, int d, char* out) {
    if (special(v, w, out)) return;
    char buf[400];
    int len = snprintf(buf, sizeof buf, "%.*f", d, v);
    const char* s = buf;
    if (len == w + 1 && buf[0] == '0') {  // "0.ddd" one too wide: ".ddd"
        s = buf + 1;
        len -= 1;
    } else if (len == w + 1 && buf[0] == '-' && buf[1] == '0') {  // "-0.ddd" -> "-.ddd"
        buf[1] = '-';
        s = buf + 1;
        len -= 1;
    }
    right_justify(s, len, w, out);
}

// ESw.d
void format_es(double v, int w, int d, char* out) {
    if (special(v, w, out)) return;
    char buf[64];
    snprintf(buf, sizeof buf, "%.*E", d, v);
    char* e = strchr(buf, 'E');
    const int ex = atoi(e + 1);
    const int ax = ex < 0 ? -ex : ex;
    char tail[16];
    if (ax <= 99) snprintf(tail, sizeof tail, "E%c%02d", ex < 0 ? '-' : '+', ax);
    else if (ax <= 999) snprintf(tail, sizeof tail, "%c%03d", ex < 0 ? '-' : '+', ax);
    else {
        right_justify("", w + 1, w, out);
        return;
    }
    *e = '\0';
    std::string s = std::string(buf) + tail;
    right_justify(s.c_str(), (int)s.size(), w, out);
}

// one snapshot record, output.f90:206-209 (without the newline)
static int snapshot_record(double x, double y, double p, double vx, double vy, double con,
                           char* line) {
    char f[6][24];
    format_f(x, 10, 3, f[0]);
    format_f(y, 10, 3, f[1]);
    format_f(p, 10, 3, f[2]);
    format_es(vx, 10, 3, f[3]);
    format_es(vy, 10, 3, f[4]);
    format_f(con + 0.50, 10, 5, f[5]);
    return snprintf(line, 80, "%s %s %s %s %s %s", f[0], f[1], f[2], f[3], f[4], f[5]);
}

}  // namespace host
}  // namespace isph

using namespace isph;
using namespace isph::host;

extern "C" {

int isph_format_f(double v, int32_t w, int32_t d, char* out) {
    if (!out || w <= 0 || w > 300 || d < 0 || d > 60) return ISPH_E_ARG;
    format_f(v, w, d, out);
    return ISPH_OK;
}

int isph_format_es(double v, int32_t w, int32_t d, char* out) {
    if (!out || w <= 0 || w > 300 || d < 0 || d > 40) return ISPH_E_ARG;
    format_es(v, w, d, out);
    return ISPH_OK;
}

int isph_snapshot_record(double x, double y, double p, double vx, double vy, double con,
                         char* out) {
    if (!out) return ISPH_E_ARG;
    snapshot_record(x, y, p, vx, vy, con, out);
    return ISPH_OK;
}

int isph_combined(isph_ctx* c, const char* dir, int32_t index) {
    if (!c || !dir) return ISPH_E_ARG;
    const int n = c->n;
    StateView& s = c->st[c->cur];
    std::vector<double2> pos(n), vel(n);
    std::vector<double> p(n), con(n);
    std::vector<int> tid(n), fl(n);
    auto D2H = hipMemcpyDeviceToHost;
    if (n) {
        HIPCHK(c, hipMemcpyAsync(pos.data(), s.pos, sizeof(double2) * n, D2H, c->stream));
        HIPCHK(c, hipMemcpyAsync(vel.data(), s.vel, sizeof(double2) * n, D2H, c->stream));
        HIPCHK(c, hipMemcpyAsync(p.data(), s.p, sizeof(double) * n, D2H, c->stream));
        HIPCHK(c, hipMemcpyAsync(con.data(), s.con, sizeof(double) * n, D2H, c->stream));
        HIPCHK(c, hipMemcpyAsync(tid.data(), s.tid, sizeof(int) * n, D2H, c->stream));
        HIPCHK(c, hipMemcpyAsync(fl.data(), s.flags, sizeof(int) * n, D2H, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const std::string d(dir);
    const std::string idx = std::to_string(index);
    const std::string names[3] = {d + "/fluid/fluid_" + idx + ".txt",
                                  d + "/buffer/buffer_" + idx + ".txt",
                                  d + "/fluid/dead_" + idx + ".txt"};
    FILE* f[3] = {nullptr, nullptr, nullptr};
    for (int k = 0; k < 3; ++k) {
        f[k] = fopen(names[k].c_str(), "w");
        if (!f[k]) {
            for (int q = 0; q < k; ++q) fclose(f[q]);
            return fail(c, ISPH_E_ARG, ("cannot open " + names[k]).c_str());
        }
    }
    char line[96];
    for (int i = 0; i < n; ++i) {
        if (tid[i] != 3) continue;
        const bool buf = fl[i] & ISPH_FLAG_BUFFER, dead = fl[i] & ISPH_FLAG_DEAD;
        int k;
        if (!buf && !dead) k = 0;
        else if (buf && !dead) k = 1;
        else if (dead && !buf) k = 2;
        else continue;
        const int len = snapshot_record(pos[i].x, pos[i].y, p[i], vel[i].x, vel[i].y, con[i], line);
        line[len] = '\n';
        fwrite(line, 1, (size_t)len + 1, f[k]);
    }
    bool bad = false;
    for (int k = 0; k < 3; ++k) bad |= fclose(f[k]) != 0;
    if (bad) return fail(c, ISPH_E_ARG, "snapshot write failed");
    return ISPH_OK;
}

int isph_probe(isph_ctx* c, int32_t np, const double* px, const double* py, double* out) {
    if (!c || np < 0 || (np && (!px || !py || !out))) return ISPH_E_ARG;
    if (np == 0) return ISPH_OK;
    double *d = nullptr;
    HIPCHK(c, hipMalloc((void**)&d, sizeof(double) * 3 * (size_t)np));
    int rc = ISPH_OK;
    if (hipMemcpyAsync(d, px, sizeof(double) * np, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(d + np, py, sizeof(double) * np, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = fail(c, ISPH_E_HIP, "probe upload failed");
    if (!rc && c->n) {
        k_probe<<<(np + 63) / 64, 64, 0, c->stream>>>(np, d, d + np, c->grid, kconsts(c),
                                                     c->st[c->cur], c->bincnt, d + 2 * np, c->err);
        if (hipGetLastError() != hipSuccess) rc = fail(c, ISPH_E_HIP, "probe launch failed");
    } else if (!rc) {
        (void)hipMemsetAsync(d + 2 * np, 0, sizeof(double) * np, c->stream);
    }
    if (!rc && hipMemcpyAsync(out, d + 2 * np, sizeof(double) * np, hipMemcpyDeviceToHost,
                              c->stream) != hipSuccess)
        rc = fail(c, ISPH_E_HIP, "probe download failed");
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (rc) return rc;
    return sync_check(c);
}

int isph_probe_noalloc(isph_ctx* c, int32_t np, const double* px, const double* py, double* out) {
    if (!c || np < 0 || (np && (!px || !py || !out))) return ISPH_E_ARG;
    if (np > ISPH_PROBE_NOALLOC_MAX)
        return fail(c, ISPH_E_ARG, "isph_probe_noalloc: np exceeds ISPH_PROBE_NOALLOC_MAX (4096)");
    if (!c->probe_scratch)
        return fail(c, ISPH_E_STATE, "isph_probe_noalloc: no scratch address for this context");
    if (np == 0) return ISPH_OK;
    // the GPU's scratch array (its device address, taken at isph_create):
    // x at [0, np), y at [np, 2np), the values at [2np, 3np)
    double* d = c->probe_scratch;
    HIPCHK(c, hipMemcpyAsync(d, px, sizeof(double) * np, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d + np, py, sizeof(double) * np, hipMemcpyHostToDevice, c->stream));
    if (c->n) {
        k_probe<<<(np + 63) / 64, 64, 0, c->stream>>>(np, d, d + np, c->grid, kconsts(c),
                                                     c->st[c->cur], c->bincnt, d + 2 * np, c->err);
    } else {
        // no particles: every value is 0 (written by a kernel, not a memset)
        k_probe_zero<<<(np + 63) / 64, 64, 0, c->stream>>>(np, d + 2 * np);
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(out, d + 2 * np, sizeof(double) * np, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return sync_check(c);
}

int isph_write_probe_series(const char* path, int32_t n, const double* a, const double* b) {
    if (!path || n < 0 || (n && (!a || !b))) return ISPH_E_ARG;
    FILE* f = fopen(path, "w");
    if (!f) return ISPH_E_ARG;
    char fa[24], fb[24];
    for (int i = 0; i < n; ++i) {
        format_f(a[i], 10, 3, fa);
        format_f(b[i], 10, 3, fb);
        fprintf(f, "%s %s\n", fa, fb);
    }
    return fclose(f) == 0 ? ISPH_OK : ISPH_E_ARG;
}

}  // extern "C"
