// isph_host.h -- host-side context of libisph.so, shared by the translation
// units of the library (forward step and C ABI: isph_capi.hip; discrete
// adjoint: isph_adjoint.hip).  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/isph_capi.h"
#include "isph_kernels.h"

namespace isph {
namespace host {

struct EventPair {
    hipEvent_t a, b;
};

// HIP events around every stride-th launch of the timed kernel slots: the
// four BiCGSTAB SpMV kernels (FP64 and mixed, p- and s-SpMV) and scalart
constexpr int PROF_SLOTS = 5;
constexpr int PROF_SCALART = 4;
struct Profiler {
    bool on = false;
    int stride = 1;        // record every stride-th launch of a slot
    int64_t seen[PROF_SLOTS] = {};
    bool cur[PROF_SLOTS] = {};
    std::vector<EventPair> pool[PROF_SLOTS];
    std::vector<char> real[PROF_SLOTS];      // per recorded launch: did real work
    size_t used[PROF_SLOTS] = {};
    int64_t launches[PROF_SLOTS] = {};
    double total_ms[PROF_SLOTS] = {};
    int batch_recorded[PROF_SLOTS] = {};
};

struct AdjointState;  // isph_adjoint.hip
struct MGData;        // isph_mg.hip
struct GmData;        // isph_gmres.hip

// a buffer zone as installed by isph_set_buffer_zones (isph_buffers.hip)
struct BufferZone {
    int nins = 0;                  // template particles that lie in the entry cells
    double2 *pos = nullptr, *vel = nullptr;
    double* p = nullptr;
    int *flags = nullptr, *key = nullptr;  // template flags and bin keys
    unsigned char* cellmask = nullptr;     // per cell: entry cell of the zone
    double mass = 0, density = 0, ovol = 0, con = 0, period = 0, xcut = 0;
    int remove_right = 0;
    long long counter = 1;         // the reference's iter1 / iter2
};

}  // namespace host
}  // namespace isph

#define ISPH_MAX_DEVICES 64  // GPUs per process the library keeps per-device addresses for

struct isph_ctx {
    isph_params prm;
    int32_t cap = 0, cap_soil = 0, n = 0, ns = 0;
    isph::GridInfo grid;
    hipStream_t stream = nullptr;
    isph::StateView st[2];
    int cur = 0;
    double* L = nullptr;
    double* divr = nullptr;
    int *key = nullptr, *key_s = nullptr, *rank = nullptr, *perm = nullptr;
    int *rs_hist = nullptr, *rs_val = nullptr;  // radix sort: tile digit counts, values
    int *bincnt = nullptr, *scanbuf = nullptr;
    int *nbr = nullptr, *cnt = nullptr;
    bool have_nbr = false;
    double *val64 = nullptr, *rhs = nullptr, *rsum = nullptr;
    float* val32 = nullptr;
    int* dpos = nullptr;
    short* nbr16 = nullptr;
    bool assembled = false;
    bool assembled32 = false;
    bool narrow_idx = false;
    double* v64[12] = {};
    float* v32[10] = {};
    double2 *rq64 = nullptr, *rv64 = nullptr;
    float2 *rq32 = nullptr, *rv32 = nullptr;
    double* part = nullptr;
    uint32_t* counters = nullptr;
    isph::SolverScalars* sc = nullptr;
    isph::SolverScalars* sc_host = nullptr;  // pinned
    isph::SimScalars* ss = nullptr;
    isph::SimScalars* ss_host = nullptr;  // pinned
    uint32_t* err = nullptr;
    double2* pos_nb = nullptr;     // positions at the last neighbour search
    double* cdiff = nullptr;       // salinity rate of the last scalart
    double4* salt_coef = nullptr;  // per particle (Dx, Dy, Dm + nut/ts, nut)
    double2* spos = nullptr;
    double *smass = nullptr, *srho = nullptr;
    int* sstart = nullptr;
    int iters = 0, outer = 0;
    double relres = 0.0;
    isph::host::Profiler prof;
    struct Slot {
        isph::StateView v{};
        int* key = nullptr;
        isph::SimScalars ss{};
        double* fsol = nullptr;
        int* bins = nullptr;  // bin starts of the last sort (the neighbour search reads them)
        int n = 0;
        bool used = false;
        // the buffer schedule: each zone's next event, the time shift and
        // the next inserted pid, so that steps after a restore insert and
        // remove as a fresh run from the same state does
        long long zone_counter[ISPH_MAX_ZONES] = {};
        int nzones = 0;
        double zone_shift = 0.0;
        int64_t next_pid = 0;
    } slots[64];
    isph::host::AdjointState* adj = nullptr;
    isph::host::MGData* mg = nullptr;
    isph::host::GmData* gm = nullptr;
    // the GMRES solve's warm start and result in row order, kept across
    // steps and re-binnings (fsol of initialize.f90:50, memory.f90:109-144)
    double* fsol = nullptr;
    isph::host::BufferZone zones[ISPH_MAX_ZONES];
    int nzones = 0;
    double zone_shift = 0.0;
    int64_t next_pid = 0;          // pid of the next inserted particle (int64: max pid + 1 may exceed int)
    double* probe_scratch = nullptr;  // this GPU's isph_probe_noalloc scratch (device address)
    isph_progress_fn progress = nullptr;  // isph_set_progress
    // the adjoint's recorded PPE solutions: run_solve copies its solution to
    // solve_record when set, and starts from solve_replay instead of the
    // particles' pressures when set (device arrays of n doubles, row order)
    double* solve_record = nullptr;
    const double* solve_replay = nullptr;
    void* progress_user = nullptr;
    std::string msg;
};

#define HIPCHK(ctx, expr)                                                             \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) {                                                       \
            if (ctx) (ctx)->msg = std::string(#expr) + ": " + hipGetErrorString(e_);  \
            return ISPH_E_HIP;                                                        \
        }                                                                             \
    } while (0)

namespace isph {
namespace host {

// blocks for n items, at least one: with no particles every per-particle
// kernel still launches (its threads skip i >= n) and every reduction sees
// one zero partial, so the stages keep their scalar bookkeeping
inline int nblk(int64_t n, int b = BLOCK) { return n > 0 ? (int)((n + b - 1) / b) : 1; }
inline int nblk8(int64_t n) { return ((nblk(n) + 7) / 8) * 8; }

int fail(isph_ctx* c, int code, const char* m);
int probe_scratch_attach(isph_ctx* c);
PorousParams porous_params(const isph_params& p);
KernelConsts kconsts(const isph_ctx* c);

template <class T>
int dalloc(isph_ctx* c, T** p, size_t count) {
    HIPCHK(c, hipMalloc((void**)p, sizeof(T) * (count ? count : 1)));
    HIPCHK(c, hipMemsetAsync(*p, 0, sizeof(T) * (count ? count : 1), c->stream));
    return ISPH_OK;
}
int alloc_state(isph_ctx* c, StateView& s, size_t cap);
void free_state(StateView& s);
int copy_state(isph_ctx* c, const StateView& a, const StateView& b, int n);
int take_errors(isph_ctx* c);
int sync_check(isph_ctx* c);
int read_sc(isph_ctx* c);
void prof_collect(isph_ctx* c);
void prof_record(isph_ctx* c, int k, bool start);

// stages of the step, enqueue only (main.f90:41-66 order)
int sort_particles(isph_ctx* c, bool with_sink, bool first);
int run_projection(isph_ctx* c);
int run_neighbour(isph_ctx* c);
int run_effpor(isph_ctx* c);
int run_eddyvis(isph_ctx* c);
int run_eddyvis_lagged(isph_ctx* c);  // pair geometry of the last search
int run_int_vel(isph_ctx* c);
int run_assemble(isph_ctx* c);
int run_solve(isph_ctx* c);
int run_comp_vel(isph_ctx* c);
int run_comp_pos(isph_ctx* c);
int run_freesurf(isph_ctx* c);
int run_shift(isph_ctx* c);
int run_timestep(isph_ctx* c);
int run_step(isph_ctx* c);

// FP64 BiCGSTAB on the current lists with the given off-diagonal values
// (unit diagonal at dpos) and right-hand side; x holds the initial guess
// on entry and the solution on return.
int solve_fp64_on(isph_ctx* c, const double* val, const double* rhs, double* x);

// solver pieces shared with the multigrid solve (isph_mg.hip)
__global__ void k_fill(int n, double* x, double v);
SolveCtl solve_ctl(const isph_ctx* c, int mixed);
ResidArgs resid_args(isph_ctx* c, double* x, double* r, const double* val = nullptr,
                     const double* rhs = nullptr);
int solve_mg(isph_ctx* c);
void mg_free(isph_ctx* c);

// a linear system for the multigrid solve: the FP64 and FP32 values and
// FP64 row sums of a Jacobi-scaled matrix on the current lists, the
// right-hand side and the solution (holding the initial guess)
struct MGSystem {
    const double* val64;
    const float* val32;
    const double* rsum;
    const double* rhs;
    double* x;
};
int solve_mg_on(isph_ctx* c, const MGSystem& sys);

// restarted GMRES, the reference's live solve (isph_gmres.hip); converged is
// false when the iteration limit was reached first
int solve_gmres(isph_ctx* c, double* x, bool& converged);
bool prec_mixed(int precision);  // the modes the assembly gives FP32 values (and row sums)
bool prec_mg(int precision);     // the multigrid-preconditioned modes
void gm_free(isph_ctx* c);

// discrete adjoint (isph_adjoint.hip)
void adj_free(isph_ctx* c);

// salinity transport (isph_scalar.hip)
int run_scalart(isph_ctx* c);
int run_scalarupdate(isph_ctx* c, double stime, bool from_ss);
int run_densityupdate(isph_ctx* c);
int run_salt_substeps(isph_ctx* c);

// buffer zones (isph_buffers.hip)
int run_buffer_schedule(isph_ctx* c, double told, double tnew);
void free_zones(isph_ctx* c);
int scan(isph_ctx* c, int* data, int64_t count);

}  // namespace host
}  // namespace isph
