// Host-side state of the solver library: one device context per process
// (device, stream, grid descriptor, grow-only workspaces, pinned state
// mirror, per-kernel event timing, tuned launch shapes).
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <utility>
#include <vector>

#include "../../include/mfix_leq.h"
#include "leq_kernels.h"

namespace mfix {

struct DevBuf {
    void *ptr = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes);
    void release();
};

// Event timing of one kernel family (0: stencil kernels, 1: line phases).
struct Prof {
    bool on = false;
    int64_t launches = 0;
    double total_ms = 0.0;
    double alg_bytes = 0.0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    std::vector<double> pending_bytes;
    size_t pool_used = 0;
    void begin(hipStream_t s);
    void end(hipStream_t s, double bytes);
    void collect();
};
constexpr int kNumProf = 2;
constexpr int kHostStates = 1024;
constexpr int kMaxPartials = 4;   // partial-sum arrays any reduction kernel writes (NV <= 4)
constexpr int kSizeClasses = 12;  // 8K, 32K, ..., 4^11 * 8K cells

// Launch shape of the streaming kernels: threads per block, cells per
// thread, and threads (= lines) per block of the line-solve kernels.
struct Shape {
    int bs = 0, cpt = 0, lines_per_block = 0;
};

int size_class(int64_t n3);

struct Context {
    bool initialised = false;
    bool have_grid = false;
    int device = 0;
    hipStream_t stream = nullptr;
    mfix_leq_grid grid = {};
    int iter_tot[10] = {0};
    DevBuf d_a, d_b, d_x;          // host-entry staging of A_M, B_M, VAR
    DevBuf ws64, ws32, a32, xalt;  // Krylov workspaces
    DevBuf line_scratch;           // LINE preconditioner line scratch (4 planes)
    DevBuf partials, state, misc;
    SolverState *host_state = nullptr;  // pinned mirror(s) of device states
    Prof prof[kNumProf];
    Shape tuned[2][kSizeClasses];  // [fp32][size class]; bs == 0 means untuned
    hipError_t last_hip_error = hipSuccess;
    char last_msg[256] = {0};
    const double *host_noise = nullptr;
    uint64_t noise_seed = 0;
    bool write_back_scaled = true;
    mfix_leq_stats last_stats = {};
    int init(int dev);
    int ensure_init();
    void release_all();
    Shape shape_for(int64_t n3, bool fp32) const;
};

Context &ctx();
int hip_fail(hipError_t e, const char *what);
int parse_pc(const char *pc);
DGrid make_dgrid(const mfix_leq_grid &h);
double alg_bytes_per_cell(int pc, int sa, int sv, int sx);

// Pointers of one solve's vectors (T = double or float).  x is VAR (FP64) or
// the FP32 correction accumulator (mixed); phat/shat exist only for LINE.
template <typename T>
struct SolveBufs {
    DGrid g;
    const T *a = nullptr;
    T *x = nullptr, *r = nullptr, *rt = nullptr, *p = nullptr, *phat = nullptr;
    T *s = nullptr, *shat = nullptr, *t = nullptr, *v = nullptr;
    T *line_scratch = nullptr;
    SolverState *st = nullptr;
    double *partials = nullptr;
    bool mixed = false;
};

template <int F, int NV>
void launch_finish(SolverState *st, const double *partials, int64_t nblk, hipStream_t s, double *out);
template <typename T>
int launch_msolve_line(const DGrid &g, const Shape &sh, const char *cmethod, const T *a, const T *r, T *z,
                       T *scratch, const SolverState *st, int skip_on_sexit, int *fatal_flag, hipStream_t s);
template <typename T>
void enqueue_iteration(const SolveBufs<T> &w, const Shape &sh, int pc, const char *cmethod, hipStream_t s);
int drive(const std::function<void()> &one_iteration, SolverState *d_state, SolverState *h_state, int itmax,
          hipStream_t s);
void fill_stats(const SolverState &h, mfix_leq_stats *st);
int solve_fp64(const DGrid &g, int pc, const char *cmethod, double tol, int itmax, double *x, double *a, double *b,
               const mfix_leq_opts *opts, mfix_leq_stats *stats, hipStream_t s, int minimize, int icheck);
int solve_mixed(const DGrid &g, int pc, const char *cmethod, double tol, int itmax, double *x, double *a, double *b,
                const mfix_leq_opts *opts, mfix_leq_stats *stats, hipStream_t s);
int run_matvec64(const DGrid &g, const double *a, const double *x, double *y, hipStream_t s);
int run_matvec32(const DGrid &g, const float *a, const float *x, float *y, hipStream_t s);
int run_diag(const DGrid &g, const double *a, const double *r, double *z, hipStream_t s);
int run_dot(const DGrid &g, const double *x, const double *y, double *d_out, hipStream_t s);
int run_scale(const DGrid &g, double *a, double *b, int *zero_row, hipStream_t s);
int run_copy(const double *src, double *dst, int64_t n, hipStream_t s);
int autotune(int imax, int jmax, int kmax, int mixed, Shape *best);

}  // namespace mfix
