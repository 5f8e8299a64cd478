// Host-side state of the solver library: one device context per process
// (device, stream, grid descriptor, grow-only workspaces, pinned state
// mirror, per-kernel event timing, tuned launch shapes) and the system
// tables that every solve (single, batched, served) runs through.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <functional>
#include <utility>
#include <vector>

#include "../../include/mfix_leq.h"
#include "leq_phases.h"

namespace mfix {

struct DevBuf {
    void *ptr = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes);
    void release();
};

// Event timing of one kernel family (0: stencil kernels, 1: line phases).
// Launches after the solve was done return at once; collect(working) counts
// only the first `working` pending launches (in stream order) as work and
// the rest as idle, so the byte rate is over launches that did the work.
struct Prof {
    bool on = false;
    int64_t launches = 0;
    double total_ms = 0.0;
    double alg_bytes = 0.0;
    int64_t idle_launches = 0;
    double idle_ms = 0.0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    std::vector<double> pending_bytes;
    size_t pool_used = 0;
    // the event pair of the next timed launch, for launch_timed (nullptrs when off)
    std::pair<hipEvent_t, hipEvent_t> take(double bytes);
    void collect(int64_t working = -1);
};
constexpr int kNumProf = 2;
constexpr int kHostStates = 1024; // pinned state mirrors (max systems read back at once)
constexpr int kMaxPartials = 4;   // partial-sum arrays any reduction kernel writes (NV <= 4)
constexpr int kSizeClasses = 12;  // 8K, 32K, ..., 4^11 * 8K cells
constexpr int kMaxRPasses = 4;    // 'R' characters of a 4-character LINE sweep string

// Launch shape of the streaming kernels: threads per block, cells per
// thread, and threads (= lines) per block of the line-solve kernels.
struct Shape {
    int bs = 0, cpt = 0, lines_per_block = 0;
};

int size_class(int64_t n3);
void launch_resid_finish(const DGrid &g, int kind, int debug, double norm, double *partials, int64_t nblk,
                         mfix_leq_resid *d_out, hipStream_t s);
int default_block_size(int64_t n3, bool fp32);
int default_lines_per_block(int64_t n3);
double mixed_delta(int64_t n3);

using Sweep = std::array<char, 4>;

// Block maps of one launch sequence over the systems `ids` of a table:
// stream kernels (smap), finish kernels (fmap), reliable updates (umap,
// update_blocks() entries per system), batched red-black line
// phases (lmap[lines-per-block variant][pass]); LINE systems whose sweep string is not made of R/S
// passes are swept one by one (line_seq).
struct MapSet {
    std::vector<int> ids;
    std::vector<int2> smap, fmap, umap, lmap[kLineLpbVariants][kMaxRPasses];
    // FP32 tables: red-black passes of the several-lanes-per-line kernel,
    // one map per kernel variant (line length class) and pass
    std::vector<int2> pmap[kPcrVariants][kMaxRPasses];
    int nlpass = 0;
    std::vector<int> line_seq;
    double cells = 0.0;           // sum of N3 over ids (timing bytes)
    double k1_bytes = 0.0;        // algorithmic bytes of one first-stencil launch over all ids
    int n_line = 0, n_fused = 0;  // LINE systems / DIAG+NONE systems among ids
    DevBuf dev;
    int *d_ids = nullptr;
    int2 *d_smap = nullptr, *d_fmap = nullptr, *d_umap = nullptr, *d_lmap[kLineLpbVariants][kMaxRPasses] = {};
    int2 *d_pmap[kPcrVariants][kMaxRPasses] = {};
    uint64_t version = 0;         // bumped whenever the device copy changes
};

// A table of systems of one storage type and its device copy.
template <typename T>
struct Group {
    std::vector<Sys<T>> sys;
    std::vector<Sweep> sweep;     // LINE sweep string of each system
    int bs = 256;                 // threads per block of the stream kernels
    bool mixed = false;
    DevBuf dev;
    Sys<T> *d_sys = nullptr;
    MapSet iter;                  // systems being iterated
    // captured chunk of iterations (replayed until every system is done)
    hipGraphExec_t exec = nullptr;
    int exec_iters = 0;
    uint64_t exec_version = 0;
    const void *exec_sys = nullptr;
    void drop_graph();
};

struct Context {
    bool initialised = false;
    double mixed_delta = 0.0;      // mixed: reliable-update trigger, 0 = by size (mfix_leq_set_mixed_delta)
    bool have_grid = false;
    int device = 0;
    hipStream_t stream = nullptr;
    mfix_leq_grid grid = {};
    int iter_tot[10] = {0};
    DevBuf d_a, d_b, d_x;          // host-entry staging of A_M, B_M, VAR
    DevBuf line1;                  // one-system table, state and block map of a standalone LINE application
    DevBuf ws;                     // single-solve workspace (carved per solve)
    DevBuf bws;                    // batch workspace
    DevBuf partials, state, misc, tmp;
    SolverState *host_state = nullptr;  // pinned mirrors of device states
    int *host_count = nullptr;          // pinned running-system count
    Group<double> g64;             // single-solve tables
    Group<float> g32;
    Group<double> b64;             // batch tables
    Group<float> b32;
    Group<double> gna;             // table of the no-allocation entry (device side in its fixed array)
    Prof prof[kNumProf];
    Shape tuned[2][kSizeClasses];  // [fp32][size class]; bs == 0 means untuned
    int use_graphs = 1;            // capture chunks of iterations in HIP graphs
    int unit_diag = 1;             // mixed DIAG: negate instead of reading a 1/A0 plane of -1s
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
// Device address of the no-allocation entry's fixed array on device `dev`
// (the current device), taken once at set-up (leq_noalloc.hip).
int noalloc_prepare(int dev);
int parse_pc(const char *pc);
DGrid make_dgrid(const mfix_leq_grid &h);
double alg_bytes_per_cell(int pc, int sa, int sv, int sx);

// Description of one system to solve (device arrays are the caller's).
struct SysDesc {
    DGrid g;
    int pc;
    Sweep sweep;
    double tol;
    int itmax;
    int minimize, icheck;
    double delta;                 // mixed: reliable-update trigger
    int max_updates;              // mixed: reliable updates allowed (0 = no limit)
    double *x, *a, *b;            // VAR, A_M, B_M (FP64, device)
    const double *noise;
    uint64_t seed;
    Prologue pro;                 // residual / under-relaxation in the set-up pass
    mfix_leq_resid *resid_out = nullptr;  // host: the residual's result (pro.resid_kind >= 0)
    int resid_debug = 0;
    double resid_norm = 1.0;
};

// Workspace bytes of one system and the carving of that workspace.
size_t sys_bytes(const DGrid &g, const Part &P, int pc, bool mixed);
template <typename T>
void sys_carve(Sys<T> &S, const SysDesc &d, const Part &P, char *base);

// With ext (a device region of ext_cap bytes), build_maps and upload_table
// write the block maps / the table there instead of into their own
// grow-only buffers: nothing is allocated.
template <typename T>
int build_maps(const Group<T> &G, const std::vector<int> &ids, MapSet &M, hipStream_t s, char *ext = nullptr,
               size_t ext_cap = 0);
template <typename T>
int upload_table(Group<T> &G, hipStream_t s, char *ext = nullptr, size_t ext_cap = 0);
template <typename T>
void enqueue_setup(Group<T> &G, const MapSet &M, hipStream_t s);
template <typename T>
void enqueue_iterations(Group<T> &G, const MapSet &M, int n, hipStream_t s);
// d_count (kCountBytes of device memory; NULL: the context's own buffer)
// receives the running-system count and flags polled between chunks.
constexpr size_t kCountBytes = 256 + 2 * sizeof(int) * kHostStates;
template <typename T>
int run_group(Group<T> &G, int max_iters, hipStream_t s, bool graphs, int *d_count = nullptr);
template <typename T>
int read_states(const Group<T> &G, const std::vector<int> &ids, SolverState *h, hipStream_t s);

void fill_stats(const SolverState &h, mfix_leq_stats *st);
int solve_fp64(const DGrid &g, int pc, const char *cmethod, double tol, int itmax, double *x, double *a, double *b,
               const mfix_leq_opts *opts, mfix_leq_stats *stats, hipStream_t s, int minimize, int icheck);
int solve_mixed(const DGrid &g, int pc, const char *cmethod, double tol, int itmax, double *x, double *a, double *b,
                const mfix_leq_opts *opts, mfix_leq_stats *stats, hipStream_t s);
int solve_with_prologue(const DGrid &g, int pc, const char *cmethod, double tol, int itmax, double *x, double *a,
                        double *b, const mfix_leq_opts *opts, mfix_leq_stats *stats, hipStream_t s, int mixed,
                        int minimize, int icheck, const Prologue &pro, int resid_debug, double resid_norm,
                        mfix_leq_resid *resid_out);
// grids: one grid per item (NULL: a 3-D serial grid from each item's sizes)
int solve_batch(int nsys, const mfix_leq_batch_item *items, mfix_leq_stats *stats, hipStream_t s,
                const mfix_leq_grid *grids = nullptr);
template <typename T>
int launch_msolve_line(const DGrid &g, const Shape &sh, const char *cmethod, const T *a, const T *r, T *z,
                       T *scratch, const SolverState *st, int skip_on_sexit, int *fatal_flag, hipStream_t s,
                       bool copy_first);
int run_matvec64(const DGrid &g, const double *a, const double *x, double *y, hipStream_t s);
int run_matvec32(const DGrid &g, const float *a, const float *x, float *y, hipStream_t s);
int run_diag(const DGrid &g, const double *a, const double *r, double *z, hipStream_t s);
int run_dot(const DGrid &g, const double *x, const double *y, double *d_out, hipStream_t s);
int run_scale(const DGrid &g, double *a, double *b, int *zero_row, hipStream_t s);
int run_copy(const double *src, double *dst, int64_t n, hipStream_t s);
// LEQ_MSOLVE (LINE) with a sweep string of R/S passes through the batched
// red-black kernels (FP64: DGTSV per line; FP32 difference layout: several
// lanes per line); returns 1 (nothing launched) for other sweep strings.
template <typename T>
int run_msolve_line_rb(const DGrid &g, const char *cmethod, const T *a, const T *r, T *z, T *scratch,
                       hipStream_t s);
int autotune(int imax, int jmax, int kmax, int mixed, Shape *best);
// CALC_RESID_PP of device arrays on the current grid (leq_resid.hip).
int resid_pp(const double *b, const int *flag, double norm, int debug, mfix_leq_resid *out, hipStream_t s);

}  // namespace mfix
