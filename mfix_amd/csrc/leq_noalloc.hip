// LEQ_BICGS on device arrays without any allocation: mfix_leq_bicgs_dev's
// FP64 solve with its workspace (Krylov vectors, line-solve scratch, partial
// sums, solver state), its one-row system table, block maps and poll words
// carved from a fixed device array of kNoAllocBytes instead of the library's
// grow-only buffers.  For callers that cannot allocate device memory (code
// under stream capture, tight serving loops, harnesses that forbid it).
//
// The array is a file-scope __device__ symbol, so every GPU the library runs
// on holds its own copy until the process ends.  Host code never uses the
// symbol's name as an address: at set-up (Context::init, once per device,
// with that device current) hipGetSymbolAddress gives each device's
// address, kept per device, and only that address reaches the kernels.
// Every call on a GPU shares that GPU's array: calls that overlap on one
// GPU (different streams or threads) race and must be serialised by the
// caller.  Nothing carried over from the previous call is read before it is
// rewritten: the solver state is reset by the set-up kernel, and every
// partial sum and poll word is written before it is read.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#include "../../include/mfix_leq.h"
#include "leq_context.h"

namespace mfix {

namespace {

constexpr size_t kNoAllocBytes = MFIX_LEQ_NOALLOC_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) char scratch[kNoAllocBytes];

char *g_scratch[kMaxDevices] = {};

inline size_t up256(size_t v) { return (v + 255) & ~(size_t)255; }

// Upper bound of the block-map bytes build_maps writes for one system: the
// stream map (hardware blocks), finish map, reliable-update map and up to
// four red-black line passes, each region padded to 256 bytes.
size_t map_bytes(const DGrid &g, const Part &P) {
    const LineGeom L = line_geom(g, AX_J);
    const int64_t lines = ((int64_t)L.nu * L.nv + 1) / 2;
    const size_t e = sizeof(int2);
    size_t b = up256(sizeof(int));                                  // ids
    b += up256(e * (size_t)xcd_grid(P.nblk));                        // smap
    b += up256(e * kFinishBlocks);                                   // fmap
    b += up256(e * (size_t)update_blocks(P.nblk));                   // umap
    b += kMaxRPasses * up256(e * (size_t)((lines + 15) / 16 + kNumXcd));  // lmap of the variant in use
    b += (kLineLpbVariants - 1) * kMaxRPasses * 256;                 // the other variants' empty maps
    b += kPcrVariants * kMaxRPasses * 256;                           // empty FP32 maps
    return b;
}

}  // namespace

// Called by Context::init with device `dev` current.
int noalloc_prepare(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return MFIX_LEQ_EINVAL;
    if (g_scratch[dev]) return 0;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch[dev] = (char *)p;
    return 0;
}

// The FP64 solve of solve_fp64 on a table of one held in the fixed array;
// direct launches only (no graph, no event, no timing).
static int solve_fp64_noalloc(char *base, const SysDesc &d, const Shape &sh, mfix_leq_stats *stats, hipStream_t s) {
    Context &C = ctx();
    const Part P = make_part(d.g.n3, sh.bs, sh.cpt);
    const size_t o_tab = 0;
    const size_t o_cnt = o_tab + up256(sizeof(Sys<double>));
    const size_t o_map = o_cnt + up256(kCountBytes);
    const size_t maps = map_bytes(d.g, P);
    const size_t o_ws = o_map + maps;
    const size_t need = o_ws + sys_bytes(d.g, P, d.pc, false);
    if (need > kNoAllocBytes) {
        std::snprintf(C.last_msg, sizeof(C.last_msg),
                      "system of %lld cells needs %zu bytes of workspace; the no-allocation entry holds %zu",
                      (long long)d.g.n3, need, kNoAllocBytes);
        return MFIX_LEQ_ETOOBIG;
    }
    Group<double> &G = C.gna;
    G.sys.resize(1);
    G.sweep.assign(1, d.sweep);
    G.bs = sh.bs;
    G.mixed = false;
    sys_carve<double>(G.sys[0], d, P, base + o_ws);
    int rc;
    if ((rc = upload_table<double>(G, s, base + o_tab, up256(sizeof(Sys<double>))))) return rc;
    if ((rc = build_maps<double>(G, std::vector<int>{0}, G.iter, s, base + o_map, maps))) return rc;
    enqueue_setup<double>(G, G.iter, s);
    rc = run_group<double>(G, d.itmax, s, false, (int *)(base + o_cnt));
    SolverState *h = C.host_state;
    if (rc || (rc = read_states<double>(G, std::vector<int>{0}, h, s))) return rc;
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    fill_stats(*h, stats);
    if (stats) stats->device_ms = 0.0;
    if (h->msolve_fatal) return MFIX_LEQ_EDGTSV;
    return 0;
}

}  // namespace mfix

using namespace mfix;

extern "C" {

int mfix_leq_bicgs_dev_noalloc(const char *cmethod, double tol, const char *pc, int itmax, double *d_var,
                               double *d_a_m, double *d_b_m, const mfix_leq_opts *opts, mfix_leq_stats *stats,
                               void *stream) {
    Context &C = ctx();
    const int p = parse_pc(pc);
    if (p < 0) return MFIX_LEQ_EBADPC;
    // initialising the library allocates (stream, pinned state): it happened
    // before, in mfix_leq_init or mfix_leq_grid_init, which this entry needs
    if (!C.initialised || !C.have_grid) return MFIX_LEQ_ENOGRID;
    if (!g_scratch[C.device]) return MFIX_LEQ_EHIP;
    if (!d_var || !d_a_m || !d_b_m) return MFIX_LEQ_EINVAL;
    const hipStream_t s = stream ? (hipStream_t)stream : C.stream;
    const DGrid g = make_dgrid(C.grid);
    char cm[5] = "RSRS";
    if (cmethod)
        for (int q = 0; q < 4 && cmethod[q]; ++q) cm[q] = cmethod[q];
    SysDesc d;
    d.g = g;
    d.pc = p;
    d.sweep = {cm[0], cm[1], cm[2], cm[3]};
    d.tol = tol;
    d.itmax = itmax;
    d.minimize = C.grid.minimize_dotproducts;
    d.icheck = C.grid.icheck_bicgs;
    d.delta = 0.0;
    d.max_updates = 0;
    d.x = d_var;
    d.a = d_a_m;
    d.b = d_b_m;
    d.noise = opts ? opts->noise : nullptr;
    d.seed = opts ? opts->noise_seed : 0;
    return solve_fp64_noalloc(g_scratch[C.device], d, C.shape_for(g.n3, false), stats, s);
}

}  // extern "C"
