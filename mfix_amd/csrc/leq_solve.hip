// Solver engine of the gfx950 LEQ_BICGS path: device context, grow-only
// workspaces, launch shapes (picked per grid size by the autotuner), system
// tables and their block maps, the launch sequence of LEQ_BICGS0
// (model/leq_bicgs.f:107-509) over a table, the driver that polls the
// running-system count between chunks of iterations (replaying a captured
// HIP graph per chunk), and the single and batched solve entries built on
// it.  The mixed-precision inner solve (FP32 recurrence, FP64 reliable
// updates) runs the same sequence on a float table.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/mfix_leq.h"
#include "leq_context.h"

namespace mfix {

Context &ctx() {
    static Context c;
    return c;
}

int hip_fail(hipError_t e, const char *what) {
    ctx().last_hip_error = e;
    std::snprintf(ctx().last_msg, sizeof(ctx().last_msg), "%s: %s", what, hipGetErrorString(e));
    if (std::getenv("MFIX_LEQ_VERBOSE")) std::fprintf(stderr, "mfix_leq: %s\n", ctx().last_msg);
    return MFIX_LEQ_EHIP;
}

int parse_pc(const char *pc) {
    if (!pc) return -1;
    if (std::strncmp(pc, "LINE", 4) == 0) return PC_LINE;
    if (std::strncmp(pc, "DIAG", 4) == 0) return PC_DIAG;
    if (std::strncmp(pc, "NONE", 4) == 0) return PC_NONE;
    return -1;
}

DGrid make_dgrid(const mfix_leq_grid &h) {
    DGrid g;
    g.ni = h.imax + 2;
    g.nj = h.jmax + 2;
    g.nk = h.no_k ? 1 : h.kmax + 2;
    g.do_k = !h.no_k;
    g.cyc_x = h.cyclic_x;
    g.cyc_y = h.cyclic_y;
    g.cyc_z = h.cyclic_z && !h.no_k;
    g.dot_all = h.is_serial ? 1 : 0;
    g.c1 = g.nj;
    g.c2 = (int64_t)g.nj * g.ni;
    g.n3 = g.c2 * g.nk;
    div_magic((uint32_t)g.c2, g.mc2, g.sc2);
    div_magic((uint32_t)g.nj, g.mnj, g.snj);
    return g;
}

// ---- device buffers -------------------------------------------------------

int DevBuf::ensure(size_t bytes) {
    if (bytes <= cap) return 0;
    if (ptr) {
        hipError_t e = hipFree(ptr);
        ptr = nullptr;
        cap = 0;
        if (e != hipSuccess) return hip_fail(e, "hipFree");
    }
    hipError_t e = hipMalloc(&ptr, bytes);
    if (e != hipSuccess) {
        ptr = nullptr;
        hip_fail(e, "hipMalloc");
        return MFIX_LEQ_ENOMEM;
    }
    cap = bytes;
    return 0;
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
}

template <typename T>
void Group<T>::drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    exec_iters = 0;
}
template struct Group<double>;
template struct Group<float>;

int Context::ensure_init() {
    if (initialised) return 0;
    return init(0);
}

int Context::init(int dev) {
    if (initialised) return 0;
    hipError_t e = hipSetDevice(dev);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    device = dev;
    int rc;
    if ((rc = noalloc_prepare(dev))) return rc;
    e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
    e = hipHostMalloc((void **)&host_state, sizeof(SolverState) * kHostStates, hipHostMallocDefault);
    if (e != hipSuccess) return hip_fail(e, "hipHostMalloc");
    e = hipHostMalloc((void **)&host_count, sizeof(int) * 4, hipHostMallocDefault);
    if (e != hipSuccess) return hip_fail(e, "hipHostMalloc");
    if (const char *v = std::getenv("MFIX_LEQ_GRAPHS")) use_graphs = std::atoi(v);
    initialised = true;
    return 0;
}

void Context::release_all() {
    for (DevBuf *b : {&d_a, &d_b, &d_x, &ws, &bws, &partials, &state, &misc, &tmp, &line1}) b->release();
    g64.drop_graph();
    g32.drop_graph();
    b64.drop_graph();
    b32.drop_graph();
    for (DevBuf *b : {&g64.dev, &g32.dev, &b64.dev, &b32.dev, &g64.iter.dev, &g32.iter.dev, &b64.iter.dev,
                      &b32.iter.dev})
        b->release();
    g64.iter.version = g32.iter.version = b64.iter.version = b32.iter.version = 0;
    for (auto &p : prof) {
        for (auto &ev : p.events) {
            (void)hipEventDestroy(ev.first);
            (void)hipEventDestroy(ev.second);
        }
        p.events.clear();
        p.pending_bytes.clear();
        p.pool_used = 0;
    }
}

// ---- per-kernel event timing ----------------------------------------------

std::pair<hipEvent_t, hipEvent_t> Prof::take(double bytes) {
    if (!on) return {nullptr, nullptr};
    if (pool_used >= events.size()) {
        std::pair<hipEvent_t, hipEvent_t> ev;
        (void)hipEventCreate(&ev.first);
        (void)hipEventCreate(&ev.second);
        events.push_back(ev);
    }
    pending_bytes.push_back(bytes);
    return events[pool_used++];
}

void Prof::collect(int64_t working) {
    for (size_t q = 0; q < pool_used; ++q) {
        float ms = 0.f;
        if (hipEventSynchronize(events[q].second) == hipSuccess &&
            hipEventElapsedTime(&ms, events[q].first, events[q].second) == hipSuccess) {
            if (working >= 0 && (int64_t)q >= working) {
                idle_ms += ms;
                ++idle_launches;
                continue;
            }
            total_ms += ms;
            alg_bytes += pending_bytes[q];
            ++launches;
        }
    }
    pool_used = 0;
    pending_bytes.clear();
}

// ---- launch shapes -----------------------------------------------------------

// Shape of the streaming kernels for a grid of n3 cells.  The autotuner
// (mfix_leq_autotune) times the candidates on the device and records the
// fastest per size class and precision; untuned classes use
// default_block_size.
Shape Context::shape_for(int64_t n3, bool fp32) const {
    const int cls = size_class(n3);
    Shape s = tuned[fp32 ? 1 : 0][cls];
    if (!s.bs) s = Shape{default_block_size(n3, fp32), 1, 0};
    if (!s.lines_per_block) s.lines_per_block = default_lines_per_block(n3);
    return s;
}

// Cells per logical block of the stencil passes when none was tuned: 256,
// and 512 for FP32 systems of more than 2^23 cells.  Measured on the 256^3
// mixed DIAG solve (two solves each): 476.0-478.8 us per iteration with 256
// cells, 471.1-471.2 with 512, 471.8-472.1 with 1024 (the finish of 512- and
// 1024-cell blocks sums 33.5K / 16.8K partials in two levels; the single-block
// finish over 16.8K partials ran 474.6-474.9).
int default_block_size(int64_t n3, bool fp32) { return fp32 && n3 > (int64_t(1) << 23) ? 512 : 256; }

// Lines per block of the red-black line kernel when none was tuned: 64.
// A colour phase is bound by the chain of dependent memory round trips of
// one line (tile loads forward, scratch reads back), which fewer lines per
// block do not shorten: measured at 128^3 (8,450 lines per colour) 73 / 120 /
// 168 us per phase at 64 / 32 / 16 lines per block, although 64 leaves half
// the CUs idle; at 256^3 476 / 575 / 937 us.
int default_lines_per_block(int64_t) { return 64; }

// Reliable-update trigger of mixed solves that set none (opts->inner_tol):
// the recursive FP32 residual fell by 1e-4 since the last FP64 update on
// systems below 2^23 cells, by 0.1 from there on.  Measured iterations to
// 1e-10 (FP64 / delta 0.1 / 1e-3 / 1e-4): pressure 64^3 527 / 879 / 610 /
// 604, 128^3 851 / 1410 / 1084 / 1027, 160^3 956 / 1333 / 1531 / 954;
// void-fraction pressure 128^3 776 / 840 / 822 / 854, 256^3 1165 / 1054 /
// 1411 / 1678; upwind momentum 64^3 13 iterations with 7 / 4 / 3 FP64
// updates.  On large systems the frequent restarts from the FP64 residual
// pay for themselves; below, each restart discards Krylov information the
// FP32 recurrence could still use.
double mixed_delta(int64_t n3) {
    if (ctx().mixed_delta > 0.0) return ctx().mixed_delta;
    return n3 < (int64_t(1) << 23) ? 1e-4 : 0.1;
}

int size_class(int64_t n3) {
    int c = 0;
    while (c < kSizeClasses - 1 && (int64_t(1) << (2 * c + 13)) < n3) ++c;
    return c;  // 8K, 32K, 128K, ... cells
}

#define MFIX_DISPATCH_BS(bs, BSV, ...)                 \
    switch (bs) {                                      \
    case 512: { constexpr int BSV = 512; __VA_ARGS__; } break; \
    case 1024: { constexpr int BSV = 1024; __VA_ARGS__; } break; \
    default: { constexpr int BSV = 256; __VA_ARGS__; } break; \
    }

// Map entries of system `id`'s nb logical blocks in XCD slab order: slot j
// of an 8-aligned range runs on XCD j % 8 and holds the logical block that
// gives each XCD one contiguous run of the system (padding slots: -1).
static void push_slab_map(std::vector<int2> &out, int id, int64_t nb) {
    const int64_t per = (nb + kNumXcd - 1) / kNumXcd;
    const int64_t full = nb - (per - 1) * kNumXcd;
    for (int64_t j = 0; j < per * kNumXcd; ++j) {
        const int64_t x = j % kNumXcd, r = j / kNumXcd;
        const bool valid = r < (x < full ? per : per - 1);
        const int64_t start = x < full ? x * per : full * per + (x - full) * (per - 1);
        out.push_back(valid ? make_int2(id, (int)(start + r)) : make_int2(-1, -1));
    }
}

static inline dim3 grid_of(const Part &P) { return dim3((unsigned)xcd_grid(P.nblk)); }
static inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// Finish blocks of a reduction over nblk partials.  Up to kFinishSingle
// partials one block sums them all (up to 16 per thread, one memory round
// trip): no device atomic, no acquire fence and no second level, whose
// latency cost ~3.5 us per finish at 64^3 where two blocks had split the
// 1,123 partials (mixed iteration 52.8 -> 42.4 us, FP64 50.8 -> 39.7 us;
// 96^3 72.0 -> 61.5 / 86.5 -> 76.2; 128^3, 8,582 partials, 104 -> 98 /
// 153 -> 148).  Larger grids split into about one partial per thread over at
// most kFinishBlocks blocks (256^3: 67 K partials, where one block took
// 16-35 us).
constexpr int64_t kFinishSingle = 16384;
static inline int finish_blocks(int64_t nblk) {
    if (nblk <= kFinishSingle) return 1;
    const int64_t want = (nblk + kFinishBlock - 1) / kFinishBlock;
    return (int)std::max<int64_t>(1, std::min<int64_t>(kFinishBlocks, want));
}

// ---- LINE preconditioner, one system (LEQ_MSOLVE, leqsol_mod.f:286-598) ----

template <typename T>
int launch_msolve_line(const DGrid &g, const Shape &sh, const char *cmethod, const T *a, const T *r, T *z,
                       T *scratch, const SolverState *st, int skip_on_sexit, int *fatal_flag, hipStream_t s,
                       bool copy_first) {
    const int64_t n3 = g.n3;
    Prof &pf = ctx().prof[1];
    if (copy_first) {
        const Part P = make_part(n3, sh.bs, sh.cpt);
        MFIX_DISPATCH_BS(sh.bs, BS,
                         hipLaunchKernelGGL((k_copy_if<BS, T>), grid_of(P), dim3(BS), 0, s, P, r, z, st,
                                            skip_on_sexit));
    }
    // at most 64 lines per block: the kernel's launch bound and the scratch
    // size (line_scratch_elems) assume it
    const int lpb = sh.lines_per_block > 0 && sh.lines_per_block < 64 ? sh.lines_per_block : 64;
    auto run = [&](const LinePhase &ph, int64_t nthreads) {
        const unsigned blocks = (unsigned)((nthreads + lpb - 1) / lpb);
        const LineGeom L = line_geom(g, ph.axis);
        // algorithmic bytes of one phase: each line solved reads A (7
        // planes), the right-hand side and four neighbour values and writes
        // the line once
        const double cells = (double)std::min<int64_t>(nthreads, (int64_t)L.nu * L.nv) * L.n *
                             (ph.kind == PH_COLOR ? 0.5 : 1.0);
        const auto ev = pf.take(cells * (7.0 * sizeof(T) + 6.0 * sizeof(T)));
        launch_timed(ev.first, ev.second, k_line_phase<T>, dim3(blocks ? blocks : 1), dim3(lpb), s, g, ph, a, r, z,
                     scratch, st, skip_on_sexit, fatal_flag);
    };
    for (int it = 0; it < 4; ++it) {
        const char ch = cmethod ? cmethod[it] : 'R';
        const bool do_i = ch == 'I' || ch == 'i', do_k = ch == 'K' || ch == 'k';
        const bool do_all = ch == 'A' || ch == 'a', do_rb = ch == 'R' || ch == 'r';
        if (!g.do_k) {
            // two-dimensional (leqsol_mod.f:367-392)
            const LineGeom L = line_geom(g, AX_J);
            if (do_i)
                for (int d = 0; d < L.nu; ++d) run(LinePhase{AX_J, PH_DIAG, d, 0, 0, 0, 0}, 1);
            if (do_rb) {
                run(LinePhase{AX_J, PH_COLOR, 1, 0, 0, 0, 0}, (int64_t)L.nu * L.nv);
                run(LinePhase{AX_J, PH_COLOR, 0, 0, 0, 0, 0}, (int64_t)L.nu * L.nv);
            }
            continue;
        }
        if (do_all) {
            // JK lines (along I) and IJ lines (along K): DGTSV failure is fatal;
            // then IK lines (along J).  The reference's enumeration is a proper
            // red-black colouring iff the fast size is even (leqsol_mod.f:399-469).
            const int axes[3] = {AX_I, AX_K, AX_J};
            const int fast[3] = {g.nj, g.ni, g.ni};
            const int fat[3] = {1, 1, 0};
            for (int q = 0; q < 3; ++q) {
                const LineGeom L = line_geom(g, axes[q]);
                if (fast[q] % 2 == 0) {
                    run(LinePhase{axes[q], PH_COLOR, 1, 1, fat[q], 0, 0}, (int64_t)L.nu * L.nv);
                    run(LinePhase{axes[q], PH_COLOR, 0, 1, fat[q], 0, 0}, (int64_t)L.nu * L.nv);
                } else {
                    run(LinePhase{axes[q], PH_SEQ, 0, 1, fat[q], fast[q], 1}, 1);
                    run(LinePhase{axes[q], PH_SEQ, 0, 1, fat[q], fast[q], 2}, 1);
                }
            }
        }
        if (do_rb) {
            // leqsol_mod.f:497-522: (i+k) odd lines first, then (i+k) even
            const LineGeom L = line_geom(g, AX_J);
            run(LinePhase{AX_J, PH_COLOR, 1, 1, 0, 0, 0}, (int64_t)L.nu * L.nv);
            run(LinePhase{AX_J, PH_COLOR, 0, 1, 0, 0, 0}, (int64_t)L.nu * L.nv);
        }
        if (do_i || do_k) {
            // lexicographic line Gauss-Seidel over (i,k) (leqsol_mod.f:570-585):
            // the I and K orders have the same dependencies, so both become
            // anti-diagonal wavefronts
            const LineGeom L = line_geom(g, AX_J);
            for (int d = 0; d <= L.nu + L.nv - 2; ++d) {
                const int lo = std::max(0, d - (L.nv - 1)), hi = std::min(d, L.nu - 1);
                run(LinePhase{AX_J, PH_DIAG, d, 0, 0, 0, 0}, hi - lo + 1);
            }
        }
    }
    return 0;
}

// LEQ_MSOLVE (leqsol_mod.f:286-598) with a sweep string whose characters are
// R passes or characters that do nothing here (S; J; K and A in 2-D),
// through the red-black kernels the solver uses: z = r (SETGUESS,
// :341-357), then per R pass the (i+k)-odd colour and the even one.  FP64:
// k_line_color_sys (DGTSV per line, the reference's bits); FP32 (A in the
// inner solve's difference layout, A0 in plane 7): k_line_pcr_sys where the
// line fits one of its variants.  A one-system table carries the arrays; its
// state is blank, so the kernels never return early.  Returns 1, launching
// nothing, for a sweep string with I/K/A passes (the caller runs
// k_line_phase instead).
template <typename T>
static void launch_line_color(int lv, dim3 gs, hipStream_t s, const Sys<T> *tab, const int2 *mp, int param,
                              int which, std::pair<hipEvent_t, hipEvent_t> ev = {nullptr, nullptr}) {
    const dim3 lb(kLineBlock);
    switch (lv) {
    case 0: launch_timed(ev.first, ev.second, k_line_color_sys<T, 64>, gs, lb, s, tab, mp, param, which); break;
    case 1: launch_timed(ev.first, ev.second, k_line_color_sys<T, 32>, gs, lb, s, tab, mp, param, which); break;
    default: launch_timed(ev.first, ev.second, k_line_color_sys<T, 16>, gs, lb, s, tab, mp, param, which); break;
    }
}

template <typename T>
int run_msolve_line_rb(const DGrid &g, const char *cmethod, const T *a, const T *r, T *z, T *scratch,
                       hipStream_t s) {
    Sweep sw = {'R', 'S', 'R', 'S'};
    if (cmethod)
        for (int q = 0; q < 4 && cmethod[q]; ++q) sw[q] = cmethod[q];
    int np = 0;
    for (char ch : sw) {
        if (ch == 'R' || ch == 'r') ++np;
        else if (ch == 'I' || ch == 'i') return 1;
        else if (g.do_k && (ch == 'K' || ch == 'k' || ch == 'A' || ch == 'a')) return 1;
    }
    Context &C = ctx();
    const LineGeom L = line_geom(g, AX_J);
    const int64_t lines = ((int64_t)L.nu * L.nv + 1) / 2;
    const int var = sizeof(T) == sizeof(float) ? line_pcr_variant(L.n) : -1;
    const int lv = line_lpb_variant(C.shape_for(g.n3, sizeof(T) == sizeof(float)).lines_per_block);
    const int lpb = var >= 0 ? line_pcr_lpb(var) : line_lpb_of(lv);
    const int64_t lb = (lines + lpb - 1) / lpb;
    const size_t o_st = align256(sizeof(Sys<T>));
    const size_t o_map = o_st + align256(sizeof(SolverState));
    int rc;
    if ((rc = C.line1.ensure(o_map + sizeof(int2) * (size_t)(lb + kNumXcd)))) return rc;
    char *base = (char *)C.line1.ptr;
    Sys<T> S = {};
    S.g = g;
    S.pc = PC_LINE;
    S.a = a;
    S.p = const_cast<T *>(r);
    S.phat = z;
    S.line_scratch = scratch;
    S.st = (SolverState *)(base + o_st);
    std::vector<int2> map;
    push_slab_map(map, 0, lb);
    hipError_t e = hipMemcpyAsync(base, &S, sizeof(S), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(base + o_map, map.data(), sizeof(int2) * map.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(z, r, sizeof(T) * (size_t)g.n3, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "line table upload");
    hipLaunchKernelGGL(k_init_state, dim3(1), dim3(1), 0, s, S.st);
    Prof &pf = C.prof[1];
    const Sys<T> *tab = (const Sys<T> *)base;
    const int2 *mp = (const int2 *)(base + o_map);
    const dim3 gs((unsigned)map.size());
    for (int q = 0; q < np; ++q)
        for (int param : {1, 0}) {
            const auto ev = pf.take(0.5 * (double)g.n3 * 13.0 * sizeof(T));
            if constexpr (sizeof(T) == sizeof(float)) {
                const Sys<float> *tf = (const Sys<float> *)tab;
                if (var >= 0) launch_line_pcr(var, gs, s, tf, mp, param, 0, ev.first, ev.second);
                else launch_line_color<float>(lv, gs, s, tf, mp, param, 0, ev);
            } else {
                launch_line_color<T>(lv, gs, s, tab, mp, param, 0, ev);
            }
        }
    // the table must outlive the launches before it is rewritten
    e = hipStreamSynchronize(s);
    return e == hipSuccess ? 0 : hip_fail(e, "line solve");
}
template int run_msolve_line_rb<double>(const DGrid &, const char *, const double *, const double *, double *,
                                        double *, hipStream_t);
template int run_msolve_line_rb<float>(const DGrid &, const char *, const float *, const float *, float *, float *,
                                       hipStream_t);

template int launch_msolve_line<double>(const DGrid &, const Shape &, const char *, const double *, const double *,
                                        double *, double *, const SolverState *, int, int *, hipStream_t, bool);
template int launch_msolve_line<float>(const DGrid &, const Shape &, const char *, const float *, const float *,
                                       float *, float *, const SolverState *, int, int *, hipStream_t, bool);

// Per-iteration algorithmic bytes per cell of the fused three-pass schedule.
double alg_bytes_per_cell(int pc, int sa, int sv, int sx) {
    if (pc == PC_DIAG) return 15.0 * sa + 15.0 * sv + 2.0 * sx;
    if (pc == PC_NONE) return 14.0 * sa + 15.0 * sv + 2.0 * sx;
    return 42.0 * sa + 30.0 * sv + 2.0 * sx;
}

// ---- system workspaces --------------------------------------------------------

// FP64: R, Rtilde, P, S, T, V (+ Phat, Shat and the DGTSV scratch for LINE).
// Mixed: the FP32 r, rt, p, s, t, v (+ LINE arrays), the FP64 correction
// accumulator y, the 8-plane FP32 matrix, the FP64 residual and second
// solution buffer.  Both: the
// partial sums and the solver state.
size_t sys_bytes(const DGrid &g, const Part &P, int pc, bool mixed) {
    const size_t n3 = (size_t)g.n3;
    const size_t sv = mixed ? sizeof(float) : sizeof(double);
    const bool line = pc == PC_LINE;
    size_t b = 0;
    b += (size_t)6 * align256(n3 * sv);
    if (mixed) b += align256(n3 * sizeof(double));  // y
    if (line) b += 2 * align256(n3 * sv) + align256((size_t)line_scratch_elems(g) * sv);
    if (mixed) b += align256(8 * n3 * sizeof(float)) + 2 * align256(n3 * sizeof(double));
    b += align256(sizeof(double) * kMaxPartials * (size_t)partial_slots(P.nblk));
    b += align256(sizeof(SolverState));
    return b;
}

template <typename T>
void sys_carve(Sys<T> &S, const SysDesc &d, const Part &P, char *base) {
    const bool mixed = sizeof(T) == sizeof(float);
    const size_t n3 = (size_t)d.g.n3;
    char *q = base;
    auto take = [&](size_t bytes) {
        char *p = q;
        q += align256(bytes);
        return p;
    };
    S = Sys<T>{};
    S.g = d.g;
    S.P = P;
    S.nfin = finish_blocks(P.nblk);
    S.pc = d.pc;
    S.itmax = d.itmax;
    S.minimize = d.minimize;
    S.icheck = d.icheck;
    S.mixed = mixed ? 1 : 0;
    S.tol = d.tol;
    S.delta = d.delta;
    S.max_updates = d.max_updates;
    S.unit_diag = ctx().unit_diag;
    S.pro = d.pro;
    S.r = (T *)take(n3 * sizeof(T));
    S.rt = (T *)take(n3 * sizeof(T));
    S.p = (T *)take(n3 * sizeof(T));
    S.s = (T *)take(n3 * sizeof(T));
    S.t = (T *)take(n3 * sizeof(T));
    S.v = (T *)take(n3 * sizeof(T));
    if (mixed) {
        S.x = nullptr;
        S.y = (double *)take(n3 * sizeof(double));
    } else {
        S.x = (T *)d.x;
        S.y = nullptr;
    }
    if (d.pc == PC_LINE) {
        S.phat = (T *)take(n3 * sizeof(T));
        S.shat = (T *)take(n3 * sizeof(T));
        S.line_scratch = (T *)take((size_t)line_scratch_elems(d.g) * sizeof(T));
    }
    if (mixed) {
        S.a = (const T *)take(8 * n3 * sizeof(float));
        S.r64 = (double *)take(n3 * sizeof(double));
        S.xb = (double *)take(n3 * sizeof(double));
    } else {
        S.a = (const T *)d.a;
        S.r64 = (double *)S.r;
        S.xb = nullptr;
    }
    S.partials = (double *)take(sizeof(double) * kMaxPartials * (size_t)partial_slots(P.nblk));
    S.st = (SolverState *)take(sizeof(SolverState));
    S.a64 = d.a;
    S.b64 = d.b;
    S.xa = d.x;
    S.noise = d.noise;
    S.seed = d.seed;
}
template void sys_carve<double>(Sys<double> &, const SysDesc &, const Part &, char *);
template void sys_carve<float>(Sys<float> &, const SysDesc &, const Part &, char *);

// ---- block maps and table upload ----------------------------------------------

// Which LINE sweeps of a system run as batched red-black passes: every
// character is R (a pass) or a character that does nothing in this
// dimension (S, and J; K and A too in 2-D).  Returns the number of passes,
// or -1 when the system needs the one-by-one sweep (I/K wavefronts, A).
static int rb_passes(const Sweep &sw, bool do_k) {
    int n = 0;
    for (char ch : sw) {
        if (ch == 'R' || ch == 'r') ++n;
        else if (ch == 'I' || ch == 'i') return -1;
        else if (do_k && (ch == 'K' || ch == 'k' || ch == 'A' || ch == 'a')) return -1;
    }
    return n;
}

template <typename T>
int build_maps(const Group<T> &G, const std::vector<int> &ids, MapSet &M, hipStream_t s, char *ext, size_t ext_cap) {
    std::vector<int2> smap, fmap, umap, lmap[kLineLpbVariants][kMaxRPasses], pmap[kPcrVariants][kMaxRPasses];
    std::vector<int> seq;
    int nlpass = 0, n_line = 0, n_fused = 0;
    double cells = 0.0, k1 = 0.0;
    const double sa = sizeof(T), sv = sizeof(T);
    for (int id : ids) {
        const Sys<T> &S = G.sys[id];
        cells += (double)S.g.n3;
        // first stencil pass: the seven coefficient planes, the operand, the
        // dot partner and the result (an FP32 DIAG system whose 1/A0 is not
        // uniformly -1 also reads its 1/A0 plane: not counted, so the rate
        // derived from these bytes is a lower bound there)
        if (S.pc == PC_LINE) ++n_line;
        else ++n_fused;
        k1 += (double)S.g.n3 * (7.0 * sa + 3.0 * sv);
        // stream kernels: an 8-aligned range whose slot j holds the logical
        // block that puts one contiguous slab of the system on each XCD
        const int64_t nb = S.P.nblk;
        push_slab_map(smap, id, nb);
        for (int f = 0; f < S.nfin; ++f) fmap.push_back(make_int2(id, f));
        for (int64_t u = 0; u < update_blocks(nb); ++u) umap.push_back(make_int2(id, (int)u));
        if (S.pc == PC_LINE) {
            const int np = rb_passes(G.sweep[id], S.g.do_k != 0);
            if (np < 0) {
                seq.push_back(id);
            } else {
                const LineGeom L = line_geom(S.g, AX_J);
                const int64_t lines = ((int64_t)L.nu * L.nv + 1) / 2;
                // FP32 (the mixed path's inner solve): several lanes per line
                // where the line fits a variant; FP64: DGTSV lane per line
                const int var = sizeof(T) == sizeof(float) ? line_pcr_variant(L.n) : -1;
                const int lv = line_lpb_variant(ctx().shape_for(S.g.n3, sizeof(T) == sizeof(float)).lines_per_block);
                const int lpb = var >= 0 ? line_pcr_lpb(var) : line_lpb_of(lv);
                const int64_t lb = (lines + lpb - 1) / lpb;
                // the same XCD slab order as the stream kernels: XCD x takes a
                // contiguous run of line blocks, so the I/K neighbour lines its
                // blocks read (the other colour's, about two blocks away along
                // K) come through one L2
                for (int q = 0; q < np; ++q) {
                    if (var >= 0) push_slab_map(pmap[var][q], id, lb);
                    else push_slab_map(lmap[lv][q], id, lb);
                }
                nlpass = std::max(nlpass, np);
            }
        }
    }
    bool same = M.ids == ids && M.smap.size() == smap.size() && M.fmap.size() == fmap.size() &&
                M.nlpass == nlpass && M.line_seq == seq && M.dev.ptr;
    if (same) same = !std::memcmp(M.smap.data(), smap.data(), smap.size() * sizeof(int2)) &&
                     !std::memcmp(M.fmap.data(), fmap.data(), fmap.size() * sizeof(int2)) &&
                     M.umap.size() == umap.size() &&
                     !std::memcmp(M.umap.data(), umap.data(), umap.size() * sizeof(int2));
    for (int q = 0; q < kMaxRPasses && same; ++q) {
        for (int v = 0; v < kLineLpbVariants && same; ++v)
            same = M.lmap[v][q].size() == lmap[v][q].size() &&
                   !std::memcmp(M.lmap[v][q].data(), lmap[v][q].data(), lmap[v][q].size() * sizeof(int2));
        for (int v = 0; v < kPcrVariants && same; ++v)
            same = M.pmap[v][q].size() == pmap[v][q].size() &&
                   !std::memcmp(M.pmap[v][q].data(), pmap[v][q].data(), pmap[v][q].size() * sizeof(int2));
    }
    M.cells = cells;
    M.k1_bytes = k1;
    M.n_line = n_line;
    M.n_fused = n_fused;
    if (same && !ext) return 0;
    M.ids = ids;
    M.smap.swap(smap);
    M.fmap.swap(fmap);
    M.umap.swap(umap);
    for (int q = 0; q < kMaxRPasses; ++q) {
        for (int v = 0; v < kLineLpbVariants; ++v) M.lmap[v][q].swap(lmap[v][q]);
        for (int v = 0; v < kPcrVariants; ++v) M.pmap[v][q].swap(pmap[v][q]);
    }
    M.nlpass = nlpass;
    M.line_seq.swap(seq);
    // one device block: ids | smap | fmap | umap | lmap[variant][pass] | pmap[variant][pass]
    std::vector<char> blob;
    auto put = [&](const void *p, size_t bytes) {
        const size_t off = blob.size();
        blob.resize(align256(off + bytes));
        if (bytes) std::memcpy(blob.data() + off, p, bytes);
        return off;
    };
    const size_t o_ids = put(M.ids.data(), M.ids.size() * sizeof(int));
    const size_t o_s = put(M.smap.data(), M.smap.size() * sizeof(int2));
    const size_t o_f = put(M.fmap.data(), M.fmap.size() * sizeof(int2));
    const size_t o_u = put(M.umap.data(), M.umap.size() * sizeof(int2));
    size_t o_l[kLineLpbVariants][kMaxRPasses], o_p[kPcrVariants][kMaxRPasses];
    for (int v = 0; v < kLineLpbVariants; ++v)
        for (int q = 0; q < kMaxRPasses; ++q) o_l[v][q] = put(M.lmap[v][q].data(), M.lmap[v][q].size() * sizeof(int2));
    for (int v = 0; v < kPcrVariants; ++v)
        for (int q = 0; q < kMaxRPasses; ++q) o_p[v][q] = put(M.pmap[v][q].data(), M.pmap[v][q].size() * sizeof(int2));
    int rc;
    char *base = ext;
    if (ext) {
        if (blob.size() > ext_cap) return MFIX_LEQ_ENOMEM;
    } else {
        if ((rc = M.dev.ensure(std::max<size_t>(blob.size(), 256)))) return rc;
        base = (char *)M.dev.ptr;
    }
    hipError_t e = hipMemcpyAsync(base, blob.data(), blob.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "block map upload");
    M.d_ids = (int *)(base + o_ids);
    M.d_smap = (int2 *)(base + o_s);
    M.d_fmap = (int2 *)(base + o_f);
    M.d_umap = (int2 *)(base + o_u);
    for (int v = 0; v < kLineLpbVariants; ++v)
        for (int q = 0; q < kMaxRPasses; ++q) M.d_lmap[v][q] = (int2 *)(base + o_l[v][q]);
    for (int v = 0; v < kPcrVariants; ++v)
        for (int q = 0; q < kMaxRPasses; ++q) M.d_pmap[v][q] = (int2 *)(base + o_p[v][q]);
    M.version += 1;
    return 0;
}
template int build_maps<double>(const Group<double> &, const std::vector<int> &, MapSet &, hipStream_t, char *,
                                size_t);
template int build_maps<float>(const Group<float> &, const std::vector<int> &, MapSet &, hipStream_t, char *, size_t);

template <typename T>
int upload_table(Group<T> &G, hipStream_t s, char *ext, size_t ext_cap) {
    int rc;
    const size_t bytes = G.sys.size() * sizeof(Sys<T>);
    Sys<T> *dst = (Sys<T> *)ext;
    if (ext) {
        if (bytes > ext_cap) return MFIX_LEQ_ENOMEM;
    } else {
        if ((rc = G.dev.ensure(std::max<size_t>(bytes, 256)))) return rc;
        dst = (Sys<T> *)G.dev.ptr;
    }
    if (G.d_sys != dst) G.drop_graph();
    G.d_sys = dst;
    if (!bytes) return 0;
    hipError_t e = hipMemcpyAsync(G.d_sys, G.sys.data(), bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? 0 : hip_fail(e, "system table upload");
}
template int upload_table<double>(Group<double> &, hipStream_t, char *, size_t);
template int upload_table<float>(Group<float> &, hipStream_t, char *, size_t);

// ---- launch sequences -------------------------------------------------------------

template <typename T, int F, int NV>
static void finish(const Group<T> &G, const MapSet &M, hipStream_t s) {
    hipLaunchKernelGGL((k_finish_sys<T, F, NV>), dim3((unsigned)M.fmap.size()), dim3(kFinishBlock), 0, s, G.d_sys,
                       M.d_fmap);
}

template <int BS, typename T>
static void setup_bs(Group<T> &G, const MapSet &M, hipStream_t s) {
    const dim3 gs((unsigned)M.smap.size()), blk(BS);
    const int nids = (int)M.ids.size();
    hipLaunchKernelGGL(k_init_states<T>, dim3((unsigned)((nids + 255) / 256)), dim3(256), 0, s, G.d_sys, M.d_ids,
                       nids);
    hipLaunchKernelGGL((k_setup_residual<BS, T>), gs, blk, 0, s, G.d_sys, M.d_smap);
    finish<T, FIN_SETUP, 1>(G, M, s);
    if (G.mixed) hipLaunchKernelGGL((k_prepare_a32<BS>), gs, blk, 0, s, (const Sys<float> *)G.d_sys, M.d_smap);
    hipLaunchKernelGGL((k_shadow<BS, T>), gs, blk, 0, s, G.d_sys, M.d_smap);
    finish<T, FIN_SHADOW, 1>(G, M, s);
}

template <typename T>
void enqueue_setup(Group<T> &G, const MapSet &M, hipStream_t s) {
    if (M.ids.empty()) return;
    MFIX_DISPATCH_BS(G.bs, BS, setup_bs<BS, T>(G, M, s));
}
template void enqueue_setup<double>(Group<double> &, const MapSet &, hipStream_t);
template void enqueue_setup<float>(Group<float> &, const MapSet &, hipStream_t);

// LINE preconditioner of every LINE system of the table: batched red-black
// passes, then the systems swept one by one.  The start value (Phat = P,
// Shat = S) was written by the producing kernel.
template <typename T>
static void line_sweeps(const Group<T> &G, const MapSet &M, int which, hipStream_t s) {
    Prof &pf = ctx().prof[1];
    for (int q = 0; q < M.nlpass; ++q) {
        for (int param : {1, 0}) {
            // one timed launch per colour and kernel family; the bytes are
            // those of every LINE system of the table (a single solve: its own)
            for (int v = 0; v < kLineLpbVariants; ++v) {
                if (M.lmap[v][q].empty()) continue;
                const auto ev = pf.take(0.5 * M.cells * 13.0 * sizeof(T));
                launch_line_color<T>(v, dim3((unsigned)M.lmap[v][q].size()), s, G.d_sys, M.d_lmap[v][q], param, which,
                                     ev);
            }
            if constexpr (sizeof(T) == sizeof(float)) {
                const Sys<float> *tab = (const Sys<float> *)G.d_sys;
                // the short-line variants in one launch when more than one holds systems
                int nshort = 0;
                PcrMaps pm{};
                for (int v = 0; v < 4; ++v) {
                    pm.m[v] = M.d_pmap[v][q];
                    pm.n[v] = (int)M.pmap[v][q].size();
                    nshort += pm.n[v] > 0 ? 1 : 0;
                }
                if (nshort > 1) {
                    const auto ev = pf.take(0.5 * M.cells * 13.0 * sizeof(T));
                    launch_timed(ev.first, ev.second, k_line_pcr_short,
                                 dim3((unsigned)(pm.n[0] + pm.n[1] + pm.n[2] + pm.n[3])), dim3(256), s, tab, pm,
                                 param, which);
                }
                for (int v = nshort > 1 ? 4 : 0; v < kPcrVariants; ++v) {
                    if (M.pmap[v][q].empty()) continue;
                    const auto ev = pf.take(0.5 * M.cells * 13.0 * sizeof(T));
                    launch_line_pcr(v, dim3((unsigned)M.pmap[v][q].size()), s, tab, M.d_pmap[v][q], param, which,
                                    ev.first, ev.second);
                }
            }
        }
    }
    for (int id : M.line_seq) {
        const Sys<T> &S = G.sys[id];
        const Shape sh{G.bs, 1, 64};
        launch_msolve_line<T>(S.g, sh, G.sweep[id].data(), S.a, which == 0 ? S.p : S.s,
                              which == 0 ? S.phat : S.shat, S.line_scratch, S.st, which, &S.st->msolve_fatal, s,
                              false);
    }
}

// One BiCGSTAB iteration (leq_bicgs.f:316-466) over the table.
template <int BS, typename T>
static void iteration_bs(const Group<T> &G, const MapSet &M, hipStream_t s) {
    const dim3 gs((unsigned)M.smap.size()), blk(BS);
    Prof &pf = ctx().prof[0];
    const bool line = M.n_line > 0;
    const dim3 gblk(BS / Group16<T>::N);  // grouped update kernels: 16 bytes per lane and array
    hipLaunchKernelGGL((k_p_update<BS, T>), gs, gblk, 0, s, G.d_sys, M.d_smap);
    if (line) line_sweeps(G, M, 0, s);
    const auto ev = pf.take(M.k1_bytes);
    if constexpr (sizeof(T) == sizeof(float))
        launch_timed(ev.first, ev.second, k_matvec_dots4<BS, 0>, gs, dim3(BS / 4), s, (const Sys<float> *)G.d_sys,
                     M.d_smap);
    else
        launch_timed(ev.first, ev.second, k_matvec_dots<BS, T, 0>, gs, blk, s, G.d_sys, M.d_smap);
    finish<T, FIN_ALPHA, 1>(G, M, s);
    hipLaunchKernelGGL((k_s_update<BS, T>), gs, gblk, 0, s, G.d_sys, M.d_smap);
    if (line) line_sweeps(G, M, 1, s);
    if constexpr (sizeof(T) == sizeof(float))
        hipLaunchKernelGGL((k_matvec_dots4<BS, 1>), gs, dim3(BS / 4), 0, s, (const Sys<float> *)G.d_sys, M.d_smap);
    else
        hipLaunchKernelGGL((k_matvec_dots<BS, T, 1>), gs, blk, 0, s, G.d_sys, M.d_smap);
    finish<T, FIN_OMEGA, 3>(G, M, s);
    hipLaunchKernelGGL((k_xr_update<BS, T>), gs, gblk, 0, s, G.d_sys, M.d_smap);
    if (G.mixed) {
        finish<T, FIN_XR_MIXED, 2>(G, M, s);
        // (its last block applies the update's scalar logic, FIN_UPDATE)
        hipLaunchKernelGGL((k_reliable_update<BS>), dim3((unsigned)M.umap.size()), blk, 0, s,
                           (const Sys<float> *)G.d_sys, M.d_umap);
    } else {
        finish<T, FIN_XR, 2>(G, M, s);
    }
}

template <typename T>
void enqueue_iterations(Group<T> &G, const MapSet &M, int n, hipStream_t s) {
    if (M.ids.empty()) return;
    MFIX_DISPATCH_BS(G.bs, BS, for (int q = 0; q < n; ++q) iteration_bs<BS, T>(G, M, s));
}
template void enqueue_iterations<double>(Group<double> &, const MapSet &, int, hipStream_t);
template void enqueue_iterations<float>(Group<float> &, const MapSet &, int, hipStream_t);

template <typename T>
static void enqueue_count(const Group<T> &G, const MapSet &M, int *d_out, int *d_flags, hipStream_t s) {
    hipLaunchKernelGGL(k_count_running<T>, dim3(1), dim3(256), 0, s, G.d_sys, M.d_ids, (int)M.ids.size(), d_out,
                       d_flags);
}

// Line-kernel launch of a LINE system: 1 + its variant index among the
// red-black kernels (the FP32 several-lanes-per-line variants, or the FP64
// lines-per-block variants), 0 for other systems and one-by-one sweeps.
template <typename T>
static int line_launch_of(const Group<T> &G, int id) {
    const Sys<T> &S = G.sys[id];
    if (S.pc != PC_LINE || rb_passes(G.sweep[id], S.g.do_k != 0) < 0) return 0;
    if (sizeof(T) == sizeof(float)) {
        const int var = line_pcr_variant(line_geom(S.g, AX_J).n);
        if (var >= 0) return 1 + kLineLpbVariants + var;
    }
    return 1 + line_lpb_variant(ctx().shape_for(S.g.n3, sizeof(T) == sizeof(float)).lines_per_block);
}

// Drop finished systems from a table's block maps once the running systems
// hold less than 60 % of the mapped cells, or once every system of one of
// the line-kernel launches is done (each colour phase launches every line
// kernel variant that holds a system, so a variant of finished systems would
// dispatch a launch of idle blocks 8 times per iteration): later launches
// stop dispatching their blocks (a handful of rebuilds per batch).
template <typename T>
static int compact(Group<T> &G, const int *d_flags, int running, hipStream_t s) {
    MapSet &M = G.iter;
    if (running >= (int)M.ids.size() || running == 0) return 0;
    std::vector<int> flags(M.ids.size());
    hipError_t e = hipMemcpyAsync(flags.data(), d_flags, sizeof(int) * flags.size(), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "running flags");
    std::vector<int> keep;
    double cells = 0.0;
    constexpr int kLaunches = 1 + kLineLpbVariants + kPcrVariants;
    bool mapped[kLaunches] = {}, running_in[kLaunches] = {};
    for (size_t q = 0; q < flags.size(); ++q) {
        const int l = line_launch_of(G, M.ids[q]);
        mapped[l] = true;
        if (flags[q]) {
            keep.push_back(M.ids[q]);
            cells += (double)G.sys[M.ids[q]].g.n3;
            running_in[l] = true;
        }
    }
    bool idle_launch = false;
    for (int l = 1; l < kLaunches; ++l) idle_launch |= mapped[l] && !running_in[l];
    if (cells >= 0.6 * M.cells && !idle_launch) return 0;
    return build_maps<T>(G, keep, M, s);
}

// Chunk of iterations of both tables, captured once and replayed.
struct ChunkGraph {
    hipGraphExec_t exec = nullptr;
    int iters = 0;
    uint64_t v64 = 0, v32 = 0;
    const void *s64 = nullptr, *s32 = nullptr;
    const void *m64 = nullptr, *m32 = nullptr;
};

static int capture_chunk(Group<double> *G64, Group<float> *G32, int iters, hipStream_t s, hipGraphExec_t *out) {
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
    for (int q = 0; q < iters; ++q) {
        if (G64) enqueue_iterations<double>(*G64, G64->iter, 1, s);
        if (G32) enqueue_iterations<float>(*G32, G32->iter, 1, s);
    }
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(s, &graph);
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    return 0;
}

// Drive both tables (either may be null) until every system is done.  The
// running count is polled after each chunk; chunks grow 4, 8, ... 64
// iterations (direct launches) or are a captured graph of 16, never more
// than ITMAX.  Launches after a system is done return at once, so
// overshooting is harmless, but each still dispatches its blocks.
static int run_tables(Group<double> *G64, Group<float> *G32, int max_iters, hipStream_t s, bool graphs,
                      int *d_count = nullptr) {
    Context &C = ctx();
    int rc;
    if (!d_count) {
        if ((rc = C.misc.ensure(kCountBytes))) return rc;
        d_count = (int *)C.misc.ptr;
    }
    int *d_cnt = d_count;
    int *d_flags64 = (int *)((char *)d_count + 256);
    int *d_flags32 = d_flags64 + kHostStates;
    if (G64 && G64->iter.ids.empty()) G64 = nullptr;
    if (G32 && G32->iter.ids.empty()) G32 = nullptr;
    if (!G64 && !G32) return 0;
    if (C.prof[0].on || C.prof[1].on) graphs = false;
    // one-by-one LINE sweeps bake their arrays into the launches: no replay
    if ((G64 && !G64->iter.line_seq.empty()) || (G32 && !G32->iter.line_seq.empty())) graphs = false;
    // graph owner: the float table's slot when present, else the double one
    hipGraphExec_t *exec = G32 ? &G32->exec : &G64->exec;
    int *exec_iters = G32 ? &G32->exec_iters : &G64->exec_iters;
    uint64_t *exec_version = G32 ? &G32->exec_version : &G64->exec_version;
    const void **exec_sys = G32 ? &G32->exec_sys : &G64->exec_sys;
    const void *sys_key = G32 ? (const void *)G32->d_sys : (const void *)G64->d_sys;
    // no chunk runs past ITMAX: with MFIX's LEQ_IT of 5-20 a fixed 16 would
    // dispatch every block of up to 11 idle iterations per solve
    const int cap = std::max(1, std::min(max_iters, 64));
    const int kGraphIters = std::min(16, cap);
    int64_t issued = 0;
    int chunk = std::min(4, cap);
    for (;;) {
        if (G64) enqueue_count<double>(*G64, G64->iter, d_cnt + 0, d_flags64, s);
        if (G32) enqueue_count<float>(*G32, G32->iter, d_cnt + 1, d_flags32, s);
        hipError_t e = hipMemcpyAsync(C.host_count, d_cnt, 2 * sizeof(int), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_fail(e, "running count");
        const int running = (G64 ? C.host_count[0] : 0) + (G32 ? C.host_count[1] : 0);
        if (running == 0) return 0;
        if (issued > 4 * ((int64_t)max_iters + 8)) return hip_fail(hipErrorUnknown, "solver did not terminate");
        if (G64 && (rc = compact<double>(*G64, d_flags64, C.host_count[0], s))) return rc;
        if (G32 && (rc = compact<float>(*G32, d_flags32, C.host_count[1], s))) return rc;
        if (G64 && G64->iter.ids.empty()) G64 = nullptr;
        if (G32 && G32->iter.ids.empty()) G32 = nullptr;
        const uint64_t version = (G64 ? G64->iter.version * 1000003ull : 0) ^ (G32 ? G32->iter.version : 0) ^
                                 ((G64 ? 1ull : 0) << 62) ^ ((G32 ? 1ull : 0) << 63);
        if (graphs) {
            if (!*exec || *exec_iters != kGraphIters || *exec_version != version || *exec_sys != sys_key) {
                if (*exec) (void)hipGraphExecDestroy(*exec);
                *exec = nullptr;
                if ((rc = capture_chunk(G64, G32, kGraphIters, s, exec))) return rc;
                *exec_iters = kGraphIters;
                *exec_version = version;
                *exec_sys = sys_key;
            }
            e = hipGraphLaunch(*exec, s);
            if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
            issued += kGraphIters;
        } else {
            for (int q = 0; q < chunk; ++q) {
                if (G64) enqueue_iterations<double>(*G64, G64->iter, 1, s);
                if (G32) enqueue_iterations<float>(*G32, G32->iter, 1, s);
            }
            issued += chunk;
            chunk = std::min(chunk * 2, cap);
        }
    }
}

template <typename T>
int run_group(Group<T> &G, int max_iters, hipStream_t s, bool graphs, int *d_count) {
    if constexpr (sizeof(T) == sizeof(double)) return run_tables(&G, nullptr, max_iters, s, graphs, d_count);
    else return run_tables(nullptr, &G, max_iters, s, graphs, d_count);
}
template int run_group<double>(Group<double> &, int, hipStream_t, bool, int *);
template int run_group<float>(Group<float> &, int, hipStream_t, bool, int *);

template <typename T>
int read_states(const Group<T> &G, const std::vector<int> &ids, SolverState *h, hipStream_t s) {
    for (size_t q = 0; q < ids.size(); ++q) {
        hipError_t e = hipMemcpyAsync(h + q, G.sys[ids[q]].st, sizeof(SolverState), hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) return hip_fail(e, "state readback");
    }
    hipError_t e = hipStreamSynchronize(s);
    return e == hipSuccess ? 0 : hip_fail(e, "state readback");
}
template int read_states<double>(const Group<double> &, const std::vector<int> &, SolverState *, hipStream_t);
template int read_states<float>(const Group<float> &, const std::vector<int> &, SolverState *, hipStream_t);

void fill_stats(const SolverState &h, mfix_leq_stats *st) {
    if (!st) return;
    st->ier = h.ier;
    st->exit_reason = h.exit_reason;
    if (h.exit_reason == EXIT_ZERO_ROW) st->iterations = 0;
    else if (h.mixed) st->iterations = h.iter_var - 1;  // inner iterations completed (FIN_XR_MIXED counts each)
    else if (h.exit_reason == EXIT_RHO) st->iterations = h.i - 1;
    else st->iterations = h.i;
    st->iter_tot_increment = h.iter_tot_inc;
    st->rnorm0 = h.rnorm0;
    st->rnorm = h.rnorm;
    st->true_rnorm = h.mixed ? h.true_rnorm : -1.0;
    st->refinements = h.updates;
    st->restarts = h.restarts;
    st->inner_iterations = h.mixed ? st->iterations : 0;
}

// ---- single solves (a table of one) -------------------------------------------

static Sweep make_sweep(const char *cmethod) {
    Sweep sw = {'R', 'S', 'R', 'S'};
    if (cmethod)
        for (int q = 0; q < 4 && cmethod[q]; ++q) sw[q] = cmethod[q];
    return sw;
}

template <typename T>
static int solve_one(Group<T> &G, const SysDesc &d, const Shape &sh, mfix_leq_stats *stats, hipStream_t s) {
    Context &C = ctx();
    const Part P = make_part(d.g.n3, sh.bs, sh.cpt);
    const bool mixed = sizeof(T) == sizeof(float);
    int rc;
    if ((rc = C.ws.ensure(sys_bytes(d.g, P, d.pc, mixed)))) return rc;
    G.sys.resize(1);
    G.sweep.assign(1, d.sweep);
    G.bs = sh.bs;
    G.mixed = mixed;
    sys_carve<T>(G.sys[0], d, P, (char *)C.ws.ptr);
    // residual of the equation as passed, formed by the set-up pass: its
    // partials and result in the context's partials buffer
    const bool resid = d.pro.flag && d.pro.resid_kind >= 0 && d.resid_out;
    mfix_leq_resid *d_rout = nullptr;
    if (resid) {
        const size_t pb = sizeof(double) * kResidVals * (size_t)P.nblk;
        if ((rc = C.partials.ensure(pb + 256 + sizeof(mfix_leq_resid)))) return rc;
        G.sys[0].pro.rpart = (double *)C.partials.ptr;
        d_rout = (mfix_leq_resid *)(((uintptr_t)C.partials.ptr + pb + 255) & ~(uintptr_t)255);
    }
    if ((rc = upload_table<T>(G, s))) return rc;
    if ((rc = build_maps<T>(G, std::vector<int>{0}, G.iter, s))) return rc;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, s);
    enqueue_setup<T>(G, G.iter, s);
    if (resid) launch_resid_finish(d.g, d.pro.resid_kind, d.resid_debug, d.resid_norm, G.sys[0].pro.rpart, P.nblk,
                                   d_rout, s);
    const bool graphs = C.use_graphs && d.g.n3 <= ((int64_t)1 << 22);
    rc = run_group<T>(G, d.itmax, s, graphs);
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    SolverState *h = C.host_state;
    if (rc || (rc = read_states<T>(G, std::vector<int>{0}, h, s))) {
        for (auto &p : C.prof) p.collect();
        return rc;
    }
    // a one-system table: its first launches did the work, later ones
    // returned at once (the solve was done)
    C.prof[0].collect(h->k1_launches);
    C.prof[1].collect(h->line_launches);
    if (resid) {
        hipError_t e = hipMemcpyAsync(d.resid_out, d_rout, sizeof(mfix_leq_resid), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_fail(e, "residual copy");
    }
    if (mixed && h->x_cur == 1) {
        hipError_t e = hipMemcpyAsync(d.x, G.sys[0].xb, sizeof(double) * d.g.n3, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_fail(e, "solution copy");
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    fill_stats(*h, stats);
    if (stats) stats->device_ms = ms;
    if (h->msolve_fatal) return MFIX_LEQ_EDGTSV;
    return 0;
}

static SysDesc make_desc(const DGrid &g, int pc, const char *cmethod, double tol, int itmax, double *x, double *a,
                         double *b, const mfix_leq_opts *opts, int minimize, int icheck, double delta) {
    SysDesc d;
    d.g = g;
    d.pc = pc;
    d.sweep = make_sweep(cmethod);
    d.tol = tol;
    d.itmax = itmax;
    d.minimize = minimize;
    d.icheck = icheck;
    d.delta = delta;
    d.max_updates = opts && opts->max_refine > 0 ? opts->max_refine : 0;
    d.x = x;
    d.a = a;
    d.b = b;
    d.noise = opts ? opts->noise : nullptr;
    d.seed = opts ? opts->noise_seed : 0;
    return d;
}

int solve_fp64(const DGrid &g, int pc, const char *cmethod, double tol, int itmax, double *x, double *a, double *b,
               const mfix_leq_opts *opts, mfix_leq_stats *stats, hipStream_t s, int minimize, int icheck) {
    const SysDesc d = make_desc(g, pc, cmethod, tol, itmax, x, a, b, opts, minimize, icheck, 0.0);
    return solve_one<double>(ctx().g64, d, ctx().shape_for(g.n3, false), stats, s);
}

int solve_mixed(const DGrid &g, int pc, const char *cmethod, double tol, int itmax, double *x, double *a, double *b,
                const mfix_leq_opts *opts, mfix_leq_stats *stats, hipStream_t s) {
    const double delta = opts && opts->inner_tol > 0.0 ? opts->inner_tol : mixed_delta(g.n3);
    const SysDesc d = make_desc(g, pc, cmethod, tol, itmax, x, a, b, opts, 0, 1, delta);
    return solve_one<float>(ctx().g32, d, ctx().shape_for(g.n3, true), stats, s);
}

// A solve whose set-up pass first forms the CALC_RESID row and applies
// UNDER_RELAX_S (pro), FP64 or mixed; the residual's result goes to
// *resid_out.
int solve_with_prologue(const DGrid &g, int pc, const char *cmethod, double tol, int itmax, double *x, double *a,
                        double *b, const mfix_leq_opts *opts, mfix_leq_stats *stats, hipStream_t s, int mixed,
                        int minimize, int icheck, const Prologue &pro, int resid_debug, double resid_norm,
                        mfix_leq_resid *resid_out) {
    const double delta = mixed ? (opts && opts->inner_tol > 0.0 ? opts->inner_tol : mixed_delta(g.n3)) : 0.0;
    SysDesc d = mixed ? make_desc(g, pc, cmethod, tol, itmax, x, a, b, opts, 0, 1, delta)
                      : make_desc(g, pc, cmethod, tol, itmax, x, a, b, opts, minimize, icheck, 0.0);
    d.pro = pro;
    d.resid_out = resid_out;
    d.resid_debug = resid_debug;
    d.resid_norm = resid_norm;
    if (mixed) return solve_one<float>(ctx().g32, d, ctx().shape_for(g.n3, true), stats, s);
    return solve_one<double>(ctx().g64, d, ctx().shape_for(g.n3, false), stats, s);
}

// ---- batched solves ---------------------------------------------------------------

// Every item becomes a row of the FP64 or the FP32 table; both tables
// advance together, one launch per phase covering all their systems, in
// captured chunks.  Threads per block are shared by a table; cells per
// thread follow each system's size.
int solve_batch(int nsys, const mfix_leq_batch_item *items, mfix_leq_stats *stats, hipStream_t s,
                const mfix_leq_grid *grids) {
    Context &C = ctx();
    int rc;
    std::vector<SysDesc> descs((size_t)nsys);
    std::vector<Part> parts((size_t)nsys);
    std::vector<size_t> offs((size_t)nsys);
    size_t total = 0;
    const int bs = 256;
    for (int q = 0; q < nsys; ++q) {
        const mfix_leq_batch_item &it = items[q];
        mfix_leq_grid hg = {};
        if (grids) {
            hg = grids[q];
        } else {
            hg.imax = it.imax;
            hg.jmax = it.jmax;
            hg.kmax = it.kmax;
            hg.is_serial = 1;
            hg.icheck_bicgs = 1;
        }
        const DGrid g = make_dgrid(hg);
        const int pc = parse_pc(it.pc);
        char cm[5] = {it.cmethod[0], it.cmethod[1], it.cmethod[2], it.cmethod[3], 0};
        // the FP64 rows follow the grid's MINIMIZE_DOTPRODUCTS / ICHECK_BICGS
        // as solve_fp64 does; the mixed rows use the fixed schedule of solve_mixed
        descs[q] = make_desc(g, pc, cm, it.tol, it.itmax, it.d_var, it.d_a_m, it.d_b_m, nullptr,
                             it.mixed ? 0 : hg.minimize_dotproducts, it.mixed ? 1 : std::max(1, hg.icheck_bicgs), mixed_delta(g.n3));
        descs[q].seed = it.noise_seed;
        const Shape sh = C.shape_for(g.n3, it.mixed != 0);
        parts[q] = make_part(g.n3, bs, sh.bs == bs ? sh.cpt : 1);
        offs[q] = total;
        total += sys_bytes(g, parts[q], pc, it.mixed != 0);
    }
    if ((rc = C.bws.ensure(std::max<size_t>(total, 256)))) return rc;
    Group<double> &G64 = C.b64;
    Group<float> &G32 = C.b32;
    G64.sys.clear();
    G64.sweep.clear();
    G32.sys.clear();
    G32.sweep.clear();
    G64.bs = G32.bs = bs;
    G64.mixed = false;
    G32.mixed = true;
    std::vector<std::pair<int, int>> where((size_t)nsys);  // (precision, row)
    std::vector<int> ids64, ids32;
    for (int q = 0; q < nsys; ++q) {
        char *base = (char *)C.bws.ptr + offs[q];
        if (items[q].mixed) {
            G32.sys.emplace_back();
            G32.sweep.push_back(descs[q].sweep);
            sys_carve<float>(G32.sys.back(), descs[q], parts[q], base);
            where[q] = {1, (int)G32.sys.size() - 1};
            ids32.push_back((int)G32.sys.size() - 1);
        } else {
            G64.sys.emplace_back();
            G64.sweep.push_back(descs[q].sweep);
            sys_carve<double>(G64.sys.back(), descs[q], parts[q], base);
            where[q] = {0, (int)G64.sys.size() - 1};
            ids64.push_back((int)G64.sys.size() - 1);
        }
    }
    int max_it = 1;
    for (int q = 0; q < nsys; ++q) max_it = std::max(max_it, items[q].itmax);
    if ((rc = upload_table<double>(G64, s)) || (rc = upload_table<float>(G32, s))) return rc;
    if ((rc = build_maps<double>(G64, ids64, G64.iter, s)) || (rc = build_maps<float>(G32, ids32, G32.iter, s)))
        return rc;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, s);
    enqueue_setup<double>(G64, G64.iter, s);
    enqueue_setup<float>(G32, G32.iter, s);
    rc = run_tables(&G64, &G32, max_it, s, C.use_graphs != 0);
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (auto &p : C.prof) p.collect();
    if (rc) return rc;
    if (nsys > kHostStates) return MFIX_LEQ_EINVAL;
    SolverState *h64 = C.host_state;
    SolverState *h32 = C.host_state + ids64.size();
    if ((rc = read_states<double>(G64, ids64, h64, s)) || (rc = read_states<float>(G32, ids32, h32, s))) return rc;
    int fatal = 0;
    for (int q = 0; q < nsys; ++q) {
        const SolverState &h = where[q].first ? h32[where[q].second] : h64[where[q].second];
        if (where[q].first && h.x_cur == 1) {
            const Sys<float> &S = G32.sys[where[q].second];
            hipError_t e = hipMemcpyAsync(S.xa, S.xb, sizeof(double) * S.g.n3, hipMemcpyDeviceToDevice, s);
            if (e != hipSuccess) return hip_fail(e, "solution copy");
        }
        if (stats) {
            fill_stats(h, &stats[q]);
            stats[q].device_ms = ms;
        }
        fatal |= h.msolve_fatal;
    }
    hipError_t e = hipStreamSynchronize(s);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "batch");
    return fatal ? MFIX_LEQ_EDGTSV : 0;
}

// ---- standalone kernels of the path -----------------------------------------

int run_matvec64(const DGrid &g, const double *a, const double *x, double *y, hipStream_t s) {
    const Shape sh = ctx().shape_for(g.n3, false);
    const Part P = make_part(g.n3, sh.bs, sh.cpt);
    MFIX_DISPATCH_BS(sh.bs, BS, hipLaunchKernelGGL((k_matvec<BS, double>), grid_of(P), dim3(BS), 0, s, g, P, a, x, y));
    return 0;
}

// The FP32 row of the inner solve through the solver's own first stencil
// pass (k_matvec_dots4, operand x, result y; its dot partial goes to
// scratch): a one-system table with a blank state, so the kernel never
// returns early, and every logical block mapped in order.
int run_matvec32(const DGrid &g, const float *a, const float *x, float *y, hipStream_t s) {
    Context &C = ctx();
    const Shape sh = C.shape_for(g.n3, true);
    const Part P = make_part(g.n3, sh.bs, 1);
    const size_t o_st = align256(sizeof(Sys<float>));
    const size_t o_part = o_st + align256(sizeof(SolverState));
    const size_t o_map = o_part + align256(sizeof(double) * kMaxPartials * (size_t)P.nblk);
    int rc;
    if ((rc = C.line1.ensure(o_map + sizeof(int2) * (size_t)P.nblk))) return rc;
    char *base = (char *)C.line1.ptr;
    Sys<float> S = {};
    S.g = g;
    S.P = P;
    S.pc = PC_NONE;
    S.a = a;
    S.p = const_cast<float *>(x);
    S.rt = const_cast<float *>(x);
    S.v = y;
    S.st = (SolverState *)(base + o_st);
    S.partials = (double *)(base + o_part);
    std::vector<int2> map((size_t)P.nblk);
    for (int64_t b = 0; b < P.nblk; ++b) map[(size_t)b] = make_int2(0, (int)b);
    hipError_t e = hipMemcpyAsync(base, &S, sizeof(S), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(base + o_map, map.data(), sizeof(int2) * map.size(), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "matvec table upload");
    hipLaunchKernelGGL(k_init_state, dim3(1), dim3(1), 0, s, S.st);
    const Sys<float> *tab = (const Sys<float> *)base;
    const int2 *mp = (const int2 *)(base + o_map);
    MFIX_DISPATCH_BS(sh.bs, BS,
                     hipLaunchKernelGGL((k_matvec_dots4<BS, 0>), dim3((unsigned)P.nblk), dim3(BS / 4), 0, s, tab, mp));
    // the table must outlive the launch before it is rewritten
    e = hipStreamSynchronize(s);
    return e == hipSuccess ? 0 : hip_fail(e, "matvec32");
}

int run_diag(const DGrid &g, const double *a, const double *r, double *z, hipStream_t s) {
    const Shape sh = ctx().shape_for(g.n3, false);
    const Part P = make_part(g.n3, sh.bs, sh.cpt);
    MFIX_DISPATCH_BS(sh.bs, BS, hipLaunchKernelGGL((k_diag<BS, double>), grid_of(P), dim3(BS), 0, s, g, P, a, r, z));
    return 0;
}

int run_dot(const DGrid &g, const double *x, const double *y, double *d_out, hipStream_t s) {
    Context &C = ctx();
    const Shape sh = C.shape_for(g.n3, false);
    const Part P = make_part(g.n3, sh.bs, sh.cpt);
    int rc;
    if ((rc = C.partials.ensure(sizeof(double) * P.nblk * kMaxPartials + 64))) return rc;
    if ((rc = C.state.ensure(sizeof(SolverState)))) return rc;
    SolverState *st = (SolverState *)C.state.ptr;
    hipLaunchKernelGGL(k_init_state, dim3(1), dim3(1), 0, s, st);
    MFIX_DISPATCH_BS(sh.bs, BS,
                     hipLaunchKernelGGL((k_dot<BS, double>), grid_of(P), dim3(BS), 0, s, g, P, x, y, g.dot_all ? 0 : 1,
                                        (double *)C.partials.ptr));
    hipLaunchKernelGGL((k_finish<FIN_DOT, 1>), dim3((unsigned)finish_blocks(P.nblk)), dim3(kFinishBlock), 0, s, st,
                       (const double *)C.partials.ptr, P.nblk, d_out);
    return 0;
}

// Unit scaling alone (LEQ_BICGS0's first step): the set-up kernel of a
// one-system table with x = 0; the residual it also forms is discarded.
int run_scale(const DGrid &g, double *a, double *b, int *zero_row, hipStream_t s) {
    Context &C = ctx();
    const int64_t n3 = g.n3;
    int rc;
    if ((rc = C.tmp.ensure(sizeof(double) * n3))) return rc;
    hipError_t e = hipMemsetAsync(C.tmp.ptr, 0, sizeof(double) * n3, s);
    if (e != hipSuccess) return hip_fail(e, "memset");
    const SysDesc d = make_desc(g, PC_NONE, "RSRS", 0.0, 1, (double *)C.tmp.ptr, a, b, nullptr, 0, 1, 0.0);
    const Shape sh = C.shape_for(n3, false);
    const Part P = make_part(n3, sh.bs, sh.cpt);
    Group<double> &G = C.g64;
    if ((rc = C.ws.ensure(sys_bytes(g, P, PC_NONE, false)))) return rc;
    G.sys.resize(1);
    G.sweep.assign(1, d.sweep);
    G.bs = sh.bs;
    G.mixed = false;
    sys_carve<double>(G.sys[0], d, P, (char *)C.ws.ptr);
    if ((rc = upload_table<double>(G, s)) || (rc = build_maps<double>(G, std::vector<int>{0}, G.iter, s))) return rc;
    const int nids = 1;
    hipLaunchKernelGGL(k_init_states<double>, dim3(1), dim3(256), 0, s, G.d_sys, G.iter.d_ids, nids);
    MFIX_DISPATCH_BS(sh.bs, BS,
                     hipLaunchKernelGGL((k_setup_residual<BS, double>), dim3((unsigned)G.iter.smap.size()), dim3(BS),
                                        0, s, G.d_sys, G.iter.d_smap));
    if ((rc = read_states<double>(G, G.iter.ids, C.host_state, s))) return rc;
    *zero_row = C.host_state->zero_row;
    return 0;
}

int run_copy(const double *src, double *dst, int64_t n, hipStream_t s) {
    const Part P = make_part(n, 256, 1);
    hipLaunchKernelGGL((k_copy_if<256, double>), grid_of(P), dim3(256), 0, s, P, src, dst, (const SolverState *)nullptr,
                       0);
    return 0;
}

// ---- autotuner ------------------------------------------------------------------

// Time 40 iterations of a DIAG solve (8 of a LINE solve) for each candidate
// on a seeded system of the given size and keep the fastest per (size class,
// precision): the
// threads per block of the streaming kernels on a DIAG solve of a
// void-fraction pressure system, then the lines per block of the red-black
// line kernel on a LINE solve of an upwind momentum system.  Candidates are
// timed round-robin, five rounds, and compared by their median time per
// iteration; a candidate replaces the default only when it is more than 2 %
// faster, so a pick does not follow run-to-run noise.
int autotune(int imax, int jmax, int kmax, int mixed, Shape *best_out) {
    Context &C = ctx();
    int rc;
    if ((rc = C.ensure_init())) return rc;
    mfix_leq_grid hg = {};
    hg.imax = imax;
    hg.jmax = jmax;
    hg.kmax = kmax;
    hg.is_serial = 1;
    hg.icheck_bicgs = 1;
    const DGrid g = make_dgrid(hg);
    const int64_t n3 = g.n3;
    std::vector<double> hp((size_t)(7 * n3)), hbp((size_t)n3), hm((size_t)(7 * n3)), hbm((size_t)n3);
    if ((rc = mfix_leq_gen_system(2, imax, jmax, kmax, 4242, 16, hp.data(), hbp.data()))) return rc;
    if ((rc = mfix_leq_gen_system(1, imax, jmax, kmax, 4243, 0, hm.data(), hbm.data()))) return rc;
    double *da = nullptr, *db = nullptr, *dx = nullptr, *dp = nullptr, *dm = nullptr, *dbp = nullptr, *dbm = nullptr;
    hipError_t e;
    const size_t A = sizeof(double) * 7 * (size_t)n3, V = sizeof(double) * (size_t)n3;
    if ((e = hipMalloc(&da, A)) != hipSuccess || (e = hipMalloc(&db, V)) != hipSuccess ||
        (e = hipMalloc(&dx, V)) != hipSuccess || (e = hipMalloc(&dp, A)) != hipSuccess ||
        (e = hipMalloc(&dm, A)) != hipSuccess || (e = hipMalloc(&dbp, V)) != hipSuccess ||
        (e = hipMalloc(&dbm, V)) != hipSuccess) {
        for (double *q : {da, db, dx, dp, dm, dbp, dbm}) (void)hipFree(q);
        return hip_fail(e, "autotune alloc");
    }
    hipStream_t s = C.stream;
    (void)hipMemcpyAsync(dp, hp.data(), A, hipMemcpyHostToDevice, s);
    (void)hipMemcpyAsync(dbp, hbp.data(), V, hipMemcpyHostToDevice, s);
    (void)hipMemcpyAsync(dm, hm.data(), A, hipMemcpyHostToDevice, s);
    (void)hipMemcpyAsync(dbm, hbm.data(), V, hipMemcpyHostToDevice, s);
    const int cls = size_class(n3);
    Shape &slot = C.tuned[mixed ? 1 : 0][cls];
    const Shape saved = slot;
    // median time per iteration of each candidate, timed round-robin
    auto pick = [&](int pc, const double *a0, const double *b0, int ncand, auto set_cand) -> int {
        constexpr int kRounds = 5;
        const int iters = pc == PC_LINE ? 8 : 40;  // set-up under 10 % of a DIAG candidate at 256^3
        std::vector<std::vector<double>> t((size_t)ncand);
        for (int round = 0; round < kRounds && !rc; ++round)
            for (int c = 0; c < ncand && !rc; ++c) {
                set_cand(c);
                (void)hipMemcpyAsync(da, a0, A, hipMemcpyDeviceToDevice, s);
                (void)hipMemcpyAsync(db, b0, V, hipMemcpyDeviceToDevice, s);
                (void)hipMemsetAsync(dx, 0, V, s);
                mfix_leq_stats st = {};
                rc = mixed ? solve_mixed(g, pc, "RSRS", 0.0, iters, dx, da, db, nullptr, &st, s)
                           : solve_fp64(g, pc, "RSRS", 0.0, iters, dx, da, db, nullptr, &st, s, 0, 1);
                t[(size_t)c].push_back(st.device_ms / std::max(1, st.iterations));
            }
        int best = 0;
        double best_t = 0.0;
        for (int c = 0; c < ncand && !rc; ++c) {
            std::vector<double> v = t[(size_t)c];
            std::nth_element(v.begin(), v.begin() + v.size() / 2, v.end());
            const double med = v[v.size() / 2];
            if (c == 0) best_t = med;
            else if (med < 0.98 * best_t) best = c, best_t = med;
        }
        return best;
    };
    // threads per block (one logical block = one block of the stencil
    // passes; the grouped update kernels cover it with 16-byte cell groups)
    // (the size default first: a candidate replaces it only when faster)
    const int db0 = default_block_size(n3, mixed);
    const int bss[] = {db0, db0 == 256 ? 512 : 256, 1024};
    const int bi = pick(PC_DIAG, dp, dbp, 3, [&](int c) { slot = Shape{bss[c], 1, 0}; });
    Shape best = Shape{bss[bi], 1, 0};
    // lines per block of the red-black line kernel: the size default first
    const int d = default_lines_per_block(n3);
    const int lpbs[] = {d, d == 64 ? 32 : 64, d == 16 ? 32 : 16};
    if (!rc) {
        const int li = pick(PC_LINE, dm, dbm, 3, [&](int c) { slot = Shape{best.bs, 1, lpbs[c]}; });
        best.lines_per_block = lpbs[li];
    }
    slot = rc ? saved : best;
    for (double *q : {da, db, dx, dp, dm, dbp, dbm}) (void)hipFree(q);
    if (rc) return rc;
    if (best_out) *best_out = best;
    return 0;
}

}  // namespace mfix
