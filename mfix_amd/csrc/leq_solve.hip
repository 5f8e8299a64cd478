// Solver engine of the gfx950 LEQ_BICGS path: device context, grow-only
// workspaces, launch shapes (picked per grid size by the autotuner), the
// FP64 driver that follows LEQ_BICGS0 (model/leq_bicgs.f:107-509) and the
// mixed-precision driver (FP32 recurrence, FP64 reliable updates).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/mfix_leq.h"
#include "leq_context.h"
#include "leq_phases.h"

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
        return hip_fail(e, "hipMalloc");
    }
    cap = bytes;
    return 0;
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
}

int Context::ensure_init() {
    if (initialised) return 0;
    return init(0);
}

int Context::init(int dev) {
    if (initialised) return 0;
    hipError_t e = hipSetDevice(dev);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    device = dev;
    e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
    e = hipHostMalloc((void **)&host_state, sizeof(SolverState) * kHostStates, hipHostMallocDefault);
    if (e != hipSuccess) return hip_fail(e, "hipHostMalloc");
    initialised = true;
    return 0;
}

void Context::release_all() {
    for (DevBuf *b : {&d_a, &d_b, &d_x, &ws64, &ws32, &line_scratch, &partials, &state, &a32, &xalt, &misc})
        b->release();
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

void Prof::begin(hipStream_t s) {
    if (!on) return;
    std::pair<hipEvent_t, hipEvent_t> ev;
    if (pool_used < events.size()) {
        ev = events[pool_used];
    } else {
        (void)hipEventCreate(&ev.first);
        (void)hipEventCreate(&ev.second);
        events.push_back(ev);
    }
    (void)hipEventRecord(ev.first, s);
}

void Prof::end(hipStream_t s, double bytes) {
    if (!on) return;
    (void)hipEventRecord(events[pool_used].second, s);
    pending_bytes.push_back(bytes);
    ++pool_used;
}

void Prof::collect() {
    for (size_t q = 0; q < pool_used; ++q) {
        float ms = 0.f;
        if (hipEventSynchronize(events[q].second) == hipSuccess &&
            hipEventElapsedTime(&ms, events[q].first, events[q].second) == hipSuccess) {
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
// fastest per size class and precision; untuned classes use one cell per
// thread in 256-thread blocks, the shape that streamed best when measured.
Shape Context::shape_for(int64_t n3, bool fp32) const {
    const int cls = size_class(n3);
    const Shape &s = tuned[fp32 ? 1 : 0][cls];
    if (s.bs) return s;
    return Shape{256, 1, 64};
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

static inline dim3 grid_of(const Part &P) { return dim3((unsigned)xcd_grid(P.nblk)); }

template <int F, int NV>
void launch_finish(SolverState *st, const double *partials, int64_t nblk, hipStream_t s, double *out) {
    // about one partial per finish thread, at most kFinishBlocks blocks:
    // small grids finish in a single block without atomics
    const int64_t want = (nblk + kFinishBlock - 1) / kFinishBlock;
    const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>(kFinishBlocks, want));
    hipLaunchKernelGGL((k_finish<F, NV>), dim3(blocks), dim3(kFinishBlock), 0, s, st, partials, nblk, out);
}

// ---- LINE preconditioner (LEQ_MSOLVE, leqsol_mod.f:286-598) ---------------

template <typename T>
int launch_msolve_line(const DGrid &g, const Shape &sh, const char *cmethod, const T *a, const T *r, T *z,
                       T *scratch, const SolverState *st, int skip_on_sexit, int *fatal_flag, hipStream_t s) {
    const int64_t n3 = g.n3;
    T *sdl = scratch, *sd = scratch + n3, *sdu = scratch + 2 * n3, *sb = scratch + 3 * n3;
    Prof &pf = ctx().prof[1];
    const Part P = make_part(n3, sh.bs, sh.cpt);
    MFIX_DISPATCH_BS(sh.bs, BS,
                     hipLaunchKernelGGL((k_copy_if<BS, T>), grid_of(P), dim3(BS), 0, s, P, r, z, st, skip_on_sexit));
    const int lpb = sh.lines_per_block > 0 ? sh.lines_per_block : 64;
    auto run = [&](const LinePhase &ph, int64_t nthreads) {
        const unsigned blocks = (unsigned)((nthreads + lpb - 1) / lpb);
        const LineGeom L = line_geom(g, ph.axis);
        // algorithmic bytes of one phase: each line solved reads A (7
        // planes), the right-hand side and four neighbour values and writes
        // the line once
        const double cells = (double)std::min<int64_t>(nthreads, (int64_t)L.nu * L.nv) * L.n *
                             (ph.kind == PH_COLOR ? 0.5 : 1.0);
        pf.begin(s);
        hipLaunchKernelGGL(k_line_phase<T>, dim3(blocks ? blocks : 1), dim3(lpb), 0, s, g, ph, a, r, z, sdl, sd, sdu,
                           sb, st, skip_on_sexit, fatal_flag);
        pf.end(s, cells * (7.0 * sizeof(T) + 6.0 * sizeof(T)));
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

template int launch_msolve_line<double>(const DGrid &, const Shape &, const char *, const double *, const double *,
                                        double *, double *, const SolverState *, int, int *, hipStream_t);
template int launch_msolve_line<float>(const DGrid &, const Shape &, const char *, const float *, const float *,
                                       float *, float *, const SolverState *, int, int *, hipStream_t);

// Per-iteration algorithmic bytes per cell of the fused three-pass schedule.
double alg_bytes_per_cell(int pc, int sa, int sv, int sx) {
    if (pc == PC_DIAG) return 15.0 * sa + 15.0 * sv + 2.0 * sx;
    if (pc == PC_NONE) return 14.0 * sa + 15.0 * sv + 2.0 * sx;
    return 42.0 * sa + 30.0 * sv + 2.0 * sx;
}

// One BiCGSTAB iteration (leq_bicgs.f:316-466) of the plain FP64 or the inner
// FP32 solve.
template <int BS, typename T>
static void iteration(const SolveBufs<T> &w, const Shape &sh, int pc, const char *cmethod, hipStream_t s) {
    const DGrid &g = w.g;
    const Part P = make_part(g.n3, BS, sh.cpt);
    const dim3 grid = grid_of(P), blk(BS);
    Prof &pf = ctx().prof[0];
    const T *a0 = w.a + (int64_t)DiagPlane<T>::value * g.n3;
    // algorithmic bytes of one stencil launch: 7 coefficient planes, the
    // operand, the dot partner and the result, once each
    const double mv_bytes = (double)g.n3 * (10.0 * sizeof(T));
    hipLaunchKernelGGL((k_p_update<BS, T>), grid, blk, 0, s, g, P, w.r, w.p, w.v, w.rt, w.st);
    const T *phat = w.p;
    if (pc == PC_LINE) {
        launch_msolve_line<T>(g, sh, cmethod, w.a, w.p, w.phat, w.line_scratch, w.st, 0, &w.st->msolve_fatal, s);
        phat = w.phat;
    }
    pf.begin(s);
    if (pc == PC_DIAG)
        hipLaunchKernelGGL((k_matvec_dots<BS, T, 0, true>), grid, blk, 0, s, g, P, w.a, phat, w.v, w.rt, w.st,
                           w.partials);
    else
        hipLaunchKernelGGL((k_matvec_dots<BS, T, 0, false>), grid, blk, 0, s, g, P, w.a, phat, w.v, w.rt, w.st,
                           w.partials);
    pf.end(s, mv_bytes);
    launch_finish<FIN_ALPHA, 1>(w.st, w.partials, P.nblk, s, nullptr);
    hipLaunchKernelGGL((k_s_update<BS, T>), grid, blk, 0, s, g, P, w.r, w.v, w.s, w.st, w.partials);
    launch_finish<FIN_SNORM, 1>(w.st, w.partials, P.nblk, s, nullptr);
    const T *shat = w.s;
    if (pc == PC_LINE) {
        launch_msolve_line<T>(g, sh, cmethod, w.a, w.s, w.shat, w.line_scratch, w.st, 1, &w.st->msolve_fatal, s);
        shat = w.shat;
    }
    pf.begin(s);
    if (pc == PC_DIAG)
        hipLaunchKernelGGL((k_matvec_dots<BS, T, 1, true>), grid, blk, 0, s, g, P, w.a, shat, w.t, w.s, w.st,
                           w.partials);
    else
        hipLaunchKernelGGL((k_matvec_dots<BS, T, 1, false>), grid, blk, 0, s, g, P, w.a, shat, w.t, w.s, w.st,
                           w.partials);
    pf.end(s, mv_bytes);
    launch_finish<FIN_OMEGA, 2>(w.st, w.partials, P.nblk, s, nullptr);
    if (pc == PC_DIAG)
        hipLaunchKernelGGL((k_xr_update<BS, T, true>), grid, blk, 0, s, g, P, w.x, (const T *)w.p, (const T *)w.s, a0,
                           w.s, w.t, w.r, w.rt, w.st, w.partials);
    else
        hipLaunchKernelGGL((k_xr_update<BS, T, false>), grid, blk, 0, s, g, P, w.x, phat, shat, a0, w.s, w.t, w.r,
                           w.rt, w.st, w.partials);
    if (w.mixed) launch_finish<FIN_XR_MIXED, 2>(w.st, w.partials, P.nblk, s, nullptr);
    else launch_finish<FIN_XR, 2>(w.st, w.partials, P.nblk, s, nullptr);
}

template <typename T>
void enqueue_iteration(const SolveBufs<T> &w, const Shape &sh, int pc, const char *cmethod, hipStream_t s) {
    MFIX_DISPATCH_BS(sh.bs, BS, iteration<BS, T>(w, sh, pc, cmethod, s));
}
template void enqueue_iteration<double>(const SolveBufs<double> &, const Shape &, int, const char *, hipStream_t);
template void enqueue_iteration<float>(const SolveBufs<float> &, const Shape &, int, const char *, hipStream_t);

// Poll the device state every few iterations until the solve is done.  The
// chunk doubles up to 64; launches after convergence return immediately.
int drive(const std::function<void()> &one_iteration, SolverState *d_state, SolverState *h_state, int itmax,
          hipStream_t s) {
    int issued = 0;
    int chunk = 4;
    for (;;) {
        hipError_t e = hipMemcpyAsync(h_state, d_state, sizeof(SolverState), hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) return hip_fail(e, "state readback");
        e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
        if (h_state->done) return 0;
        if (issued > 4 * (itmax + 8)) return hip_fail(hipErrorUnknown, "solver did not terminate");
        const int todo = std::min(chunk, std::max(1, itmax + 2 - h_state->i));
        for (int q = 0; q < todo; ++q) one_iteration();
        issued += todo;
        chunk = std::min(chunk * 2, 64);
    }
}

void fill_stats(const SolverState &h, mfix_leq_stats *st) {
    if (!st) return;
    st->ier = h.ier;
    st->exit_reason = h.exit_reason;
    if (h.exit_reason == EXIT_ZERO_ROW) st->iterations = 0;
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

// ---- FP64 LEQ_BICGS on device arrays --------------------------------------

int solve_fp64(const DGrid &g, int pc, const char *cmethod, double tol, int itmax, double *x, double *a, double *b,
               const mfix_leq_opts *opts, mfix_leq_stats *stats, hipStream_t s, int minimize, int icheck) {
    Context &C = ctx();
    const int64_t n3 = g.n3;
    const Shape sh = C.shape_for(n3, false);
    const Part P = make_part(n3, sh.bs, sh.cpt);
    int rc;
    // R, Rtilde, P, S, T, V, and for LINE Phat, Shat
    const int nvec = pc == PC_LINE ? 8 : 6;
    if ((rc = C.ws64.ensure(sizeof(double) * n3 * nvec))) return rc;
    if ((rc = C.partials.ensure(sizeof(double) * P.nblk * kMaxPartials + 64))) return rc;
    if ((rc = C.state.ensure(sizeof(SolverState)))) return rc;
    if (pc == PC_LINE && (rc = C.line_scratch.ensure(sizeof(double) * n3 * 4))) return rc;
    double *W = (double *)C.ws64.ptr;
    SolveBufs<double> w;
    w.g = g;
    w.a = a;
    w.x = x;
    w.r = W + 0 * n3;
    w.rt = W + 1 * n3;
    w.p = W + 2 * n3;
    w.s = W + 3 * n3;
    w.t = W + 4 * n3;
    w.v = W + 5 * n3;
    w.phat = pc == PC_LINE ? W + 6 * n3 : nullptr;
    w.shat = pc == PC_LINE ? W + 7 * n3 : nullptr;
    w.st = (SolverState *)C.state.ptr;
    w.partials = (double *)C.partials.ptr;
    w.line_scratch = pc == PC_LINE ? (double *)C.line_scratch.ptr : nullptr;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, s);
    // k_init_state rewrites the whole state, so nothing carries over
    hipLaunchKernelGGL(k_init_state, dim3(1), dim3(1), 0, s, w.st, tol, itmax, minimize, icheck, 0, 0.0);
    const double *noise = opts ? opts->noise : nullptr;
    const uint64_t seed = opts ? opts->noise_seed : 0;
    MFIX_DISPATCH_BS(sh.bs, BS, {
        hipLaunchKernelGGL((k_setup_residual<BS>), grid_of(P), dim3(BS), 0, s, g, P, a, b, (const double *)x, w.r,
                           w.st, w.partials);
        launch_finish<FIN_SETUP, 1>(w.st, w.partials, P.nblk, s, nullptr);
        hipLaunchKernelGGL((k_shadow<BS, double>), grid_of(P), dim3(BS), 0, s, g, P, (const double *)w.r, w.r, w.rt,
                           (double *)nullptr, noise, seed, w.st, w.partials);
        launch_finish<FIN_SHADOW, 1>(w.st, w.partials, P.nblk, s, nullptr);
    });
    auto one = [&]() { enqueue_iteration<double>(w, sh, pc, cmethod, s); };
    SolverState *h = C.host_state;
    rc = drive(one, w.st, h, itmax, s);
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (auto &p : C.prof) p.collect();
    if (rc) return rc;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    fill_stats(*h, stats);
    if (stats) stats->device_ms = ms;
    if (h->msolve_fatal) return MFIX_LEQ_EDGTSV;
    return 0;
}

// ---- mixed precision (FP32 recurrence, FP64 reliable updates) -------------

int solve_mixed(const DGrid &g, int pc, const char *cmethod, double tol, int itmax, double *x, double *a, double *b,
                const mfix_leq_opts *opts, mfix_leq_stats *stats, hipStream_t s) {
    Context &C = ctx();
    const int64_t n3 = g.n3;
    const Shape sh = C.shape_for(n3, true);
    const Part P = make_part(n3, sh.bs, sh.cpt);
    int rc;
    // FP32: r, rt, p, s, t, v, y (+ phat, shat for LINE); FP64: r64, second x
    const int nvec = pc == PC_LINE ? 9 : 7;
    if ((rc = C.ws32.ensure(sizeof(float) * n3 * nvec))) return rc;
    if ((rc = C.a32.ensure(sizeof(float) * n3 * 8))) return rc;
    if ((rc = C.xalt.ensure(sizeof(double) * n3 * 2))) return rc;
    if ((rc = C.partials.ensure(sizeof(double) * P.nblk * kMaxPartials + 64))) return rc;
    if ((rc = C.state.ensure(sizeof(SolverState)))) return rc;
    if (pc == PC_LINE && (rc = C.line_scratch.ensure(sizeof(float) * n3 * 4))) return rc;
    float *W = (float *)C.ws32.ptr;
    float *A32 = (float *)C.a32.ptr;
    double *r64 = (double *)C.xalt.ptr;
    double *xb = r64 + n3;  // second FP64 solution buffer
    SolveBufs<float> w;
    w.g = g;
    w.a = A32;
    w.r = W + 0 * n3;
    w.rt = W + 1 * n3;
    w.p = W + 2 * n3;
    w.s = W + 3 * n3;
    w.t = W + 4 * n3;
    w.v = W + 5 * n3;
    w.x = W + 6 * n3;  // y: FP32 correction accumulator
    w.phat = pc == PC_LINE ? W + 7 * n3 : nullptr;
    w.shat = pc == PC_LINE ? W + 8 * n3 : nullptr;
    w.st = (SolverState *)C.state.ptr;
    w.partials = (double *)C.partials.ptr;
    w.line_scratch = pc == PC_LINE ? (float *)C.line_scratch.ptr : nullptr;
    w.mixed = true;
    const double delta = opts && opts->inner_tol > 0.0 ? opts->inner_tol : 0.1;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, s);
    hipLaunchKernelGGL(k_init_state, dim3(1), dim3(1), 0, s, w.st, tol, itmax, 0, 1, 1, delta);
    const double *noise = opts ? opts->noise : nullptr;
    const uint64_t seed = opts ? opts->noise_seed : 0;
    MFIX_DISPATCH_BS(sh.bs, BS, {
        hipLaunchKernelGGL((k_setup_residual<BS>), grid_of(P), dim3(BS), 0, s, g, P, a, b, (const double *)x, r64,
                           w.st, w.partials);
        launch_finish<FIN_SETUP, 1>(w.st, w.partials, P.nblk, s, nullptr);
        hipLaunchKernelGGL((k_prepare_a32<BS>), grid_of(P), dim3(BS), 0, s, g, P, (const double *)a, A32);
        hipLaunchKernelGGL((k_shadow<BS, float>), grid_of(P), dim3(BS), 0, s, g, P, (const double *)r64, w.r, w.rt,
                           w.x, noise, seed, w.st, w.partials);
        launch_finish<FIN_SHADOW, 1>(w.st, w.partials, P.nblk, s, nullptr);
    });
    auto one = [&]() {
        enqueue_iteration<float>(w, sh, pc, cmethod, s);
        MFIX_DISPATCH_BS(sh.bs, BS, {
            hipLaunchKernelGGL((k_reliable_update<BS>), grid_of(P), dim3(BS), 0, s, g, P, (const double *)a,
                               (const double *)b, (const double *)x, (const double *)xb, x, xb, (const float *)w.x,
                               w.r, (const float *)w.rt, w.st, w.partials);
        });
        launch_finish<FIN_UPDATE, 3>(w.st, w.partials, P.nblk, s, nullptr);
    };
    SolverState *h = C.host_state;
    rc = drive(one, w.st, h, itmax, s);
    if (!rc && h->x_cur == 1) {
        hipError_t e = hipMemcpyAsync(x, xb, sizeof(double) * n3, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) rc = hip_fail(e, "hipMemcpyAsync");
    }
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (auto &p : C.prof) p.collect();
    if (rc) return rc;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    fill_stats(*h, stats);
    if (stats) stats->device_ms = ms;
    if (h->msolve_fatal) return MFIX_LEQ_EDGTSV;
    return 0;
}

// ---- standalone kernels of the path -----------------------------------------

int run_matvec64(const DGrid &g, const double *a, const double *x, double *y, hipStream_t s) {
    const Shape sh = ctx().shape_for(g.n3, false);
    const Part P = make_part(g.n3, sh.bs, sh.cpt);
    MFIX_DISPATCH_BS(sh.bs, BS, hipLaunchKernelGGL((k_matvec<BS, double>), grid_of(P), dim3(BS), 0, s, g, P, a, x, y));
    return 0;
}

int run_matvec32(const DGrid &g, const float *a, const float *x, float *y, hipStream_t s) {
    const Shape sh = ctx().shape_for(g.n3, true);
    const Part P = make_part(g.n3, sh.bs, sh.cpt);
    MFIX_DISPATCH_BS(sh.bs, BS, hipLaunchKernelGGL((k_matvec<BS, float>), grid_of(P), dim3(BS), 0, s, g, P, a, x, y));
    return 0;
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
    MFIX_DISPATCH_BS(sh.bs, BS,
                     hipLaunchKernelGGL((k_dot<BS, double>), grid_of(P), dim3(BS), 0, s, g, P, x, y, g.dot_all ? 0 : 1,
                                        (double *)C.partials.ptr));
    launch_finish<FIN_DOT, 1>((SolverState *)C.state.ptr, (const double *)C.partials.ptr, P.nblk, s, d_out);
    return 0;
}

int run_scale(const DGrid &g, double *a, double *b, int *zero_row, hipStream_t s) {
    Context &C = ctx();
    const int64_t n3 = g.n3;
    const Shape sh = C.shape_for(n3, false);
    const Part P = make_part(n3, sh.bs, sh.cpt);
    int rc;
    if ((rc = C.partials.ensure(sizeof(double) * P.nblk * kMaxPartials + 64))) return rc;
    if ((rc = C.state.ensure(sizeof(SolverState)))) return rc;
    if ((rc = C.ws64.ensure(sizeof(double) * n3 * 2))) return rc;
    SolverState *st = (SolverState *)C.state.ptr;
    double *x0 = (double *)C.ws64.ptr, *r = x0 + n3;
    hipError_t e = hipMemsetAsync(x0, 0, sizeof(double) * n3, s);
    if (e != hipSuccess) return hip_fail(e, "memset");
    hipLaunchKernelGGL(k_init_state, dim3(1), dim3(1), 0, s, st, 0.0, 1, 0, 1, 0, 0.0);
    MFIX_DISPATCH_BS(sh.bs, BS,
                     hipLaunchKernelGGL((k_setup_residual<BS>), grid_of(P), dim3(BS), 0, s, g, P, a, b,
                                        (const double *)x0, r, st, (double *)C.partials.ptr));
    e = hipMemcpyAsync(C.host_state, st, sizeof(SolverState), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return hip_fail(e, "copy");
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "scale");
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

// Time a few iterations of the solve for each candidate shape on a seeded
// system of the given size and keep the fastest per (size class, precision).
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
    std::vector<double> ha((size_t)(7 * n3)), hb((size_t)n3);
    if ((rc = mfix_leq_gen_system(2, imax, jmax, kmax, 4242, 16, ha.data(), hb.data()))) return rc;
    double *da = nullptr, *db = nullptr, *dx = nullptr, *da0 = nullptr;
    hipError_t e;
    if ((e = hipMalloc(&da, sizeof(double) * 7 * n3)) != hipSuccess || (e = hipMalloc(&db, sizeof(double) * n3)) != hipSuccess ||
        (e = hipMalloc(&dx, sizeof(double) * n3)) != hipSuccess || (e = hipMalloc(&da0, sizeof(double) * 7 * n3)) != hipSuccess)
        return hip_fail(e, "autotune alloc");
    hipStream_t s = C.stream;
    (void)hipMemcpyAsync(da0, ha.data(), sizeof(double) * 7 * n3, hipMemcpyHostToDevice, s);
    (void)hipMemcpyAsync(db, hb.data(), sizeof(double) * n3, hipMemcpyHostToDevice, s);
    const Shape cands[] = {{256, 1, 64}, {256, 2, 64}, {512, 1, 64}, {512, 2, 64}, {1024, 1, 64}, {256, 4, 64}};
    const int cls = size_class(n3);
    Shape best = cands[0];
    double best_ms = 1e30;
    const Shape saved = C.tuned[mixed ? 1 : 0][cls];
    const int iters = 12;
    std::vector<double> hb_copy = hb;
    for (const Shape &cand : cands) {
        C.tuned[mixed ? 1 : 0][cls] = cand;
        double t = 1e30;
        for (int rep = 0; rep < 2; ++rep) {
            (void)hipMemcpyAsync(da, da0, sizeof(double) * 7 * n3, hipMemcpyDeviceToDevice, s);
            (void)hipMemcpyAsync(db, hb_copy.data(), sizeof(double) * n3, hipMemcpyHostToDevice, s);
            (void)hipMemsetAsync(dx, 0, sizeof(double) * n3, s);
            mfix_leq_stats st = {};
            rc = mixed ? solve_mixed(g, PC_DIAG, "RSRS", 0.0, iters, dx, da, db, nullptr, &st, s)
                       : solve_fp64(g, PC_DIAG, "RSRS", 0.0, iters, dx, da, db, nullptr, &st, s, 0, 1);
            if (rc) break;
            t = std::min(t, st.device_ms / std::max(1, st.iterations));
        }
        if (rc) break;
        if (t < best_ms) {
            best_ms = t;
            best = cand;
        }
    }
    C.tuned[mixed ? 1 : 0][cls] = rc ? saved : best;
    (void)hipFree(da);
    (void)hipFree(db);
    (void)hipFree(dx);
    (void)hipFree(da0);
    if (rc) return rc;
    if (best_out) *best_out = best;
    return 0;
}

}  // namespace mfix
