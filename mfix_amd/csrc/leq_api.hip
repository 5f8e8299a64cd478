// C ABI of the gfx950 LEQ_BICGS solver (include/mfix_leq.h).  Every entry
// validates its arguments, selects the context's stream and calls the solver
// engine (leq_solve.hip); none of them computes on the host.
#include <hip/hip_runtime.h>

#include <cmath>

#include <cstring>
#include <vector>

#include "../../include/mfix_leq.h"
#include "leq_context.h"

using namespace mfix;

namespace {

const char *err_name(int code) {
    switch (code) {
    case MFIX_LEQ_OK: return "ok";
    case MFIX_LEQ_EBADPC: return "preconditioner option not found (PC must be LINE, DIAG or NONE)";
    case MFIX_LEQ_ENOGRID: return "grid not initialised (call mfix_leq_grid_init)";
    case MFIX_LEQ_EHIP: return "HIP runtime error";
    case MFIX_LEQ_EINVAL: return "invalid argument";
    case MFIX_LEQ_ENOMEM: return "device memory allocation failed or over budget";
    case MFIX_LEQ_EDGTSV: return "DGTSV failed in a JK/IJ line sweep";
    case MFIX_LEQ_EBUSY: return "serving queue full";
    case MFIX_LEQ_EMETHOD: return "only LEQ_METHOD=2 (BiCGSTAB) is provided";
    case MFIX_LEQ_ETOOBIG: return "system larger than the no-allocation entry's fixed workspace (MFIX_LEQ_NOALLOC_BYTES)";
    default: return "unknown error";
    }
}

int ready(hipStream_t *s, void *stream) {
    Context &C = ctx();
    int rc;
    if ((rc = C.ensure_init())) return rc;
    if (!C.have_grid) return MFIX_LEQ_ENOGRID;
    *s = stream ? (hipStream_t)stream : C.stream;
    return 0;
}

int sync_status(hipStream_t s, const char *what) {
    hipError_t e = hipStreamSynchronize(s);
    if (e == hipSuccess) e = hipGetLastError();
    return e == hipSuccess ? 0 : hip_fail(e, what);
}

// SOLVE_LIN_EQ calls deferred to the next sweep flush, in call order.
struct Deferred {
    mfix_leq_batch_item item;
    mfix_leq_grid grid;   // the grid current at the call
    int vno;
    int *ier;
};
std::vector<Deferred> &sweep_queue() {
    static std::vector<Deferred> q;
    return q;
}

int copy_sync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
    int rc;
    if ((rc = ctx().ensure_init())) return rc;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, ctx().stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx().stream);
    return e == hipSuccess ? 0 : hip_fail(e, "hipMemcpy");
}

}  // namespace

extern "C" {

int mfix_leq_init(int device) { return ctx().init(device); }

int mfix_leq_finalize(void) {
    Context &C = ctx();
    if (!C.initialised) return 0;
    (void)hipStreamSynchronize(C.stream);
    C.release_all();
    if (C.host_state) (void)hipHostFree(C.host_state);
    if (C.host_count) (void)hipHostFree(C.host_count);
    C.host_state = nullptr;
    C.host_count = nullptr;
    (void)hipStreamDestroy(C.stream);
    C.stream = nullptr;
    C.initialised = false;
    return 0;
}

int mfix_leq_grid_init(const mfix_leq_grid *g) {
    if (!g) return MFIX_LEQ_EINVAL;
    if (g->imax < 1 || g->jmax < 1 || (!g->no_k && g->kmax < 1)) return MFIX_LEQ_ENOGRID;
    const double n3 = (double)(g->imax + 2) * (g->jmax + 2) * (g->no_k ? 1 : g->kmax + 2);
    if (n3 >= 2147483647.0) return MFIX_LEQ_EINVAL;  // kernels index cells with 32-bit (i,j,k) math
    Context &C = ctx();
    C.grid = *g;
    if (C.grid.icheck_bicgs < 1) C.grid.icheck_bicgs = 1;
    C.have_grid = true;
    return 0;
}

int64_t mfix_leq_dimension_3(void) {
    Context &C = ctx();
    if (!C.have_grid) return MFIX_LEQ_ENOGRID;
    return make_dgrid(C.grid).n3;
}

int mfix_leq_iter_tot(int vno) {
    if (vno < 1 || vno > 10) return MFIX_LEQ_EINVAL;
    return ctx().iter_tot[vno - 1];
}

void mfix_leq_reset_iter_tot(void) { std::memset(ctx().iter_tot, 0, sizeof(ctx().iter_tot)); }

const char *mfix_leq_error_string(int code) {
    if ((code == MFIX_LEQ_EHIP || code == MFIX_LEQ_ETOOBIG) && ctx().last_msg[0]) return ctx().last_msg;
    return err_name(code);
}

int mfix_leq_bicgs_dev(const char *cmethod, double tol, const char *pc, int itmax, double *d_var, double *d_a_m,
                       double *d_b_m, const mfix_leq_opts *opts, mfix_leq_stats *stats, void *stream) {
    hipStream_t s;
    int rc;
    const int p = parse_pc(pc);
    if (p < 0) return MFIX_LEQ_EBADPC;
    if ((rc = ready(&s, stream))) return rc;
    if (!d_var || !d_a_m || !d_b_m) return MFIX_LEQ_EINVAL;
    const Context &C = ctx();
    return solve_fp64(make_dgrid(C.grid), p, cmethod ? cmethod : "RSRS", tol, itmax, d_var, d_a_m, d_b_m, opts, stats,
                      s, C.grid.minimize_dotproducts, C.grid.icheck_bicgs);
}

int mfix_leq_bicgs_mixed_dev(const char *cmethod, double tol, const char *pc, int itmax, double *d_var,
                             double *d_a_m, double *d_b_m, const mfix_leq_opts *opts, mfix_leq_stats *stats,
                             void *stream) {
    hipStream_t s;
    int rc;
    const int p = parse_pc(pc);
    if (p < 0) return MFIX_LEQ_EBADPC;
    if ((rc = ready(&s, stream))) return rc;
    if (!d_var || !d_a_m || !d_b_m) return MFIX_LEQ_EINVAL;
    return solve_mixed(make_dgrid(ctx().grid), p, cmethod ? cmethod : "RSRS", tol, itmax, d_var, d_a_m, d_b_m, opts,
                       stats, s);
}

int mfix_leq_resid_relax_bicgs_dev(int resid_kind, const int *d_flag, double resid_tol, const double *d_ep_s,
                                   double dil_ep_s, double norm, int debug_resid, mfix_leq_resid *resid, double ur,
                                   const char *cmethod, double tol, const char *pc, int itmax, int mixed,
                                   double *d_var, double *d_a_m, double *d_b_m, const mfix_leq_opts *opts,
                                   mfix_leq_stats *stats, void *stream) {
    hipStream_t s;
    int rc;
    const int p = parse_pc(pc);
    if (p < 0) return MFIX_LEQ_EBADPC;
    if ((rc = ready(&s, stream))) return rc;
    if (!d_var || !d_a_m || !d_b_m || !d_flag) return MFIX_LEQ_EINVAL;
    if (resid_kind < -1 || resid_kind > RESID_PP || (resid_kind >= 0 && !resid)) return MFIX_LEQ_EINVAL;
    if (!std::isfinite(ur)) return MFIX_LEQ_EINVAL;
    Prologue pro;
    pro.flag = d_flag;
    pro.resid_kind = resid_kind;
    pro.resid_tol = resid_tol;
    pro.ep_s = d_ep_s;
    pro.dil_ep_s = dil_ep_s;
    if (ur != 0.0) {
        pro.relax = 1;
        pro.f1 = 1.0 / ur;  // F1 = ONE/UR, F2 = F1 - ONE (under_relax.f:67-68)
        pro.f2 = pro.f1 - 1.0;
    }
    const Context &C = ctx();
    return solve_with_prologue(make_dgrid(C.grid), p, cmethod ? cmethod : "RSRS", tol, itmax, d_var, d_a_m, d_b_m,
                               opts, stats, s, mixed != 0, C.grid.minimize_dotproducts, C.grid.icheck_bicgs, pro,
                               debug_resid, norm, resid_kind >= 0 ? resid : nullptr);
}

// Host-array drop-in for LEQ_BICGS (model/leq_bicgs.f:23-90).
int mfix_leq_bicgs(const char *vname, int vname_len, int vno, double *var, double *a_m, double *b_m,
                   const char *cmethod, int cmethod_len, double tol, const char *pc, int itmax, int *ier) {
    (void)vname;
    (void)vname_len;
    Context &C = ctx();
    hipStream_t s;
    int rc;
    const int p = parse_pc(pc);
    if (p < 0) return MFIX_LEQ_EBADPC;
    if ((rc = ready(&s, nullptr))) return rc;
    if (!var || !a_m || !b_m || !ier) return MFIX_LEQ_EINVAL;
    char cm[5] = "    ";
    for (int q = 0; q < 4; ++q) cm[q] = (cmethod && q < cmethod_len) ? cmethod[q] : ' ';
    const DGrid g = make_dgrid(C.grid);
    const int64_t n3 = g.n3;
    if ((rc = C.d_a.ensure(sizeof(double) * 7 * n3)) || (rc = C.d_b.ensure(sizeof(double) * n3)) ||
        (rc = C.d_x.ensure(sizeof(double) * n3)))
        return rc;
    double *da = (double *)C.d_a.ptr, *db = (double *)C.d_b.ptr, *dx = (double *)C.d_x.ptr;
    hipError_t e;
    if ((e = hipMemcpyAsync(da, a_m, sizeof(double) * 7 * n3, hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(db, b_m, sizeof(double) * n3, hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(dx, var, sizeof(double) * n3, hipMemcpyHostToDevice, s)) != hipSuccess)
        return hip_fail(e, "host to device copy");
    mfix_leq_opts o = {};
    o.write_back_scaled = 1;
    if (C.host_noise) {
        if ((rc = C.misc.ensure(sizeof(double) * n3))) return rc;
        if ((e = hipMemcpyAsync(C.misc.ptr, C.host_noise, sizeof(double) * n3, hipMemcpyHostToDevice, s)) !=
            hipSuccess)
            return hip_fail(e, "noise copy");
        o.noise = (const double *)C.misc.ptr;
    }
    o.noise_seed = C.noise_seed;
    mfix_leq_stats st = {};
    rc = solve_fp64(g, p, cm, tol, itmax, dx, da, db, &o, &st, s, C.grid.minimize_dotproducts, C.grid.icheck_bicgs);
    if (rc && rc != MFIX_LEQ_EDGTSV) return rc;
    if ((e = hipMemcpyAsync(var, dx, sizeof(double) * n3, hipMemcpyDeviceToHost, s)) != hipSuccess)
        return hip_fail(e, "device to host copy");
    if (C.write_back_scaled) {
        if ((e = hipMemcpyAsync(a_m, da, sizeof(double) * 7 * n3, hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (e = hipMemcpyAsync(b_m, db, sizeof(double) * n3, hipMemcpyDeviceToHost, s)) != hipSuccess)
            return hip_fail(e, "device to host copy");
    }
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    *ier = st.ier;
    if (vno >= 1 && vno <= 10) C.iter_tot[vno - 1] += st.iter_tot_increment;
    C.last_stats = st;
    return rc;
}

// SOLVE_LIN_EQ (model/solve_lin_eq.f:17-195), METHOD = 2 only.
int mfix_solve_lin_eq(const char *vname, int vname_len, int vno, double *var, double *a_m, double *b_m, int m,
                      int itmax, int method, const char *sweep, double tol1, const char *pc, int *ier) {
    if (method != 2) return MFIX_LEQ_EMETHOD;
    Context &C = ctx();
    if (!C.have_grid) return MFIX_LEQ_ENOGRID;
    if (m < 0) return MFIX_LEQ_EINVAL;
    const int64_t n3 = make_dgrid(C.grid).n3;
    // A_M(:,:,M) and B_M(:,M) are contiguous slices (M is the last dimension)
    return mfix_leq_bicgs(vname, vname_len, vno, var, a_m + (int64_t)m * 7 * n3, b_m + (int64_t)m * n3, sweep, 4,
                          tol1, pc, itmax, ier);
}

int mfix_leq_matvec_dev(const double *d_a_m, const double *d_x, double *d_y, void *stream) {
    hipStream_t s;
    int rc;
    if ((rc = ready(&s, stream))) return rc;
    if (!d_a_m || !d_x || !d_y) return MFIX_LEQ_EINVAL;
    run_matvec64(make_dgrid(ctx().grid), d_a_m, d_x, d_y, s);
    return sync_status(s, "matvec");
}

int mfix_leq_msolve32_dev(const char *cmethod, const float *d_a_m, const float *d_r, float *d_z, void *stream) {
    hipStream_t s;
    int rc;
    if ((rc = ready(&s, stream))) return rc;
    if (!d_a_m || !d_r || !d_z) return MFIX_LEQ_EINVAL;
    Context &C = ctx();
    const DGrid g = make_dgrid(C.grid);
    if ((rc = C.tmp.ensure(sizeof(float) * (size_t)line_scratch_elems(g)))) return rc;
    if ((rc = C.misc.ensure(sizeof(int) * 4 + sizeof(double) * 8))) return rc;
    const char *cm = cmethod ? cmethod : "RSRS";
    rc = run_msolve_line_rb<float>(g, cm, d_a_m, d_r, d_z, (float *)C.tmp.ptr, s);
    if (rc < 0) return rc;
    if (rc == 1) {
        int *flag = (int *)C.misc.ptr;
        hipError_t e = hipMemsetAsync(flag, 0, sizeof(int), s);
        if (e != hipSuccess) return hip_fail(e, "memset");
        launch_msolve_line<float>(g, C.shape_for(g.n3, true), cm, d_a_m, d_r, d_z, (float *)C.tmp.ptr, nullptr, 0,
                                  flag, s, true);
        int hflag = 0;
        e = hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) return hip_fail(e, "copy");
        if ((rc = sync_status(s, "msolve32"))) return rc;
        for (auto &pp : C.prof) pp.collect();
        return hflag ? MFIX_LEQ_EDGTSV : 0;
    }
    if ((rc = sync_status(s, "msolve32"))) return rc;
    for (auto &pp : C.prof) pp.collect();
    return 0;
}

int mfix_leq_matvec32_dev(const float *d_a_m, const float *d_x, float *d_y, void *stream) {
    hipStream_t s;
    int rc;
    if ((rc = ready(&s, stream))) return rc;
    if (!d_a_m || !d_x || !d_y) return MFIX_LEQ_EINVAL;
    run_matvec32(make_dgrid(ctx().grid), d_a_m, d_x, d_y, s);
    return sync_status(s, "matvec32");
}

int mfix_leq_msolve_dev(const char *pc, const char *cmethod, const double *d_a_m, const double *d_r, double *d_z,
                        void *stream) {
    hipStream_t s;
    int rc;
    const int p = parse_pc(pc);
    if (p < 0) return MFIX_LEQ_EBADPC;
    if ((rc = ready(&s, stream))) return rc;
    if (!d_a_m || !d_r || !d_z) return MFIX_LEQ_EINVAL;
    Context &C = ctx();
    const DGrid g = make_dgrid(C.grid);
    hipError_t e;
    if (p == PC_NONE) {
        e = hipMemcpyAsync(d_z, d_r, sizeof(double) * g.n3, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return hip_fail(e, "copy");
        return sync_status(s, "msolve");
    }
    if (p == PC_DIAG) {
        run_diag(g, d_a_m, d_r, d_z, s);
        return sync_status(s, "msolve");
    }
    if ((rc = C.tmp.ensure(sizeof(double) * (size_t)line_scratch_elems(g)))) return rc;
    if ((rc = C.misc.ensure(sizeof(int) * 4 + sizeof(double) * 8))) return rc;
    int *flag = (int *)C.misc.ptr;
    e = hipMemsetAsync(flag, 0, sizeof(int), s);
    if (e != hipSuccess) return hip_fail(e, "memset");
    // R/S sweep strings: the solver's own red-black kernel; others the
    // per-line phases
    rc = run_msolve_line_rb<double>(g, cmethod ? cmethod : "RSRS", d_a_m, d_r, d_z, (double *)C.tmp.ptr, s);
    if (rc < 0) return rc;
    if (rc == 1)
        launch_msolve_line<double>(g, C.shape_for(g.n3, false), cmethod ? cmethod : "RSRS", d_a_m, d_r, d_z,
                                   (double *)C.tmp.ptr, nullptr, 0, flag, s, true);
    int hflag = 0;
    e = hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return hip_fail(e, "copy");
    if ((rc = sync_status(s, "msolve"))) return rc;
    for (auto &pp : C.prof) pp.collect();
    return hflag ? MFIX_LEQ_EDGTSV : 0;
}

int mfix_leq_dot_dev(const double *d_x, const double *d_y, double *result, void *stream) {
    hipStream_t s;
    int rc;
    if ((rc = ready(&s, stream))) return rc;
    if (!result || !d_x || !d_y) return MFIX_LEQ_EINVAL;
    Context &C = ctx();
    if ((rc = C.misc.ensure(sizeof(double) * 8))) return rc;
    double *out = (double *)C.misc.ptr;
    if ((rc = run_dot(make_dgrid(C.grid), d_x, d_y, out, s))) return rc;
    hipError_t e = hipMemcpyAsync(result, out, sizeof(double), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return hip_fail(e, "copy");
    return sync_status(s, "dot");
}

int mfix_leq_scale_dev(double *d_a_m, double *d_b_m, int *ier, void *stream) {
    hipStream_t s;
    int rc;
    if ((rc = ready(&s, stream))) return rc;
    if (!ier || !d_a_m || !d_b_m) return MFIX_LEQ_EINVAL;
    int zero = 0;
    if ((rc = run_scale(make_dgrid(ctx().grid), d_a_m, d_b_m, &zero, s))) return rc;
    *ier = zero ? -2 : 0;
    return 0;
}

int mfix_leq_batch_solve(int nsys, const mfix_leq_batch_item *items, mfix_leq_stats *stats, void *stream) {
    if (nsys < 0 || (nsys > 0 && !items)) return MFIX_LEQ_EINVAL;
    if (nsys > kHostStates) return MFIX_LEQ_EINVAL;
    for (int q = 0; q < nsys; ++q) {
        const mfix_leq_batch_item &it = items[q];
        if (parse_pc(it.pc) < 0) return MFIX_LEQ_EBADPC;
        if (it.imax < 1 || it.jmax < 1 || it.kmax < 1 || !it.d_var || !it.d_a_m || !it.d_b_m)
            return MFIX_LEQ_EINVAL;
        if ((double)(it.imax + 2) * (it.jmax + 2) * (it.kmax + 2) >= 2147483647.0) return MFIX_LEQ_EINVAL;
    }
    int rc;
    if ((rc = ctx().ensure_init())) return rc;
    if (nsys == 0) return 0;
    return solve_batch(nsys, items, stats, stream ? (hipStream_t)stream : ctx().stream);
}

int mfix_solve_lin_eq_deferred(const char *vname, int vname_len, int vno, double *d_var, double *d_a_m,
                               double *d_b_m, int itmax, int method, const char *sweep, double tol1, const char *pc,
                               int mixed, int *ier) {
    (void)vname;
    (void)vname_len;
    // the checks SOLVE_LIN_EQ / LEQ_BICGS make at the call (solve_lin_eq.f:
    // 139-190, leq_bicgs.f:83-87) happen here, before anything is queued
    if (method != 2) return MFIX_LEQ_EMETHOD;
    if (parse_pc(pc) < 0) return MFIX_LEQ_EBADPC;
    if (!d_var || !d_a_m || !d_b_m || !ier || !sweep || vno < 1 || vno > 10 || itmax < 0) return MFIX_LEQ_EINVAL;
    Context &C = ctx();
    if (!C.have_grid) return MFIX_LEQ_ENOGRID;
    std::vector<Deferred> &q = sweep_queue();
    if ((int)q.size() >= kHostStates) return MFIX_LEQ_EBUSY;
    Deferred d = {};
    d.item.imax = C.grid.imax;
    d.item.jmax = C.grid.jmax;
    d.item.kmax = C.grid.no_k ? 1 : C.grid.kmax;
    std::memcpy(d.item.pc, pc, 4);
    // a C string shorter than 4 characters is blank-padded, never read past its NUL
    bool ended = false;
    for (int i = 0; i < 4; ++i) {
        ended = ended || sweep[i] == '\0';
        d.item.cmethod[i] = ended ? ' ' : sweep[i];
    }
    d.item.tol = tol1;
    d.item.itmax = itmax;
    d.item.mixed = mixed != 0;
    d.item.d_var = d_var;
    d.item.d_a_m = d_a_m;
    d.item.d_b_m = d_b_m;
    d.item.noise_seed = 0;   // the internal generator's default, as a single solve without options
    d.grid = C.grid;
    d.vno = vno;
    d.ier = ier;
    q.push_back(d);
    return 0;
}

int mfix_leq_sweep_pending(void) { return (int)sweep_queue().size(); }

void mfix_leq_sweep_discard(void) { sweep_queue().clear(); }

int mfix_leq_sweep_flush(mfix_leq_stats *stats, void *stream) {
    std::vector<Deferred> q;
    q.swap(sweep_queue());   // the queue is empty whatever the outcome
    if (q.empty()) return 0;
    int rc;
    Context &C = ctx();
    if ((rc = C.ensure_init())) return rc;
    const int n = (int)q.size();
    std::vector<mfix_leq_batch_item> items((size_t)n);
    std::vector<mfix_leq_grid> grids((size_t)n);
    std::vector<mfix_leq_stats> st((size_t)n);
    for (int i = 0; i < n; ++i) {
        items[(size_t)i] = q[(size_t)i].item;
        grids[(size_t)i] = q[(size_t)i].grid;
    }
    rc = solve_batch(n, items.data(), st.data(), stream ? (hipStream_t)stream : C.stream, grids.data());
    if (rc && rc != MFIX_LEQ_EDGTSV) return rc;
    for (int i = 0; i < n; ++i) {
        *q[(size_t)i].ier = st[(size_t)i].ier;
        C.iter_tot[q[(size_t)i].vno - 1] += st[(size_t)i].iter_tot_increment;
        if (stats) stats[i] = st[(size_t)i];
    }
    return rc;
}

int mfix_leq_autotune(int imax, int jmax, int kmax, int mixed, int *block, int *lines_per_block) {
    if (imax < 1 || jmax < 1 || kmax < 1) return MFIX_LEQ_EINVAL;
    Shape best;
    int rc = autotune(imax, jmax, kmax, mixed, &best);
    if (rc) return rc;
    if (block) *block = best.bs * 100 + best.cpt;
    if (lines_per_block) *lines_per_block = best.lines_per_block;
    return 0;
}

int mfix_leq_set_shape(int imax, int jmax, int kmax, int mixed, int block_threads, int cells_per_thread) {
    if (imax < 1 || jmax < 1 || kmax < 1) return MFIX_LEQ_EINVAL;
    if (block_threads != 256 && block_threads != 512 && block_threads != 1024) return MFIX_LEQ_EINVAL;
    if (cells_per_thread != 1 && cells_per_thread != 2 && cells_per_thread != 4) return MFIX_LEQ_EINVAL;
    const int64_t n3 = (int64_t)(imax + 2) * (jmax + 2) * (kmax + 2);
    Shape &sh = ctx().tuned[mixed ? 1 : 0][size_class(n3)];
    sh = Shape{block_threads, cells_per_thread, sh.lines_per_block};
    return 0;
}

int mfix_leq_set_mixed_delta(double delta) {
    if (!(delta >= 0.0 && delta < 1.0)) return MFIX_LEQ_EINVAL;
    ctx().mixed_delta = delta;
    return 0;
}

int mfix_leq_set_line_block(int imax, int jmax, int kmax, int mixed, int lines_per_block) {
    if (imax < 1 || jmax < 1 || kmax < 1) return MFIX_LEQ_EINVAL;
    if (lines_per_block != 0 && lines_per_block != 16 && lines_per_block != 32 && lines_per_block != 64)
        return MFIX_LEQ_EINVAL;
    const int64_t n3 = (int64_t)(imax + 2) * (jmax + 2) * (kmax + 2);
    Shape &sh = ctx().tuned[mixed ? 1 : 0][size_class(n3)];
    if (!sh.bs) sh = Shape{256, 1, 0};
    sh.lines_per_block = lines_per_block;
    return 0;
}

int mfix_leq_prof_enable(int which, int on) {
    if (which < 0 || which >= kNumProf) return MFIX_LEQ_EINVAL;
    Prof &p = ctx().prof[which];
    p.on = on != 0;
    p.launches = 0;
    p.total_ms = 0.0;
    p.alg_bytes = 0.0;
    p.idle_launches = 0;
    p.idle_ms = 0.0;
    return 0;
}

int mfix_leq_prof_idle(int which, int64_t *launches, double *total_ms) {
    if (which < 0 || which >= kNumProf) return MFIX_LEQ_EINVAL;
    const Prof &p = ctx().prof[which];
    if (launches) *launches = p.idle_launches;
    if (total_ms) *total_ms = p.idle_ms;
    return 0;
}

int mfix_leq_prof_read(int which, int64_t *launches, double *total_ms, double *alg_bytes) {
    if (which < 0 || which >= kNumProf) return MFIX_LEQ_EINVAL;
    const Prof &p = ctx().prof[which];
    if (launches) *launches = p.launches;
    if (total_ms) *total_ms = p.total_ms;
    if (alg_bytes) *alg_bytes = p.alg_bytes;
    return 0;
}

int mfix_leq_set_noise(const double *host_noise, uint64_t seed) {
    ctx().host_noise = host_noise;
    ctx().noise_seed = seed;
    return 0;
}

int mfix_leq_set_write_back(int on) {
    ctx().write_back_scaled = on != 0;
    return 0;
}

int mfix_leq_set_graphs(int on) {
    ctx().use_graphs = on != 0;
    return 0;
}

int mfix_leq_set_unit_diag(int on) {
    ctx().unit_diag = on != 0;
    return 0;
}

int mfix_leq_last_stats(mfix_leq_stats *st) {
    if (!st) return MFIX_LEQ_EINVAL;
    *st = ctx().last_stats;
    return 0;
}

int mfix_leq_dev_alloc(size_t bytes, void **out) {
    int rc;
    if ((rc = ctx().ensure_init())) return rc;
    if (!out) return MFIX_LEQ_EINVAL;
    *out = nullptr;
    hipError_t e = hipMalloc(out, bytes ? bytes : 1);
    if (e != hipSuccess) {
        hip_fail(e, "hipMalloc");
        return MFIX_LEQ_ENOMEM;
    }
    return 0;
}

int mfix_leq_dev_free(void *p) {
    if (!p) return 0;
    hipError_t e = hipFree(p);
    return e == hipSuccess ? 0 : hip_fail(e, "hipFree");
}

int mfix_leq_copy_to_dev(void *dst, const void *src, size_t bytes) {
    return copy_sync(dst, src, bytes, hipMemcpyHostToDevice);
}
int mfix_leq_copy_from_dev(void *dst, const void *src, size_t bytes) {
    return copy_sync(dst, src, bytes, hipMemcpyDeviceToHost);
}
int mfix_leq_copy_dev(void *dst, const void *src, size_t bytes) {
    return copy_sync(dst, src, bytes, hipMemcpyDeviceToDevice);
}

int mfix_leq_copy_kernel_dev(double *dst, const double *src, int64_t n) {
    int rc;
    if ((rc = ctx().ensure_init())) return rc;
    if (!dst || !src || n < 0) return MFIX_LEQ_EINVAL;
    run_copy(src, dst, n, ctx().stream);
    return sync_status(ctx().stream, "copy kernel");
}

int mfix_leq_sync(void) {
    int rc;
    if ((rc = ctx().ensure_init())) return rc;
    hipError_t e = hipDeviceSynchronize();
    return e == hipSuccess ? 0 : hip_fail(e, "hipDeviceSynchronize");
}

}  // extern "C"
