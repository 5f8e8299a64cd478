// Residual norms of an assembled equation: CALC_RESID_C, CALC_RESID_S and
// CALC_RESID_PP (model/calc_resid.f:16-178, :195-371, :391-534), serial
// semantics, on the device.  One pass reads A_M, B_M, VAR and the cell
// FLAG array once; each block writes its partial sums (NUM, DEN, fluid
// cells) and its candidate for the largest residual; one finish block adds
// the partials in a fixed order and applies the reference's normalisation.
#include <hip/hip_runtime.h>

#include "../../include/mfix_leq.h"
#include "leq_context.h"

namespace mfix {

constexpr double kUndefined = 9.87654321e31;   // UNDEFINED, model/param1_mod.f:18
constexpr double kLargeNumber = 1.0e32;        // LARGE_NUMBER, model/param1_mod.f:23
constexpr int kResidBlock = 256;
constexpr int kResidCells = 8;                 // cells per thread: 2048 per block and partial
constexpr int kResidFinish = 1024;

struct ResidArgs {
    DGrid g;
    int kind, debug;
    double tol, dil_ep_s, norm;
    const double *var, *a, *b, *ep_s;
    const int *flag;
    double *partials;                          // kResidVals x nblk
    int64_t nblk;
    mfix_leq_resid *out;
};

__device__ __forceinline__ void better(double &v, double &i, double ov, double oi) { resid_better(v, i, ov, oi); }

__global__ void __launch_bounds__(kResidBlock) k_resid(ResidArgs R) {
    __shared__ double lds[(kResidBlock / 64) * 3];
    __shared__ double lv[kResidBlock / 64], li[kResidBlock / 64];
    const DGrid &g = R.g;
    // candidates must beat the scan's start value: RESID_IJK(1) = 0 for C
    // and S (:114-115), -1 for PP (:423)
    const double init = R.kind == RESID_PP ? -1.0 : 0.0;
    const double none = 1.0e300;
    double sums[3] = {0.0, 0.0, 0.0};
    double bv = init, bi = none;
    const auto *__restrict__ a = to_global(R.a);
    const auto *__restrict__ var = to_global(R.var);
    const auto *__restrict__ b = to_global(R.b);
    const auto *__restrict__ flag = to_global(R.flag);
    const int64_t c0 = (int64_t)blockIdx.x * (kResidBlock * kResidCells) + threadIdx.x;
    for (int q = 0; q < kResidCells; ++q) {
        const int64_t c = c0 + (int64_t)q * kResidBlock;   // increasing per thread
        double num1, den1;
        if (!resid_row(g, R.kind, R.tol, R.ep_s, R.dil_ep_s, a, b, var, flag, c, num1, den1)) continue;
        sums[0] += num1;
        sums[1] += den1;
        sums[2] += 1.0;
        if (num1 > init) better(bv, bi, num1, (double)c);
    }
    block_sum<kResidBlock, 3>(sums, lds);
    // block candidate: wave shuffles, then the waves in order
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_down(bv, off, 64), oi = __shfl_down(bi, off, 64);
        better(bv, bi, ov, oi);
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        lv[wv] = bv;
        li[wv] = bi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = lv[0], i = li[0];
        for (int u = 1; u < kResidBlock / 64; ++u) better(v, i, lv[u], li[u]);
        auto *__restrict__ p = to_global(R.partials);
        const int64_t nb = R.nblk, blk = blockIdx.x;
        p[0 * nb + blk] = sums[0];
        p[1 * nb + blk] = sums[1];
        p[2 * nb + blk] = sums[2];
        p[3 * nb + blk] = v;
        p[4 * nb + blk] = i;
    }
}

// One block: fixed-order sums of the partials and the reference's
// normalisation (calc_resid.f:155-171 for C, :348-362 for S, :462-524 for
// PP).  ijk_resid is 1-based like the reference's IJK.
__global__ void __launch_bounds__(kResidFinish) k_resid_finish(ResidArgs R) {
    __shared__ double lds[(kResidFinish / 64) * 3];
    __shared__ double lv[kResidFinish / 64], li[kResidFinish / 64];
    const double init = R.kind == RESID_PP ? -1.0 : 0.0;
    const double none = 1.0e300;
    double sums[3] = {0.0, 0.0, 0.0};
    double bv = init, bi = none;
    const auto *__restrict__ p = to_global((const double *)R.partials);
    const int64_t nb = R.nblk;
    // partials tid, tid + kResidFinish, ... in that order, the loads of
    // eight of them issued before the first add (a solver's set-up pass
    // leaves one partial per logical block: 67K at 256^3 with 256-cell
    // blocks, which a rolled loop would wait on one round trip at a time)
    constexpr int U = 8;
    for (int64_t q0 = threadIdx.x; q0 < nb; q0 += (int64_t)U * kResidFinish) {
        double pv[U][kResidVals];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t q = q0 + (int64_t)u * kResidFinish;
#pragma unroll
            for (int f = 0; f < kResidVals; ++f) pv[u][f] = q < nb ? p[f * nb + q] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (q0 + (int64_t)u * kResidFinish < nb) {
                sums[0] += pv[u][0];
                sums[1] += pv[u][1];
                sums[2] += pv[u][2];
                better(bv, bi, pv[u][3], pv[u][4]);
            }
    }
    block_sum<kResidFinish, 3>(sums, lds);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_down(bv, off, 64), oi = __shfl_down(bi, off, 64);
        better(bv, bi, ov, oi);
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        lv[wv] = bv;
        li[wv] = bi;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double maxv = lv[0], maxi = li[0];
    for (int u = 1; u < kResidFinish / 64; ++u) better(maxv, maxi, lv[u], li[u]);
    const double num = sums[0], den = sums[1];
    const int ncells = (int)sums[2];
    int ijk = maxi < none ? (int)maxi + 1 : 1;   // no candidate: the scan's start cell 1
    mfix_leq_resid o;
    o.num = num;
    o.den = den;
    o.ncells = ncells;
    if (R.kind == RESID_PP) {
        double r, m = maxv;
        if (den * R.norm > 0.0) {
            r = num / (den * R.norm);
            if (R.debug) m = ncells * maxv / (den * R.norm);
        } else if (num == 0.0) {
            r = 0.0;
            if (R.debug) {
                m = 0.0;
                ijk = 0;
            }
        } else {
            r = kLargeNumber;
            if (R.debug) m = kLargeNumber;
        }
        o.resid = r;
        o.max_resid = m;
        o.ijk_resid = ijk;
    } else if (!R.debug) {
        // the reference returns NUM and DEN only (calc_resid.f:113, :306)
        o.resid = 0.0;
        o.max_resid = -1.0;
        o.ijk_resid = 0;
    } else if (den > 0.0) {
        o.resid = num / den;
        o.max_resid = ncells * maxv / den;
        o.ijk_resid = ijk;
    } else if (num == 0.0) {
        o.resid = 0.0;
        o.max_resid = 0.0;
        o.ijk_resid = 0;
    } else {
        o.resid = kUndefined;
        o.max_resid = kUndefined;
        o.ijk_resid = ijk;
    }
    *R.out = o;
}

// Launches both passes on s and copies the result to *out (blocking).
static int run_resid(ResidArgs R, mfix_leq_resid *out, hipStream_t s) {
    Context &C = ctx();
    int rc;
    R.nblk = (R.g.n3 + kResidBlock * kResidCells - 1) / (kResidBlock * kResidCells);
    if ((rc = C.partials.ensure(sizeof(double) * kResidVals * (size_t)R.nblk + 256))) return rc;
    R.partials = (double *)C.partials.ptr;
    R.out = (mfix_leq_resid *)((char *)C.partials.ptr + sizeof(double) * kResidVals * (size_t)R.nblk);
    R.out = (mfix_leq_resid *)(((uintptr_t)R.out + 63) & ~(uintptr_t)63);
    hipLaunchKernelGGL(k_resid, dim3((unsigned)R.nblk), dim3(kResidBlock), 0, s, R);
    hipLaunchKernelGGL(k_resid_finish, dim3(1), dim3(kResidFinish), 0, s, R);
    hipError_t e = hipMemcpyAsync(out, R.out, sizeof(mfix_leq_resid), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess) e = hipGetLastError();
    return e == hipSuccess ? 0 : hip_fail(e, "residual");
}

// The finish of residual partials a solver's set-up pass wrote (kResidVals
// x nblk slots, one per logical block): the same fixed-order sums and
// normalisation, into *d_out on the device.
void launch_resid_finish(const DGrid &g, int kind, int debug, double norm, double *partials, int64_t nblk,
                         mfix_leq_resid *d_out, hipStream_t s) {
    ResidArgs R = {};
    R.g = g;
    R.kind = kind;
    R.debug = debug != 0;
    R.norm = norm;
    R.partials = partials;
    R.nblk = nblk;
    R.out = d_out;
    hipLaunchKernelGGL(k_resid_finish, dim3(1), dim3(kResidFinish), 0, s, R);
}

static int resid_ready(hipStream_t *s, void *stream) {
    Context &C = ctx();
    int rc;
    if ((rc = C.ensure_init())) return rc;
    if (!C.have_grid) return MFIX_LEQ_ENOGRID;
    *s = stream ? (hipStream_t)stream : C.stream;
    return 0;
}

static ResidArgs resid_args(int kind, int debug, const double *var, const double *a, const double *b, const int *flag,
                            double tol, const double *ep_s, double dil_ep_s, double norm) {
    ResidArgs R = {};
    R.g = make_dgrid(ctx().grid);
    R.kind = kind;
    R.debug = debug != 0;
    R.var = var;
    R.a = a;
    R.b = b;
    R.flag = flag;
    R.tol = tol;
    R.ep_s = ep_s;
    R.dil_ep_s = dil_ep_s;
    R.norm = norm;
    return R;
}

int resid_pp(const double *b, const int *flag, double norm, int debug, mfix_leq_resid *out, hipStream_t s) {
    return run_resid(resid_args(RESID_PP, debug, nullptr, nullptr, b, flag, 0.0, nullptr, 0.0, norm), out, s);
}

// Host arrays: staged through the context's device buffers.
static int resid_host(int kind, const double *var, const double *a_m, const double *b_m, const int *flag, double tol,
                      const double *ep_s, double dil_ep_s, double norm, int debug, mfix_leq_resid *out) {
    Context &C = ctx();
    hipStream_t s;
    int rc;
    if ((rc = resid_ready(&s, nullptr))) return rc;
    const int64_t n3 = make_dgrid(C.grid).n3;
    const bool need_av = kind != RESID_PP;
    if ((rc = C.d_b.ensure(sizeof(double) * n3)) || (rc = C.tmp.ensure(sizeof(int) * n3 + sizeof(double) * n3 + 256)))
        return rc;
    if (need_av && ((rc = C.d_a.ensure(sizeof(double) * 7 * n3)) || (rc = C.d_x.ensure(sizeof(double) * n3))))
        return rc;
    int *dflag = (int *)C.tmp.ptr;
    double *deps = (double *)(((uintptr_t)((char *)C.tmp.ptr + sizeof(int) * n3) + 255) & ~(uintptr_t)255);
    hipError_t e = hipMemcpyAsync(C.d_b.ptr, b_m, sizeof(double) * n3, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dflag, flag, sizeof(int) * n3, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && need_av) e = hipMemcpyAsync(C.d_a.ptr, a_m, sizeof(double) * 7 * n3, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && need_av) e = hipMemcpyAsync(C.d_x.ptr, var, sizeof(double) * n3, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && ep_s) e = hipMemcpyAsync(deps, ep_s, sizeof(double) * n3, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "residual host to device copy");
    const ResidArgs R = resid_args(kind, debug, (const double *)C.d_x.ptr, (const double *)C.d_a.ptr,
                                   (const double *)C.d_b.ptr, dflag, tol, ep_s ? deps : nullptr, dil_ep_s, norm);
    return run_resid(R, out, s);
}

}  // namespace mfix

using namespace mfix;

extern "C" {

int mfix_calc_resid_c_dev(const double *d_var, const double *d_a_m, const double *d_b_m, const int *d_flag,
                          int debug_resid, mfix_leq_resid *out, void *stream) {
    hipStream_t s;
    int rc;
    if ((rc = resid_ready(&s, stream))) return rc;
    if (!d_var || !d_a_m || !d_b_m || !d_flag || !out) return MFIX_LEQ_EINVAL;
    return run_resid(resid_args(RESID_C, debug_resid, d_var, d_a_m, d_b_m, d_flag, 0.0, nullptr, 0.0, 1.0), out, s);
}

int mfix_calc_resid_s_dev(const double *d_var, const double *d_a_m, const double *d_b_m, const int *d_flag,
                          double tol, const double *d_ep_s, double dil_ep_s, int debug_resid, mfix_leq_resid *out,
                          void *stream) {
    hipStream_t s;
    int rc;
    if ((rc = resid_ready(&s, stream))) return rc;
    if (!d_var || !d_a_m || !d_b_m || !d_flag || !out) return MFIX_LEQ_EINVAL;
    return run_resid(resid_args(RESID_S, debug_resid, d_var, d_a_m, d_b_m, d_flag, tol, d_ep_s, dil_ep_s, 1.0), out,
                     s);
}

int mfix_calc_resid_pp_dev(const double *d_b_m, const int *d_flag, double norm, int debug_resid,
                           mfix_leq_resid *out, void *stream) {
    hipStream_t s;
    int rc;
    if ((rc = resid_ready(&s, stream))) return rc;
    if (!d_b_m || !d_flag || !out) return MFIX_LEQ_EINVAL;
    return run_resid(resid_args(RESID_PP, debug_resid, nullptr, nullptr, d_b_m, d_flag, 0.0, nullptr, 0.0, norm),
                     out, s);
}

int mfix_calc_resid_c(const double *var, const double *a_m, const double *b_m, const int *flag, int debug_resid,
                      mfix_leq_resid *out) {
    if (!var || !a_m || !b_m || !flag || !out) return MFIX_LEQ_EINVAL;
    return resid_host(RESID_C, var, a_m, b_m, flag, 0.0, nullptr, 0.0, 1.0, debug_resid, out);
}

int mfix_calc_resid_s(const double *var, const double *a_m, const double *b_m, const int *flag, double tol,
                      const double *ep_s, double dil_ep_s, int debug_resid, mfix_leq_resid *out) {
    if (!var || !a_m || !b_m || !flag || !out) return MFIX_LEQ_EINVAL;
    return resid_host(RESID_S, var, a_m, b_m, flag, tol, ep_s, dil_ep_s, 1.0, debug_resid, out);
}

int mfix_calc_resid_pp(const double *b_m, const int *flag, double norm, int debug_resid, mfix_leq_resid *out) {
    if (!b_m || !flag || !out) return MFIX_LEQ_EINVAL;
    return resid_host(RESID_PP, nullptr, nullptr, b_m, flag, 0.0, nullptr, 0.0, norm, debug_resid, out);
}

}  // extern "C"
