// Streaming phase kernels of one BiCGSTAB iteration (LEQ_BICGS0,
// model/leq_bicgs.f:316-466) and its set-up (:194-305).
//
// Each kernel is a pass over the cells with BS threads per block; a
// reduction phase writes one partial per block and is followed by a
// k_finish launch (leq_kernels.h).  With the DIAG preconditioner
// (LEQ_MSOLVE1, leqsol_mod.f:729-744) Phat = P/A0 and Shat = S/A0 are never
// stored: the stencil kernels divide each operand by its A0 as they load it
// and the solution update divides again, which rounds exactly like the
// reference's separate division pass and saves two vector writes and reads
// per preconditioner application.
#pragma once
#include "leq_kernels.h"

namespace mfix {

// Fresh solver state: the whole struct is rewritten, so nothing of a previous
// solve carries over.
__global__ void k_init_state(SolverState *st, double tol, int itmax, int minimize, int icheck, int mixed,
                             double delta) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        SolverState s = {};
        s.tol = tol;
        s.itmax = itmax;
        s.i = 1;
        s.iter_var = 1;
        s.minimize = minimize;
        s.icheck = icheck > 0 ? icheck : 1;
        s.mixed = mixed;
        s.delta = delta;
        s.true_rnorm = -1.0;
        *st = s;
    }
}

// Unit scaling (leq_bicgs.f:194-233) fused with R = B - A*Var (:283-287) and
// the partial sums of <R,R> (:291).  Each row is scaled by its own max-norm,
// so the residual row uses only this thread's coefficients.
template <int BS>
__global__ void __launch_bounds__(BS) k_setup_residual(DGrid g, Part P, double *__restrict__ a,
                                                       double *__restrict__ b, const double *__restrict__ x,
                                                       double *__restrict__ r, SolverState *st, double *partials) {
    __shared__ double lds[(BS / 64) * 1];
    if (!xcd_valid(P.nblk)) return;
    const int64_t blk = xcd_block(P.nblk);
    double v[1] = {0.0};
    const int64_t n3 = g.n3;
    int zero = 0;
    const Operand<double, false> X{x, nullptr};
    FOR_CELLS(BS, P, blk, c) {
        double amax = 0.0;
#pragma unroll
        for (int p = 0; p < 7; ++p) amax = fmax(amax, fabs(a[p * n3 + c]));
        if (amax > 0.0) {
            const double oam = 1.0 / amax;
#pragma unroll
            for (int p = 0; p < 7; ++p) a[p * n3 + c] = a[p * n3 + c] * oam;
            b[c] = b[c] * oam;
        } else {
            zero = 1;
        }
        const Nb n = neighbours_fast(g, c);
        const double rr = b[c] - stencil_row<double>(g, a, X, c, n);
        r[c] = rr;
        if (g.dot_all || n.interior) v[0] += rr * rr;
    }
    if (zero) atomicOr(&st->zero_row, 1);
    block_partial<BS, 1>(v, partials, P.nblk, blk, lds);
}

// Rtilde = R + (2U-1)*1e-6*Rnorm0 (leq_bicgs.f:298-305) and rho_0 = <Rtilde,R>
// (:318).  For the mixed path (T = float) R and Rtilde are rounded from the
// FP64 residual and the FP32 correction accumulator is cleared.
template <int BS, typename T>
__global__ void __launch_bounds__(BS) k_shadow(DGrid g, Part P, const double *__restrict__ r64, T *__restrict__ r,
                                               T *__restrict__ rt, T *__restrict__ y,
                                               const double *__restrict__ noise, uint64_t seed,
                                               const SolverState *st, double *partials) {
    __shared__ double lds[(BS / 64) * 1];
    if (st->done) return;
    if (!xcd_valid(P.nblk)) return;
    const int64_t blk = xcd_block(P.nblk);
    double v[1] = {0.0};
    const double rn0 = st->rnorm0;
    FOR_CELLS(BS, P, blk, c) {
        const double u = noise ? noise[c] : urand(seed, 100, (uint64_t)c);
        const double rc = r64[c];
        const T t = (T)(rc + (2.0 * u - 1.0) * 1.0e-6 * rn0);
        rt[c] = t;
        const T rT = (T)rc;
        if (y) {
            r[c] = rT;
            y[c] = (T)0;
        }
        if (g.dot_all || is_interior(g, c)) v[0] += (double)t * (double)rT;
    }
    block_partial<BS, 1>(v, partials, P.nblk, blk, lds);
}

// P = R (i=1, or after a restart) or P = R + beta*(P - omega*V)
// (leq_bicgs.f:334-343).  A restart also resets Rtilde = R.
template <int BS, typename T>
__global__ void __launch_bounds__(BS) k_p_update(DGrid g, Part P, const T *__restrict__ r, T *__restrict__ p,
                                                 const T *__restrict__ v, T *__restrict__ rt, const SolverState *st) {
    if (st->done) return;
    if (!xcd_valid(P.nblk)) return;
    const int64_t blk = xcd_block(P.nblk);
    const bool restart = st->restart;
    const bool first = st->i == 1 || restart;
    const double beta = st->beta, omega = st->omega;
    FOR_CELLS(BS, P, blk, c) {
        const T rc = r[c];
        if (restart) rt[c] = rc;
        double pv;
        if (first) pv = (double)rc;
        else pv = (double)rc + beta * ((double)p[c] - omega * (double)v[c]);
        p[c] = (T)pv;
    }
}

// MODE 0: V = A*Phat (leq_bicgs.f:355) and <Rtilde,V> (:357).
// MODE 1: T = A*Shat (:409), <T,S>, <T,T> (:411-421).
// DIV: the operand is x/A0 (DIAG), x otherwise (NONE: x = P or S; LINE:
// x = the preconditioned array).
template <int BS, typename T, int MODE, bool DIV>
__global__ void __launch_bounds__(BS) k_matvec_dots(DGrid g, Part P, const T *__restrict__ a,
                                                    const T *__restrict__ x, T *__restrict__ y,
                                                    const T *__restrict__ w, const SolverState *st,
                                                    double *partials) {
    __shared__ double lds[(BS / 64) * 2];
    if (st->done) return;
    if (MODE == 1 && st->s_exit) return;
    if (!xcd_valid(P.nblk)) return;
    const int64_t blk = xcd_block(P.nblk);
    double v[2] = {0.0, 0.0};
    const bool all = MODE == 1 ? (g.dot_all && !st->minimize) : g.dot_all;
    const Operand<T, DIV> X{x, a + (int64_t)DiagPlane<T>::value * g.n3};
    FOR_CELLS(BS, P, blk, c) {
        const Nb n = neighbours_fast(g, c);
        const T yy = Stencil<T>::row(g, a, X, c, n);
        y[c] = yy;
        if (all || n.interior) {
            const double yd = (double)yy, wd = (double)w[c];
            v[0] += (MODE == 0) ? wd * yd : yd * wd;
            if (MODE == 1) v[1] += yd * yd;
        }
    }
    if (MODE == 0) {
        double v1[1] = {v[0]};
        block_partial<BS, 1>(v1, partials, P.nblk, blk, lds);
    } else {
        block_partial<BS, 2>(v, partials, P.nblk, blk, lds);
    }
}

// S = R - alpha*V (leq_bicgs.f:365-367) and ||S||^2 (:373).
template <int BS, typename T>
__global__ void __launch_bounds__(BS) k_s_update(DGrid g, Part P, const T *__restrict__ r, const T *__restrict__ v,
                                                 T *__restrict__ s, const SolverState *st, double *partials) {
    __shared__ double lds[(BS / 64) * 1];
    if (st->done) return;
    if (!xcd_valid(P.nblk)) return;
    const int64_t blk = xcd_block(P.nblk);
    double acc[1] = {0.0};
    const double alpha = st->alpha;
    FOR_CELLS(BS, P, blk, c) {
        const T sv = (T)((double)r[c] - alpha * (double)v[c]);
        s[c] = sv;
        if (g.dot_all || is_interior(g, c)) acc[0] += (double)sv * (double)sv;
    }
    block_partial<BS, 1>(acc, partials, P.nblk, blk, lds);
}

// Var += alpha*Phat + omega*Shat; R = S - omega*T (leq_bicgs.f:432-436);
// ||R||^2 (:440) and the next <Rtilde,R> (:318).  After the S-norm exit only
// Var += alpha*Phat is applied (:374-376).  ph/sh are P/S (DIV: divided by
// A0 here) or the preconditioned arrays.  In the mixed path x is the FP32
// correction accumulator, restarted at 0 after a reliable update.
template <int BS, typename T, bool DIV>
__global__ void __launch_bounds__(BS) k_xr_update(DGrid g, Part P, T *__restrict__ x, const T *__restrict__ ph,
                                                  const T *__restrict__ sh, const T *__restrict__ a0,
                                                  const T *__restrict__ s, const T *__restrict__ t,
                                                  T *__restrict__ r, const T *__restrict__ rt,
                                                  const SolverState *st, double *partials) {
    __shared__ double lds[(BS / 64) * 2];
    if (st->done) return;
    if (!xcd_valid(P.nblk)) return;
    const int64_t blk = xcd_block(P.nblk);
    double acc[2] = {0.0, 0.0};
    const double alpha = st->alpha, omega = st->omega;
    const bool sexit = st->s_exit;
    const bool yr = st->y_reset;
    FOR_CELLS(BS, P, blk, c) {
        const double x0 = yr ? 0.0 : (double)x[c];
        double phv, d0 = 1.0;
        if (DIV) {
            d0 = (double)a0[c];
            phv = (double)(T)((double)ph[c] / d0);
        } else {
            phv = (double)ph[c];
        }
        if (sexit) {
            x[c] = (T)(x0 + alpha * phv);
        } else {
            const double shv = DIV ? (double)(T)((double)sh[c] / d0) : (double)sh[c];
            x[c] = (T)(x0 + alpha * phv + omega * shv);
            const T rv = (T)((double)s[c] - omega * (double)t[c]);
            r[c] = rv;
            if (g.dot_all || is_interior(g, c)) {
                acc[0] += (double)rv * (double)rv;
                acc[1] += (double)rt[c] * (double)rv;
            }
        }
    }
    block_partial<BS, 2>(acc, partials, P.nblk, blk, lds);
}

// Reliable update (mixed path): xo = xi + y, r64 = b - A64*(xi + y) with the
// neighbours' xi + y formed on the fly (xi is read-only here, xo is the other
// buffer), r = (float)r64; partial ||r64||^2, <Rtilde, r> and <r, r>.
template <int BS>
__global__ void __launch_bounds__(BS) k_reliable_update(DGrid g, Part P, const double *__restrict__ a,
                                                        const double *__restrict__ b, const double *__restrict__ x0,
                                                        const double *__restrict__ x1, double *xo0, double *xo1,
                                                        const float *__restrict__ y, float *__restrict__ r,
                                                        const float *__restrict__ rt, const SolverState *st,
                                                        double *partials) {
    __shared__ double lds[(BS / 64) * 3];
    if (st->done || !st->need_update) return;
    if (!xcd_valid(P.nblk)) return;
    const int64_t blk = xcd_block(P.nblk);
    const int cur = st->x_cur;
    const double *__restrict__ xi = cur ? x1 : x0;
    double *__restrict__ xo = cur ? xo0 : xo1;
    double acc[3] = {0.0, 0.0, 0.0};
    struct XY {
        const double *__restrict__ xi;
        const float *__restrict__ y;
        __device__ __forceinline__ double operator()(int64_t q) const { return xi[q] + (double)y[q]; }
    } X{xi, y};
    FOR_CELLS(BS, P, blk, c) {
        const Nb n = neighbours_fast(g, c);
        const double rr = b[c] - stencil_row<double>(g, a, X, c, n);
        xo[c] = X(c);
        const float r32 = (float)rr;
        r[c] = r32;
        if (g.dot_all || n.interior) {
            acc[0] += rr * rr;
            acc[1] += (double)rt[c] * (double)r32;
            acc[2] += (double)r32 * (double)r32;
        }
    }
    block_partial<BS, 3>(acc, partials, P.nblk, blk, lds);
}

// FP32 difference-layout copy of the (scaled) FP64 matrix: off-diagonals
// rounded, plane 3 = A0 + sum(off-diagonals) formed in FP64, plane 7 = A0.
// A coupling clamped to the cell itself contributes a_l*(x_c - x_c) = 0 in
// the difference form and a_l*x_c through the excess, as in LEQ_MATVEC.
template <int BS>
__global__ void __launch_bounds__(BS) k_prepare_a32(DGrid g, Part P, const double *__restrict__ a,
                                                    float *__restrict__ a32) {
    if (!xcd_valid(P.nblk)) return;
    const int64_t blk = xcd_block(P.nblk);
    const int64_t n3 = g.n3;
    FOR_CELLS(BS, P, blk, c) {
        double e = a[3 * n3 + c];
#pragma unroll
        for (int p = 0; p < 7; ++p) {
            if (p == 3) continue;
            const double v = a[p * n3 + c];
            a32[p * n3 + c] = (float)v;
            e += v;
        }
        a32[3 * n3 + c] = (float)e;
        a32[7 * n3 + c] = (float)a[3 * n3 + c];
    }
}

// ---- LEQ_MATVEC / LEQ_MSOLVE1 / dot / copy as standalone kernels ----------

template <int BS, typename T>
__global__ void __launch_bounds__(BS) k_matvec(DGrid g, Part P, const T *__restrict__ a, const T *__restrict__ x,
                                               T *__restrict__ y) {
    if (!xcd_valid(P.nblk)) return;
    const int64_t blk = xcd_block(P.nblk);
    const Operand<T, false> X{x, nullptr};
    FOR_CELLS(BS, P, blk, c) {
        const Nb n = neighbours_fast(g, c);
        y[c] = stencil_row<T>(g, a, X, c, n);
    }
}

template <int BS, typename T>
__global__ void __launch_bounds__(BS) k_diag(DGrid g, Part P, const T *__restrict__ a, const T *__restrict__ r,
                                             T *__restrict__ z) {
    if (!xcd_valid(P.nblk)) return;
    const int64_t blk = xcd_block(P.nblk);
    FOR_CELLS(BS, P, blk, c) z[c] = (T)((double)r[c] / (double)a[3 * g.n3 + c]);
}

template <int BS, typename T>
__global__ void __launch_bounds__(BS) k_dot(DGrid g, Part P, const T *__restrict__ x, const T *__restrict__ y,
                                            int interior_only, double *partials) {
    __shared__ double lds[(BS / 64) * 1];
    if (!xcd_valid(P.nblk)) return;
    const int64_t blk = xcd_block(P.nblk);
    double v[1] = {0.0};
    FOR_CELLS(BS, P, blk, c) {
        if (!interior_only || is_interior(g, c)) v[0] += (double)x[c] * (double)y[c];
    }
    block_partial<BS, 1>(v, partials, P.nblk, blk, lds);
}

template <int BS, typename T>
__global__ void __launch_bounds__(BS) k_copy_if(Part P, const T *__restrict__ src, T *__restrict__ dst,
                                                const SolverState *st, int skip_on_sexit) {
    if (st) {
        if (st->done) return;
        if (skip_on_sexit && st->s_exit) return;
    }
    if (!xcd_valid(P.nblk)) return;
    const int64_t blk = xcd_block(P.nblk);
    FOR_CELLS(BS, P, blk, c) dst[c] = src[c];
}

}  // namespace mfix
