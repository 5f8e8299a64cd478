// Streaming phase kernels of one BiCGSTAB iteration (LEQ_BICGS0,
// model/leq_bicgs.f:316-466) and its set-up (:194-305), for one or many
// independent systems per launch.
//
// A launch covers a table of systems (Sys<T>) through a block map: hardware
// block b works on system map[b].x, logical block map[b].y (or returns at
// once for a padding entry).  Each system owns a contiguous, 8-aligned range
// of hardware blocks whose logical blocks are permuted so that XCD x (blocks
// b % 8 == x) holds one contiguous slab of that system's cells.  Every
// system keeps its own SolverState and partial sums and its blocks return as
// soon as it is done, so converged systems cost nothing in later launches
// and a solve gives the same bits alone or in a batch.  A single solve is a
// table of one.
//
// With the DIAG preconditioner (LEQ_MSOLVE1, leqsol_mod.f:729-744) Phat =
// P/A0 and Shat = S/A0 are never stored: the stencil kernels divide each
// operand by its A0 as they load it and the solution update divides again,
// which rounds exactly like the reference's separate division pass and saves
// two vector writes and reads per preconditioner application.
#pragma once
#include "leq_kernels.h"

namespace mfix {

// One system of a launch.  T is the Krylov storage type (double: plain FP64
// solve; float: inner solve of the mixed path).  The FP64 arrays (matrix,
// rhs, solution buffers, true residual) sit beside the T ones; for a plain
// FP64 solve a64 aliases a, b64 is B_M, xa is VAR and r64 is r.
template <typename T>
struct Sys {
    DGrid g;
    Part P;                // cell split: P.nblk logical blocks of P.chunk cells
    int nfin;              // blocks of this system in finish launches
    int pc;                // PC_NONE / PC_DIAG / PC_LINE
    int itmax, minimize, icheck, mixed;
    double tol, delta;
    const T *a;            // reference layout (FP64) or difference layout (FP32)
    T *x, *r, *rt, *p, *s, *t, *v, *phat, *shat;  // x: VAR (FP64 solve only)
    double *y;             // mixed: FP64 correction accumulator since the last reliable update
    T *line_scratch;       // 4 planes of DGTSV scratch (LINE)
    SolverState *st;
    double *partials;      // kMaxPartials * P.nblk
    double *a64, *b64, *xa, *xb, *r64;
    const double *noise;   // optional U[0,1) array for Rtilde
    uint64_t seed;
};

// Map entry -> (system, logical block); returns from the kernel for padding.
#define SYS_BLOCK(T)                             \
    const int2 e_ = map[blockIdx.x];             \
    if (e_.x < 0) return;                        \
    const Sys<T> &S = sys[e_.x];                 \
    const int64_t blk = e_.y;                    \
    const DGrid &g = S.g;                        \
    const Part &P = S.P;

// Fresh solver state of the listed systems (the whole struct is rewritten,
// so nothing of a previous solve carries over).
template <typename T>
__global__ void k_init_states(const Sys<T> *__restrict__ sys, const int *__restrict__ ids, int nids) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nids) return;
    const Sys<T> &S = sys[ids[q]];
    SolverState s = {};
    s.tol = S.tol;
    s.itmax = S.itmax;
    s.i = 1;
    s.iter_var = 1;
    s.minimize = S.minimize;
    s.icheck = S.icheck > 0 ? S.icheck : 1;
    s.mixed = S.mixed;
    s.delta = S.delta;
    s.true_rnorm = -1.0;
    *S.st = s;
}

// Unit scaling (leq_bicgs.f:194-233) fused with R = B - A*Var (:283-287) and
// the partial sums of <R,R> (:291).  Each row is scaled by its own max-norm,
// so the residual row uses only this thread's coefficients.  FP64 arrays.
template <int BS, typename T>
__global__ void __launch_bounds__(BS) k_setup_residual(const Sys<T> *__restrict__ sys, const int2 *__restrict__ map) {
    __shared__ double lds[(BS / 64) * 1];
    SYS_BLOCK(T)
    double *__restrict__ a = S.a64;
    double *__restrict__ b = S.b64;
    double *__restrict__ r = S.r64;
    double v[1] = {0.0};
    const int64_t n3 = g.n3;
    int zero = 0;
    const Operand<double, false> X{S.xa, nullptr};
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
    if (zero) atomicOr(&S.st->zero_row, 1);
    block_partial<BS, 1>(v, S.partials, P.nblk, blk, lds);
}

// Rtilde = R + (2U-1)*1e-6*Rnorm0 (leq_bicgs.f:298-305) and rho_0 = <Rtilde,R>
// (:318).  For the mixed path (T = float) R and Rtilde are rounded from the
// FP64 residual and the FP32 correction accumulator is cleared.
template <int BS, typename T>
__global__ void __launch_bounds__(BS) k_shadow(const Sys<T> *__restrict__ sys, const int2 *__restrict__ map) {
    __shared__ double lds[(BS / 64) * 1];
    SYS_BLOCK(T)
    if (S.st->done) return;
    double v[1] = {0.0};
    const double rn0 = S.st->rnorm0;
    const double *__restrict__ r64 = S.r64;
    T *__restrict__ rt = S.rt;
    const bool mixed = S.mixed;
    FOR_CELLS(BS, P, blk, c) {
        const double u = S.noise ? S.noise[c] : urand(S.seed, 100, (uint64_t)c);
        const double rc = r64[c];
        const T t = (T)(rc + (2.0 * u - 1.0) * 1.0e-6 * rn0);
        rt[c] = t;
        const T rT = (T)rc;
        if (mixed) {
            S.r[c] = rT;
            S.y[c] = 0.0;
        }
        if (g.dot_all || is_interior(g, c)) v[0] += (double)t * (double)rT;
    }
    block_partial<BS, 1>(v, S.partials, P.nblk, blk, lds);
}

// P = R (i=1, or after a restart) or P = R + beta*(P - omega*V)
// (leq_bicgs.f:334-343).  A restart also resets Rtilde = R.  For LINE the
// preconditioner's start value Phat = P (z = r, leqsol_mod.f:341-357) is
// written in the same pass.
template <int BS, typename T>
__global__ void __launch_bounds__(BS) k_p_update(const Sys<T> *__restrict__ sys, const int2 *__restrict__ map) {
    SYS_BLOCK(T)
    const SolverState *st = S.st;
    if (st->done) return;
    const bool restart = st->restart;
    const bool first = st->i == 1 || restart;
    const double beta = st->beta, omega = st->omega;
    const T *__restrict__ r = S.r;
    T *__restrict__ p = S.p;
    const T *__restrict__ v = S.v;
    T *__restrict__ ph = S.pc == PC_LINE ? S.phat : nullptr;
    FOR_CELLS(BS, P, blk, c) {
        const T rc = r[c];
        if (restart) S.rt[c] = rc;
        double pv;
        if (first) pv = (double)rc;
        else pv = (double)rc + beta * ((double)p[c] - omega * (double)v[c]);
        p[c] = (T)pv;
        if (ph) ph[c] = (T)pv;
    }
}

// MODE 0: V = A*Phat (leq_bicgs.f:355) and <Rtilde,V> (:357).
// MODE 1: T = A*Shat (:409), <T,S>, <T,T> (:411-421).
// The operand is P/A0 or S/A0 for DIAG, the preconditioned array for LINE,
// and P or S for NONE.
template <int BS, typename T, int MODE>
__global__ void __launch_bounds__(BS) k_matvec_dots(const Sys<T> *__restrict__ sys, const int2 *__restrict__ map) {
    __shared__ double lds[(BS / 64) * 2];
    SYS_BLOCK(T)
    const SolverState *st = S.st;
    if (st->done) return;
    if (MODE == 1 && st->s_exit) return;
    double v[2] = {0.0, 0.0};
    const bool all = MODE == 1 ? (g.dot_all && !st->minimize) : g.dot_all;
    const T *__restrict__ a = S.a;
    const T *__restrict__ a0 = a + (int64_t)DiagPlane<T>::value * g.n3;
    T *__restrict__ y = MODE == 0 ? S.v : S.t;
    const T *__restrict__ w = MODE == 0 ? S.rt : S.s;
    const T *__restrict__ xin = MODE == 0 ? (S.pc == PC_LINE ? S.phat : S.p) : (S.pc == PC_LINE ? S.shat : S.s);
    auto body = [&](auto X) {
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
    };
    if (S.pc == PC_DIAG) body(Operand<T, true>{xin, a0});
    else body(Operand<T, false>{xin, nullptr});
    if (MODE == 0) {
        double v1[1] = {v[0]};
        block_partial<BS, 1>(v1, S.partials, P.nblk, blk, lds);
    } else {
        block_partial<BS, 2>(v, S.partials, P.nblk, blk, lds);
    }
}

// S = R - alpha*V (leq_bicgs.f:365-367) and ||S||^2 (:373); for LINE also
// the preconditioner's start value Shat = S.
template <int BS, typename T>
__global__ void __launch_bounds__(BS) k_s_update(const Sys<T> *__restrict__ sys, const int2 *__restrict__ map) {
    __shared__ double lds[(BS / 64) * 1];
    SYS_BLOCK(T)
    const SolverState *st = S.st;
    if (st->done) return;
    double acc[1] = {0.0};
    const double alpha = st->alpha;
    const T *__restrict__ r = S.r;
    const T *__restrict__ v = S.v;
    T *__restrict__ s = S.s;
    T *__restrict__ sh = S.pc == PC_LINE ? S.shat : nullptr;
    FOR_CELLS(BS, P, blk, c) {
        const T sv = (T)((double)r[c] - alpha * (double)v[c]);
        s[c] = sv;
        if (sh) sh[c] = sv;
        if (g.dot_all || is_interior(g, c)) acc[0] += (double)sv * (double)sv;
    }
    block_partial<BS, 1>(acc, S.partials, P.nblk, blk, lds);
}

// Var += alpha*Phat + omega*Shat; R = S - omega*T (leq_bicgs.f:432-436);
// ||R||^2 (:440) and the next <Rtilde,R> (:318).  After the S-norm exit only
// Var += alpha*Phat is applied (:374-376).  In the mixed path the update goes
// to the FP64 correction accumulator y (restarted at 0 after a reliable
// update) with the same FP32-rounded Phat/Shat the stencil used: keeping y
// in FP64 stops its rounding from polluting every reliable update.
template <int BS, typename T>
__global__ void __launch_bounds__(BS) k_xr_update(const Sys<T> *__restrict__ sys, const int2 *__restrict__ map) {
    __shared__ double lds[(BS / 64) * 2];
    SYS_BLOCK(T)
    const SolverState *st = S.st;
    if (st->done) return;
    double acc[2] = {0.0, 0.0};
    const double alpha = st->alpha, omega = st->omega;
    const bool sexit = st->s_exit;
    const bool yr = st->y_reset;
    const bool div = S.pc == PC_DIAG;
    const T *__restrict__ ph = S.pc == PC_LINE ? S.phat : S.p;
    const T *__restrict__ sh = S.pc == PC_LINE ? S.shat : S.s;
    const T *__restrict__ a0 = S.a + (int64_t)DiagPlane<T>::value * g.n3;
    T *__restrict__ x = S.x;
    double *__restrict__ y = S.y;
    const bool mixed = S.mixed;
    const T *__restrict__ s = S.s;
    const T *__restrict__ t = S.t;
    T *__restrict__ r = S.r;
    const T *__restrict__ rt = S.rt;
    FOR_CELLS(BS, P, blk, c) {
        const double x0 = mixed ? (yr ? 0.0 : y[c]) : (double)x[c];
        double phv, d0 = 1.0;
        if (div) {
            d0 = (double)a0[c];
            phv = (double)(T)((double)ph[c] / d0);
        } else {
            phv = (double)ph[c];
        }
        if (sexit) {
            if (mixed) y[c] = x0 + alpha * phv;
            else x[c] = (T)(x0 + alpha * phv);
        } else {
            const double shv = div ? (double)(T)((double)sh[c] / d0) : (double)sh[c];
            if (mixed) y[c] = x0 + alpha * phv + omega * shv;
            else x[c] = (T)(x0 + alpha * phv + omega * shv);
            const T rv = (T)((double)s[c] - omega * (double)t[c]);
            r[c] = rv;
            if (g.dot_all || is_interior(g, c)) {
                acc[0] += (double)rv * (double)rv;
                acc[1] += (double)rt[c] * (double)rv;
            }
        }
    }
    block_partial<BS, 2>(acc, S.partials, P.nblk, blk, lds);
}

// Reliable update (mixed path): xo = xi + y, r64 = b - A64*(xi + y) with the
// neighbours' xi + y formed on the fly (xi is read-only here, xo is the other
// buffer), r = (float)r64; partial ||r64||^2, <Rtilde, r> and <r, r>.  After
// a breakdown (discard_y) y is dropped: xo = xi.
template <int BS>
__global__ void __launch_bounds__(BS) k_reliable_update(const Sys<float> *__restrict__ sys,
                                                        const int2 *__restrict__ map) {
    __shared__ double lds[(BS / 64) * 3];
    SYS_BLOCK(float)
    const SolverState *st = S.st;
    if (st->done || !st->need_update) return;
    const int cur = st->x_cur;
    const double *__restrict__ xi = cur ? S.xb : S.xa;
    double *__restrict__ xo = cur ? S.xa : S.xb;
    const double *__restrict__ a = S.a64;
    const double *__restrict__ b = S.b64;
    float *__restrict__ r = S.r;
    const float *__restrict__ rt = S.rt;
    double acc[3] = {0.0, 0.0, 0.0};
    struct XY {
        const double *__restrict__ xi;
        const double *__restrict__ y;
        __device__ __forceinline__ double operator()(int64_t q) const { return y ? xi[q] + y[q] : xi[q]; }
    } X{xi, st->discard_y ? nullptr : S.y};
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
    block_partial<BS, 3>(acc, S.partials, P.nblk, blk, lds);
}

// FP32 difference-layout copy of the (scaled) FP64 matrix: off-diagonals
// rounded, plane 3 = A0 + sum(off-diagonals) formed in FP64, plane 7 = A0.
// A coupling clamped to the cell itself contributes a_l*(x_c - x_c) = 0 in
// the difference form and a_l*x_c through the excess, as in LEQ_MATVEC.
template <int BS>
__global__ void __launch_bounds__(BS) k_prepare_a32(const Sys<float> *__restrict__ sys, const int2 *__restrict__ map) {
    SYS_BLOCK(float)
    if (S.st->done) return;
    const double *__restrict__ a = S.a64;
    float *__restrict__ a32 = const_cast<float *>(S.a);
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

// Finish kernels over a table: finish block b serves system map[b].x as its
// block map[b].y of S.nfin.
template <typename T, int F, int NV>
__global__ void __launch_bounds__(kFinishBlock) k_finish_sys(const Sys<T> *__restrict__ sys,
                                                             const int2 *__restrict__ map) {
    __shared__ double lds[(kFinishBlock / 64) * 4];
    const int2 e = map[blockIdx.x];
    if (e.x < 0) return;
    const Sys<T> &S = sys[e.x];
    SolverState *st = S.st;
    if (fin_skips<F>(st)) return;
    double tot[NV];
    if (finish_sum<NV>(S.partials, S.P.nblk, e.y, S.nfin, st->l2, &st->fin_counter, lds, tot) && threadIdx.x == 0)
        fin_apply<F, NV>(st, tot, nullptr);
}

// Red-black line phase along J for every LINE system of a table (RSRS and
// the 2-D R sweep): block b solves lines [map[b].y*64, +64) of colour
// `param` of system map[b].x, numbered densely by color_line.  which = 0
// preconditions P into Phat, 1 S into Shat (skipped after the S-norm exit).
template <typename T>
__global__ void __launch_bounds__(64) k_line_color_sys(const Sys<T> *__restrict__ sys, const int2 *__restrict__ map,
                                                       int param, int which) {
    const int2 e = map[blockIdx.x];
    if (e.x < 0) return;
    const Sys<T> &S = sys[e.x];
    const SolverState *st = S.st;
    if (st->done || (which == 1 && st->s_exit)) return;
    const DGrid &g = S.g;
    const LineGeom L = line_geom(g, AX_J);
    int u, v;
    const int64_t lane = (int64_t)e.y * 64 + threadIdx.x;
    if (!color_line(L.nu, L.nv, g.do_k ? 1 : 0, param, lane, u, v)) return;
    const int64_t c0 = (int64_t)u * L.su + (int64_t)v * L.sv;
    // lanes of this system: 64 per block, blocks for the larger colour
    const int64_t nl = (((int64_t)L.nu * L.nv + 1) / 2 + 63) / 64 * 64;
    solve_line<T>(g, L, AX_J, c0, S.a, which == 0 ? S.p : S.s, which == 0 ? S.phat : S.shat, S.line_scratch, lane,
                  nl, 0, nullptr);
}

// Number of the listed systems still running.
template <typename T>
__global__ void k_count_running(const Sys<T> *__restrict__ sys, const int *__restrict__ ids, int nids, int *out) {
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    for (int q = threadIdx.x; q < nids; q += blockDim.x)
        if (!sys[ids[q]].st->done) atomicAdd(&cnt, 1);
    __syncthreads();
    if (threadIdx.x == 0) *out = cnt;
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

// Copy (LINE start value z = r of the standalone LEQ_MSOLVE entry,
// leqsol_mod.f:341-357, and the traffic-calibration copy); with `st` it is
// skipped once the solve is done or, with skip_on_sexit, after the S-norm
// exit.
template <int BS, typename T>
__global__ void __launch_bounds__(BS) k_copy_if(Part P, const T *__restrict__ src, T *__restrict__ dst,
                                                const SolverState *st, int skip_on_sexit) {
    if (st && (st->done || (skip_on_sexit && st->s_exit))) return;
    if (!xcd_valid(P.nblk)) return;
    const int64_t blk = xcd_block(P.nblk);
    FOR_CELLS(BS, P, blk, c) dst[c] = src[c];
}

// One-system finish (standalone dot product).
template <int F, int NV>
__global__ void __launch_bounds__(kFinishBlock) k_finish(SolverState *st, const double *__restrict__ partials,
                                                         int64_t nblk, double *out) {
    __shared__ double lds[(kFinishBlock / 64) * 4];
    if (fin_skips<F>(st)) return;
    double tot[NV];
    if (finish_sum<NV>(partials, nblk, (int)blockIdx.x, (int)gridDim.x, st->l2, &st->fin_counter, lds, tot) &&
        threadIdx.x == 0)
        fin_apply<F, NV>(st, tot, out);
}

// Blank state for the standalone dot/scale entries.
static __global__ void k_init_state(SolverState *st) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        SolverState s = {};
        s.icheck = 1;
        s.itmax = 1;
        s.true_rnorm = -1.0;
        *st = s;
    }
}

}  // namespace mfix
