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
#include <hip/hip_ext.h>

#include "leq_kernels.h"

namespace mfix {

// Launch `kernel`; given a pair of events the dispatch itself records the
// kernel's start and end in them (hipExtLaunchKernelGGL), so a per-launch
// time is the kernel's, without the launch latency that events recorded on
// the stream around it would add.  No events: a plain launch (the only form
// used inside a captured graph).
template <typename... Args, typename... P>
inline void launch_timed(hipEvent_t e0, hipEvent_t e1, void (*kernel)(Args...), dim3 grid, dim3 block, hipStream_t s,
                         P... args) {
    if (e0) hipExtLaunchKernelGGL(kernel, grid, block, 0, s, e0, e1, 0, static_cast<Args>(args)...);
    else hipLaunchKernelGGL(kernel, grid, block, 0, s, static_cast<Args>(args)...);
}

// Work the set-up pass does on a row before it scales it
// (mfix_leq_resid_relax_bicgs_dev): the CALC_RESID row of the equation as
// passed (partials of NUM, DEN, fluid cells and the largest NUM1 per logical
// block), then UNDER_RELAX_S.
struct Prologue {
    const int *flag = nullptr;  // FLAG; null: no prologue
    int resid_kind = -1;        // RESID_C / RESID_S / RESID_PP; -1: no residual
    double resid_tol = 0.0, dil_ep_s = 0.0;
    const double *ep_s = nullptr;
    double *rpart = nullptr;    // kResidVals x P.nblk partial slots
    int relax = 0;              // UNDER_RELAX_S: A0 *= f1, B += A0 VAR f2 (f1 = 1/UR, f2 = f1 - 1)
    double f1 = 1.0, f2 = 0.0;
};

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
    int itmax, minimize, icheck, mixed, max_updates;
    int unit_diag;         // mixed DIAG: negate where every 1/A0 is -1 (else read the plane)
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
    Prologue pro;          // residual and under-relaxation in the set-up pass
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
    s.max_updates = S.max_updates;
    s.restart_always = S.mixed && S.g.n3 >= (int64_t(1) << 18);
    s.nonunit_diag = !S.mixed || !S.unit_diag;  // the FP64 solver always divides, like the reference
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
    auto *__restrict__ a = to_global(S.a64);
    auto *__restrict__ b = to_global(S.b64);
    auto *__restrict__ r = to_global(S.r64);
    double v[1] = {0.0};
    const int64_t n3 = g.n3;
    int zero = 0;
    const gptr<const double> xv = to_global((const double *)S.xa);
    const Operand<double, OP_PLAIN, gptr<const double>> X{xv, nullptr};
    // prologue (uniform over the block: one system)
    const Prologue &pr = S.pro;
    const bool do_res = pr.flag && pr.resid_kind >= 0, do_rel = pr.flag && pr.relax;
    const gptr<const int> flag = to_global(pr.flag);
    const double rinit = pr.resid_kind == RESID_PP ? -1.0 : 0.0;  // RESID_IJK scan start (calc_resid.f:114, :423)
    double rs[3] = {0.0, 0.0, 0.0}, bv = rinit, bi = 1.0e300;
    FOR_CELLS(BS, P, blk, c) {
        // the row, B and VAR(c) in registers for all three steps
        double ar[7];
#pragma unroll
        for (int p = 0; p < 7; ++p) ar[p] = a[p * n3 + c];
        double bc = b[c];
        const double vc = xv[c];
        if (do_res) {
            // CALC_RESID row of the equation as passed
            double num1, den1;
            if (resid_row_vals(g, pr.resid_kind, pr.resid_tol, pr.ep_s, pr.dil_ep_s, ar, bc, vc, xv, flag, c, num1,
                               den1)) {
                rs[0] += num1;
                rs[1] += den1;
                rs[2] += 1.0;
                if (num1 > rinit) resid_better(bv, bi, num1, (double)c);
            }
        }
        if (do_rel) {
            // UNDER_RELAX_S (under_relax.f:64-76), as k_under_relax_s
            const double ap = ar[3];
            const double na = ap * pr.f1, nb = bc + ap * vc * pr.f2;
            if (fluid_at(flag[c]) && ap != -1.0) {
                ar[3] = na;
                bc = nb;
            }
        }
        double amax = 0.0;
#pragma unroll
        for (int p = 0; p < 7; ++p) amax = fmax(amax, fabs(ar[p]));
        if (amax > 0.0) {
            const double oam = 1.0 / amax;
#pragma unroll
            for (int p = 0; p < 7; ++p) ar[p] = ar[p] * oam;
            bc = bc * oam;
        } else {
            zero = 1;
        }
        if (amax > 0.0 || do_rel) {
#pragma unroll
            for (int p = 0; p < 7; ++p) a[p * n3 + c] = ar[p];
            b[c] = bc;
        }
        const Nb n = neighbours_fast(g, c);
        const double rr = bc - stencil_combine<double>(g.do_k, ar[0], ar[1], ar[2], ar[3], ar[4], ar[5], ar[6], X(n.km),
                                                       X(n.im), X(n.jm), vc, X(n.jp), X(n.ip), X(n.kp));
        r[c] = rr;
        if (g.dot_all || n.interior) v[0] += rr * rr;
    }
    if (zero) atomicOr((int *)&to_global(S.st)->zero_row, 1);
    if (do_res) {
        // the block's residual partials: sums in block order, the largest
        // NUM1 by wave shuffles then the waves in order (k_resid's reduction)
        __shared__ double lr[(BS / 64) * 3], lbv[BS / 64], lbi[BS / 64];
        block_sum<BS, 3>(rs, lr);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_down(bv, off, 64), oi = __shfl_down(bi, off, 64);
            resid_better(bv, bi, ov, oi);
        }
        if ((threadIdx.x & 63) == 0) {
            lbv[threadIdx.x >> 6] = bv;
            lbi[threadIdx.x >> 6] = bi;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double bvv = lbv[0], bii = lbi[0];
            for (int u = 1; u < BS / 64; ++u) resid_better(bvv, bii, lbv[u], lbi[u]);
            auto *__restrict__ rp = to_global(pr.rpart);
            const int64_t nb = P.nblk;
            rp[0 * nb + blk] = rs[0];
            rp[1 * nb + blk] = rs[1];
            rp[2 * nb + blk] = rs[2];
            rp[3 * nb + blk] = bvv;
            rp[4 * nb + blk] = bii;
        }
    }
    block_partial<BS, 1>(v, to_global(S.partials), P.nblk, blk, lds);
}

// Rtilde = R + (2U-1)*1e-6*Rnorm0 (leq_bicgs.f:298-305) and rho_0 = <Rtilde,R>
// (:318).  For the mixed path (T = float) R and Rtilde are rounded from the
// FP64 residual and the FP32 correction accumulator is cleared.
template <int BS, typename T>
__global__ void __launch_bounds__(BS) k_shadow(const Sys<T> *__restrict__ sys, const int2 *__restrict__ map) {
    __shared__ double lds[(BS / 64) * 1];
    SYS_BLOCK(T)
    if (to_global(S.st)->done) return;
    double v[1] = {0.0};
    const double rn0 = to_global(S.st)->rnorm0;
    auto *__restrict__ r64 = to_global(S.r64);
    auto *__restrict__ rt = to_global(S.rt);
    const bool mixed = S.mixed;
    FOR_CELLS(BS, P, blk, c) {
        const double u = to_global(S.noise) ? to_global(S.noise)[c] : urand(S.seed, 100, (uint64_t)c);
        const double rc = r64[c];
        const T t = (T)(rc + (2.0 * u - 1.0) * 1.0e-6 * rn0);
        rt[c] = t;
        const T rT = (T)rc;
        if (mixed) {
            to_global(S.r)[c] = rT;
            to_global(S.y)[c] = 0.0;
        }
        if (g.dot_all || is_interior(g, c)) v[0] += (double)t * (double)rT;
    }
    block_partial<BS, 1>(v, to_global(S.partials), P.nblk, blk, lds);
}

// The three update kernels below run BS/VN threads per logical block of BS
// cells, each thread one 16-byte group of VN cells (VN = 4 FP32, 2 FP64);
// a logical block of another size falls back to one cell per step.
//
// P = R (i=1, or after a restart) or P = R + beta*(P - omega*V)
// (leq_bicgs.f:334-343).  A restart also resets Rtilde = R.  For LINE the
// preconditioner's start value Phat = P (z = r, leqsol_mod.f:341-357) is
// written in the same pass.
template <int BS, typename T>
__global__ void __launch_bounds__(BS / Group16<T>::N) k_p_update(const Sys<T> *__restrict__ sys,
                                                                 const int2 *__restrict__ map) {
    SYS_BLOCK(T)
    auto *__restrict__ st = to_global(S.st);
    if (st->done) return;
    const bool restart = st->restart;
    const bool first = st->i == 1 || restart;
    const double beta = st->beta, omega = st->omega;
    auto *__restrict__ r = to_global(S.r);
    auto *__restrict__ p = to_global(S.p);
    auto *__restrict__ v = to_global(S.v);
    auto *__restrict__ rt = to_global(S.rt);
    const auto ph = S.pc == PC_LINE ? to_global(S.phat) : nullptr;
    auto cell = [&](int64_t c) {
        const T rc = r[c];
        if (restart) rt[c] = rc;
        double pv;
        if (first) pv = (double)rc;
        else pv = (double)rc + beta * ((double)p[c] - omega * (double)v[c]);
        p[c] = (T)pv;
        if (ph) ph[c] = (T)pv;
    };
    constexpr int VN = Group16<T>::N, TB = BS / VN;
    if (P.chunk == (int64_t)BS) {
        const int64_t c0 = blk * P.chunk + (int64_t)VN * threadIdx.x;
        if (c0 + VN <= P.n) {
            const vec_t<T, VN> rv = ldv<T, VN>(r + c0);
            if (restart) stv<T, VN>(rt + c0, rv);
            vec_t<T, VN> pv = rv;
            if (!first) {
                const vec_t<T, VN> po = ldv<T, VN>(p + c0), vo = ldv<T, VN>(v + c0);
#pragma unroll
                for (int l = 0; l < VN; ++l)
                    pv[l] = (T)((double)rv[l] + beta * ((double)po[l] - omega * (double)vo[l]));
            }
            stv<T, VN>(p + c0, pv);
            if (ph) stv<T, VN>(ph + c0, pv);
        } else {
            for (int64_t c = c0; c < P.n; ++c) cell(c);
        }
        return;
    }
    FOR_CELLS(TB, P, blk, c) cell(c);
}

// MODE 0: V = A*Phat (leq_bicgs.f:355) and <Rtilde,V> (:357).
// MODE 1: T = A*Shat (:409), <T,S>, <T,T> (:411-421).
// The operand is P/A0 or S/A0 for DIAG, the preconditioned array for LINE,
// and P or S for NONE.  The S-norm exit (:372-376) is decided with omega,
// after this pass: T is then unused, so the iterates do not change.
// One cell per thread (measured faster here than 16-byte cell groups: the
// operand groups at c +- c1 straddle the 16-byte grid, and a group that
// crosses the end of a J row serializes its cells, stalling its wave).
template <int BS, typename T, int MODE>
__global__ void __launch_bounds__(BS) k_matvec_dots(const Sys<T> *__restrict__ sys, const int2 *__restrict__ map) {
    __shared__ double lds[(BS / 64) * 2];
    __shared__ unsigned arrived;
    SYS_BLOCK(T)
    auto *__restrict__ st = to_global(S.st);
    if (st->done) return;
    if (MODE == 1 && st->s_exit) return;
    if (MODE == 0 && blk == 0 && threadIdx.x == 0) st->k1_launches += 1;  // (timing: working launches)
    partial_begin<BS>(&arrived);
    double v[2] = {0.0, 0.0};
    const bool all = MODE == 1 ? (g.dot_all && !st->minimize) : g.dot_all;
    auto *__restrict__ a = to_global(S.a);
    const auto a0 = a + (int64_t)DiagPlane<T>::value * g.n3;
    auto *__restrict__ y = MODE == 0 ? to_global(S.v) : to_global(S.t);
    auto *__restrict__ w = MODE == 0 ? to_global(S.rt) : to_global(S.s);
    const auto xin = MODE == 0 ? (S.pc == PC_LINE ? to_global(S.phat) : to_global(S.p)) : (S.pc == PC_LINE ? to_global(S.shat) : to_global(S.s));
    auto body = [&](auto X) {
        auto cell = [&](int64_t c) {
            const Nb n = neighbours_fast(g, c);
            const T wc = ld_stream(&w[c]);
            const T yy = Stencil<T>::row(g, a, X, c, n);
            y[c] = yy;
            if (all || n.interior) {
                const double yd = (double)yy, wd = (double)wc;
                v[0] += (MODE == 0) ? wd * yd : yd * wd;
                if (MODE == 1) v[1] += yd * yd;
            }
        };
        FOR_CELLS(BS, P, blk, c) cell(c);
    };
    if (S.pc == PC_DIAG && st->nonunit_diag) body(Operand<T, OP_DIAG, gptr<const T>>{xin, a0});
    else if (S.pc == PC_DIAG) body(Operand<T, OP_NEG, gptr<const T>>{xin, nullptr});
    else body(Operand<T, OP_PLAIN, gptr<const T>>{xin, nullptr});
    if (MODE == 0) {
        double v1[1] = {v[0]};
        block_partial_last<BS, 1>(v1, to_global(S.partials), P.nblk, blk, lds, &arrived);
    } else {
        block_partial_last<BS, 2>(v, to_global(S.partials), P.nblk, blk, lds, &arrived);
    }
}

// k_matvec_dots for the FP32 inner solve with four consecutive cells per
// thread (BS/4 threads over a BS-cell logical block) and 16-byte accesses:
// the seven coefficient planes, the dot partner and the result move as one
// dwordx4 per lane each, and the operand as five (at c, c -+ c1, c -+ c2:
// every I and K neighbour of the thread's four cells lies in the same
// relative position of the neighbouring row and plane).  The J neighbours
// come from the adjacent lanes' groups by lane shuffles; lanes 0 and 63 load
// the one value past their wave's run.  Per four cells that is 14 memory
// instructions where one cell per thread needs 16 per cell; one cell per
// thread streamed at 4.8-5.1 TB/s against 6.3-6.7 TB/s for the update
// passes, whose lanes already move 16 bytes.  A cell on the I or J edge of
// the grid takes itself as the clamped neighbour (set_increments.f:72-112),
// as neighbours_fast does, by a select on its (i, j).  Groups on the first
// or last K plane, past the end of the grid, of a 2-D grid or of a grid
// with a cyclic axis (whose wrapped neighbours are not at fixed offsets)
// take the per-cell path (Stencil<float>::row).  The row arithmetic is
// stencil_combine_diff either way, so every row is bit-identical to the
// per-cell path's; a thread adds its four cells before the wave tree.
// Measured at 256^3 (the void-fraction pressure system, 256-cell blocks):
// 133.5 / 127.8 us per pass against 133.7 / 129.3 us for two cells per
// thread with 4-byte accesses, so the instruction count was not the limit;
// both move 5.6 TB/s of measured HBM traffic.  The seven planes of a cell
// stored together (one 32-byte record per cell, one stream instead of
// seven) ran 290 / 279 us: each record load touched a cache line per lane.
template <int BS, int MODE>
__global__ void __launch_bounds__(BS / 4) k_matvec_dots4(const Sys<float> *__restrict__ sys,
                                                          const int2 *__restrict__ map) {
    constexpr int NT = BS / 4;
    __shared__ double lds[(NT / 64 > 0 ? NT / 64 : 1) * 2];
    __shared__ unsigned arrived;
    SYS_BLOCK(float)
    auto *__restrict__ st = to_global(S.st);
    if (st->done) return;
    if (MODE == 1 && st->s_exit) return;
    if (MODE == 0 && blk == 0 && threadIdx.x == 0) st->k1_launches += 1;  // (timing: working launches)
    partial_begin<NT>(&arrived);
    double v[2] = {0.0, 0.0};
    const bool all = MODE == 1 ? (g.dot_all && !st->minimize) : g.dot_all;
    const int64_t n3 = g.n3;
    auto *__restrict__ a = to_global(S.a);
    const auto a0p = a + (int64_t)DiagPlane<float>::value * n3;
    auto *__restrict__ y = MODE == 0 ? to_global(S.v) : to_global(S.t);
    auto *__restrict__ w = MODE == 0 ? to_global(S.rt) : to_global(S.s);
    const auto xin = MODE == 0 ? (S.pc == PC_LINE ? to_global(S.phat) : to_global(S.p))
                               : (S.pc == PC_LINE ? to_global(S.shat) : to_global(S.s));
    const int op = S.pc != PC_DIAG ? OP_PLAIN : (st->nonunit_diag ? OP_DIAG : OP_NEG);
    const int64_t c0 = blk * P.chunk + 4 * (int64_t)threadIdx.x;
    const int64_t c1 = g.c1, c2 = g.c2;
    const int lane = threadIdx.x & 63;
    // fast path: all four cells exist and lie off the first and last K
    // planes (so every load below is in bounds), on a 3-D grid without
    // cyclic axes
    const bool fast = g.do_k && !(g.cyc_x | g.cyc_y | g.cyc_z) && c0 >= c2 && c0 + 3 + c2 < n3;
    f4v X = 0.f, XIm = 0.f, XIp = 0.f, XKm = 0.f, XKp = 0.f, A[7], W = 0.f;
    auto xop = [&](int64_t q) -> float {
        const float e = xin[q];
        return op == OP_DIAG ? diag_apply(e, (float)a0p[q]) : op == OP_NEG ? -e : e;
    };
    if (fast) {
        // the operand with plain loads: its values are read again as other
        // cells' I and K neighbours, from L2
        X = ldv_keep<float, 4>(xin + c0);
        XIm = ldv_keep<float, 4>(xin + c0 - c1);
        XIp = ldv_keep<float, 4>(xin + c0 + c1);
        XKm = ldv_keep<float, 4>(xin + c0 - c2);
        XKp = ldv_keep<float, 4>(xin + c0 + c2);
#pragma unroll
        for (int p = 0; p < 7; ++p) A[p] = ldv<float, 4>(a + (int64_t)p * n3 + c0);
        W = ldv<float, 4>(w + c0);
        if (op == OP_DIAG) {
            X = X * ldv_keep<float, 4>(a0p + c0);
            XIm = XIm * ldv_keep<float, 4>(a0p + c0 - c1);
            XIp = XIp * ldv_keep<float, 4>(a0p + c0 + c1);
            XKm = XKm * ldv_keep<float, 4>(a0p + c0 - c2);
            XKp = XKp * ldv_keep<float, 4>(a0p + c0 + c2);
        } else if (op == OP_NEG) {
            X = -X;
            XIm = -XIm;
            XIp = -XIp;
            XKm = -XKm;
            XKp = -XKp;
        }
    }
    // J neighbours across a group's edges: the adjacent lane's group when
    // that lane is on the fast path too, else one load (every lane takes part
    // in the shuffles)
    float xl = __shfl_up(X.w, 1, 64), xr = __shfl_down(X.x, 1, 64);
    const int fl = __shfl_up((int)fast, 1, 64), fr = __shfl_down((int)fast, 1, 64);
    if (fast && (lane == 0 || !fl)) xl = xop(c0 - 1);
    if (fast && (lane == 63 || !fr)) xr = xop(c0 + 4);
    if (fast) {
        int i, j, k;
        cell_ijk(g, c0, i, j, k);
        f4v Y;
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            const float xc = X[l];
            const float xjm = j == 0 ? xc : (l == 0 ? xl : X[l - 1]);
            const float xjp = j == g.nj - 1 ? xc : (l == 3 ? xr : X[l + 1]);
            const float xim = i == 0 ? xc : XIm[l];
            const float xip = i == g.ni - 1 ? xc : XIp[l];
            const float yy = stencil_combine_diff(true, A[0][l], A[1][l], A[2][l], A[3][l], A[4][l], A[5][l], A[6][l],
                                                  XKm[l], xim, xjm, xc, xjp, xip, XKp[l]);
            Y[l] = yy;
            const bool interior = i > 0 && i < g.ni - 1 && j > 0 && j < g.nj - 1;  // k is interior here
            if (all || interior) {
                const double yd = (double)yy, wd = (double)W[l];
                v[0] += (MODE == 0) ? wd * yd : yd * wd;
                if (MODE == 1) v[1] += yd * yd;
            }
            if (++j == g.nj) {
                j = 0;
                if (++i == g.ni) i = 0;
            }
        }
        stv<float, 4>(y + c0, Y);
    } else {
        // per-cell path
        const int64_t end = blk * P.chunk + P.chunk < P.n ? blk * P.chunk + P.chunk : P.n;
        auto cells = [&](auto X) {
            for (int64_t c = c0; c < c0 + 4 && c < end; ++c) {
                const Nb nb = neighbours_fast(g, c);
                const float wc = ld_stream(&w[c]);
                const float yy = Stencil<float>::row(g, a, X, c, nb);
                y[c] = yy;
                if (all || nb.interior) {
                    const double yd = (double)yy, wd = (double)wc;
                    v[0] += (MODE == 0) ? wd * yd : yd * wd;
                    if (MODE == 1) v[1] += yd * yd;
                }
            }
        };
        if (op == OP_DIAG) cells(Operand<float, OP_DIAG, gptr<const float>>{xin, a0p});
        else if (op == OP_NEG) cells(Operand<float, OP_NEG, gptr<const float>>{xin, nullptr});
        else cells(Operand<float, OP_PLAIN, gptr<const float>>{xin, nullptr});
    }
    if (MODE == 0) {
        double v1[1] = {v[0]};
        block_partial_last<NT, 1>(v1, to_global(S.partials), P.nblk, blk, lds, &arrived);
    } else {
        block_partial_last<NT, 2>(v, to_global(S.partials), P.nblk, blk, lds, &arrived);
    }
}

// S = R - alpha*V (leq_bicgs.f:365-367) and ||S||^2 (:373); for LINE also
// the preconditioner's start value Shat = S.  The S-norm exit test is applied
// by the finish after the second stencil pass.
template <int BS, typename T>
__global__ void __launch_bounds__(BS / Group16<T>::N) k_s_update(const Sys<T> *__restrict__ sys,
                                                                 const int2 *__restrict__ map) {
    constexpr int VN = Group16<T>::N, TB = BS / VN;
    __shared__ double lds[(TB / 64) * 1];
    SYS_BLOCK(T)
    auto *__restrict__ st = to_global(S.st);
    if (st->done) return;
    double acc[1] = {0.0};
    const double alpha = st->alpha;
    auto *__restrict__ r = to_global(S.r);
    auto *__restrict__ v = to_global(S.v);
    auto *__restrict__ s = to_global(S.s);
    const auto sh = S.pc == PC_LINE ? to_global(S.shat) : nullptr;
    auto cell = [&](int64_t c) {
        const T sv = (T)((double)r[c] - alpha * (double)v[c]);
        s[c] = sv;
        if (sh) sh[c] = sv;
        if (g.dot_all || is_interior(g, c)) acc[0] += (double)sv * (double)sv;
    };

    if (P.chunk == (int64_t)BS) {
        const int64_t c0 = blk * P.chunk + (int64_t)VN * threadIdx.x;
        if (c0 + VN <= P.n) {
            const vec_t<T, VN> rv = ldv<T, VN>(r + c0), vv = ldv<T, VN>(v + c0);
            vec_t<T, VN> sv;
#pragma unroll
            for (int l = 0; l < VN; ++l) sv[l] = (T)((double)rv[l] - alpha * (double)vv[l]);
            stv<T, VN>(s + c0, sv);
            if (sh) stv<T, VN>(sh + c0, sv);
#pragma unroll
            for (int l = 0; l < VN; ++l)
                if (g.dot_all || is_interior(g, c0 + l)) acc[0] += (double)(T)sv[l] * (double)(T)sv[l];
        } else {
            for (int64_t c = c0; c < P.n; ++c) cell(c);
        }
    } else {
        FOR_CELLS(TB, P, blk, c) cell(c);
    }
    // slot 2: the second stencil pass writes slots 0 and 1, and one finish
    // launch after it takes all three
    block_partial<TB, 1>(acc, to_global(S.partials) + 2 * P.nblk, P.nblk, blk, lds);
}

// Var += alpha*Phat + omega*Shat; R = S - omega*T (leq_bicgs.f:432-436);
// ||R||^2 (:440) and the next <Rtilde,R> (:318).  After the S-norm exit only
// Var += alpha*Phat is applied (:374-376).  In the mixed path the update goes
// to the FP64 correction accumulator y (restarted at 0 after a reliable
// update) with the same FP32-rounded Phat/Shat the stencil used: keeping y
// in FP64 stops its rounding from polluting every reliable update.
template <int BS, typename T>
__global__ void __launch_bounds__(BS / Group16<T>::N) k_xr_update(const Sys<T> *__restrict__ sys,
                                                                  const int2 *__restrict__ map) {
    constexpr int VN = Group16<T>::N, TB = BS / VN;
    __shared__ double lds[(TB / 64) * 2];
    SYS_BLOCK(T)
    auto *__restrict__ st = to_global(S.st);
    if (st->done) return;
    double acc[2] = {0.0, 0.0};
    const double alpha = st->alpha, omega = st->omega;
    const bool sexit = st->s_exit;
    const bool yr = st->y_reset;
    const bool div = S.pc == PC_DIAG && st->nonunit_diag;
    const bool neg = S.pc == PC_DIAG && !st->nonunit_diag;
    const auto ph = S.pc == PC_LINE ? to_global(S.phat) : to_global(S.p);
    const auto sh = S.pc == PC_LINE ? to_global(S.shat) : to_global(S.s);
    auto *__restrict__ a0 = to_global(S.a) + (int64_t)DiagPlane<T>::value * g.n3;
    auto *__restrict__ x = to_global(S.x);
    auto *__restrict__ y = to_global(S.y);
    const bool mixed = S.mixed;
    auto *__restrict__ s = to_global(S.s);
    auto *__restrict__ t = to_global(S.t);
    auto *__restrict__ r = to_global(S.r);
    auto *__restrict__ rt = to_global(S.rt);
    // preconditioned Phat/Shat of one cell, rounded as the stencil used them
    auto pre = [&](T z, T d0) -> double { return div ? (double)diag_apply(z, d0) : neg ? -(double)z : (double)z; };
    auto cell = [&](int64_t c) {
        const T rtc = rt[c];
        const double x0 = mixed ? (yr ? 0.0 : y[c]) : (double)x[c];
        const T d0 = div ? (T)a0[c] : (T)1;
        const double phv = pre((T)ph[c], d0);
        if (sexit) {
            if (mixed) y[c] = x0 + alpha * phv;
            else x[c] = (T)(x0 + alpha * phv);
        } else {
            const double shv = pre((T)sh[c], d0);
            if (mixed) y[c] = x0 + alpha * phv + omega * shv;
            else x[c] = (T)(x0 + alpha * phv + omega * shv);
            const T rv = (T)((double)s[c] - omega * (double)t[c]);
            r[c] = rv;
            if (g.dot_all || is_interior(g, c)) {
                acc[0] += (double)rv * (double)rv;
                acc[1] += (double)rtc * (double)rv;
            }
        }
    };

    // (an FP32 table is always the inner solve of the mixed path)
    if (P.chunk == (int64_t)BS && (mixed || sizeof(T) == sizeof(double))) {
        const int64_t c0 = blk * P.chunk + (int64_t)VN * threadIdx.x;
        if (c0 + VN <= P.n) {
            const vec_t<T, VN> PH = ldv<T, VN>(ph + c0);
            vec_t<T, VN> D0 = (T)1;
            if (div) D0 = ldv<T, VN>(a0 + c0);
            vec_t<double, VN> X0 = 0.0;
            if (mixed) {
                if (!yr) X0 = ldv<double, VN>(y + c0);
            } else if constexpr (sizeof(T) == sizeof(double)) {
                X0 = ldv<double, VN>(x + c0);
            }
            vec_t<double, VN> XN;
            if (sexit) {
#pragma unroll
                for (int l = 0; l < VN; ++l) XN[l] = X0[l] + alpha * pre(PH[l], div ? (T)D0[l] : (T)1);
            } else {
                const vec_t<T, VN> SH = ldv<T, VN>(sh + c0), SV = ldv<T, VN>(s + c0), TV = ldv<T, VN>(t + c0);
                const vec_t<T, VN> RT = ldv<T, VN>(rt + c0);
                vec_t<T, VN> RV;
#pragma unroll
                for (int l = 0; l < VN; ++l) {
                    const T d0 = div ? (T)D0[l] : (T)1;
                    XN[l] = X0[l] + alpha * pre(PH[l], d0) + omega * pre(SH[l], d0);
                    RV[l] = (T)((double)SV[l] - omega * (double)TV[l]);
                }
                stv<T, VN>(r + c0, RV);
#pragma unroll
                for (int l = 0; l < VN; ++l)
                    if (g.dot_all || is_interior(g, c0 + l)) {
                        acc[0] += (double)(T)RV[l] * (double)(T)RV[l];
                        acc[1] += (double)(T)RT[l] * (double)(T)RV[l];
                    }
            }
            if (mixed) {
                stv<double, VN>(y + c0, XN);
            } else if constexpr (sizeof(T) == sizeof(double)) {
                stv<double, VN>(x + c0, XN);
            }
        } else {
            for (int64_t c = c0; c < P.n; ++c) cell(c);
        }
    } else {
        FOR_CELLS(TB, P, blk, c) cell(c);
    }
    block_partial<TB, 2>(acc, to_global(S.partials), P.nblk, blk, lds);
}

// Reliable update (mixed path): xo = xi + y, r64 = b - A64*(xi + y) with the
// neighbours' xi + y formed on the fly (xi is read-only here, xo is the other
// buffer), r = (float)r64; partial ||r64||^2, <Rtilde, r> (rho of the
// continued recurrence) and <r, r> (rho of a restarted one, whose Rtilde is
// r).  After a breakdown (discard_y) y is dropped: xo = xi.
//
// The kernel is launched every iteration but works only every ~100, so each
// system gets at most kUpdateBlocks blocks (map entry y = 0..ub-1) that loop
// over the logical blocks of their XCD's slab: an idle launch retires in a
// few microseconds.  Each block keeps one partial per sum over all its
// logical blocks (a fixed order for a given system size) and counts itself
// in with a release increment; the last block to arrive sums the ub
// partials in block order and applies the update's scalar logic (FIN_UPDATE)
// itself, so no finish launch follows the (mostly idle) update.
template <int BS>
__global__ void __launch_bounds__(BS) k_reliable_update(const Sys<float> *__restrict__ sys,
                                                        const int2 *__restrict__ map) {
    __shared__ double lds[(BS / 64) * 3];
    const int2 e_ = map[blockIdx.x];
    if (e_.x < 0) return;
    const Sys<float> &S = sys[e_.x];
    const DGrid &g = S.g;
    const Part &P = S.P;
    auto *__restrict__ st = to_global(S.st);
    if (st->done || !st->need_update) return;
    const int cur = st->x_cur;
    const auto xi = cur ? to_global(S.xb) : to_global(S.xa);
    const auto xo = cur ? to_global(S.xa) : to_global(S.xb);
    auto *__restrict__ a = to_global(S.a64);
    auto *__restrict__ b = to_global(S.b64);
    auto *__restrict__ r = to_global(S.r);
    auto *__restrict__ rt = to_global(S.rt);
    struct XY {
        gptr<const double> xi;
        gptr<const double> y;
        __device__ __forceinline__ double operator()(int64_t q) const { return y ? xi[q] + y[q] : xi[q]; }
    } X{xi, st->discard_y ? nullptr : to_global(S.y)};
    const int64_t ub = update_blocks(P.nblk), nper = ub / kNumXcd;
    const int64_t per = (P.nblk + kNumXcd - 1) / kNumXcd;
    const int64_t lo = (e_.y % kNumXcd) * per, hi = lo + per < P.nblk ? lo + per : P.nblk;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t blk = lo + e_.y / kNumXcd; blk < hi; blk += nper) {
        FOR_CELLS(BS, P, blk, c) {
            const Nb n = neighbours_fast(g, c);
            const float rtc = rt[c];
            const double bc = b[c];
            const double rr = bc - stencil_row<double>(g, a, X, c, n);
            xo[c] = X(c);
            const float r32 = (float)rr;
            r[c] = r32;
            if (g.dot_all || n.interior) {
                acc[0] += rr * rr;
                acc[1] += (double)rtc * (double)r32;
                acc[2] += (double)r32 * (double)r32;
            }
        }
    }
    auto *__restrict__ part = to_global(S.partials);  // 3 x ub slots (partial_slots)
    __shared__ bool last;
    block_sum<BS, 3>(acc, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < 3; ++q) part[q * ub + e_.y] = acc[q];
        const unsigned prev = __hip_atomic_fetch_add(&st->upd_counter, 1u, __ATOMIC_RELEASE,
                                                     __HIP_MEMORY_SCOPE_AGENT);
        last = prev == (unsigned)(ub - 1);
    }
    __syncthreads();
    if (!last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    double tot[3] = {0.0, 0.0, 0.0};
    for (int64_t u = threadIdx.x; u < ub; u += BS) {
#pragma unroll
        for (int q = 0; q < 3; ++q) tot[q] += load_agent(&part[q * ub + u]);
    }
    __syncthreads();  // lds is reused
    block_sum<BS, 3>(tot, lds);
    if (threadIdx.x == 0) {
        st->upd_counter = 0u;
        fin_apply<FIN_UPDATE, 3>(st, tot, nullptr);
    }
}

// FP32 difference-layout copy of the (scaled) FP64 matrix: off-diagonals
// rounded, plane 3 = A0 + sum(off-diagonals) formed in FP64, plane 7 = 1/A0
// (DIAG) or A0 (LINE, NONE).
// A coupling clamped to the cell itself contributes a_l*(x_c - x_c) = 0 in
// the difference form and a_l*x_c through the excess, as in LEQ_MATVEC.
template <int BS>
__global__ void __launch_bounds__(BS) k_prepare_a32(const Sys<float> *__restrict__ sys, const int2 *__restrict__ map) {
    SYS_BLOCK(float)
    if (to_global(S.st)->done) return;
    auto *__restrict__ a = to_global(S.a64);
    auto *__restrict__ a32 = to_global(const_cast<float *>(S.a));
    const int64_t n3 = g.n3;
    int nonunit = 0;
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
        // DIAG systems keep 1/A0 (the inner DIAG preconditioner multiplies),
        // LINE systems A0 (the line solves' diagonal)
        const float d = S.pc == PC_DIAG ? (float)(1.0 / a[3 * n3 + c]) : (float)a[3 * n3 + c];
        a32[7 * n3 + c] = d;
        nonunit |= d != -1.0f;
    }
    // a unit-scaled A0 within 2^-25 of -1 has 1/A0 = -1 in FP32, and the
    // DIAG kernels then negate instead of reading this plane
    if (__any(nonunit) && (threadIdx.x & 63) == 0 && S.pc == PC_DIAG)
        atomicOr((int *)&to_global(S.st)->nonunit_diag, 1);
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
    auto *__restrict__ st = to_global(S.st);
    if (fin_skips<F>(st)) return;
    double tot[NV];
    if (finish_sum<NV>(to_global(S.partials), S.P.nblk, e.y, S.nfin, st->l2, &st->fin_counter, lds, tot) && threadIdx.x == 0)
        fin_apply<F, NV>(st, tot, nullptr);
}

// Red-black line phase along J for every LINE system of a table (RSRS and
// the 2-D R sweep).  Block b takes lines [map[b].y*LPB, +LPB) of colour
// `param` of system map[b].x, numbered densely by color_line, with five
// waves: lane l < LPB of wave 0 runs DGTSV on line l (the operations and
// order of solve_line, so the result is bit-identical); waves 1-4 feed it.
// Lines along J are contiguous, so a loader wave reads rows in runs: a tile
// is RT = 1024/LPB rows of every line, loader w takes the lines of quarter
// w-1, and element k of its lane is row lane % RT of line (quarter) +
// lane / RT + (LPB/16) k, each load instruction fetching LPB/16 lines x RT
// contiguous rows (16 rows, 128 bytes of FP64, per line at 64 lines per
// block); it forms the right-hand side (B minus the I/K neighbour terms,
// whose offsets are constant along a J-line) and stages it in LDS, while
// wave 0 eliminates the previous tile: one memory round trip per tile (16
// rows at 64 lines per block, 32 at 32, 64 at 16).
// Finished rows go to scratch in [row][lane] order; for the back solve
// waves 1-4 stage them in LDS in groups of 8 rows, two groups ahead, and
// waves 1-2 write the solution out through LDS, coalesced.  Fewer lines per block give more blocks (128^3: 8,450 lines per
// colour are 133 blocks of 64 but 265 of 32); the launch shape picks LPB
// per grid size.  Four loaders of four elements each keep the kernel at two
// blocks per CU (a loader with eight elements needs 256 registers, one
// block per CU).  which = 0 preconditions P into Phat, 1 S into Shat.
template <typename T, int LPB>
__global__ void __launch_bounds__(320) k_line_color_sys(const Sys<T> *__restrict__ sys, const int2 *__restrict__ map,
                                                        int param, int which) {
    static_assert(LPB == 64 || LPB == 32 || LPB == 16, "lines per block");
    constexpr int RL = 1024 / LPB; // rows per run a loader instruction reads of each line
    constexpr int LPI = 64 / RL;   // lines per load instruction
    constexpr int K = 4;           // tile elements per loader lane
    constexpr int RT = RL;         // rows per elimination tile
    constexpr int RB = 8;          // rows per back-solve group
    constexpr int KS = LPB / 16;   // solution elements per lane of each storing wave
    // LDS rows padded so that a loader's staging store (RL rows x LPI lines)
    // and a storing wave's read (8 rows x 8 lines) each touch every bank
    // pair exactly twice, the minimum for 64 lanes of 8 bytes (unpadded,
    // the rows share banks: 8- to 32-way conflicts)
    constexpr int SL = LPB + LPI, SX = LPB + 8;
    __shared__ T tl[2][4][RT][SL]; // dl, d, du, bb: RT rows x LPB lines, double buffered
    __shared__ int linfo[LPB];      // DGTSV INFO per line
    const int2 e = map[blockIdx.x];
    if (e.x < 0) return;
    const Sys<T> &S = sys[e.x];
    auto *__restrict__ st = to_global(S.st);
    if (st->done || (which == 1 && st->s_exit)) return;
    if (e.y == 0 && threadIdx.x == 0) st->line_launches += 1;  // (timing: working launches)
    const DGrid &g = S.g;
    const LineGeom L = line_geom(g, AX_J);
    const int n = L.n;
    const int64_t n3 = g.n3;
    const int wave = threadIdx.x >> 6;  // 0: elimination and back solve; 1-4: loads (1, 2: stores)
    const int lane = threadIdx.x & 63;
    const bool mine = lane < LPB;       // wave 0: this lane solves line `lane` of the block
    const int ro = 0;                                      // a loader's first row within a tile
    const int lh = wave >= 1 ? (LPB / 4) * (wave - 1) : 0;  // its first line
    const int do_k = g.do_k;
    constexpr int pd = DiagPlane<T>::value;
    const gptr<const T> a = to_global(S.a);
    const gptr<const T> pI_m = a + 2 * n3, pI_p = a + 4 * n3, pK_m = a, pK_p = a + 6 * n3;
    const gptr<const T> pD = a + pd * n3, pJ_m = a + n3, pJ_p = a + 5 * n3;
    const gptr<const T> bm = to_global((const T *)(which == 0 ? S.p : S.s));
    const gptr<T> var = to_global(which == 0 ? S.phat : S.shat);
    const gptr<T> sc = to_global(S.line_scratch + (int64_t)e.y * 4 * n * LPB);
    const uint32_t c1 = (uint32_t)g.c1, c2 = (uint32_t)g.c2;
    const uint32_t pl = (uint32_t)n * LPB;  // finished rows: 4 arrays of n rows x LPB lanes
    const uint32_t pb = pl * (uint32_t)sizeof(T);  // (bytes)
    const gptr<char> scb = reinterpret_cast<gptr<char>>(sc);
    int u = 0, v = 0;
    const int vld = mine && color_line(L.nu, L.nv, do_k ? 1 : 0, param, (int64_t)e.y * LPB + lane, u, v) ? 1 : 0;
    const int rsub = lane % RL;
    // loader view: element k is row rsub of line lh + lane / RL + LPI k, with
    // its start cell (~0u: no line) and the I/K neighbour offsets, constant
    // along a J-line (set_increments.f:72-112)
    uint32_t jc[K], jim[K], jip[K], jkm[K], jkp[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int src = lh + lane / RL + LPI * k;
        const int iu = __shfl(u, src, 64), kv = __shfl(v, src, 64);
        jc[k] = __shfl(vld, src, 64) ? (uint32_t)iu * c1 + (uint32_t)kv * c2 : ~0u;
        jim[k] = (uint32_t)(nb_minus(iu, g.ni, g.cyc_x) - iu) * c1;
        jip[k] = (uint32_t)(nb_plus(iu, g.ni, g.cyc_x) - iu) * c1;
        jkm[k] = do_k ? (uint32_t)(nb_minus(kv, g.nk, g.cyc_z) - kv) * c2 : 0u;
        jkp[k] = do_k ? (uint32_t)(nb_plus(kv, g.nk, g.cyc_z) - kv) * c2 : 0u;
    }
    // Load the loader's rows [Q0, Q0+RL) of its elements (all loads issued
    // before any is used) and stage them in LDS buffer BUF: the right-hand
    // side in LEQ_IKSWEEP's order (leqsol_mod.f:891-902).
    T tb[K], t1m[K], t1p[K], t2m[K], t2p[K], v1m[K], v1p[K], v2m[K], v2p[K], td[K], tdl[K], tdu[K];
    auto tile = [&](int q0, int buf) __attribute__((always_inline)) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int q = q0 + rsub;
            const bool ok = jc[k] != ~0u && q < n;
            const uint32_t c = ok ? jc[k] + (uint32_t)q : 0u;  // in bounds even if loaded
            tb[k] = ok ? bm[c] : (T)0;
            t1m[k] = ok ? pI_m[c] : (T)0;
            t1p[k] = ok ? pI_p[c] : (T)0;
            v1m[k] = ok ? var[c + jim[k]] : (T)0;
            v1p[k] = ok ? var[c + jip[k]] : (T)0;
            t2m[k] = ok && do_k ? pK_m[c] : (T)0;
            t2p[k] = ok && do_k ? pK_p[c] : (T)0;
            v2m[k] = ok && do_k ? var[c + jkm[k]] : (T)0;
            v2p[k] = ok && do_k ? var[c + jkp[k]] : (T)0;
            td[k] = ok ? pD[c] : (T)0;
            tdl[k] = ok && q > 0 ? pJ_m[c] : (T)0;
            tdu[k] = ok && q < n - 1 ? pJ_p[c] : (T)0;
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            T bb;
            if (do_k) bb = tb[k] - t1m[k] * v1m[k] - t1p[k] * v1p[k] - t2m[k] * v2m[k] - t2p[k] * v2p[k];
            else bb = tb[k] - t1m[k] * v1m[k] - t1p[k] * v1p[k];
            const int j = lh + lane / RL + LPI * k;
            tl[buf][0][ro + rsub][j] = tdl[k];
            tl[buf][1][ro + rsub][j] = td[k];
            tl[buf][2][ro + rsub][j] = tdu[k];
            tl[buf][3][ro + rsub][j] = bb;
        }
    };
    const int ntiles = (n + RT - 1) / RT;
    if (wave >= 1) tile(ro, 0);
    __syncthreads();
    T d0 = (T)0, du0 = (T)0, b0 = (T)0;
    int info = 0;
    for (int tt = 0; tt < ntiles; ++tt) {
        const int buf = tt & 1;
        if (wave >= 1) {
            if (tt + 1 < ntiles) tile((tt + 1) * RT + ro, buf ^ 1);
        } else if (mine) {
            // DGTSV elimination of this lane's line over the tile
            // (DGTSV.f:107-163, the operations of solve_line in the same order)
            const int q0 = tt * RT;
            const int re = n - q0 < RT ? n - q0 : RT;  // rows of this tile
            int r = 0;
            if (tt == 0) {  // row 0 only starts the recurrence
                d0 = tl[0][1][0][lane];
                du0 = tl[0][2][0][lane];
                b0 = tl[0][3][0][lane];
                r = 1;
            }
            // finished row i = q0 + r - 1 of the scratch arrays, as a byte
            // offset from the block's base (32-bit offsets, one add per row)
            uint32_t so = ((uint32_t)(q0 + r - 1) * LPB + (uint32_t)lane) * (uint32_t)sizeof(T);
            // eliminate with row i+1 = (dl1, d1, du1, b1), store finished row i
            auto row = [&](const T dl1, const T d1, const T du1, const T b1, const int i)
                           __attribute__((always_inline)) {
                const bool keep = fabs((double)d0) >= fabs((double)dl1);  // no interchange
                auto put = [&](const T fd, const T fu, const T fu2, const T fb) __attribute__((always_inline)) {
                    *reinterpret_cast<gptr<T>>(scb + so) = fd;
                    *reinterpret_cast<gptr<T>>(scb + (so + pb)) = fu;
                    *reinterpret_cast<gptr<T>>(scb + (so + 2 * pb)) = fu2;
                    *reinterpret_cast<gptr<T>>(scb + (so + 3 * pb)) = fb;
                };
                if (__all(keep)) {
                    // no lane interchanges rows (every diagonally dominant line)
                    if (d0 == (T)0 && !info) info = i + 1;
                    const T fact = dl1 / d0;
                    put(d0, du0, (T)0, b0);
                    d0 = d1 - fact * du0;
                    du0 = du1;
                    b0 = b1 - fact * b0;
                } else {
                    // per lane, row i stays or trades places with row i+1: one
                    // division and the same products, on selected operands
                    if (keep && d0 == (T)0 && !info) info = i + 1;
                    const T fact = (keep ? dl1 : d0) / (keep ? d0 : dl1);
                    const T pm = keep ? du0 : d1;  // row i's DU after the interchange
                    const T bm = keep ? b0 : b1;   // row i's B after the interchange
                    const bool sh = !keep && i < n - 2;
                    put(keep ? d0 : dl1, pm, sh ? du1 : (T)0, bm);
                    d0 = (keep ? d1 : du0) - fact * pm;
                    du0 = sh ? -fact * du1 : du1;
                    b0 = (keep ? b1 : b0) - fact * bm;
                }
                so += LPB * (uint32_t)sizeof(T);
            };
            // two rows per step in two register sets, each read from LDS one
            // row ahead of its elimination (the LDS latency stays off the
            // serial chain of divisions, and no values move between sets)
            T adl = tl[buf][0][r][lane], ad = tl[buf][1][r][lane], adu = tl[buf][2][r][lane];
            T ab = tl[buf][3][r][lane];
            for (; r + 1 < re; r += 2) {
                const T bdl = tl[buf][0][r + 1][lane], bd = tl[buf][1][r + 1][lane];
                const T bdu = tl[buf][2][r + 1][lane], bb = tl[buf][3][r + 1][lane];
                row(adl, ad, adu, ab, q0 + r - 1);
                const int r2 = r + 2 < RT ? r + 2 : RT - 1;  // in the tile; unused past its end
                adl = tl[buf][0][r2][lane];
                ad = tl[buf][1][r2][lane];
                adu = tl[buf][2][r2][lane];
                ab = tl[buf][3][r2][lane];
                row(bdl, bd, bdu, bb, q0 + r);
            }
            if (r < re) row(adl, ad, adu, ab, q0 + r - 1);
        }
        __syncthreads();
    }
    if (wave == 0 && mine) {
        if (!info && d0 == (T)0) info = n;  // DGTSV.f:160-163
        linfo[lane] = info;
    }
    __syncthreads();
    // storing view (waves 1 and 2): element k is row lane & 7 of line
    // (lane >> 3) + 8 (KS (wave - 1) + k); a line whose elimination hit a zero
    // pivot keeps its values (leqsol_mod.f:910-914)
    uint32_t sjc[KS];
#pragma unroll
    for (int k = 0; k < KS; ++k) {
        const int src = (lane >> 3) + 8 * (KS * (wave == 2 ? 1 : 0) + k);
        const int su = __shfl(u, src, 64), sv = __shfl(v, src, 64);
        sjc[k] = __shfl(vld, src, 64) && !linfo[src] ? (uint32_t)su * c1 + (uint32_t)sv * c2 : ~0u;
    }
    // back solve (DGTSV.f:233-246) in groups of RB rows from the top row
    // down; rows n-2 .. 0 come from scratch (row n-1 is still in registers).
    // Waves 1-4 stage the scratch rows of group t+1 in LDS (the tiles'
    // buffers, free now) while wave 0 solves group t, with the loads of the
    // next DB groups in flight in their registers: the back solve waits on
    // memory once, not once per group.
    constexpr int GE = 4 * RB * LPB;  // scratch values of one group
    constexpr int FE = GE / 256;      // per lane of the four staging waves
    constexpr int DB = 2;             // groups in flight per staging lane
    static_assert(2 * GE <= 2 * 4 * RT * SL, "two staged groups fit in the tile buffers");
    T(*bs)[4][RB][LPB] = reinterpret_cast<T(*)[4][RB][LPB]>(&tl[0][0][0][0]);  // slot t & 1
    // solution rows of group t in slot t & 1, row hi - r in xs[t & 1][r]
    T(*xs)[RB][SX] = reinterpret_cast<T(*)[RB][SX]>(&tl[0][0][0][0] + 2 * GE);
    static_assert(2 * GE + 2 * RB * SX <= 2 * 4 * RT * SL, "staged groups and solution rows fit in the tile buffers");
    const int ngrp = (n + RB - 1) / RB;
    const bool stager = wave >= 1;
    const int fl = (wave - 1) * 64 + lane;  // 0..255 on the staging waves
    T fr[DB][FE];
    auto issue = [&](int t, T(&dst)[FE]) __attribute__((always_inline)) {
        const int hi = n - 1 - t * RB, lo = hi - RB + 1 > 0 ? hi - RB + 1 : 0;
#pragma unroll
        for (int k = 0; k < FE; ++k) {
            const int e = fl + 256 * k, vq = e / (RB * LPB), r = (e / LPB) % RB, ln = e % LPB;
            const int q = hi - r;
            const bool ok = t < ngrp && q >= lo && q <= n - 2;
            dst[k] = ok ? sc[(uint32_t)vq * pl + (uint32_t)q * LPB + ln] : (T)0;
        }
    };
    auto stage = [&](int t, const T(&src)[FE]) __attribute__((always_inline)) {
#pragma unroll
        for (int k = 0; k < FE; ++k) {
            const int e = fl + 256 * k, vq = e / (RB * LPB), r = (e / LPB) % RB, ln = e % LPB;
            bs[t & 1][vq][r][ln] = src[k];
        }
    };
    T x1 = mine ? b0 / d0 : (T)0, x2 = (T)0;
    // group t: solve it from LDS (wave 0) into solution slot t & 1, stage
    // group t+1 and put group t+1+DB in flight in the register set group t+1
    // held (waves 1-4; set (t+1) % DB, so the loop is unrolled by DB), then
    // store its solution rows (waves 1-2).  One barrier per group: the
    // staging of group t+2 into bs[t & 1] and the solving of group t+2 into
    // xs[t & 1] both come after the next group's barrier, which every
    // reader of group t's slots has passed.
    auto group = [&](int t, T(&set)[FE], bool first) __attribute__((always_inline)) {
        const int hi = n - 1 - t * RB, lo = hi - RB + 1 > 0 ? hi - RB + 1 : 0;
        T(&xo)[RB][SX] = xs[t & 1];
        if (wave == 0 && mine) {
            // row hi - r; rows below lo (last group only) are zeros in LDS and
            // give values that are neither stored nor used
#pragma unroll
            for (int r = 0; r < RB; ++r) {
                const T bq = bs[t & 1][3][r][lane], uq = bs[t & 1][1][r][lane];
                const T u2q = bs[t & 1][2][r][lane], dq = bs[t & 1][0][r][lane];
                if (first && r == 0) {
                    xo[0][lane] = x1;  // row n-1: B/D of the elimination
                } else {
                    // row n-2 has U2 = 0 and x2 = 0 here, so this is its
                    // (B - DU x1) / D bit for bit (DGTSV.f:233-246)
                    const T x = (bq - uq * x1 - u2q * x2) / dq;
                    x2 = x1;
                    x1 = x;
                    xo[r][lane] = x;
                }
            }
        } else if (stager && t + 1 < ngrp) {
            stage(t + 1, set);
            issue(t + 1 + DB, set);
        }
        __syncthreads();
        if (wave == 1 || wave == 2) {
#pragma unroll
            for (int k = 0; k < KS; ++k) {
                const int q = hi - (lane & 7);
                const int j = (lane >> 3) + 8 * (KS * (wave - 1) + k);
                if (sjc[k] != ~0u && q >= lo) var[sjc[k] + (uint32_t)q] = xo[lane & 7][j];
            }
        }
    };
    if (stager) {
        issue(0, fr[0]);
        issue(1, fr[1]);
        stage(0, fr[0]);
        issue(DB, fr[0]);
    }
    __syncthreads();
    for (int t0 = 0; t0 < ngrp; t0 += DB) {
        group(t0, fr[1], t0 == 0);
        if (t0 + 1 < ngrp) group(t0 + 1, fr[0], false);
    }
}

constexpr int kLineBlock = 320;  // threads of k_line_color_sys: one solving wave, four loaders

// Lines-per-block variants of k_line_color_sys (a launch shape picks one per
// grid size).
constexpr int kLineLpbVariants = 3;
__host__ __device__ constexpr int line_lpb_of(int var) { return var == 0 ? 64 : var == 1 ? 32 : 16; }
__host__ __device__ inline int line_lpb_variant(int lpb) { return lpb >= 64 ? 0 : lpb >= 32 ? 1 : 2; }

// Red-black line phase along J for the FP32 inner solve of the mixed path:
// several lanes per line.  The FP64 kernel above must repeat DGTSV's
// elimination order to reproduce the reference bit for bit, one serial chain
// per line; the FP32 recurrence is corrected by the FP64 reliable updates, so
// here each line of n rows is split over P lanes of m = ceil(n/P) <= M
// consecutive rows (partition method, H. H. Wang, ACM TOMS 7 (1981) 170-183):
//   1. each lane eliminates its chunk's rows but the last (the separator),
//      writing every one as x_j = U_j - V_j G_l - W_j G_{l-1} in terms of
//      its own separator G_l and the previous lane's G_{l-1};
//   2. substituting into the separator rows leaves a tridiagonal system in
//      the P separators, solved by parallel cyclic reduction across the lanes
//      (log2 P steps of lane shuffles; Hockney 1965);
//   3. each lane recovers its rows from G_l and G_{l-1}.
// No pivoting: MFIX's rows are diagonally dominant, and a line whose result
// is not finite keeps its values, as a failed DGTSV leaves it
// (leqsol_mod.f:910-914).  Rows are loaded and stored in interleaved order
// (lane gl takes rows gl, gl+P, ...: one load instruction reads P
// consecutive rows of a line) and the right-hand side is formed there; the
// four coefficients of the line's tridiagonal system go through LDS to the
// chunk order the elimination needs, and the solution back the same way.
// In LDS row q of chunk l = q / m, offset q - l*m, sits at l*s + offset with
// a chunk stride s = m | 1 (odd, so the lanes' chunk reads fall in distinct
// banks).  Rows past n are identity rows (x = 0, no coupling: the last real
// row has no upper coupling).  Block b takes lines
// [map[b].y*LPB, +LPB) of colour `param` (numbered densely by color_line), four
// waves of 64/P lines.  which = 0 preconditions P into Phat, 1 S into Shat.
template <int P, int M>
__host__ __device__ constexpr int pcr_lds_floats() {
    return 4 * (64 / P) * (4 * P * (M + 1) + P);  // lines per block x LS (below)
}

template <int P, int M>
__device__ __forceinline__ void line_pcr_block(const Sys<float> *__restrict__ sys, const int2 e, int param,
                                               int which, float *__restrict__ lds) {
    constexpr int LPW = 64 / P;
    constexpr int LPB = 4 * LPW;
    if (e.x < 0) return;
    const Sys<float> &S = sys[e.x];
    auto *__restrict__ st = to_global(S.st);
    if (st->done || (which == 1 && st->s_exit)) return;
    if (e.y == 0 && threadIdx.x == 0) st->line_launches += 1;  // (timing: working launches)
    const DGrid &g = S.g;
    const int n = g.nj;
    const int m = (n + P - 1) / P;  // rows per lane, <= M (the host picks the variant)
    const int64_t n3 = g.n3;
    const int lane = threadIdx.x & 63;
    const int grp = lane / P, gl = lane - grp * P, base = grp * P;
    const int64_t t = (int64_t)e.y * LPB + (threadIdx.x >> 6) * LPW + grp;
    int u = 0, v = 0;
    const bool ok = color_line(g.ni, g.nk, g.do_k ? 1 : 0, param, t, u, v);
    if (!ok) u = v = 0;  // a group past the colour's last line addresses line (0, 0) and stores nothing
    const uint32_t c1 = (uint32_t)g.c1, c2 = (uint32_t)g.c2;
    const uint32_t c0 = (uint32_t)u * c1 + (uint32_t)v * c2;
    const uint32_t oim = (uint32_t)(nb_minus(u, g.ni, g.cyc_x) - u) * c1;
    const uint32_t oip = (uint32_t)(nb_plus(u, g.ni, g.cyc_x) - u) * c1;
    const uint32_t okm = g.do_k ? (uint32_t)(nb_minus(v, g.nk, g.cyc_z) - v) * c2 : 0u;
    const uint32_t okp = g.do_k ? (uint32_t)(nb_plus(v, g.nk, g.cyc_z) - v) * c2 : 0u;
    const gptr<const float> a = to_global(S.a);
    const gptr<const float> pK_m = a, pJ_m = a + n3, pI_m = a + 2 * n3, pI_p = a + 4 * n3, pJ_p = a + 5 * n3;
    const gptr<const float> pK_p = a + 6 * n3, pD = a + (int64_t)DiagPlane<float>::value * n3;
    const gptr<const float> bm = to_global((const float *)(which == 0 ? S.p : S.s));
    const gptr<float> var = to_global(which == 0 ? S.phat : S.shat);
    const int do_k = g.do_k;
    // LDS: per line, the four coefficient arrays of NS slots; lines LS apart
    // (LS = P mod 64: the groups of a wave start in different banks)
    constexpr int NS = P * (M + 1), LS = 4 * NS + P;
    static_assert(LPB * LS == pcr_lds_floats<P, M>(), "LDS size");
    float *__restrict__ ln = lds + ((threadIdx.x >> 6) * LPW + grp) * LS;
    const int sc = m | 1;                 // chunk stride in LDS
    const float rm = 1.f / (float)m;      // (q + 0.5) * rm truncates to q / m for q < 4096, m <= 8
    auto slot = [&](int q) {
        const int l = (int)(((float)q + 0.5f) * rm);
        return l * sc + (q - l * m);
    };
    // rows q = gl + P*j: loads issued together, then the right-hand side in
    // LEQ_IKSWEEP's order (leqsol_mod.f:891-902), and the line's tridiagonal
    // coefficients to LDS
    {
        float tb[M], t1m[M], t1p[M], v1m[M], v1p[M], t2m[M], t2p[M], v2m[M], v2p[M], ta[M], tbd[M], tc[M];
#pragma unroll
        for (int j = 0; j < M; ++j) {
            const int q = gl + P * j;
            const bool in = ok && q < n;
            const uint32_t c = in ? c0 + (uint32_t)q : c0;  // in bounds even when not used
            tb[j] = in ? bm[c] : 0.f;
            t1m[j] = in ? pI_m[c] : 0.f;
            t1p[j] = in ? pI_p[c] : 0.f;
            v1m[j] = in ? var[c + oim] : 0.f;
            v1p[j] = in ? var[c + oip] : 0.f;
            t2m[j] = in && do_k ? pK_m[c] : 0.f;
            t2p[j] = in && do_k ? pK_p[c] : 0.f;
            v2m[j] = in && do_k ? var[c + okm] : 0.f;
            v2p[j] = in && do_k ? var[c + okp] : 0.f;
            tbd[j] = in ? pD[c] : 1.f;
            ta[j] = in && q > 0 ? pJ_m[c] : 0.f;
            tc[j] = in && q < n - 1 ? pJ_p[c] : 0.f;
        }
#pragma unroll
        for (int j = 0; j < M; ++j) {
            const int q = gl + P * j;
            if (ok && q < n) {
                const float d = do_k ? tb[j] - t1m[j] * v1m[j] - t1p[j] * v1p[j] - t2m[j] * v2m[j] - t2p[j] * v2p[j]
                                     : tb[j] - t1m[j] * v1m[j] - t1p[j] * v1p[j];
                const int o = slot(q);
                ln[o] = ta[j];
                ln[NS + o] = tbd[j];
                ln[2 * NS + o] = tc[j];
                ln[3 * NS + o] = d;
            }
        }
    }
    __syncthreads();
    // this lane's chunk: rows q = gl*m + j
    float ra[M], rb[M], rc[M], rd[M];
#pragma unroll
    for (int j = 0; j < M; ++j) {
        const int q = gl * m + j;
        const bool in = ok && j < m && q < n;
        const int o = gl * sc + j;
        ra[j] = in ? ln[o] : 0.f;
        rb[j] = in ? ln[NS + o] : 1.f;
        rc[j] = in ? ln[2 * NS + o] : 0.f;
        rd[j] = in ? ln[3 * NS + o] : 0.f;
    }
    // 1. rows 0 .. m-2 of the chunk: forward elimination carrying the
    // coupling to the previous lane's separator, then backward substitution.
    // Reciprocals here and in the reduction are the hardware's (1 ulp, one
    // instruction) instead of correctly rounded divisions (about ten): this
    // FP32 solve is a preconditioner whose result the FP64 updates correct.
    float cp[M], wp[M], dp[M], U[M], V[M], W[M];
#pragma unroll
    for (int j = 0; j < M - 1; ++j) {
        if (j < m - 1) {
            if (j == 0) {
                const float r = __builtin_amdgcn_rcpf(rb[0]);
                cp[0] = rc[0] * r;
                wp[0] = ra[0] * r;
                dp[0] = rd[0] * r;
            } else {
                const float r = __builtin_amdgcn_rcpf(rb[j] - ra[j] * cp[j - 1]);
                cp[j] = rc[j] * r;
                wp[j] = -ra[j] * wp[j - 1] * r;
                dp[j] = (rd[j] - ra[j] * dp[j - 1]) * r;
            }
        }
    }
    float Ul = 0.f, Vl = 0.f, Wl = 0.f;  // row m-2 (the one next to the separator)
#pragma unroll
    for (int j = M - 2; j >= 0; --j) {
        if (j < m - 1) {
            if (j == m - 2) {
                U[j] = dp[j];
                V[j] = cp[j];
                W[j] = wp[j];
                Ul = U[j];
                Vl = V[j];
                Wl = W[j];
            } else {
                U[j] = dp[j] - cp[j] * U[j + 1];
                V[j] = -cp[j] * V[j + 1];
                W[j] = wp[j] - cp[j] * W[j + 1];
            }
        }
    }
    // the separator row (row m-1 of the chunk)
    float as = 0.f, bs = 1.f, cs = 0.f, ds = 0.f;
#pragma unroll
    for (int j = 0; j < M; ++j)
        if (j == m - 1) {
            as = ra[j];
            bs = rb[j];
            cs = rc[j];
            ds = rd[j];
        }
    // 2. reduced system A G_{l-1} + B G_l + C G_{l+1} = D over the P lanes
    const int nxt = base + (gl + 1 < P ? gl + 1 : gl);
    float U0 = m > 1 ? U[0] : 0.f, V0 = m > 1 ? V[0] : 0.f, W0 = m > 1 ? W[0] : 0.f;
    float Un = __shfl(U0, nxt, 64), Vn = __shfl(V0, nxt, 64), Wn = __shfl(W0, nxt, 64);
    if (gl + 1 >= P) Un = Vn = Wn = 0.f;
    float A, B, Cc, D;
    if (m > 1) {
        A = -as * Wl;
        B = bs - as * Vl - cs * Wn;
        Cc = -cs * Vn;
        D = ds - as * Ul - cs * Un;
    } else {
        A = as;
        B = bs;
        Cc = cs;
        D = ds;
    }
#pragma unroll
    for (int sft = 1; sft < P; sft <<= 1) {
        const bool hl = gl >= sft, hr = gl + sft < P;
        const int sl = base + (hl ? gl - sft : gl), sr = base + (hr ? gl + sft : gl);
        const float Al = __shfl(A, sl, 64), Bl = __shfl(B, sl, 64), Cl = __shfl(Cc, sl, 64), Dl = __shfl(D, sl, 64);
        const float Ar = __shfl(A, sr, 64), Br = __shfl(B, sr, 64), Cr = __shfl(Cc, sr, 64), Dr = __shfl(D, sr, 64);
        const float al = hl ? -A * __builtin_amdgcn_rcpf(Bl) : 0.f;
        const float ga = hr ? -Cc * __builtin_amdgcn_rcpf(Br) : 0.f;
        const float nA = hl ? al * Al : 0.f;
        const float nC = hr ? ga * Cr : 0.f;
        const float nB = B + (hl ? al * Cl : 0.f) + (hr ? ga * Ar : 0.f);
        const float nD = D + (hl ? al * Dl : 0.f) + (hr ? ga * Dr : 0.f);
        A = nA;
        B = nB;
        Cc = nC;
        D = nD;
    }
    const float G = D * __builtin_amdgcn_rcpf(B);
    float Gp = __shfl(G, base + (gl > 0 ? gl - 1 : 0), 64);
    if (gl == 0) Gp = 0.f;
    // 3. the chunk's rows, back to LDS (this lane's own slots of the
    // right-hand side array, read above by this lane only); a line with a
    // non-finite value keeps its values
    int bad = 0;
#pragma unroll
    for (int j = 0; j < M; ++j) {
        const float x = j < m - 1 ? U[j] - V[j] * G - W[j] * Gp : G;
        const int q = gl * m + j;
        if (ok && j < m && q < n) {
            bad |= !isfinite(x);
            ln[3 * NS + gl * sc + j] = x;
        }
    }
    const uint64_t bm_bad = __ballot(bad);
    const uint64_t gmask = P == 64 ? ~0ull : (((1ull << P) - 1ull) << base);
    const bool store = ok && !(bm_bad & gmask);
    __syncthreads();
    if (!store) return;
#pragma unroll
    for (int j = 0; j < M; ++j) {
        const int q = gl + P * j;
        if (q < n) var[c0 + (uint32_t)q] = ln[3 * NS + slot(q)];
    }
}

template <int P, int M>
__global__ void __launch_bounds__(256) k_line_pcr_sys(const Sys<float> *__restrict__ sys,
                                                      const int2 *__restrict__ map, int param, int which) {
    __shared__ float lds[pcr_lds_floats<P, M>()];
    line_pcr_block<P, M>(sys, map[blockIdx.x], param, which, lds);
}

// The four short-line variants (lines of up to 160 rows) of one colour in
// one launch: a table of systems with lines of several lengths (a batch)
// then issues one launch per colour phase instead of one per variant, and
// the variants' blocks run side by side.  Blocks [0, n[3]) take map m[3]
// (P 32, M 5), the next n[2] map m[2], and so on: the longest first.
struct PcrMaps {
    const int2 *m[4];
    int n[4];
};
constexpr int kPcrSmallLds = pcr_lds_floats<32, 5>();
static_assert(kPcrSmallLds >= pcr_lds_floats<32, 4>() && kPcrSmallLds >= pcr_lds_floats<16, 4>() &&
                  kPcrSmallLds >= pcr_lds_floats<16, 2>(),
              "LDS of the short-line variants");
static __global__ void __launch_bounds__(256) k_line_pcr_short(const Sys<float> *__restrict__ sys,
                                                               const PcrMaps pm, int param, int which) {
    __shared__ float lds[kPcrSmallLds];
    int b = blockIdx.x;
    if (b < pm.n[3]) return line_pcr_block<32, 5>(sys, pm.m[3][b], param, which, lds);
    b -= pm.n[3];
    if (b < pm.n[2]) return line_pcr_block<32, 4>(sys, pm.m[2][b], param, which, lds);
    b -= pm.n[2];
    if (b < pm.n[1]) return line_pcr_block<16, 4>(sys, pm.m[1][b], param, which, lds);
    b -= pm.n[1];
    if (b < pm.n[0]) line_pcr_block<16, 2>(sys, pm.m[0][b], param, which, lds);
}

// Variant of k_line_pcr_sys for a line of n rows: lanes per line P and most
// rows per lane M; -1 when n is too long for the registers of one lane.
__host__ __device__ inline int line_pcr_variant(int n) {
    if (n <= 32) return 0;   // P 16, M 2
    if (n <= 64) return 1;   // P 16, M 4
    if (n <= 128) return 2;  // P 32, M 4
    if (n <= 160) return 3;  // P 32, M 5
    if (n <= 256) return 4;  // P 64, M 4
    if (n <= 320) return 5;  // P 64, M 5
    if (n <= 512) return 6;  // P 64, M 8
    return -1;
}
constexpr int kPcrVariants = 7;
__host__ __device__ inline int line_pcr_lanes(int var) { return var <= 1 ? 16 : var <= 3 ? 32 : 64; }
// lines per block of a variant (four waves of 64/P lines)
__host__ __device__ inline int line_pcr_lpb(int var) { return 4 * (64 / line_pcr_lanes(var)); }

// Launch the k_line_pcr_sys variant `var` (line_pcr_variant) over `grid` blocks.
inline void launch_line_pcr(int var, dim3 grid, hipStream_t s, const Sys<float> *tab, const int2 *map,
                            int param, int which, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    const dim3 bs(256);
    switch (var) {
    case 0: launch_timed(e0, e1, k_line_pcr_sys<16, 2>, grid, bs, s, tab, map, param, which); break;
    case 1: launch_timed(e0, e1, k_line_pcr_sys<16, 4>, grid, bs, s, tab, map, param, which); break;
    case 2: launch_timed(e0, e1, k_line_pcr_sys<32, 4>, grid, bs, s, tab, map, param, which); break;
    case 3: launch_timed(e0, e1, k_line_pcr_sys<32, 5>, grid, bs, s, tab, map, param, which); break;
    case 4: launch_timed(e0, e1, k_line_pcr_sys<64, 4>, grid, bs, s, tab, map, param, which); break;
    case 5: launch_timed(e0, e1, k_line_pcr_sys<64, 5>, grid, bs, s, tab, map, param, which); break;
    default: launch_timed(e0, e1, k_line_pcr_sys<64, 8>, grid, bs, s, tab, map, param, which); break;
    }
}

// Number of the listed systems still running, and a running flag per
// listed system.
template <typename T>
__global__ void k_count_running(const Sys<T> *__restrict__ sys, const int *__restrict__ ids, int nids, int *out,
                                int *flags) {
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    for (int q = threadIdx.x; q < nids; q += blockDim.x) {
        const int run = !sys[ids[q]].st->done;
        flags[q] = run;
        if (run) atomicAdd(&cnt, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) *out = cnt;
}

// ---- LEQ_MATVEC / LEQ_MSOLVE1 / dot / copy as standalone kernels ----------

template <int BS, typename T>
__global__ void __launch_bounds__(BS) k_matvec(DGrid g, Part P, const T *__restrict__ a, const T *__restrict__ x,
                                               T *__restrict__ y) {
    if (!xcd_valid(P.nblk)) return;
    const int64_t blk = xcd_block(P.nblk);
    const Operand<T, OP_PLAIN> X{x, nullptr};
    FOR_CELLS(BS, P, blk, c) {
        const Nb n = neighbours_fast(g, c);
        y[c] = Stencil<T>::row(g, a, X, c, n);  // FP32: the inner solve's difference form
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
