// Kernels of the BiCGSTAB path (LEQ_BICGS0, model/leq_bicgs.f:107-509).
//
// One solve keeps its Krylov scalars and control flags in a device-resident
// SolverState.  Each reduction phase is a streaming kernel that writes one
// partial sum per block, followed by a one-block finish kernel that sums the
// partials in a fixed order and applies the reference's scalar logic (alpha,
// the S-norm exit, omega, the convergence test, IER).  No iteration needs a
// host round trip, a chunk of iterations can be captured in a HIP graph, and
// every kernel returns at once when the solve is done, which makes extra
// launches after convergence harmless.
//
// Every streaming kernel runs a fixed number of blocks; logical block b owns
// the contiguous cells [b*chunk, (b+1)*chunk) and XCD x owns a contiguous
// range of logical blocks, so the I/K neighbours a block reads were loaded
// into the same XCD's L2 by its neighbouring blocks.
//
// Vector/matrix storage type T is double (plain FP64 solve) or float (inner
// solve of the mixed-precision path); element-wise updates are evaluated in
// FP64 registers and rounded once to T, reductions accumulate in FP64.
// FP32 matrices use the difference layout (8 planes: six off-diagonals, the
// row excess A0+sum(off) in plane 3, A0 in plane 7), see stencil_row_diff.
#pragma once
#include "leq_device.h"

namespace mfix {

enum Pc : int { PC_NONE = 0, PC_DIAG = 1, PC_LINE = 2 };

struct SolverState {
    double rnorm0, rnorm, snorm, rho, rho_prev, alpha, omega, beta;
    double tol, rtv, txs, txt;
    double true_rnorm;      // mixed: last FP64 ||b - A x||
    double upd_rnorm;       // mixed: ||r|| at the last reliable update
    double delta;           // mixed: reliable-update trigger factor
    int max_updates;        // mixed: reliable updates allowed (0 = no limit)
    int i;                  // current loop index (1-based), leq_bicgs.f:316
    int iter_var;           // the reference's `iter` counter
    int itmax;
    int done, ier, exit_reason, s_exit, iter_tot_inc, isconverged;
    int minimize, icheck;
    int zero_row;
    int need_update;        // mixed: reliable update requested by the last iteration
    int msolve_fatal;       // DGTSV INFO>0 in a JK/IJ line sweep (fatal in the reference)
    int updates;            // mixed: reliable updates performed
    int mixed;              // 1 = inner solve of the mixed-precision path
    int final_update;       // mixed: the pending update ends the solve
    int y_reset;            // mixed: the FP64 correction accumulator restarts from 0
    int discard_y;          // mixed: a non-finite scalar appeared; the next update drops y
    int restart_always;     // mixed: restart the recurrence at every reliable update (large systems)
    int beta_invalid;       // mixed: this update cannot continue the recurrence (S-norm exit, breakdown)
    int x_cur;              // mixed: which FP64 solution buffer is current
    int restart;            // mixed: restart the recurrence (p = r, Rtilde = r) at the next iteration
    int restarts;           // mixed: recurrence restarts (one per reliable update that did not end the solve)
    unsigned fin_counter;   // arrivals of the current finish launch (reset by its last block)
    unsigned upd_counter;   // mixed: arrivals of the current reliable-update launch (reset by its last block)
    int nonunit_diag;       // some unit-scaled A0 differs from -1 (DIAG must read the A0 plane)
    int k1_launches;        // launches of the first stencil pass that worked on this system (timing)
    int line_launches;      // colour-phase launches that worked on this system (timing)
    double l2[4 * kFinishBlocks];  // level-2 sums of the finish launches
};

enum Exit : int { EXIT_CONVERGED = 0, EXIT_SNORM = 1, EXIT_RHO = 2, EXIT_ITMAX = 3, EXIT_ZERO_ROW = 4 };

constexpr double kTolMin = 2.220446049250313e-16;   // EPSILON(1d0), leq_bicgs.f:277
constexpr double kSmallNumber = 1.0e-15;            // SMALL_NUMBER, param1_mod.f:24
constexpr double kRatioTol = (double)0.2f;          // ratiotol = 0.2 (default-real literal), leq_bicgs.f:165

// Plane holding A0: 3 in the reference layout (FP64), 7 in the FP32 difference layout.
template <typename T> struct DiagPlane { static constexpr int value = 3; };
template <> struct DiagPlane<float> { static constexpr int value = 7; };

template <typename T> struct Stencil;
template <> struct Stencil<double> {
    template <typename X, typename PA>
    __device__ __forceinline__ static double row(const DGrid &g, PA a, const X &x, int64_t c, const Nb &n) {
        return stencil_row<double>(g, a, x, c, n);
    }
    template <typename... V>
    __device__ __forceinline__ static double combine(bool do_k, V... v) {
        return stencil_combine<double>(do_k, v...);
    }
};
template <> struct Stencil<float> {
    template <typename X, typename PA>
    __device__ __forceinline__ static float row(const DGrid &g, PA a, const X &x, int64_t c, const Nb &n) {
        return stencil_row_diff(g, a, x, c, n);
    }
    template <typename... V>
    __device__ __forceinline__ static float combine(bool do_k, V... v) {
        return stencil_combine_diff(do_k, v...);
    }
};

// Loop over the cells of logical block BLK (block size BS).
#define FOR_CELLS(BS, P, BLK, C)                                                    \
    for (int64_t C = (BLK) * (P).chunk + threadIdx.x,                               \
                 _end = ((BLK) + 1) * (P).chunk < (P).n ? ((BLK) + 1) * (P).chunk : (P).n; \
         C < _end; C += (BS))

// Final IER classification after ITMAX trips, leq_bicgs.f:486-495.
template <typename PS>
__device__ __forceinline__ void finish_itmax(PS st) {
    int conv = st->isconverged;
    const double rn = (double)(float)st->rnorm;
    if (!conv) conv = rn <= st->tol * st->rnorm0;
    st->ier = 0;
    if (!conv) {
        st->ier = -1;
        st->iter_tot_inc = st->iter_var;
        if (rn >= kRatioTol * (double)(float)st->rnorm0) st->ier = -2;
    }
    st->isconverged = conv;
    st->exit_reason = EXIT_ITMAX;
    st->done = 1;
}

// ---- scalar logic applied by the finish kernels ---------------------------

enum Fin : int {
    FIN_SETUP = 0,      // Rnorm0 (leq_bicgs.f:291), zero-row IER=-2 (:217-233)
    FIN_SHADOW,         // rho_0 (:318) and its zero test at i=1 (:320-332)
    FIN_ALPHA,          // alpha = rho/<Rtilde,V> (:357-361)
    FIN_OMEGA,          // ||S|| <= EPSILON exit (:372-374); omega = <T,S>/<T,T>, TxT=0 -> SMALL_NUMBER (:411-429)
    FIN_XR,             // ||R||, convergence, next rho and beta (:318-343, :439-495)
    FIN_XR_MIXED,       // same for the FP32 inner recurrence, reliable-update trigger
    FIN_UPDATE,         // reliable update: FP64 true residual, convergence, restart
    FIN_DOT,            // standalone dot product
};

// True when the phase kernel that wrote the partials returned early (its
// partials are stale): the finish kernel then does nothing either.
template <int F, typename PS>
__device__ __forceinline__ bool fin_skips(PS st) {
    if (F == FIN_SETUP || F == FIN_DOT) return false;
    if (st->done) return true;
    if (F == FIN_OMEGA) return st->s_exit != 0;
    if (F == FIN_UPDATE) return !st->need_update;
    return false;
}

template <int F, int NV, typename PS>
__device__ __forceinline__ void fin_apply(PS st, const double (&tot)[NV], double *out) {
    auto &S = *st;
    if (F == FIN_SETUP) {
        S.rnorm0 = sqrt(tot[0]);
        S.true_rnorm = S.rnorm0;
        S.upd_rnorm = S.rnorm0;
        if (S.zero_row) {
            S.ier = -2;
            S.exit_reason = EXIT_ZERO_ROW;
            S.done = 1;
        }
    } else if (F == FIN_SHADOW) {
        S.rho = tot[0];
        if (S.itmax < 1) {
            finish_itmax(st);
        } else if (tot[0] == 0.0) {
            S.ier = 0;
            S.exit_reason = EXIT_RHO;
            S.done = 1;
        }
    } else if (F == FIN_ALPHA) {
        S.rtv = tot[0];
        S.alpha = S.rho / tot[0];
    } else if (F == FIN_OMEGA) {
        if (NV > 2 && !S.minimize) {
            // S-norm exit (leq_bicgs.f:372-376); ||S||^2 from the S update
            S.snorm = sqrt(tot[NV > 2 ? 2 : 0]);
            if (S.snorm <= kTolMin) {
                S.s_exit = 1;
                return;
            }
        }
        double txt = tot[1];
        S.txs = tot[0];
        if (txt == 0.0) txt = kSmallNumber;
        S.txt = txt;
        S.omega = tot[0] / txt;
    } else if (F == FIN_XR) {
        if (S.s_exit) {
            S.isconverged = 1;
            S.ier = 0;
            S.exit_reason = EXIT_SNORM;
            S.done = 1;
            return;
        }
        const bool checked = !S.minimize || (S.iter_var % S.icheck) == 0;
        if (checked) {
            S.rnorm = sqrt(tot[0]);
            S.isconverged = S.rnorm <= S.tol * S.rnorm0;
            if (S.isconverged) {
                S.iter_tot_inc = S.iter_var + 1;
                S.ier = 0;
                S.exit_reason = EXIT_CONVERGED;
                S.done = 1;
                return;
            }
        }
        S.iter_var += 1;
        if (S.i >= S.itmax) {
            finish_itmax(st);
            return;
        }
        S.i += 1;
        const double rho_new = tot[1];
        if (rho_new == 0.0) {
            S.ier = -2;
            S.exit_reason = EXIT_RHO;
            S.done = 1;
            return;
        }
        S.beta = (rho_new / S.rho) * (S.alpha / S.omega);
        S.rho_prev = S.rho;
        S.rho = rho_new;
    } else if (F == FIN_XR_MIXED) {
        // FP32 inner recurrence.  A reliable update (FP64 true residual) is
        // requested when the recursive residual fell by delta since the last
        // one, reached the tolerance, the S-norm exit or rho = 0 occurred, or a
        // non-finite scalar appeared (then the correction y is dropped).
        S.y_reset = 0;
        S.restart = 0;
        S.iter_var += 1;
        const bool last = S.i >= S.itmax;
        if (!last) S.i += 1;
        if (!isfinite(tot[0]) || !isfinite(tot[1]) || !isfinite(S.alpha) || !isfinite(S.omega)) {
            S.need_update = 1;
            S.discard_y = 1;
            S.beta_invalid = 1;
        } else if (S.s_exit) {
            // the iteration ended at its S-norm exit: omega of this iteration
            // is stale, so the recurrence cannot continue
            S.need_update = 1;
            S.beta_invalid = 1;
        } else {
            S.rnorm = sqrt(tot[0]);
            const double rho_new = tot[1];
            if (S.rnorm <= S.tol * S.rnorm0 || S.rnorm <= S.delta * S.upd_rnorm || rho_new == 0.0) {
                S.need_update = 1;
                if (rho_new == 0.0) S.beta_invalid = 1;
            } else {
                S.beta = (rho_new / S.rho) * (S.alpha / S.omega);
                S.rho_prev = S.rho;
                S.rho = rho_new;
            }
        }
        if (last || (S.need_update && S.max_updates > 0 && S.updates + 1 >= S.max_updates)) {
            S.need_update = 1;
            S.final_update = 1;
            S.exit_reason = EXIT_ITMAX;
        }
    } else if (F == FIN_UPDATE) {
        // x <- x + y and the FP64 true residual were formed by the update
        // kernel.  Converged on ||b - A x|| <= TOL*Rnorm0.  Otherwise the FP32
        // recurrence continues from the true residual (residual replacement)
        // or restarts from it with a fresh shadow residual (P = R,
        // Rtilde = R, rho = <R,R>): always on systems of 2^18 cells and
        // more, where continuing lets rho = <Rtilde,R> decay below FP32's
        // rounding level in the hard phase (256^3: 6671 iterations against
        // 1099; 90^3: 1632 against 995), and on smaller ones only when the
        // update shows drift (the true residual grew since the last update
        // or is 10x the recursive one), since there restarting every cycle
        // costs up to 2x the iterations (32^3: 560 against 349).
        const bool discarded = S.discard_y;
        S.discard_y = 0;
        S.s_exit = 0;
        S.x_cur = 1 - S.x_cur;
        S.y_reset = 1;
        S.need_update = 0;
        S.updates += 1;
        const double prev_true = S.upd_rnorm;
        const double rec = S.rnorm;
        S.true_rnorm = sqrt(tot[0]);
        S.upd_rnorm = S.true_rnorm;
        S.rnorm = S.true_rnorm;
        if (!isfinite(S.true_rnorm)) {
            S.ier = -2;
            S.exit_reason = EXIT_RHO;
            S.done = 1;
            return;
        }
        if (S.true_rnorm <= S.tol * S.rnorm0) {
            S.isconverged = 1;
            S.ier = 0;
            S.exit_reason = EXIT_CONVERGED;
            S.done = 1;
            return;
        }
        if (S.final_update) {
            // ITMAX reached: classify like leq_bicgs.f:486-495 on the true residual
            S.ier = S.true_rnorm < kRatioTol * S.rnorm0 ? -1 : -2;
            S.done = 1;
            return;
        }
        const double rho_cont = tot[1];
        const bool drift = discarded || S.true_rnorm > prev_true || S.true_rnorm > 10.0 * rec;
        if (S.restart_always || drift || rho_cont == 0.0 || S.beta_invalid) {
            S.restart = 1;
            S.restarts += 1;
            S.rho = tot[2];
            if (S.rho == 0.0) {
                S.ier = -2;
                S.exit_reason = EXIT_RHO;
                S.done = 1;
            }
        } else {
            S.beta = (rho_cont / S.rho) * (S.alpha / S.omega);
            S.rho_prev = S.rho;
            S.rho = rho_cont;
        }
        S.beta_invalid = 0;
    } else if (F == FIN_DOT) {
        *out = tot[0];
    }
}

// ---- LINE preconditioner (LEQ_MSOLVE, leqsol_mod.f:286-598) ---------------
//
// One thread solves one line with the DGTSV algorithm of model/DGTSV.f:107-163
// (elimination with partial pivoting) and :233-246 (back solve), in the same
// operation order, so the result is the reference's bit for bit.  The
// elimination carries the two active rows in registers and reads each row's
// inputs (three tridiagonal coefficients and the right-hand side formed from
// the cross-line neighbours, LEQ_IKSWEEP leqsol_mod.f:891-902 and its JK/IJ
// twins) once, as it reaches them; only the finished rows (diagonal, first
// and second super-diagonal, right-hand side) go to scratch, each written
// once and read once by the back solve.  The thread's scratch rows are
// contiguous (one cache line serves 16 FP64 rows), so no step waits on a
// store it just made.  Lines of one phase are independent (red-black
// colours, anti-diagonal wavefronts), so a phase is one launch.

enum LineAxis : int { AX_J = 0, AX_I = 1, AX_K = 2 };
enum PhaseKind : int { PH_COLOR = 0, PH_DIAG = 1, PH_SEQ = 2 };

struct LinePhase {
    int axis;        // direction of the lines
    int kind;        // PH_COLOR: (u+v*vw) parity == param; PH_DIAG: u+v == param; PH_SEQ: reference order
    int param;
    int vw;          // weight of v in the colour (0 for the 2-D I-colouring)
    int fatal;       // INFO>0 is fatal (JK/IJ sweeps) instead of leaving the line unchanged
    int seq_s1;      // PH_SEQ: fast size and pass (icase) of the reference enumeration
    int seq_icase;
};

// Axis geometry of a line family: length n, stride along the line, and the
// (u, v) coordinates that label the lines (1-based like the reference).
struct LineGeom {
    int n, nu, nv;
    int64_t stride, su, sv;
};

__host__ __device__ inline LineGeom line_geom(const DGrid &g, int axis) {
    LineGeom L;
    if (axis == AX_J) { L.n = g.nj; L.stride = 1; L.nu = g.ni; L.su = g.c1; L.nv = g.nk; L.sv = g.c2; }
    else if (axis == AX_I) { L.n = g.ni; L.stride = g.c1; L.nu = g.nj; L.su = 1; L.nv = g.nk; L.sv = g.c2; }
    else { L.n = g.nk; L.stride = g.c2; L.nu = g.ni; L.su = g.c1; L.nv = g.nj; L.sv = 1; }
    return L;
}

// Scratch elements a line launch needs: 4 arrays of (lanes x line length);
// every launch here uses at most n3/n + 64 lanes of at most 64 threads.
__host__ __device__ inline int64_t line_scratch_elems(const DGrid &g) {
    const int nmax = g.ni > g.nj ? (g.ni > g.nk ? g.ni : g.nk) : (g.nj > g.nk ? g.nj : g.nk);
    return 4 * (g.n3 + 64 * (int64_t)nmax);
}

// Position p (1-based) of the reference's red-black enumeration of a 2-D
// index set (leqsol_mod.f:404-418) -> 1-based (fast, slow).
__device__ __forceinline__ void rb_position(int p, int s1, int *f, int *s) {
    int slow = (p % s1 != 0) ? p / s1 + 1 : p / s1;
    int fast = (p - 1 - (slow - 1) * s1) + 1 + (slow % 2);
    if (fast > s1) fast = fast - s1;
    *f = fast;
    *s = slow;
}

// Inputs of row q of a line: DL (coupling to row q-1, CC; 0 at q = 0),
// D, DU (coupling to row q+1, EE; 0 at the last row) and the right-hand
// side BB = B - cross-line neighbour terms in the reference's order.
template <typename T>
struct LineRow {
    T dl, d, du, b;
};

template <typename T>
__device__ __forceinline__ LineRow<T> line_row(const DGrid &g, const LineGeom &L, int axis, int64_t c0, int q,
                                               const T *__restrict__ a, const T *__restrict__ bm,
                                               const T *__restrict__ var) {
    const int64_t n3 = g.n3;
    const int pd = DiagPlane<T>::value;
    int plo, pup, x1m, x1p, x2m, x2p;
    if (axis == AX_J) { plo = 1; pup = 5; x1m = 2; x1p = 4; x2m = 0; x2p = 6; }
    else if (axis == AX_I) { plo = 2; pup = 4; x1m = 1; x1p = 5; x2m = 0; x2p = 6; }
    else { plo = 0; pup = 6; x1m = 1; x1p = 5; x2m = 2; x2p = 4; }
    const int64_t c = c0 + (int64_t)q * L.stride;
    const Nb nb = neighbours(g, c);
    int64_t m1, p1, m2, p2;
    if (axis == AX_J) { m1 = nb.im; p1 = nb.ip; m2 = nb.km; p2 = nb.kp; }
    else if (axis == AX_I) { m1 = nb.jm; p1 = nb.jp; m2 = nb.km; p2 = nb.kp; }
    else { m1 = nb.jm; p1 = nb.jp; m2 = nb.im; p2 = nb.ip; }
    LineRow<T> r;
    if (axis == AX_J && !g.do_k)
        r.b = bm[c] - a[x1m * n3 + c] * var[m1] - a[x1p * n3 + c] * var[p1];
    else
        r.b = bm[c] - a[x1m * n3 + c] * var[m1] - a[x1p * n3 + c] * var[p1] - a[x2m * n3 + c] * var[m2] -
              a[x2p * n3 + c] * var[p2];
    r.d = a[pd * n3 + c];
    r.dl = q == 0 ? (T)0 : a[plo * n3 + c];
    r.du = q == L.n - 1 ? (T)0 : a[pup * n3 + c];
    return r;
}

// Solve the line starting at cell c0 in place of var.  `scratch` holds 4
// arrays of nl x L.n elements; this thread uses slot `lane` of each.
template <typename T>
__device__ void solve_line(const DGrid &g, const LineGeom &L, int axis, int64_t c0, const T *__restrict__ a,
                           const T *__restrict__ bm, T *__restrict__ var, T *__restrict__ scratch, int64_t lane,
                           int64_t nl, int fatal, int *fatal_flag) {
    const int n = L.n;
    const int64_t plane = nl * n;
    T *__restrict__ sD = scratch + lane * n;       // final diagonal
    T *__restrict__ sU = sD + plane;               // first super-diagonal
    T *__restrict__ sU2 = sD + 2 * plane;          // second super-diagonal (pivoting fill, DL(i) after :127)
    T *__restrict__ sB = sD + 3 * plane;           // right-hand side
    LineRow<T> r0 = line_row<T>(g, L, axis, c0, 0, a, bm, var);
    T d0 = r0.d, du0 = r0.du, b0 = r0.b;
    int info = 0;
    LineRow<T> nx;
    if (n > 1) nx = line_row<T>(g, L, axis, c0, 1, a, bm, var);
    for (int i = 0; i < n - 1; ++i) {
        // row i+1 (its inputs were read one step ahead)
        const T dl1 = nx.dl;
        T d1 = nx.d, du1 = nx.du, b1 = nx.b;
        if (i + 2 < n) nx = line_row<T>(g, L, axis, c0, i + 2, a, bm, var);
        T fd, fu, fu2, fb;
        if (fabs((double)d0) >= fabs((double)dl1)) {
            // no row interchange (DGTSV.f:111-122 / :139-147)
            if (d0 == (T)0) {
                info = i + 1;
                break;
            }
            const T fact = dl1 / d0;
            d1 = d1 - fact * du0;
            b1 = b1 - fact * b0;
            fd = d0;
            fu = du0;
            fu2 = (T)0;
            fb = b0;
        } else {
            // interchange rows i and i+1 (DGTSV.f:123-136 / :148-158)
            const T fact = d0 / dl1;
            fd = dl1;
            T temp = d1;
            d1 = du0 - fact * temp;
            if (i < n - 2) {
                fu2 = du1;
                du1 = -fact * fu2;
            } else {
                fu2 = (T)0;  // the last interchange leaves no fill (never read)
            }
            fu = temp;
            temp = b0;
            fb = b1;
            b1 = temp - fact * b1;
        }
        sD[i] = fd;
        sU[i] = fu;
        sU2[i] = fu2;
        sB[i] = fb;
        d0 = d1;
        du0 = du1;
        b0 = b1;
    }
    if (!info && d0 == (T)0) info = n;  // DGTSV.f:160-163
    if (info) {
        if (fatal) atomicOr(fatal_flag, 1);
        return;  // the line keeps its current values (leqsol_mod.f:910-914)
    }
    // back solve, DGTSV.f:233-246
    const int64_t st = L.stride;
    T x1 = b0 / d0;
    var[c0 + (int64_t)(n - 1) * st] = x1;
    if (n == 1) return;
    T x0 = (sB[n - 2] - sU[n - 2] * x1) / sD[n - 2];
    var[c0 + (int64_t)(n - 2) * st] = x0;
    T x2 = x1;
    x1 = x0;
    for (int q = n - 3; q >= 0; --q) {
        const T x = (sB[q] - sU[q] * x1 - sU2[q] * x2) / sD[q];
        var[c0 + (int64_t)q * st] = x;
        x2 = x1;
        x1 = x;
    }
}

// One phase of line solves.  `st` (nullable) makes the launch a no-op once
// the solve is done, or, with skip_on_sexit, after the S-norm exit.
template <typename T>
__global__ void __launch_bounds__(64) k_line_phase(DGrid g, LinePhase ph, const T *__restrict__ a,
                                                   const T *__restrict__ bm, T *__restrict__ var, T *scratch,
                                                   const SolverState *st, int skip_on_sexit, int *fatal_flag) {
    if (st) {
        if (st->done) return;
        if (skip_on_sexit && st->s_exit) return;
    }
    const LineGeom L = line_geom(g, ph.axis);
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nl = (int64_t)gridDim.x * blockDim.x;
    if (ph.kind == PH_SEQ) {
        // reference enumeration of one red-black pass when the colouring is
        // not proper (odd fast size): one thread walks the lines in order
        if (tid != 0) return;
        const int total = L.nu * L.nv;
        for (int p = ph.seq_icase; p <= total; p += 2) {
            int f, s;
            rb_position(p, ph.seq_s1, &f, &s);
            const int64_t c0 = (int64_t)(f - 1) * L.su + (int64_t)(s - 1) * L.sv;
            solve_line<T>(g, L, ph.axis, c0, a, bm, var, scratch, 0, nl, ph.fatal, fatal_flag);
        }
        return;
    }
    int u, v;  // 0-based line coordinates
    if (ph.kind == PH_COLOR) {
        if (tid >= (int64_t)L.nu * L.nv) return;
        u = (int)(tid % L.nu);
        v = (int)(tid / L.nu);
        if ((((u + 1) + (v + 1) * ph.vw) & 1) != ph.param) return;
    } else {
        const int d = ph.param;
        const int lo = d - (L.nv - 1) > 0 ? d - (L.nv - 1) : 0;
        const int hi = d < L.nu - 1 ? d : L.nu - 1;
        if (tid > hi - lo) return;
        u = lo + (int)tid;
        v = d - u;
    }
    const int64_t c0 = (int64_t)u * L.su + (int64_t)v * L.sv;
    solve_line<T>(g, L, ph.axis, c0, a, bm, var, scratch, tid, nl, ph.fatal, fatal_flag);
}

// Lines of one red-black colour along J, numbered densely: line t of the
// colour -> (u, v) = (i-1, k-1).  The colour of line (u, v) is the parity of
// (u+1) + (v+1)*vw (vw = 1 in 3-D, leqsol_mod.f:497-522; vw = 0 in 2-D,
// where only I alternates, :367-392).  Row v holds the u = u0(v), u0+2, ...
// In 3-D two consecutive rows hold nu lines together.
__host__ __device__ inline bool color_line(int nu, int nv, int vw, int param, int64_t t, int &u, int &v) {
    if (vw == 0) {
        const int u0 = (param + 1) & 1;
        const int cnt = (nu - u0 + 1) / 2;
        if (cnt <= 0) return false;
        v = (int)(t / cnt);
        if (v >= nv) return false;
        u = u0 + 2 * (int)(t % cnt);
        return true;
    }
    const int64_t m = t / nu;
    const int r = (int)(t - m * nu);
    v = (int)(2 * m);
    const int u0 = (param + v) & 1;
    const int cnt0 = (nu - u0 + 1) / 2;
    if (r < cnt0) {
        u = u0 + 2 * r;
    } else {
        v += 1;
        u = ((param + v) & 1) + 2 * (r - cnt0);
    }
    return v < nv;
}

// Number of lines of a colour (an upper bound over both colours is
// (nu*nv+1)/2).
__host__ __device__ inline int64_t color_count(int nu, int nv, int vw, int param) {
    int64_t n = 0;
    for (int v = 0; v < nv; ++v) {
        const int u0 = vw ? ((param + v) & 1) : ((param + 1) & 1);
        n += (nu - u0 + 1) / 2;
    }
    return n;
}

}  // namespace mfix
