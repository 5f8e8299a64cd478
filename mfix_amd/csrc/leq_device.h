// Device-side building blocks of the gfx950 LEQ_BICGS solver.
//
// Storage follows the reference's serial layout (include/mfix_leq.h): J is
// the stride-1 axis, seven SoA coefficient planes, one ghost layer.  Plane
// index p = l + 3 for stencil slot l (model/param_mod.f:28-50):
//   p0 = bottom K-, p1 = south J-, p2 = west I-, p3 = diagonal,
//   p4 = east I+, p5 = north J+, p6 = top K+.
//
// Floating-point contraction is off in every translation unit that includes
// this header, so FP64 products and sums round exactly as the reference's
// x86-64 build does (separate multiply and add, left-to-right order of
// model/leqsol_mod.f:195-204).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace mfix {

constexpr int kNumXcd = 8;       // MI355X: 8 XCDs, blocks dispatched round-robin
constexpr int kNumCus = 256;     // MI355X: 256 CUs (32 per XCD)
constexpr int kFinishBlock = 1024;

// Grid descriptor passed by value to every kernel.
struct DGrid {
    int ni, nj, nk;       // cells per axis including the two ghosts (nk=1 in 2-D)
    int do_k;             // 3-D run
    int cyc_x, cyc_y, cyc_z;
    int dot_all;          // dot_product_par domain: 1 = every cell (IS_SERIAL), 0 = interior
    int64_t c1, c2, n3;   // strides of I and K, DIMENSION_3
    uint64_t mc2, mnj;    // reciprocals of c2 and nj for the cell decode (div_magic)
    int sc2, snj;
};

// Division by an invariant divisor d >= 1 as a multiply and a shift
// (Granlund and Montgomery, "Division by invariant integers using
// multiplication", 1994, Thm. 4.2): with l = ceil(log2 d) and
// m = ceil(2^(32+l) / d), floor(n / d) = (n * m) >> (32 + l) for every
// 0 <= n < 2^32; with n < 2^31 (DIMENSION_3 < 2^31 is checked at grid set-up)
// the product fits in 64 bits.
inline void div_magic(uint32_t d, uint64_t &m, int &s) {
    int l = 0;
    while ((uint64_t(1) << l) < d) ++l;
    s = 32 + l;
    m = ((uint64_t(1) << s) + d - 1) / d;
}
__device__ __forceinline__ uint32_t div_by(uint32_t n, uint64_t m, int s) {
    return (uint32_t)(((uint64_t)n * m) >> s);
}

// Neighbour along one axis (0-based), set_increments.f:72-112: clamped to the
// cell itself at the array ends; wrapped for cyclic axes on interior cells.
__device__ __forceinline__ int nb_minus(int i, int n, int cyc) {
    const bool shift = !(i == 0 || i == n - 1);
    if (cyc && shift) { const int t = i - 1; return t == 0 ? n - 2 : t; }
    return i > 0 ? i - 1 : 0;
}
__device__ __forceinline__ int nb_plus(int i, int n, int cyc) {
    const bool shift = !(i == 0 || i == n - 1);
    if (cyc && shift) { const int t = i + 1; return t == n - 1 ? 1 : t; }
    return i < n - 1 ? i + 1 : n - 1;
}

struct Nb {
    int64_t im, ip, jm, jp, km, kp;
    bool interior;
};

// (i,j,k) of storage index c (0-based), by the precomputed reciprocals
// (DIMENSION_3 < 2^31, checked at grid set-up).
__device__ __forceinline__ void cell_ijk(const DGrid &g, int64_t c, int &i, int &j, int &k) {
    const uint32_t cc = (uint32_t)c, c2 = (uint32_t)g.c2, nj = (uint32_t)g.nj;
    const uint32_t kk = div_by(cc, g.mc2, g.sc2);
    const uint32_t m = cc - kk * c2;
    const uint32_t ii = div_by(m, g.mnj, g.snj);
    i = (int)ii;
    j = (int)(m - ii * nj);
    k = (int)kk;
}

__device__ __forceinline__ bool is_interior(const DGrid &g, int64_t c) {
    int i, j, k;
    cell_ijk(g, c, i, j, k);
    const bool ij = i > 0 && i < g.ni - 1 && j > 0 && j < g.nj - 1;
    return g.do_k ? (ij && k > 0 && k < g.nk - 1) : ij;
}

// FLAG predicates, model/include/functions.inc:409-414 (FLUID_AT),
// :417-423 (P_FLOW_AT) and :469-474 (WALL_AT).
__device__ __forceinline__ bool fluid_at(int f) { return f == 1; }
__device__ __forceinline__ bool wall_at(int f) { return f >= 100; }
__device__ __forceinline__ bool pflow_at(int f) { return f == 10 || f == 11; }

__device__ __forceinline__ Nb neighbours(const DGrid &g, int64_t c) {
    int i, j, k;
    cell_ijk(g, c, i, j, k);
    Nb n;
    n.jm = c + (nb_minus(j, g.nj, g.cyc_y) - j);
    n.jp = c + (nb_plus(j, g.nj, g.cyc_y) - j);
    n.im = c + (int64_t)(nb_minus(i, g.ni, g.cyc_x) - i) * g.c1;
    n.ip = c + (int64_t)(nb_plus(i, g.ni, g.cyc_x) - i) * g.c1;
    if (g.do_k) {
        n.km = c + (int64_t)(nb_minus(k, g.nk, g.cyc_z) - k) * g.c2;
        n.kp = c + (int64_t)(nb_plus(k, g.nk, g.cyc_z) - k) * g.c2;
        n.interior = i > 0 && i < g.ni - 1 && j > 0 && j < g.nj - 1 && k > 0 && k < g.nk - 1;
    } else {
        n.km = n.kp = c;
        n.interior = i > 0 && i < g.ni - 1 && j > 0 && j < g.nj - 1;
    }
    return n;
}

// Neighbours with a fast path: away from the two outermost layers every
// cell has the fixed increments -1/+1, -c1/+c1, -c2/+c2 (cyclic wraps and
// clamps only touch layers 0, 1, n-2, n-1), so only a thin shell of cells
// evaluates the per-axis rules.
__device__ __forceinline__ Nb neighbours_fast(const DGrid &g, int64_t c) {
    int i, j, k;
    cell_ijk(g, c, i, j, k);
    const bool core = i > 1 && i < g.ni - 2 && j > 1 && j < g.nj - 2 && (!g.do_k || (k > 1 && k < g.nk - 2));
    Nb n;
    if (core) {
        n.jm = c - 1;
        n.jp = c + 1;
        n.im = c - g.c1;
        n.ip = c + g.c1;
        n.km = g.do_k ? c - g.c2 : c;
        n.kp = g.do_k ? c + g.c2 : c;
        n.interior = true;
        return n;
    }
    n.jm = c + (nb_minus(j, g.nj, g.cyc_y) - j);
    n.jp = c + (nb_plus(j, g.nj, g.cyc_y) - j);
    n.im = c + (int64_t)(nb_minus(i, g.ni, g.cyc_x) - i) * g.c1;
    n.ip = c + (int64_t)(nb_plus(i, g.ni, g.cyc_x) - i) * g.c1;
    if (g.do_k) {
        n.km = c + (int64_t)(nb_minus(k, g.nk, g.cyc_z) - k) * g.c2;
        n.kp = c + (int64_t)(nb_plus(k, g.nk, g.cyc_z) - k) * g.c2;
        n.interior = i > 0 && i < g.ni - 1 && j > 0 && j < g.nj - 1 && k > 0 && k < g.nk - 1;
    } else {
        n.km = n.kp = c;
        n.interior = i > 0 && i < g.ni - 1 && j > 0 && j < g.nj - 1;
    }
    return n;
}

// Pointers read from a system table are generic; these casts tell the
// compiler they address global memory, so loads and stores use the global
// (not flat) instructions, which do not make LDS waits also wait for them.
template <typename T>
using gptr = __attribute__((address_space(1))) T *;
template <typename T>
__device__ __forceinline__ gptr<T> to_global(T *p) {
    return (gptr<T>)p;
}


// Streaming loads of the coefficient planes (read once per sweep): the
// non-temporal hint keeps them from displacing the reused vector in L2.
template <typename P>
__device__ __forceinline__ auto ld_stream(P p) {
    return __builtin_nontemporal_load(p);
}

// DIAG preconditioner applied to one value.  FP64 (the reference's
// LEQ_MSOLVE1): z = r / A0, one correctly rounded division.  FP32 inner
// solve: the diagonal plane holds 1/A0 rounded to FP32 and z = r * (1/A0),
// one multiply where a division costs about ten FP64 instructions; the
// same product is used by the stencil and by the solution update, so the
// preconditioned operator stays consistent.
__device__ __forceinline__ double diag_apply(double r, double a0) { return r / a0; }
__device__ __forceinline__ float diag_apply(float r, float inv_a0) { return r * inv_a0; }

// ---- 16-byte cell groups -----------------------------------------------------
//
// The fused streaming kernels give each thread VN = 16/sizeof(T) consecutive
// cells (4 FP32, 2 FP64) and move each array with one 16-byte access per
// lane: a quarter of the load instructions and of the wave launches of one
// cell per thread.  The vectors need only element alignment (neighbour
// groups at c +- c1 and plane starts p*n3 are generally not 16-byte
// aligned); gfx950 serves such accesses as dwordx4.
typedef float f4v __attribute__((ext_vector_type(4), aligned(4)));
typedef double d2v __attribute__((ext_vector_type(2), aligned(8)));
typedef double d4v __attribute__((ext_vector_type(4), aligned(8)));

template <typename E, int N> struct VecOf;
template <> struct VecOf<float, 4> { using type = f4v; };
template <> struct VecOf<double, 2> { using type = d2v; };
template <> struct VecOf<double, 4> { using type = d4v; };

template <typename T> struct Group16 { static constexpr int N = 16 / sizeof(T); };
template <typename E, int N> using vec_t = typename VecOf<E, N>::type;

// The update passes read each array once, so their loads carry the
// non-temporal hint; their stores stay plain: the next pass reads what they
// store, and at 256^3 (FP32) plain stores ran the solve 0.7 % faster than
// non-temporal ones (574.8 against 578.8 ms, same iterations).
template <typename E, int N, typename P>
__device__ __forceinline__ vec_t<E, N> ldv(P p) {
    return __builtin_nontemporal_load((const __attribute__((address_space(1))) vec_t<E, N> *)p);
}
// Plain (cached) vector load, for values other threads read again.
template <typename E, int N, typename P>
__device__ __forceinline__ vec_t<E, N> ldv_keep(P p) {
    return *(const __attribute__((address_space(1))) vec_t<E, N> *)p;
}
template <typename E, int N, typename P>
__device__ __forceinline__ void stv(P p, vec_t<E, N> v) {
    *(__attribute__((address_space(1))) vec_t<E, N> *)p = v;
}

// Operand fetch of the stencil: x itself, or x/A0 of the neighbour (DIAG
// preconditioner applied on the fly, LEQ_MSOLVE1 leqsol_mod.f:729-744, so
// Phat and Shat never exist as arrays).  Both forms round exactly like the
// reference's separate division and matvec.  OP_NEG is DIAG on a system
// whose every unit-scaled A0 is exactly -1 (each row's largest coefficient
// is its diagonal, as in MFIX's diagonally dominant equations): x/(-1) and
// x*(1/(-1)) are exactly -x, so the A0 plane is not read at all.
enum OperandOp : int { OP_PLAIN = 0, OP_DIAG = 1, OP_NEG = 2 };

template <typename T, int OP, typename PT = const T *>
struct Operand {
    PT x;
    PT a0;
    __device__ __forceinline__ T operator()(int64_t q) const {
        if (OP == OP_DIAG) return diag_apply(x[q], a0[q]);
        if (OP == OP_NEG) return -x[q];
        return x[q];
    }
};

// Row arithmetic of LEQ_MATVEC in the reference's summation order, on
// values already loaded (a_l = plane l+3 of the row, x.. = operand values).
// A 2-D grid (km = kp = c) passes its K-plane values too and the select
// drops their terms, leaving the reference's sum untouched.
template <typename T>
__device__ __forceinline__ T stencil_combine(bool do_k, T a0, T a1, T a2, T a3, T a4, T a5, T a6, T xkm, T xim,
                                             T xjm, T xc, T xjp, T xip, T xkp) {
    T v = a1 * xjm + a2 * xim;
    v = v + a3 * xc;
    v = v + a4 * xip;
    v = v + a5 * xjp;
    T vk = v + a0 * xkm;
    vk = vk + a6 * xkp;
    return do_k ? vk : v;
}

// FP32 stencil in difference form: with the diagonal plane replaced by the
// row excess e = A0 + sum(off-diagonals) (formed in FP64, then rounded),
// y = sum_l a_l (x_nb - x_c) + e x_c is algebraically LEQ_MATVEC but avoids
// the cancellation between A0 x_c and the neighbour terms that costs a
// diffusion operator's smooth modes most of FP32's digits.  a3 = e here.
__device__ __forceinline__ float stencil_combine_diff(bool do_k, float a0, float a1, float a2, float a3, float a4,
                                                      float a5, float a6, float xkm, float xim, float xjm, float xc,
                                                      float xjp, float xip, float xkp) {
    double v = (double)(a1 * (xjm - xc)) + (double)(a2 * (xim - xc));
    v += (double)(a4 * (xip - xc));
    v += (double)(a5 * (xjp - xc));
    double vk = v + (double)(a0 * (xkm - xc));
    vk += (double)(a6 * (xkp - xc));
    v = do_k ? vk : v;
    v += (double)a3 * (double)xc;
    return (float)v;
}

// y(c) of LEQ_MATVEC.  Every load of the row is issued before its
// arithmetic, so a thread waits on memory once per row.
template <typename T, typename X, typename PA>
__device__ __forceinline__ T stencil_row(const DGrid &g, PA a, const X &x, int64_t c, const Nb &n) {
    const int64_t n3 = g.n3;
    const T a1 = ld_stream(&a[1 * n3 + c]), a2 = ld_stream(&a[2 * n3 + c]), a3 = ld_stream(&a[3 * n3 + c]);
    const T a4 = ld_stream(&a[4 * n3 + c]), a5 = ld_stream(&a[5 * n3 + c]);
    const T a0 = ld_stream(&a[0 * n3 + c]), a6 = ld_stream(&a[6 * n3 + c]);
    const T xjm = x(n.jm), xim = x(n.im), xc = x(c), xip = x(n.ip), xjp = x(n.jp);
    const T xkm = x(n.km), xkp = x(n.kp);
    return stencil_combine<T>(g.do_k, a0, a1, a2, a3, a4, a5, a6, xkm, xim, xjm, xc, xjp, xip, xkp);
}

template <typename X, typename PA>
__device__ __forceinline__ float stencil_row_diff(const DGrid &g, PA a, const X &x, int64_t c, const Nb &n) {
    const int64_t n3 = g.n3;
    const float a1 = ld_stream(&a[1 * n3 + c]), a2 = ld_stream(&a[2 * n3 + c]), a3 = ld_stream(&a[3 * n3 + c]);
    const float a4 = ld_stream(&a[4 * n3 + c]), a5 = ld_stream(&a[5 * n3 + c]);
    const float a0 = ld_stream(&a[0 * n3 + c]), a6 = ld_stream(&a[6 * n3 + c]);
    const float xc = x(c), xjm = x(n.jm), xim = x(n.im), xip = x(n.ip), xjp = x(n.jp);
    const float xkm = x(n.km), xkp = x(n.kp);
    return stencil_combine_diff(g.do_k, a0, a1, a2, a3, a4, a5, a6, xkm, xim, xjm, xc, xjp, xip, xkp);
}

// ---- CALC_RESID rows ---------------------------------------------------------

enum ResidKind : int { RESID_C = 0, RESID_S = 1, RESID_PP = 2 };

// One row of CALC_RESID_C / _S / _PP (model/calc_resid.f:86-110, :280-303,
// :440-447), serial semantics: false for a cell the sums skip (ghost layer,
// not FLUID_AT; for S |VAR| <= tol or a dilute solids cell).  NUM1 =
// |B - sum A*VAR| with a wall neighbour replaced by the cell itself for C
// (EAST_OF and the like, model/set_index1a.f:93-97), DEN1 = |A0*VAR|; for PP
// NUM1 = |B|, DEN1 = 1.
template <typename PV, typename PF>
__device__ __forceinline__ bool resid_row_vals(const DGrid &g, int kind, double tol, const double *ep_s,
                                               double dil_ep_s, const double (&ar)[7], double bc, double vc, PV var,
                                               PF flag, int64_t c, double &num1, double &den1) {
    if (c >= g.n3 || !is_interior(g, c) || !fluid_at(flag[c])) return false;  // IS_ON_myPE_wobnd, FLUID_AT
    if (kind == RESID_PP) {
        num1 = fabs(bc);
        den1 = 1.0;
        return true;
    }
    if (kind == RESID_S) {
        if (!(fabs(vc) > tol)) return false;
        if (ep_s && to_global(ep_s)[c] <= dil_ep_s) return false;
    }
    const Nb n = neighbours_fast(g, c);
    int64_t e = n.ip, w = n.im, no = n.jp, so = n.jm, t = n.kp, bo = n.km;
    if (kind == RESID_C) {
        e = wall_at(flag[e]) ? c : e;
        w = wall_at(flag[w]) ? c : w;
        no = wall_at(flag[no]) ? c : no;
        so = wall_at(flag[so]) ? c : so;
        t = wall_at(flag[t]) ? c : t;
        bo = wall_at(flag[bo]) ? c : bo;
    }
    // B - (A0 V + AE VE + AW VW + AN VN + AS VS), then the top and bottom
    // pair (calc_resid.f:94-103), in that order
    double r = bc - (ar[3] * vc + ar[4] * var[e] + ar[2] * var[w] + ar[5] * var[no] + ar[1] * var[so]);
    if (g.do_k) r = r - (ar[6] * var[t] + ar[0] * var[bo]);
    num1 = fabs(r);
    den1 = fabs(ar[3] * vc);
    return true;
}

// resid_row_vals with the row, B and VAR loaded here (PP reads only B).
template <typename PA, typename PV, typename PF>
__device__ __forceinline__ bool resid_row(const DGrid &g, int kind, double tol, const double *ep_s, double dil_ep_s,
                                          PA a, PA b, PV var, PF flag, int64_t c, double &num1, double &den1) {
    if (c >= g.n3 || !is_interior(g, c) || !fluid_at(flag[c])) return false;
    double ar[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double vc = 0.0;
    if (kind != RESID_PP) {
        const int64_t n3 = g.n3;
#pragma unroll
        for (int p = 0; p < 7; ++p) ar[p] = a[p * n3 + c];
        vc = var[c];
    }
    return resid_row_vals(g, kind, tol, ep_s, dil_ep_s, ar, b[c], vc, var, flag, c, num1, den1);
}

// The larger residual candidate (value, 0-based cell); equal values keep
// the first cell, as the reference's scan with a strict '>' does.
__device__ __forceinline__ void resid_better(double &v, double &i, double ov, double oi) {
    if (ov > v || (ov == v && oi < i)) {
        v = ov;
        i = oi;
    }
}
constexpr int kResidVals = 5;  // partial slots per block: NUM, DEN, fluid cells, best value, best cell

// ---- launch shape ---------------------------------------------------------
//
// A launch has nblk logical blocks of BS threads; logical block b owns the
// contiguous cells [b*chunk, (b+1)*chunk), chunk = BS * cells-per-thread.
// Hardware places block b on XCD b % 8; the remap below gives XCD x one
// contiguous range of logical blocks, i.e. one slab of the grid, so the
// I/K neighbours its blocks read were brought into the same L2.
struct Part {
    int64_t nblk, chunk, n;
};

inline Part make_part(int64_t n, int bs, int cpt) {
    Part p;
    p.n = n;
    p.chunk = (int64_t)bs * (cpt > 0 ? cpt : 1);
    p.nblk = (n + p.chunk - 1) / p.chunk;
    if (p.nblk < 1) p.nblk = 1;
    return p;
}

// Blocks given to one system by the reliable-update launch: a multiple of
// the XCD count (so every system's map range starts on XCD 0), at most
// kUpdateBlocks.
constexpr int kUpdateBlocks = 1024;
__host__ __device__ __forceinline__ int64_t update_blocks(int64_t nblk) {
    const int64_t r = (nblk + kNumXcd - 1) / kNumXcd * kNumXcd;
    return r < kUpdateBlocks ? r : kUpdateBlocks;
}
// Partial-sum slots per array of a system: one per logical block of the
// streaming kernels, and at least one per block of the reliable-update launch.
__host__ __device__ __forceinline__ int64_t partial_slots(int64_t nblk) {
    return nblk > kUpdateBlocks ? nblk : (int64_t)kUpdateBlocks;
}

__device__ __forceinline__ int64_t xcd_block(int64_t nblocks) {
    const int64_t b = blockIdx.x;
    const int64_t per = (nblocks + kNumXcd - 1) / kNumXcd;
    const int64_t x = b % kNumXcd, r = b / kNumXcd;
    const int64_t full = nblocks - (per - 1) * kNumXcd;  // XCDs that get `per` blocks
    const int64_t start = x < full ? x * per : full * per + (x - full) * (per - 1);
    return start + r;
}
// Hardware blocks to launch for nblocks logical ones (every XCD gets
// ceil(nblocks/8) slots; surplus slots exit immediately).
__host__ __device__ inline int64_t xcd_grid(int64_t nblocks) {
    const int64_t per = (nblocks + kNumXcd - 1) / kNumXcd;
    return per * kNumXcd;
}
__device__ __forceinline__ bool xcd_valid(int64_t nblocks) {
    const int64_t b = blockIdx.x;
    const int64_t per = (nblocks + kNumXcd - 1) / kNumXcd;
    const int64_t x = b % kNumXcd, r = b / kNumXcd;
    const int64_t full = nblocks - (per - 1) * kNumXcd;
    return r < (x < full ? per : per - 1);
}

// Plane-marching order for the one-pass assembly kernels, which read many
// fields at K-1/K+1 and I-1/I+1.  The in-plane indices [0, c2) are cut into
// chunks of bs cells; kMarchGroup consecutive chunks form a group, and a
// group's blocks run plane after plane (K fastest).  The K-1 plane a block
// reads was loaded by the block kMarchGroup logical blocks earlier, and the
// I-1/I+1 rows lie in the same group except at its two edges, so both are
// L2 hits; with the XCD remap above each XCD marches its own groups.
constexpr int kMarchGroup = 8;
__host__ __device__ inline int64_t march_blocks(int64_t c2, int nk, int bs) {
    const int64_t groups = ((c2 + bs - 1) / bs + kMarchGroup - 1) / kMarchGroup;
    return groups * kMarchGroup * nk;
}
// Cell of thread t of logical block L, or -1 past the plane's last cell.
// Logical blocks number fewer than 2^31 (DIMENSION_3 < 2^31).
__device__ __forceinline__ int64_t march_cell(const DGrid &g, int64_t L, int bs, int t) {
    const uint32_t per_group = (uint32_t)kMarchGroup * (uint32_t)g.nk;
    const uint32_t l = (uint32_t)L, grp = l / per_group, r = l - grp * per_group;
    const uint32_t k = r / kMarchGroup, q = r - k * kMarchGroup;
    const int64_t m = ((int64_t)grp * kMarchGroup + q) * bs + t;
    return m < g.c2 ? (int64_t)k * g.c2 + m : -1;
}

// ---- deterministic reductions (fixed tree, wave64 shuffles) ----------------
// block_sum leaves the block totals in thread 0; lds holds NV * (BS/64).
template <int BS, int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *lds) {
    constexpr int NW = BS / 64;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double s = v[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
        v[q] = s;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) lds[q * NW + w] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double s = lds[q * NW + 0];
            for (int u = 1; u < NW; ++u) s += lds[q * NW + u];
            v[q] = s;
        }
    }
}

// Grid-wide sums in two steps.  Each block stores its NV partial sums with
// plain stores; the finish kernel launched next sums them in a fixed order
// (finish_sum below).  The kernel boundary makes the partials visible across
// the eight XCDs' L2s once per launch, where a per-block release fence
// (last-block-done pattern) would write back each XCD's L2 for every block.
template <int BS, int NV, typename PD>
__device__ __forceinline__ void block_partial(double (&v)[NV], PD partials, int64_t nblk, int64_t blk,
                                              double *lds) {
    block_sum<BS, NV>(v, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) partials[q * nblk + blk] = v[q];
    }
}

// block_partial without a barrier at the block's end: each wave leaves its
// sums in LDS and counts itself in; the last wave of the block to arrive adds
// the wave sums in wave order (the order block_sum uses, so the partials are
// the same bits) and stores them.  Waves that finish early retire at once
// instead of holding their slots until the slowest wave of the block reaches
// a barrier.  *cnt must be zero before any wave arrives: partial_begin, at
// the kernel's start.
template <int BS>
__device__ __forceinline__ void partial_begin(unsigned *cnt) {
    if (BS > 64) {
        if (threadIdx.x == 0) *cnt = 0u;
        __syncthreads();
    }
}

template <int BS, int NV, typename PD>
__device__ __forceinline__ void block_partial_last(double (&v)[NV], PD partials, int64_t nblk, int64_t blk,
                                                   double *lds, unsigned *cnt) {
    constexpr int NW = BS / 64;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double s = v[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
        v[q] = s;
    }
    if ((threadIdx.x & 63) != 0) return;
    if (NW == 1) {
#pragma unroll
        for (int q = 0; q < NV; ++q) partials[q * nblk + blk] = v[q];
        return;
    }
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) lds[q * NW + w] = v[q];
    const unsigned before = __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (before != (unsigned)(NW - 1)) return;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double s = lds[q * NW + 0];
        for (int u = 1; u < NW; ++u) s += lds[q * NW + u];
        partials[q * nblk + blk] = s;
    }
}

// Finish step, run by nfb <= kFinishBlocks blocks of kFinishBlock threads
// (fb = this block's index among them).  Block fb sums the partials of
// slice fb in a fixed order into level-2 sums; the
// last block to arrive (counted with one device-scope release increment
// per block; the last takes an acquire fence) adds the level-2 sums in slice
// order, so the result does not depend on arrival order.  Only the few
// level-2 values are written before each release, so it has almost nothing
// to write back.  Returns true in the last block, with the totals in thread
// 0; the counter is reset there for the next launch.
constexpr int kFinishBlocks = 64;

template <typename P>
__device__ __forceinline__ double load_agent(P p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int NV, typename PP, typename PL, typename PC>
__device__ __forceinline__ bool finish_sum(PP partials, int64_t nblk, int fb, int nfb, PL l2, PC counter,
                                           double *lds, double (&tot)[NV]) {
    __shared__ bool last;
    const int64_t per = (nblk + nfb - 1) / nfb;
    const int64_t lo = (int64_t)fb * per;
    const int64_t hi = lo + per < nblk ? lo + per : nblk;
    // the NV sums advance together, so their loads share one round trip;
    // each still adds its partials in block order
#pragma unroll
    for (int q = 0; q < NV; ++q) tot[q] = 0.0;
    for (int64_t b = lo + threadIdx.x; b < hi; b += kFinishBlock) {
        double pv[NV];
#pragma unroll
        for (int q = 0; q < NV; ++q) pv[q] = partials[q * nblk + b];
#pragma unroll
        for (int q = 0; q < NV; ++q) tot[q] += pv[q];
    }
    block_sum<kFinishBlock, NV>(tot, lds);
    if (nfb == 1) return true;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) l2[q * nfb + fb] = tot[q];
        // release: this block's level-2 sums are visible before its count
        const unsigned prev = __hip_atomic_fetch_add((unsigned *)counter, 1u, __ATOMIC_RELEASE,
                                                     __HIP_MEMORY_SCOPE_AGENT);
        last = prev == (unsigned)nfb - 1;
    }
    __syncthreads();
    if (!last) return false;
    // acquire: every other block's sums, released before its count, are read
    // after the count that made this block the last
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
#pragma unroll
    for (int q = 0; q < NV; ++q) tot[q] = threadIdx.x < nfb ? load_agent(&l2[q * nfb + threadIdx.x]) : 0.0;
    __syncthreads();
    block_sum<kFinishBlock, NV>(tot, lds);
    if (threadIdx.x == 0) *counter = 0u;
    return true;
}

// Counter-based U[0,1) (same construction as the host generator).
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double urand(uint64_t seed, uint64_t stream, uint64_t idx) {
    uint64_t h = mix64(seed * 0x9E3779B97F4A7C15ULL + 0x632BE59BD9B4E019ULL);
    h = mix64(h ^ (stream * 0xD1B54A32D192ED03ULL));
    h = mix64(h + idx * 0xAEF17502108EF2D9ULL);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

}  // namespace mfix
