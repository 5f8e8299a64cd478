// The gas pressure-correction equation assembled and solved on the device:
// INIT_AB_M + CONV_PP_G + SOURCE_PP_G (model/init_ab_m.f:20-78,
// model/conv_pp_g.f:33-182, model/source_pp_g.f:18-233) for phase 0 with
// MMAX = 0, RO_G0 defined (no COMPRESSIBLE_PP_G), no cut cells, no shear,
// no point or user sources; and SOLVE_PP_G (model/solve_pp_g.f:25-159)
// around it.  A_M/B_M never cross PCIe: the fields the assembly reads are
// resident, and the solver runs on the assembled arrays in place.
//
// CONV_PP_G scatters: a fluid cell writes its east/north/top face value
// ROP_g?(c)*A??(c) into its own row and into the west/south/bottom entry of
// its IP/JP/KP neighbour, and a fluid cell whose IM/JM/KM neighbour is not
// fluid writes that neighbour's face value into its own west/south/bottom
// entry.  For a fluid row both routes give the same number, the face value
// of the IM/JM/KM neighbour (IP_OF(IM_OF(c)) = c for every fluid cell), so
// the kernel gathers: one thread per cell, no write conflicts.
#include <hip/hip_runtime.h>

#include "../../include/mfix_leq.h"
#include "leq_context.h"

namespace mfix {

constexpr int kAsmBlock = 256;

struct AsmArgs {
    DGrid g;
    mfix_pp_g_fields f;
    double *a, *b, *bmmax;
    unsigned *bad;
};

__global__ void __launch_bounds__(kAsmBlock) k_assemble_pp_g(AsmArgs R) {
    const DGrid &g = R.g;
    const int64_t c = (int64_t)blockIdx.x * kAsmBlock + threadIdx.x;
    if (c >= g.n3) return;
    const int64_t n3 = g.n3;
    const mfix_pp_g_fields &F = R.f;
    auto *__restrict__ a = to_global(R.a);
    auto *__restrict__ b = to_global(R.b);
    const auto *__restrict__ flag = to_global(F.flag);
    double ae = 0.0, aw = 0.0, an = 0.0, as = 0.0, at = 0.0, ab = 0.0, a0 = -1.0, bc = 0.0;
    if (flag[c] == 1) {
        const Nb n = neighbours(g, c);   // IM_OF.. of a fluid cell (set_index1a.f:85-222)
        const int64_t im = n.im, jm = n.jm, km = n.km;
        const auto *__restrict__ rop_ge = to_global(F.rop_ge);
        const auto *__restrict__ rop_gn = to_global(F.rop_gn);
        const auto *__restrict__ rop_gt = to_global(F.rop_gt);
        const auto *__restrict__ ayz = to_global(F.ayz);
        const auto *__restrict__ axz = to_global(F.axz);
        const auto *__restrict__ axy = to_global(F.axy);
        const auto *__restrict__ vol = to_global(F.vol);
        // CONV_PP_G (conv_pp_g.f:69-113); INIT_AB_M leaves 0 where it writes nothing
        ae = rop_ge[c] * ayz[c];
        aw = rop_ge[im] * ayz[im];
        an = rop_gn[c] * axz[c];
        as = rop_gn[jm] * axz[jm];
        if (g.do_k) {
            at = rop_gt[c] * axy[c];
            ab = rop_gt[km] * axy[km];
        }
        // SOURCE_PP_G, gas terms (source_pp_g.f:77-101)
        const auto *__restrict__ u = to_global(F.u_g);
        const auto *__restrict__ v = to_global(F.v_g);
        const auto *__restrict__ w = to_global(F.w_g);
        const double bma = (to_global(F.rop_g)[c] - to_global(F.rop_go)[c]) * vol[c] * F.odt;
        const double bme = ae * u[c];
        const double bmw = aw * u[im];
        const double bmn = an * v[c];
        const double bms = as * v[jm];
        const double bmt = at * w[c];
        const double bmb = ab * w[km];
        const double bmr = to_global(F.sum_r_g)[c] * vol[c];
        bc = -((-(bma + bme - bmw + bmn - bms + bmt - bmb)) + bmr);
        double m = fabs(bma);
        m = fmax(m, fabs(bme));
        m = fmax(m, fabs(bmw));
        m = fmax(m, fabs(bmn));
        m = fmax(m, fabs(bms));
        m = fmax(m, fabs(bmt));
        m = fmax(m, fabs(bmb));
        m = fmax(m, fabs(bmr));
        to_global(R.bmmax)[c] = m;
        const auto *__restrict__ d_e = to_global(F.d_e);
        const auto *__restrict__ d_n = to_global(F.d_n);
        const auto *__restrict__ d_t = to_global(F.d_t);
        ae = ae * d_e[c];
        aw = aw * d_e[im];
        an = an * d_n[c];
        as = as * d_n[jm];
        at = at * d_t[c];
        ab = ab * d_t[km];
        a0 = -(ae + aw + an + as + at + ab);
        if (fabs(a0) < kSmallNumber) {
            if (fabs(bc) < kSmallNumber) {
                a0 = -1.0;
                bc = 0.0;
            } else {
                // the reference reports the cell and carries on
                atomicAdd((unsigned *)R.bad, 1u);
            }
        }
        // faces toward pressure-flow cells are cut after A0 is formed
        // (source_pp_g.f:193-212)
        if (pflow_at(flag[im])) aw = 0.0;
        if (pflow_at(flag[n.ip])) ae = 0.0;
        if (pflow_at(flag[jm])) as = 0.0;
        if (pflow_at(flag[n.jp])) an = 0.0;
        if (pflow_at(flag[km])) ab = 0.0;
        if (pflow_at(flag[n.kp])) at = 0.0;
    }
    if (c == (int64_t)F.ijk_p_g - 1) {   // reference pressure cell (source_pp_g.f:214-218)
        ae = aw = an = as = at = ab = 0.0;
        a0 = -1.0;
        bc = 0.0;
    }
    a[0 * n3 + c] = ab;
    a[1 * n3 + c] = as;
    a[2 * n3 + c] = aw;
    a[3 * n3 + c] = a0;
    a[4 * n3 + c] = ae;
    a[5 * n3 + c] = an;
    a[6 * n3 + c] = at;
    b[c] = bc;
}

static int assemble(const mfix_pp_g_fields *f, double *a, double *b, double *bmmax, int *n_singular, hipStream_t s) {
    Context &C = ctx();
    int rc;
    AsmArgs R;
    R.g = make_dgrid(C.grid);
    R.f = *f;
    R.a = a;
    R.b = b;
    R.bmmax = bmmax;
    if ((rc = C.state.ensure(sizeof(SolverState) + 256))) return rc;
    R.bad = (unsigned *)((char *)C.state.ptr + sizeof(SolverState));
    hipError_t e = hipMemsetAsync(R.bad, 0, sizeof(unsigned), s);
    if (e != hipSuccess) return hip_fail(e, "assembly counter");
    hipLaunchKernelGGL(k_assemble_pp_g, dim3((unsigned)((R.g.n3 + kAsmBlock - 1) / kAsmBlock)), dim3(kAsmBlock), 0, s,
                       R);
    unsigned bad = 0;
    e = hipMemcpyAsync(&bad, R.bad, sizeof(unsigned), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "pressure-correction assembly");
    if (n_singular) *n_singular = (int)bad;
    return 0;
}

static bool fields_ok(const mfix_pp_g_fields *f) {
    return f && f->rop_ge && f->rop_gn && f->rop_gt && f->ayz && f->axz && f->axy && f->vol && f->rop_g &&
           f->rop_go && f->u_g && f->v_g && f->w_g && f->sum_r_g && f->d_e && f->d_n && f->d_t && f->flag;
}

static int asm_ready(hipStream_t *s, void *stream) {
    Context &C = ctx();
    int rc;
    if ((rc = C.ensure_init())) return rc;
    if (!C.have_grid) return MFIX_LEQ_ENOGRID;
    *s = stream ? (hipStream_t)stream : C.stream;
    return 0;
}

}  // namespace mfix

using namespace mfix;

extern "C" {

int mfix_assemble_pp_g_dev(const mfix_pp_g_fields *fields, double *d_a_m, double *d_b_m, double *d_b_mmax,
                           int *n_singular, void *stream) {
    hipStream_t s;
    int rc;
    if ((rc = asm_ready(&s, stream))) return rc;
    if (!fields_ok(fields) || !d_a_m || !d_b_m || !d_b_mmax) return MFIX_LEQ_EINVAL;
    if (fields->ijk_p_g < 0 || fields->ijk_p_g > make_dgrid(ctx().grid).n3) return MFIX_LEQ_EINVAL;
    return assemble(fields, d_a_m, d_b_m, d_b_mmax, n_singular, s);
}

int mfix_solve_pp_g_dev(const mfix_pp_g_fields *fields, double normg, double *d_pp_g, double *d_a_m, double *d_b_m,
                        double *d_b_mmax, int itmax, int method, const char *sweep, double tol, const char *pc,
                        int mixed, double *resg, int *ier, mfix_leq_stats *stats, void *stream) {
    hipStream_t s;
    int rc;
    if (method != 2) return MFIX_LEQ_EMETHOD;
    const int p = parse_pc(pc);
    if (p < 0) return MFIX_LEQ_EBADPC;
    if ((rc = asm_ready(&s, stream))) return rc;
    if (!fields_ok(fields) || !d_pp_g || !d_a_m || !d_b_m || !d_b_mmax || !resg || !ier) return MFIX_LEQ_EINVAL;
    Context &C = ctx();
    const DGrid g = make_dgrid(C.grid);
    if (fields->ijk_p_g < 0 || fields->ijk_p_g > g.n3) return MFIX_LEQ_EINVAL;
    // PP_G(:) = ZERO (solve_pp_g.f:75), then the assembly
    hipError_t e = hipMemsetAsync(d_pp_g, 0, sizeof(double) * g.n3, s);
    if (e != hipSuccess) return hip_fail(e, "PP_G reset");
    int bad = 0;
    if ((rc = assemble(fields, d_a_m, d_b_m, d_b_mmax, &bad, s))) return rc;
    // normalisation of the residual (solve_pp_g.f:100-112): NORMG = 0 takes
    // CALC_RESID_PP of B_MMAX over DEN = 10
    mfix_leq_resid r;
    double normloc = normg;
    if (normg == 0.0) {
        if ((rc = resid_pp(d_b_mmax, fields->flag, 1.0, 1, &r, s))) return rc;
        normloc = r.resid / 1.0e1;
    }
    if ((rc = resid_pp(d_b_m, fields->flag, normloc, 1, &r, s))) return rc;
    *resg = r.resid;
    // SOLVE_LIN_EQ('Pp_g', 1, PP_G, A_M, B_M, 0, ...) -> LEQ_BICGS (solve_lin_eq.f:151-152)
    mfix_leq_stats st = {};
    char cm[5] = "RSRS";
    for (int q = 0; q < 4 && sweep && sweep[q]; ++q) cm[q] = sweep[q];
    rc = mixed ? solve_mixed(g, p, cm, tol, itmax, d_pp_g, d_a_m, d_b_m, nullptr, &st, s)
               : solve_fp64(g, p, cm, tol, itmax, d_pp_g, d_a_m, d_b_m, nullptr, &st, s, C.grid.minimize_dotproducts,
                            C.grid.icheck_bicgs);
    if (rc) return rc;
    *ier = st.ier;
    C.iter_tot[0] += st.iter_tot_increment;   // ITER_TOT(1), VNO = 1 for 'Pp_g'
    if (stats) *stats = st;
    (void)bad;
    return 0;
}

}  // extern "C"
