// The scalar transport equation's assembly on the device, first-order
// upwinding, phase M's slice of A_M/B_M:
//   INIT_AB_M + CONV_DIF_PHI with DISC = 0 and DEF_COR off
//     (model/init_ab_m.f:20-78, model/conv_dif_phi.f:17-80): CONV_DIF_PHI0
//     (:214-339) with GET_PHICELL_DIFF_TERMS (:87-198, no cut cells) and
//     DIF_PHI_IS (:722-836), which adds nothing without internal surfaces;
//   SOURCE_PHI (model/source_phi.f:25-139, no point sources), in place;
//   UNDER_RELAX_S (model/under_relax.f:20-81), in place.
// BC_PHI (model/bc_phi.f) runs between the first two in the reference; it
// stays with the caller, which is why the passes are separate.
//
// CONV_DIF_PHI0 scatters: a fluid cell writes its east/north/top face into
// its own row and into the west/south/bottom entry of its IP/JP/KP
// neighbour, and writes its own west/south/bottom entry when the IM/JM/KM
// neighbour is not fluid.  Here every row gathers instead, one thread per
// row: a fluid row's west entry is the face it shares with IM_OF, computed
// by one expression for both routes, and a non-fluid row's west entry is
// the east face of a fluid cell whose IP_OF it is.
#include <hip/hip_runtime.h>

#include "../../include/mfix_leq.h"
#include "leq_context.h"

namespace mfix {

constexpr int kPhiBlock = 256;

struct PhiArgs {
    DGrid g;
    mfix_phi_fields f;
    double *a, *b;
};

struct SourceArgs {
    DGrid g;
    const double *s_p, *s_c, *ep, *phi;
    const int *flag;
    double dil_ep_s;
    double *a, *b;
};

struct RelaxArgs {
    DGrid g;
    const double *var;
    const int *flag;
    double f1, f2;
    double *a, *b;
};

// AVG_X_H / AVG_Y_H / AVG_Z_H (model/include/fun_avg.inc:93-129): the
// harmonic mean of the two cells' values at a face, m on the minus side.
__device__ __forceinline__ double avg_h(double m, double p, double f, double fbar) {
    return m * p / fmax(kSmallNumber, f * m + fbar * p);
}

__global__ void __launch_bounds__(kPhiBlock) k_conv_dif_phi(PhiArgs R, int64_t nblk) {
    const DGrid &g = R.g;
    if (!xcd_valid(nblk)) return;
    const int64_t c = march_cell(g, xcd_block(nblk), kPhiBlock, threadIdx.x);
    if (c < 0) return;
    const int64_t n3 = g.n3;
    const mfix_phi_fields &F = R.f;
    const auto *__restrict__ flag = to_global(F.flag);
    const auto *__restrict__ dif = to_global(F.dif);
    const auto *__restrict__ u = to_global(F.u);
    const auto *__restrict__ v = to_global(F.v);
    const auto *__restrict__ w = to_global(F.w);
    const auto *__restrict__ fe = to_global(F.flux_e);
    const auto *__restrict__ fn = to_global(F.flux_n);
    const auto *__restrict__ ft = to_global(F.flux_t);
    const auto *__restrict__ ayz = to_global(F.ayz);
    const auto *__restrict__ axz = to_global(F.axz);
    const auto *__restrict__ axy = to_global(F.axy);
    const auto *__restrict__ odx_e = to_global(F.odx_e);
    const auto *__restrict__ ox = to_global(F.ox);
    const auto *__restrict__ fx_e = to_global(F.fx_e);
    const auto *__restrict__ fx_eb = to_global(F.fx_e_bar);
    const auto *__restrict__ ody_n = to_global(F.ody_n);
    const auto *__restrict__ fy_n = to_global(F.fy_n);
    const auto *__restrict__ fy_nb = to_global(F.fy_n_bar);
    const auto *__restrict__ odz_t = to_global(F.odz_t);
    const auto *__restrict__ fz_t = to_global(F.fz_t);
    const auto *__restrict__ fz_tb = to_global(F.fz_t_bar);
    int i, j, k;
    cell_ijk(g, c, i, j, k);
    // IM_OF.. as a fluid cell has them (set_index1a.f:85-222).  They stay
    // inside the grid for any cell, so a fluid row's coefficients are formed
    // for every cell from loads issued together, without a branch the
    // compiler could sink them into, and kept only on fluid rows (in 2-D
    // KM = KP = the cell and the K faces are dropped).
    const Nb n = neighbours(g, c);
    const int iw = nb_minus(i, g.ni, g.cyc_x), js = nb_minus(j, g.nj, g.cyc_y), kb = nb_minus(k, g.nk, g.cyc_z);
    const int fc = flag[c];
    const int f_ip = flag[n.ip], f_jp = flag[n.jp], f_kp = flag[n.kp];
    const int f_im = flag[n.im], f_jm = flag[n.jm], f_km = flag[n.km];
    const double dc = dif[c];
    const double d_ip = dif[n.ip], d_jp = dif[n.jp], d_kp = dif[n.kp];
    const double d_im = dif[n.im], d_jm = dif[n.jm], d_km = dif[n.km];
    const double uc = u[c], vc = v[c], wc = w[c], u_im = u[n.im], v_jm = v[n.jm], w_km = w[n.km];
    const double fec = fe[c], fnc = fn[c], ftc = ft[c], fe_im = fe[n.im], fn_jm = fn[n.jm], ft_km = ft[n.km];
    const double ayc = ayz[c], axzc = axz[c], axyc = axy[c];
    const double ay_im = ayz[n.im], axz_jm = axz[n.jm], axy_km = axy[n.km];
    const double gx = fx_e[i], gxb = fx_eb[i], ox_e = odx_e[i], gxw = fx_e[iw], gxwb = fx_eb[iw], ox_w = odx_e[iw];
    const double gy = fy_n[j], gyb = fy_nb[j], oy_n = ody_n[j], gys = fy_n[js], gysb = fy_nb[js], oy_s = ody_n[js];
    const double gz = fz_t[k], gzb = fz_tb[k], oz_t = odz_t[k], gzb_ = fz_t[kb], gzbb = fz_tb[kb], oz_b = odz_t[kb];
    const double oxi = ox[i];
    // own east/north/top faces; EAST_OF is IP_OF unless that is a wall
    const double de = wall_at(f_ip) ? dc : d_ip;
    const double dn = wall_at(f_jp) ? dc : d_jp;
    const double dt = wall_at(f_kp) ? dc : d_kp;
    const double dfe = avg_h(dc, de, gx, gxb) * (ox_e * ayc);
    const double dfn = avg_h(dc, dn, gy, gyb) * (oy_n * axzc);
    const double dft = avg_h(dc, dt, gz, gzb) * (oxi * oz_t * axyc);
    // west/south/bottom: a fluid IM_OF writes this face as its own east face
    // (EAST_OF(IM_OF) is this cell); otherwise this cell writes it with
    // WEST_OF, which is the cell itself when IM_OF is a wall.  Both are the
    // same expression.
    const double dw = wall_at(f_im) ? dc : d_im;
    const double ds = wall_at(f_jm) ? dc : d_jm;
    const double db = wall_at(f_km) ? dc : d_km;
    const double dfw = avg_h(dw, dc, gxw, gxwb) * (ox_w * ay_im);
    const double dfs = avg_h(ds, dc, gys, gysb) * (oy_s * axz_jm);
    const double dfb = avg_h(db, dc, gzb_, gzbb) * (oxi * oz_b * axy_km);
    const bool fluid = fluid_at(fc), fk = fluid && g.do_k;
    double ae = fluid ? (uc >= 0.0 ? dfe : dfe - fec) : 0.0;
    double an = fluid ? (vc >= 0.0 ? dfn : dfn - fnc) : 0.0;
    double at = fk ? (wc >= 0.0 ? dft : dft - ftc) : 0.0;
    double aw = fluid ? (u_im >= 0.0 ? dfw + fe_im : dfw) : 0.0;
    double as = fluid ? (v_jm >= 0.0 ? dfs + fn_jm : dfs) : 0.0;
    double ab = fk ? (w_km >= 0.0 ? dfb + ft_km : dfb) : 0.0;
    if (!fluid) {
        // a non-fluid row holds what a fluid cell whose IP_OF/JP_OF/KP_OF it
        // is scatters into it: that cell's east/north/top face, with
        // EAST_OF = the fluid cell itself when this row is a wall.  The
        // sources, in storage order (the later write stands): the plain -1
        // neighbour, and on a cyclic axis the last interior cell, whose
        // IP_OF wraps to the first interior layer.
        auto from = [&](int p, int n, int cyc, int64_t st, auto o, auto fx, auto fxb, auto area, auto vel, auto flux,
                        double oxi) {
            double out = 0.0;
            const int src[2] = {(p > 0 && nb_plus(p - 1, n, cyc) == p) ? p - 1 : -1,
                                (cyc && p == 1 && n >= 3) ? n - 2 : -1};
            for (int q = 0; q < 2; ++q) {
                const int t = src[q];
                if (t < 0) continue;
                const int64_t d = c + (int64_t)(t - p) * st;
                if (!fluid_at(flag[d])) continue;
                const int64_t e = wall_at(flag[c]) ? d : c;
                const double df = avg_h(dif[d], dif[e], fx[t], fxb[t]) * (oxi * o[t] * area[d]);
                out = vel[d] >= 0.0 ? df + flux[d] : df;
            }
            return out;
        };
        // C_AE = ODX_E(I)*AYZ, C_AN = ODY_N(J)*AXZ, C_AT = OX(I)*ODZ_T(K)*AXY
        // (1.0 * ODX_E is ODX_E exactly)
        aw = from(i, g.ni, g.cyc_x, g.c1, odx_e, fx_e, fx_eb, ayz, u, fe, 1.0);
        as = from(j, g.nj, g.cyc_y, 1, ody_n, fy_n, fy_nb, axz, v, fn, 1.0);
        if (g.do_k) ab = from(k, g.nk, g.cyc_z, g.c2, odz_t, fz_t, fz_tb, axy, w, ft, ox[i]);
    }
    auto *__restrict__ a = to_global(R.a);
    a[0 * n3 + c] = ab;
    a[1 * n3 + c] = as;
    a[2 * n3 + c] = aw;
    a[3 * n3 + c] = -1.0;   // INIT_AB_M; SOURCE_PHI forms A0
    a[4 * n3 + c] = ae;
    a[5 * n3 + c] = an;
    a[6 * n3 + c] = at;
    to_global(R.b)[c] = 0.0;
}

// SOURCE_PHI on the fluid rows.  Its dilute branch asks IS_AT_E(IMJK) etc.
// Without internal surfaces SET_FLAGS (model/set_flags.f:403-470) leaves a
// face flagged below 2000 only where a wall cell lies on it (a cyclic
// boundary face is set back to 2000 when the ghost layer carries FLAG
// 106/107, as SET_FLAGS gives it); a wall neighbour makes WEST_OF
// (EAST_OF ..) the cell itself, whose EP is at most DIL_EP_s, so the test
// cannot change the result and is left out.
__global__ void __launch_bounds__(kPhiBlock) k_source_phi(SourceArgs R) {
    const DGrid &g = R.g;
    const int64_t c = (int64_t)blockIdx.x * kPhiBlock + threadIdx.x;
    if (c >= g.n3) return;
    const int64_t n3 = g.n3;
    const auto *__restrict__ flag = to_global(R.flag);
    const auto *__restrict__ ep = to_global(R.ep);
    const auto *__restrict__ phi = to_global(R.phi);
    auto *__restrict__ a = to_global(R.a);
    auto *__restrict__ b = to_global(R.b);
    const double dil = R.dil_ep_s;
    // every load of a dense fluid row in one round; the row is formed for
    // every cell and stored only where SOURCE_PHI writes it
    const int fc = flag[c];
    const double epc = ep[c], ph = phi[c], sp = to_global(R.s_p)[c], sc = to_global(R.s_c)[c], b0 = b[c];
    const double ab = a[0 * n3 + c], as = a[1 * n3 + c], aw = a[2 * n3 + c];
    const double ae = a[4 * n3 + c], an = a[5 * n3 + c], at = a[6 * n3 + c];
    const double a0 = -(ae + aw + an + as + at + ab + sp);
    const bool keep_b = b0 < sc || ph == 0.0;
    const double a0d = keep_b ? a0 : a0 - b0 / ph;
    const double bd = keep_b ? -sc + b0 : -sc;
    if (!fluid_at(fc)) return;
    if (epc > dil) {
        a[3 * n3 + c] = a0d;
        b[c] = bd;
        return;
    }
    const Nb n = neighbours(g, c);
    auto side = [&](int64_t mp) { return ep[wall_at(flag[mp]) ? c : mp] > dil ? 1.0 : 0.0; };
    const double sw = side(n.im), se = side(n.ip), ss = side(n.jm), sn = side(n.jp);
    const double sb = g.do_k ? side(n.km) : 0.0, st = g.do_k ? side(n.kp) : 0.0;
    double d0 = -1.0, db = 0.0;
    if ((sw + se + ss + sn + sb + st) == 0.0)
        db = -ph;
    else
        d0 = -(se + sw + sn + ss + st + sb);
    a[0 * n3 + c] = sb;
    a[1 * n3 + c] = ss;
    a[2 * n3 + c] = sw;
    a[3 * n3 + c] = d0;
    a[4 * n3 + c] = se;
    a[5 * n3 + c] = sn;
    a[6 * n3 + c] = st;
    b[c] = db;
}

__global__ void __launch_bounds__(kPhiBlock) k_under_relax_s(RelaxArgs R) {
    const int64_t c = (int64_t)blockIdx.x * kPhiBlock + threadIdx.x;
    if (c >= R.g.n3) return;
    auto *__restrict__ a0 = to_global(R.a) + 3 * R.g.n3;
    auto *__restrict__ b = to_global(R.b);
    const int fc = to_global(R.flag)[c];
    const double ap = a0[c], bc = b[c], var = to_global(R.var)[c];
    const double na = ap * R.f1, nb = bc + ap * var * R.f2;
    if (fluid_at(fc) && ap != -1.0) {
        a0[c] = na;
        b[c] = nb;
    }
}

static int phi_ready(hipStream_t *s, void *stream, DGrid *g) {
    Context &C = ctx();
    int rc;
    if ((rc = C.ensure_init())) return rc;
    if (!C.have_grid) return MFIX_LEQ_ENOGRID;
    *s = stream ? (hipStream_t)stream : C.stream;
    *g = make_dgrid(C.grid);
    return 0;
}

static unsigned phi_blocks(const DGrid &g) { return (unsigned)((g.n3 + kPhiBlock - 1) / kPhiBlock); }

static int launched(const char *what) {
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : hip_fail(e, what);
}

}  // namespace mfix

using namespace mfix;

extern "C" {

int mfix_conv_dif_phi_dev(const mfix_phi_fields *f, double *d_a_m, double *d_b_m, void *stream) {
    hipStream_t s;
    PhiArgs R;
    int rc;
    if ((rc = phi_ready(&s, stream, &R.g))) return rc;
    if (!f || !f->dif || !f->u || !f->v || !f->w || !f->flux_e || !f->flux_n || !f->flux_t || !f->ayz ||
        !f->axz || !f->axy || !f->odx_e || !f->ox || !f->fx_e || !f->fx_e_bar || !f->ody_n || !f->fy_n ||
        !f->fy_n_bar || !f->odz_t || !f->fz_t || !f->fz_t_bar || !f->flag || !d_a_m || !d_b_m)
        return MFIX_LEQ_EINVAL;
    R.f = *f;
    R.a = d_a_m;
    R.b = d_b_m;
    const int64_t nblk = march_blocks(R.g.c2, R.g.nk, kPhiBlock);
    hipLaunchKernelGGL(k_conv_dif_phi, dim3((unsigned)xcd_grid(nblk)), dim3(kPhiBlock), 0, s, R, nblk);
    return launched("CONV_DIF_PHI");
}

int mfix_source_phi_dev(const double *d_s_p, const double *d_s_c, const double *d_ep, const double *d_phi,
                        const int *d_flag, double dil_ep_s, double *d_a_m, double *d_b_m, void *stream) {
    hipStream_t s;
    SourceArgs R;
    int rc;
    if ((rc = phi_ready(&s, stream, &R.g))) return rc;
    if (!d_s_p || !d_s_c || !d_ep || !d_phi || !d_flag || !d_a_m || !d_b_m) return MFIX_LEQ_EINVAL;
    R.s_p = d_s_p;
    R.s_c = d_s_c;
    R.ep = d_ep;
    R.phi = d_phi;
    R.flag = d_flag;
    R.dil_ep_s = dil_ep_s;
    R.a = d_a_m;
    R.b = d_b_m;
    hipLaunchKernelGGL(k_source_phi, dim3(phi_blocks(R.g)), dim3(kPhiBlock), 0, s, R);
    return launched("SOURCE_PHI");
}

int mfix_under_relax_s_dev(const double *d_var, const int *d_flag, double ur, double *d_a_m, double *d_b_m,
                           void *stream) {
    hipStream_t s;
    RelaxArgs R;
    int rc;
    if ((rc = phi_ready(&s, stream, &R.g))) return rc;
    if (!d_var || !d_flag || !d_a_m || !d_b_m || !(ur != 0.0)) return MFIX_LEQ_EINVAL;
    R.var = d_var;
    R.flag = d_flag;
    R.f1 = 1.0 / ur;   // F1 = ONE/UR, F2 = F1 - ONE (under_relax.f:67-68)
    R.f2 = R.f1 - 1.0;
    R.a = d_a_m;
    R.b = d_b_m;
    hipLaunchKernelGGL(k_under_relax_s, dim3(phi_blocks(R.g)), dim3(kPhiBlock), 0, s, R);
    return launched("UNDER_RELAX_S");
}

}  // extern "C"
