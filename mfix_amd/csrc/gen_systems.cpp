// Synthetic A_M/B_M systems in MFIX's serial storage layout.
//
// Layout (model/include/functions.inc:37-42, model/dmp_modules/gridmap_mod.f:553-556):
// one ghost layer per side, J stride 1, I stride nj, K stride nj*ni; seven
// SoA coefficient planes l=-3..3 (model/param_mod.f:28-50).  Sign convention
// of the assembled MFIX systems: off-diagonals >= 0, diagonal negative
// (model/conv_pp_g.f:82-88, model/source_pp_g.f:146-153,
// model/conv_dif_phi.f:276-290).  Ghost rows are identity-like: A0=-1,
// off-diagonals 0, B=0 (model/init_ab_m.f:56-76).
//
// Every random number is a pure function of (seed, stream, index), so the
// systems are identical whatever the thread count.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

#include "../../include/mfix_leq.h"

namespace {

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Counter-based uniform in [0,1) with 53 random bits.
inline double urand(uint64_t seed, uint64_t stream, uint64_t idx) {
    uint64_t h = mix64(seed * 0x9E3779B97F4A7C15ULL + 0x632BE59BD9B4E019ULL);
    h = mix64(h ^ (stream * 0xD1B54A32D192ED03ULL));
    h = mix64(h + idx * 0xAEF17502108EF2D9ULL);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

enum Stream : uint64_t {
    S_FACE_X = 1, S_FACE_Y = 2, S_FACE_Z = 3, S_DIAG = 4, S_RHS = 5,
    S_SPHERE = 6, S_MODE = 7, S_DIFF_X = 8, S_DIFF_Y = 9, S_DIFF_Z = 10,
};

struct Dims {
    int64_t ni, nj, nk, n3;
    int imax, jmax, kmax;
    int64_t idx(int i, int j, int k) const {  // 1-based cell indices
        return (int64_t)(j - 1) + (int64_t)(i - 1) * nj + (int64_t)(k - 1) * nj * ni;
    }
    bool interior(int i, int j, int k) const {
        return i >= 2 && i <= imax + 1 && j >= 2 && j <= jmax + 1 && k >= 2 && k <= kmax + 1;
    }
};

// Face value holder: coefficient of the face between cell c and its +axis
// neighbour, stored at the lower cell's ijk; 0 on faces touching a ghost.
struct Faces {
    std::vector<double> fx, fy, fz;
};

void ghost_rows(const Dims &d, double *a, double *b) {
    // zero everything, then identity rows everywhere; interior rows are
    // overwritten by the caller
    std::memset(a, 0, sizeof(double) * 7 * d.n3);
    for (int64_t c = 0; c < d.n3; ++c) {
        a[3 * d.n3 + c] = -1.0;
        b[c] = 0.0;
    }
}

// Symmetric (kinds 0 and 2): interior faces 10^U[-1,1], optionally scaled by
// the face-averaged void fraction.
void gen_pressure(const Dims &d, uint64_t seed, const std::vector<float> *eps,
                  double rhs_amp, double *a, double *b) {
    ghost_rows(d, a, b);
    const int64_t n3 = d.n3;
    auto face = [&](uint64_t stream, int64_t c0, int64_t c1) -> double {
        double u = urand(seed, stream, (uint64_t)c0);
        double v = std::pow(10.0, 2.0 * u - 1.0);
        if (eps) v *= 0.5 * ((double)(*eps)[c0] + (double)(*eps)[c1]);
        return v;
    };
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 2; k <= d.kmax + 1; ++k) {
        for (int i = 2; i <= d.imax + 1; ++i) {
            for (int j = 2; j <= d.jmax + 1; ++j) {
                const int64_t c = d.idx(i, j, k);
                // east (I+) / west (I-)
                double ae = (i < d.imax + 1) ? face(S_FACE_X, c, d.idx(i + 1, j, k)) : 0.0;
                double aw = (i > 2) ? face(S_FACE_X, d.idx(i - 1, j, k), c) : 0.0;
                double an = (j < d.jmax + 1) ? face(S_FACE_Y, c, d.idx(i, j + 1, k)) : 0.0;
                double as = (j > 2) ? face(S_FACE_Y, d.idx(i, j - 1, k), c) : 0.0;
                double at = (k < d.kmax + 1) ? face(S_FACE_Z, c, d.idx(i, j, k + 1)) : 0.0;
                double ab = (k > 2) ? face(S_FACE_Z, d.idx(i, j, k - 1), c) : 0.0;
                a[0 * n3 + c] = ab;
                a[1 * n3 + c] = as;
                a[2 * n3 + c] = aw;
                a[4 * n3 + c] = ae;
                a[5 * n3 + c] = an;
                a[6 * n3 + c] = at;
                double sum = ab + as + aw + ae + an + at;
                a[3 * n3 + c] = -sum - 1.0e-3 * urand(seed, S_DIAG, (uint64_t)c);
                b[c] = rhs_amp * (2.0 * urand(seed, S_RHS, (uint64_t)c) - 1.0);
            }
        }
    }
}

// eps_g: 1.0 background, 0.4 inside nspheres seeded spheres (centre uniform
// in the interior box, radius U[3,12] cells).  Stored per cell (ghosts 1.0).
std::vector<float> void_fraction(const Dims &d, uint64_t seed, int nspheres) {
    std::vector<float> eps((size_t)d.n3, 1.0f);
    for (int s = 0; s < nspheres; ++s) {
        double cx = d.imax * urand(seed, S_SPHERE, 4ULL * s + 0);
        double cy = d.jmax * urand(seed, S_SPHERE, 4ULL * s + 1);
        double cz = d.kmax * urand(seed, S_SPHERE, 4ULL * s + 2);
        double r = 3.0 + 9.0 * urand(seed, S_SPHERE, 4ULL * s + 3);
        // interior cell (i,j,k) has centre ((i-2)+0.5, (j-2)+0.5, (k-2)+0.5)
        int i0 = std::max(2, (int)std::floor(cx - r + 1.5));
        int i1 = std::min(d.imax + 1, (int)std::ceil(cx + r + 1.5));
        int j0 = std::max(2, (int)std::floor(cy - r + 1.5));
        int j1 = std::min(d.jmax + 1, (int)std::ceil(cy + r + 1.5));
        int k0 = std::max(2, (int)std::floor(cz - r + 1.5));
        int k1 = std::min(d.kmax + 1, (int)std::ceil(cz + r + 1.5));
        for (int k = k0; k <= k1; ++k)
            for (int i = i0; i <= i1; ++i)
                for (int j = j0; j <= j1; ++j) {
                    double dx = (i - 1.5) - cx, dy = (j - 1.5) - cy, dz = (k - 1.5) - cz;
                    if (dx * dx + dy * dy + dz * dz <= r * r) eps[(size_t)d.idx(i, j, k)] = 0.4f;
                }
    }
    return eps;
}

// Nonsymmetric first-order upwind momentum-like system (kind 1).
// Face fluxes are the discrete curl of a vector potential psi made of 8
// seeded Fourier modes sin(pi p x/L) sin(pi q y/M) sin(pi r z/N) that vanish on
// the box boundary, so every cell's net outflow is zero (div curl = 0) and
// boundary faces carry no flux.  Upwind coefficients follow
// model/conv_dif_phi.f:276-290: A_E = D_e + max(-F_e,0), A_W = D_w + max(F_w,0).
void gen_momentum(const Dims &d, uint64_t seed, double *a, double *b) {
    ghost_rows(d, a, b);
    const int NM = 8;
    const int nx = d.imax, ny = d.jmax, nz = d.kmax;
    double amp[NM][3];
    int px[NM], py[NM], pz[NM];
    for (int m = 0; m < NM; ++m) {
        px[m] = 1 + (int)(4.0 * urand(seed, S_MODE, 8ULL * m + 0));
        py[m] = 1 + (int)(4.0 * urand(seed, S_MODE, 8ULL * m + 1));
        pz[m] = 1 + (int)(4.0 * urand(seed, S_MODE, 8ULL * m + 2));
        for (int c = 0; c < 3; ++c) amp[m][c] = 2.0 * urand(seed, S_MODE, 8ULL * m + 3 + c) - 1.0;
    }
    // separable tables: sin(pi p t / n) at t = h/2 for h = 0..2n (half-integer grid)
    auto table = [](int n, int p) {
        std::vector<double> t((size_t)(2 * n + 1));
        for (int h = 0; h <= 2 * n; ++h) t[(size_t)h] = (h == 0 || h == 2 * n) ? 0.0 : std::sin(M_PI * p * (0.5 * h) / n);
        return t;
    };
    std::vector<std::vector<double>> tx(NM), ty(NM), tz(NM);
    for (int m = 0; m < NM; ++m) { tx[m] = table(nx, px[m]); ty[m] = table(ny, py[m]); tz[m] = table(nz, pz[m]); }
    // psi_c at half-integer coordinates (2x, 2y, 2z)
    auto psi = [&](int c, int hx, int hy, int hz) {
        double s = 0.0;
        for (int m = 0; m < NM; ++m) s += amp[m][c] * tx[m][(size_t)hx] * ty[m][(size_t)hy] * tz[m][(size_t)hz];
        return s;
    };
    // face fluxes in box coordinates: x-face at X=f (f=0..nx) of cell column (y,z)
    // interior cell indices ix=0..nx-1 etc; cell centre at ix+0.5
    const int64_t nfx = (int64_t)(nx + 1) * ny * nz, nfy = (int64_t)nx * (ny + 1) * nz, nfz = (int64_t)nx * ny * (nz + 1);
    std::vector<double> Fx((size_t)nfx), Fy((size_t)nfy), Fz((size_t)nfz);
    auto ix_ = [&](int f, int y, int z) { return ((int64_t)z * ny + y) * (nx + 1) + f; };
    auto iy_ = [&](int x, int f, int z) { return ((int64_t)z * (ny + 1) + f) * nx + x; };
    auto iz_ = [&](int x, int y, int f) { return ((int64_t)f * ny + y) * nx + x; };
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int f = 0; f <= nx; ++f) {
                // x-face at (X=f, y+.5, z+.5): d(psi_z)/dy - d(psi_y)/dz
                double v = (psi(2, 2 * f, 2 * y + 2, 2 * z + 1) - psi(2, 2 * f, 2 * y, 2 * z + 1))
                         - (psi(1, 2 * f, 2 * y + 1, 2 * z + 2) - psi(1, 2 * f, 2 * y + 1, 2 * z));
                Fx[(size_t)ix_(f, y, z)] = (f == 0 || f == nx) ? 0.0 : v;
            }
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int f = 0; f <= ny; ++f)
            for (int x = 0; x < nx; ++x) {
                // y-face at (x+.5, Y=f, z+.5): d(psi_x)/dz - d(psi_z)/dx
                double v = (psi(0, 2 * x + 1, 2 * f, 2 * z + 2) - psi(0, 2 * x + 1, 2 * f, 2 * z))
                         - (psi(2, 2 * x + 2, 2 * f, 2 * z + 1) - psi(2, 2 * x, 2 * f, 2 * z + 1));
                Fy[(size_t)iy_(x, f, z)] = (f == 0 || f == ny) ? 0.0 : v;
            }
#pragma omp parallel for collapse(2) schedule(static)
    for (int f = 0; f <= nz; ++f)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                // z-face at (x+.5, y+.5, Z=f): d(psi_y)/dx - d(psi_x)/dy
                double v = (psi(1, 2 * x + 2, 2 * y + 1, 2 * f) - psi(1, 2 * x, 2 * y + 1, 2 * f))
                         - (psi(0, 2 * x + 1, 2 * y + 2, 2 * f) - psi(0, 2 * x + 1, 2 * y, 2 * f));
                Fz[(size_t)iz_(x, y, f)] = (f == 0 || f == nz) ? 0.0 : v;
            }
    // diffusion per face (interior faces only; walls carry D = 0)
    auto Dx = [&](int f, int y, int z) { return (f == 0 || f == nx) ? 0.0 : 0.5 * std::pow(4.0, urand(seed, S_DIFF_X, (uint64_t)ix_(f, y, z))); };
    auto Dy = [&](int x, int f, int z) { return (f == 0 || f == ny) ? 0.0 : 0.5 * std::pow(4.0, urand(seed, S_DIFF_Y, (uint64_t)iy_(x, f, z))); };
    auto Dz = [&](int x, int y, int f) { return (f == 0 || f == nz) ? 0.0 : 0.5 * std::pow(4.0, urand(seed, S_DIFF_Z, (uint64_t)iz_(x, y, f))); };
    // amplitude: max cell Peclet |F|/D = 10
    double pe = 0.0, fmax = 0.0;
#pragma omp parallel for collapse(2) reduction(max : pe)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                if (x + 1 < nx) pe = std::max(pe, std::fabs(Fx[(size_t)ix_(x + 1, y, z)]) / Dx(x + 1, y, z));
                if (y + 1 < ny) pe = std::max(pe, std::fabs(Fy[(size_t)iy_(x, y + 1, z)]) / Dy(x, y + 1, z));
                if (z + 1 < nz) pe = std::max(pe, std::fabs(Fz[(size_t)iz_(x, y, z + 1)]) / Dz(x, y, z + 1));
            }
    const double scale = pe > 0.0 ? 10.0 / pe : 0.0;
#pragma omp parallel for reduction(max : fmax)
    for (int64_t q = 0; q < nfx; ++q) fmax = std::max(fmax, std::fabs(Fx[(size_t)q]) * scale);
#pragma omp parallel for reduction(max : fmax)
    for (int64_t q = 0; q < nfy; ++q) fmax = std::max(fmax, std::fabs(Fy[(size_t)q]) * scale);
#pragma omp parallel for reduction(max : fmax)
    for (int64_t q = 0; q < nfz; ++q) fmax = std::max(fmax, std::fabs(Fz[(size_t)q]) * scale);
    // transient term rho*V/dt with rho = V = dx = 1 and dt = 0.5*dx/max|u|
    const double transient = fmax > 0.0 ? fmax / 0.5 : 1.0;
    const int64_t n3 = d.n3;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < ny; ++y) {
                const int64_t c = d.idx(x + 2, y + 2, z + 2);
                double fe = Fx[(size_t)ix_(x + 1, y, z)] * scale, fw = Fx[(size_t)ix_(x, y, z)] * scale;
                double fn = Fy[(size_t)iy_(x, y + 1, z)] * scale, fs = Fy[(size_t)iy_(x, y, z)] * scale;
                double ft = Fz[(size_t)iz_(x, y, z + 1)] * scale, fb = Fz[(size_t)iz_(x, y, z)] * scale;
                double ae = Dx(x + 1, y, z) + std::max(-fe, 0.0);
                double aw = Dx(x, y, z) + std::max(fw, 0.0);
                double an = Dy(x, y + 1, z) + std::max(-fn, 0.0);
                double as = Dy(x, y, z) + std::max(fs, 0.0);
                double at = Dz(x, y, z + 1) + std::max(-ft, 0.0);
                double ab = Dz(x, y, z) + std::max(fb, 0.0);
                a[0 * n3 + c] = ab;
                a[1 * n3 + c] = as;
                a[2 * n3 + c] = aw;
                a[4 * n3 + c] = ae;
                a[5 * n3 + c] = an;
                a[6 * n3 + c] = at;
                a[3 * n3 + c] = -(ab + as + aw + ae + an + at + transient);
                b[c] = 2.0 * urand(seed, S_RHS, (uint64_t)c) - 1.0;
            }
}

}  // namespace

extern "C" int mfix_leq_gen_system(int kind, int imax, int jmax, int kmax, uint64_t seed,
                                   int nspheres, double *a_m, double *b_m) {
    if (imax < 1 || jmax < 1 || kmax < 1 || !a_m || !b_m) return MFIX_LEQ_EINVAL;
    Dims d;
    d.imax = imax; d.jmax = jmax; d.kmax = kmax;
    d.ni = imax + 2; d.nj = jmax + 2; d.nk = kmax + 2;
    d.n3 = d.ni * d.nj * d.nk;
    switch (kind) {
    case 0: gen_pressure(d, seed, nullptr, 1.0, a_m, b_m); return 0;
    case 1: gen_momentum(d, seed, a_m, b_m); return 0;
    case 2: {
        std::vector<float> eps = void_fraction(d, seed, nspheres < 0 ? 2000 : nspheres);
        gen_pressure(d, seed, &eps, 1.0e-3, a_m, b_m);
        return 0;
    }
    default: return MFIX_LEQ_EINVAL;
    }
}

extern "C" int mfix_leq_gen_uniform(int64_t n, uint64_t seed, uint64_t stream, double *out) {
    if (n < 0 || (!out && n > 0)) return MFIX_LEQ_EINVAL;
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < n; ++q) out[q] = urand(seed, 100 + stream, (uint64_t)q);
    return 0;
}
