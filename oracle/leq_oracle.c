/*
 * leq_oracle.c - CPU restatement of MFIX's LEQ_BICGS path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * solver: only tests, the smoke check and the CPU-baseline leg of bench.py
 * load it.  The product library never links or calls it.
 *
 * It follows the reference's serial semantics (numPEs=1, RE_INDEXING off):
 *   layout/increments   model/include/functions.inc:37-42,
 *                       model/dmp_modules/gridmap_mod.f:259-351,511-556,
 *                       model/set_increments.f:72-112,173-178
 *   unit scaling        model/leq_bicgs.f:194-233
 *   LEQ_BICGS0          model/leq_bicgs.f:107-509
 *   LEQ_MATVEC          model/leqsol_mod.f:103-265 (summation order :195-204)
 *   LEQ_MSOLVE (LINE)   model/leqsol_mod.f:286-598, sweeps :767-1086
 *   LEQ_MSOLVE0 (NONE)  model/leqsol_mod.f:616-663
 *   LEQ_MSOLVE1 (DIAG)  model/leqsol_mod.f:681-749
 *   DGTSV (NRHS=1)      model/DGTSV.f:107-163,233-246
 *   dot_product_par(2)  model/leqsol_mod.f:1094-1295
 * Pinning (see tests/test_oracle_*.py): golden vectors written by the
 * compiled reference (tests/golden/), scipy's LAPACK dgtsv, brute-force
 * numpy stencils, a sparse direct solve, and closed-form special cases.
 *
 * Threading follows the reference's OpenMP structure, so the timed CPU
 * baseline is the reference's parallel algorithm and not a serial one:
 * LEQ_MATVEC is a parallel loop over cells (leqsol_mod.f:188-189), the row
 * scaling a parallel loop over rows (leq_bicgs.f:197-226), the vector
 * updates parallel workshares (leq_bicgs.f:285-436), DIAG a parallel loop
 * (leqsol_mod.f:730), and each red-black pass of LINE a parallel loop over
 * the lines of one colour (leqsol_mod.f:497-522, 2-D :383-391).  The dot
 * products stay single-threaded in storage order: with IS_SERIAL (the
 * default) the reference calls the serial intrinsic dot_product
 * (leqsol_mod.f:1125-1136).  Every parallel loop is element-wise or over
 * independent lines, so the results are the same bits at any thread count.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Threads of the parallel loops (0 or less: leave the OpenMP default);
 * returns the count in effect before the call. */
int oracle_set_threads(int n) {
#ifdef _OPENMP
    int prev = omp_get_max_threads();
    if (n > 0) omp_set_num_threads(n);
    return prev;
#else
    (void)n;
    return 1;
#endif
}

static double og_now(void) {
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

typedef struct {
    int imax, jmax, kmax, no_k;
    int cyclic_x, cyclic_y, cyclic_z;
    int is_serial, minimize_dotproducts, icheck_bicgs;
} og_grid;

typedef struct {
    int ni, nj, nk;
    int64_t c1, c2, n3;
    int do_k;
    const og_grid *g;
} og_dims;

static og_dims dims_of(const og_grid *g) {
    og_dims d;
    d.g = g;
    d.ni = g->imax + 2;
    d.nj = g->jmax + 2;
    d.nk = g->no_k ? 1 : g->kmax + 2;
    d.do_k = !g->no_k;
    d.c1 = d.nj;
    d.c2 = (int64_t)d.nj * d.ni;
    d.n3 = d.c2 * d.nk;
    return d;
}

/* 0-based storage index of 1-based (i,j,k): functions.inc:41 with
 * c0 = -c1 - c2 in the serial layout (gridmap_mod.f:553-556). */
static inline int64_t IJK(const og_dims *d, int i, int j, int k) {
    return (int64_t)(j - 1) + (int64_t)(i - 1) * d->c1 + (int64_t)(k - 1) * d->c2;
}

/* Neighbour index along one axis: set_increments.f:72-112.  n = number of
 * cells on the axis including the two ghosts (IMIN3=1 .. IMAX3=n). */
static inline int nb_minus(int i, int n, int cyclic) {
    int shift = !(i == 1 || i == n);
    if (cyclic && shift) {
        /* IMAP_C(IMAP_C(i)-1): i-1 == 1 (ghost IMIN2) maps to IMAX1 = n-1 */
        int t = i - 1;
        return (t == 1) ? n - 1 : t;
    }
    return i - 1 < 1 ? 1 : i - 1;
}
static inline int nb_plus(int i, int n, int cyclic) {
    int shift = !(i == 1 || i == n);
    if (cyclic && shift) {
        int t = i + 1;
        return (t == n) ? 2 : t;
    }
    return i + 1 > n ? n : i + 1;
}

/* INCREMENT_FOR_MP(1..6) targets of cell (i,j,k): IM, IP, JM, JP, KM, KP. */
static void neighbours(const og_dims *d, int i, int j, int k, int64_t nb[6]) {
    const og_grid *g = d->g;
    nb[0] = IJK(d, nb_minus(i, d->ni, g->cyclic_x), j, k);
    nb[1] = IJK(d, nb_plus(i, d->ni, g->cyclic_x), j, k);
    nb[2] = IJK(d, i, nb_minus(j, d->nj, g->cyclic_y), k);
    nb[3] = IJK(d, i, nb_plus(j, d->nj, g->cyclic_y), k);
    if (d->do_k) {
        nb[4] = IJK(d, i, j, nb_minus(k, d->nk, g->cyclic_z));
        nb[5] = IJK(d, i, j, nb_plus(k, d->nk, g->cyclic_z));
    } else {
        nb[4] = nb[5] = IJK(d, i, j, k);
    }
}

#define A(l, c) a[((int64_t)((l) + 3)) * n3 + (c)]

/* LEQ_MATVEC leqsol_mod.f:103-265, non-RE_INDEXING branch.  In the serial
 * layout istart..iend etc. cover every cell (gridmap_mod.f:511-532).  The
 * core-cell loop (:186-209) uses the first core cell's class, which equals
 * each core cell's own increments, so one per-cell loop reproduces both. */
void oracle_matvec(const og_grid *g, const double *a, const double *x, double *y) {
    og_dims d = dims_of(g);
    const int64_t n3 = d.n3;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= d.nk; ++k)
        for (int i = 1; i <= d.ni; ++i)
            for (int j = 1; j <= d.nj; ++j) {
                int64_t c = IJK(&d, i, j, k), nb[6];
                neighbours(&d, i, j, k, nb);
                double v = A(-2, c) * x[nb[2]] + A(-1, c) * x[nb[0]] + A(0, c) * x[c]
                         + A(1, c) * x[nb[1]] + A(2, c) * x[nb[3]];
                if (d.do_k) {
                    v = v + A(-3, c) * x[nb[4]];
                    v = v + A(3, c) * x[nb[5]];
                }
                y[c] = v;
            }
}

/* DGTSV with NRHS=1, DGTSV.f:107-163 (elimination with partial pivoting)
 * and :233-246 (back solve).  dl has n-1 entries, d and du n, b n.
 * Returns INFO (0 or the 1-based index of a zero pivot). */
int oracle_dgtsv(int n, double *dl, double *d, double *du, double *b) {
    if (n == 0) return 0;
    for (int i = 0; i < n - 2; ++i) {
        if (fabs(d[i]) >= fabs(dl[i])) {
            if (d[i] != 0.0) {
                double fact = dl[i] / d[i];
                d[i + 1] = d[i + 1] - fact * du[i];
                b[i + 1] = b[i + 1] - fact * b[i];
            } else {
                return i + 1;
            }
            dl[i] = 0.0;
        } else {
            double fact = d[i] / dl[i];
            d[i] = dl[i];
            double temp = d[i + 1];
            d[i + 1] = du[i] - fact * temp;
            dl[i] = du[i + 1];
            du[i + 1] = -fact * dl[i];
            du[i] = temp;
            temp = b[i];
            b[i] = b[i + 1];
            b[i + 1] = temp - fact * b[i + 1];
        }
    }
    if (n > 1) {
        int i = n - 2;
        if (fabs(d[i]) >= fabs(dl[i])) {
            if (d[i] != 0.0) {
                double fact = dl[i] / d[i];
                d[i + 1] = d[i + 1] - fact * du[i];
                b[i + 1] = b[i + 1] - fact * b[i];
            } else {
                return i + 1;
            }
        } else {
            double fact = d[i] / dl[i];
            d[i] = dl[i];
            double temp = d[i + 1];
            d[i + 1] = du[i] - fact * temp;
            du[i] = temp;
            temp = b[i];
            b[i] = b[i + 1];
            b[i + 1] = temp - fact * b[i + 1];
        }
    }
    if (d[n - 1] == 0.0) return n;
    b[n - 1] = b[n - 1] / d[n - 1];
    if (n > 1) b[n - 2] = (b[n - 2] - du[n - 2] * b[n - 1]) / d[n - 2];
    for (int i = n - 3; i >= 0; --i)
        b[i] = (b[i] - du[i] * b[i + 1] - dl[i] * b[i + 2]) / d[i];
    return 0;
}

/* Line along J at (i,k): LEQ_IKSWEEP leqsol_mod.f:851-921 (3-D) and
 * LEQ_ISWEEP :767-834 (2-D, k=1, no K terms).  On DGTSV failure the line
 * keeps its current values (:910-914).  Returns INFO. */
static int sweep_j_line(const og_dims *d, int i, int k, const double *a, const double *bm, double *var,
                        double *cc, double *dd, double *ee, double *bb) {
    const int64_t n3 = d->n3;
    const int n = d->nj;
    for (int j = 1; j <= n; ++j) {
        int64_t c = IJK(d, i, j, k), nb[6];
        neighbours(d, i, j, k, nb);
        dd[j - 1] = A(0, c);
        cc[j - 1] = A(-2, c);
        ee[j - 1] = A(2, c);
        if (d->do_k)
            bb[j - 1] = bm[c] - A(-1, c) * var[nb[0]] - A(1, c) * var[nb[1]] - A(-3, c) * var[nb[4]]
                      - A(3, c) * var[nb[5]];
        else
            bb[j - 1] = bm[c] - A(-1, c) * var[nb[0]] - A(1, c) * var[nb[1]];
    }
    cc[0] = 0.0;
    ee[n - 1] = 0.0;
    int info = oracle_dgtsv(n, cc + 1, dd, ee, bb);
    if (info != 0) return info;
    for (int j = 1; j <= n; ++j) var[IJK(d, i, j, k)] = bb[j - 1];
    return 0;
}

/* Line along I at (j,k): LEQ_JKSWEEP leqsol_mod.f:939-1004 (failure is
 * fatal in the reference: returns INFO to the caller). */
static int sweep_i_line(const og_dims *d, int j, int k, const double *a, const double *bm, double *var,
                        double *cc, double *dd, double *ee, double *bb) {
    const int64_t n3 = d->n3;
    const int n = d->ni;
    for (int i = 1; i <= n; ++i) {
        int64_t c = IJK(d, i, j, k), nb[6];
        neighbours(d, i, j, k, nb);
        dd[i - 1] = A(0, c);
        cc[i - 1] = A(-1, c);
        ee[i - 1] = A(1, c);
        bb[i - 1] = bm[c] - A(-2, c) * var[nb[2]] - A(2, c) * var[nb[3]] - A(-3, c) * var[nb[4]]
                  - A(3, c) * var[nb[5]];
    }
    cc[0] = 0.0;
    ee[n - 1] = 0.0;
    int info = oracle_dgtsv(n, cc + 1, dd, ee, bb);
    if (info != 0) return info;
    for (int i = 1; i <= n; ++i) var[IJK(d, i, j, k)] = bb[i - 1];
    return 0;
}

/* Line along K at (i,j): LEQ_IJSWEEP leqsol_mod.f:1021-1086 (fatal on failure). */
static int sweep_k_line(const og_dims *d, int i, int j, const double *a, const double *bm, double *var,
                        double *cc, double *dd, double *ee, double *bb) {
    const int64_t n3 = d->n3;
    const int n = d->nk;
    for (int k = 1; k <= n; ++k) {
        int64_t c = IJK(d, i, j, k), nb[6];
        neighbours(d, i, j, k, nb);
        dd[k - 1] = A(0, c);
        cc[k - 1] = A(-3, c);
        ee[k - 1] = A(3, c);
        bb[k - 1] = bm[c] - A(-2, c) * var[nb[2]] - A(2, c) * var[nb[3]] - A(-1, c) * var[nb[0]]
                  - A(1, c) * var[nb[1]];
    }
    cc[0] = 0.0;
    ee[n - 1] = 0.0;
    int info = oracle_dgtsv(n, cc + 1, dd, ee, bb);
    if (info != 0) return info;
    for (int k = 1; k <= n; ++k) var[IJK(d, i, j, k)] = bb[k - 1];
    return 0;
}

/* Position p (1-based) of the reference's red-black enumeration of a
 * 2-D index set of size s1 (fast) x s2 (slow), leqsol_mod.f:404-418:
 * returns the (fast, slow) 1-based offsets added to the range starts. */
static void rb_position(int p, int s1, int lo2, int *f, int *s) {
    int slow;
    if (p % s1 != 0) slow = p / s1 + lo2;
    else slow = p / s1 + lo2 - 1;
    int fast = (p - 1 - (slow - lo2) * s1) + 1 + (slow % 2); /* + j1 with j1 = 1 */
    if (fast > s1) fast = fast - s1 + 1 - 1;                 /* j - j2 + j1 - 1 with j1=1, j2=s1 */
    *f = fast;
    *s = slow;
}

/* One red-black pass of J lines.  3-D (leqsol_mod.f:497-522): first = 1
 * takes k odd -> i = 2, 4, ..; k even -> i = 1, 3, .. (1-based), first = 0
 * the other colour.  2-D (:383-391): first = 1 takes i = 1, 3, .., first = 0
 * i = 2, 4, ..  The lines of one pass read only lines of the other colour,
 * so they are independent: the pass is a parallel loop, as in the
 * reference, with private line buffers.  DGTSV failures leave the line
 * unchanged (:910-914). */
static void rb_pass(const og_dims *d, int three_d, int first, const double *a, const double *bm, double *var,
                    int nmax) {
#pragma omp parallel
    {
        double *buf = malloc(sizeof(double) * 4 * (size_t)nmax);
        if (three_d) {
#pragma omp for schedule(static)
            for (int k = 1; k <= d->nk; ++k) {
                const int i0 = ((k % 2 != 0) == (first != 0)) ? 2 : 1;
                for (int i = i0; i <= d->ni; i += 2)
                    sweep_j_line(d, i, k, a, bm, var, buf, buf + nmax, buf + 2 * nmax, buf + 3 * nmax);
            }
        } else {
#pragma omp for schedule(static)
            for (int i = first ? 1 : 2; i <= d->ni; i += 2)
                sweep_j_line(d, i, 1, a, bm, var, buf, buf + nmax, buf + 2 * nmax, buf + 3 * nmax);
        }
        free(buf);
    }
}

/* LEQ_MSOLVE leqsol_mod.f:286-598.  var = b (SETGUESS, :341-357), then one
 * pass per character of the 4-character CMETHOD.  Returns 0, or a positive
 * DGTSV INFO from a JK/IJ sweep, which the reference treats as fatal. */
int oracle_msolve_line(const og_grid *g, const double *bm, const double *a, double *var, const char *cmethod) {
    og_dims d = dims_of(g);
    int nmax = d.ni > d.nj ? d.ni : d.nj;
    if (d.nk > nmax) nmax = d.nk;
    double *cc = malloc(sizeof(double) * nmax), *dd = malloc(sizeof(double) * nmax);
    double *ee = malloc(sizeof(double) * nmax), *bb = malloc(sizeof(double) * nmax);
    int fatal = 0;
    memcpy(var, bm, sizeof(double) * d.n3);
    for (int it = 0; it < 4 && !fatal; ++it) {
        char ch = cmethod[it];
        int do_i = (ch == 'I' || ch == 'i'), do_k = (ch == 'K' || ch == 'k');
        int do_all = (ch == 'A' || ch == 'a'), do_rb = (ch == 'R' || ch == 'r');
        if (!d.do_k) {
            /* two-dimensional (:367-392) */
            if (do_i)
                for (int i = 1; i <= d.ni; ++i) sweep_j_line(&d, i, 1, a, bm, var, cc, dd, ee, bb);
            if (do_rb) {
                rb_pass(&d, 0, 1, a, bm, var, nmax);
                rb_pass(&d, 0, 0, a, bm, var, nmax);
            }
            continue;
        }
        if (do_all) {
            /* JK lines (along I), :399-421 */
            int jsize = d.nj, ksize = d.nk;
            for (int icase = 1; icase <= 2 && !fatal; ++icase)
                for (int p = icase; p <= ksize * jsize; p += 2) {
                    int j, k;
                    rb_position(p, jsize, 1, &j, &k);
                    int info = sweep_i_line(&d, j, k, a, bm, var, cc, dd, ee, bb);
                    if (info) { fatal = info; break; }
                }
            /* IJ lines (along K), :425-446 */
            int isize = d.ni;
            jsize = d.nj;
            for (int icase = 1; icase <= 2 && !fatal; ++icase)
                for (int p = icase; p <= jsize * isize; p += 2) {
                    int i, j;
                    rb_position(p, isize, 1, &i, &j);
                    int info = sweep_k_line(&d, i, j, a, bm, var, cc, dd, ee, bb);
                    if (info) { fatal = info; break; }
                }
            /* IK lines (along J), :450-469 */
            isize = d.ni;
            ksize = d.nk;
            for (int icase = 1; icase <= 2 && !fatal; ++icase)
                for (int p = icase; p <= ksize * isize; p += 2) {
                    int i, k;
                    rb_position(p, isize, 1, &i, &k);
                    sweep_j_line(&d, i, k, a, bm, var, cc, dd, ee, bb);
                }
        }
        if (do_rb && !fatal) {
            /* :497-522: k odd -> i = istart+1, +3..; k even -> i = istart, +2.. */
            rb_pass(&d, 1, 1, a, bm, var, nmax);
            rb_pass(&d, 1, 0, a, bm, var, nmax);
        }
        if (!fatal) {
            /* USE_IKLOOP = .FALSE. branch, :570-585 */
            if (do_i)
                for (int k = 1; k <= d.nk; ++k)
                    for (int i = 1; i <= d.ni; ++i) sweep_j_line(&d, i, k, a, bm, var, cc, dd, ee, bb);
            if (do_k)
                for (int i = 1; i <= d.ni; ++i)
                    for (int k = 1; k <= d.nk; ++k) sweep_j_line(&d, i, k, a, bm, var, cc, dd, ee, bb);
        }
    }
    free(cc); free(dd); free(ee); free(bb);
    return fatal;
}

/* LEQ_MSOLVE1 leqsol_mod.f:681-749: var = 0, then b/A0 over *start2..*end2
 * (every cell in the serial layout). */
void oracle_msolve_diag(const og_grid *g, const double *bm, const double *a, double *var) {
    og_dims d = dims_of(g);
    const int64_t n3 = d.n3;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n3; ++c) var[c] = bm[c] / A(0, c);
}

/* LEQ_MSOLVE0 leqsol_mod.f:616-663. */
void oracle_msolve_none(const og_grid *g, const double *bm, double *var) {
    og_dims d = dims_of(g);
    memcpy(var, bm, sizeof(double) * d.n3);
}

/* dot_product_par leqsol_mod.f:1094-1201.  Serial default: the intrinsic
 * over all DIMENSION_3 cells in storage order (:1125-1136).  Otherwise the
 * interior cells *start1..*end1 in (k,i,j) loop order (:1152-1162). */
double oracle_dot(const og_grid *g, const double *x, const double *y) {
    og_dims d = dims_of(g);
    double s = 0.0;
    if (g->is_serial) {
        for (int64_t c = 0; c < d.n3; ++c) s += x[c] * y[c];
        return s;
    }
    int k0 = d.do_k ? 2 : 1, k1 = d.do_k ? d.nk - 1 : 1;
    for (int k = k0; k <= k1; ++k)
        for (int i = 2; i <= d.ni - 1; ++i)
            for (int j = 2; j <= d.nj - 1; ++j) {
                int64_t c = IJK(&d, i, j, k);
                s += x[c] * y[c];
            }
    return s;
}

/* dot_product_par2 leqsol_mod.f:1210-1295 (interior cells, no dead cells
 * without a cut-cell grid). */
void oracle_dot2(const og_grid *g, const double *r1, const double *r2, const double *r3, const double *r4,
                 double out[2]) {
    og_dims d = dims_of(g);
    double p0 = 0.0, p1 = 0.0;
    int k0 = d.do_k ? 2 : 1, k1 = d.do_k ? d.nk - 1 : 1;
    for (int k = k0; k <= k1; ++k)
        for (int i = 2; i <= d.ni - 1; ++i)
            for (int j = 2; j <= d.nj - 1; ++j) {
                int64_t c = IJK(&d, i, j, k);
                p0 += r1[c] * r2[c];
                p1 += r3[c] * r4[c];
            }
    out[0] = p0;
    out[1] = p1;
}

/* Row max-norm unit scaling, leq_bicgs.f:194-233 (loop over *start2..*end2,
 * every cell in serial).  Returns 0 or -2 on a zero row. */
int oracle_unit_scale(const og_grid *g, double *a, double *bm) {
    og_dims d = dims_of(g);
    const int64_t n3 = d.n3;
    int zero = 0;
#pragma omp parallel for schedule(static) reduction(| : zero)
    for (int64_t c = 0; c < n3; ++c) {
        double amax = 0.0;
        for (int l = -3; l <= 3; ++l) {
            double v = fabs(A(l, c));
            if (v > amax) amax = v;
        }
        if (amax > 0.0) {
            double oam = 1.0 / amax;
            for (int l = -3; l <= 3; ++l) A(l, c) = A(l, c) * oam;
            bm[c] = bm[c] * oam;
        } else {
            zero = 1;
        }
    }
    return zero ? -2 : 0;
}

typedef struct {
    int iterations, iter_tot_increment, ier, exit_reason;
    double rnorm0, rnorm;
    int msolve_fatal;
    double seconds_setup;   /* wall time from the call to the end of the set-up (scaling, R, Rtilde) */
    double seconds_total;   /* wall time of the whole call */
} og_result;

/* LEQ_BICGS (PC dispatch, leq_bicgs.f:74-87) + LEQ_BICGS0 (:107-509).
 * noise: the U[0,1) array random_number would fill Rtilde with (:298).
 * rnorm_trace (optional, itmax+1 entries) receives Rnorm0 then each Rnorm;
 * time_trace (optional, itmax+1 entries) the wall time since the call at the
 * end of the set-up, then at the end of each iteration.
 * Returns 0, or -1 for an unknown PC (the reference aborts). */
int oracle_leq_bicgs(const og_grid *g, const char *pc, const char *cmethod, double tol, int itmax,
                     double *var, double *a, double *bm, const double *noise, og_result *res,
                     double *rnorm_trace, double *time_trace) {
    const double t_call = og_now();
    int mode;
    if (strncmp(pc, "LINE", 4) == 0) mode = 2;
    else if (strncmp(pc, "DIAG", 4) == 0) mode = 1;
    else if (strncmp(pc, "NONE", 4) == 0) mode = 0;
    else return -1;
    const int use_pc = mode != 0;
    og_dims d = dims_of(g);
    const int64_t n = d.n3;
    memset(res, 0, sizeof(*res));
    res->ier = oracle_unit_scale(g, a, bm);
    if (res->ier != 0) {
        res->exit_reason = 4;
        res->seconds_total = og_now() - t_call;
        return 0;
    }

    double *R = calloc(n, 8), *Rt = calloc(n, 8), *P = calloc(n, 8), *Pp = calloc(n, 8);
    double *S = calloc(n, 8), *Sp = calloc(n, 8), *T = calloc(n, 8), *V = calloc(n, 8);
    double *alpha = calloc(itmax + 2, 8), *omega = calloc(itmax + 2, 8), *rho = calloc(itmax + 2, 8);
    const double tolmin = 2.220446049250313e-16; /* EPSILON(1d0) */
    const double ratiotol = (double)0.2f;          /* DOUBLE PRECISION :: ratiotol = 0.2 (default-real literal) */
    double rnorm = 0.0, rnorm0, snorm;
    int isconverged = 0, iter = 1, done = 0;

    oracle_matvec(g, a, var, R);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n; ++c) R[c] = bm[c] - R[c];
    rnorm0 = sqrt(oracle_dot(g, R, R));
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n; ++c) Rt[c] = R[c] + (2.0 * noise[c] - 1.0) * 1.0e-6 * rnorm0;
    if (rnorm_trace) rnorm_trace[0] = rnorm0;
    res->seconds_setup = og_now() - t_call;
    if (time_trace) time_trace[0] = res->seconds_setup;

    double *Phat = P, *Shat = S;
    int i;
    for (i = 1; i <= itmax; ++i) {
        rho[i - 1] = oracle_dot(g, Rt, R);
        if (rho[i - 1] == 0.0) {
            res->ier = (i != 1) ? -2 : 0;
            res->exit_reason = 2;
            res->iterations = i - 1;
            done = 1;
            break;
        }
        if (i == 1) {
#pragma omp parallel for schedule(static)
            for (int64_t c = 0; c < n; ++c) P[c] = R[c];
        } else {
            double beta = (rho[i - 1] / rho[i - 2]) * (alpha[i - 1] / omega[i - 1]);
#pragma omp parallel for schedule(static)
            for (int64_t c = 0; c < n; ++c) P[c] = R[c] + beta * (P[c] - omega[i - 1] * V[c]);
        }
        if (use_pc) {
            if (mode == 2) res->msolve_fatal |= oracle_msolve_line(g, P, a, Pp, cmethod);
            else oracle_msolve_diag(g, P, a, Pp);
            Phat = Pp;
        } else {
            Phat = P;
        }
        oracle_matvec(g, a, Phat, V);
        double rtv = oracle_dot(g, Rt, V);
        alpha[i] = rho[i - 1] / rtv;
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < n; ++c) S[c] = R[c] - alpha[i] * V[c];
        if (!g->minimize_dotproducts) {
            snorm = sqrt(oracle_dot(g, S, S));
            if (snorm <= tolmin) {
#pragma omp parallel for schedule(static)
                for (int64_t c = 0; c < n; ++c) var[c] = var[c] + alpha[i] * Phat[c];
                isconverged = 1;
                res->exit_reason = 1;
                res->iterations = i;
                break;
            }
        }
        if (use_pc) {
            if (mode == 2) res->msolve_fatal |= oracle_msolve_line(g, S, a, Sp, cmethod);
            else oracle_msolve_diag(g, S, a, Sp);
            Shat = Sp;
        } else {
            Shat = S;
        }
        oracle_matvec(g, a, Shat, T);
        double txs, txt;
        if (!g->minimize_dotproducts) {
            txs = oracle_dot(g, T, S);
            txt = oracle_dot(g, T, T);
        } else {
            double two[2];
            oracle_dot2(g, T, S, T, T, two);
            txs = two[0];
            txt = two[1];
        }
        if (txt == 0.0) txt = 1.0e-15; /* SMALL_NUMBER param1_mod.f:24 */
        omega[i] = txs / txt;
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < n; ++c) var[c] = var[c] + alpha[i] * Phat[c] + omega[i] * Shat[c];
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < n; ++c) R[c] = S[c] - omega[i] * T[c];
        int icheck = g->icheck_bicgs > 0 ? g->icheck_bicgs : 1;
        if (!g->minimize_dotproducts || (iter % icheck) == 0) {
            rnorm = sqrt(oracle_dot(g, R, R));
            if (rnorm_trace) rnorm_trace[i] = rnorm;
            isconverged = (rnorm <= tol * rnorm0);
            if (isconverged) {
                res->iter_tot_increment = iter + 1;
                res->exit_reason = 0;
                res->iterations = i;
                if (time_trace) time_trace[i] = og_now() - t_call;
                break;
            }
        }
        if (time_trace) time_trace[i] = og_now() - t_call;
        iter = iter + 1;
    }
    if (!done) {
        if (i > itmax) { res->iterations = itmax; res->exit_reason = 3; }
        if (!isconverged) isconverged = ((double)(float)rnorm <= tol * rnorm0);
        res->ier = 0;
        if (!isconverged) {
            res->ier = -1;
            res->iter_tot_increment = iter;
            if ((double)(float)rnorm >= ratiotol * (double)(float)rnorm0) res->ier = -2;
        }
    }
    res->rnorm0 = rnorm0;
    res->rnorm = rnorm;
    res->seconds_total = og_now() - t_call;
    free(R); free(Rt); free(P); free(Pp); free(S); free(Sp); free(T); free(V);
    free(alpha); free(omega); free(rho);
    return 0;
}

/* ---- residual norms: CALC_RESID_C / _S / _PP (model/calc_resid.f) --------- */

typedef struct {
    double num, den, resid, max_resid;
    int ijk_resid, ncells;
} og_resid;

#define OG_UNDEFINED 9.87654321e31   /* UNDEFINED, model/param1_mod.f:18 */
#define OG_LARGE_NUMBER 1.0e32       /* LARGE_NUMBER, model/param1_mod.f:23 */

/* FLAG predicates, model/include/functions.inc:409-414 (FLUID_AT),
 * :461-466 (FLOW_AT), :469-474 (WALL_AT). */
static inline int og_fluid(int f) { return f == 1; }
static inline int og_wall(int f) { return f >= 100; }
static inline int og_flow(int f) { return f >= 10 && f <= 31; }

/* Neighbours of cell (i,j,k) as SET_INDEX1A builds them
 * (model/set_index1a.f:85-222): mp = IM, IP, JM, JP, KM, KP with the
 * true-corner rule (a wall or flow cell on the first/last interior layer is
 * its own neighbour across that face); nb = E, W, S, N, B, T, the mp
 * neighbour unless that one is a wall cell, then the cell itself. */
static void og_index1a(const og_dims *d, const int *flag, int i, int j, int k, int64_t mp[6], int64_t nb[6]) {
    const og_grid *g = d->g;
    const int64_t c = IJK(d, i, j, k);
    const int wf = og_wall(flag[c]) || og_flow(flag[c]);
    const int kmin1 = d->do_k ? 2 : 1, kmax1 = d->do_k ? d->nk - 1 : 1;
    mp[0] = (wf && i == 2) ? c : IJK(d, nb_minus(i, d->ni, g->cyclic_x), j, k);
    mp[1] = (wf && i == d->ni - 1) ? c : IJK(d, nb_plus(i, d->ni, g->cyclic_x), j, k);
    mp[2] = (wf && j == 2) ? c : IJK(d, i, nb_minus(j, d->nj, g->cyclic_y), k);
    mp[3] = (wf && j == d->nj - 1) ? c : IJK(d, i, nb_plus(j, d->nj, g->cyclic_y), k);
    if (d->do_k) {
        mp[4] = (wf && k == kmin1) ? c : IJK(d, i, j, nb_minus(k, d->nk, g->cyclic_z));
        mp[5] = (wf && k == kmax1) ? c : IJK(d, i, j, nb_plus(k, d->nk, g->cyclic_z));
    } else {
        mp[4] = mp[5] = c;
    }
    /* E <- IP, W <- IM, S <- JM, N <- JP, B <- KM, T <- KP */
    const int src[6] = {1, 0, 2, 3, 4, 5};
    for (int q = 0; q < 6; ++q) nb[q] = og_wall(flag[mp[src[q]]]) ? c : mp[src[q]];
}

/* kind 0: CALC_RESID_C (model/calc_resid.f:16-178), kind 1: CALC_RESID_S
 * (:195-371; tol = TOL, ep_s = EP_S(:,M) or NULL for M = 0, dil_ep_s =
 * DIL_EP_s), kind 2: CALC_RESID_PP (:391-534, b only, norm = NORM).  Serial
 * run: the loop over IJK is the storage order, IS_ON_myPE_wobnd
 * (functions.inc:121-133) keeps the cells of the first interior layers
 * ISTART1..IEND1 etc. (set_max2.f:53-69), and the DMP gathers reduce to the
 * local values.  debug = DEBUG_RESID (model/run_mod.f:244).  With debug off,
 * C and S return NUM and DEN only: RESID and IJK_RESID are left as 0 and
 * MAX_RESID as -1 (the reference leaves RESID and IJK_RESID unset). */
void oracle_calc_resid(const og_grid *g, int kind, const double *var, const double *a, const double *b,
                       const int *flag, double tol, const double *ep_s, double dil_ep_s, double norm, int debug,
                       og_resid *out) {
    og_dims d = dims_of(g);
    const int64_t n3 = d.n3;
    double num = 0.0, den = 0.0, maxr = -1.0;
    int ncells = 0;
    int64_t ijkr = 1;                 /* 1-based IJK_RESID */
    double *rijk = kind != 2 ? (double *)calloc((size_t)n3, sizeof(double)) : NULL;
    const int kmin1 = d.do_k ? 2 : 1, kmax1 = d.do_k ? d.nk - 1 : 1;
    for (int k = 1; k <= d.nk; ++k)
        for (int i = 1; i <= d.ni; ++i)
            for (int j = 1; j <= d.nj; ++j) {
                const int64_t c = IJK(&d, i, j, k);
                if (!(i >= 2 && i <= d.ni - 1 && j >= 2 && j <= d.nj - 1 && k >= kmin1 && k <= kmax1)) continue;
                if (!og_fluid(flag[c])) continue;
                if (kind == 2) {
                    const double num1 = fabs(b[c]);
                    if (num1 > maxr) { maxr = num1; ijkr = c + 1; }
                    ncells += 1;
                    num = num + num1;
                    den = den + 1.0;
                    continue;
                }
                if (kind == 1) {
                    if (!(fabs(var[c]) > tol)) continue;
                    if (ep_s && ep_s[c] <= dil_ep_s) continue;
                }
                int64_t mp[6], nb[6];
                og_index1a(&d, flag, i, j, k, mp, nb);
                int64_t e, w, s, n, bt, t;
                if (kind == 0) { e = nb[0]; w = nb[1]; s = nb[2]; n = nb[3]; bt = nb[4]; t = nb[5]; }
                else { w = mp[0]; e = mp[1]; s = mp[2]; n = mp[3]; bt = mp[4]; t = mp[5]; }
                double num1 = b[c] - (A(0, c) * var[c] + A(1, c) * var[e] + A(-1, c) * var[w] +
                                      A(2, c) * var[n] + A(-2, c) * var[s]);
                if (d.do_k) num1 = num1 - (A(3, c) * var[t] + A(-3, c) * var[bt]);
                num1 = fabs(num1);
                const double den1 = fabs(A(0, c) * var[c]);
                rijk[c] = num1;
                ncells += 1;
                num = num + num1;
                den = den + den1;
            }
    out->num = num;
    out->den = den;
    out->ncells = ncells;
    if (kind == 2) {
        double r, m = maxr;
        int64_t ij = ijkr;
        if (den * norm > 0.0) {
            r = num / (den * norm);
            if (debug) m = ncells * maxr / (den * norm);
        } else if (num == 0.0) {
            r = 0.0;
            if (debug) { m = 0.0; ij = 0; }
        } else {
            r = OG_LARGE_NUMBER;
            if (debug) m = OG_LARGE_NUMBER;
        }
        out->resid = r;
        out->max_resid = m;
        out->ijk_resid = (int)ij;
        return;
    }
    if (!debug) {
        out->resid = 0.0;
        out->max_resid = -1.0;
        out->ijk_resid = 0;
        free(rijk);
        return;
    }
    /* first cell of the largest residual, from RESID_IJK(1) (:114-122) */
    ijkr = 1;
    maxr = rijk[0];
    for (int k = 1; k <= d.nk; ++k)
        for (int i = 1; i <= d.ni; ++i)
            for (int j = 1; j <= d.nj; ++j) {
                const int64_t c = IJK(&d, i, j, k);
                if (!(i >= 2 && i <= d.ni - 1 && j >= 2 && j <= d.nj - 1 && k >= kmin1 && k <= kmax1)) continue;
                if (rijk[c] > maxr) { ijkr = c + 1; maxr = rijk[c]; }
            }
    free(rijk);
    if (den > 0.0) {
        out->resid = num / den;
        out->max_resid = ncells * maxr / den;
        out->ijk_resid = (int)ijkr;
    } else if (num == 0.0) {
        out->resid = 0.0;
        out->max_resid = 0.0;
        out->ijk_resid = 0;
    } else {
        out->resid = OG_UNDEFINED;
        out->max_resid = OG_UNDEFINED;
        out->ijk_resid = (int)ijkr;
    }
}

/* ---- pressure-correction assembly: INIT_AB_M + CONV_PP_G + SOURCE_PP_G ----
 * Gas phase only (MMAX = 0), incompressible (RO_G0 defined: no
 * COMPRESSIBLE_PP_G), no cut cells, no shear, no point or user sources.
 *   INIT_AB_M    model/init_ab_m.f:20-78
 *   CONV_PP_G    model/conv_pp_g.f:33-182 (gas loop :69-113), in its own
 *                scatter form: a fluid cell writes its east/north/top face
 *                coefficient into its own row and into the neighbour's
 *                west/south/bottom entry
 *   SOURCE_PP_G  model/source_pp_g.f:18-233 (gas terms :77-178, pressure
 *                flow neighbours :193-212, reference cell :214-218)
 * Fields are DIMENSION_3 arrays; ijk_p_g is 1-based (0 = UNDEFINED_I).
 * Returns the number of fluid cells whose A0 vanished with B != 0 (the
 * reference writes an error message there and carries on). */
typedef struct {
    const double *rop_ge, *rop_gn, *rop_gt, *ayz, *axz, *axy, *vol;
    const double *rop_g, *rop_go, *u_g, *v_g, *w_g, *sum_r_g, *d_e, *d_n, *d_t;
    const int *flag;
    double odt;
    int ijk_p_g;
} og_pp_fields;

static inline int og_pflow(int f) { return f == 10 || f == 11; }   /* P_FLOW_AT, functions.inc:417-423 */

int oracle_assemble_pp_g(const og_grid *g, const og_pp_fields *F, double *a, double *b, double *bmmax) {
    og_dims d = dims_of(g);
    const int64_t n3 = d.n3;
    const int *flag = F->flag;
    int bad = 0;
    /* INIT_AB_M */
    for (int64_t c = 0; c < n3; ++c) {
        for (int l = -3; l <= 3; ++l) A(l, c) = 0.0;
        A(0, c) = -1.0;
        b[c] = 0.0;
    }
    /* CONV_PP_G, gas loop */
    for (int k = 1; k <= d.nk; ++k)
        for (int i = 1; i <= d.ni; ++i)
            for (int j = 1; j <= d.nj; ++j) {
                const int64_t c = IJK(&d, i, j, k);
                if (!og_fluid(flag[c])) continue;
                int64_t mp[6], nb[6];
                og_index1a(&d, flag, i, j, k, mp, nb);
                double am = F->rop_ge[c] * F->ayz[c];
                A(1, c) = am;
                A(-1, mp[1]) = am;
                am = F->rop_gn[c] * F->axz[c];
                A(2, c) = am;
                A(-2, mp[3]) = am;
                if (d.do_k) {
                    am = F->rop_gt[c] * F->axy[c];
                    A(3, c) = am;
                    A(-3, mp[5]) = am;
                }
                if (!og_fluid(flag[mp[0]])) A(-1, c) = F->rop_ge[mp[0]] * F->ayz[mp[0]];
                if (!og_fluid(flag[mp[2]])) A(-2, c) = F->rop_gn[mp[2]] * F->axz[mp[2]];
                if (d.do_k && !og_fluid(flag[mp[4]])) A(-3, c) = F->rop_gt[mp[4]] * F->axy[mp[4]];
            }
    /* SOURCE_PP_G, gas phase */
    for (int k = 1; k <= d.nk; ++k)
        for (int i = 1; i <= d.ni; ++i)
            for (int j = 1; j <= d.nj; ++j) {
                const int64_t c = IJK(&d, i, j, k);
                if (og_fluid(flag[c])) {
                    int64_t mp[6], nb[6];
                    og_index1a(&d, flag, i, j, k, mp, nb);
                    const int64_t im = mp[0], jm = mp[2], km = mp[4];
                    const double bma = (F->rop_g[c] - F->rop_go[c]) * F->vol[c] * F->odt;
                    const double bme = A(1, c) * F->u_g[c];
                    const double bmw = A(-1, c) * F->u_g[im];
                    const double bmn = A(2, c) * F->v_g[c];
                    const double bms = A(-2, c) * F->v_g[jm];
                    const double bmt = A(3, c) * F->w_g[c];
                    const double bmb = A(-3, c) * F->w_g[km];
                    const double bmr = F->sum_r_g[c] * F->vol[c];
                    b[c] = -((-(bma + bme - bmw + bmn - bms + bmt - bmb)) + bmr);
                    double m = fabs(bma);
                    const double rest[7] = {fabs(bme), fabs(bmw), fabs(bmn), fabs(bms), fabs(bmt), fabs(bmb),
                                            fabs(bmr)};
                    for (int q = 0; q < 7; ++q) m = rest[q] > m ? rest[q] : m;
                    bmmax[c] = m;
                    A(1, c) = A(1, c) * F->d_e[c];
                    A(-1, c) = A(-1, c) * F->d_e[im];
                    A(2, c) = A(2, c) * F->d_n[c];
                    A(-2, c) = A(-2, c) * F->d_n[jm];
                    A(3, c) = A(3, c) * F->d_t[c];
                    A(-3, c) = A(-3, c) * F->d_t[km];
                    A(0, c) = -(A(1, c) + A(-1, c) + A(2, c) + A(-2, c) + A(3, c) + A(-3, c));
                    if (fabs(A(0, c)) < 1.0e-15) {   /* SMALL_NUMBER, param1_mod.f:24 */
                        if (fabs(b[c]) < 1.0e-15) {
                            A(0, c) = -1.0;
                            b[c] = 0.0;
                        } else {
                            bad += 1;
                        }
                    }
                } else {
                    for (int l = -3; l <= 3; ++l) A(l, c) = 0.0;
                    A(0, c) = -1.0;
                    b[c] = 0.0;
                }
            }
    for (int k = 1; k <= d.nk; ++k)
        for (int i = 1; i <= d.ni; ++i)
            for (int j = 1; j <= d.nj; ++j) {
                const int64_t c = IJK(&d, i, j, k);
                if (!og_fluid(flag[c])) continue;
                int64_t mp[6], nb[6];
                og_index1a(&d, flag, i, j, k, mp, nb);
                if (og_pflow(flag[mp[0]])) A(-1, c) = 0.0;
                if (og_pflow(flag[mp[1]])) A(1, c) = 0.0;
                if (og_pflow(flag[mp[2]])) A(-2, c) = 0.0;
                if (og_pflow(flag[mp[3]])) A(2, c) = 0.0;
                if (og_pflow(flag[mp[4]])) A(-3, c) = 0.0;
                if (og_pflow(flag[mp[5]])) A(3, c) = 0.0;
            }
    if (F->ijk_p_g > 0) {
        const int64_t c = F->ijk_p_g - 1;
        b[c] = 0.0;
        for (int l = -3; l <= 3; ++l) A(l, c) = 0.0;
        A(0, c) = -1.0;
    }
    return bad;
}

/* ---- scalar transport, first-order upwinding ----
 *   INIT_AB_M            model/init_ab_m.f:20-78
 *   CONV_DIF_PHI         model/conv_dif_phi.f:17-80 with DISC = 0 and DEF_COR
 *                        off: CONV_DIF_PHI0 (:214-339, the reference's own
 *                        scatter order) with GET_PHICELL_DIFF_TERMS (:87-198,
 *                        no cut cells) and DIF_PHI_IS (:722-836), which adds
 *                        nothing without internal surfaces
 *   SOURCE_PHI           model/source_phi.f:25-139 (no point sources)
 *   UNDER_RELAX_S        model/under_relax.f:20-81
 * The 1-D geometry arrays hold Fortran index I = 1..IMAX2 at [I-1]. */
typedef struct {
    const double *dif, *u, *v, *w, *flux_e, *flux_n, *flux_t, *ayz, *axz, *axy;
    const double *odx_e, *ox, *fx_e, *fx_e_bar, *ody_n, *fy_n, *fy_n_bar, *odz_t, *fz_t, *fz_t_bar;
    const int *flag;
} og_phi_fields;

/* AVG_X_H / AVG_Y_H / AVG_Z_H, model/include/fun_avg.inc:93-129 */
static inline double og_avg_h(double m, double p, double f, double fbar) {
    const double den = f * m + fbar * p;
    return m * p / (den > 1.0e-15 ? den : 1.0e-15);   /* MAX(SMALL_NUMBER, .) */
}

/* GET_PHICELL_DIFF_TERMS for fluid cell (i,j,k): D = fe, fw, fn, fs, ft, fb */
static void og_phi_diff(const og_dims *d, const og_phi_fields *F, int i, int j, int k, const int64_t mp[6],
                        const int64_t nb[6], double D[6]) {
    const og_grid *g = d->g;
    const int64_t c = IJK(d, i, j, k);
    const int im = nb_minus(i, d->ni, g->cyclic_x), jm = nb_minus(j, d->nj, g->cyclic_y);
    const int km = d->do_k ? nb_minus(k, d->nk, g->cyclic_z) : k;
    const double *dif = F->dif;
    const double c_ae = F->odx_e[i - 1] * F->ayz[c];
    const double c_aw = F->odx_e[im - 1] * F->ayz[mp[0]];
    const double c_an = F->ody_n[j - 1] * F->axz[c];
    const double c_as = F->ody_n[jm - 1] * F->axz[mp[2]];
    const double c_at = F->ox[i - 1] * F->odz_t[k - 1] * F->axy[c];
    const double c_ab = F->ox[i - 1] * F->odz_t[km - 1] * F->axy[mp[4]];
    /* nb = E, W, S, N, B, T */
    D[0] = og_avg_h(dif[c], dif[nb[0]], F->fx_e[i - 1], F->fx_e_bar[i - 1]) * c_ae;
    D[1] = og_avg_h(dif[nb[1]], dif[c], F->fx_e[im - 1], F->fx_e_bar[im - 1]) * c_aw;
    D[2] = og_avg_h(dif[c], dif[nb[3]], F->fy_n[j - 1], F->fy_n_bar[j - 1]) * c_an;
    D[3] = og_avg_h(dif[nb[2]], dif[c], F->fy_n[jm - 1], F->fy_n_bar[jm - 1]) * c_as;
    D[4] = D[5] = 0.0;
    if (d->do_k) {
        D[4] = og_avg_h(dif[c], dif[nb[5]], F->fz_t[k - 1], F->fz_t_bar[k - 1]) * c_at;
        D[5] = og_avg_h(dif[nb[4]], dif[c], F->fz_t[km - 1], F->fz_t_bar[km - 1]) * c_ab;
    }
}

void oracle_conv_dif_phi0(const og_grid *g, const og_phi_fields *F, double *a, double *b) {
    og_dims d = dims_of(g);
    const int64_t n3 = d.n3;
    const int *flag = F->flag;
    for (int64_t c = 0; c < n3; ++c) {
        for (int l = -3; l <= 3; ++l) A(l, c) = 0.0;
        A(0, c) = -1.0;
        b[c] = 0.0;
    }
    for (int k = 1; k <= d.nk; ++k)
        for (int i = 1; i <= d.ni; ++i)
            for (int j = 1; j <= d.nj; ++j) {
                const int64_t c = IJK(&d, i, j, k);
                if (!og_fluid(flag[c])) continue;
                int64_t mp[6], nb[6];
                og_index1a(&d, flag, i, j, k, mp, nb);
                double D[6];
                og_phi_diff(&d, F, i, j, k, mp, nb, D);
                /* mp = IM, IP, JM, JP, KM, KP */
                if (F->u[c] >= 0.0) {
                    A(1, c) = D[0];
                    A(-1, mp[1]) = D[0] + F->flux_e[c];
                } else {
                    A(1, c) = D[0] - F->flux_e[c];
                    A(-1, mp[1]) = D[0];
                }
                if (!og_fluid(flag[mp[0]])) A(-1, c) = F->u[mp[0]] >= 0.0 ? D[1] + F->flux_e[mp[0]] : D[1];
                if (F->v[c] >= 0.0) {
                    A(2, c) = D[2];
                    A(-2, mp[3]) = D[2] + F->flux_n[c];
                } else {
                    A(2, c) = D[2] - F->flux_n[c];
                    A(-2, mp[3]) = D[2];
                }
                if (!og_fluid(flag[mp[2]])) A(-2, c) = F->v[mp[2]] >= 0.0 ? D[3] + F->flux_n[mp[2]] : D[3];
                if (d.do_k) {
                    if (F->w[c] >= 0.0) {
                        A(3, c) = D[4];
                        A(-3, mp[5]) = D[4] + F->flux_t[c];
                    } else {
                        A(3, c) = D[4] - F->flux_t[c];
                        A(-3, mp[5]) = D[4];
                    }
                    if (!og_fluid(flag[mp[4]])) A(-3, c) = F->w[mp[4]] >= 0.0 ? D[5] + F->flux_t[mp[4]] : D[5];
                }
            }
}

/* FLAG_E/N/T as SET_FLAGS leaves them without internal surfaces
 * (model/set_flags.f:403-470, storage order; UNDEFINED_I from
 * model/param1_mod.f:19; CYCLIC_AT = FLAG 106 or 107, functions.inc:524-530). */
static void og_face_flags(const og_dims *d, const int *flag, int *fe, int *fn, int *ft) {
    const int und = 987654321;
    for (int64_t c = 0; c < d->n3; ++c) fe[c] = fn[c] = ft[c] = und;
    for (int k = 1; k <= d->nk; ++k)
        for (int i = 1; i <= d->ni; ++i)
            for (int j = 1; j <= d->nj; ++j) {
                const int64_t c = IJK(d, i, j, k);
                int64_t mp[6], nb[6];
                og_index1a(d, flag, i, j, k, mp, nb);
                if (og_wall(flag[c])) {
                    fe[c] = fn[c] = ft[c] = 0;
                    fe[mp[0]] = 0;
                    fn[mp[2]] = 0;
                    ft[mp[4]] = 0;
                    if (flag[c] == 106 || flag[c] == 107) {
                        const int jin = j != 1 && j != d->nj, iin = i != 1 && i != d->ni;
                        const int kin = !d->do_k || (k != 1 && k != d->nk);
                        if (i == d->ni && jin && kin && !og_wall(flag[mp[0]])) fe[mp[0]] = 2000;
                        if (j == d->nj && iin && kin && !og_wall(flag[mp[2]])) fn[mp[2]] = 2000;
                        if (k == d->nk && jin && iin && !og_wall(flag[mp[4]])) ft[mp[4]] = 2000;
                    }
                } else if (og_fluid(flag[c])) {
                    if (!og_wall(flag[mp[0]]) && fe[mp[0]] == und) fe[mp[0]] = 2000 + flag[mp[0]];
                    if (!og_wall(flag[mp[2]]) && fn[mp[2]] == und) fn[mp[2]] = 2000 + flag[mp[2]];
                    if (!og_wall(flag[mp[4]]) && ft[mp[4]] == und) ft[mp[4]] = 2000 + flag[mp[4]];
                    if (!og_wall(flag[mp[1]]) && fe[c] == und) fe[c] = 2000 + flag[mp[1]];
                    if (!og_wall(flag[mp[3]]) && fn[c] == und) fn[c] = 2000 + flag[mp[3]];
                    if (!og_wall(flag[mp[5]]) && ft[c] == und) ft[c] = 2000 + flag[mp[5]];
                }
            }
}

/* SOURCE_PHI on the fluid rows, in place; IS_AT_E/N/T (functions.inc:760-781)
 * read the face flags above. */
void oracle_source_phi(const og_grid *g, const double *s_p, const double *s_c, const double *ep, const double *phi,
                       const int *flag, double dil_ep_s, double *a, double *b) {
    og_dims d = dims_of(g);
    const int64_t n3 = d.n3;
    int *fe = malloc(sizeof(int) * 3 * n3);
    int *fn = fe + n3, *ft = fe + 2 * n3;
    og_face_flags(&d, flag, fe, fn, ft);
    for (int k = 1; k <= d.nk; ++k)
        for (int i = 1; i <= d.ni; ++i)
            for (int j = 1; j <= d.nj; ++j) {
                const int64_t c = IJK(&d, i, j, k);
                if (!og_fluid(flag[c])) continue;
                int64_t mp[6], nb[6];
                og_index1a(&d, flag, i, j, k, mp, nb);
                if (ep[c] <= dil_ep_s) {
                    for (int l = -3; l <= 3; ++l) A(l, c) = 0.0;
                    A(0, c) = -1.0;
                    b[c] = 0.0;
                    /* nb = E, W, S, N, B, T; mp = IM, IP, JM, JP, KM, KP */
                    if (ep[nb[1]] > dil_ep_s && !(fe[mp[0]] < 2000)) A(-1, c) = 1.0;
                    if (ep[nb[0]] > dil_ep_s && !(fe[c] < 2000)) A(1, c) = 1.0;
                    if (ep[nb[2]] > dil_ep_s && !(fn[mp[2]] < 2000)) A(-2, c) = 1.0;
                    if (ep[nb[3]] > dil_ep_s && !(fn[c] < 2000)) A(2, c) = 1.0;
                    if (d.do_k) {
                        if (ep[nb[4]] > dil_ep_s && !(ft[mp[4]] < 2000)) A(-3, c) = 1.0;
                        if (ep[nb[5]] > dil_ep_s && !(ft[c] < 2000)) A(3, c) = 1.0;
                    }
                    if ((A(-1, c) + A(1, c) + A(-2, c) + A(2, c) + A(-3, c) + A(3, c)) == 0.0)
                        b[c] = -phi[c];
                    else
                        A(0, c) = -(A(1, c) + A(-1, c) + A(2, c) + A(-2, c) + A(3, c) + A(-3, c));
                } else {
                    A(0, c) = -(A(1, c) + A(-1, c) + A(2, c) + A(-2, c) + A(3, c) + A(-3, c) + s_p[c]);
                    if (b[c] < s_c[c] || phi[c] == 0.0) {
                        b[c] = -s_c[c] + b[c];
                    } else {
                        A(0, c) = A(0, c) - b[c] / phi[c];
                        b[c] = -s_c[c];
                    }
                }
            }
    free(fe);
}

void oracle_under_relax_s(const og_grid *g, const double *var, const int *flag, double ur, double *a, double *b) {
    og_dims d = dims_of(g);
    const int64_t n3 = d.n3;
    const double f1 = 1.0 / ur, f2 = f1 - 1.0;
    for (int64_t c = 0; c < n3; ++c) {
        if (!og_fluid(flag[c])) continue;
        const double ap = A(0, c);
        if (ap != -1.0) {
            A(0, c) = ap * f1;
            b[c] = b[c] + ap * var[c] * f2;
        }
    }
}
