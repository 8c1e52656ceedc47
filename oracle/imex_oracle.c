/*
 * imex_oracle.c -- FP64 CPU restatement of the IMEXLBM per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by the
 * tests, by the smoke check and by the CPU-baseline leg of the benchmark, and
 * never by the product library.  The product path (libimexlbm.so) neither
 * links nor calls anything here.
 *
 * It keeps the reference's structure on purpose, because it is both the
 * parity oracle and the reference-equivalent CPU baseline:
 *   - populations f, ft as [27][lx][ly][lz] doubles, k (z) stride-1, with three
 *     ghost layers on each side (src/lbm.hpp:22-24,107; src/lbm.cpp:43-44);
 *   - a collide pass in place, then the 26-channel periodic ghost fill, then a
 *     pull-stream pass into ft and a swap (src/lbm.cpp:131-207);
 *   - R "ranks" run one after another in this process, each with its own
 *     ghosted subdomain, using the reference's Cartesian factorisation and
 *     block offsets (src/mpi_view_transfer.cpp:53-195); every channel copy
 *     goes through a pack buffer, as MPI would, so R = 1 is the 26
 *     self-channel case (src/mpi_view_transfer.cpp:349-370).
 * Arithmetic follows the reference source expression by expression, left to
 * right, and must be compiled with -ffp-contract=off (see oracle/Makefile).
 *
 * How each function is pinned (tests/test_oracle.py):
 *   lattice tables        Σw=1, Σw e=0, Σw e_a e_b=δ_ab/3, 4th-moment isotropy,
 *                         e[opp[i]] = -e[i]
 *   oracle_tau            closed form with the literal 3.1415926
 *   oracle_decompose      hand-evaluated factorisations (1,2,4,8,12 ranks)
 *   oracle_initialize     Σf = 3p, Σf e = u against the analytic field, mass =
 *                         Σ 3p, KE_0 against the analytic integral
 *   oracle_collision      conserves Σf and Σf e; equilibrium is a fixed point;
 *                         second moment of feq = p δ + u u; non-equilibrium
 *                         part scales by exactly (1 - 1/(tau0+0.5))
 *   oracle_streaming      unique-value tagging: exact periodic permutation
 *   multi-rank ghost fill R = 1, 2, 4, 8 give bitwise-identical populations
 *   whole path            the KE transient 9.852 -> 10.219 -> 9.889 quoted by
 *                         tests/regression.sh:28-30 for the 32^3 Re=1600 deck;
 *                         mass drift and momentum bounds of regression.sh:41-42;
 *                         viscous KE decay rate against 6 nu / L^2
 *   oracle_conserved      the summation order of Kokkos parallel_reduce is not
 *                         reproducible without Kokkos: it is compared with a
 *                         compensated long-double sum at 1e-12 relative
 *                         ("order unpinned", as in the reference's own
 *                         cross-rank tolerance, tests/regression.sh:43).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OQ 27
#define OG 3            /* ghost layers per side, src/lbm.hpp:24            */
#define OPI 3.1415926   /* the literal used for pi, src/lbm.cpp:3           */

/* D3Q27 weights, velocities and opposites: src/lbm.cpp:9-35 (same order). */
static const double OW[OQ] = {
    8.0 / 27.0,
    2.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0,
    1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0,
    1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0,
    1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0,
    1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0,
    1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0};
static const int OE[OQ][3] = {
    {0, 0, 0},
    {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
    {1, 1, 0}, {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0},
    {1, 0, 1}, {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1},
    {0, 1, 1}, {0, -1, -1}, {0, 1, -1}, {0, -1, 1},
    {1, 1, 1}, {-1, -1, -1}, {1, -1, 1}, {-1, 1, -1},
    {1, 1, -1}, {-1, -1, 1}, {1, -1, -1}, {-1, 1, 1}};
static const int OOPP[OQ] = {0, 2, 1, 4, 3, 6, 5,
                             8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17,
                             20, 19, 22, 21, 24, 23, 26, 25};

typedef struct {
    int p[3];        /* rank coordinates px, py, pz                        */
    int l[3];        /* interior extents l_l                               */
    int n[3];        /* ghosted extents lx, ly, lz                         */
    int lo[3];       /* global offsets x_lo, y_lo, z_lo                    */
    size_t plane;    /* lx*ly*lz, the per-direction stride                 */
    double *f, *ft;
} orank;

typedef struct {
    int gl[3];
    int nranks, r[3];
    double tau0, rho0, u0, cs2;
    orank *rk;
    double *buf;     /* pack buffers: 26 channels per rank, sized to fit   */
    size_t chan_cap; /* doubles per channel slot                           */
} oracle_t;

/* ------------------------------------------------------------------------ */
/* Scalars                                                                  */
/* ------------------------------------------------------------------------ */

/* System::System, src/System.cpp:13,28: tau = u0*sx/Re/cs2/2.0/3.1415926,
 * evaluated left to right with sx and Re ints. */
double oracle_tau(double u0, int sx, int Re)
{
    const double cs2 = 1.0 / 3.0;
    return u0 * sx / Re / cs2 / 2.0 / 3.1415926;
}

void oracle_tables(double *w, int *e, int *opp)
{
    for (int a = 0; a < OQ; ++a) {
        w[a] = OW[a];
        opp[a] = OOPP[a];
        for (int d = 0; d < 3; ++d) e[3 * a + d] = OE[a][d];
    }
}

/* LBM::setup_Cartesian factorisation, src/mpi_view_transfer.cpp:55-91.
 * The local extents are int quotients converted to double; the first strict
 * minimum of the squared deviation wins; loops run x outer, y inner. */
void oracle_decompose(int nranks, int glx, int gly, int glz, int *out)
{
    double min_e = 1e16;
    out[0] = out[1] = out[2] = 1;
    for (int i = 1; i < nranks + 1; i++) {
        if (nranks % i != 0) continue;
        for (int j = 1; j < nranks / i + 1; j++) {
            if ((nranks / i) % j != 0) continue;
            const int zt = nranks / i / j;
            const double loc_x = glx / i;
            const double loc_y = gly / j;
            const double loc_z = glz / zt;
            const double average = (loc_x + loc_y + loc_z) / 3.0;
            const double u2 = (loc_x - average) * (loc_x - average) +
                              (loc_y - average) * (loc_y - average) +
                              (loc_z - average) * (loc_z - average);
            if (u2 < min_e) {
                min_e = u2;
                out[0] = i;
                out[1] = j;
                out[2] = zt;
            }
        }
    }
}

/* setup_Local, src/mpi_view_transfer.cpp:165-167,185-188. */
int oracle_local_len(int gl, int r, int p)
{
    return (p - gl % r >= 0) ? gl / r : gl / r + 1;
}
int oracle_block_lo(int gl, int r, int p)
{
    const int base = gl / r, rem = gl % r;
    return p * base + (p < rem ? p : rem);
}

/* Rank linearisation me = px + rx*pz + rx*rz*py (y slowest),
 * src/mpi_view_transfer.cpp:96-98. */
static int rank_of(const oracle_t *o, int px, int py, int pz)
{
    px = (px % o->r[0] + o->r[0]) % o->r[0];
    py = (py % o->r[1] + o->r[1]) % o->r[1];
    pz = (pz % o->r[2] + o->r[2]) % o->r[2];
    return px + o->r[0] * pz + o->r[0] * o->r[2] * py;
}

static inline size_t IDX(const orank *R, int ii, int i, int j, int k)
{
    return (((size_t)ii * R->n[0] + i) * R->n[1] + j) * R->n[2] + k;
}

/* ------------------------------------------------------------------------ */
/* Construction                                                             */
/* ------------------------------------------------------------------------ */

void oracle_destroy(void *h)
{
    oracle_t *o = (oracle_t *)h;
    if (!o) return;
    if (o->rk) {
        for (int r = 0; r < o->nranks; ++r) {
            free(o->rk[r].f);
            free(o->rk[r].ft);
        }
        free(o->rk);
    }
    free(o->buf);
    free(o);
}

/* LBM::LBM + setup_MPI geometry (src/lbm.hpp:179-183,
 * src/mpi_view_transfer.cpp:41-195).  Returns NULL on allocation failure. */
void *oracle_create(int glx, int gly, int glz, double tau0, double rho0,
                    double u0, int nranks)
{
    if (glx < 1 || gly < 1 || glz < 1 || nranks < 1) return NULL;
    oracle_t *o = (oracle_t *)calloc(1, sizeof(oracle_t));
    if (!o) return NULL;
    o->gl[0] = glx; o->gl[1] = gly; o->gl[2] = glz;
    o->tau0 = tau0; o->rho0 = rho0; o->u0 = u0;
    o->cs2 = 1.0 / 3.0;
    o->nranks = nranks;
    oracle_decompose(nranks, glx, gly, glz, o->r);
    o->rk = (orank *)calloc((size_t)nranks, sizeof(orank));
    if (!o->rk) { oracle_destroy(o); return NULL; }
    size_t maxface = 0;
    for (int me = 0; me < nranks; ++me) {
        orank *R = &o->rk[me];
        R->p[0] = me % o->r[0];
        R->p[2] = (me / o->r[0]) % o->r[2];
        R->p[1] = me / o->r[0] / o->r[2];
        for (int d = 0; d < 3; ++d) {
            R->l[d] = oracle_local_len(o->gl[d], o->r[d], R->p[d]);
            if (R->l[d] < 1) { oracle_destroy(o); return NULL; }
            R->n[d] = R->l[d] + 2 * OG;
            R->lo[d] = oracle_block_lo(o->gl[d], o->r[d], R->p[d]);
        }
        R->plane = (size_t)R->n[0] * R->n[1] * R->n[2];
        R->f = (double *)calloc(OQ * R->plane, sizeof(double));
        R->ft = (double *)calloc(OQ * R->plane, sizeof(double));
        if (!R->f || !R->ft) { oracle_destroy(o); return NULL; }
        size_t a = (size_t)R->n[1] * R->n[2], b = (size_t)R->n[0] * R->n[2],
               c = (size_t)R->n[0] * R->n[1];
        size_t m = a > b ? a : b;
        m = m > c ? m : c;
        if (m > maxface) maxface = m;
    }
    o->chan_cap = OQ * maxface;
    o->buf = (double *)malloc(sizeof(double) * o->chan_cap * 26 * (size_t)nranks);
    if (!o->buf) { oracle_destroy(o); return NULL; }
    return o;
}

void oracle_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int oracle_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Initialize: init_macro + initf, src/lbm.cpp:78-108, every cell including  */
/* ghosts.                                                                  */
/* ------------------------------------------------------------------------ */
void oracle_initialize(void *h)
{
    oracle_t *o = (oracle_t *)h;
    const double rho0 = o->rho0, u0 = o->u0, cs2 = o->cs2;
    for (int me = 0; me < o->nranks; ++me) {
        orank *R = &o->rk[me];
        const int lx = R->n[0], ly = R->n[1], lz = R->n[2];
#pragma omp parallel for schedule(static)
        for (int i = 0; i < lx; ++i)
            for (int j = 0; j < ly; ++j)
                for (int k = 0; k < lz; ++k) {
                    const double xx = (double)(i - OG + R->lo[0]) / (double)o->gl[0] * 2.0 * OPI;
                    const double yy = (double)(j - OG + R->lo[1]) / (double)o->gl[1] * 2.0 * OPI;
                    const double zz = (double)(k - OG + R->lo[2]) / (double)o->gl[2] * 2.0 * OPI;
                    const double sx = sin(xx), cx = cos(xx);
                    const double sy = sin(yy), cy = cos(yy);
                    const double cz = cos(zz);
                    const double u = u0 * sx * cy * cz;
                    const double v = -u0 * cx * sy * cz;
                    const double w = 0.0;
                    const double p = rho0 * cs2 +
                                     rho0 * u0 * u0 / 16.0 * (cos(2.0 * xx) + cos(2.0 * yy)) *
                                         (cos(2.0 * zz) + 2.0);
                    for (int ii = 0; ii < OQ; ++ii) {
                        const double edu = OE[ii][0] * u + OE[ii][1] * v + OE[ii][2] * w;
                        const double udu = u * u + v * v + w * w;
                        R->f[IDX(R, ii, i, j, k)] =
                            OW[ii] * (3.0 * p + 3.0 * edu + 4.5 * edu * edu - 1.5 * udu);
                        R->ft[IDX(R, ii, i, j, k)] = 0.0;
                    }
                }
    }
}

/* ------------------------------------------------------------------------ */
/* Collision: src/lbm.cpp:131-166, in place over the interior.              */
/* ------------------------------------------------------------------------ */
void oracle_collision(void *h)
{
    oracle_t *o = (oracle_t *)h;
    const double inv_tau = 1.0 / (o->tau0 + 0.5);
    for (int me = 0; me < o->nranks; ++me) {
        orank *R = &o->rk[me];
#pragma omp parallel for schedule(static)
        for (int i = OG; i < OG + R->l[0]; ++i)
            for (int j = OG; j < OG + R->l[1]; ++j)
                for (int k = OG; k < OG + R->l[2]; ++k) {
                    double fl[OQ];
                    double pl = 0.0, ul = 0.0, vl = 0.0, wl = 0.0;
                    for (int ii = 0; ii < OQ; ++ii) {
                        const double fv = R->f[IDX(R, ii, i, j, k)];
                        fl[ii] = fv;
                        pl += fv;
                        ul += fv * OE[ii][0];
                        vl += fv * OE[ii][1];
                        wl += fv * OE[ii][2];
                    }
                    const double u = ul, v = vl, w = wl;
                    const double pp = pl * (1.0 / 3.0);
                    const double udu = u * u + v * v + w * w;
                    for (int ii = 0; ii < OQ; ++ii) {
                        const double edu = OE[ii][0] * u + OE[ii][1] * v + OE[ii][2] * w;
                        const double feq =
                            OW[ii] * (3.0 * pp + 3.0 * edu + 4.5 * edu * edu - 1.5 * udu);
                        R->f[IDX(R, ii, i, j, k)] = fl[ii] - (fl[ii] - feq) * inv_tau;
                    }
                }
    }
}

/* ------------------------------------------------------------------------ */
/* Ghost fill: pack_f / exchange_f / unpack_f, src/mpi_view_transfer.cpp:    */
/* 260-456.  A channel is named by the offset (cx,cy,cz) of the rank it is   */
/* sent to.  The sender packs its first (c=-1) or last (c=+1) interior layer */
/* on each non-zero axis; on zero axes a face spans the full ghosted extent  */
/* and an edge the interior only.  The receiver, the rank at +c, writes it   */
/* into its innermost ghost layer on the opposite side.  Faces are unpacked  */
/* before edges and edges before corners, so the innermost ghost shell ends  */
/* up exact, as in the reference.                                           */
/* ------------------------------------------------------------------------ */
typedef struct { int c[3]; } ochan;
static int g_nchan = 0;
static ochan g_chan[26];

static void build_channels(void)
{
    if (g_nchan) return;
    int n = 0;
    /* faces, then edges, then corners */
    for (int nnz = 1; nnz <= 3; ++nnz)
        for (int cx = -1; cx <= 1; ++cx)
            for (int cy = -1; cy <= 1; ++cy)
                for (int cz = -1; cz <= 1; ++cz) {
                    const int z = (cx != 0) + (cy != 0) + (cz != 0);
                    if (z != nnz) continue;
                    g_chan[n].c[0] = cx; g_chan[n].c[1] = cy; g_chan[n].c[2] = cz;
                    ++n;
                }
    g_nchan = n;
}

/* Range along one axis for the sender (recv = 0) or the receiver (recv = 1). */
static void axis_range(int c, int is_face, int recv, int lint, int nfull,
                       int *a, int *b)
{
    if (c == 0) {
        if (is_face) { *a = 0; *b = nfull; }
        else { *a = OG; *b = OG + lint; }
    } else if (!recv) {
        *a = (c < 0) ? OG : OG + lint - 1;
        *b = *a + 1;
    } else {
        /* receiver is the rank at +c; the data came from its -c side */
        *a = (c < 0) ? OG + lint : OG - 1;
        *b = *a + 1;
    }
}

static void oracle_passf(oracle_t *o, int which)
{
    build_channels();
    /* pack every rank's 26 channels */
    for (int me = 0; me < o->nranks; ++me) {
        orank *R = &o->rk[me];
        double *F = which ? R->ft : R->f;
        for (int ch = 0; ch < 26; ++ch) {
            const int *c = g_chan[ch].c;
            const int is_face = (c[0] != 0) + (c[1] != 0) + (c[2] != 0) == 1;
            int a[3], b[3];
            for (int d = 0; d < 3; ++d)
                axis_range(c[d], is_face, 0, R->l[d], R->n[d], &a[d], &b[d]);
            double *dst = o->buf + ((size_t)me * 26 + ch) * o->chan_cap;
            size_t m = 0;
            for (int ii = 0; ii < OQ; ++ii)
                for (int i = a[0]; i < b[0]; ++i)
                    for (int j = a[1]; j < b[1]; ++j)
                        for (int k = a[2]; k < b[2]; ++k)
                            dst[m++] = F[IDX(R, ii, i, j, k)];
        }
    }
    /* unpack: receiver `me` takes channel ch from the rank at -c */
    for (int me = 0; me < o->nranks; ++me) {
        orank *R = &o->rk[me];
        double *F = which ? R->ft : R->f;
        for (int ch = 0; ch < 26; ++ch) {
            const int *c = g_chan[ch].c;
            const int is_face = (c[0] != 0) + (c[1] != 0) + (c[2] != 0) == 1;
            const int src = rank_of(o, R->p[0] - c[0], R->p[1] - c[1], R->p[2] - c[2]);
            int a[3], b[3];
            for (int d = 0; d < 3; ++d)
                axis_range(c[d], is_face, 1, R->l[d], R->n[d], &a[d], &b[d]);
            /* The sender's slab and this slab have equal shape: on zero axes
             * both ranks share the same l and n (same column of the grid). */
            const double *s = o->buf + ((size_t)src * 26 + ch) * o->chan_cap;
            size_t m = 0;
            for (int ii = 0; ii < OQ; ++ii)
                for (int i = a[0]; i < b[0]; ++i)
                    for (int j = a[1]; j < b[1]; ++j)
                        for (int k = a[2]; k < b[2]; ++k)
                            F[IDX(R, ii, i, j, k)] = s[m++];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Streaming: passf, then pull ft(ii,x) = f(ii, x - e_ii), then swap,        */
/* src/lbm.cpp:168-207.                                                     */
/* ------------------------------------------------------------------------ */
void oracle_streaming(void *h)
{
    oracle_t *o = (oracle_t *)h;
    oracle_passf(o, 0);
    for (int me = 0; me < o->nranks; ++me) {
        orank *R = &o->rk[me];
#pragma omp parallel for schedule(static)
        for (int i = OG; i < OG + R->l[0]; ++i)
            for (int j = OG; j < OG + R->l[1]; ++j)
                for (int k = OG; k < OG + R->l[2]; ++k)
                    for (int ii = 0; ii < OQ; ++ii)
                        R->ft[IDX(R, ii, i, j, k)] =
                            R->f[IDX(R, ii, i - OE[ii][0], j - OE[ii][1], k - OE[ii][2])];
        double *t = R->f;
        R->f = R->ft;
        R->ft = t;
    }
}

void oracle_step(void *h, int nsteps)
{
    for (int s = 0; s < nsteps; ++s) {
        oracle_collision(h);
        oracle_streaming(h);
    }
}

/* ------------------------------------------------------------------------ */
/* Conserved: src/lbm.cpp:266-306.  Per-rank sums in iteration order, then  */
/* summed over ranks in rank order (the 5-double Allreduce).                */
/* ------------------------------------------------------------------------ */
void oracle_conserved(void *h, double *out)
{
    oracle_t *o = (oracle_t *)h;
    double tot[5] = {0, 0, 0, 0, 0};
    for (int me = 0; me < o->nranks; ++me) {
        orank *R = &o->rk[me];
        double m = 0.0, mx = 0.0, my = 0.0, mz = 0.0, e_kin = 0.0;
        for (int i = OG; i < OG + R->l[0]; ++i)
            for (int j = OG; j < OG + R->l[1]; ++j)
                for (int k = OG; k < OG + R->l[2]; ++k) {
                    double pl = 0.0, ul = 0.0, vl = 0.0, wl = 0.0;
                    for (int ii = 0; ii < OQ; ++ii) {
                        const double fv = R->f[IDX(R, ii, i, j, k)];
                        pl += fv;
                        ul += fv * OE[ii][0];
                        vl += fv * OE[ii][1];
                        wl += fv * OE[ii][2];
                    }
                    m += pl;
                    mx += ul;
                    my += vl;
                    mz += wl;
                    e_kin += 0.5 * (ul * ul + vl * vl + wl * wl);
                }
        tot[0] += m; tot[1] += mx; tot[2] += my; tot[3] += mz; tot[4] += e_kin;
    }
    memcpy(out, tot, sizeof(tot));
}

/* Same per-cell quantities, summed with Neumaier compensation in long
 * double: an order-independent stand-in for the exact sum. */
void oracle_conserved_compensated(void *h, double *out)
{
    oracle_t *o = (oracle_t *)h;
    long double s[5] = {0, 0, 0, 0, 0}, c[5] = {0, 0, 0, 0, 0};
    for (int me = 0; me < o->nranks; ++me) {
        orank *R = &o->rk[me];
        for (int i = OG; i < OG + R->l[0]; ++i)
            for (int j = OG; j < OG + R->l[1]; ++j)
                for (int k = OG; k < OG + R->l[2]; ++k) {
                    double pl = 0.0, ul = 0.0, vl = 0.0, wl = 0.0;
                    for (int ii = 0; ii < OQ; ++ii) {
                        const double fv = R->f[IDX(R, ii, i, j, k)];
                        pl += fv;
                        ul += fv * OE[ii][0];
                        vl += fv * OE[ii][1];
                        wl += fv * OE[ii][2];
                    }
                    const double v[5] = {pl, ul, vl, wl,
                                         0.5 * (ul * ul + vl * vl + wl * wl)};
                    for (int q = 0; q < 5; ++q) {
                        const long double t = s[q] + v[q];
                        if (fabsl(s[q]) >= fabsl((long double)v[q]))
                            c[q] += (s[q] - t) + v[q];
                        else
                            c[q] += ((long double)v[q] - t) + s[q];
                        s[q] = t;
                    }
                }
    }
    for (int q = 0; q < 5; ++q) out[q] = (double)(s[q] + c[q]);
}

/* ------------------------------------------------------------------------ */
/* State access in the global reference order out[((ii*glx+X)*gly+Y)*glz+Z] */
/* (interior only).                                                         */
/* ------------------------------------------------------------------------ */
void oracle_get_f(void *h, double *out)
{
    oracle_t *o = (oracle_t *)h;
    const size_t gx = o->gl[0], gy = o->gl[1], gz = o->gl[2];
    for (int me = 0; me < o->nranks; ++me) {
        orank *R = &o->rk[me];
        for (int ii = 0; ii < OQ; ++ii)
            for (int i = 0; i < R->l[0]; ++i)
                for (int j = 0; j < R->l[1]; ++j)
                    for (int k = 0; k < R->l[2]; ++k)
                        out[((ii * gx + (size_t)(i + R->lo[0])) * gy + (size_t)(j + R->lo[1])) * gz +
                            (size_t)(k + R->lo[2])] = R->f[IDX(R, ii, i + OG, j + OG, k + OG)];
    }
}

void oracle_set_f(void *h, const double *in)
{
    oracle_t *o = (oracle_t *)h;
    const size_t gx = o->gl[0], gy = o->gl[1], gz = o->gl[2];
    for (int me = 0; me < o->nranks; ++me) {
        orank *R = &o->rk[me];
        for (int ii = 0; ii < OQ; ++ii)
            for (int i = 0; i < R->l[0]; ++i)
                for (int j = 0; j < R->l[1]; ++j)
                    for (int k = 0; k < R->l[2]; ++k)
                        R->f[IDX(R, ii, i + OG, j + OG, k + OG)] =
                            in[((ii * gx + (size_t)(i + R->lo[0])) * gy + (size_t)(j + R->lo[1])) * gz +
                               (size_t)(k + R->lo[2])];
    }
}

/* Geometry of one rank, for tests of the decomposition. */
void oracle_rank_geometry(void *h, int me, int *out9)
{
    oracle_t *o = (oracle_t *)h;
    const orank *R = &o->rk[me];
    for (int d = 0; d < 3; ++d) {
        out9[d] = R->p[d];
        out9[3 + d] = R->l[d];
        out9[6 + d] = R->lo[d];
    }
}

void oracle_grid(void *h, int *out3)
{
    oracle_t *o = (oracle_t *)h;
    out3[0] = o->r[0]; out3[1] = o->r[1]; out3[2] = o->r[2];
}
