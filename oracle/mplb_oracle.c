/*
 * mplb_oracle.c -- CPU oracle (test infrastructure only; see mplb_oracle.h).
 *
 * Plain C, FP64, array-of-structures storage ([node][Q]), dense mat-vecs, and the
 * reference's evaluation order in every expression (no FMA: build with
 * -ffp-contract=off).  Paths below are relative to the reference's src/.
 *
 * Pinning (what checks this restatement, tests/test_oracle_*.py, tests/test_tables.py):
 *   lattice tables      bit-exact against the reference's own latticeModel/D2Q9_*.c and
 *                       D3Q15_*.c compiled into oracle/_ref (oracle/Makefile), plus the
 *                       isotropy identities and M * invM = I.
 *   topology            neighbour tables and boundary lists bit-exact against the
 *                       reference's own latticeBox2D/{genericBoundary,periodicX,periodicY,
 *                       periodicXY}.c and latticeBox3D/{genericBoundary3D,periodicXY3D,
 *                       periodicZ3D,periodicXYZ3D}.c compiled into oracle/_ref (refbox.c);
 *                       duplicate-node periodicity also against an independent wrap formula.
 *   EOS / potential     closed forms: VdW critical point, ideal gas, CS/PR limits.
 *   liMRT / xuMRT       the textbook second-order equilibrium is a collision fixed point
 *                       (every equilibrium moment); bounceBack channels: zero momentum on
 *                       wall nodes (U = F/2rho), exact Poiseuille parabola with
 *                       nu = cs2 (1/s_nu - 1/2); shear-wave decay at that nu; flat-interface
 *                       coexistence against the Li-Luo-Li mechanical-stability condition
 *                       with eps = -16 G sigma (sigma-term); sigma has no effect where the
 *                       interaction force vanishes (liMRT uses Fi); surface-tension tensors
 *                       leave flat coexistence unchanged and scale the Laplace jump by
 *                       (1 - kappa).  Also mass conservation and momentum change = force.
 *   stream              permutation on fully periodic boxes (exact mass conservation,
 *                       inverse of the reverse stream).
 *   interForce          zero on uniform fields, antisymmetry under mirror, smooth-field
 *                       expansion F ~ -G psi cs2-weighted grad psi.
 *   boundaries          bounceBack mass conservation in closed boxes and zero wall
 *                       momentum; fixedT/fixedU reproduce the wall value when the
 *                       neighbour is at equilibrium; fixedU is the non-equilibrium
 *                       extrapolation (Guo, Zheng, Shi 2002) next to forced non-equilibrium
 *                       neighbours; outflow is the implicit-upwind convective condition on
 *                       the outlet face; glibc rand() stream for fixedRandomT taken from
 *                       the host libc.
 *   heatSource          vanishes for uniform T, U; divergence/gradient stencils exact on
 *                       linear fields; with myMRT it makes the steady flux rho dT/dy
 *                       uniform across a static density profile.
 *   myMRT / liTempMRT   frozen-flow transport: a sine temperature decays at the closed-form
 *                       diffusivity (myMRT: the heat source's conductivity) and liTempMRT
 *                       advects at the fixed-point speed of its whole-moment correction.
 *   macroPressure       closed forms on uniform, linear and quadratic potentials, wall and
 *                       corner nodes included.
 * Parity unpinned (marked where they occur): liTempMRT's source filter
 * ss = 1.5 s - 0.5 ss_old; the heat source's time levels (old T and U, new rho); the
 * fixedRandomT draw order; xuMRT's sigma-term built from the total force (no physical
 * law separates it from Fi once gravity acts); the sign of liMRTPressure's one-sided
 * gradient at walls; the spot box's lists (periodicXY3DSpot.c reads dictIO, whose
 * headers need <mpi.h>, so it is not compiled into oracle/_ref).
 */
#include "mplb_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int model, d, Q;
    double cs2;
    int vel[15][3];
    int rev[15];
    double omega[15];
    double w[15];
    double M[15][15];
    double invM[15][15];
} orc_lattice;

struct orc_case {
    orc_params p;
    orc_lattice lat;
    int64_t N;
    int* nb;                 /* [N][Q] upstream neighbour x - c_k, or -1 */
    unsigned* isOnBnd;       /* readLatticeMesh.c:263-295 */
    int nbd;
    int bdsize[ORC_MAXBD];
    int* bd[ORC_MAXBD];
    double *rho, *T, *p_, *U, *Fi, *Ft;   /* U, Fi, Ft: [N][3] */
    double *f, *g, *fswap, *gswap;        /* [N][Q] */
    double *ss;                           /* g scalarSource */
    int f_update, g_update;
    double rho0;
    /* time (timeInfo.h) */
    int current, stp;
    int frdt_local[2][ORC_MAXBD];
    int err;
    char errmsg[256];
};

static void orc_fail(orc_case* c, const char* msg) {
    if (!c->err) {
        c->err = 1;
        snprintf(c->errmsg, sizeof(c->errmsg), "%s", msg);
    }
}

/* ------------------------------------------------------------------------- */
/* Lattice tables: latticeModel/D2Q9/D2Q9_{Vel,Reverse,omega,ppw,cs2,MRT,MRTInv}.c and
 * latticeModel/D3Q15/D3Q15_*.c, assembled as in setLatticeInfo.c:16-90.          */
static void set_lattice(orc_lattice* L, int model) {
    memset(L, 0, sizeof(*L));
    L->model = model;
    if (model == ORC_D2Q9) {
        static const int v[9][3] = {{0,0,0},{1,0,0},{0,1,0},{-1,0,0},{0,-1,0},
                                    {1,1,0},{-1,1,0},{-1,-1,0},{1,-1,0}};        /* D2Q9_Vel.c:20-54 */
        static const int r[9] = {0,3,4,1,2,7,8,5,6};                                 /* D2Q9_Reverse.c:11-19 */
        static const int Mi[9][9] = {                                                 /* D2Q9_MRT.c:20-108 */
            { 1, 1, 1, 1, 1, 1, 1, 1, 1},
            {-4,-1,-1,-1,-1, 2, 2, 2, 2},
            { 4,-2,-2,-2,-2, 1, 1, 1, 1},
            { 0, 1, 0,-1, 0, 1,-1,-1, 1},
            { 0,-2, 0, 2, 0, 1,-1,-1, 1},
            { 0, 0, 1, 0,-1, 1, 1,-1,-1},
            { 0, 0,-2, 0, 2, 1, 1,-1,-1},
            { 0, 1,-1, 1,-1, 0, 0, 0, 0},
            { 0, 0, 0, 0, 0, 1,-1, 1,-1}};
        L->d = 2; L->Q = 9; L->cs2 = 1.0 / 3.0;                                      /* D2Q9_cs2.c */
        for (int k = 0; k < 9; k++) {
            for (int j = 0; j < 3; j++) L->vel[k][j] = v[k][j];
            L->rev[k] = r[k];
            for (int j = 0; j < 9; j++) L->M[k][j] = Mi[k][j];
        }
        L->omega[0] = 4.0 / 9.0;                                                      /* D2Q9_omega.c:12-20 */
        for (int k = 1; k <= 4; k++) L->omega[k] = 1.0 / 9.0;
        for (int k = 5; k <= 8; k++) L->omega[k] = 1.0 / 36.0;
        L->w[0] = 0;                                                                  /* D2Q9_ppw.c:12-20 */
        for (int k = 1; k <= 4; k++) L->w[k] = 1.0 / 3;
        for (int k = 5; k <= 8; k++) L->w[k] = 1.0 / 12;
        /* D2Q9_MRTInv.c:20-115, written with c1 = 1/9, c2 = 1/6, c3 = 1/4 */
        const double c1 = 1.0 / 9.0, c2 = 1.0 / 6.0, c3 = 1.0 / 4.0;
        const double inv[9][9] = {
            {c1, -c1,    c1,     0,    0,      0,    0,      0,   0},
            {c1, -c1/4, -c1/2,  c2,  -c2,      0,    0,     c3,   0},
            {c1, -c1/4, -c1/2,   0,    0,     c2,  -c2,    -c3,   0},
            {c1, -c1/4, -c1/2, -c2,   c2,      0,    0,     c3,   0},
            {c1, -c1/4, -c1/2,   0,    0,    -c2,   c2,    -c3,   0},
            {c1,  c1/2,  c1/4,  c2,  c2/2,    c2,  c2/2,     0,  c3},
            {c1,  c1/2,  c1/4, -c2, -c2/2,    c2,  c2/2,     0, -c3},
            {c1,  c1/2,  c1/4, -c2, -c2/2,   -c2, -c2/2,     0,  c3},
            {c1,  c1/2,  c1/4,  c2,  c2/2,   -c2, -c2/2,     0, -c3}};
        for (int k = 0; k < 9; k++)
            for (int j = 0; j < 9; j++) L->invM[k][j] = inv[k][j];
    } else {
        static const int v[15][3] = {{0,0,0},{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1},
                                     {1,1,1},{-1,1,1},{1,-1,1},{-1,-1,1},
                                     {1,1,-1},{-1,1,-1},{1,-1,-1},{-1,-1,-1}};     /* D3Q15_Vel.c:20-78 */
        static const int r[15] = {0,2,1,4,3,6,5,14,13,12,11,10,9,8,7};             /* D3Q15_Reverse.c:11-25 */
        L->d = 3; L->Q = 15; L->cs2 = 1.0 / 3.0;                                      /* D3Q15_cs2.c */
        for (int k = 0; k < 15; k++) {
            for (int j = 0; j < 3; j++) L->vel[k][j] = v[k][j];
            L->rev[k] = r[k];
        }
        L->omega[0] = 2.0 / 9.0;                                                      /* D3Q15_omega.c:12-28 */
        for (int k = 1; k <= 6; k++) L->omega[k] = 1.0 / 9.0;
        for (int k = 7; k < 15; k++) L->omega[k] = 1.0 / 72.0;
        L->w[0] = 0;                                                                  /* D3Q15_ppw.c:12-28 */
        for (int k = 1; k <= 6; k++) L->w[k] = 1.0 / 3.0;
        for (int k = 7; k < 15; k++) L->w[k] = 1.0 / 24.0;
        /* D3Q15_MRT.c:20-258, row by row as in the reference table */
        static const int Mi[15][15] = {
            { 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
            {-2,-1,-1,-1,-1,-1,-1, 1, 1, 1, 1, 1, 1, 1, 1},
            {16,-4,-4,-4,-4,-4,-4, 1, 1, 1, 1, 1, 1, 1, 1},
            { 0, 1,-1, 0, 0, 0, 0, 1,-1, 1,-1, 1,-1, 1,-1},
            { 0,-4, 4, 0, 0, 0, 0, 1,-1, 1,-1, 1,-1, 1,-1},
            { 0, 0, 0, 1,-1, 0, 0, 1, 1,-1,-1, 1, 1,-1,-1},
            { 0, 0, 0,-4, 4, 0, 0, 1, 1,-1,-1, 1, 1,-1,-1},
            { 0, 0, 0, 0, 0, 1,-1, 1, 1, 1, 1,-1,-1,-1,-1},
            { 0, 0, 0, 0, 0,-4, 4, 1, 1, 1, 1,-1,-1,-1,-1},
            { 0, 2, 2,-1,-1,-1,-1, 0, 0, 0, 0, 0, 0, 0, 0},
            { 0, 0, 0, 1, 1,-1,-1, 0, 0, 0, 0, 0, 0, 0, 0},
            { 0, 0, 0, 0, 0, 0, 0, 1,-1,-1, 1, 1,-1,-1, 1},
            { 0, 0, 0, 0, 0, 0, 0, 1, 1,-1,-1,-1,-1, 1, 1},
            { 0, 0, 0, 0, 0, 0, 0, 1,-1, 1,-1,-1, 1,-1, 1},
            { 0, 0, 0, 0, 0, 0, 0, 1,-1,-1, 1,-1, 1, 1,-1}};
        for (int k = 0; k < 15; k++)
            for (int j = 0; j < 15; j++) L->M[k][j] = Mi[k][j];
        /* D3Q15_MRTInv.c:21-259: column j of invM, literal constants of the reference */
        for (int k = 0; k < 15; k++) {
            L->invM[k][0] = 1.0 / 15.0;
            L->invM[k][1] = (k == 0) ? -1.0 / 9.0 : (k <= 6 ? -1.0 / 18.0 : 1.0 / 18.0);
            L->invM[k][2] = (k == 0) ? 2.0 / 45 : (k <= 6 ? -1.0 / 90.0 : 1.0 / 360.0);
        }
        for (int k = 0; k < 15; k++) {
            for (int a = 0; a < 3; a++) {
                int col = 3 + 2 * a;         /* columns 3,5,7: j_a ; 4,6,8: q_a */
                int c = v[k][a];
                if (k == 0) { L->invM[k][col] = 0; L->invM[k][col + 1] = 0; continue; }
                if (k <= 6) {
                    L->invM[k][col] = c == 0 ? 0 : (c > 0 ? 0.1 : -0.1);
                    L->invM[k][col + 1] = c == 0 ? 0 : (c > 0 ? -0.1 : 0.1);
                } else {
                    L->invM[k][col] = c > 0 ? 0.1 : -0.1;
                    L->invM[k][col + 1] = c > 0 ? 0.025 : -0.025;
                }
            }
            /* column 9: 3cx^2 - c^2 moment, column 10: cy^2 - cz^2 */
            if (k == 1 || k == 2) { L->invM[k][9] = 1.0 / 6.0; L->invM[k][10] = 0; }
            else if (k >= 3 && k <= 6) {
                L->invM[k][9] = -1.0 / 12.0;
                L->invM[k][10] = (k <= 4) ? 0.25 : -0.25;
            } else { L->invM[k][9] = 0; L->invM[k][10] = 0; }
            /* columns 11..14: products cx*cy, cy*cz, cx*cz, cx*cy*cz on the corner links */
            if (k >= 7) {
                int cx = v[k][0], cy = v[k][1], cz = v[k][2];
                L->invM[k][11] = (cx * cy > 0) ? 0.125 : -0.125;
                L->invM[k][12] = (cy * cz > 0) ? 0.125 : -0.125;
                L->invM[k][13] = (cx * cz > 0) ? 0.125 : -0.125;
                L->invM[k][14] = (cx * cy * cz > 0) ? 0.125 : -0.125;
            } else {
                L->invM[k][11] = L->invM[k][12] = L->invM[k][13] = L->invM[k][14] = 0;
            }
        }
    }
}

void orc_lattice_tables(int model, int* Q, int* d, double* cs2, int* vel, int* rev,
                        double* omega, double* w, double* M, double* invM) {
    orc_lattice L;
    set_lattice(&L, model);
    *Q = L.Q; *d = L.d; *cs2 = L.cs2;
    for (int k = 0; k < L.Q; k++) {
        for (int j = 0; j < 3; j++) vel[k * 3 + j] = L.vel[k][j];
        rev[k] = L.rev[k]; omega[k] = L.omega[k]; w[k] = L.w[k];
        for (int j = 0; j < L.Q; j++) { M[k * L.Q + j] = L.M[k][j]; invM[k * L.Q + j] = L.invM[k][j]; }
    }
}

/* basic/matVecMult.c:4-20: res[i] = sum_j M[i][j] v[j], j ascending from 0 */
static void matvec(const double M[15][15], const double* v, double* res, int n) {
    for (int i = 0; i < n; i++) {
        double r = 0;
        for (int j = 0; j < n; j++) r += M[i][j] * v[j];
        res[i] = r;
    }
}

/* basic/vectorMag2.c:4-19 */
static double mag2(const double v[3]) {
    double m = 0;
    for (int k = 0; k < 3; k++) m += v[k] * v[k];
    return m;
}

/* ------------------------------------------------------------------------- */
/* EOS/p_eos.c:7-64 */
static double p_eos(const orc_params* p, double rho, double T) {
    double pr = 0, alpha, beta;
    switch (p->eos) {
    case ORC_IDEAL:
        pr = rho * p->R * T;
        break;
    case ORC_CARSTAR:
        alpha = p->b * rho / 4.0;
        beta = 1 - alpha;
        pr = rho * p->R * T * ((1 + alpha + alpha * alpha - alpha * alpha * alpha) / (beta * beta * beta))
             - p->a * rho * rho;
        break;
    case ORC_PENGROB:
        alpha = 0.170151343 * p->a / (p->b * p->R);
        beta = (1 + (0.37464 + 1.5422 * p->w_pr - 0.26992 * p->w_pr * p->w_pr) * (1 - sqrt(T / alpha)));
        pr = rho * p->R * T / (1 - p->b * rho)
             - p->a * beta * beta * rho * rho / (1 + 2 * p->b * rho - p->b * rho * p->b * rho);
        break;
    case ORC_VDW:
        pr = rho * p->R * T / (p->M - rho * p->b) - p->a * rho * rho / (p->M * p->M);
        break;
    default:
        break;
    }
    return pr;
}

double orc_p_eos(const orc_params* p, double rho, double T) { return p_eos(p, rho, T); }

/* lbEquation/pseudoPot/potential.c:7-13 */
static double potential(const orc_case* c, double rho, double T) {
    return sqrt(2 * (p_eos(&c->p, rho, T) - rho * c->lat.cs2) / (c->p.G));
}

/* ------------------------------------------------------------------------- */
/* Topology: applications/mesh/latticeBox2D/latticeBox2D.c:105-239 and
 * latticeBox3D/latticeBox3D.c:108-227 (neighbour of node x along k is the upstream
 * node x - c_k, checked by coordinate distance on the box surface).             */
static void build_nb(orc_case* c) {
    const orc_lattice* L = &c->lat;
    const int nx = c->p.nx, ny = c->p.ny, nz = c->lat.d == 2 ? 1 : c->p.nz;
    const int Q = L->Q;
    const int64_t N = c->N;
    int* nb = c->nb;
    for (int64_t i = 0; i < N * Q; i++) nb[i] = -1;
    if (L->d == 2) {
        for (int j = 1; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++) {
                int id = i + j * nx;
                for (int v = 0; v < Q; v++)
                    nb[id * Q + v] = id + L->vel[L->rev[v]][0] + L->vel[L->rev[v]][1] * nx;
            }
        /* rows j = 0 and j = ny-1 (latticeBox2D.c:186-211), then columns (214-239) */
        for (int pass = 0; pass < 2; pass++) {
            int jlo = pass == 0 ? 0 : 1, jhi = pass == 0 ? ny : ny - 1;
            int jst = pass == 0 ? (ny - 1) : 1;
            if (jst == 0) jst = 1;
            for (int j = jlo; j < jhi; j += jst) {
                int ilo = 0, ist = pass == 0 ? 1 : (nx - 1);
                if (ist == 0) ist = 1;
                for (int i = ilo; i < nx; i += ist) {
                    int id = i + j * nx;
                    for (int v = 0; v < Q; v++) {
                        int nid = id + L->vel[L->rev[v]][0] + L->vel[L->rev[v]][1] * nx;
                        if (nid >= 0 && nid <= nx * ny - 1) {
                            int pi = id % nx, pj = id / nx, qi = nid % nx, qj = nid / nx;
                            if (abs(pi - qi) <= 1 && abs(pj - qj) <= 1) nb[id * Q + v] = nid;
                        }
                    }
                }
            }
        }
    } else {
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) {
                    int64_t id = i + (int64_t)j * nx + (int64_t)k * nx * ny;
                    int onb = (i == 0) || (i == nx - 1) || (j == 0) || (j == ny - 1) || (k == 0) || (k == nz - 1);
                    for (int v = 0; v < Q; v++) {
                        const int* cr = L->vel[L->rev[v]];
                        int64_t nid = id + cr[0] + (int64_t)cr[1] * nx + (int64_t)cr[2] * nx * ny;
                        if (!onb) { nb[id * Q + v] = (int)nid; continue; }
                        if (nid >= 0 && nid <= N - 1) {
                            int qi = (int)(nid % nx), qj = (int)((nid / nx) % ny), qk = (int)(nid / ((int64_t)nx * ny));
                            if (abs(i - qi) <= 1 && abs(j - qj) <= 1 && abs(k - qk) <= 1) nb[id * Q + v] = (int)nid;
                        }
                    }
                }
    }
}

static void copy_missing(int* nb, int Q, int64_t a, int64_t b, int vfirst) {
    for (int v = vfirst; v < Q; v++) {
        if (nb[a * Q + v] == -1) nb[a * Q + v] = nb[b * Q + v];
        if (nb[b * Q + v] == -1) nb[b * Q + v] = nb[a * Q + v];
    }
}

static void fix_corners(int* nb, int Q, const int64_t cid[4]) {
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            if (i != j)
                for (int v = 1; v < Q; v++)
                    if (nb[cid[i] * Q + v] == -1 && nb[cid[j] * Q + v] != -1) nb[cid[i] * Q + v] = nb[cid[j] * Q + v];
}

static int* bd_alloc(orc_case* c, int bid, int n) {
    c->bdsize[bid] = n;
    c->bd[bid] = (int*)calloc(n > 0 ? n : 1, sizeof(int));
    return c->bd[bid];
}

/* Boundary lists and periodic links: latticeBox2D/{genericBoundary,periodicX,periodicY,
 * periodicXY}.c and latticeBox3D/{genericBoundary3D,periodicXY3D,periodicZ3D,periodicXYZ3D,
 * periodicXY3DSpot}.c.  Known reference bugs are reproduced where their result is
 * well defined (periodicY/genericBoundary use (nx-1)*ny; periodicZ3D uses j*ny).      */
static void build_boundaries(orc_case* c) {
    const int nx = c->p.nx, ny = c->p.ny, nz = c->lat.d == 2 ? 1 : c->p.nz, Q = c->lat.Q;
    int* nb = c->nb;
    const int64_t nxy = (int64_t)nx * ny;
    c->nbd = 0;
    if (c->lat.d == 2) {
        switch (c->p.btype) {
        case ORC_GENERIC: {              /* genericBoundary.c:6-49 */
            c->nbd = 4;
            int *x0 = bd_alloc(c, 0, ny), *x1 = bd_alloc(c, 1, ny);
            int *y0 = bd_alloc(c, 2, nx - 2), *y1 = bd_alloc(c, 3, nx - 2);
            for (int j = 0; j < ny; j++) { x0[j] = j * nx; x1[j] = (nx - 1) + j * nx; }
            for (int j = 1; j < nx - 1; j++) { y0[j - 1] = j; y1[j - 1] = j + (nx - 1) * ny; }
            break;
        }
        case ORC_PERIODIC_X: {           /* periodicX.c:6-73 (effective lists j = 1..ny-2) */
            c->nbd = 4;
            int *x0 = bd_alloc(c, 0, ny - 2), *x1 = bd_alloc(c, 1, ny - 2);
            int *y0 = bd_alloc(c, 2, nx), *y1 = bd_alloc(c, 3, nx);
            for (int j = 0; j < ny; j++) {
                if (j >= 1 && j <= ny - 2) { x0[j - 1] = j * nx; x1[j - 1] = (nx - 1) + j * nx; }
                copy_missing(nb, Q, (int64_t)j * nx, (int64_t)nx - 1 + (int64_t)j * nx, 0);
            }
            for (int j = 0; j < nx; j++) { y0[j] = j; y1[j] = j + (ny - 1) * nx; }
            break;
        }
        case ORC_PERIODIC_Y: {           /* periodicY.c:6-71 */
            if ((int64_t)(nx - 1) * ny + (nx - 1) >= nxy) {
                orc_fail(c, "Boundary list leaves the lattice for this box shape");
                return;
            }
            c->nbd = 4;
            int *x0 = bd_alloc(c, 0, ny - 2), *x1 = bd_alloc(c, 1, ny - 2);
            int *y0 = bd_alloc(c, 2, nx), *y1 = bd_alloc(c, 3, nx);
            for (int j = 1; j < ny - 1; j++) { x0[j - 1] = j * nx; x1[j - 1] = (nx - 1) + j * nx; }
            for (int j = 0; j < nx; j++) {
                y0[j] = j; y1[j] = j + (ny - 1) * nx;
                copy_missing(nb, Q, j, (int64_t)j + (int64_t)(nx - 1) * ny, 0);
            }
            break;
        }
        case ORC_PERIODIC_XY: {          /* periodicXY.c:6-101 */
            c->nbd = 0;
            for (int j = 1; j < ny - 1; j++) copy_missing(nb, Q, (int64_t)j * nx, (int64_t)nx - 1 + (int64_t)j * nx, 0);
            for (int j = 1; j < nx - 1; j++) copy_missing(nb, Q, j, (int64_t)j + (int64_t)(ny - 1) * nx, 0);
            int64_t cid[4] = {0, nx - 1, (int64_t)nx * (ny - 1), (int64_t)nx * ny - 1};
            fix_corners(nb, Q, cid);
            break;
        }
        default:
            orc_fail(c, "Unrecognized boundary type");
        }
        /* the (nx-1)*ny indexing leaves the lattice when nx > ny: refused, as the
         * reference would write past its arrays */
        for (int b = 0; b < c->nbd; b++)
            for (int i = 0; i < c->bdsize[b]; i++)
                if (c->bd[b][i] < 0 || c->bd[b][i] >= nxy) {
                    orc_fail(c, "Boundary list leaves the lattice for this box shape");
                    for (int q = 0; q < c->nbd; q++) {
                        free(c->bd[q]);
                        c->bd[q] = NULL;
                        c->bdsize[q] = 0;
                    }
                    c->nbd = 0;
                    return;
                }
        return;
    }
    switch (c->p.btype) {
    case ORC_GENERIC: {                  /* genericBoundary3D.c:6-93 */
        c->nbd = 6;
        int *x0 = bd_alloc(c, 0, ny * nz), *x1 = bd_alloc(c, 1, ny * nz);
        int *y0 = bd_alloc(c, 2, (nx - 2) * nz), *y1 = bd_alloc(c, 3, (nx - 2) * nz);
        int *z0 = bd_alloc(c, 4, (nx - 2) * (ny - 2)), *z1 = bd_alloc(c, 5, (nx - 2) * (ny - 2));
        int id = 0;
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++) { x0[id] = j * nx + k * nxy; x1[id] = (nx - 1) + j * nx + k * nxy; id++; }
        id = 0;
        for (int k = 0; k < nz; k++)
            for (int i = 1; i < nx - 1; i++) { y0[id] = i + k * nxy; y1[id] = i + (ny - 1) * nx + k * nxy; id++; }
        id = 0;
        for (int j = 1; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++) { z0[id] = i + j * nx; z1[id] = i + j * nx + (nz - 1) * nxy; id++; }
        break;
    }
    case ORC_PERIODIC_XY:                /* periodicXY3D.c:6-226 */
    case ORC_PERIODIC_XY_SPOT: {         /* periodicXY3DSpot.c:7-313 */
        if (c->p.btype == ORC_PERIODIC_XY) {
            c->nbd = 2;
            int *z0 = bd_alloc(c, 0, nx * ny), *z1 = bd_alloc(c, 1, nx * ny);
            int k = 0;
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) { z0[k] = i + j * nx; z1[k] = i + j * nx + (nz - 1) * nxy; k++; }
        } else {
            c->nbd = 3;
            int count = 0;
            double rr = c->p.spot_radius * c->p.spot_radius;
            for (int i = 0; i < nx; i++)
                for (int j = 0; j < ny; j++) {
                    int x0 = (int)(c->p.spot_centre[0] - (int)i), y0 = (int)(c->p.spot_centre[1] - (int)j);
                    if ((x0 * x0 + y0 * y0) <= rr) count++;
                }
            int *z0 = bd_alloc(c, 0, nx * ny - count), *z1 = bd_alloc(c, 1, nx * ny), *sp = bd_alloc(c, 2, count);
            int k = 0, k0 = 0, ks = 0;
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) {
                    z1[k] = i + j * nx + (nz - 1) * nxy;
                    int x0 = (int)(c->p.spot_centre[0] - (int)i), y0 = (int)(c->p.spot_centre[1] - (int)j);
                    if ((x0 * x0 + y0 * y0) <= rr) sp[ks++] = i + j * nx;
                    else z0[k0++] = i + j * nx;
                    k++;
                }
        }
        for (int k = 1; k < nz - 1; k++)
            for (int j = 0; j < ny; j++) copy_missing(nb, Q, j * nx + k * nxy, j * nx + k * nxy + nx - 1, 0);
        for (int k = 1; k < nz - 1; k++)
            for (int i = 0; i < nx; i++) copy_missing(nb, Q, i + k * nxy, i + k * nxy + (int64_t)(ny - 1) * nx, 0);
        int64_t cid[4] = {0, nx - 1, (int64_t)nx * (ny - 1), nxy - 1};
        fix_corners(nb, Q, cid);
        int64_t cz[4] = {(nz - 1) * nxy, nx - 1 + (nz - 1) * nxy, (int64_t)nx * (ny - 1) + (nz - 1) * nxy, nxy - 1 + (nz - 1) * nxy};
        fix_corners(nb, Q, cz);
        break;
    }
    case ORC_PERIODIC_Z: {               /* periodicZ3D.c:6-251; well defined only for nz == nx */
        if (nz != nx) { orc_fail(c, "periodicZ boundary lists are only defined for Nz == Nx"); return; }
        /* Y-list entries beyond nz*(nx-2) stay unset for ny > nz; node j*ny + (nz-1)*nx*ny
         * leaves the lattice for ny >= nx + 2 */
        if (ny > nz) { orc_fail(c, "periodicZ boundary lists are only defined for Ny <= Nz"); return; }
        c->nbd = 4;
        int *x0 = bd_alloc(c, 0, nx * ny), *x1 = bd_alloc(c, 1, nx * ny);
        int *y0 = bd_alloc(c, 2, (nx - 2) * ny), *y1 = bd_alloc(c, 3, (nx - 2) * ny);
        int i = 0;
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++) { x0[i] = j * nx + k * nxy; x1[i] = (nx - 1) + j * nx + k * nxy; i++; }
        int jj = 0;
        for (int k = 0; k < nz; k++)
            for (int ii = 1; ii < nx - 1; ii++) {
                if (jj < (nx - 2) * ny) { y0[jj] = ii + k * nxy; y1[jj] = ii + (ny - 1) * nx + k * nxy; }
                jj++;
            }
        for (int j = 1; j < ny - 1; j++)
            for (int ii = 1; ii < nx - 1; ii++) copy_missing(nb, Q, ii + j * nx, ii + j * nx + (nz - 1) * nxy, 0);
        for (int ii = 0; ii < nx; ii++) {
            copy_missing(nb, Q, ii, ii + (nz - 1) * nxy, 0);
            copy_missing(nb, Q, ii + (ny - 1) * nx, ii + (ny - 1) * nx + (nz - 1) * nxy, 0);
        }
        for (int j = 1; j < ny - 1; j++) {
            copy_missing(nb, Q, (int64_t)j * ny, (int64_t)j * ny + (nz - 1) * nxy, 0);
            copy_missing(nb, Q, (nx - 1) + (int64_t)j * nx, (nx - 1) + (int64_t)j * nx + (nz - 1) * nxy, 0);
        }
        break;
    }
    case ORC_PERIODIC_XYZ: {             /* periodicXYZ3D.c:6-159 */
        c->nbd = 0;
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++) copy_missing(nb, Q, j * nx + k * nxy, j * nx + k * nxy + nx - 1, 0);
        for (int k = 0; k < nz; k++)
            for (int i = 0; i < nx; i++) copy_missing(nb, Q, i + k * nxy, i + k * nxy + (int64_t)(ny - 1) * nx, 0);
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) copy_missing(nb, Q, i + (int64_t)j * nx, i + (int64_t)j * nx + (nz - 1) * nxy, 0);
        break;
    }
    default:
        orc_fail(c, "Unrecognized boundary type");
    }
}

/* ------------------------------------------------------------------------- */
orc_case* orc_create(const orc_params* p) {
    orc_case* c = (orc_case*)calloc(1, sizeof(orc_case));
    c->p = *p;
    set_lattice(&c->lat, p->lbmodel);
    const int nz = c->lat.d == 2 ? 1 : p->nz;
    c->N = (int64_t)p->nx * p->ny * nz;
    const int Q = c->lat.Q;
    const int64_t N = c->N;
    c->nb = (int*)malloc(sizeof(int) * N * Q);
    build_nb(c);
    build_boundaries(c);
    c->isOnBnd = (unsigned*)malloc(sizeof(unsigned) * N);
    for (int64_t i = 0; i < N; i++) {                    /* readLatticeMesh.c:263-295 */
        unsigned m = 0;
        for (int k = 0; k < Q; k++) if (c->nb[i * Q + k] == -1) m++;
        c->isOnBnd[i] = m ? 1 : 0;
    }
    c->rho = (double*)calloc(N, sizeof(double));
    c->T = (double*)calloc(N, sizeof(double));
    c->p_ = (double*)calloc(N, sizeof(double));
    c->ss = (double*)calloc(N, sizeof(double));
    c->U = (double*)calloc(N * 3, sizeof(double));
    c->Fi = (double*)calloc(N * 3, sizeof(double));
    c->Ft = (double*)calloc(N * 3, sizeof(double));
    c->f = (double*)calloc(N * Q, sizeof(double));
    c->g = (double*)calloc(N * Q, sizeof(double));
    c->fswap = (double*)calloc(N * Q, sizeof(double));
    c->gswap = (double*)calloc(N * Q, sizeof(double));
    /* multiPhasePP.c:104-116: field update flags from --frozen / --nht */
    c->f_update = (p->frozen == 0) ? 0 : 1;
    c->g_update = (p->ht == 0) ? 0 : 1;
    c->rho0 = p->rho0;
    c->current = p->start;                              /* readTimeInfo.c */
    c->stp = 0;
    for (int b = 0; b < ORC_MAXBD; b++) { c->frdt_local[0][b] = p->f_localT0[b]; c->frdt_local[1][b] = p->g_localT0[b]; }
#ifdef _OPENMP
    if (p->nthreads > 0) omp_set_num_threads(p->nthreads);
#endif
    return c;
}

void orc_destroy(orc_case* c) {
    if (!c) return;
    free(c->nb); free(c->isOnBnd);
    for (int b = 0; b < ORC_MAXBD; b++) free(c->bd[b]);
    free(c->rho); free(c->T); free(c->p_); free(c->ss); free(c->U); free(c->Fi); free(c->Ft);
    free(c->f); free(c->g); free(c->fswap); free(c->gswap);
    free(c);
}

int orc_error(const orc_case* c) { return c->err; }
const char* orc_error_msg(const orc_case* c) { return c->errmsg; }
int64_t orc_npoints(const orc_case* c) { return c->N; }
int orc_Q(const orc_case* c) { return c->lat.Q; }
int orc_nbd(const orc_case* c) { return c->nbd; }
int orc_bd_size(const orc_case* c, int bid) { return c->bdsize[bid]; }
void orc_get_nb(const orc_case* c, int32_t* out) { for (int64_t i = 0; i < c->N * c->lat.Q; i++) out[i] = c->nb[i]; }
void orc_get_bd(const orc_case* c, int bid, int32_t* out) { for (int i = 0; i < c->bdsize[bid]; i++) out[i] = c->bd[bid][i]; }
double orc_get_rho0(const orc_case* c) { return c->rho0; }
void orc_set_rho0(orc_case* c, double r) { c->rho0 = r; }
int orc_current_time(const orc_case* c) { return c->current; }

static double* field_ptr(const orc_case* c, const char* name, int64_t* n) {
    const int64_t N = c->N, Q = c->lat.Q;
    if (!strcmp(name, "rho")) { *n = N; return c->rho; }
    if (!strcmp(name, "T")) { *n = N; return c->T; }
    if (!strcmp(name, "p")) { *n = N; return c->p_; }
    if (!strcmp(name, "ss")) { *n = N; return c->ss; }
    if (!strcmp(name, "U")) { *n = 3 * N; return c->U; }
    if (!strcmp(name, "Fi")) { *n = 3 * N; return c->Fi; }
    if (!strcmp(name, "Ft")) { *n = 3 * N; return c->Ft; }
    if (!strcmp(name, "f")) { *n = Q * N; return c->f; }
    if (!strcmp(name, "g")) { *n = Q * N; return c->g; }
    *n = 0;
    return NULL;
}

int orc_set_field(orc_case* c, const char* name, const double* v) {
    int64_t n;
    double* d = field_ptr(c, name, &n);
    if (!d) return -1;
    memcpy(d, v, sizeof(double) * n);
    return 0;
}

int orc_get_field(const orc_case* c, const char* name, double* v) {
    int64_t n;
    double* d = field_ptr(c, name, &n);
    if (!d) return -1;
    memcpy(v, d, sizeof(double) * n);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Equilibria.
 * liMRTEquilibrium.c:3-29 / xuMRTEquilibrium.c:3-29 (identical second-order form). */
static void ns_equilibrium(const orc_lattice* L, double rho, const double v[3], double* f) {
    for (int k = 0; k < L->Q; k++) {
        double alpha = 0, beta = 0;
        for (int j = 0; j < 3; j++) {
            alpha += L->vel[k][j] * v[j];
            beta += v[j] * v[j];
        }
        f[k] = rho * L->omega[k] * (1 + alpha / L->cs2 + 0.5 * alpha * alpha / (L->cs2 * L->cs2) - 0.5 * beta / L->cs2);
    }
}

/* myMRTEquilibrium.c:5-95 */
static void my_equilibrium(const orc_lattice* L, double T, const double U[3], double a1, double a2, double* f) {
    double n_eq[15];
    memset(n_eq, 0, sizeof(n_eq));
    n_eq[0] = T;
    n_eq[1] = a1 * T;
    n_eq[2] = a2 * T;
    if (L->model == ORC_D2Q9) {
        n_eq[3] = T * U[0];
        n_eq[4] = T * (-U[0]);
        n_eq[5] = T * U[1];
        n_eq[6] = T * (-U[1]);
    } else {
        n_eq[3] = T * U[0];
        n_eq[4] = -(T * U[0]);
        n_eq[5] = T * U[1];
        n_eq[6] = -(T * U[1]);
        n_eq[7] = T * U[2];
        n_eq[8] = -(T * U[2]);
    }
    matvec(L->invM, n_eq, f, L->Q);
}

/* liTempMRTEquilibrium.c:5-44 (D2Q9 only; leaves f untouched otherwise) */
static void litemp_equilibrium(const orc_lattice* L, double T, const double U[3], double* f) {
    if (L->model != ORC_D2Q9) return;
    double n_eq[15];
    n_eq[0] = T;
    n_eq[1] = -2.0 * T;
    n_eq[2] = 2.0 * T;
    n_eq[3] = T * U[0];
    n_eq[4] = T * (-U[0]);
    n_eq[5] = T * U[1];
    n_eq[6] = T * (-U[1]);
    n_eq[7] = 0;
    n_eq[8] = 0;
    matvec(L->invM, n_eq, f, L->Q);
}

static int field_model(const orc_case* c, int field) { return field == 0 ? c->p.f_model : c->p.g_model; }
static double* field_data(const orc_case* c, int field) { return field == 0 ? c->f : c->g; }
static int field_update(const orc_case* c, int field) { return field == 0 ? c->f_update : c->g_update; }

/* lbEquation/general/equilibrium.c:6-90 */
void orc_equilibrium(orc_case* c, int field) {
    if (!field_update(c, field)) return;
    const int Q = c->lat.Q;
    double* fv = field_data(c, field);
    const int model = field_model(c, field);
    const double a1 = field == 0 ? 0 : c->p.g_alpha1, a2 = field == 0 ? 0 : c->p.g_alpha2;
    #pragma omp parallel for schedule(static)
    for (int64_t id = 0; id < c->N; id++) {
        switch (model) {
        case ORC_LIMRT:
        case ORC_XUMRT: ns_equilibrium(&c->lat, c->rho[id], &c->U[id * 3], &fv[id * Q]); break;
        case ORC_MYMRT: my_equilibrium(&c->lat, c->T[id], &c->U[id * 3], a1, a2, &fv[id * Q]); break;
        case ORC_LITEMPMRT: litemp_equilibrium(&c->lat, c->T[id], &c->U[id * 3], &fv[id * Q]); break;
        default: break;
        }
    }
    if (model < 0 || model > 3) orc_fail(c, "Unrecognized collision model in equilibrium distribution");
}

/* ------------------------------------------------------------------------- */
/* interForce.c:7-52 + interactionForce.c:8-84 */
void orc_inter_force(orc_case* c) {
    const orc_lattice* L = &c->lat;
    const int Q = L->Q;
    #pragma omp parallel for schedule(static)
    for (int64_t id = 0; id < c->N; id++) {
        double Ff[3] = {0, 0, 0};
        double F[3] = {0.0, 0.0, 0.0};
        if (c->isOnBnd[id] == 0) {
            for (int k = 1; k < Q; k++) {
                int nid = c->nb[id * Q + L->rev[k]];
                double alpha = L->w[k] * potential(c, c->rho[nid], c->T[nid]);
                for (int i = 0; i < 3; i++) F[i] += alpha * (double)L->vel[k][i];
            }
            double beta = -c->p.G * potential(c, c->rho[id], c->T[id]);
            for (int i = 0; i < 3; i++) Ff[i] = F[i] * beta;
        }
        for (int j = 0; j < 3; j++) {
            c->Fi[id * 3 + j] = Ff[j];
            c->Ft[id * 3 + j] = Ff[j] + (c->rho[id] - c->rho0) * c->p.g[j];
            if (c->isOnBnd[id] == 0) c->Ft[id * 3 + j] += c->p.fext[j];
        }
    }
}

/* ------------------------------------------------------------------------- */
/* liMRTSurfaceTension.c:6-154 (D2Q9) */
static void li_surface_tension(const orc_case* c, int64_t id, double* C) {
    const orc_lattice* L = &c->lat;
    const double* Lam = c->p.f_Lambda;
    double psi = potential(c, c->rho[id], c->T[id]);
    int noneigh = 0;
    for (int k = 1; k < L->Q; k++) if (c->nb[id * L->Q + k] == -1) noneigh++;
    if (noneigh == 0) {
        double Qt[2][2] = {{0, 0}, {0, 0}};
        for (int k = 1; k < 9; k++) {
            int nid = c->nb[id * L->Q + L->rev[k]];
            double alpha = L->w[k] * (potential(c, c->rho[nid], c->T[nid]) - psi);
            for (int i = 0; i < 2; i++)
                for (int j = 0; j < 2; j++) Qt[i][j] += alpha * L->vel[k][i] * L->vel[k][j];
        }
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) Qt[i][j] = Qt[i][j] * 0.5 * psi * c->p.G * c->p.f_kappa;
        C[0] = 0;
        C[1] = 1.5 * Lam[1] * (Qt[0][0] + Qt[1][1]);
        C[2] = -1.5 * Lam[2] * (Qt[0][0] + Qt[1][1]);
        C[3] = 0; C[4] = 0; C[5] = 0; C[6] = 0;
        C[7] = -Lam[8] * (Qt[0][0] - Qt[1][1]);
        C[8] = -Lam[8] * Qt[0][1];
    } else {
        for (int k = 0; k < 9; k++) C[k] = 0;
    }
}

/* xuMRTSurfaceTension.c:6-172 (D3Q15) */
static void xu_surface_tension(const orc_case* c, int64_t id, double* C) {
    const orc_lattice* L = &c->lat;
    const double* Lam = c->p.f_Lambda;
    double psi = potential(c, c->rho[id], c->T[id]);
    int noneigh = 0;
    for (int k = 1; k < L->Q; k++) if (c->nb[id * L->Q + k] == -1) noneigh++;
    if (noneigh == 0) {
        double Qt[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int k = 1; k < L->Q; k++) {
            int nid = c->nb[id * L->Q + L->rev[k]];
            double alpha = L->w[k] * (potential(c, c->rho[nid], c->T[nid]) - psi);
            for (int i = 0; i < 3; i++)
                for (int j = i; j < 3; j++) Qt[i][j] += alpha * L->vel[k][i] * L->vel[k][j];
        }
        for (int i = 0; i < 3; i++)
            for (int j = i; j < 3; j++) Qt[i][j] = Qt[i][j] * 0.5 * psi * c->p.G * c->p.f_kappa;
        for (int k = 0; k < 15; k++) C[k] = 0;
        C[1] = 4.0 * Lam[1] * (Qt[0][0] + Qt[1][1] + Qt[2][2]) / 5.0;
        C[9] = -Lam[10] * (2 * Qt[0][0] - Qt[1][1] - Qt[2][2]);
        C[10] = -Lam[10] * (Qt[1][1] - Qt[2][2]);
        C[11] = -Lam[10] * Qt[0][1];
        C[12] = -Lam[10] * Qt[1][2];
        C[13] = -Lam[10] * Qt[0][2];
    } else {
        for (int k = 0; k < L->Q; k++) C[k] = 0;
    }
}

/* liMRTCollision.c:7-136 with liMRTForce.c:10-56 */
static void limrt_collision(orc_case* c) {
    const orc_lattice* L = &c->lat;
    const int Q = L->Q;
    if (Q != 9) { orc_fail(c, "MRT force squeme not implemented for D3Q15"); return; }
    const double* Lam = c->p.f_Lambda;
    const double sigma = c->p.f_sigma;
    #pragma omp parallel for schedule(static)
    for (int64_t id = 0; id < c->N; id++) {
        double m[15], m_eq[15], S[15], C[15];
        const double rho = c->rho[id];
        const double* U = &c->U[id * 3];
        double umag = 0;
        for (int k = 0; k < 3; k++) umag += U[k] * U[k];
        m_eq[0] = rho;
        m_eq[1] = rho * (-2 + 3 * umag);
        m_eq[2] = rho * (1 - 3 * umag);
        m_eq[3] = rho * U[0];
        m_eq[4] = rho * (-U[0]);
        m_eq[5] = rho * U[1];
        m_eq[6] = rho * (-U[1]);
        m_eq[7] = rho * (U[0] * U[0] - U[1] * U[1]);
        m_eq[8] = rho * U[0] * U[1];
        matvec(L->M, &c->f[id * Q], m, Q);
        /* liMRTForce.c:13-42 */
        const double F[3] = {c->Ft[id * 3 + 0], c->Ft[id * 3 + 1], c->Ft[id * 3 + 2]};
        const double* Fi = &c->Fi[id * 3];
        double psi = potential(c, rho, c->T[id]);
        S[0] = 0;
        S[1] = 6 * (U[0] * F[0] + U[1] * F[1]) + 12 * sigma * (Fi[0] * Fi[0] + Fi[1] * Fi[1]) / (psi * psi * ((1 / Lam[1]) - 0.5));
        S[2] = -6 * (U[0] * F[0] + U[1] * F[1]) - 12 * sigma * (Fi[0] * Fi[0] + Fi[1] * Fi[1]) / (psi * psi * ((1 / Lam[2]) - 0.5));
        S[3] = F[0];
        S[4] = -F[0];
        S[5] = F[1];
        S[6] = -F[1];
        S[7] = 2 * (U[0] * F[0] - U[1] * F[1]);
        S[8] = U[0] * F[1] + U[1] * F[0];
        if (c->p.f_st) li_surface_tension(c, id, C);
        for (int k = 0; k < Q; k++) {
            m[k] = m[k] - Lam[k] * (m[k] - m_eq[k]) + (1 - 0.5 * Lam[k]) * S[k];
            if (c->p.f_st) m[k] = m[k] + C[k];
        }
        matvec(L->invM, m, &c->f[id * Q], Q);
    }
}

/* xuMRTCollision.c:8-118 with xuMRTEquilibriumMS.c:4-57 and xuMRTForce.c:6-87 */
static void xumrt_collision(orc_case* c) {
    const orc_lattice* L = &c->lat;
    const int Q = L->Q;
    if (L->model != ORC_D3Q15) { orc_fail(c, "xuMRT is only defined for D3Q15"); return; }
    const double* Lam = c->p.f_Lambda;
    const double sigma = c->p.f_sigma;
    #pragma omp parallel for schedule(static)
    for (int64_t id = 0; id < c->N; id++) {
        double m[15], m_eq[15], S[15], C[15];
        const double U[3] = {c->U[id * 3], c->U[id * 3 + 1], c->U[id * 3 + 2]};
        const double rho = c->rho[id];
        double umag = mag2(U);
        m_eq[0] = rho;
        m_eq[1] = rho * (-1.0 + umag);
        m_eq[2] = rho * (1.0 - 5.0 * umag);
        m_eq[3] = rho * U[0];
        m_eq[4] = rho * -7.0 * U[0] / 3.0;
        m_eq[5] = rho * U[1];
        m_eq[6] = rho * -7.0 * U[1] / 3.0;
        m_eq[7] = rho * U[2];
        m_eq[8] = rho * -7.0 * U[2] / 3.0;
        m_eq[9] = rho * (2.0 * U[0] * U[0] - U[1] * U[1] - U[2] * U[2]);
        m_eq[10] = rho * (U[1] * U[1] - U[2] * U[2]);
        m_eq[11] = rho * U[0] * U[1];
        m_eq[12] = rho * U[1] * U[2];
        m_eq[13] = rho * U[0] * U[2];
        m_eq[14] = 0;
        matvec(L->M, &c->f[id * Q], m, Q);
        /* xuMRTForce.c:18-69: F = Fi + (rho - rho_0) g, also in the sigma-term (parity
         * unpinned where gravity acts: the sigma-term is pinned with g = 0 only) */
        double F[3];
        for (int j = 0; j < 3; j++) F[j] = c->Fi[id * 3 + j] + (rho - c->rho0) * c->p.g[j];
        double psi = potential(c, rho, c->T[id]);
        double uDotF = U[0] * F[0] + U[1] * F[1] + U[2] * F[2];
        S[0] = 0;
        S[1] = 2.0 * uDotF + (6.0 * sigma * mag2(F)) / (psi * psi * ((1 / Lam[1]) - 0.5));
        S[2] = -10.0 * uDotF;
        S[3] = F[0];
        S[4] = -7.0 * F[0] / 3.0;
        S[5] = F[1];
        S[6] = -7.0 * F[1] / 3.0;
        S[7] = F[2];
        S[8] = -7.0 * F[2] / 3.0;
        S[9] = 4.0 * U[0] * F[0] - 2.0 * U[1] * F[1] - 2.0 * U[2] * F[2];
        S[10] = 2.0 * U[1] * F[1] - 2.0 * U[2] * F[2];
        S[11] = U[0] * F[1] + U[1] * F[0];
        S[12] = U[1] * F[2] + U[2] * F[1];
        S[13] = U[0] * F[2] + U[2] * F[0];
        S[14] = 0;
        if (c->p.f_st) xu_surface_tension(c, id, C);
        for (int k = 0; k < Q; k++) {
            m[k] = m[k] - Lam[k] * (m[k] - m_eq[k]) + (1 - 0.5 * Lam[k]) * S[k];
            if (c->p.f_st) m[k] = m[k] + C[k];
        }
        matvec(L->invM, m, &c->f[id * Q], Q);
    }
}

/* liTempMRTCollision.c:8-102 with liTempMRTEquilibriumMS.c:4-55, liTempMRTSource.c:4-60 */
static void litempmrt_collision(orc_case* c, int field) {
    const orc_lattice* L = &c->lat;
    const int Q = L->Q;
    if (L->model != ORC_D2Q9) { orc_fail(c, "Lattice model not available for liTempMRTModel"); return; }
    const double* Lam = field == 0 ? c->p.f_Lambda : c->p.g_Lambda;
    double* fv = field_data(c, field);
    #pragma omp parallel for schedule(static)
    for (int64_t id = 0; id < c->N; id++) {
        double n[15] = {0}, n_eq[15], Gam[15];
        const double T = c->T[id];
        const double U[3] = {c->U[id * 3], c->U[id * 3 + 1], c->U[id * 3 + 2]};
        n_eq[0] = T;
        n_eq[1] = -2.0 * T;
        n_eq[2] = 2.0 * T;
        n_eq[3] = T * U[0];
        n_eq[4] = T * (-U[0]);
        n_eq[5] = T * U[1];
        n_eq[6] = T * (-U[1]);
        n_eq[7] = 0;
        n_eq[8] = 0;
        matvec(L->M, &fv[id * Q], n, Q);
        for (int i = 0; i < Q; i++) Gam[i] = 0;
        Gam[0] = c->ss[id];
        double n4 = n[4], n6 = n[6];
        for (int k = 0; k < Q; k++) n[k] = n[k] - Lam[k] * (n[k] - n_eq[k]) + Gam[k];
        n[3] = n[3] + (1 - Lam[3] / 2) * Lam[4] * n4;
        n[5] = n[5] + (1 - Lam[5] / 2) * Lam[6] * n6;
        matvec(L->invM, n, &fv[id * Q], Q);
    }
}

/* myMRTCollision.c:8-260 with myMRTEquilibriumMS.c:4-83, myMRTSource.c:4-60 and
 * basic/sparseMatVecMult.c:4-30 (diagonal first, then the off-diagonal pairs).     */
static void mymrt_collision(orc_case* c, int field) {
    const orc_lattice* L = &c->lat;
    const int Q = L->Q;
    const double* Lam = field == 0 ? c->p.f_Lambda : c->p.g_Lambda;
    const double a1 = c->p.g_alpha1, a2 = c->p.g_alpha2;
    double* fv = field_data(c, field);
    double Qm[15][15], Qa[15][15];
    int nz;
    int nzp[3][2] = {{3, 4}, {5, 6}, {7, 8}};
    memset(Qm, 0, sizeof(Qm));
    for (int k = 0; k < Q; k++) Qm[k][k] = Lam[k];
    if (L->model == ORC_D2Q9) {
        nz = 2;
        Qm[3][4] = Qm[4][4] * (Qm[3][3] / 2.0 - 1.0);
        Qm[5][6] = Qm[6][6] * (Qm[5][5] / 2.0 - 1.0);
    } else {
        nz = 3;
        Qm[3][4] = Qm[4][4] * (Qm[3][3] / 2.0 - 1.0);
        Qm[5][6] = Qm[6][6] * (Qm[5][5] / 2.0 - 1.0);
        Qm[7][8] = Qm[8][8] * (Qm[7][7] / 2.0 - 1.0);
    }
    for (int j = 0; j < Q; j++)
        for (int k = 0; k < Q; k++) {
            Qa[j][k] = -0.5 * Qm[j][k];
            if (j == k) Qa[j][k] = 1.0 + Qa[j][k];
        }
    #pragma omp parallel for schedule(static)
    for (int64_t id = 0; id < c->N; id++) {
        double n[15], n_eq[15], Gam[15], aux1[15], aux2[15];
        const double T = c->T[id];
        const double U[3] = {c->U[id * 3], c->U[id * 3 + 1], c->U[id * 3 + 2]};
        for (int k = 0; k < 15; k++) n_eq[k] = 0;
        n_eq[0] = T;
        n_eq[1] = a1 * T;
        n_eq[2] = a2 * T;
        if (L->model == ORC_D2Q9) {
            n_eq[3] = T * U[0]; n_eq[4] = T * (-U[0]); n_eq[5] = T * U[1]; n_eq[6] = T * (-U[1]);
        } else {
            n_eq[3] = T * U[0]; n_eq[4] = -(T * U[0]); n_eq[5] = T * U[1]; n_eq[6] = -(T * U[1]);
            n_eq[7] = T * U[2]; n_eq[8] = -(T * U[2]);
        }
        matvec(L->M, &fv[id * Q], n, Q);
        for (int i = 0; i < Q; i++) Gam[i] = 0;
        Gam[0] = c->ss[id];
        for (int k = 0; k < Q; k++) aux2[k] = n[k] - n_eq[k];
        for (int i = 0; i < Q; i++) aux1[i] = Qm[i][i] * aux2[i];
        for (int i = 0; i < nz; i++) { int j = nzp[i][0], k = nzp[i][1]; aux1[j] = aux1[j] + Qm[j][k] * aux2[k]; }
        for (int i = 0; i < Q; i++) aux2[i] = Qa[i][i] * Gam[i];
        for (int i = 0; i < nz; i++) { int j = nzp[i][0], k = nzp[i][1]; aux2[j] = aux2[j] + Qa[j][k] * Gam[k]; }
        for (int k = 0; k < Q; k++) n[k] = n[k] - aux1[k] + aux2[k];
        matvec(L->invM, n, &fv[id * Q], Q);
    }
}

/* lbEquation/general/collision.c:6-67 */
void orc_collision(orc_case* c, int field) {
    if (!field_update(c, field)) return;
    switch (field_model(c, field)) {
    case ORC_LIMRT: limrt_collision(c); break;
    case ORC_LITEMPMRT: litempmrt_collision(c, field); break;
    case ORC_MYMRT: mymrt_collision(c, field); break;
    case ORC_XUMRT: xumrt_collision(c); break;
    default: orc_fail(c, "Collision model is not yet implemented");
    }
}

/* lbEquation/general/lbstream.c:5-72 (three passes: copy, gather, copy back) */
void orc_stream(orc_case* c, int field) {
    if (!field_update(c, field)) return;
    const int Q = c->lat.Q;
    double* v = field_data(c, field);
    double* sw = field == 0 ? c->fswap : c->gswap;
    const int64_t N = c->N;
    #pragma omp parallel for schedule(static)
    for (int64_t id = 0; id < N; id++)
        for (int k = 0; k < Q; k++) sw[id * Q + k] = v[id * Q + k];
    #pragma omp parallel for schedule(static)
    for (int64_t id = 0; id < N; id++)
        for (int k = 0; k < Q; k++) {
            int nid = c->nb[id * Q + k];
            if (nid != -1) sw[id * Q + k] = v[(int64_t)nid * Q + k];
        }
    #pragma omp parallel for schedule(static)
    for (int64_t id = 0; id < N; id++)
        for (int k = 0; k < Q; k++) v[id * Q + k] = sw[id * Q + k];
}

/* ------------------------------------------------------------------------- */
/* Boundary conditions (sequential, in place, list order, k ascending).            */

/* boundary/bounceBack/bounceBack.c:7-77 */
static void bc_bounceback(orc_case* c, double* fv, int bid) {
    const orc_lattice* L = &c->lat;
    const int Q = L->Q;
    for (int i = 0; i < c->bdsize[bid]; i++) {
        int64_t id = c->bd[bid][i];
        for (int k = 0; k < Q; k++) {
            if (c->nb[id * Q + k] == -1) {
                fv[id * Q + k] = fv[id * Q + L->rev[k]];
            } else if (c->nb[id * Q + L->rev[k]] != -1) {
                int axis = (Q == 9) ? (k >= 1 && k <= 4) : (k >= 1 && k <= 6);
                if (axis) {
                    fv[id * Q + k] = 0.5 * fv[id * Q + k] + 0.5 * fv[id * Q + L->rev[k]];
                    fv[id * Q + L->rev[k]] = fv[id * Q + k];
                }
            }
        }
    }
}

/* boundary/fixedT/fixedT.c:15-142 */
static void bc_fixedT(orc_case* c, double* fv, int bid, int model, double Tw) {
    const orc_lattice* L = &c->lat;
    const int Q = L->Q;
    double feq[15];
    for (int i = 0; i < c->bdsize[bid]; i++) {
        int64_t id = c->bd[bid][i];
        for (int k = 0; k < Q; k++) {
            if (c->nb[id * Q + k] != -1) continue;
            int nbid = c->nb[id * Q + L->rev[k]];
            if (nbid == -1) continue;
            double eq_nb, eq_bnd;
            if (model == ORC_MYMRT) my_equilibrium(L, c->T[nbid], &c->U[nbid * 3], c->p.g_alpha1, c->p.g_alpha2, feq);
            else if (model == ORC_LITEMPMRT) litemp_equilibrium(L, c->T[nbid], &c->U[nbid * 3], feq);
            else { orc_fail(c, "Unable to apply boundary condition on temperature"); return; }
            eq_nb = feq[k];
            if (model == ORC_MYMRT) my_equilibrium(L, Tw, &c->U[nbid * 3], c->p.g_alpha1, c->p.g_alpha2, feq);
            else litemp_equilibrium(L, Tw, &c->U[nbid * 3], feq);
            eq_bnd = feq[k];
            fv[id * Q + k] = eq_bnd + (fv[(int64_t)nbid * Q + k] - eq_nb);
        }
    }
}

/* boundary/fixedRandomT/fixedRandomT.c:14-132: srand(pid) on every call (pid = 0 on a
 * single rank), one rand() per missing link while localTStep < psteps.              */
/* parity unpinned: the draw order (srand(pid) on every call, one rand() per missing
 * link in list and link order); the generator itself is pinned against the host libc. */
static void bc_fixedRandomT(orc_case* c, double* fv, int field, int bid, int model, double Tw, double pert, int psteps) {
    const orc_lattice* L = &c->lat;
    const int Q = L->Q;
    if (model != ORC_MYMRT) { orc_fail(c, "Unable to apply boundary condition on temperature"); return; }
    double feq[15];
    srand(0);
    double minT = Tw * (100.0 - pert) / 100.0;
    double maxT = Tw * (100.0 + pert) / 100.0;
    double r = 1;
    int* local = &c->frdt_local[field][bid];
    for (int i = 0; i < c->bdsize[bid]; i++) {
        int64_t id = c->bd[bid][i];
        for (int k = 0; k < Q; k++) {
            if (c->nb[id * Q + k] != -1) continue;
            int nbid = c->nb[id * Q + L->rev[k]];
            if (nbid == -1) continue;
            my_equilibrium(L, c->T[nbid], &c->U[nbid * 3], c->p.g_alpha1, c->p.g_alpha2, feq);
            double eq_nb = feq[k];
            if (*local < psteps) {
                r = (double)rand() / RAND_MAX;
                double Tb = minT + (maxT - minT) * r;
                my_equilibrium(L, Tb, &c->U[nbid * 3], c->p.g_alpha1, c->p.g_alpha2, feq);
            } else {
                my_equilibrium(L, Tw, &c->U[nbid * 3], c->p.g_alpha1, c->p.g_alpha2, feq);
            }
            double eq_bnd = feq[k];
            fv[id * Q + k] = eq_bnd + (fv[(int64_t)nbid * Q + k] - eq_nb);
        }
    }
    (*local)++;
}

/* boundary/fixedU/fixedU.c:7-143 (liMRT only; other models: no-op) */
static void bc_fixedU(orc_case* c, double* fv, int bid, int model, const double Uw[3]) {
    if (model != ORC_LIMRT) return;
    const orc_lattice* L = &c->lat;
    const int Q = L->Q;
    double feq_nb[15], feq_bnd[15];
    for (int i = 0; i < c->bdsize[bid]; i++) {
        int64_t id = c->bd[bid][i];
        for (int k = 0; k < Q; k++) {
            if (c->nb[id * Q + k] != -1) continue;
            int nbid = c->nb[id * Q + L->rev[k]];
            if (nbid == -1) continue;
            double lv[3] = {0, 0, 0};
            double nbrho = 0;
            double F[3];
            for (int jj = 0; jj < 3; jj++) F[jj] = c->Fi[nbid * 3 + jj] + (c->rho[nbid] - c->rho0) * c->p.g[jj];
            for (int jj = 0; jj < 3; jj++)
                for (int kk = 0; kk < Q; kk++) lv[jj] += L->vel[kk][jj] * fv[(int64_t)nbid * Q + kk];
            for (int kk = 0; kk < Q; kk++) nbrho += fv[(int64_t)nbid * Q + kk];
            for (int jj = 0; jj < 3; jj++) lv[jj] = (lv[jj] + F[jj] * 0.5) / nbrho;
            ns_equilibrium(L, nbrho, lv, feq_nb);
            ns_equilibrium(L, nbrho, Uw, feq_bnd);
            fv[id * Q + k] = feq_bnd[k] + (fv[(int64_t)nbid * Q + k] - feq_nb[k]);
        }
    }
}

/* boundary/outflow/outflow.c:12-208 (normal accumulated in float, as the reference: at
 * an outlet corner the normal's components truncate to 0, the node is its own neighbour
 * and keeps its populations) */
static void bc_outflow(orc_case* c, double* fv, int bid) {
    const orc_lattice* L = &c->lat;
    const int Q = L->Q;
    for (int i = 0; i < c->bdsize[bid]; i++) {
        int64_t id = c->bd[bid][i];
        float normal[3] = {0, 0, 0};
        for (int k = 0; k < Q; k++)
            if (c->nb[id * Q + L->rev[k]] == -1)
                for (int j = 0; j < 3; j++) normal[j] += L->vel[k][j];
        float nmag = 0;
        for (int j = 0; j < 3; j++) nmag += normal[j] * normal[j];
        if (nmag != 0) {
            nmag = sqrt(nmag);
            for (int j = 0; j < 3; j++) normal[j] = normal[j] / nmag;
        } else {
            orc_fail(c, "Unable to detect normal in outflow boundary");
            return;
        }
        unsigned ln = 0;
        for (int k = 0; k < Q; k++)
            if (L->vel[k][0] == (int)normal[0] && L->vel[k][1] == (int)normal[1] && L->vel[k][2] == (int)normal[2]) ln = k;
        unsigned neigh = (unsigned)c->nb[id * Q + ln];
        if (neigh != (unsigned)-1) {
            double uAdv = 0;
            double nbv[3] = {0, 0, 0};
            double nbrho = 0;
            double F[3];
            for (int jj = 0; jj < 3; jj++) F[jj] = c->Fi[neigh * 3 + jj] + (c->rho[neigh] - c->rho0) * c->p.g[jj];
            for (int jj = 0; jj < 3; jj++)
                for (int kk = 0; kk < Q; kk++) nbv[jj] += L->vel[kk][jj] * fv[(int64_t)neigh * Q + kk];
            for (int kk = 0; kk < Q; kk++) nbrho += fv[(int64_t)neigh * Q + kk];
            for (int jj = 0; jj < 3; jj++) nbv[jj] = (nbv[jj] + F[jj] * 0.5) / nbrho;
            for (int jj = 0; jj < 3; jj++) uAdv += normal[jj] * nbv[jj];
            for (int k = 0; k < Q; k++)
                if (c->nb[id * Q + k] == -1)
                    fv[id * Q + k] = (fv[id * Q + k] + uAdv * fv[(int64_t)neigh * Q + k]) / (1 + uAdv);
        }
    }
}

/* lbEquation/general/updateBoundaries.c:6-75 */
void orc_update_boundaries(orc_case* c, int field) {
    if (!field_update(c, field)) return;
    double* fv = field_data(c, field);
    const int model = field_model(c, field);
    const int* bt = field == 0 ? c->p.f_bc : c->p.g_bc;
    for (int b = 0; b < c->nbd; b++) {
        const double* val = field == 0 ? c->p.f_bcval[b] : c->p.g_bcval[b];
        switch (bt[b]) {
        case ORC_PERIODIC: break;
        case ORC_BBACK: bc_bounceback(c, fv, b); break;
        case ORC_FU: bc_fixedU(c, fv, b, model, val); break;
        case ORC_FT: bc_fixedT(c, fv, b, model, val[0]); break;
        case ORC_OF: bc_outflow(c, fv, b); break;
        case ORC_FEP: case ORC_UHF: break;
        case ORC_FRDT:
            bc_fixedRandomT(c, fv, field, b, model, val[0],
                            field == 0 ? c->p.f_pert[b] : c->p.g_pert[b],
                            field == 0 ? c->p.f_psteps[b] : c->p.g_psteps[b]);
            break;
        default: orc_fail(c, "Unrecognized boundary condition"); return;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* macroDensity.c:7-75 with liMRTDensity.c / xuMRTDensity.c */
void orc_macro_density(orc_case* c) {
    if (!c->f_update) return;
    const int Q = c->lat.Q;
    if (c->p.f_model != ORC_LIMRT && c->p.f_model != ORC_XUMRT) { orc_fail(c, "Density update not available"); return; }
    #pragma omp parallel for schedule(static)
    for (int64_t id = 0; id < c->N; id++) {
        double rho = 0;
        for (int k = 0; k < Q; k++) rho += c->f[id * Q + k];
        c->rho[id] = rho;
    }
}

/* macroVelocity.c:6-65 with liMRTVelocity.c:6-55 (F = Ft) / xuMRTVelocity.c:6-55
 * (F = Fi + (rho - rho0) g) */
void orc_macro_velocity(orc_case* c) {
    if (!c->f_update) return;
    const orc_lattice* L = &c->lat;
    const int Q = L->Q;
    const int model = c->p.f_model;
    if (model != ORC_LIMRT && model != ORC_XUMRT) { orc_fail(c, "Velocity update not available"); return; }
    #pragma omp parallel for schedule(static)
    for (int64_t id = 0; id < c->N; id++) {
        double F[3], lv[3] = {0, 0, 0};
        for (int j = 0; j < 3; j++)
            F[j] = model == ORC_LIMRT ? c->Ft[id * 3 + j] : c->Fi[id * 3 + j] + (c->rho[id] - c->rho0) * c->p.g[j];
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < Q; k++) lv[j] += L->vel[k][j] * c->f[id * Q + k];
        for (int j = 0; j < 3; j++) lv[j] = (lv[j] + F[j] * 0.5) / c->rho[id];
        for (int j = 0; j < 3; j++) c->U[id * 3 + j] = lv[j];
    }
}

/* macroTemperature.c:7-68 with myMRTTemperature.c:5-39 / liTempMRTTemperature.c:5-26 */
void orc_macro_temperature(orc_case* c) {
    if (!c->g_update) return;
    const int Q = c->lat.Q;
    const int model = c->p.g_model;
    if (model != ORC_MYMRT && model != ORC_LITEMPMRT) { orc_fail(c, "Temperature update not available"); return; }
    #pragma omp parallel for schedule(static)
    for (int64_t id = 0; id < c->N; id++) {
        double T = 0;
        for (int k = 0; k < Q; k++) T += c->g[id * Q + k];
        if (model == ORC_MYMRT) T = T + 0.5 * c->ss[id];
        c->T[id] = T;
    }
}

/* finiteDifference/scalarGradient.c, invScalarGradient.c, vectorDivergence.c: central
 * differences 0.5 (q[a] - q[b]) with a = nb[id][rev axis+], b = nb[id][axis+], and
 * one-sided differences where one neighbour is missing.                              */
static const int fd_axes2[2][2] = {{3, 1}, {4, 2}};
static const int fd_axes3[3][2] = {{2, 1}, {4, 3}, {6, 5}};

static double fd_diff(const orc_case* c, const double* q, int stride, int comp, int64_t id, int a, int b, int inv, int* has) {
    double qa = 0, qb = 0, qi = 0;
    *has = 1;
    if (a != -1) qa = inv ? 1.0 / q[(int64_t)a * stride + comp] : q[(int64_t)a * stride + comp];
    if (b != -1) qb = inv ? 1.0 / q[(int64_t)b * stride + comp] : q[(int64_t)b * stride + comp];
    qi = inv ? 1.0 / q[id * stride + comp] : q[id * stride + comp];
    if (a != -1 && b != -1) return 0.5 * (qa - qb);
    if (a == -1 && b != -1) return (qi - qb);
    if (a != -1 && b == -1) return (qa - qi);
    *has = 0;
    return 0;
}

/* parity unpinned: the time levels (previous step's T and U, this step's rho) follow
 * the call order of multiPhasePP.c:176-240 as read. */
/* heatSource.c:7-202 */
void orc_heat_source(orc_case* c) {
    if (!c->g_update) return;
    const orc_lattice* L = &c->lat;
    const int Q = L->Q;
    const int model = c->p.g_model;
    if (model != ORC_MYMRT && model != ORC_LITEMPMRT) { orc_fail(c, "Heat source not defined"); return; }
    const double* Lam = c->p.g_Lambda;
    const double a1 = c->p.g_alpha1, a2 = c->p.g_alpha2, Cv = c->p.Cv;
    const int naxes = L->d;
    const int current = c->current;
    #pragma omp parallel for schedule(static)
    for (int64_t id = 0; id < c->N; id++) {
        double lambda;
        if (model == ORC_MYMRT) {
            if (L->model == ORC_D2Q9)
                lambda = c->rho[id] * Cv * (1 / Lam[3] - 0.5) * (4.0 + 3.0 * a1 + 2.0 * a2) / 6.0;
            else
                lambda = c->rho[id] * Cv * (1 / Lam[3] - 0.5) * (6.0 + 11.0 * a1 + a2) / 9.0;
        } else {
            lambda = c->rho[id] * Cv * (1 / Lam[3] - 0.5);
        }
        double gradT[3] = {0, 0, 0}, gradR[3] = {0, 0, 0}, divU = 0;
        int has;
        for (int ax = 0; ax < naxes; ax++) {
            int ka = naxes == 2 ? fd_axes2[ax][0] : fd_axes3[ax][0];
            int kb = naxes == 2 ? fd_axes2[ax][1] : fd_axes3[ax][1];
            int a = c->nb[id * Q + ka], b = c->nb[id * Q + kb];
            gradT[ax] = fd_diff(c, c->T, 1, 0, id, a, b, 0, &has);
            gradR[ax] = fd_diff(c, c->rho, 1, 0, id, a, b, 1, &has);
            double dv = fd_diff(c, c->U, 3, ax, id, a, b, 0, &has);
            if (has) divU += dv;
        }
        double dot = 0;
        for (int k = 0; k < 3; k++) dot += gradT[k] * gradR[k];
        double dpdT = p_eos(&c->p, c->rho[id], 1.1 * c->T[id]) - p_eos(&c->p, c->rho[id], 0.9 * c->T[id]);
        dpdT = dpdT / (0.2 * c->T[id]);
        if (model == ORC_MYMRT) {
            c->ss[id] = -lambda * dot / Cv + divU * c->T[id] * (1.0 - dpdT / (c->rho[id] * Cv));
        } else {
            /* parity unpinned: the filter ss = 1.5 s - 0.5 ss_old (heatSource.c:149-165) */
            double oldss = current == 0 ? 0 : c->ss[id];
            double s = -lambda * dot / Cv + divU * c->T[id] * (1.0 - dpdT / (c->rho[id] * Cv));
            c->ss[id] = s + 0.5 * (s - oldss);
        }
    }
}

/* macroPressure.c:8-69 with liMRTPressure.c:6-253 / xuMRTPressure.c:12,350 */
void orc_macro_pressure(orc_case* c) {
    if (!c->f_update) return;
    const int Q = c->lat.Q;
    const int model = c->p.f_model;
    const double G = c->p.G, cs2 = c->lat.cs2, sigma = c->p.f_sigma;
    if (model != ORC_LIMRT && model != ORC_XUMRT) { orc_fail(c, "Pressure update not available"); return; }
    #pragma omp parallel for schedule(static)
    for (int64_t id = 0; id < c->N; id++) {
        double psi = potential(c, c->rho[id], c->T[id]);
        double pr = c->rho[id] * cs2 + 0.5 * G * psi * psi;
        if (model == ORC_LIMRT) {
            double lap = 0;
            if (c->lat.model == ORC_D2Q9) {
                for (int ax = 0; ax < 2; ax++) {
                    int ka = ax == 0 ? 3 : 4, kb = ax == 0 ? 1 : 2;
                    int a = c->nb[id * Q + ka], b = c->nb[id * Q + kb];
                    double pa = 0, pb = 0, pc = 0, pi = potential(c, c->rho[id], c->T[id]);
                    if (a != -1 && b != -1) {
                        pa = potential(c, c->rho[a], c->T[a]);
                        pb = potential(c, c->rho[b], c->T[b]);
                        if (ax == 0) lap = pa + pb - 2 * pi;
                        else lap = lap + pa + pb - 2 * pi;
                    } else if (a == -1 && b != -1) {
                        int cc = c->nb[(int64_t)b * Q + kb];
                        pc = potential(c, c->rho[cc], c->T[cc]);
                        pb = potential(c, c->rho[b], c->T[b]);
                        if (ax == 0) lap = pc - 2 * pb + pi;
                        else lap = lap + pc - 2 * pb + pi;
                    } else if (a != -1 && b == -1) {
                        int cc = c->nb[(int64_t)a * Q + ka];
                        pc = potential(c, c->rho[cc], c->T[cc]);
                        pa = potential(c, c->rho[a], c->T[a]);
                        if (ax == 0) lap = pc - 2 * pa + pi;
                        else lap = lap + pc - 2 * pa + pi;
                    } else if (ax == 0) {
                        lap = 0;
                    }
                }
            }
            pr = pr + G * psi * lap / 12;
            /* parity unpinned: the sign of the one-sided gradient at walls; it reaches no
             * output (only |grad psi|^2 is used, which the linear-potential pin fixes) */
            double grad[3] = {0, 0, 0};
            if (c->lat.model == ORC_D2Q9) {
                for (int ax = 0; ax < 2; ax++) {
                    int ka = ax == 0 ? 3 : 4, kb = ax == 0 ? 1 : 2;
                    int a = c->nb[id * Q + ka], b = c->nb[id * Q + kb];
                    if (a != -1 && b != -1) grad[ax] = 0.5 * (potential(c, c->rho[a], c->T[a]) - potential(c, c->rho[b], c->T[b]));
                    else if (a == -1 && b != -1) grad[ax] = (potential(c, c->rho[id], c->T[id]) - potential(c, c->rho[b], c->T[b]));
                    else if (a != -1 && b == -1) grad[ax] = (potential(c, c->rho[a], c->T[a]) - potential(c, c->rho[id], c->T[id]));
                    else grad[ax] = 0;
                }
                grad[2] = 0.0;
            }
            pr = pr + 2.0 * G * G * sigma * mag2(grad);
        }
        c->p_[id] = pr;
    }
}

/* basic/averageScalarField.c:4-78 (single rank) */
double orc_average_rho(orc_case* c) {
    double avg = 0;
    if (c->rho0 <= 0) {
        avg = 0;
    } else {
        double s = 0;
        for (int64_t i = 0; i < c->N; i++) s += c->rho[i];
        avg = s / (double)(int)c->N;
    }
    return avg;
}

/* ------------------------------------------------------------------------- */
/* multiPhasePP.c:72-147 */
void orc_init(orc_case* c) {
    c->rho0 = orc_average_rho(c);
    orc_equilibrium(c, 0);
    orc_equilibrium(c, 1);
    orc_inter_force(c);
    orc_heat_source(c);
}

/* IO/time/timeOptions.c:35-64 */
static int update_time(orc_case* c) {
    int upd = 1;
    ++c->current;
    ++c->stp;
    if (c->current > c->p.end) upd = 0;
    if (c->stp > c->p.write_interval) c->stp = 1;
    return upd;
}

static int has_nan(const double* v, int64_t n) {
    for (int64_t i = 0; i < n; i++) if (isnan(v[i])) return 1;
    return 0;
}

/* multiPhasePP.c:157-318 */
int orc_run(orc_case* c, int nsteps) {
    int done = 0;
    const int64_t N = c->N, Q = c->lat.Q;
    while (done < nsteps && update_time(c)) {
        orc_collision(c, 0);
        orc_stream(c, 0);
        orc_update_boundaries(c, 0);
        orc_collision(c, 1);
        orc_stream(c, 1);
        orc_update_boundaries(c, 1);
        orc_macro_density(c);
        if (c->p.ht != 0) orc_heat_source(c);
        orc_macro_temperature(c);
        if (c->p.frozen != 0) orc_inter_force(c);
        orc_macro_velocity(c);
        done++;
        if (c->err) return -2;
        if (c->stp == c->p.write_interval) {          /* timeOptions.c:18-30 writeFlag */
            orc_macro_pressure(c);
            c->rho0 = orc_average_rho(c);
            if (has_nan(c->rho, N) || has_nan(c->T, N) || has_nan(c->p_, N) || has_nan(c->U, 3 * N) ||
                has_nan(c->f, Q * N) || has_nan(c->g, Q * N)) {
                orc_fail(c, "Floating point exception. NaN solution");
                return -1;
            }
        }
    }
    return done;
}
