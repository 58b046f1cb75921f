/*
 * mplb_oracle.h -- CPU oracle for the multiphaseLB per-timestep lattice update.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64, array-of-structures
 * restatement of the reference solver's hot path (multiPhasePP time loop and the
 * routines it calls), written to follow the reference's order of floating-point
 * operations term by term so that it can serve as the parity checker of the HIP
 * library.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product library never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference's src/ directory).  Pinning of each part (golden checks) is
 * listed in the header comment of mplb_oracle.c.
 */
#ifndef MPLB_ORACLE_H
#define MPLB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* lattice models: latticeModel/latticeInfo.h DdQq */
enum { ORC_D2Q9 = 0, ORC_D3Q15 = 1 };
/* lbeModel.h */
enum { ORC_LIMRT = 0, ORC_MYMRT = 1, ORC_XUMRT = 2, ORC_LITEMPMRT = 3, ORC_NOMODEL = -1 };
/* EOS/EOSInfo.h eosType */
enum { ORC_IDEAL = 0, ORC_CARSTAR = 1, ORC_PENGROB = 2, ORC_VDW = 3 };
/* boundary/bdType.h */
enum { ORC_BBACK = 0, ORC_FT = 1, ORC_FRDT = 2, ORC_FGT = 3, ORC_FU = 4,
       ORC_UHF = 5, ORC_FEP = 6, ORC_OF = 7, ORC_PERIODIC = 8 };
/* latticeBox2D/3D boundaryType */
enum { ORC_GENERIC = 0, ORC_PERIODIC_X = 1, ORC_PERIODIC_Y = 2, ORC_PERIODIC_XY = 3,
       ORC_PERIODIC_Z = 4, ORC_PERIODIC_XYZ = 5, ORC_PERIODIC_XY_SPOT = 6 };

#define ORC_MAXBD 6

typedef struct {
    int lbmodel;           /* ORC_D2Q9 / ORC_D3Q15 */
    int nx, ny, nz;        /* nz = 1 for D2Q9 */
    int btype;             /* boundaryType */
    double spot_radius, spot_centre[2];  /* periodicXYSpot only */

    int eos;
    double a, b, w_pr, R, G, M, Cv, rho0, g[3], fext[3];

    int f_model;
    double f_Lambda[15], f_sigma, f_kappa;
    int f_st;              /* 0 none, 1 liSurfTen / xuSurfTen */

    int g_model;
    double g_Lambda[15], g_alpha1, g_alpha2;

    int ht;                /* 1 unless --nht   (checkMpArgs.c) */
    int frozen;            /* 1 unless --frozen (checkMpArgs.c: mp.frozen) */

    int f_bc[ORC_MAXBD], g_bc[ORC_MAXBD];
    double f_bcval[ORC_MAXBD][3], g_bcval[ORC_MAXBD][3];
    double g_pert[ORC_MAXBD];
    int g_psteps[ORC_MAXBD], g_localT0[ORC_MAXBD];
    double f_pert[ORC_MAXBD];
    int f_psteps[ORC_MAXBD], f_localT0[ORC_MAXBD];

    int start, end, write_interval;
    int nthreads;          /* OpenMP threads for node loops (results are thread-count independent) */
} orc_params;

typedef struct orc_case orc_case;

orc_case* orc_create(const orc_params* p);
void orc_destroy(orc_case* c);
int orc_error(const orc_case* c);              /* non-zero after a reference errorMsg() path */
const char* orc_error_msg(const orc_case* c);

int64_t orc_npoints(const orc_case* c);
int orc_Q(const orc_case* c);
int orc_nbd(const orc_case* c);
int orc_bd_size(const orc_case* c, int bid);
void orc_get_nb(const orc_case* c, int32_t* out);            /* [N][Q] */
void orc_get_bd(const orc_case* c, int bid, int32_t* out);   /* bd list */

/* name in rho, T, p, U (N x 3), Fi, Ft, f (N x Q), g, ss (g scalar source) */
int orc_set_field(orc_case* c, const char* name, const double* v);
int orc_get_field(const orc_case* c, const char* name, double* v);
double orc_get_rho0(const orc_case* c);
void orc_set_rho0(orc_case* c, double r);
int orc_current_time(const orc_case* c);

/* solver initialisation, multiPhasePP.c:72-147 */
void orc_init(orc_case* c);
/* the time loop, multiPhasePP.c:157-315; returns number of steps taken, <0 on NaN at a write step */
int orc_run(orc_case* c, int nsteps);

/* fine-grained routines; field 0 = f, 1 = g */
void orc_collision(orc_case* c, int field);
void orc_stream(orc_case* c, int field);
void orc_update_boundaries(orc_case* c, int field);
void orc_macro_density(orc_case* c);
void orc_macro_velocity(orc_case* c);
void orc_macro_temperature(orc_case* c);
void orc_inter_force(orc_case* c);
void orc_heat_source(orc_case* c);
void orc_equilibrium(orc_case* c, int field);
void orc_macro_pressure(orc_case* c);
double orc_average_rho(orc_case* c);

/* lattice constant tables and EOS, for the table and closed-form checks */
void orc_lattice_tables(int model, int* Q, int* d, double* cs2, int* vel /*Q*3*/, int* rev,
                        double* omega, double* w, double* M /*Q*Q*/, double* invM);
double orc_p_eos(const orc_params* p, double rho, double T);

#ifdef __cplusplus
}
#endif

#endif
