/*
 * mplb.h -- C ABI of the MI355X lattice-update library (libmplb.so).
 *
 * Drop-in replacement for the per-timestep path of the multiphaseLB solver: the time
 * loop body of applications/solvers/multiPhasePP/multiPhasePP.c:157-315 and the
 * routines it calls.  A case is described exactly as for the reference (the case
 * directory's properties/{latticeProperties,macroProperties,simulation} and
 * start/boundaries dictionaries, read with the reference's dictIO lookup rules);
 * the lattice tables are those latticeBox2D/latticeBox3D + latticeMeshPartition
 * (one rank) would write.  Plain C types only.  Every call returns 0 on success or
 * a negative MPLB_E* code; mplb_last_error() gives the reference's message text
 * where the reference would have called errorMsg() and exited.
 *
 * Threading: a case handle is used by one host thread at a time.  All calls are
 * synchronous with respect to the host unless stated otherwise.
 */
#ifndef MPLB_H
#define MPLB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPLB_OK 0
#define MPLB_E_ARG (-1)        /* bad argument */
#define MPLB_E_IO (-2)         /* a dictionary or file could not be read */
#define MPLB_E_MODEL (-3)      /* the reference rejects this model/lattice/boundary combination */
#define MPLB_E_HIP (-4)        /* HIP runtime error, or no usable GPU */
#define MPLB_E_NAN (-5)        /* NaN found at a write step (writeScalarField.c:16-24) */
#define MPLB_E_STATE (-6)      /* call out of order (e.g. step before init) */

/* precision of the device state */
#define MPLB_FP64 0            /* parity mode: reference arithmetic, bit-exact */
#define MPLB_FP32 1
#define MPLB_FP16S 2           /* populations stored as FP16 deviations f - w_k rho_ref, g - geq_k(T_ref) (lattice
                                  means of the initial fields), FP32 arithmetic */
#define MPLB_FP16N 3           /* as MPLB_FP16S, but f stored as (f - w_k R(x)) / (w_k R(x)), R(x) the node's own
                                  density when written (per node, FP16): accurate in both phases, 2 B more per node */

/* extension flags (behaviour the reference does not have; all off by default) */
#define MPLB_EXT_LIMRT_D3Q15 0x1   /* accept Collision liMRT on D3Q15 as the xuMRT D3Q15 scheme */
#define MPLB_EXT_NO_FIXEDGRADT_ERROR 0x2  /* treat fixedGradT as a no-op instead of the reference's fatal error */

typedef struct mplb_opts {
    int precision;         /* MPLB_FP64 / MPLB_FP32 / MPLB_FP16S / MPLB_FP16N */
    int heat_transfer;     /* 1 = default, 0 = multiPhasePP --nht */
    int frozen;            /* 0 = default, 1 = multiPhasePP --frozen */
    int fused;             /* 1: use the fused structured-box kernels when the case allows it */
    int extensions;        /* MPLB_EXT_* bits */
    int device;            /* HIP device ordinal; -1 opens the case host-only (dictionaries and tables, no GPU) */
} mplb_opts;

typedef struct mplb_info {
    int lbmodel;           /* 0 D2Q9, 1 D3Q15 */
    int Q, d;
    int nx, ny, nz;
    int64_t npoints;       /* stored nodes (duplicate periodic end nodes included) */
    int nbd;               /* number of boundary lists */
    int f_model, g_model;  /* 0 liMRT, 1 myMRT, 2 xuMRT, 3 liTempMRT, -1 none */
    int start_time, end_time, write_interval, current_time;
    int precision;
    int fused_active;      /* 1 if mplb_step runs the fused kernels */
    double rho0;           /* EOS.rho_0 currently in use */
} mplb_info;

typedef struct mplb_case mplb_case;

/* Default options (FP64, heat transfer on, not frozen, fused off, device 0). */
void mplb_default_opts(mplb_opts* opts);

/* Parse a reference case directory and build the lattice on the device.
 * Replaces readLatticeMesh (IO/readLatticeMesh.c:101-660), setLatticeInfo
 * (latticeModel/setLatticeInfo.c:16-90), readEOSInfo (EOS/readEOSInfo.c:10-111),
 * readLbeField (IO/readLbeField.c:13-259) and latticeBox2D/latticeBox3D. */
int mplb_case_open(const char* case_dir, const mplb_opts* opts, mplb_case** out);
int mplb_case_close(mplb_case* c);
const char* mplb_last_error(const mplb_case* c);
int mplb_case_info(const mplb_case* c, mplb_info* out);
/* The values the case's dictionaries gave, as the library read them (readEOSInfo.c,
 * readLbeField.c, readTimeInfo.c rules): key "EOS/model" (0 idealGas, 1
 * Carnahan-Starling, 2 Peng-Robinson, 3 vanDerWaals), "EOS/a", "EOS/b", "EOS/w" (w_pr),
 * "EOS/R", "EOS/G", "EOS/M", "EOS/Cv", "EOS/rho_0", "EOS/g" (3), "EOS/extForce" (3),
 * "time/start", "time/end", "time/writeInterval", and per field ("f/" or "g/"):
 * "Collision" (mplb_info's model codes), "Lambda" (Q), "sigma", "SurfaceTension"
 * (1 = liSurfTen/xuSurfTen), "kappa_st", "alpha_1", "alpha_2", and "bc/<list>/type"
 * (0 bounceBack, 1 fixedT, 2 fixedRandomT, 3 fixedGradT, 4 fixedU, 5 uniformHeatFlux,
 * 6 fixedEOSp, 7 outflow, 8 periodic), "bc/<list>/value" (3), "bc/<list>/pert",
 * "bc/<list>/duration".  Writes up to n values to out (may be NULL) and returns how many
 * the key has, or MPLB_E_ARG for an unknown key.  No counterpart in the reference (its
 * readers keep the values in their structs); it lets a caller check a case's parse. */
int mplb_case_param(const mplb_case* c, const char* key, double* out, size_t n);

/* Field transfer, double precision host arrays.  name: "rho", "T", "p", "U" (N x 3,
 * node-major as the reference's U[id][j]), "Fi", "Ft", "f", "g" (N x Q, node-major),
 * "ss" (g scalar source).  n is the number of doubles in the host buffer. */
int mplb_set_field(mplb_case* c, const char* name, const double* host, size_t n);
int mplb_get_field(mplb_case* c, const char* name, double* host, size_t n);

/* Lattice tables: nb is N x Q node-major (nb[id*Q + k], upstream node x - c_k or -1),
 * exactly the processor0/lattice/neighbours table of a one-rank case.
 * mplb_get_boundary copies boundary list bid (file order) and its name. */
int mplb_get_neighbours(const mplb_case* c, int32_t* nb, size_t n);
int mplb_get_boundary(const mplb_case* c, int bid, int32_t* nodes, size_t n, char* name, size_t name_len);

/* Solver initialisation, multiPhasePP.c:72-147: rho_0 = averageScalarField(rho),
 * equilibrium(f), equilibrium(g), interForce, heatSource. */
int mplb_init(mplb_case* c);

/* Advance up to nsteps time steps of multiPhasePP.c:157-315 (stops at endTime).
 * At write steps: macroPressure, rho_0 update and the NaN check (MPLB_E_NAN). */
int mplb_step(mplb_case* c, int nsteps, int* steps_done);

/* One routine of the step, for isolated parity checks; field 0 = f, 1 = g. */
int mplb_collision(mplb_case* c, int field);              /* lbEquation/general/collision.c:6-67 */
int mplb_stream(mplb_case* c, int field);                 /* lbEquation/general/lbstream.c:5-72 */
int mplb_update_boundaries(mplb_case* c, int field);      /* lbEquation/general/updateBoundaries.c:6-75 */
int mplb_macro_density(mplb_case* c);                     /* lbEquation/general/macroDensity.c:7-75 */
int mplb_macro_velocity(mplb_case* c);                    /* lbEquation/general/macroVelocity.c:6-65 */
int mplb_macro_temperature(mplb_case* c);                 /* lbEquation/general/macroTemperature.c:7-68 */
int mplb_inter_force(mplb_case* c);                       /* lbEquation/pseudoPot/interForce.c:7-52 */
int mplb_heat_source(mplb_case* c);                       /* lbEquation/general/heatSource.c:7-202 */
int mplb_equilibrium(mplb_case* c, int field);            /* lbEquation/general/equilibrium.c:6-90 */
int mplb_macro_pressure(mplb_case* c);                    /* lbEquation/general/macroPressure.c:8-69 */
/* basic/averageScalarField.c:4-78 on rho (0 when rho_0 <= 0, as the reference). */
int mplb_average_density(mplb_case* c, double* out);

/* Timing: runs nsteps steps and reports the elapsed device time of the whole region
 * and of each kernel class measured with HIP events on the library's stream.
 * kernel_ms has room for MPLB_NTIMERS doubles (average ms per launch), names via
 * mplb_timer_name(). */
#define MPLB_NTIMERS 8
int mplb_time_steps(mplb_case* c, int nsteps, double* total_ms, double* kernel_ms, int* kernel_launches);
/* (returns the number of steps taken, or a negative MPLB_E* code) */
const char* mplb_timer_name(int i);

/* Constant tables of the lattice models (for the bit-exact table check). */
int mplb_lattice_tables(int lbmodel, int* vel, int* rev, double* omega, double* w, double* M, double* invM);

/* glibc-compatible rand() stream used to precompute the fixedRandomT draws
 * (fixedRandomT.c:37,81): writes n successive values after srand(seed). */
int mplb_glibc_rand(unsigned seed, int n, int32_t* out);

/* Files in the reference's formats.  A case directory holding the processor0/lattice files
 * (written by latticeBox2D/3D + latticeMeshPartition for one rank) is opened from those
 * tables, as the solver's readLatticeMesh.c does; otherwise the box is built from
 * properties/latticeProperties. */
/* processor0/lattice/{points,ghosts,neighbours,recvGhosts,sendGhosts,boundary,vtkCells}
 * under dir (IO/writeLatticeMesh.c:20-226, one rank); with dataFormat ensight also
 * dir/lattice.geo and dir/lbm.case first (writeLatticeMesh.c:17-22,
 * IO/ensight/writeMeshToEnsight.c, updateCaseFile.c). */
int mplb_write_lattice(const mplb_case* c, const char* dir);
/* rho, T, p, U, f, g of the current time level into dir/processor0/<time>/ in the case's
 * dataFormat (multiPhasePP.c:276-309 with writeScalarField.c / writeVectorField.c,
 * including their NaN check: MPLB_E_NAN): asciiRaw or binaryRaw, one file per field; or
 * pvtu (IO/pvtu/writeMeshToPvtu.c, writeScalarToPvtu.c, writeArrayToPvtu.c,
 * writePvtuExtra.c, writeMainPvd.c): processor0/<time>/fields.vtu, <time>/fields.pvtu and
 * case.pvd under dir, with the field names of start/initialFields; or binary EnSight Gold
 * (multiPhasePP.c:283-285 updateCaseFile, IO/ensight/write{Scalar,Vector,LbeField}To
 * Ensight.c): the current time appended to dir/lbm.case, then dir/lattice.<name>_<index>
 * per field (f and g: lattice.f<k>_<index> per population), index = the time's position
 * in the case file's list. */
int mplb_write_fields(mplb_case* c, const char* dir);
/* Field <name> of time level <time> under dir, comps values per node, in the case's
 * dataFormat; asciiRaw/binaryRaw values pass through float as the reference's readers do
 * (readScalarFromAsciiRaw.c); pvtu values are read in double (readScalarFromPvtu.c /
 * readArrayFromPvtu.c); ensight files are float32, found through lbm.case's time list
 * (IO/ensight/read*FromEnsight.c, timeToIndex.c: index 0 for a time not listed). */
int mplb_read_field(const mplb_case* c, const char* dir, int time, const char* name, int comps, double* out,
                    size_t n);

/* ---------------------------------------------------------------------------------
 * Batched ensembles: many independent small cases advanced together on one GPU.
 * Replaces running multiPhasePP (multiPhasePP.c:157-241) once per case: each case
 * keeps the reference's time loop with --nht, and the batch advances every case in
 * the same launches (a work table of (case, node range) per workgroup), captures runs
 * of steps in a HIP graph and drops each case from the table once it is steady.
 * Cases: D2Q9 periodicXY boxes, f Collision liMRT without liSurfTen, EOS/rho_0 <= 0,
 * opened with heat_transfer = 0 and frozen = 0; all cases share the EOS block and the
 * f model constants (Lambda, sigma); sizes differ freely.  The same case handle may be
 * passed for several entries (it is read for its mesh and dictionaries only; a
 * host-only handle, opts.device = -1, suffices).  Temperature is uniform per case.
 */
typedef struct mplb_batch mplb_batch;
int mplb_batch_create(mplb_case* const* cases, int n, int precision, int device, mplb_batch** out);
int mplb_batch_close(mplb_batch* b);
const char* mplb_batch_last_error(const mplb_batch* b);
/* Initial density (node-major, the case's N values) and temperature of case i. */
int mplb_batch_set_case(mplb_batch* b, int i, const double* rho, size_t n, double T);
/* equilibrium + first collision of every case (multiPhasePP.c:122-147 and step 1's
 * collision); afterwards every case is at time level 1. */
int mplb_batch_init(mplb_batch* b);
/* n more steps of every case (plain launches). */
int mplb_batch_step(mplb_batch* b, int n);
/* Steps the cases still changing in graph-captured runs of check_every (even) steps;
 * after each run a case whose max_x |rho(x) - rho_prev(x)| / rho_l over the last
 * check_every steps is below threshold is steady (rho_l = largest initial density of
 * the case) and leaves the work table, a case whose density turned NaN is diverged.
 * Stops when no case is left or after max_steps.  Per case (arrays of n, may be NULL):
 * time level reached, last relative change, state (0 still running, 1 steady,
 * 2 diverged).  Returns the number of steady cases or a negative MPLB_E* code. */
int mplb_batch_run_to_steady(mplb_batch* b, double threshold, int check_every, int64_t max_steps,
                             int64_t* case_steps, double* case_delta, int* case_state);
/* "rho" or "p" (macroPressure.c:20-65, liMRT form) of case i at its current level. */
int mplb_batch_get_field(mplb_batch* b, int i, const char* name, double* out, size_t n);
int64_t mplb_batch_case_steps(const mplb_batch* b, int i);
/* Autotuner: times the single-pass kernel at 16/32/64 rows per work item (when every
 * case is at most 256 nodes wide) and the two-pass kernels at 64..1024 nodes per
 * workgroup, each with and without the XCD-ordered work table, over steps_per_trial
 * steps each on the current cases (which advance meanwhile), and keeps the fastest.
 * chosen_block is the two-pass block size (also used by the measurement passes).
 * mplb_batch_set_launch selects the two-pass kernels by hand, mplb_batch_set_march the
 * single-pass kernel (rows per item; 0 = two-pass); both kernels give identical results.
 * New batches use the single-pass kernel with 32 rows when every case fits. */
int mplb_batch_tune(mplb_batch* b, int steps_per_trial, int* chosen_block, int* chosen_xcd_order);
int mplb_batch_set_launch(mplb_batch* b, int block, int xcd_order);
int mplb_batch_set_march(mplb_batch* b, int rows);
int mplb_batch_march_rows(const mplb_batch* b);
/* on != 0: runs of steps (mplb_batch_step, each graph run of mplb_batch_run_to_steady)
 * go out as one persistent cooperative launch of the single-pass kernel, the grid
 * meeting at a barrier between steps, instead of one launch per step; results are
 * identical.  Needs the single-pass kernel (rows > 0; mplb_batch_tune may pick the
 * two-pass kernels, which then run per step).  No counterpart in the reference (its
 * cases run one multiPhasePP process each, multiPhasePP.c:157-241). */
int mplb_batch_set_persistent(mplb_batch* b, int on);
int mplb_batch_launch_info(const mplb_batch* b, int* block, int* xcd_order, int* workgroups, int64_t* nodes);

/* ---------------------------------------------------------------------------------
 * Runs to steady state and the learned warm start.  No counterpart in the reference
 * (its cases start from setInitialFields generators, setInitialFields.c:131-200, and
 * run a fixed number of steps); these calls let a driver stop at a steady criterion
 * and start a case from a proposed near-steady density.
 */
/* max over nodes of |rho - rho at the previous call| for the current time level (device
 * reduction; NaN if a density is NaN).  The first call after mplb_init only takes the
 * snapshot and returns NaN. */
int mplb_density_change(mplb_case* c, double* max_abs);
/* rho(x) = prof(|x - centre| - R) on the stored lattice (x = node coordinates), prof
 * sampled at d0 + k dd, k = 0..nprof-1, linear interpolation, clamped at both ends;
 * writes the case's density (call mplb_init afterwards) and returns in *mass the mass of
 * the distinct sites: on a periodic axis of a built box the last node duplicates the
 * first (period n - 1) and is left out of the sum. */
int mplb_set_radial_density(mplb_case* c, const double* prof, int nprof, double d0, double dd, const double* centre,
                            double R, double* mass);
/* mplb_set_radial_density without any device allocation (no hipMalloc/hipFree): the
 * profile table and the partial masses live in a fixed device scratch array of the
 * library (552 KiB per GPU, held until the process ends).  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP
 * graph capture, allocation-free serving or stepping loops).  Limits: nprof <= 4096 and
 * at most 2^24 stored nodes (MPLB_E_ARG beyond them).  Results are identical to
 * mplb_set_radial_density's.  All calls on one GPU share that GPU's scratch: calls that
 * overlap on one GPU (several host threads) race. */
int mplb_set_radial_density_noalloc(mplb_case* c, const double* prof, int nprof, double d0, double dd,
                                    const double* centre, double R, double* mass);

/* Multilayer perceptron (tanh hidden layers, linear output, mean squared error, Adam)
 * whose products run on the matrix cores with BF16 operands and FP32 accumulation;
 * FP32 master parameters.  widths[0..nlayers]: input width (<= 32), then layer widths
 * (multiples of 32).  Parameters are laid out per layer as W [in][out] then b [out].
 * Every product processes `batch` rows (a multiple of 32). */
typedef struct mplb_mlp mplb_mlp;
int mplb_mlp_create(int nlayers, const int* widths, int batch, int device, mplb_mlp** out);
int mplb_mlp_close(mplb_mlp* n);
const char* mplb_mlp_last_error(const mplb_mlp* n);
int64_t mplb_mlp_num_params(const mplb_mlp* n);
/* also resets the Adam moments */
int mplb_mlp_set_params(mplb_mlp* n, const float* params, size_t count);
int mplb_mlp_get_params(mplb_mlp* n, float* params, size_t count);
/* y [rows][out] = net(x [rows][in]) for any number of rows */
int mplb_mlp_forward(mplb_mlp* n, const float* x, int rows, float* y);
/* loss and gradient of one batch (x [batch][in], t [batch][out]); grad may be NULL */
int mplb_mlp_gradient(mplb_mlp* n, const float* x, const float* t, float* loss, float* grad, size_t ngrad);
/* iters Adam steps on minibatches idx[it * batch .. +batch) of the samples X/T, replayed
 * from one captured graph; loss_hist[it] (may be NULL) is each minibatch's loss.  Each
 * call starts a fresh Adam run: moments zeroed, bias-correction step count from 1. */
int mplb_mlp_train(mplb_mlp* n, const float* X, const float* T, int nsamples, const int32_t* idx, int iters,
                   float lr, float beta1, float beta2, float eps, float* loss_hist);
/* intermediate tensors of the last gradient call as FP32: "a" (level j BF16 input/
 * activation), "a32", "y", "dz", "dzb", "wb" (layer j BF16 weights [in_padded][out]) */
int mplb_mlp_tensor(mplb_mlp* n, const char* name, int j, float* out, size_t count);

#ifdef __cplusplus
}
#endif

#endif
