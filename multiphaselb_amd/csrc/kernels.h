// Launch interface of the per-routine HIP kernels (kernels_generic.hip).  Device arrays
// are structure-of-arrays: a population field is [Q][N] (population k of node i at
// k*N + i), a vector field is [3][N], the neighbour table is [Q][N] int32.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "physics.h"

namespace mplb {

enum LbeModel : int { M_LIMRT = 0, M_MYMRT = 1, M_XUMRT = 2, M_LITEMPMRT = 3 };
enum BdType : int { BD_BBACK = 0, BD_FT = 1, BD_FRDT = 2, BD_FGT = 3, BD_FU = 4, BD_UHF = 5,
                    BD_FEP = 6, BD_OF = 7, BD_PERIODIC = 8 };

// Model constants of one population field (lbParameters.h union, flattened).
struct ModelParams {
    int model;
    double Lambda[15];
    double sigma;
    int surf;          // liSurfTen / xuSurfTen
    double kappa;
    double alpha1, alpha2;
};

// Device view of the mesh and the macroscopic fields.
template <typename T>
struct DevState {
    int64_t N;
    const int* nb;               // [Q][N]
    const uint8_t* onbnd;        // isOnBnd
    T* rho;
    T* temp;
    T* p;
    T* U;                        // [3][N]
    T* Fi;                       // [3][N]
    T* Ft;                       // [3][N]
    T* ss;                       // scalar source of the g field
    T* psi;                      // potential(rho, T) scratch
};

// One boundary list prepared for launch.
struct BdList {
    const int* nodes;            // device list
    int n;
    int type;
    int sequential;              // list has read-after-write hazards: apply in order
    double value[3];             // fixedT T / fixedU U / fixedRandomT T
    double pert;                 // fixedRandomT percentage
    const double* rnd;           // fixedRandomT draws, link order
    const int* rnd_off;          // fixedRandomT first draw of each list node
    int draw;                    // fixedRandomT: 1 while localTStep < psteps
};

template <typename T>
struct GenericOps {
    static void psi(int model, hipStream_t s, const EosParams& e, DevState<T> st);
    static void inter_force(int model, hipStream_t s, const EosParams& e, DevState<T> st, double rho0);
    static void collide(int model, hipStream_t s, const EosParams& e, const ModelParams& mp, DevState<T> st,
                        T* f, double rho0);
    static void stream(int model, hipStream_t s, int64_t N, const int* nb, const T* in, T* out);
    static void boundary(int model, hipStream_t s, const EosParams& e, const ModelParams& mp, DevState<T> st,
                         T* f, const BdList& b, double rho0);
    static void macro_density(int model, hipStream_t s, DevState<T> st, const T* f);
    static void macro_velocity(int model, hipStream_t s, const EosParams& e, int fmodel, DevState<T> st,
                               const T* f, double rho0);
    static void macro_temperature(int model, hipStream_t s, int gmodel, DevState<T> st, const T* g);
    static void heat_source(int model, hipStream_t s, const EosParams& e, const ModelParams& mp, DevState<T> st,
                            int first_call);
    static void equilibrium(int model, hipStream_t s, const ModelParams& mp, DevState<T> st, T* f);
    static void pressure(int model, hipStream_t s, const EosParams& e, int fmodel, double sigma, DevState<T> st);
    // sum of x[0..n) in ascending order with one accumulator (bit-identical to a serial
    // loop), or a tree sum with FP64 partials when ordered == 0.
    static void sum(hipStream_t s, const T* x, int64_t n, double* out_dev, int ordered, double* scratch);
    static void nan_count(hipStream_t s, const T* x, int64_t n, unsigned int* out_dev);
};

}  // namespace mplb
