// Fused structured-box path: FP32 arithmetic, FP32 or FP16-deviation population storage.
//
// The state between steps is the post-collision population set of the current step.
// Isothermal D3Q15 boxes (FP32 or FP16 storage) whose links are all the periodic wrap take
// one single-pass kernel per step (k_fused_march: pull-stream, psi in an LDS ring of
// planes, force, velocity and collision).  Other boxes take two node-parallel passes (three with heat
// transfer):
//   K1  pull-stream f (and g) from the post-collision buffers, apply the boundary
//       conditions of the node in registers, and write what the neighbours' stencils
//       need: psi = potential(rho, T) (isothermal) or rho and sum g (thermal);
//   KH  (thermal) heat source from the rho / T / U stencils, new T, psi;
//   K2  pull again (the same addresses K1 touched), psi stencil -> interaction force ->
//       velocity, then the MRT collisions of the next step, written to the other buffer.
// Interior nodes compute the pull source as a fixed offset; nodes on the box surface
// read the neighbour table, so the reference's duplicate-node periodic links and wall
// links are followed exactly.  Block b of K1/K2 covers the 256 consecutive node ids
// [256 b, 256 b + 256).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "case.h"
#include "fused.h"
#include "physics.h"

namespace mplb {

namespace {

enum : uint8_t { FB_NONE = 0, FB_BBACK = 1, FB_FT = 2, FB_FRDT = 3 };

struct FArgs {
    int nx, ny, nz;
    int64_t N, nxy;
    int64_t P;                                  // pitch between population planes (>= N)
    int nblocks;                                // 256-node blocks
    int face_table;                             // faces (x0 x1 y0 y1 z0 z1 bits) read through the table
    const int* nb;                              // [Q][N]
    const uint8_t* onbnd;
    const uint8_t* bcf;                         // per-node boundary code, f (surface nodes)
    const uint8_t* bcg;                         // low 4 bits: code; fixedRandomT: list slot in the high 4
    const float* bcTw;                          // per-node wall temperature (g)
    const float* bcPert;                        // per-node fixedRandomT percentage
    const int* rnd_off;                         // per-node first draw
    const double* rnd;
    int draw;                                   // fixedRandomT lists drawing this step (bit per list slot)
    EosParams eos;
    float rho0;
    int fmodel, gmodel, surf;
    float fl[15], gl[15];
    float sigma, kappa, a1, a2;
    const void* fin;
    void* fout;
    const void* gin;
    void* gout;
    float* rho;
    float* sumg;
    float* psi;
    float* S;
    const float* Told;
    float* Tnew;
    const float* Uold;
    float* Unew;
    float Tconst;
    int Tuniform;
    // loop-invariant factors of the FP32 arithmetic, set on the host: the surface-tension
    // source coefficients 12 sigma / (1/L1 - 1/2), 12 sigma / (1/L2 - 1/2) (D2Q9) or
    // 6 sigma / (1/L1 - 1/2) (D3Q15), and for the van der Waals potential R, b, a / M^2,
    // 1 / G and M (vdw = 1)
    float sf1, sf2;
    int vdw;
    float vR, vb, va_m2, vinv_g, vM;
    // optional outputs (materialisation)
    float* out_rho;
    float* out_U;
    float* out_Fi;
    float* out_Ft;
    int first_call;                             // liTempMRT heat source S_old = 0
    float fshift[15], gshift[15];               // FP16 storage shifts (f: FP16S only)
    // single-pass kernel: column tiles in x, tiles per plane, work items, items per XCD,
    // planes per chunk
    int mtx, mtiles, mitems, mper, mzc;
    // FP16N storage of f: per-node reference densities of the input and output levels
    const _Float16* rin;
    _Float16* rout;
};

template <typename S>
__device__ inline float ld(const void* p, int64_t i, float shift) {
    if constexpr (sizeof(S) == 2) {
        return (float)((const _Float16*)p)[i] + shift;
    } else {
        return ((const float*)p)[i];
    }
}

#ifndef MPLB_NT_STORE
#define MPLB_NT_STORE 1
#endif

template <typename S>
__device__ inline void st(void* p, int64_t i, float v, float shift) {
    if constexpr (sizeof(S) == 2) {
        if (MPLB_NT_STORE) __builtin_nontemporal_store((_Float16)(v - shift), (_Float16*)p + i);
        else ((_Float16*)p)[i] = (_Float16)(v - shift);
    } else {
        if (MPLB_NT_STORE) __builtin_nontemporal_store(v, (float*)p + i);
        else ((float*)p)[i] = v;
    }
}

// Two FP16 storage modes for f (FP32 arithmetic in both):
// * MPLB_FP16S (storage type _Float16): f_q - w_q rho_ref, rho_ref the lattice mean of
//   the initial density (one shift per population, fshift);
// * MPLB_FP16N (storage type h16n): f_q at node x as its relative deviation
//   (f_q - w_q R(x)) / (w_q R(x)), R(x) being x's density when the population was written
//   (the collision's density, rounded to FP16 and kept per node beside the populations,
//   one array per level like the populations).  The reference follows each node's own
//   density and the deviation is scaled by it, so both phases of a two-phase field keep
//   the 11-bit mantissa (with one shift the light phase's deviations are of the order of
//   the density jump; unscaled, they would fall among the FP16 subnormals).
// g keeps one global shift per population in both: its rest equilibrium depends on T only.
struct h16n {
    _Float16 v;
};
template <typename S>
constexpr bool kNodeRef = std::is_same<S, h16n>::value;
static_assert(sizeof(h16n) == 2, "FP16 storage element");

constexpr float kRefMin = 6.103515625e-05f;     // smallest normal FP16 (R never rounds to 0)

// FP16 storage: w_q R of the node a population was stored at (FP16N), or the global
// shift of population q (FP16S)
template <class L, typename S>
__device__ inline float f_wref(const FArgs& a, int64_t node, int q) {
    if constexpr (kNodeRef<S>) {
        return (float)L::t.omega[q] * (float)a.rin[node];
    } else if constexpr (sizeof(S) == 2) {
        return a.fshift[q];
    } else {
        return 0.f;
    }
}

// a stored f population; wref from f_wref
template <typename S>
__device__ inline float ldf(const void* p, int64_t i, float wref) {
    if constexpr (kNodeRef<S>) {
        return fmaf((float)((const _Float16*)p)[i], wref, wref);
    } else {
        return ld<S>(p, i, wref);
    }
}

// FP16N: writes node id's reference for the output level (the node's density, at least
// kRefMin; NaN stays NaN) and returns it with its reciprocal; 0 in the other modes
template <typename S>
__device__ inline float put_ref(const FArgs& a, int64_t id, float rho, float& rinv) {
    if constexpr (kNodeRef<S>) {
        const float r = (rho >= kRefMin || rho != rho) ? rho : kRefMin;
        const _Float16 h = (_Float16)r;
        a.rout[id] = h;
        const float R = (float)h;
        rinv = 1.f / R;
        return R;
    } else {
        rinv = 0.f;
        return 0.f;
    }
}

// FP16N: stores f population q of a node (reference R and 1 / R = rinv from put_ref)
template <class L>
__device__ inline void stf_ref(void* p, int64_t i, int q, float v, float R, float rinv) {
    const float wref = (float)L::t.omega[q] * R;
    const float dev = (v - wref) * ((float)(1.0 / L::t.omega[q]) * rinv);
    if (MPLB_NT_STORE) __builtin_nontemporal_store((_Float16)dev, (_Float16*)p + i);
    else ((_Float16*)p)[i] = (_Float16)dev;
}

struct NodePos {
    int i, j, k;
    int64_t id;
    bool surface;       // on the box surface
    bool table;         // on a face whose links must be read from the neighbour table
    bool valid;
    unsigned X[3], Y[3], Z[3];   // wrapped coordinates u-1, u, u+1 premultiplied by the strides
};

// psi = potential(rho, T) (potential.c:7-13) in FP32 with the loop-invariant factors taken
// from the host for the van der Waals EOS (one division instead of three); other EOS
// types use the shared form.
// The fused FP32 kernels take quotients and square roots from the hardware reciprocal
// and square root (about 1 ulp) instead of the correctly rounded instruction sequences
// (about ten instructions each): the single passes spend half their time issuing VALU
// instructions, and their results carry FP32 rounding anyway.
__device__ inline float rcp_fast(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline float sqrt_fast(float x) { return __builtin_amdgcn_sqrtf(x); }

__device__ inline float fpot(const FArgs& a, float rho, float T) {
    if (a.vdw) {
        const float p = rho * a.vR * T * rcp_fast(a.vM - rho * a.vb) - a.va_m2 * rho * rho;
        return sqrt_fast(2.f * (p - rho * float(kCs2)) * a.vinv_g);
    }
    return potential<float>(a.eos, rho, T);
}

// Block b covers the 256 consecutive node ids [256 b, 256 b + 256).
template <class L>
__device__ inline NodePos node_pos(const FArgs& a) {
    NodePos p;
    const unsigned id = blockIdx.x * 256u + threadIdx.x;     // N < 2^31
    const unsigned nxy = (unsigned)a.nxy, nx = (unsigned)a.nx;
    p.id = id;
    p.valid = p.id < a.N;
    const unsigned k = id / nxy, r = id - k * nxy, j = r / nx;
    p.k = (int)k;
    p.j = (int)j;
    p.i = (int)(r - j * nx);
    const bool x0 = p.i == 0, x1 = p.i == a.nx - 1, y0 = p.j == 0, y1 = p.j == a.ny - 1;
    const bool z0 = L::D == 3 && p.k == 0, z1 = L::D == 3 && p.k == a.nz - 1;
    p.surface = x0 || x1 || y0 || y1 || z0 || z1;
    const int m = a.face_table;
    p.table = (x0 && (m & 1)) || (x1 && (m & 2)) || (y0 && (m & 4)) || (y1 && (m & 8)) || (z0 && (m & 16)) ||
              (z1 && (m & 32));
    // periodic wrap with duplicate end nodes: u - 1 at 0 is n - 2, u + 1 at n - 1 is 1
    p.X[0] = x0 ? nx - 2 : p.i - 1;
    p.X[1] = p.i;
    p.X[2] = x1 ? 1 : p.i + 1;
    p.Y[0] = (y0 ? a.ny - 2 : p.j - 1) * nx;
    p.Y[1] = j * nx;
    p.Y[2] = (y1 ? 1 : p.j + 1) * nx;
    if (L::D == 3) {
        p.Z[0] = (z0 ? a.nz - 2 : p.k - 1) * nxy;
        p.Z[1] = k * nxy;
        p.Z[2] = (z1 ? 1 : p.k + 1) * nxy;
    } else {
        p.Z[0] = p.Z[1] = p.Z[2] = 0;
    }
    return p;
}

// upstream node x - c_q (lbstream.c:16-62: a missing link keeps the node's own value)
template <class L>
__device__ inline unsigned upstream(const FArgs& a, const NodePos& p, int q) {
    if (p.table) {
        const int s = a.nb[q * a.N + p.id];
        return s == -1 ? (unsigned)p.id : (unsigned)s;
    }
    return p.X[1 - L::t.c[q][0]] + p.Y[1 - L::t.c[q][1]] + p.Z[1 - L::t.c[q][2]];
}

// downstream node x + c_q (the interaction-force and stencil neighbour), -1 if missing
template <class L>
__device__ inline int64_t downstream(const FArgs& a, const NodePos& p, int q) {
    if (p.table) return a.nb[L::t.rev[q] * a.N + p.id];
    return (int64_t)(p.X[1 + L::t.c[q][0]] + p.Y[1 + L::t.c[q][1]] + p.Z[1 + L::t.c[q][2]]);
}

// REF: f stored against per-node references (FP16N); otherwise shift holds the storage
// shifts
template <class L, typename S, bool REF>
__device__ inline void pull(const FArgs& a, const NodePos& p, const void* src, const float* shift, float* v) {
#pragma unroll
    for (int q = 0; q < L::Q; q++) {
        const void* plane = (const char*)src + (size_t)q * a.P * sizeof(S);
        if constexpr (REF && kNodeRef<S>) {
            const unsigned up = upstream<L>(a, p, q);
            v[q] = ldf<S>(plane, up, f_wref<L, S>(a, up, q));
        } else {
            v[q] = ld<S>(plane, upstream<L>(a, p, q), shift[q]);
        }
    }
}

// bounceBack.c:14-75 in registers
template <class L>
__device__ inline void bounce_back(const FArgs& a, const NodePos& p, float* v) {
    bool miss[L::Q];
#pragma unroll
    for (int q = 0; q < L::Q; q++) miss[q] = a.nb[q * a.N + p.id] == -1;
#pragma unroll
    for (int q = 0; q < L::Q; q++) {
        const int r = L::t.rev[q];
        if (miss[q]) {
            v[q] = v[r];
        } else if (!miss[r] && q >= 1 && q <= L::naxis) {
            v[q] = 0.5f * v[q] + 0.5f * v[r];
            v[r] = v[q];
        }
    }
}

// fixedT.c:45-122 / fixedRandomT.c:48-112 in registers; the neighbour's post-stream
// population is pulled directly and its T and U are the previous step's.
template <class L, typename S>
__device__ inline void fixed_temperature(const FArgs& a, const NodePos& p, float* v, uint8_t code, bool draws) {
    int draw = code == FB_FRDT ? a.rnd_off[p.id] : 0;
    const float Tw0 = a.bcTw[p.id];
    float minT = 0, maxT = 0;
    if (code == FB_FRDT) {
        const float pert = a.bcPert[p.id];
        minT = Tw0 * (100.0f - pert) / 100.0f;
        maxT = Tw0 * (100.0f + pert) / 100.0f;
    }
    for (int q = 0; q < L::Q; q++) {
        if (a.nb[q * a.N + p.id] != -1) continue;
        const int nbid = a.nb[L::t.rev[q] * a.N + p.id];
        if (nbid == -1) continue;
        const int s2 = a.nb[q * a.N + nbid];
        const float gnb = ld<S>(a.gin, q * a.P + (s2 == -1 ? (int64_t)nbid : (int64_t)s2), a.gshift[q]);
        const float U[3] = {a.Uold[nbid], a.Uold[a.N + nbid], a.Uold[2 * a.N + nbid]};
        const float Tnb = a.Tuniform ? a.Tconst : a.Told[nbid];
        float geq[L::Q];
        if (a.gmodel == 1) my_geq<L, float>(Tnb, U, a.a1, a.a2, geq);
        else litemp_geq<L, float>(Tnb, U, geq);
        const float eq_nb = geq[q];
        float Tw = Tw0;
        if (code == FB_FRDT && draws) Tw = minT + (maxT - minT) * (float)a.rnd[draw++];
        if (a.gmodel == 1) my_geq<L, float>(Tw, U, a.a1, a.a2, geq);
        else litemp_geq<L, float>(Tw, U, geq);
        v[q] = geq[q] + (gnb - eq_nb);
    }
}

template <class L, typename S>
__device__ inline void pull_f(const FArgs& a, const NodePos& p, float* v) {
    pull<L, S, kNodeRef<S>>(a, p, a.fin, a.fshift, v);
    if (p.table && a.bcf[p.id] == FB_BBACK) bounce_back<L>(a, p, v);
}

template <class L, typename S>
__device__ inline void pull_g(const FArgs& a, const NodePos& p, float* v) {
    pull<L, S, false>(a, p, a.gin, a.gshift, v);
    if (!p.table) return;
    const uint8_t bc = a.bcg[p.id];
    const uint8_t code = bc & 0xF;
    if (code == FB_BBACK) bounce_back<L>(a, p, v);
    else if (code == FB_FT || code == FB_FRDT) fixed_temperature<L, S>(a, p, v, code, ((a.draw >> (bc >> 4)) & 1) != 0);
}

// K1
template <class L, typename S, bool THERMAL>
__global__ void __launch_bounds__(256) k_fused_density(FArgs a) {
    const NodePos p = node_pos<L>(a);
    if (!p.valid) return;
    float v[L::Q];
    pull_f<L, S>(a, p, v);
    const float rho = rho_sym<L, float>(v);
    if (THERMAL) {
        pull_g<L, S>(a, p, v);
        const float sg = rho_sym<L, float>(v);
        a.rho[p.id] = rho;
        a.sumg[p.id] = sg;
    } else {
        const float T = a.Tuniform ? a.Tconst : a.Told[p.id];
        a.psi[p.id] = fpot(a, rho, T);
        if (a.out_rho) a.out_rho[p.id] = rho;
    }
}

// heatSource.c:26-170 after the stencils: conductivity, dp/dT by central difference and
// the source from dot = grad T . grad(1/rho) and divU (shared by k_fused_heat and the
// thermal single pass, so both round alike)
template <class L>
__device__ inline float heat_tail(const FArgs& a, float rho, float Told, float dot, float divU) {
    // multiply-adds fused within an expression only, so every kernel that inlines this
    // rounds alike
#pragma clang fp contract(on)
    const float Cv = (float)a.eos.Cv;
    float lambda;
    if (a.gmodel == 1) {
        lambda = (L::D == 2) ? rho * Cv * (rcp_fast(a.gl[3]) - 0.5f) * (4.0f + 3.0f * a.a1 + 2.0f * a.a2) / 6.0f
                             : rho * Cv * (rcp_fast(a.gl[3]) - 0.5f) * (6.0f + 11.0f * a.a1 + a.a2) / 9.0f;
    } else {
        lambda = rho * Cv * (rcp_fast(a.gl[3]) - 0.5f);
    }
    // dp/dT by the central difference of heatSource.c; the van der Waals pressure is linear
    // in T, so there the difference quotient is rho R / (M - rho b) exactly
    float dpdT;
    if (a.vdw) {
        dpdT = rho * a.vR * rcp_fast(a.vM - rho * a.vb);
    } else {
        dpdT = p_eos<float>(a.eos, rho, 1.1f * Told) - p_eos<float>(a.eos, rho, 0.9f * Told);
        dpdT = dpdT * rcp_fast(0.2f * Told);
    }
    return -lambda * dot * rcp_fast(Cv) + divU * Told * (1.0f - dpdT * rcp_fast(rho * Cv));
}

// central-difference terms of one axis (heatSource.c:60-80 with both neighbours present)
__device__ inline void heat_axis(float Tp, float Tm, float rp, float rm, float up, float um, float& dot, float& divU) {
#pragma clang fp contract(on)
    const float gT = 0.5f * (Tp - Tm);
    const float gR = 0.5f * (rcp_fast(rp) - rcp_fast(rm));
    const float dU = 0.5f * (up - um);
    dot += gT * gR;
    divU += dU;
}

// KH: heatSource.c:26-170 with macroTemperature (myMRTTemperature.c:16-30)
template <class L>
__global__ void __launch_bounds__(256) k_fused_heat(FArgs a) {
    const NodePos p = node_pos<L>(a);
    if (!p.valid) return;
    const int64_t id = p.id;
    const float rho = a.rho[id], Told = a.Told[id];
    float dot = 0.f, divU = 0.f;
#pragma unroll
    for (int ax = 0; ax < L::D; ax++) {
        // a = nb[id][fd_plus] = x + e_ax, b = nb[id][fd_minus] = x - e_ax
        const int64_t na = downstream<L>(a, p, L::fd_minus[ax]);
        const int64_t nbb = downstream<L>(a, p, L::fd_plus[ax]);
        const float* Uax = a.Uold + ax * a.N;
        if (na != -1 && nbb != -1) {
            heat_axis(a.Told[na], a.Told[nbb], a.rho[na], a.rho[nbb], Uax[na], Uax[nbb], dot, divU);
        } else {
            float gT, gR, dU;
            if (na == -1 && nbb != -1) {
                gT = Told - a.Told[nbb];
                gR = 1.0f / rho - 1.0f / a.rho[nbb];
                dU = Uax[id] - Uax[nbb];
            } else if (na != -1 && nbb == -1) {
                gT = a.Told[na] - Told;
                gR = 1.0f / a.rho[na] - 1.0f / rho;
                dU = Uax[na] - Uax[id];
            } else {
                gT = 0.f;
                gR = 0.f;
                dU = 0.f;
            }
            dot += gT * gR;
            divU += dU;
        }
    }
    const float s = heat_tail<L>(a, rho, Told, dot, divU);
    float T;
    float Sv;
    if (a.gmodel == 1) {
        Sv = s;
        T = a.sumg[id] + 0.5f * Sv;
    } else {
        const float old = a.first_call ? 0.f : a.S[id];
        Sv = s + 0.5f * (s - old);
        T = a.sumg[id];
    }
    a.S[id] = Sv;
    a.Tnew[id] = T;
    a.psi[id] = fpot(a, rho, T);
}

// Interaction force (interactionForce.c:47-79), velocity (liMRTVelocity.c /
// xuMRTVelocity.c) and the MRT collision of f (liMRTCollision.c:33-106 /
// xuMRTCollision.c:42-92, surface tension terms of liMRTSurfaceTension.c /
// xuMRTSurfaceTension.c) of one node: v holds its pulled populations on entry and the
// post-collision populations on exit; psin[q] is psi at the downstream neighbour
// x + c_q (psin[0] the node's own psi).  interior: the node has no missing link
// (isOnBnd == 0); full: surface tension applies (no missing link).
template <class L>
__device__ inline void collide_node(const FArgs& a, float* v, float rho, const float* psin, bool interior, bool full,
                                    float* U, float* Fi, float* Ft) {
    const float psic = psin[0];
    Fi[0] = Fi[1] = Fi[2] = 0.f;
    constexpr Pairs<L> P = make_pairs<L>();
    if (interior) {
        // sum over q of w_q psi(x + c_q) c_q, as pair differences (odd under inversion)
        float F[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < Pairs<L>::n; k++) {
            const int q = P.q[k];
            const float dpsi = (float)L::t.w[q] * (psin[q] - psin[P.r[k]]);
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const int c = L::t.c[q][d];
                if (c == 1) F[d] += dpsi;
                else if (c == -1) F[d] -= dpsi;
            }
        }
        const float beta = -(float)a.eos.G * psic;
#pragma unroll
        for (int d = 0; d < 3; d++) Fi[d] = F[d] * beta;
    }
    float Fu[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const float grav = (rho - a.rho0) * (float)a.eos.g[d];
        Ft[d] = Fi[d] + grav + (interior ? (float)a.eos.fext[d] : 0.f);
        Fu[d] = a.fmodel == 0 ? Ft[d] : Fi[d] + grav;
    }
    const float irho = rcp_fast(rho);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        float j = 0.f;
#pragma unroll
        for (int k = 0; k < Pairs<L>::n; k++) {
            const int c = L::t.c[P.q[k]][d];
            if (c == 1) j += v[P.q[k]] - v[P.r[k]];
            else if (c == -1) j -= v[P.q[k]] - v[P.r[k]];
        }
        U[d] = (j + 0.5f * Fu[d]) * irho;
    }
    // MRT collision of f (liMRTCollision.c:33-106 / xuMRTCollision.c:42-92)
    float m[L::Q], meq[L::Q], Sf[L::Q];
    moments_sym<L, float>(v, m);
    const float umag = U[0] * U[0] + U[1] * U[1] + U[2] * U[2];
    const float psi2 = psic * psic;
    if (L::D == 2) {
        meq[0] = rho;
        meq[1] = rho * (-2.f + 3.f * umag);
        meq[2] = rho * (1.f - 3.f * umag);
        meq[3] = rho * U[0];
        meq[4] = -rho * U[0];
        meq[5] = rho * U[1];
        meq[6] = -rho * U[1];
        meq[7] = rho * (U[0] * U[0] - U[1] * U[1]);
        meq[8] = rho * U[0] * U[1];
        const float uF = U[0] * Fu[0] + U[1] * Fu[1];
        const float fi2 = Fi[0] * Fi[0] + Fi[1] * Fi[1];
        Sf[0] = 0.f;
        const float fp = fi2 * rcp_fast(psi2);
        Sf[1] = 6.f * uF + a.sf1 * fp;
        Sf[2] = -6.f * uF - a.sf2 * fp;
        Sf[3] = Fu[0];
        Sf[4] = -Fu[0];
        Sf[5] = Fu[1];
        Sf[6] = -Fu[1];
        Sf[7] = 2.f * (U[0] * Fu[0] - U[1] * Fu[1]);
        Sf[8] = U[0] * Fu[1] + U[1] * Fu[0];
    } else {
        meq[0] = rho;
        meq[1] = rho * (-1.f + umag);
        meq[2] = rho * (1.f - 5.f * umag);
        meq[3] = rho * U[0];
        meq[4] = rho * (-7.f / 3.f) * U[0];
        meq[5] = rho * U[1];
        meq[6] = rho * (-7.f / 3.f) * U[1];
        meq[7] = rho * U[2];
        meq[8] = rho * (-7.f / 3.f) * U[2];
        meq[9] = rho * (2.f * U[0] * U[0] - U[1] * U[1] - U[2] * U[2]);
        meq[10] = rho * (U[1] * U[1] - U[2] * U[2]);
        meq[11] = rho * U[0] * U[1];
        meq[12] = rho * U[1] * U[2];
        meq[13] = rho * U[0] * U[2];
        meq[14] = 0.f;
        const float uF = U[0] * Fu[0] + U[1] * Fu[1] + U[2] * Fu[2];
        const float F2 = Fu[0] * Fu[0] + Fu[1] * Fu[1] + Fu[2] * Fu[2];
        Sf[0] = 0.f;
        Sf[1] = 2.f * uF + a.sf1 * F2 * rcp_fast(psi2);
        Sf[2] = -10.f * uF;
        Sf[3] = Fu[0];
        Sf[4] = (-7.f / 3.f) * Fu[0];
        Sf[5] = Fu[1];
        Sf[6] = (-7.f / 3.f) * Fu[1];
        Sf[7] = Fu[2];
        Sf[8] = (-7.f / 3.f) * Fu[2];
        Sf[9] = 4.f * U[0] * Fu[0] - 2.f * U[1] * Fu[1] - 2.f * U[2] * Fu[2];
        Sf[10] = 2.f * U[1] * Fu[1] - 2.f * U[2] * Fu[2];
        Sf[11] = U[0] * Fu[1] + U[1] * Fu[0];
        Sf[12] = U[1] * Fu[2] + U[2] * Fu[1];
        Sf[13] = U[0] * Fu[2] + U[2] * Fu[0];
        Sf[14] = 0.f;
    }
    // m becomes the moment change of the collision, added to v at the end
#pragma unroll
    for (int q = 0; q < L::Q; q++) m[q] = (1.f - 0.5f * a.fl[q]) * Sf[q] - a.fl[q] * (m[q] - meq[q]);
    if (a.surf) {
        // liMRTSurfaceTension.c / xuMRTSurfaceTension.c (zero next to missing links)
        if (full) {
            float Qt[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
            for (int k = 0; k < Pairs<L>::n; k++) {
                const int q = P.q[k];
                const float al = (float)L::t.w[q] * ((psin[q] - psic) + (psin[P.r[k]] - psic));
#pragma unroll
                for (int r = 0; r < L::D; r++)
#pragma unroll
                    for (int s = r; s < L::D; s++) Qt[r][s] += al * (float)(L::t.c[q][r] * L::t.c[q][s]);
            }
            const float sc = 0.5f * psic * (float)a.eos.G * a.kappa;
            if (L::D == 2) {
                const float tr = (Qt[0][0] + Qt[1][1]) * sc;
                m[1] += 1.5f * a.fl[1] * tr;
                m[2] += -1.5f * a.fl[2] * tr;
                m[7] += -a.fl[8] * (Qt[0][0] - Qt[1][1]) * sc;
                m[8] += -a.fl[8] * Qt[0][1] * sc;
            } else {
                m[1] += 4.f * a.fl[1] * (Qt[0][0] + Qt[1][1] + Qt[2][2]) * sc / 5.f;
                const float l10 = -a.fl[10] * sc;
                m[9] += l10 * (2.f * Qt[0][0] - Qt[1][1] - Qt[2][2]);
                m[10] += l10 * (Qt[1][1] - Qt[2][2]);
                m[11] += l10 * Qt[0][1];
                m[12] += l10 * Qt[1][2];
                m[13] += l10 * Qt[0][2];
            }
        }
    }
    populations_add_sym<L, float>(m, v);
}

// MRT collision of g (myMRTCollision.c:35-212 / liTempMRTCollision.c:31-84) of one node:
// v holds its pulled populations on entry and the post-collision ones on exit; T and Sv
// are the node's new temperature and heat source, U its new velocity.
template <class L>
__device__ inline void collide_g(const FArgs& a, float* v, float T, float Sv, const float* U) {
    float n[L::Q], neq[L::Q];
    moments_sym<L, float>(v, n);
#pragma unroll
    for (int q = 0; q < L::Q; q++) neq[q] = 0.f;
    neq[0] = T;
    if (a.gmodel == 1) {
        neq[1] = a.a1 * T;
        neq[2] = a.a2 * T;
    } else {
        neq[1] = -2.f * T;
        neq[2] = 2.f * T;
    }
    neq[3] = T * U[0];
    neq[4] = -T * U[0];
    neq[5] = T * U[1];
    neq[6] = -T * U[1];
    if (L::D == 3) {
        neq[7] = T * U[2];
        neq[8] = -T * U[2];
    }
    if (a.gmodel == 1) {
        // myMRTCollision.c: n <- n - Q (n - neq) + (I - Q/2) Gamma
        float d[L::Q];
#pragma unroll
        for (int q = 0; q < L::Q; q++) d[q] = n[q] - neq[q];
        float r[L::Q];
#pragma unroll
        for (int q = 0; q < L::Q; q++) r[q] = a.gl[q] * d[q];
        const int np = L::D == 2 ? 2 : 3;
#pragma unroll
        for (int pr = 0; pr < np; pr++) {
            const int i0 = 3 + 2 * pr, i1 = 4 + 2 * pr;
            r[i0] += a.gl[i1] * (a.gl[i0] / 2.f - 1.f) * d[i1];
        }
        // n becomes the moment change, added to v below
#pragma unroll
        for (int q = 0; q < L::Q; q++) n[q] = -r[q];
        n[0] += (1.f - 0.5f * a.gl[0]) * Sv;
    } else {
        const float n4 = n[4], n6 = n[6];
#pragma unroll
        for (int q = 0; q < L::Q; q++) n[q] = -a.gl[q] * (n[q] - neq[q]);
        n[0] += Sv;
        n[3] += (1.f - a.gl[3] / 2.f) * a.gl[4] * n4;
        n[5] += (1.f - a.gl[5] / 2.f) * a.gl[6] * n6;
    }
    populations_add_sym<L, float>(n, v);
}

// K2
template <class L, typename S, bool THERMAL>
__global__ void __launch_bounds__(256) k_fused_collide(FArgs a) {
    const NodePos p = node_pos<L>(a);
    if (!p.valid) return;
    const int64_t id = p.id, N = a.N;
    float v[L::Q];
    pull_f<L, S>(a, p, v);
    const float rho = rho_sym<L, float>(v);
    const float psic = a.psi[id];
    const bool interior = !p.table || a.onbnd[id] == 0;
    float psin[L::Q];
    psin[0] = psic;
    if (interior || a.surf) {
#pragma unroll
        for (int q = 1; q < L::Q; q++) {
            const int64_t n = downstream<L>(a, p, q);
            psin[q] = n == -1 ? psic : a.psi[n];
        }
    }
    bool full = true;
    if (a.surf && p.table) {
#pragma unroll
        for (int q = 1; q < L::Q; q++)
            if (a.nb[q * N + id] == -1) full = false;
    }
    float U[3], Fi[3], Ft[3];
    collide_node<L>(a, v, rho, psin, interior, full, U, Fi, Ft);
    if constexpr (kNodeRef<S>) {
        float rinv;
        const float R = put_ref<S>(a, id, rho, rinv);
#pragma unroll
        for (int q = 0; q < L::Q; q++) stf_ref<L>(a.fout, q * a.P + id, q, v[q], R, rinv);
    } else {
#pragma unroll
        for (int q = 0; q < L::Q; q++) st<S>(a.fout, q * a.P + id, v[q], a.fshift[q]);
    }

    if (THERMAL) {
        pull_g<L, S>(a, p, v);
        collide_g<L>(a, v, a.Tnew[id], a.S[id], U);
#pragma unroll
        for (int q = 0; q < L::Q; q++) st<S>(a.gout, q * a.P + id, v[q], a.gshift[q]);
        a.Unew[id] = U[0];
        a.Unew[N + id] = U[1];
        a.Unew[2 * N + id] = U[2];
    }
    if (a.out_U) {
        a.out_U[id] = U[0];
        a.out_U[N + id] = U[1];
        a.out_U[2 * N + id] = U[2];
    }
    if (a.out_rho) a.out_rho[id] = rho;
    if (a.out_Fi) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            a.out_Fi[d * N + id] = Fi[d];
            a.out_Ft[d * N + id] = Ft[d];
        }
    }
}


// ---------------------------------------------------------------------------
// Single-pass isothermal step on fully periodic D3Q15 boxes.  A workgroup owns a
// MX x MY column of (x, y) sites over a chunk of z planes and walks up the chunk: at
// plane z it pulls the populations of plane z + 1 for its sites and for the one-site
// ring around them, turns them into psi in an LDS ring of four planes, then collides
// plane z from the populations it pulled one plane earlier (kept in registers) and the
// psi stencil read from LDS.  psi never goes to memory, and per node the step reads
// the populations once and writes them once; the ring re-pulls hit the L2, because the
// workgroups that own those sites run on the same XCD at the same time (work items are
// handed out so that each XCD takes a contiguous range of chunk-major items).  The
// column is 64 x 4 (256-byte rows per population plane): on the 256^3 bubble it runs 6%
// faster than 32 x 8, 48% faster than 16 x 16 and 6% faster than 128 x 2.  One thread
// per ring site: MRING <= 256.
constexpr int MX = 64, MY = 4, MHX = MX + 2, MHY = MY + 2, MSLOT = MHX * MHY;
constexpr int MRING = 2 * MHX + 2 * MY;
static_assert(MX * MY == 256 && MRING <= 256, "one thread per column site and per ring site");

__device__ inline int wrapd(int u, int n) { return u < 0 ? u + (n - 1) : (u > n - 1 ? u - (n - 1) : u); }

// s_waitcnt immediates (vmcnt in bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8): every
// counter at zero; only the LDS counter at zero; only the vector-memory counter at zero
constexpr int kWaitAll = 0;
constexpr int kWaitLds = 0xF | 0x7 << 4 | 3 << 14;
constexpr int kWaitVm = 0x7 << 4 | 0xF << 8;

// Element index of population q of node i in 32 bits, for the thermal single pass: with
// 64-bit plane offsets the compiler keeps thirty 64-bit plane addresses (f and g) in
// scalar registers, which spill and come back through v_readlane before every pull.  The
// thermal single pass runs only on boxes whose Q x pitch is below 2^32.
__device__ inline unsigned pidx(const FArgs& a, int q, unsigned i) { return (unsigned)q * (unsigned)a.P + i; }

// populations of node (x, y, z) pulled from x - c_q (duplicate-node periodic wrap) and
// their sum (REF: f stored against per-node references, FP16N)
template <class L, typename S, bool REF>
__device__ inline float march_pull_from(const FArgs& a, const void* src, const float* shift, int x, int y, int z,
                                       float* v) {
    const unsigned nx = (unsigned)a.nx, nxy = (unsigned)a.nxy;
    const unsigned X[3] = {(unsigned)wrapd(x - 1, a.nx), (unsigned)x, (unsigned)wrapd(x + 1, a.nx)};
    const unsigned Y[3] = {(unsigned)wrapd(y - 1, a.ny) * nx, (unsigned)y * nx, (unsigned)wrapd(y + 1, a.ny) * nx};
    const unsigned Z[3] = {(unsigned)wrapd(z - 1, a.nz) * nxy, (unsigned)z * nxy, (unsigned)wrapd(z + 1, a.nz) * nxy};
#pragma unroll
    for (int q = 0; q < L::Q; q++) {
        const void* plane = (const char*)src + (size_t)q * a.P * sizeof(S);
        if constexpr (REF && kNodeRef<S>) {
            const unsigned up = X[1 - L::t.c[q][0]] + Y[1 - L::t.c[q][1]] + Z[1 - L::t.c[q][2]];
            v[q] = ldf<S>(plane, up, f_wref<L, S>(a, up, q));
        } else {
            v[q] = ld<S>(plane, X[1 - L::t.c[q][0]] + Y[1 - L::t.c[q][1]] + Z[1 - L::t.c[q][2]], shift[q]);
        }
    }
    return rho_sym<L, float>(v);
}

template <class L, typename S>
__device__ inline float march_pull(const FArgs& a, int x, int y, int z, float* v) {
    return march_pull_from<L, S, kNodeRef<S>>(a, a.fin, a.fshift, x, y, z, v);
}

template <class L, typename S>
__device__ inline float march_psi(const FArgs& a, int x, int y, int z, float* v) {
    const float rho = march_pull<L, S>(a, x, y, z, v);
    const float T = a.Tuniform ? a.Tconst : a.Told[(int64_t)x + (int64_t)y * a.nx + (int64_t)z * a.nxy];
    return fpot(a, rho, T);
}

// Pulls split into issue and use, so that a lane's pulls for several sites are in flight
// together: issue_pulls loads the stored values of node (x, y, z)'s pulled populations
// (and, for f stored against per-node references, the references of their source nodes)
// into registers without using them; stored_pull turns one into the population as
// march_pull_from does.  FP16 values are held zero-extended in 32-bit registers: held as
// halves, the compiler packs two per register right after the loads, which waits for
// them.
template <typename S>
using raw_t = typename std::conditional<sizeof(S) == 2, unsigned, float>::type;

__device__ inline float half_bits(unsigned u) { return (float)__builtin_bit_cast(_Float16, (unsigned short)u); }

template <class L, typename S, bool REF, bool I32 = false>
__device__ inline void issue_pulls(const FArgs& a, const void* src, int x, int y, int z, raw_t<S>* v, unsigned* r) {
    const unsigned nx = (unsigned)a.nx, nxy = (unsigned)a.nxy;
    const unsigned X[3] = {(unsigned)wrapd(x - 1, a.nx), (unsigned)x, (unsigned)wrapd(x + 1, a.nx)};
    const unsigned Y[3] = {(unsigned)wrapd(y - 1, a.ny) * nx, (unsigned)y * nx, (unsigned)wrapd(y + 1, a.ny) * nx};
    const unsigned Z[3] = {(unsigned)wrapd(z - 1, a.nz) * nxy, (unsigned)z * nxy, (unsigned)wrapd(z + 1, a.nz) * nxy};
#pragma unroll
    for (int q = 0; q < L::Q; q++) {
        const unsigned up = X[1 - L::t.c[q][0]] + Y[1 - L::t.c[q][1]] + Z[1 - L::t.c[q][2]];
        // I32: 32-bit element index (pidx), else a 64-bit plane offset plus the node
        const int64_t e = I32 ? (int64_t)pidx(a, q, up) : (int64_t)q * a.P + up;
        if constexpr (sizeof(S) == 2) {
            v[q] = ((const unsigned short*)src)[e];
        } else {
            v[q] = ((const float*)src)[e];
        }
        if constexpr (REF && kNodeRef<S>) r[q] = ((const unsigned short*)a.rin)[up];
    }
}

template <class L, typename S, bool REF>
__device__ inline float stored_pull(raw_t<S> v, const unsigned* r, int q, const float* shift) {
    if constexpr (REF && kNodeRef<S>) {
        const float wref = (float)L::t.omega[q] * half_bits(r[q]);
        return fmaf(half_bits(v), wref, wref);
    } else if constexpr (sizeof(S) == 2) {
        return half_bits(v) + shift[q];
    } else {
        return v;
    }
}

// The populations pulled for psi at plane z + 1 stay in registers (fn) until plane z + 1
// collides; the register budget is MPLB_MARCH_WAVES waves per SIMD (three: about 150
// VGPRs, no spills; four waves spill and run 13% slower on the 256^3 bubble).
#ifndef MPLB_MARCH_WAVES
#define MPLB_MARCH_WAVES 3
#endif
// FP16 storage (FP16S, FP16N) runs four waves per SIMD (128 VGPRs, seven spilled): on the
// 256^3 bubble FP16S's march takes 0.319 against 0.330 ms at three waves (no spills),
// FP16N's 0.425 at either
#ifndef MPLB_MARCH_WAVES_H
#define MPLB_MARCH_WAVES_H 4
#endif
// wave priority while a kernel issues its population pulls (0 = lowest, 3 = highest)
#ifndef MPLB_PULL_PRIO
#define MPLB_PULL_PRIO 2
#endif

template <class L, typename S, bool OUT>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(sizeof(S) == 2 ? MPLB_MARCH_WAVES_H : MPLB_MARCH_WAVES))) k_fused_march(FArgs a) {
    __shared__ float ring[4][MSLOT];
    const unsigned wi = (blockIdx.x & 7u) * (unsigned)a.mper + (blockIdx.x >> 3);
    if (wi >= (unsigned)a.mitems) return;
    const int chunk = (int)(wi / (unsigned)a.mtiles), tile = (int)(wi - (unsigned)chunk * (unsigned)a.mtiles);
    const int x0 = (tile % a.mtx) * MX, y0 = (tile / a.mtx) * MY;
    const int z0 = chunk * a.mzc, z1 = min(z0 + a.mzc, a.nz);
    const int t = threadIdx.x, tx = t % MX, ty = t / MX;
    const int x = x0 + tx, y = y0 + ty;
    const bool owner = x < a.nx && y < a.ny;
    const bool edge = x <= a.nx && y <= a.ny;          // raw n wraps onto site 1
    const int ex = wrapd(x, a.nx), ey = wrapd(y, a.ny);
    // this thread's site on the ring around the column (threads < MRING)
    int rhx = 0, rhy = 0;
    if (t < MHX) {
        rhx = t;
    } else if (t < 2 * MHX) {
        rhx = t - MHX;
        rhy = MHY - 1;
    } else if (t < 2 * MHX + MY) {
        rhy = 1 + (t - 2 * MHX);
    } else {
        rhx = MHX - 1;
        rhy = 1 + (t - 2 * MHX - MY);
    }
    const bool ring_site = t < MRING && x0 + rhx - 1 <= a.nx && y0 + rhy - 1 <= a.ny;
    const int rx = wrapd(x0 + rhx - 1, a.nx), ry = wrapd(y0 + rhy - 1, a.ny);
    float fc[L::Q], fn[L::Q], tmp[L::Q];
    // psi of raw plane zr (z0 - 1 <= zr <= z1) into its ring slot; own populations -> fn.
    // A wave pulling a plane runs at raised priority (s_setprio), so the SIMD issues its
    // loads ahead of other waves' collision arithmetic and more loads are in flight
    // (8% faster on the 256^3 bubble).
    // FP16S issues the lane's column and ring pulls of a plane together (0.330 against
    // 0.340 ms per step on the 256^3 bubble); FP32 and FP16N pull and use the column site
    // first (together: 0.468 against 0.446, and 0.467 against 0.434)
    auto plane = [&](int zr) {
        if constexpr (std::is_same<S, _Float16>::value) {
            const int zw = wrapd(zr, a.nz);
            float* sl = ring[(zr - z0 + 1) & 3];
            raw_t<S> pe[L::Q], pr[L::Q];
            float te = a.Tconst, tr = a.Tconst;
            __builtin_amdgcn_s_setprio(MPLB_PULL_PRIO);
            if (edge) {
                issue_pulls<L, S, false>(a, a.fin, ex, ey, zw, pe, nullptr);
                if (!a.Tuniform) te = a.Told[(int64_t)ex + (int64_t)ey * a.nx + (int64_t)zw * a.nxy];
            }
            if (ring_site) {
                issue_pulls<L, S, false>(a, a.fin, rx, ry, zw, pr, nullptr);
                if (!a.Tuniform) tr = a.Told[(int64_t)rx + (int64_t)ry * a.nx + (int64_t)zw * a.nxy];
            }
            __builtin_amdgcn_s_setprio(0);
            if (edge) {
#pragma unroll
                for (int q = 0; q < L::Q; q++) fn[q] = stored_pull<L, S, false>(pe[q], nullptr, q, a.fshift);
                sl[(ty + 1) * MHX + tx + 1] = fpot(a, rho_sym<L, float>(fn), te);
            }
            if (ring_site) {
#pragma unroll
                for (int q = 0; q < L::Q; q++) tmp[q] = stored_pull<L, S, false>(pr[q], nullptr, q, a.fshift);
                sl[rhy * MHX + rhx] = fpot(a, rho_sym<L, float>(tmp), tr);
            }
            __builtin_amdgcn_s_waitcnt(kWaitVm);
        } else {
            const int zw = wrapd(zr, a.nz);
            float* sl = ring[(zr - z0 + 1) & 3];
            __builtin_amdgcn_s_setprio(MPLB_PULL_PRIO);
            if (edge) sl[(ty + 1) * MHX + tx + 1] = march_psi<L, S>(a, ex, ey, zw, fn);
            if (ring_site) sl[rhy * MHX + rhx] = march_psi<L, S>(a, rx, ry, zw, tmp);
            __builtin_amdgcn_s_setprio(0);
        }
    };
    plane(z0 - 1);
    plane(z0);
#pragma unroll
    for (int q = 0; q < L::Q; q++) fc[q] = fn[q];
    for (int z = z0; z < z1; z++) {
        plane(z + 1);
        __syncthreads();
        if (owner) {
            const float rho = rho_sym<L, float>(fc);
            float psin[L::Q];
#pragma unroll
            for (int q = 0; q < L::Q; q++)
                psin[q] = ring[(z + L::t.c[q][2] - z0 + 1) & 3][(ty + 1 + L::t.c[q][1]) * MHX + tx + 1 + L::t.c[q][0]];
            float U[3], Fi[3], Ft[3];
            collide_node<L>(a, fc, rho, psin, true, true, U, Fi, Ft);
            const int64_t id = (int64_t)x + (int64_t)y * a.nx + (int64_t)z * a.nxy;
            if constexpr (kNodeRef<S>) {
                float rinv;
                const float R = put_ref<S>(a, id, rho, rinv);
#pragma unroll
                for (int q = 0; q < L::Q; q++) stf_ref<L>(a.fout, q * a.P + id, q, fc[q], R, rinv);
            } else {
#pragma unroll
                for (int q = 0; q < L::Q; q++) st<S>(a.fout, q * a.P + id, fc[q], a.fshift[q]);
            }
            if (OUT) {
                const int64_t N = a.N;
                a.out_rho[id] = rho;
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    a.out_U[d * N + id] = U[d];
                    a.out_Fi[d * N + id] = Fi[d];
                    a.out_Ft[d * N + id] = Ft[d];
                }
            }
        }
#pragma unroll
        for (int q = 0; q < L::Q; q++) fc[q] = fn[q];
    }
}

// The same single pass with its pulls one plane ahead (software-pipelined, FP32 and
// FP16S storage): the loads of plane z + 2 (column and ring sites, plus T when it is not
// uniform) are issued before plane z collides and land in registers as stored values, so
// their latency runs under the collision of plane z; psi of plane z + 2 is formed before
// plane z is stored.  The arithmetic per node is that of k_fused_march.  Two waves per
// SIMD (the in-flight pulls hold 30 more registers).
#ifndef MPLB_MARCH_PF_WAVES
#define MPLB_MARCH_PF_WAVES 2
#endif

template <class L, typename S, bool OUT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MPLB_MARCH_PF_WAVES)))
k_fused_march_pf(FArgs a) {
    static_assert(!kNodeRef<S>, "FP32 or FP16S storage");
    using R = raw_t<S>;
    using M = typename std::conditional<sizeof(S) == 2, unsigned short, float>::type;   // stored element
    __shared__ float ring[4][MSLOT];
    const unsigned wi = (blockIdx.x & 7u) * (unsigned)a.mper + (blockIdx.x >> 3);
    if (wi >= (unsigned)a.mitems) return;
    const int chunk = (int)(wi / (unsigned)a.mtiles), tile = (int)(wi - (unsigned)chunk * (unsigned)a.mtiles);
    const int x0 = (tile % a.mtx) * MX, y0 = (tile / a.mtx) * MY;
    const int z0 = chunk * a.mzc, z1 = min(z0 + a.mzc, a.nz);
    const int t = threadIdx.x, tx = t % MX, ty = t / MX;
    const int x = x0 + tx, y = y0 + ty;
    const bool owner = x < a.nx && y < a.ny;
    const bool edge = x <= a.nx && y <= a.ny;          // raw n wraps onto site 1
    const int ex = wrapd(x, a.nx), ey = wrapd(y, a.ny);
    int rhx = 0, rhy = 0;
    if (t < MHX) {
        rhx = t;
    } else if (t < 2 * MHX) {
        rhx = t - MHX;
        rhy = MHY - 1;
    } else if (t < 2 * MHX + MY) {
        rhy = 1 + (t - 2 * MHX);
    } else {
        rhx = MHX - 1;
        rhy = 1 + (t - 2 * MHX - MY);
    }
    const bool ring_site = t < MRING && x0 + rhx - 1 <= a.nx && y0 + rhy - 1 <= a.ny;
    const int rx = wrapd(x0 + rhx - 1, a.nx), ry = wrapd(y0 + rhy - 1, a.ny);
    const unsigned nx = (unsigned)a.nx, nxy = (unsigned)a.nxy;
    const unsigned EX[3] = {(unsigned)wrapd(ex - 1, a.nx), (unsigned)ex, (unsigned)wrapd(ex + 1, a.nx)};
    const unsigned EY[3] = {(unsigned)wrapd(ey - 1, a.ny) * nx, (unsigned)ey * nx, (unsigned)wrapd(ey + 1, a.ny) * nx};
    const unsigned RX[3] = {(unsigned)wrapd(rx - 1, a.nx), (unsigned)rx, (unsigned)wrapd(rx + 1, a.nx)};
    const unsigned RY[3] = {(unsigned)wrapd(ry - 1, a.ny) * nx, (unsigned)ry * nx, (unsigned)wrapd(ry + 1, a.ny) * nx};
    float fc[L::Q], fn[L::Q];
    R pe[L::Q], pr[L::Q];              // pulls in flight: column site, ring site
    float te = a.Tconst, tr = a.Tconst;
    // issues the pulls of raw plane zr (stored values; nothing waits on them here)
    auto issue = [&](int zr) {
        const int zw = wrapd(zr, a.nz);
        const unsigned Z[3] = {(unsigned)wrapd(zw - 1, a.nz) * nxy, (unsigned)zw * nxy, (unsigned)wrapd(zw + 1, a.nz) * nxy};
        __builtin_amdgcn_s_setprio(MPLB_PULL_PRIO);
        if (edge) {
#pragma unroll
            for (int q = 0; q < L::Q; q++)
                pe[q] = ((const M*)a.fin)[(size_t)q * a.P +
                                          (EX[1 - L::t.c[q][0]] + EY[1 - L::t.c[q][1]] + Z[1 - L::t.c[q][2]])];
            if (!a.Tuniform) te = a.Told[(int64_t)ex + (int64_t)ey * a.nx + (int64_t)zw * a.nxy];
        }
        if (ring_site) {
#pragma unroll
            for (int q = 0; q < L::Q; q++)
                pr[q] = ((const M*)a.fin)[(size_t)q * a.P +
                                          (RX[1 - L::t.c[q][0]] + RY[1 - L::t.c[q][1]] + Z[1 - L::t.c[q][2]])];
            if (!a.Tuniform) tr = a.Told[(int64_t)rx + (int64_t)ry * a.nx + (int64_t)zw * a.nxy];
        }
        __builtin_amdgcn_s_setprio(0);
    };
    auto stored = [&](R v, int q) -> float { return stored_pull<L, S, false>(v, nullptr, q, a.fshift); };
    // psi of raw plane zr from the landed pulls into its ring slot; column populations -> fo.
    // It ends with an explicit wait for every outstanding memory operation (the pulls, and
    // the previous plane's stores, issued before this plane's collision): the compiler
    // cannot tell that a lane which skipped the pulls also skipped their use, and would
    // otherwise make the next plane's address arithmetic wait on the stores issued in
    // between (one counter for loads and stores, which complete in order).
    auto finish = [&](int zr, float* fo) {
        float* sl = ring[(zr - z0 + 1) & 3];
        float v[L::Q];
#pragma unroll
        for (int q = 0; q < L::Q; q++) {
            fo[q] = stored(pe[q], q);
            v[q] = stored(pr[q], q);
        }
        const float rho_e = rho_sym<L, float>(fo), rho_r = rho_sym<L, float>(v);
        if (edge) sl[(ty + 1) * MHX + tx + 1] = fpot(a, rho_e, te);
        if (ring_site) sl[rhy * MHX + rhx] = fpot(a, rho_r, tr);
        __builtin_amdgcn_s_waitcnt(kWaitVm);
    };
    issue(z0 - 1);
    finish(z0 - 1, fn);
    issue(z0);
    finish(z0, fc);
    issue(z0 + 1);
    finish(z0 + 1, fn);
    __syncthreads();
    for (int z = z0; z < z1; z++) {
        // ring slots: z - 1, z, z + 1 are read by this collision; z + 2 takes the slot of
        // z - 2, which every wave finished reading before the barrier that ended plane z - 1.
        // psi of z + 2 is formed before plane z is stored, so that waiting for its pulls
        // does not wait for those stores as well (one counter covers loads and stores).
        const bool more = z + 2 <= z1;
        if (more) issue(z + 2);
        float fnn[L::Q];
        float rho = 0.f, U[3], Fi[3], Ft[3];
        if (owner) {
            rho = rho_sym<L, float>(fc);
            float psin[L::Q];
#pragma unroll
            for (int q = 0; q < L::Q; q++)
                psin[q] = ring[(z + L::t.c[q][2] - z0 + 1) & 3][(ty + 1 + L::t.c[q][1]) * MHX + tx + 1 + L::t.c[q][0]];
            collide_node<L>(a, fc, rho, psin, true, true, U, Fi, Ft);
        }
        if (more) finish(z + 2, fnn);
        if (owner) {
            const int64_t id = (int64_t)x + (int64_t)y * a.nx + (int64_t)z * a.nxy;
#pragma unroll
            for (int q = 0; q < L::Q; q++) st<S>(a.fout, q * a.P + id, fc[q], a.fshift[q]);
            if (OUT) {
                const int64_t N = a.N;
                a.out_rho[id] = rho;
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    a.out_U[d * N + id] = U[d];
                    a.out_Fi[d * N + id] = Fi[d];
                    a.out_Ft[d * N + id] = Ft[d];
                }
            }
        }
#pragma unroll
        for (int q = 0; q < L::Q; q++) {
            fc[q] = fn[q];
            fn[q] = fnn[q];
        }
        __syncthreads();
    }
}


// ---------------------------------------------------------------------------
// Single-pass thermal step (xuMRT f + myMRT g with heat transfer) on fully periodic
// D3Q15 boxes: the three passes (density, heat source, collide) of multiPhasePP.c:176-240
// in one z-march, so f and g are pulled from memory once per step.  A workgroup owns an
// MX x MY column of sites over a chunk of z planes; for collision plane z it
//   1. pulls f on the column and a two-site ring around it at plane z + 2 (rho), and loads
//      the previous step's T and U there (LDS rings of three planes),
//   2. at plane z + 1, on the column and a one-site ring: heat source from the rho, T and U
//      stencils in LDS, g pulled for its sum, new temperature and psi (LDS ring),
//   3. collides plane z: f pulled two planes earlier (kept in LDS), psi stencil from LDS,
//      force and velocity, f collision; the own site's g pulled one plane earlier (kept in
//      registers), g collision; writes both, the new T and U.
// Per node and step HBM sees f and g read once and written once and T, U read once and
// written once.  The arithmetic of every node is that of k_fused_density, k_fused_heat
// and k_fused_collide (shared device functions); the compiler may still fuse a
// multiply-add differently in the two kernels, so results agree to FP32 rounding.
// Column of TMX x TMY sites, one thread each (64 x 8: 512 threads; the rings around the
// column hold 304 and 148 sites, one extra site for the first threads).  On the 256^3
// thermal bubble this kernel runs 64 x 8 at 1.095 ms per step, 32 x 16 at 1.174 and 32 x 8
// (256 threads, two workgroups per CU) at 1.437; 64 x 4 (three f pulls per thread and
// plane) ran at 1.86 against 1.48 for 64 x 8 in an earlier form of it.
#ifndef MPLB_TMX
#define MPLB_TMX 64
#define MPLB_TMY 8
#endif
constexpr int TMX = MPLB_TMX, TMY = MPLB_TMY, TNT = TMX * TMY;
constexpr int TRX = TMX + 4, TRY = TMY + 4, TNR = TRX * TRY;    // rho/T/U planes (two-site ring)
constexpr int TPX = TMX + 2, TPY = TMY + 2, TNP = TPX * TPY;    // psi planes (one-site ring)
constexpr int TRING2 = TNR - TNT, TRING1 = TNP - TNT;           // ring sites (64 x 8: 304 and 148)
static_assert(TRING2 <= TNT && TRING1 <= TNT, "at most one ring site per thread");

#ifndef MPLB_TMARCH_WAVES
#define MPLB_TMARCH_WAVES 2
#endif

// ring site r of the two-site ring -> (sx, sy) in the TRX x TRY plane
__device__ inline void tring2(int r, int& sx, int& sy) {
    if (r < 2 * TRX) {
        sy = r / TRX;
        sx = r % TRX;
    } else if (r < 2 * TRX + 4 * TMY) {
        const int rr = r - 2 * TRX;
        sy = 2 + rr / 4;
        const int c = rr % 4;
        sx = c < 2 ? c : TMX + c;
    } else {
        const int rr = r - 2 * TRX - 4 * TMY;
        sy = 2 + TMY + rr / TRX;
        sx = rr % TRX;
    }
}

// ring site r of the one-site ring -> (px, py) in the TPX x TPY plane
__device__ inline void tring1(int r, int& px, int& py) {
    if (r < TPX) {
        py = 0;
        px = r;
    } else if (r < TPX + 2 * TMY) {
        const int rr = r - TPX;
        py = 1 + rr / 2;
        px = (rr % 2) ? TPX - 1 : 0;
    } else {
        py = TPY - 1;
        px = r - TPX - 2 * TMY;
    }
}

__device__ inline int tslot(int zr) { return (zr + 3) % 3; }

template <class L, typename S, bool OUT>
__global__ void __launch_bounds__(TNT) __attribute__((amdgpu_waves_per_eu(MPLB_TMARCH_WAVES))) k_fused_tmarch(FArgs a) {
    __shared__ float s_rho[3][TNR], s_T[3][TNR], s_U[3][3][TNR];
    __shared__ float s_psi[3][TNP];
    const unsigned wi = (blockIdx.x & 7u) * (unsigned)a.mper + (blockIdx.x >> 3);
    if (wi >= (unsigned)a.mitems) return;
    const int chunk = (int)(wi / (unsigned)a.mtiles), tile = (int)(wi - (unsigned)chunk * (unsigned)a.mtiles);
    const int x0 = (tile % a.mtx) * TMX, y0 = (tile / a.mtx) * TMY;
    const int z0 = chunk * a.mzc, z1 = min(z0 + a.mzc, a.nz);
    const int t = threadIdx.x, tx = t % TMX, ty = t / TMX;
    const int x = x0 + tx, y = y0 + ty;
    const bool owner = x < a.nx && y < a.ny;
    const int64_t N = a.N;
    // ring sites go to the first threads (whole waves skip the extra work); sites needed
    // by some owner lie at raw x <= nx + 1, y <= ny + 1 (the period n - 1 wraps them)
    int r2x = 0, r2y = 0, r1x = 0, r1y = 0;
    if (t < TRING2) tring2(t, r2x, r2y);
    if (t < TRING1) tring1(t, r1x, r1y);
    const bool r2v = t < TRING2 && x0 - 2 + r2x <= a.nx + 1 && y0 - 2 + r2y <= a.ny + 1;
    const bool r1v = t < TRING1 && x0 - 1 + r1x <= a.nx && y0 - 1 + r1y <= a.ny;
    const bool own2 = x <= a.nx + 1 && y <= a.ny + 1;
    const bool own1 = x <= a.nx && y <= a.ny;

    // the own site's f of the planes waiting for their collision (each thread reads only
    // the slots it wrote)
    __shared__ float s_f[2][L::Q][TNT];
    float f0[L::Q], g0[L::Q], g1[L::Q], tmp[L::Q];
    float s0 = 0.f, T0 = 0.f, s1 = 0.f, T1 = 0.f;

    // step 1 for plane zr, in two halves: issue_plane loads the pulled f of the column's
    // sites and of the two-site ring with their previous T and U (every load of the lane
    // issued before any is used); finish_plane turns them into rho, T and U in the LDS
    // planes and the own site's populations into its s_f slot.  In the march the loads of
    // plane z + 3 are issued before plane z collides, so they land under the collision.
    constexpr int QR = kNodeRef<S> ? L::Q : 1;
    raw_t<S> po[L::Q], pr[L::Q];
    unsigned ro[QR], rr[QR];
    float To = 0.f, Uo0 = 0.f, Uo1 = 0.f, Uo2 = 0.f, Tr = 0.f, Ur0 = 0.f, Ur1 = 0.f, Ur2 = 0.f;
    const int ogx = wrapd(x0 + tx, a.nx), ogy = wrapd(y0 + ty, a.ny);
    const int rgx = wrapd(x0 - 2 + r2x, a.nx), rgy = wrapd(y0 - 2 + r2y, a.ny);
    auto issue_plane = [&](int zr) __attribute__((always_inline)) {
        const int gz = wrapd(zr, a.nz);
        const int64_t ido = (int64_t)ogx + (int64_t)ogy * a.nx + (int64_t)gz * a.nxy;
        const int64_t idr = (int64_t)rgx + (int64_t)rgy * a.nx + (int64_t)gz * a.nxy;
        __builtin_amdgcn_s_setprio(MPLB_PULL_PRIO);
        if (own2) {
            issue_pulls<L, S, true, true>(a, a.fin, ogx, ogy, gz, po, ro);
            To = a.Told[ido];
            Uo0 = a.Uold[ido];
            Uo1 = a.Uold[N + ido];
            Uo2 = a.Uold[2 * N + ido];
        }
        if (r2v) {
            issue_pulls<L, S, true, true>(a, a.fin, rgx, rgy, gz, pr, rr);
            Tr = a.Told[idr];
            Ur0 = a.Uold[idr];
            Ur1 = a.Uold[N + idr];
            Ur2 = a.Uold[2 * N + idr];
        }
        __builtin_amdgcn_s_setprio(0);
    };
    auto finish_plane = [&](int zr) __attribute__((always_inline)) {
        const int sl = tslot(zr);
        if (own2) {
            float f2[L::Q];
#pragma unroll
            for (int q = 0; q < L::Q; q++) {
                f2[q] = stored_pull<L, S, true>(po[q], ro, q, a.fshift);
                s_f[zr & 1][q][t] = f2[q];
            }
            const int i = (ty + 2) * TRX + tx + 2;
            s_rho[sl][i] = rho_sym<L, float>(f2);
            s_T[sl][i] = To;
            s_U[sl][0][i] = Uo0;
            s_U[sl][1][i] = Uo1;
            s_U[sl][2][i] = Uo2;
        }
        if (r2v) {
#pragma unroll
            for (int q = 0; q < L::Q; q++) tmp[q] = stored_pull<L, S, true>(pr[q], rr, q, a.fshift);
            const int i = r2y * TRX + r2x;
            s_rho[sl][i] = rho_sym<L, float>(tmp);
            s_T[sl][i] = Tr;
            s_U[sl][0][i] = Ur0;
            s_U[sl][1][i] = Ur1;
            s_U[sl][2][i] = Ur2;
        }
        // nothing of this phase is in flight any more; saying so keeps the compiler (which
        // cannot pair a lane's skipped loads with its skipped uses) from waiting on the
        // collision's stores when it next reuses these registers
        __builtin_amdgcn_s_waitcnt(kWaitVm);
    };
    auto load_plane = [&](int zr) __attribute__((always_inline)) {
        issue_plane(zr);
        finish_plane(zr);
    };
    // heat source of site (px, py) of the one-site plane zr from the rho, T and U stencils
    // in LDS
    auto heat_site = [&](int px, int py, int zr, float& rho) __attribute__((always_inline)) {
        const int sl = tslot(zr), sm = tslot(zr - 1), sp = tslot(zr + 1);
        const int i = (py + 1) * TRX + px + 1;
        rho = s_rho[sl][i];
        const float Told = s_T[sl][i];
        float dot = 0.f, divU = 0.f;
        heat_axis(s_T[sl][i + 1], s_T[sl][i - 1], s_rho[sl][i + 1], s_rho[sl][i - 1], s_U[sl][0][i + 1],
                  s_U[sl][0][i - 1], dot, divU);
        heat_axis(s_T[sl][i + TRX], s_T[sl][i - TRX], s_rho[sl][i + TRX], s_rho[sl][i - TRX], s_U[sl][1][i + TRX],
                  s_U[sl][1][i - TRX], dot, divU);
        heat_axis(s_T[sp][i], s_T[sm][i], s_rho[sp][i], s_rho[sm][i], s_U[sp][2][i], s_U[sm][2][i], dot, divU);
        return heat_tail<L>(a, rho, Told, dot, divU);
    };
    // step 2 for plane zr: heat source, new temperature and psi of the column's sites and
    // of the one-site ring; the own site's g populations, source and temperature go to g1,
    // s1, T1.  The lane's g pulls (own and ring site) are issued before the heat sources
    // are formed, so their latency runs under that arithmetic.
    auto psi_plane = [&](int zr) __attribute__((always_inline)) {
        const int gz = wrapd(zr, a.nz), sl = tslot(zr);
        raw_t<S> go[L::Q], gr[L::Q];
        __builtin_amdgcn_s_setprio(MPLB_PULL_PRIO);
        if (own1) issue_pulls<L, S, false, true>(a, a.gin, wrapd(x0 + tx, a.nx), wrapd(y0 + ty, a.ny), gz, go, nullptr);
        if (r1v) issue_pulls<L, S, false, true>(a, a.gin, wrapd(x0 - 1 + r1x, a.nx), wrapd(y0 - 1 + r1y, a.ny), gz, gr, nullptr);
        __builtin_amdgcn_s_setprio(0);
        float rho_o = 0.f, rho_r = 0.f, so = 0.f, sr = 0.f;
        if (own1) so = heat_site(tx + 1, ty + 1, zr, rho_o);
        if (r1v) sr = heat_site(r1x, r1y, zr, rho_r);
        if (own1) {
#pragma unroll
            for (int q = 0; q < L::Q; q++) g1[q] = stored_pull<L, S, false>(go[q], nullptr, q, a.gshift);
            const float T = rho_sym<L, float>(g1) + 0.5f * so;
            s_psi[sl][(ty + 1) * TPX + tx + 1] = fpot(a, rho_o, T);
            s1 = so;
            T1 = T;
        }
        if (r1v) {
#pragma unroll
            for (int q = 0; q < L::Q; q++) tmp[q] = stored_pull<L, S, false>(gr[q], nullptr, q, a.gshift);
            const float T = rho_sym<L, float>(tmp) + 0.5f * sr;
            s_psi[sl][r1y * TPX + r1x] = fpot(a, rho_r, T);
        }
        __builtin_amdgcn_s_waitcnt(kWaitVm);
    };

    load_plane(z0 - 2);
    load_plane(z0 - 1);
    load_plane(z0);
    __syncthreads();
    psi_plane(z0 - 1);
    __syncthreads();
    load_plane(z0 + 1);
    __syncthreads();
    psi_plane(z0);
#pragma unroll
    for (int q = 0; q < L::Q; q++) g0[q] = g1[q];
    s0 = s1;
    T0 = T1;
    __syncthreads();
    issue_plane(z0 + 2);
    for (int z = z0; z < z1; z++) {
        // plane z's populations leave their slot before plane z + 2's take it
#pragma unroll
        for (int q = 0; q < L::Q; q++) f0[q] = s_f[z & 1][q][t];
        finish_plane(z + 2);
        __syncthreads();
        psi_plane(z + 1);
        __syncthreads();
        if (z + 3 <= z1 + 1) issue_plane(z + 3);
        if (owner) {
            const float rho = rho_sym<L, float>(f0);
            float psin[L::Q];
#pragma unroll
            for (int q = 0; q < L::Q; q++)
                psin[q] = s_psi[tslot(z + L::t.c[q][2])][(ty + 1 + L::t.c[q][1]) * TPX + tx + 1 + L::t.c[q][0]];
            float U[3], Fi[3], Ft[3];
            collide_node<L>(a, f0, rho, psin, true, true, U, Fi, Ft);
            const int64_t id = (int64_t)x + (int64_t)y * a.nx + (int64_t)z * a.nxy;
            if constexpr (kNodeRef<S>) {
                float rinv;
                const float R = put_ref<S>(a, id, rho, rinv);
#pragma unroll
                for (int q = 0; q < L::Q; q++) stf_ref<L>(a.fout, pidx(a, q, (unsigned)id), q, f0[q], R, rinv);
            } else {
#pragma unroll
                for (int q = 0; q < L::Q; q++) st<S>(a.fout, pidx(a, q, (unsigned)id), f0[q], a.fshift[q]);
            }
            collide_g<L>(a, g0, T0, s0, U);
#pragma unroll
            for (int q = 0; q < L::Q; q++) st<S>(a.gout, pidx(a, q, (unsigned)id), g0[q], a.gshift[q]);
            a.Tnew[id] = T0;
            a.Unew[id] = U[0];
            a.Unew[N + id] = U[1];
            a.Unew[2 * N + id] = U[2];
            if (OUT) {
                a.S[id] = s0;
                a.out_rho[id] = rho;
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    a.out_U[d * N + id] = U[d];
                    a.out_Fi[d * N + id] = Fi[d];
                    a.out_Ft[d * N + id] = Ft[d];
                }
            }
        }
#pragma unroll
        for (int q = 0; q < L::Q; q++) g0[q] = g1[q];
        s0 = s1;
        T0 = T1;
    }
}

// ---------------------------------------------------------------------------
// The isothermal single pass with its pulls staged in LDS (FP32 storage).  Plane z + 2's
// populations for the column and its ring go straight from memory into LDS
// (global_load_lds_dword, no registers held while in flight) and arrive while the
// workgroup turns plane z + 1's stage into psi and collides plane z.  Two stages of
// 15 x 396 floats (47.5 KB) and the psi ring (6.3 KB) leave room for three workgroups
// per CU.  The workgroup meets with a bare s_barrier: a fence would wait for the stage
// in flight (the LDS writes of the direct loads count as vector-memory operations).  The
// wait before a stage is read is vmcnt(0), the previous plane's stores included: with
// loads and stores mixed the vector-memory counter need not count down in issue order.
constexpr int LST = MX * MY + MRING;       // staged sites: the column, then its ring
#ifndef MPLB_MARCH_LDS_WAVES
#define MPLB_MARCH_LDS_WAVES 3
#endif

typedef __attribute__((address_space(3))) void lds_void;

// the workgroup meets; the memory clobber keeps the compiler from moving LDS accesses
// across it
__device__ inline void wg_barrier() { asm volatile("s_barrier" ::: "memory"); }

// direct loads of node (x, y, z)'s pulled populations into stage[q][slot]; the wave's
// 64 lanes fill 64 consecutive floats from the wave's first slot
template <class L>
__device__ inline void stage_pull(const FArgs& a, int x, int y, int z, float (*stage)[LST], int wave_slot) {
    const unsigned nx = (unsigned)a.nx, nxy = (unsigned)a.nxy;
    const unsigned X[3] = {(unsigned)wrapd(x - 1, a.nx), (unsigned)x, (unsigned)wrapd(x + 1, a.nx)};
    const unsigned Y[3] = {(unsigned)wrapd(y - 1, a.ny) * nx, (unsigned)y * nx, (unsigned)wrapd(y + 1, a.ny) * nx};
    const unsigned Z[3] = {(unsigned)wrapd(z - 1, a.nz) * nxy, (unsigned)z * nxy, (unsigned)wrapd(z + 1, a.nz) * nxy};
#pragma unroll
    for (int q = 0; q < L::Q; q++) {
        const float* src = (const float*)a.fin + (size_t)q * a.P +
                           (X[1 - L::t.c[q][0]] + Y[1 - L::t.c[q][1]] + Z[1 - L::t.c[q][2]]);
        __builtin_amdgcn_global_load_lds((const void*)src, (lds_void*)&stage[q][wave_slot], 4, 0, 0);
    }
}

template <class L, bool OUT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MPLB_MARCH_LDS_WAVES)))
k_fused_march_lds(FArgs a) {
    __shared__ float stage0[L::Q][LST], stage1[L::Q][LST];
    __shared__ float ring[4][MSLOT];
    const unsigned wi = (blockIdx.x & 7u) * (unsigned)a.mper + (blockIdx.x >> 3);
    if (wi >= (unsigned)a.mitems) return;
    const int chunk = (int)(wi / (unsigned)a.mtiles), tile = (int)(wi - (unsigned)chunk * (unsigned)a.mtiles);
    const int x0 = (tile % a.mtx) * MX, y0 = (tile / a.mtx) * MY;
    const int z0 = chunk * a.mzc, z1 = min(z0 + a.mzc, a.nz);
    const int t = threadIdx.x, tx = t % MX, ty = t / MX;
    const int x = x0 + tx, y = y0 + ty;
    const bool owner = x < a.nx && y < a.ny;
    const bool edge = x <= a.nx && y <= a.ny;          // raw n wraps onto site 1
    // every lane of a wave takes part in a direct load (its LDS slot follows from its lane
    // number), so threads off the lattice load site 0's populations and never use them
    const int ex = edge ? wrapd(x, a.nx) : 0, ey = edge ? wrapd(y, a.ny) : 0;
    int rhx = 0, rhy = 0;
    if (t < MHX) {
        rhx = t;
    } else if (t < 2 * MHX) {
        rhx = t - MHX;
        rhy = MHY - 1;
    } else if (t < 2 * MHX + MY) {
        rhy = 1 + (t - 2 * MHX);
    } else {
        rhx = MHX - 1;
        rhy = 1 + (t - 2 * MHX - MY);
    }
    const bool ring_site = t < MRING && x0 + rhx - 1 <= a.nx && y0 + rhy - 1 <= a.ny;
    const int rx = ring_site ? wrapd(x0 + rhx - 1, a.nx) : 0, ry = ring_site ? wrapd(y0 + rhy - 1, a.ny) : 0;
    const int own_slot = t & ~63, ring_slot = MX * MY + (t & ~63);
    float fc[L::Q], fn[L::Q], tmp[L::Q];

    // temperatures of the staged planes (own and ring site), loaded with their stage:
    // a load issued after the stage's direct loads would make its first use wait for
    // all of them (the counter counts down in order)
    float Te0 = a.Tconst, Tr0 = a.Tconst, Te1 = a.Tconst, Tr1 = a.Tconst;
    // direct loads of raw plane zr (own site; ring site for threads < MRING)
    auto issue = [&](int zr, float (*stage)[LST], float& Te, float& Tr) {
        const int zw = wrapd(zr, a.nz);
        if (!a.Tuniform) {
            Te = a.Told[(int64_t)ex + (int64_t)ey * a.nx + (int64_t)zw * a.nxy];
            if (t < MRING) Tr = a.Told[(int64_t)rx + (int64_t)ry * a.nx + (int64_t)zw * a.nxy];
        }
        stage_pull<L>(a, ex, ey, zw, stage, own_slot);
        if (t < MRING) stage_pull<L>(a, rx, ry, zw, stage, ring_slot);
    };
    // psi of raw plane zr from its stage into its ring slot; own populations -> fn
    auto psi_from = [&](int zr, float (*stage)[LST], float Te, float Tr) {
        float* sl = ring[(zr - z0 + 1) & 3];
#pragma unroll
        for (int q = 0; q < L::Q; q++) fn[q] = stage[q][t];
        if (edge) sl[(ty + 1) * MHX + tx + 1] = fpot(a, rho_sym<L, float>(fn), Te);
        if (ring_site) {
#pragma unroll
            for (int q = 0; q < L::Q; q++) tmp[q] = stage[q][MX * MY + t];
            sl[rhy * MHX + rhx] = fpot(a, rho_sym<L, float>(tmp), Tr);
        }
    };
    // this wave's loads and stores complete, then the workgroup meets: every wave's direct
    // loads have landed and every wave is done reading the stage that is loaded next
    auto sync_stage = [&]() {
        __builtin_amdgcn_s_waitcnt(kWaitAll);
        wg_barrier();
    };
    // LDS writes complete, then the workgroup meets; direct loads stay in flight
    auto sync_ring = [&]() {
        __builtin_amdgcn_s_waitcnt(kWaitLds);
        wg_barrier();
    };
    // plane z's collision, in place in fc (its stores wait for the next plane)
    auto collide = [&](int z) {
        if (!owner) return;
        const float rho = rho_sym<L, float>(fc);
        float psin[L::Q];
#pragma unroll
        for (int q = 0; q < L::Q; q++)
            psin[q] = ring[(z + L::t.c[q][2] - z0 + 1) & 3][(ty + 1 + L::t.c[q][1]) * MHX + tx + 1 + L::t.c[q][0]];
        float U[3], Fi[3], Ft[3];
        collide_node<L>(a, fc, rho, psin, true, true, U, Fi, Ft);
        if (OUT) {
            const int64_t id = (int64_t)x + (int64_t)y * a.nx + (int64_t)z * a.nxy;
            const int64_t N = a.N;
            a.out_rho[id] = rho;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                a.out_U[d * N + id] = U[d];
                a.out_Fi[d * N + id] = Fi[d];
                a.out_Ft[d * N + id] = Ft[d];
            }
        }
    };
    // plane z's collided populations (in fc) to memory
    auto store_post = [&](int z) {
        if (!owner) return;
        const int64_t id = (int64_t)x + (int64_t)y * a.nx + (int64_t)z * a.nxy;
#pragma unroll
        for (int q = 0; q < L::Q; q++) st<float>(a.fout, q * a.P + id, fc[q], 0.f);
    };
    // Plane zr is staged in stage0 when zr - z0 is odd (z0 - 1, z0 + 1, ...), in stage1
    // when it is even; the loop takes two planes per trip so that both are named
    // statically.  A plane's stores are issued right after the next stage's loads, so
    // both have a whole plane's work to complete before the wait that covers them.
    issue(z0 - 1, stage0, Te0, Tr0);
    issue(z0, stage1, Te1, Tr1);
    sync_stage();
    psi_from(z0 - 1, stage0, Te0, Tr0);
    psi_from(z0, stage1, Te1, Tr1);
    sync_ring();                                       // stage0 read by every wave
    issue(z0 + 1, stage0, Te0, Tr0);
    for (int z = z0; z < z1; z += 2) {
        sync_stage();                                  // plane z + 1 in stage0
        if (z + 2 <= z1) issue(z + 2, stage1, Te1, Tr1);
        if (z > z0) store_post(z - 1);
#pragma unroll
        for (int q = 0; q < L::Q; q++) fc[q] = fn[q];
        psi_from(z + 1, stage0, Te0, Tr0);
        sync_ring();
        collide(z);
        if (z + 1 >= z1) break;
        sync_stage();                                  // plane z + 2 in stage1
        if (z + 3 <= z1) issue(z + 3, stage0, Te0, Tr0);
        store_post(z);
#pragma unroll
        for (int q = 0; q < L::Q; q++) fc[q] = fn[q];
        psi_from(z + 2, stage1, Te1, Tr1);
        sync_ring();
        collide(z + 1);
    }
    store_post(z1 - 1);
}

// Post-stream, post-boundary populations of the current time level (what the
// reference holds in f.value after updateBoundaries), FP32 SoA.
template <class L, typename S, bool G>
__global__ void __launch_bounds__(256) k_fused_materialize(FArgs a, float* out) {
    const NodePos p = node_pos<L>(a);
    if (!p.valid) return;
    float v[L::Q];
    if (G) pull_g<L, S>(a, p, v);
    else pull_f<L, S>(a, p, v);
#pragma unroll
    for (int q = 0; q < L::Q; q++) out[q * a.N + p.id] = v[q];
}

// FP32 SoA g -> storage (with the global shifts)
template <typename S>
__global__ void k_to_storage(const float* __restrict__ in, void* out, int64_t N, int Q, FArgs a) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= N * Q) return;
    const int q = (int)(i / N);
    st<S>(out, q * a.P + (i - q * N), in[i], a.gshift[q]);
}

// FP32 SoA f -> storage in a.fout; FP16 storage writes each node's reference (the sum of
// its populations) to a.rout and the deviations from it
template <class L, typename S>
__global__ void __launch_bounds__(256) k_to_storage_f(const float* __restrict__ in, FArgs a) {
    const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id >= a.N) return;
    float v[L::Q];
#pragma unroll
    for (int q = 0; q < L::Q; q++) v[q] = in[q * a.N + id];
    if constexpr (kNodeRef<S>) {
        float rinv;
        const float R = put_ref<S>(a, id, rho_sym<L, float>(v), rinv);
#pragma unroll
        for (int q = 0; q < L::Q; q++) stf_ref<L>(a.fout, q * a.P + id, q, v[q], R, rinv);
    } else {
#pragma unroll
        for (int q = 0; q < L::Q; q++) st<S>(a.fout, q * a.P + id, v[q], a.fshift[q]);
    }
}

struct Fused {
    int lb = 0;
    bool thermal = false;
    bool half = false;
    void* fa = nullptr;
    void* fb = nullptr;
    void* ga = nullptr;
    void* gb = nullptr;
    float* rho = nullptr;
    float* sumg = nullptr;
    float* T[2] = {nullptr, nullptr};
    float* U[2] = {nullptr, nullptr};
    int cur = 0;                   // T[cur], U[cur] hold the current time level
    uint8_t* bcf = nullptr;
    uint8_t* bcg = nullptr;
    float* bcTw = nullptr;
    float* bcPert = nullptr;
    int* rnd_off = nullptr;
    std::vector<std::pair<int, int>> frdt_lists;   // (field, list) of fixedRandomT
    bool need_prologue = true;     // post-collision buffers must be rebuilt from the generic state
    bool generic_current = true;   // generic device arrays describe the current time level
    int last_draw = 0;             // fixedRandomT drawing lists of the last step (bit per slot)
    float rho0_built = 0.f;        // rho_0 the post-collision buffers were computed with
    int64_t P = 0;                 // population plane pitch
    int face_table = 0;            // faces whose links differ from the periodic wrap
    float fshift[15] = {0}, gshift[15] = {0};
    bool noderef = false;          // FP16N: f against per-node references
    _Float16* ra = nullptr;        // FP16N: per-node references of fa and fb
    _Float16* rb = nullptr;
    bool march = false;            // single-pass kernel (isothermal, fully periodic D3Q15)
    bool tmarch = false;           // thermal single-pass kernel (xuMRT + myMRT, fully periodic D3Q15)
    bool march_lds = false;        // the isothermal single pass with LDS-staged pulls (FP32)
    bool march_pf = false;         // the isothermal single pass with its pulls a plane ahead
    bool t_uniform = false;        // the isothermal temperature field is one value (t_const)
    float t_const = 0.f;
    int mtx = 0, mtiles = 0, mitems = 0, mper = 0, mzc = 0;
};

Fused* F(mplb_case* c) { return (Fused*)c->fz; }

#ifndef MPLB_PITCH_PAD
#define MPLB_PITCH_PAD 0
#endif
// Population planes are MPLB_PITCH_PAD elements longer than the lattice so that the
// Q concurrent streams of a pull do not sit at a power-of-two distance apart.
int64_t pitch_for(int64_t N) { return ((N + 63) / 64) * 64 + MPLB_PITCH_PAD; }

int err(mplb_case* c, int code, const std::string& m) {
    c->err = m;
    return code;
}

#define FCHK(c, call)                                                                            \
    do {                                                                                         \
        hipError_t _e = (call);                                                                  \
        if (_e != hipSuccess) return err((c), MPLB_E_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

int reverse_of(const mplb_case* c, int k) { return c->mesh.lbmodel == 0 ? D2Q9::t.rev[k] : D3Q15::t.rev[k]; }

FArgs make_args(mplb_case* c) {
    Fused* z = F(c);
    FArgs a;
    std::memset(&a, 0, sizeof(a));
    const Mesh& m = c->mesh;
    a.nx = m.nx;
    a.ny = m.ny;
    a.nz = m.nz;
    a.N = m.N;
    a.P = z->P;
    a.nxy = (int64_t)m.nx * m.ny;
    a.nblocks = (int)((m.N + 255) / 256);
    a.face_table = z->face_table;
    a.nb = c->nb;
    a.onbnd = c->onbnd;
    a.bcf = z->bcf;
    a.bcg = z->bcg;
    a.bcTw = z->bcTw;
    a.bcPert = z->bcPert;
    a.rnd_off = z->rnd_off;
    a.rnd = c->rnd;
    a.eos = c->eos;
    a.rho0 = (float)c->rho0;
    const ModelParams& fp = c->fc[0].mp;
    a.fmodel = fp.model == M_LIMRT ? 0 : 1;
    a.surf = fp.surf;
    a.sigma = (float)fp.sigma;
    a.kappa = (float)fp.kappa;
    for (int q = 0; q < 15; q++) a.fl[q] = (float)fp.Lambda[q];
    if (m.lbmodel == 0) {
        a.sf1 = (float)(12.0 * fp.sigma / (1.0 / fp.Lambda[1] - 0.5));
        a.sf2 = (float)(12.0 * fp.sigma / (1.0 / fp.Lambda[2] - 0.5));
    } else {
        a.sf1 = (float)(6.0 * fp.sigma / (1.0 / fp.Lambda[1] - 0.5));
    }
    a.vdw = c->eos.type == EOS_VDW && c->eos.G != 0.0 ? 1 : 0;
    a.vR = (float)c->eos.R;
    a.vb = (float)c->eos.b;
    a.va_m2 = (float)(c->eos.a / (c->eos.M * c->eos.M));
    a.vinv_g = c->eos.G != 0.0 ? (float)(1.0 / c->eos.G) : 0.f;
    a.vM = (float)c->eos.M;
    if (c->fc[1].present) {
        const ModelParams& gp = c->fc[1].mp;
        a.gmodel = gp.model == M_MYMRT ? 1 : 3;
        a.a1 = (float)gp.alpha1;
        a.a2 = (float)gp.alpha2;
        for (int q = 0; q < 15; q++) a.gl[q] = (float)gp.Lambda[q];
    }
    a.psi = (float*)c->psi;
    a.S = (float*)c->ss;
    a.rho = z->rho;
    a.sumg = z->sumg;
    a.Told = z->thermal ? z->T[z->cur] : (const float*)c->temp;
    a.Tnew = z->thermal ? z->T[1 - z->cur] : nullptr;
    a.Uold = z->thermal ? z->U[z->cur] : (const float*)c->U;
    a.Unew = z->thermal ? z->U[1 - z->cur] : nullptr;
    a.Tuniform = (!z->thermal && z->t_uniform) ? 1 : 0;
    a.Tconst = z->t_const;
    a.mtx = z->mtx;
    a.mtiles = z->mtiles;
    a.mitems = z->mitems;
    a.mper = z->mper;
    a.mzc = z->mzc;
    std::memcpy(a.fshift, z->fshift, sizeof(a.fshift));
    a.rin = z->ra;
    a.rout = z->rb;
    std::memcpy(a.gshift, z->gshift, sizeof(a.gshift));
    return a;
}

// The single-pass kernel (D3Q15, FP32 or FP16 storage)
template <class L, typename S, bool OUT>
void launch_march(mplb_case* c, const FArgs& a) {
    if constexpr (L::D == 3) {
        const dim3 grid(8u * (unsigned)a.mper);
        if constexpr (std::is_same<S, float>::value) {
            if (F(c)->march_lds) {
                k_fused_march_lds<L, OUT><<<grid, 256, 0, c->stream>>>(a);
                return;
            }
        }
        if constexpr (!kNodeRef<S>) {
            if (F(c)->march_pf) {
                k_fused_march_pf<L, S, OUT><<<grid, 256, 0, c->stream>>>(a);
                return;
            }
        }
        k_fused_march<L, S, OUT><<<grid, 256, 0, c->stream>>>(a);
    }
}

// The thermal single-pass kernel (D3Q15, FP32 or FP16 storage)
template <class L, typename S, bool OUT>
void launch_tmarch(mplb_case* c, const FArgs& a) {
    if constexpr (L::D == 3) k_fused_tmarch<L, S, OUT><<<dim3(8u * (unsigned)a.mper), TNT, 0, c->stream>>>(a);
}

template <class L, typename S>
void launch_step(mplb_case* c, FArgs a, bool thermal, bool outputs) {
    const dim3 grid((unsigned)a.nblocks), block(256);
    if (outputs) {
        a.out_rho = (float*)c->rho;
        a.out_Fi = (float*)c->Fi;
        a.out_Ft = (float*)c->Ft;
        if (!thermal) a.out_U = (float*)c->U;
    }
    if (thermal && F(c)->tmarch) {
        launch_tmarch<L, S, false>(c, a);
    } else if (thermal) {
        k_fused_density<L, S, true><<<grid, block, 0, c->stream>>>(a);
        k_fused_heat<L><<<grid, block, 0, c->stream>>>(a);
        k_fused_collide<L, S, true><<<grid, block, 0, c->stream>>>(a);
    } else if (F(c)->march) {
        launch_march<L, S, false>(c, a);
    } else {
        k_fused_density<L, S, false><<<grid, block, 0, c->stream>>>(a);
        k_fused_collide<L, S, false><<<grid, block, 0, c->stream>>>(a);
    }
}

// Time each kernel separately when the case is timing.
struct Ev {
    mplb_case* c;
    int idx;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    Ev(mplb_case* cc, int i) : c(cc), idx(i) {
        if (!c->timing) return;
        // timing only: a failed event leaves that kernel class untimed
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        (void)hipEventRecord(e0, c->stream);
    }
    ~Ev() {
        if (!c->timing) return;
        (void)hipEventRecord(e1, c->stream);
        c->ev_pool.push_back(e0);
        c->ev_pool.push_back(e1);
        c->ev_pool.push_back((hipEvent_t)(intptr_t)idx);
    }
};

template <class L, typename S>
void launch_step_timed(mplb_case* c, FArgs a, bool thermal, bool outputs) {
    const dim3 grid((unsigned)a.nblocks), block(256);
    if (outputs) {
        a.out_rho = (float*)c->rho;
        a.out_Fi = (float*)c->Fi;
        a.out_Ft = (float*)c->Ft;
        if (!thermal) a.out_U = (float*)c->U;
    }
    // timer slots: 6 = fused collide (dominant), 7 = density/heat passes
    if (thermal && F(c)->tmarch) {
        Ev e(c, 6);
        launch_tmarch<L, S, false>(c, a);
    } else if (thermal) {
        {
            Ev e(c, 7);
            k_fused_density<L, S, true><<<grid, block, 0, c->stream>>>(a);
        }
        {
            Ev e(c, 5);
            k_fused_heat<L><<<grid, block, 0, c->stream>>>(a);
        }
        Ev e(c, 6);
        k_fused_collide<L, S, true><<<grid, block, 0, c->stream>>>(a);
    } else if (F(c)->march) {
        Ev e(c, 6);
        launch_march<L, S, false>(c, a);
    } else {
        {
            Ev e(c, 7);
            k_fused_density<L, S, false><<<grid, block, 0, c->stream>>>(a);
        }
        Ev e(c, 6);
        k_fused_collide<L, S, false><<<grid, block, 0, c->stream>>>(a);
    }
}

template <class L, typename S>
int run_prologue(mplb_case* c) {
    // generic state (post-stream populations and macroscopic fields of the current time
    // level) -> post-collision buffers of the next step
    Fused* z = F(c);
    DevState<float> st;
    st.N = c->mesh.N;
    st.nb = c->nb;
    st.onbnd = c->onbnd;
    st.rho = (float*)c->rho;
    st.temp = (float*)c->temp;
    st.p = (float*)c->p;
    st.U = (float*)c->U;
    st.Fi = (float*)c->Fi;
    st.Ft = (float*)c->Ft;
    st.ss = (float*)c->ss;
    st.psi = (float*)c->psi;
    GenericOps<float>::psi(c->mesh.lbmodel, c->stream, c->eos, st);
    const int64_t N = c->mesh.N;
    const int Q = c->mesh.Q;
    const size_t bytes = sizeof(float) * Q * N;
    // collide copies of f and g (the generic arrays keep the current level)
    struct Scratch {
        float* p = nullptr;
        ~Scratch() {
            if (p) (void)hipFree(p);
        }
    } scratch;
    FCHK(c, hipMalloc(&scratch.p, bytes));
    float* tmp = scratch.p;
    FArgs a = make_args(c);
    for (int fld = 0; fld < (z->thermal ? 2 : 1); fld++) {
        FCHK(c, hipMemcpyAsync(tmp, c->f[fld], bytes, hipMemcpyDeviceToDevice, c->stream));
        GenericOps<float>::collide(c->mesh.lbmodel, c->stream, c->eos, c->fc[fld].mp, st, tmp, c->rho0);
        if (fld == 0) {
            // fa (with its references ra when FP16) is the input of the next step
            a.fout = z->fa;
            a.rout = z->ra;
            k_to_storage_f<L, S><<<(unsigned)((N + 255) / 256), 256, 0, c->stream>>>(tmp, a);
        } else {
            k_to_storage<S><<<(unsigned)((N * Q + 255) / 256), 256, 0, c->stream>>>(tmp, z->ga, N, Q, a);
        }
    }
    if (z->thermal) {
        FCHK(c, hipMemcpyAsync(z->T[z->cur], c->temp, sizeof(float) * N, hipMemcpyDeviceToDevice, c->stream));
        FCHK(c, hipMemcpyAsync(z->U[z->cur], c->U, sizeof(float) * 3 * N, hipMemcpyDeviceToDevice, c->stream));
    }
    FCHK(c, hipStreamSynchronize(c->stream));
    z->need_prologue = false;
    z->generic_current = true;
    z->rho0_built = (float)c->rho0;
    return 0;
}

template <class L, typename S>
int do_step(mplb_case* c) {
    Fused* z = F(c);
    if (z->need_prologue) {
        int rc = run_prologue<L, S>(c);
        if (rc) return rc;
    }
    FArgs a = make_args(c);
    a.fin = z->fa;
    a.fout = z->fb;
    a.gin = z->ga;
    a.gout = z->gb;
    a.first_call = c->current == 0 ? 1 : 0;
    // fixedRandomT (fixedRandomT.c:79): each list draws while its own localTStep < psteps
    a.draw = 0;
    for (size_t s = 0; s < z->frdt_lists.size(); s++) {
        FieldCfg& fc = c->fc[z->frdt_lists[s].first];
        const int b = z->frdt_lists[s].second;
        if (fc.localT[b] < fc.psteps[b]) a.draw |= 1 << s;
        fc.localT[b]++;
    }
    z->last_draw = a.draw;
    z->rho0_built = a.rho0;
    if (c->timing) launch_step_timed<L, S>(c, a, z->thermal, false);
    else launch_step<L, S>(c, a, z->thermal, false);
    std::swap(z->fa, z->fb);
    std::swap(z->ra, z->rb);
    std::swap(z->ga, z->gb);
    if (z->thermal) z->cur = 1 - z->cur;
    z->generic_current = false;
    return 0;
}

// After do_step the buffers hold: fa = post-collision of step n+1, fb = post-collision
// of step n (input of the last K2).  Materialising level n pulls from fb.
template <class L, typename S>
int materialize(mplb_case* c) {
    Fused* z = F(c);
    if (z->generic_current) return 0;
    FArgs a = make_args(c);
    const dim3 grid((unsigned)a.nblocks), block(256);
    // the last step's inputs: post-collision of level n in fb/gb, old T/U = previous
    a.fin = z->fb;
    a.rin = z->rb;
    a.gin = z->gb;
    a.draw = z->last_draw;
    if (z->thermal) {
        a.Told = z->T[1 - z->cur];
        a.Uold = z->U[1 - z->cur];
    }
    k_fused_materialize<L, S, false><<<grid, block, 0, c->stream>>>(a, (float*)c->f[0]);
    if (z->thermal) k_fused_materialize<L, S, true><<<grid, block, 0, c->stream>>>(a, (float*)c->f[1]);
    // macroscopic outputs: re-run the collide pass (deterministic, rewrites fa)
    a.fout = z->fa;
    a.rout = z->ra;
    a.gout = z->ga;
    a.first_call = 0;
    if (z->thermal) {
        a.Tnew = z->T[z->cur];
        a.Unew = z->U[z->cur];
    }
    a.out_rho = (float*)c->rho;
    a.out_Fi = (float*)c->Fi;
    a.out_Ft = (float*)c->Ft;
    a.out_U = (float*)c->U;
    if (z->thermal && z->tmarch) {
        // the step's own kernel with outputs on the last step's inputs: it rewrites fa, ga,
        // T and U with the values they hold
        launch_tmarch<L, S, true>(c, a);
    } else if (z->thermal) {
        k_fused_collide<L, S, true><<<grid, block, 0, c->stream>>>(a);
    } else if (z->march) {
        // the step's own kernel with outputs: it rewrites fa with the same values
        launch_march<L, S, true>(c, a);
    } else {
        k_fused_collide<L, S, false><<<grid, block, 0, c->stream>>>(a);
    }
    if (z->thermal) {
        const int64_t N = c->mesh.N;
        FCHK(c, hipMemcpyAsync(c->temp, z->T[z->cur], sizeof(float) * N, hipMemcpyDeviceToDevice, c->stream));
    }
    FCHK(c, hipStreamSynchronize(c->stream));
    z->generic_current = true;
    return 0;
}

template <typename Fn>
int dispatch(mplb_case* c, Fn fn) {
    Fused* z = F(c);
    if (z->lb == 0) {
        if (z->noderef) return fn(D2Q9{}, h16n{});
        if (z->half) return fn(D2Q9{}, _Float16{});
        return fn(D2Q9{}, float{});
    }
    if (z->noderef) return fn(D3Q15{}, h16n{});
    if (z->half) return fn(D3Q15{}, _Float16{});
    return fn(D3Q15{}, float{});
}

}  // namespace

bool fused_supported(const mplb_case* c, std::string* why) {
    auto no = [&](const char* m) {
        if (why) *why = m;
        return false;
    };
    if (c->opts.precision == MPLB_FP64) return no("the fused kernels compute in FP32");
    // node ids are 32-bit in the fused kernels (workgroup * 256 + lane)
    if (c->mesh.N > (int64_t(1) << 31) - 256) return no("more than 2^31 - 256 nodes");
    const Mesh& m = c->mesh;
    if (!m.structured) return no("lattice nodes not in structured box order");
    if (m.nz > 1 && m.d == 2) return no("2D box with nz > 1");
    const int fm = c->fc[0].mp.model;
    if (!(fm == M_LIMRT && m.lbmodel == 0) && !(fm == M_XUMRT && m.lbmodel == 1)) return no("f model/lattice");
    if (c->update[0] == 0) return no("frozen f field");
    const bool thermal = c->fc[1].present && c->update[1];
    if (thermal) {
        const int gm = c->fc[1].mp.model;
        if (gm != M_MYMRT && !(gm == M_LITEMPMRT && m.lbmodel == 0)) return no("g model");
    }
    // every node on at most one boundary list per field, only node-local or
    // interior-neighbour conditions, and no conditional on a node another condition writes
    std::vector<uint8_t> seen(m.N, 0);
    for (size_t b = 0; b < m.bd.size(); b++) {
        for (int f = 0; f < 2; f++) {
            if (f == 1 && !thermal) continue;
            const int t = c->fc[f].bctype[b];
            if (t == BD_PERIODIC) continue;
            if (f == 0 && t != BD_BBACK) return no("f boundary type other than bounceBack/periodic");
            if (f == 1 && t != BD_BBACK && t != BD_FT && t != BD_FRDT) return no("g boundary type");
            for (int32_t id : m.bd[b]) {
                if (seen[id] & (1 << f)) return no("node on two boundary lists");
                seen[id] |= (uint8_t)(1 << f);
            }
        }
    }
    if (thermal) {
        // g list of every node; fixedRandomT lists get a 4-bit slot in the kernels
        std::vector<int> glist(m.N, -1);
        int nfrdt = 0;
        for (size_t b = 0; b < m.bd.size(); b++) {
            const int t = c->fc[1].bctype[b];
            if (t == BD_FRDT) nfrdt++;
            if (t == BD_PERIODIC) continue;
            for (int32_t id : m.bd[b]) glist[id] = (int)b;
        }
        if (nfrdt > 16) return no("more than 16 fixedRandomT lists");
        const int naxis = m.lbmodel == 0 ? 4 : 6;
        for (size_t b = 0; b < m.bd.size(); b++) {
            const int t = c->fc[1].bctype[b];
            if (t != BD_FT && t != BD_FRDT) continue;
            for (int32_t id : m.bd[b])
                for (int k = 0; k < m.Q; k++) {
                    if (m.nb[(int64_t)id * m.Q + k] != -1) continue;
                    const int32_t nbid = m.nb[(int64_t)id * m.Q + reverse_of(c, k)];
                    if (nbid == -1 || !(seen[nbid] & 2)) continue;
                    if (m.nb[(int64_t)nbid * m.Q + k] == -1)
                        return no("fixedT reads a population another condition writes");
                    // bounceBack.c:34-44 averages present axis pairs: a bounceBack list applied
                    // before this one rewrites the neighbour's link k that fixedT reads
                    const int nl = glist[nbid];
                    if (nl >= 0 && nl < (int)b && c->fc[1].bctype[nl] == BD_BBACK && k >= 1 && k <= naxis &&
                        m.nb[(int64_t)nbid * m.Q + reverse_of(c, k)] != -1)
                        return no("fixedT reads a population an earlier bounceBack list averages");
                }
        }
    }
    return true;
}

int fused_create(mplb_case* c) {
    Fused* z = new Fused();
    c->fz = z;
    const Mesh& m = c->mesh;
    const int64_t N = m.N;
    const int Q = m.Q;
    z->lb = m.lbmodel;
    z->thermal = c->fc[1].present && c->update[1];
    z->half = c->opts.precision == MPLB_FP16S || c->opts.precision == MPLB_FP16N;
    z->noderef = c->opts.precision == MPLB_FP16N;
    const size_t sb = z->half ? 2 : 4;
    z->P = pitch_for(N);
    FCHK(c, hipMalloc(&z->fa, sb * Q * z->P));
    FCHK(c, hipMalloc(&z->fb, sb * Q * z->P));
    if (z->noderef) {
        FCHK(c, hipMalloc(&z->ra, sizeof(_Float16) * N));
        FCHK(c, hipMalloc(&z->rb, sizeof(_Float16) * N));
    }
    if (z->thermal) {
        FCHK(c, hipMalloc(&z->ga, sb * Q * z->P));
        FCHK(c, hipMalloc(&z->gb, sb * Q * z->P));
        FCHK(c, hipMalloc(&z->sumg, sizeof(float) * N));
        for (int i = 0; i < 2; i++) {
            FCHK(c, hipMalloc(&z->T[i], sizeof(float) * N));
            FCHK(c, hipMalloc(&z->U[i], sizeof(float) * 3 * N));
        }
    }
    FCHK(c, hipMalloc(&z->rho, sizeof(float) * N));
    // per-node boundary codes
    std::vector<uint8_t> bcf(N, 0), bcg(N, 0);
    std::vector<float> tw(N, 0.f), pert(N, 0.f);
    std::vector<int> off(N, 0);
    bool any_frdt = false;
    for (size_t b = 0; b < m.bd.size(); b++) {
        for (int f = 0; f < 2; f++) {
            if (f == 1 && !z->thermal) continue;
            const FieldCfg& fc = c->fc[f];
            const int t = fc.bctype[b];
            uint8_t code = t == BD_BBACK ? FB_BBACK : t == BD_FT ? FB_FT : t == BD_FRDT ? FB_FRDT : FB_NONE;
            if (code == FB_NONE) continue;
            int cnt = 0;
            const uint8_t slot = code == FB_FRDT ? (uint8_t)(z->frdt_lists.size() << 4) : 0;
            for (int32_t id : m.bd[b]) {
                (f == 0 ? bcf : bcg)[id] = code | slot;
                if (f == 1) {
                    tw[id] = (float)fc.bcval[b][0];
                    pert[id] = (float)fc.pert[b];
                    if (code == FB_FRDT) {
                        off[id] = cnt;
                        for (int k = 0; k < Q; k++)
                            if (m.nb[(int64_t)id * Q + k] == -1 && m.nb[(int64_t)id * Q + reverse_of(c, k)] != -1) cnt++;
                    }
                }
            }
            if (code == FB_FRDT) {
                any_frdt = true;
                z->frdt_lists.push_back({f, (int)b});
            }
        }
    }
    // Faces whose links are not the duplicate-node periodic wrap, or that carry a
    // boundary condition, are read through the neighbour table.
    {
        const int nx = m.nx, ny = m.ny, nz = m.nz;
        const int64_t nxy = (int64_t)nx * ny;
        auto wrapc = [](int u, int n) { return u < 0 ? u + (n - 1) : (u > n - 1 ? u - (n - 1) : u); };
        int mask = 0;
        for (int64_t id = 0; id < N; id++) {
            const int i = (int)(id % nx), j = (int)((id / nx) % ny), k = (int)(id / nxy);
            int faces = (i == 0 ? 1 : 0) | (i == nx - 1 ? 2 : 0) | (j == 0 ? 4 : 0) | (j == ny - 1 ? 8 : 0);
            if (m.d == 3) faces |= (k == 0 ? 16 : 0) | (k == nz - 1 ? 32 : 0);
            if (!faces || (faces & ~mask) == 0) continue;
            bool differs = bcf[id] != FB_NONE || bcg[id] != FB_NONE;
            for (int q = 0; q < Q && !differs; q++) {
                const int* cq = m.lbmodel == 0 ? D2Q9::t.c[q] : D3Q15::t.c[q];
                const int64_t want = wrapc(i - cq[0], nx) + (int64_t)wrapc(j - cq[1], ny) * nx +
                                     (m.d == 3 ? (int64_t)wrapc(k - cq[2], nz) * nxy : 0);
                if (m.nb[id * Q + q] != want) differs = true;
            }
            if (differs) mask |= faces;
        }
        z->face_table = mask;
    }
    // single-pass kernel: isothermal D3Q15 with every link the periodic wrap and no
    // boundary node (MPLB_FUSED_KERNEL=twopass keeps the two-pass kernels)
    {
        const char* env = std::getenv("MPLB_FUSED_KERNEL");
        const bool twopass = env && std::strcmp(env, "twopass") == 0;
        bool any_bnd = false;
        for (int64_t id = 0; id < N && !any_bnd; id++) any_bnd = m.onbnd[id] != 0;
        const bool periodic3 = !twopass && m.lbmodel == 1 && z->face_table == 0 && !any_bnd && m.nx >= 3 && m.ny >= 3 &&
                               m.nz >= 3;
        z->march = periodic3 && !z->thermal;
        // thermal single pass: xuMRT f with myMRT g (the D3Q15 thermal pair).  The default
        // with FP32 storage (1.10 against 1.38 ms per step for the three passes on the
        // 256^3 thermal bubble); with FP16 storage the three passes are as fast or faster
        // (FP16S 0.961 against 0.975 ms, FP16N 1.004 against 1.162).
        // MPLB_THERMAL_MARCH=1 / =0 forces the single pass / the three passes.
        const char* tenv = std::getenv("MPLB_THERMAL_MARCH");
        const bool tpick = tenv ? std::strcmp(tenv, "1") == 0 : !z->half;
        // (it indexes populations in 32 bits: Q x pitch below 2^32)
        z->tmarch = periodic3 && z->thermal && c->fc[1].mp.model == M_MYMRT && tpick &&
                    15 * pitch_for(N) <= (int64_t)0xFFFFFFFF;
        // LDS-staged pulls for the isothermal single pass, FP32 storage (opt-in, MPLB_MARCH_LDS=1)
        const char* lenv = std::getenv("MPLB_MARCH_LDS");
        z->march_lds = z->march && !z->half && lenv && std::strcmp(lenv, "1") == 0;
        // pulls issued a plane ahead (FP32 and FP16S storage; opt-in, MPLB_MARCH_PF=1)
        const char* penv = std::getenv("MPLB_MARCH_PF");
        z->march_pf = z->march && !z->march_lds && !z->noderef && penv && std::strcmp(penv, "1") == 0;
        if (z->march || z->tmarch) {
            const int tmx = z->tmarch ? TMX : MX, tmy = z->tmarch ? TMY : MY;
            z->mtx = (m.nx + tmx - 1) / tmx;
            z->mtiles = z->mtx * ((m.ny + tmy - 1) / tmy);
            // z chunks (of at least 8 planes) per column: the work items run in rounds of
            // `resident` workgroups (CUs x workgroups per CU at the kernel's occupancy), so
            // pick the count that minimises rounds x (planes per chunk + 2 halo planes)
            // (halo: 2 planes for the isothermal kernel's psi stencil, 4 for the thermal
            // kernel's rho/T/U stencils of the heat source under the psi stencil)
            int bpc = 0, cus = 0;
            const int halo = z->tmarch ? 4 : 2;
            if (z->tmarch && z->noderef)
                FCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_fused_tmarch<D3Q15, h16n, false>, TNT, 0));
            else if (z->tmarch && z->half)
                FCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_fused_tmarch<D3Q15, _Float16, false>, TNT, 0));
            else if (z->tmarch)
                FCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_fused_tmarch<D3Q15, float, false>, TNT, 0));
            else if (z->march_lds)
                FCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_fused_march_lds<D3Q15, false>, 256, 0));
            else if (z->march_pf && z->half)
                FCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_fused_march_pf<D3Q15, _Float16, false>, 256, 0));
            else if (z->march_pf)
                FCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_fused_march_pf<D3Q15, float, false>, 256, 0));
            else if (z->noderef)
                FCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_fused_march<D3Q15, h16n, false>, 256, 0));
            else if (z->half)
                FCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_fused_march<D3Q15, _Float16, false>, 256, 0));
            else
                FCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_fused_march<D3Q15, float, false>, 256, 0));
            FCHK(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->opts.device));
            const long long resident = std::max(1LL, (long long)bpc * cus);
            int nchunks = 1;
            long long best = -1;
            for (int n = 1; n <= std::max(1, m.nz / 8); n++) {
                const long long rounds = ((long long)z->mtiles * n + resident - 1) / resident;
                const long long cost = rounds * ((m.nz + n - 1) / n + halo);
                if (best < 0 || cost < best) {
                    best = cost;
                    nchunks = n;
                }
            }
            // MPLB_MARCH_CHUNKS=<n> overrides the count (1 <= n <= Nz; for timing sweeps)
            if (const char* ce = std::getenv("MPLB_MARCH_CHUNKS")) {
                const int n = std::atoi(ce);
                if (n >= 1 && n <= m.nz) nchunks = n;
            }
            z->mzc = (m.nz + nchunks - 1) / nchunks;
            nchunks = (m.nz + z->mzc - 1) / z->mzc;
            z->mitems = z->mtiles * nchunks;
            z->mper = (z->mitems + 7) / 8;
        }
    }
    FCHK(c, hipMalloc(&z->bcf, N));
    FCHK(c, hipMemcpy(z->bcf, bcf.data(), N, hipMemcpyHostToDevice));
    FCHK(c, hipMalloc(&z->bcg, N));
    FCHK(c, hipMemcpy(z->bcg, bcg.data(), N, hipMemcpyHostToDevice));
    FCHK(c, hipMalloc(&z->bcTw, sizeof(float) * N));
    FCHK(c, hipMemcpy(z->bcTw, tw.data(), sizeof(float) * N, hipMemcpyHostToDevice));
    FCHK(c, hipMalloc(&z->bcPert, sizeof(float) * N));
    FCHK(c, hipMemcpy(z->bcPert, pert.data(), sizeof(float) * N, hipMemcpyHostToDevice));
    FCHK(c, hipMalloc(&z->rnd_off, sizeof(int) * N));
    FCHK(c, hipMemcpy(z->rnd_off, off.data(), sizeof(int) * N, hipMemcpyHostToDevice));
    (void)any_frdt;
    return 0;
}

void fused_destroy(mplb_case* c) {
    Fused* z = F(c);
    if (!z) return;
    void* ptrs[] = {z->fa, z->fb, z->ra, z->rb, z->ga, z->gb, z->rho, z->sumg, z->T[0], z->T[1], z->U[0], z->U[1],
                    z->bcf, z->bcg, z->bcTw, z->bcPert, z->rnd_off};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);     // teardown: nothing left to report to
    delete z;
    c->fz = nullptr;
}

static int set_shifts(mplb_case* c) {
    Fused* z = F(c);
    if (!z->half) return 0;
    // deviation storage: f_k - w_k rho_ref (FP16S; FP16N stores f against per-node
    // references), g_k - geq_k(T_ref, 0); reference values are the lattice means of the
    // initial fields
    const int64_t N = c->mesh.N;
    std::vector<float> h(N);
    FCHK(c, hipMemcpy(h.data(), c->rho, sizeof(float) * N, hipMemcpyDeviceToHost));
    double rs = 0;
    for (float v : h) rs += v;
    const double rho_ref = rs / (double)N;
    double ts = 0;
    FCHK(c, hipMemcpy(h.data(), c->temp, sizeof(float) * N, hipMemcpyDeviceToHost));
    for (float v : h) ts += v;
    const double T_ref = ts / (double)N;
    const int Q = c->mesh.Q;
    double geq[15] = {0};
    const double U0[3] = {0, 0, 0};
    if (c->fc[1].present) {
        const ModelParams& gp = c->fc[1].mp;
        if (c->mesh.lbmodel == 0) {
            if (gp.model == M_MYMRT) my_geq<D2Q9, double>(T_ref, U0, gp.alpha1, gp.alpha2, geq);
            else litemp_geq<D2Q9, double>(T_ref, U0, geq);
        } else {
            my_geq<D3Q15, double>(T_ref, U0, gp.alpha1, gp.alpha2, geq);
        }
    }
    for (int q = 0; q < Q; q++) {
        const double w = c->mesh.lbmodel == 0 ? D2Q9::t.omega[q] : D3Q15::t.omega[q];
        z->fshift[q] = (float)(w * rho_ref);
        z->gshift[q] = (float)geq[q];
    }
    return 0;
}

int fused_after_init(mplb_case* c) {
    const int rc = set_shifts(c);
    if (rc) return rc;
    F(c)->need_prologue = true;
    F(c)->generic_current = true;
    return dispatch(c, [&](auto L, auto S) { return run_prologue<decltype(L), decltype(S)>(c); });
}

int fused_step(mplb_case* c) {
    return dispatch(c, [&](auto L, auto S) { return do_step<decltype(L), decltype(S)>(c); });
}

int fused_sync_to_generic(mplb_case* c) {
    if (!c->fz) return 0;
    return dispatch(c, [&](auto L, auto S) { return materialize<decltype(L), decltype(S)>(c); });
}

int fused_after_write(mplb_case* c) {
    // The write step may have re-averaged rho_0 (multiPhasePP.c:249-315).  rho_0 enters
    // only the gravity term.  liMRT's next collision uses Ft from the step's interForce,
    // i.e. the old rho_0, which is what the post-collision buffers already hold; xuMRT's
    // collision forms Fi + (rho - rho_0) g with the new value (xuMRTForce.c:18-69), so
    // only then is the next step's collision recomputed from the current level.
    Fused* z = F(c);
    const bool gravity = c->eos.g[0] != 0 || c->eos.g[1] != 0 || c->eos.g[2] != 0;
    if (gravity && c->fc[0].mp.model == M_XUMRT && (float)c->rho0 != z->rho0_built) z->need_prologue = true;
    return 0;
}

int fused_set_field(mplb_case* c, const char* name, const double* host, size_t n) {
    int rc = fused_sync_to_generic(c);
    if (rc) return rc;
    Fused* z = F(c);
    z->need_prologue = true;
    if (name && host && std::strcmp(name, "T") == 0 && n == (size_t)c->mesh.N) {
        // an isothermal step reads T only through psi: a uniform field is passed as one
        // value (the FP32 value the generic array holds) instead of being read per node
        const float t0 = (float)host[0];
        bool uni = true;
        for (size_t i = 1; i < n && uni; i++) uni = (float)host[i] == t0;
        z->t_uniform = uni;
        z->t_const = t0;
    }
    return 1;
}

int fused_get_field(mplb_case* c, const char*, double*, size_t) {
    int rc = fused_sync_to_generic(c);
    if (rc) return rc;
    return 1;
}

}  // namespace mplb
