// Batched ensemble of small D2Q9 liMRT cases: fully periodic boxes (periodicXY),
// uniform temperature per case, heat transfer off, no body force.  Every case keeps the
// reference's own time loop (multiPhasePP.c:157-241 with --nht); the batch only changes
// how the work reaches the GPU:
//   * the cases' nodes are concatenated into one structure-of-arrays population buffer
//     ([9][P], each case padded to a multiple of 64 nodes);
//   * a work table lists (case, first node) per workgroup, so one launch covers cases of
//     different sizes; its order can give each XCD a contiguous range of the table, so
//     the rows a workgroup pulls from are mostly in that XCD's L2;
//   * one step is one launch (k_batch_march: workgroups hold whole rows of up to four
//     cases of one size and march up the rows with psi in an LDS ring, so the
//     populations are read once per step) or two launches over the node table (psi, then
//     force + velocity + collision with a pull stream); both give identical results, and
//     runs of steps are captured in a HIP graph and replayed;
//   * every `check_every` steps one pass measures each case's largest density change; the
//     cases that meet the steady criterion leave the work table (their populations stay
//     in place) and the graph is captured again for the smaller table;
//   * an autotuner times the launch shapes (rows per march item, or nodes per workgroup
//     of the two-pass kernels) and the XCD ordering on the live table.
// Storage: FP64 (same arithmetic as the per-routine FP64 kernels: bit-identical to the
// reference restatement), FP32, or FP16 with FP32 arithmetic: deviations f - w_k rho_ref
// (FP16S) or relative deviations (f - w_k R(x)) / (w_k R(x)), R(x) the node's own density
// when it was written, kept per node in FP16 (FP16N).
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mplb.h"
#include "case.h"
#include "physics.h"

using namespace mplb;

namespace {

struct BCase {
    int nx, ny;
    int64_t off;          // first node of the case in the concatenated arrays
    int64_t n;            // nodes
    double T;             // uniform temperature
    float rho_ref;        // FP16S deviation reference (initial mean density)
    int state;            // 0 active, 1 converged, 2 diverged (NaN)
};

struct BArgs {
    const BCase* cases;
    const int2* work;     // (case, first local node) per workgroup; case -1 = padding
    int64_t P;            // plane pitch (total padded nodes)
    EosParams eos;
    double lam[9];
    double sigma;
    const void* fin;      // populations (storage type)
    void* fout;
    void* psi;            // compute type
    void* rho;            // compute type: initial density, then density snapshots
    void* p;              // compute type: pressure output
    unsigned* delta;      // per case: max |rho - snapshot| as float bits
    // FP16 storage: per-node reference densities of the input and output levels
    const _Float16* rin;
    _Float16* rout;
};

// Work item of the single-pass kernel: rows [y0, y1) of up to four cases of the same
// nx x ny (workgroup lane t serves case slot t / nx, column t % nx); ncase 0 = padding.
struct MItem {
    int c[4];
    int ncase;
    int nx;
    int y0, y1;
};

constexpr int kMarchLanes = 256;

enum Mode : int { FROM_PULL = 0, FROM_INIT = 1 };

template <typename S, typename T>
__device__ inline T ld(const void* p, int64_t i, T shift) {
    if constexpr (sizeof(S) == 2) return T((float)((const _Float16*)p)[i]) + shift;
    else return T(((const S*)p)[i]);
}
template <typename S, typename T>
__device__ inline void st(void* p, int64_t i, T v, T shift) {
    if constexpr (sizeof(S) == 2) __builtin_nontemporal_store((_Float16)(float)(v - shift), (_Float16*)p + i);
    else __builtin_nontemporal_store((S)v, (S*)p + i);
}

struct Pos {
    int c;
    int64_t gid;          // node in the concatenated arrays
    int64_t off;          // first node of the case
    bool valid;
    unsigned X[3], Y[3];  // wrapped i-1, i, i+1 and (j-1, j, j+1) * nx
};

// Duplicate-node periodic wrap (period n - 1): the neighbour of i = 0 in -x is n - 2 and
// the neighbour of n - 1 in +x is 1, as the reference's periodicXY tables give
// (latticeBox2D/periodicXY.c); the tables' corner links point at the other stored copy
// of the same site, which holds the same value.
__device__ inline Pos pos(const BArgs& a, BCase& bc) {
    Pos p;
    const int2 w = a.work[blockIdx.x];
    p.c = w.x;
    p.valid = false;
    if (w.x < 0) return p;
    bc = a.cases[w.x];
    const unsigned lid = (unsigned)w.y + threadIdx.x;
    p.valid = lid < (unsigned)bc.n;
    const unsigned nx = (unsigned)bc.nx;
    const unsigned j = lid / nx;
    const unsigned i = lid - j * nx;
    p.off = bc.off;
    p.gid = bc.off + lid;
    p.X[0] = i == 0 ? nx - 2 : i - 1;
    p.X[1] = i;
    p.X[2] = i == nx - 1 ? 1 : i + 1;
    p.Y[0] = (j == 0 ? (unsigned)bc.ny - 2 : j - 1) * nx;
    p.Y[1] = j * nx;
    p.Y[2] = (j == (unsigned)bc.ny - 1 ? 1 : j + 1) * nx;
    return p;
}

// Two FP16 storage modes (FP32 arithmetic in both):
// * MPLB_FP16S (storage type _Float16): f_q - w_q rho_ref, rho_ref the case's mean initial
//   density;
// * MPLB_FP16N (storage type h16n): f_q at node x as its relative deviation
//   (f_q - w_q R(x)) / (w_q R(x)), R(x) being x's density when the population was written
//   (rounded to FP16, one array per level beside the populations): each node's deviations
//   keep the 11-bit mantissa in both phases of a two-phase case.
struct h16n {
    _Float16 v;
};
template <typename S>
constexpr bool kNodeRef = std::is_same<S, h16n>::value;
static_assert(sizeof(h16n) == 2, "FP16 storage element");

constexpr float kRefMin = 6.103515625e-05f;     // smallest normal FP16 (R never rounds to 0)

// FP16 storage: w_q R of the node a population was stored at (FP16N), or the case's shift
// of population q (FP16S)
template <typename S, typename T>
__device__ inline T wref_of(const BArgs& a, const BCase& bc, int64_t node, int q) {
    if constexpr (kNodeRef<S>) return T(D2Q9::t.omega[q]) * T((float)a.rin[node]);
    else if constexpr (sizeof(S) == 2) return T(D2Q9::t.omega[q]) * T(bc.rho_ref);
    else return T(0);
}

template <typename S, typename T>
__device__ inline T ldf(const void* p, int64_t i, T wref) {
    if constexpr (kNodeRef<S>) return T(fmaf((float)((const _Float16*)p)[i], (float)wref, (float)wref));
    else return ld<S, T>(p, i, wref);
}

// FP16N: writes node gid's reference for the output level (its density, at least kRefMin;
// NaN stays NaN) and returns it with its reciprocal; 0 in the other modes
template <typename S, typename T>
__device__ inline T put_ref(const BArgs& a, int64_t gid, T rho, T& rinv) {
    if constexpr (kNodeRef<S>) {
        const float rf = (float)rho;
        const float r = (rf >= kRefMin || rf != rf) ? rf : kRefMin;
        const _Float16 h = (_Float16)r;
        a.rout[gid] = h;
        rinv = T(1.f / (float)h);
        return T((float)h);
    } else {
        rinv = T(0);
        return T(0);
    }
}

template <typename S, typename T>
__device__ inline void stf(const BCase& bc, void* p, int64_t i, int q, T v, T R, T rinv) {
    if constexpr (kNodeRef<S>) {
        const float wref = (float)D2Q9::t.omega[q] * (float)R;
        const float dev = ((float)v - wref) * ((float)(1.0 / D2Q9::t.omega[q]) * (float)rinv);
        __builtin_nontemporal_store((_Float16)dev, (_Float16*)p + i);
    } else {
        st<S, T>(p, i, v, sizeof(S) == 2 ? T(D2Q9::t.omega[q]) * T(bc.rho_ref) : T(0));
    }
}

// populations of the current time level: pull-stream of the post-collision buffer
// (lbstream.c:16-62: f_k(x) <- f_k(x - c_k))
template <typename S, typename T>
__device__ inline void pull(const BArgs& a, const BCase& bc, const Pos& p, const void* src, T* v) {
#pragma unroll
    for (int q = 0; q < 9; q++) {
        const int64_t node = p.off + p.X[1 - D2Q9::t.c[q][0]] + p.Y[1 - D2Q9::t.c[q][1]];
        v[q] = ldf<S, T>(src, q * a.P + node, wref_of<S, T>(a, bc, node, q));
    }
}

// density of pulled populations: the reference's index order in FP64 (parity), the
// inversion-symmetric pair form in FP32 (no drift from biased rounding)
// FP32 quotients and the pseudopotential from the hardware reciprocal and square root
// (about 1 ulp) instead of the correctly rounded instruction sequences; FP64 (the
// parity mode) keeps IEEE division and sqrt
template <typename T>
__device__ inline T bdiv(T x, T y) {
    if constexpr (sizeof(T) == 4) return x * __builtin_amdgcn_rcpf(y);
    else return x / y;
}

template <typename T>
__device__ inline T bpotential(const EosParams& e, T rho, T temp) {
    if constexpr (sizeof(T) == 4) {
        if (e.type == EOS_VDW) {
            const float M = (float)e.M;
            const float p = rho * (float)e.R * temp * __builtin_amdgcn_rcpf(M - rho * (float)e.b) -
                            (float)e.a * __builtin_amdgcn_rcpf(M * M) * rho * rho;
            return __builtin_amdgcn_sqrtf(2.f * (p - rho * (float)kCs2) * __builtin_amdgcn_rcpf((float)e.G));
        }
    }
    return potential<T>(e, rho, temp);
}

template <typename T>
__device__ inline T density(const T* v) {
    if constexpr (sizeof(T) == 8) {
        T rho = 0;
#pragma unroll
        for (int q = 0; q < 9; q++) rho += v[q];
        return rho;
    } else {
        return rho_sym<D2Q9, T>(v);
    }
}

// Populations and density for the collision of the next step.  FROM_INIT is the first
// step: populations f = f_eq(rho, 0) of the initial density (equilibrium.c:17-85 with
// U = 0), and the collision uses that density and U = 0 as read from the case
// (multiPhasePP.c:122-147); later steps use the macroscopic fields of the pulled
// populations (macroDensity.c:19-60, macroVelocity.c:17-50).
template <typename S, typename T, int MODE>
__device__ inline T load_state(const BArgs& a, const BCase& bc, const Pos& p, T* v) {
    if (MODE == FROM_PULL) {
        pull<S, T>(a, bc, p, a.fin, v);
        return density<T>(v);
    }
    const T rho = ((const T*)a.rho)[p.gid];
    const T U0[3] = {T(0), T(0), T(0)};
#pragma unroll
    for (int q = 0; q < 9; q++) v[q] = ns_feq<D2Q9, T>(q, rho, U0);
    return rho;
}

// psi = potential(rho, T) of the current time level (potential.c:7-13)
template <typename S, typename T, int MODE>
__global__ void k_batch_psi(BArgs a) {
    BCase bc;
    const Pos p = pos(a, bc);
    if (!p.valid) return;
    T v[9];
    const T rho = load_state<S, T, MODE>(a, bc, p, v);
    ((T*)a.psi)[p.gid] = bpotential<T>(a.eos, rho, T(bc.T));
}

// Interaction force (interForce.c:16-39, interactionForce.c:47-79; every node of a fully
// periodic box is interior), velocity (liMRTVelocity.c:11-50) and the liMRT collision
// (liMRTCollision.c:33-106 with liMRTForce.c:13-42), written to the other buffer.  The
// expressions keep the per-routine FP64 kernels' association order.
// One node: v holds its pulled populations (density rho), psic its psi and psin[q] the
// psi of x + c_q; the post-collision populations go to node gid of the other buffer.
template <typename S, typename T, int MODE>
__device__ inline void collide_site(const BArgs& a, const BCase& bc, T* v, const T rho, const T psic, const T* psin,
                                    int64_t gid) {
    T F[2] = {T(0.0), T(0.0)};
    constexpr Pairs<D2Q9> PR = make_pairs<D2Q9>();
    if constexpr (sizeof(T) == 8) {
#pragma unroll
        for (int q = 1; q < 9; q++) {
            const T al = T(D2Q9::t.w[q]) * psin[q];
#pragma unroll
            for (int d = 0; d < 2; d++) {
                const int c = D2Q9::t.c[q][d];
                if (c == 1) F[d] += al;
                else if (c == -1) F[d] += -al;
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < Pairs<D2Q9>::n; k++) {
            const int q = PR.q[k];
            const T dpsi = T(D2Q9::t.w[q]) * (psin[q] - psin[PR.r[k]]);
#pragma unroll
            for (int d = 0; d < 2; d++) {
                const int c = D2Q9::t.c[q][d];
                if (c == 1) F[d] += dpsi;
                else if (c == -1) F[d] -= dpsi;
            }
        }
    }
    const T beta = -T(a.eos.G) * psic;
    const T Fi[2] = {F[0] * beta, F[1] * beta};
    T Ft[2];
#pragma unroll
    for (int d = 0; d < 2; d++) Ft[d] = Fi[d] + (rho - T(0)) * T(a.eos.g[d]) + T(a.eos.fext[d]);
    T U[2] = {T(0), T(0)};
    if (MODE == FROM_PULL) {
#pragma unroll
        for (int d = 0; d < 2; d++) {
            T lv = 0;
            if constexpr (sizeof(T) == 8) {
#pragma unroll
                for (int q = 0; q < 9; q++) {
                    const int c = D2Q9::t.c[q][d];
                    if (c == 1) lv += v[q];
                    else if (c == -1) lv += -v[q];
                }
            } else {
#pragma unroll
                for (int k = 0; k < Pairs<D2Q9>::n; k++) {
                    const int c = D2Q9::t.c[PR.q[k]][d];
                    if (c == 1) lv += v[PR.q[k]] - v[PR.r[k]];
                    else if (c == -1) lv -= v[PR.q[k]] - v[PR.r[k]];
                }
            }
            U[d] = bdiv<T>(lv + Ft[d] * T(0.5), rho);
        }
    }
    T m[9], meq[9], Sf[9];
    const T umag = U[0] * U[0] + U[1] * U[1];
    meq[0] = rho;
    meq[1] = rho * (T(-2) + T(3) * umag);
    meq[2] = rho * (T(1) - T(3) * umag);
    meq[3] = rho * U[0];
    meq[4] = rho * (-U[0]);
    meq[5] = rho * U[1];
    meq[6] = rho * (-U[1]);
    meq[7] = rho * (U[0] * U[0] - U[1] * U[1]);
    meq[8] = rho * U[0] * U[1];
    if constexpr (sizeof(T) == 8) moments<D2Q9, T>(v, m);
    else moments_sym<D2Q9, T>(v, m);
    const T sigma = T(a.sigma);
    const T L1 = T(a.lam[1]), L2 = T(a.lam[2]);
    const T uF = U[0] * Ft[0] + U[1] * Ft[1];
    const T fi2 = Fi[0] * Fi[0] + Fi[1] * Fi[1];
    Sf[0] = 0;
    Sf[1] = T(6) * uF + bdiv<T>(T(12) * sigma * fi2, psic * psic * (bdiv<T>(T(1), L1) - T(0.5)));
    Sf[2] = T(-6) * uF - bdiv<T>(T(12) * sigma * fi2, psic * psic * (bdiv<T>(T(1), L2) - T(0.5)));
    Sf[3] = Ft[0];
    Sf[4] = -Ft[0];
    Sf[5] = Ft[1];
    Sf[6] = -Ft[1];
    Sf[7] = T(2) * (U[0] * Ft[0] - U[1] * Ft[1]);
    Sf[8] = U[0] * Ft[1] + U[1] * Ft[0];
    if constexpr (sizeof(T) == 8) {
        // the oracle's order: post-collision moments, then populations rebuilt from them
#pragma unroll
        for (int q = 0; q < 9; q++) {
            const T l = T(a.lam[q]);
            m[q] = m[q] - l * (m[q] - meq[q]) + (T(1) - T(0.5) * l) * Sf[q];
        }
        populations<D2Q9, T>(m, v);
    } else {
        // the moment change added to the pulled populations (no per-step rebuild)
#pragma unroll
        for (int q = 0; q < 9; q++) {
            const T l = T(a.lam[q]);
            m[q] = (T(1) - T(0.5) * l) * Sf[q] - l * (m[q] - meq[q]);
        }
        populations_add_sym<D2Q9, T>(m, v);
    }
    T rinv;
    const T R = put_ref<S, T>(a, gid, rho, rinv);
#pragma unroll
    for (int q = 0; q < 9; q++) stf<S, T>(bc, a.fout, q * a.P + gid, q, v[q], R, rinv);
}

template <typename S, typename T, int MODE>
__global__ void k_batch_collide(BArgs a) {
    BCase bc;
    const Pos p = pos(a, bc);
    if (!p.valid) return;
    T v[9];
    const T rho = load_state<S, T, MODE>(a, bc, p, v);
    const T* psi = (const T*)a.psi;
    T psin[9];
#pragma unroll
    for (int q = 1; q < 9; q++) psin[q] = psi[p.off + p.X[1 + D2Q9::t.c[q][0]] + p.Y[1 + D2Q9::t.c[q][1]]];
    collide_site<S, T, MODE>(a, bc, v, rho, psi[p.gid], psin, p.gid);
}

__device__ inline int wrap_dup(int u, int n) { return u < 0 ? u + (n - 1) : (u > n - 1 ? u - (n - 1) : u); }

// The same step in one pass.  Work item it covers rows [y0, y1) of its cases; the
// workgroup pulls row by row (rows y0 - 1 .. y1, wrapped), turns each row into psi in an
// LDS ring of four rows and collides row y from the populations it pulled one row
// earlier (kept in registers) and the psi stencil of rows y - 1 .. y + 1.  A case's row
// lies whole in the workgroup, so the x stencil wraps inside it (duplicate-node wrap);
// psi never goes to memory.  Per node and step the populations are read once and written
// once.  LDS hazards: row y + 1 goes to slot (y + 2 - y0) & 3, which no lane reads while
// colliding row y - 1 or y, and one barrier per row publishes it.
// wave priority while a row's populations are pulled (as in the 3D single pass: loads
// go out ahead of other waves' collisions; 0.9% faster in FP64, neutral in FP32)
#ifndef MPLB_BATCH_PRIO
#define MPLB_BATCH_PRIO 2
#endif
template <typename S, typename T>
__device__ inline void march_item(const BArgs& a, const MItem& it, T (*ring)[kMarchLanes]) {
    if (it.ncase == 0) return;
    const int t = threadIdx.x;
    const int nx = it.nx;
    const int slot = t / nx, x = t - slot * nx;
    const bool act = slot < it.ncase;
    BCase bc{};
    if (act) bc = a.cases[it.c[slot]];
    const int ny = act ? bc.ny : 3;
    const int sb = slot * nx;
    const unsigned X[3] = {(unsigned)(x == 0 ? nx - 2 : x - 1), (unsigned)x, (unsigned)(x == nx - 1 ? 1 : x + 1)};
    T fc[9], fn[9];
    // populations and psi of raw row yr (y0 - 1 <= yr <= y1)
    auto row = [&](int yr) {
        if (!act) return;
        if (MPLB_BATCH_PRIO) __builtin_amdgcn_s_setprio(MPLB_BATCH_PRIO);
        const int yw = wrap_dup(yr, ny);
        const unsigned Y[3] = {(unsigned)((yw == 0 ? ny - 2 : yw - 1) * nx), (unsigned)(yw * nx),
                               (unsigned)((yw == ny - 1 ? 1 : yw + 1) * nx)};
#pragma unroll
        for (int q = 0; q < 9; q++) {
            const int64_t node = bc.off + X[1 - D2Q9::t.c[q][0]] + Y[1 - D2Q9::t.c[q][1]];
            fn[q] = ldf<S, T>(a.fin, q * a.P + node, wref_of<S, T>(a, bc, node, q));
        }
        ring[(yr - it.y0 + 1) & 3][t] = bpotential<T>(a.eos, density<T>(fn), T(bc.T));
        if (MPLB_BATCH_PRIO) __builtin_amdgcn_s_setprio(0);
    };
    row(it.y0 - 1);
    row(it.y0);
#pragma unroll
    for (int q = 0; q < 9; q++) fc[q] = fn[q];
    for (int y = it.y0; y < it.y1; y++) {
        row(y + 1);
        __syncthreads();
        if (act) {
            T psin[9];
#pragma unroll
            for (int q = 1; q < 9; q++)
                psin[q] = ring[(y + D2Q9::t.c[q][1] - it.y0 + 1) & 3][sb + X[1 + D2Q9::t.c[q][0]]];
            collide_site<S, T, FROM_PULL>(a, bc, fc, density<T>(fc), ring[(y - it.y0 + 1) & 3][t], psin,
                                          bc.off + (int64_t)y * nx + x);
        }
#pragma unroll
        for (int q = 0; q < 9; q++) fc[q] = fn[q];
    }
}

template <typename S, typename T>
__global__ void __launch_bounds__(kMarchLanes) k_batch_march(BArgs a, const MItem* items, int per, int nitems) {
    __shared__ T ring[4][kMarchLanes];
    const int wi = (int)((blockIdx.x & 7u) * (unsigned)per + (blockIdx.x >> 3));
    if (wi >= nitems) return;
    march_item<S, T>(a, items[wi], ring);
}

// The same steps from one persistent launch (cooperative, so every workgroup is resident
// at once): each workgroup walks the work items of its XCD slice, item v of the launch
// order being taken by workgroup v mod G (G, the grid, a multiple of 8, so an item keeps
// the XCD the per-step launch gives it), and the grid meets at a barrier between steps.
// The level buffers alternate per step as in the per-step launches: step s reads
// f0 / r0 when s is even.  Every workgroup runs every step and reaches every barrier.
template <typename S, typename T>
__global__ void __launch_bounds__(kMarchLanes) k_batch_persist(BArgs a, const MItem* items, int per, int nitems,
                                                                int nsteps, void* f0, void* f1, _Float16* r0,
                                                                _Float16* r1) {
    __shared__ T ring[4][kMarchLanes];
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    const unsigned total = 8u * (unsigned)per;
    for (int s = 0; s < nsteps; s++) {
        a.fin = (s & 1) ? f1 : f0;
        a.fout = (s & 1) ? f0 : f1;
        a.rin = (s & 1) ? r1 : r0;
        a.rout = (s & 1) ? r0 : r1;
        for (unsigned v = blockIdx.x; v < total; v += gridDim.x) {
            const int wi = (int)((v & 7u) * (unsigned)per + (v >> 3));
            if (wi < nitems) march_item<S, T>(a, items[wi], ring);
            __syncthreads();      // the next item's rows reuse the LDS ring
        }
        if (s + 1 < nsteps) grid.sync();
    }
}


// Density of the current time level into the snapshot buffer; with `compare`, the
// per-case maximum of |rho - previous snapshot| (NaN propagates: its bits order above
// every finite and infinite non-negative float).
template <typename S, typename T>
__global__ void k_batch_measure(BArgs a, int compare) {
    BCase bc;
    const Pos p = pos(a, bc);
    __shared__ unsigned red[16];
    unsigned d = 0;
    if (p.valid) {
        T v[9];
        pull<S, T>(a, bc, p, a.fin, v);
        const T rho = density<T>(v);
        T* r = (T*)a.rho;
        if (compare) d = __float_as_uint(fabsf((float)(rho - r[p.gid])));
        r[p.gid] = rho;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) d = max(d, (unsigned)__shfl_down((int)d, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = d;
    __syncthreads();
    if (threadIdx.x == 0 && p.c >= 0 && compare) {
        unsigned m = 0;
        for (unsigned i = 0; i < (blockDim.x + 63) / 64; i++) m = max(m, red[i]);
        atomicMax(&a.delta[p.c], m);
    }
}

// Pressure of the current time level (macroPressure.c:20-65 with liMRTPressure.c, the
// all-neighbours-present branch): p = rho cs2 + G psi^2 / 2 + G psi lap(psi) / 12
// + 2 G^2 sigma |grad psi|^2.  psi and rho (snapshot) hold the current level.
template <typename T>
__global__ void k_batch_pressure(BArgs a) {
    BCase bc;
    const Pos p = pos(a, bc);
    if (!p.valid) return;
    const T* psi = (const T*)a.psi;
    const T G = T(a.eos.G);
    const T ps = psi[p.gid];
    T pr = ((const T*)a.rho)[p.gid] * T(kCs2) + T(0.5) * G * ps * ps;
    const T xa = psi[p.off + p.X[2] + p.Y[1]], xb = psi[p.off + p.X[0] + p.Y[1]];
    const T ya = psi[p.off + p.X[1] + p.Y[2]], yb = psi[p.off + p.X[1] + p.Y[0]];
    T lap = xa + xb - T(2) * ps;
    lap = lap + ya + yb - T(2) * ps;
    const T grad[3] = {T(0.5) * (xa - xb), T(0.5) * (ya - yb), T(0)};
    pr = pr + G * ps * lap / T(12);
    T g2 = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) g2 += grad[k] * grad[k];
    pr = pr + T(2.0) * G * G * T(a.sigma) * g2;
    ((T*)a.p)[p.gid] = pr;
}

// density of the current time level and psi, for the field readback
template <typename S, typename T>
__global__ void k_batch_rho_psi(BArgs a) {
    BCase bc;
    const Pos p = pos(a, bc);
    if (!p.valid) return;
    T v[9];
    pull<S, T>(a, bc, p, a.fin, v);
    const T rho = density<T>(v);
    ((T*)a.rho)[p.gid] = rho;
    ((T*)a.psi)[p.gid] = bpotential<T>(a.eos, rho, T(bc.T));
}

__global__ void k_fill_u32(unsigned* p, int n, unsigned v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

struct mplb_batch {
    int prec = MPLB_FP32;
    std::string err;
    int dev = 0;
    hipStream_t stream = nullptr;
    std::vector<BCase> cases;
    std::vector<std::vector<double>> rho_init;
    std::vector<int64_t> case_steps;
    int64_t P = 0;
    EosParams eos{};
    double lam[9] = {0};
    double sigma = 0;
    size_t sbytes = 4, cbytes = 4;
    void* fa = nullptr;                 // post-collision populations of the current level
    void* fb = nullptr;
    _Float16* ra = nullptr;             // FP16 storage: per-node references of fa and fb
    _Float16* rb = nullptr;
    void* psi = nullptr;
    void* rho = nullptr;
    void* p = nullptr;
    BCase* d_cases = nullptr;
    int2* d_work = nullptr;
    size_t work_cap = 0;
    int nwork = 0;                      // launch grid (a multiple of 8 with XCD ordering)
    unsigned* d_delta = nullptr;
    int bsize = 256;
    int xcd_order = 1;
    MItem* d_mwork = nullptr;           // single-pass work items (launch order)
    size_t mwork_cap = 0;
    int nmwork = 0;                     // launch grid of k_batch_march (8 x mper)
    int mper = 0;
    bool march_ok = true;               // every case is at most kMarchLanes wide
    int mrows = 32;                     // rows per march item; 0: the two-pass kernels
    bool persist = false;               // single pass: one persistent launch per run of steps
    bool initialized = false;
    bool fields_valid = false;
    int64_t steps = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_steps = 0;
    bool graph_stale = true;
};

namespace {

int bfail(mplb_batch* b, int code, const std::string& m) {
    if (b) b->err = m;
    return code;
}

#define BCHK(b, call)                                                                                           \
    do {                                                                                                        \
        hipError_t _e = (call);                                                                                 \
        if (_e != hipSuccess) return bfail((b), MPLB_E_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

BArgs bargs(mplb_batch* b) {
    BArgs a;
    std::memset(&a, 0, sizeof(a));
    a.cases = b->d_cases;
    a.work = b->d_work;
    a.P = b->P;
    a.eos = b->eos;
    for (int q = 0; q < 9; q++) a.lam[q] = b->lam[q];
    a.sigma = b->sigma;
    a.psi = b->psi;
    a.rho = b->rho;
    a.p = b->p;
    a.delta = b->d_delta;
    return a;
}

// Work table of the cases in state 0 (of every case with all = true).  With XCD
// ordering, workgroup w (dispatched to XCD w mod 8) takes entry (w mod 8) * per + w / 8
// of the natural order, so each XCD sweeps one contiguous range of nodes.
int rebuild_work(mplb_batch* b, bool all = false) {
    std::vector<int2> nat;
    for (size_t c = 0; c < b->cases.size(); c++) {
        if (!all && b->cases[c].state != 0) continue;
        for (int64_t s = 0; s < b->cases[c].n; s += b->bsize) nat.push_back(make_int2((int)c, (int)s));
    }
    std::vector<int2> w;
    if (b->xcd_order && !nat.empty()) {
        const size_t per = (nat.size() + 7) / 8;
        w.assign(per * 8, make_int2(-1, 0));
        for (size_t g = 0; g < w.size(); g++) {
            const size_t src = (g % 8) * per + g / 8;
            if (src < nat.size()) w[g] = nat[src];
        }
    } else {
        w = nat;
    }
    b->nwork = (int)w.size();
    // single-pass items: cases grouped by size, kMarchLanes / nx (at most 4) per item,
    // mrows rows each; XCD x takes items [x per, (x + 1) per) of the natural order (the
    // kernel's block remap), or block i takes item i without XCD ordering
    std::vector<MItem> mw;
    b->nmwork = 0;
    b->mper = 0;
    if (b->mrows > 0) {
        std::vector<std::pair<int, int>> keys;
        std::vector<std::vector<int>> members;
        for (size_t c = 0; c < b->cases.size(); c++) {
            if (!all && b->cases[c].state != 0) continue;
            const std::pair<int, int> k(b->cases[c].nx, b->cases[c].ny);
            size_t g = 0;
            while (g < keys.size() && keys[g] != k) g++;
            if (g == keys.size()) {
                keys.push_back(k);
                members.emplace_back();
            }
            members[g].push_back((int)c);
        }
        std::vector<MItem> mnat;
        for (size_t g = 0; g < keys.size(); g++) {
            const int nx = keys[g].first, ny = keys[g].second;
            const int k = std::min(4, kMarchLanes / nx);
            for (size_t i = 0; i < members[g].size(); i += k) {
                MItem it{};
                it.ncase = (int)std::min<size_t>(k, members[g].size() - i);
                for (int j = 0; j < it.ncase; j++) it.c[j] = members[g][i + j];
                it.nx = nx;
                for (int y0 = 0; y0 < ny; y0 += b->mrows) {
                    it.y0 = y0;
                    it.y1 = std::min(y0 + b->mrows, ny);
                    mnat.push_back(it);
                }
            }
        }
        if (!mnat.empty()) {
            const size_t per = (mnat.size() + 7) / 8;
            mw.assign(per * 8, MItem{});
            for (size_t i = 0; i < mnat.size(); i++) {
                if (b->xcd_order) mw[i] = mnat[i];
                else mw[(i % 8) * per + i / 8] = mnat[i];
            }
            b->mper = (int)per;
            b->nmwork = (int)mw.size();
        }
    }
    if (mw.size() > b->mwork_cap) {
        if (b->d_mwork) BCHK(b, hipFree(b->d_mwork));
        b->d_mwork = nullptr;
        BCHK(b, hipMalloc(&b->d_mwork, sizeof(MItem) * mw.size()));
        b->mwork_cap = mw.size();
    }
    if (!mw.empty())
        BCHK(b, hipMemcpyAsync(b->d_mwork, mw.data(), sizeof(MItem) * mw.size(), hipMemcpyHostToDevice, b->stream));
    if (w.size() > b->work_cap) {
        if (b->d_work) BCHK(b, hipFree(b->d_work));
        b->d_work = nullptr;
        BCHK(b, hipMalloc(&b->d_work, sizeof(int2) * w.size()));
        b->work_cap = w.size();
    }
    if (!w.empty())
        BCHK(b, hipMemcpyAsync(b->d_work, w.data(), sizeof(int2) * w.size(), hipMemcpyHostToDevice, b->stream));
    BCHK(b, hipMemcpyAsync(b->d_cases, b->cases.data(), sizeof(BCase) * b->cases.size(), hipMemcpyHostToDevice,
                           b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    b->graph_stale = true;
    return 0;
}

template <typename S, typename T>
void launch_steps(mplb_batch* b, int nsteps) {
    BArgs a = bargs(b);
    const dim3 grid((unsigned)b->nwork), block((unsigned)b->bsize);
    for (int s = 0; s < nsteps; s++) {
        a.fin = b->fa;
        a.fout = b->fb;
        a.rin = b->ra;
        a.rout = b->rb;
        if (b->mrows > 0) {
            if (b->nmwork > 0)
                k_batch_march<S, T><<<(unsigned)b->nmwork, kMarchLanes, 0, b->stream>>>(a, b->d_mwork, b->mper,
                                                                                      b->nmwork);
        } else if (b->nwork > 0) {
            k_batch_psi<S, T, FROM_PULL><<<grid, block, 0, b->stream>>>(a);
            k_batch_collide<S, T, FROM_PULL><<<grid, block, 0, b->stream>>>(a);
        }
        std::swap(b->fa, b->fb);
        std::swap(b->ra, b->rb);
    }
}

// nsteps single-pass steps from one persistent cooperative launch; the grid is the number
// of workgroups that fit the device at once (from the occupancy query), rounded down to
// a multiple of 8 and at most the item count, so the launch fails rather than wait on a
// workgroup that cannot start
template <typename S, typename T>
int launch_persist(mplb_batch* b, int nsteps) {
    if (nsteps <= 0 || b->nmwork <= 0) return 0;
    int bpc = 0, cus = 0;
    BCHK(b, hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_batch_persist<S, T>, kMarchLanes, 0));
    BCHK(b, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->dev));
    int g = std::min(bpc * cus, b->nmwork) / 8 * 8;
    if (g < 8) return bfail(b, MPLB_E_HIP, "persistent launch: fewer than 8 resident workgroups");
    BArgs a = bargs(b);
    const MItem* items = b->d_mwork;
    int per = b->mper, nitems = b->nmwork, ns = nsteps;
    void* f0 = b->fa;
    void* f1 = b->fb;
    _Float16* r0 = b->ra;
    _Float16* r1 = b->rb;
    void* args[] = {&a, &items, &per, &nitems, &ns, &f0, &f1, &r0, &r1};
    BCHK(b, hipLaunchCooperativeKernel((const void*)k_batch_persist<S, T>, dim3((unsigned)g), dim3(kMarchLanes), args, 0,
                                       b->stream));
    if (nsteps & 1) {
        std::swap(b->fa, b->fb);
        std::swap(b->ra, b->rb);
    }
    return 0;
}

// nsteps steps through a captured graph; nsteps is even, so the buffers end where they
// started and each replay continues from the previous one.  With the persistent kernel,
// one cooperative launch instead.
template <typename S, typename T>
int run_graph(mplb_batch* b, int nsteps) {
    if (b->persist && b->mrows > 0) return launch_persist<S, T>(b, nsteps);
    if (b->gexec && (b->graph_stale || b->graph_steps != nsteps)) {
        BCHK(b, hipGraphExecDestroy(b->gexec));
        BCHK(b, hipGraphDestroy(b->graph));
        b->gexec = nullptr;
        b->graph = nullptr;
    }
    if (!b->gexec) {
        BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        launch_steps<S, T>(b, nsteps);
        BCHK(b, hipStreamEndCapture(b->stream, &b->graph));
        BCHK(b, hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
        b->graph_steps = nsteps;
        b->graph_stale = false;
    }
    BCHK(b, hipGraphLaunch(b->gexec, b->stream));
    return 0;
}

template <typename S, typename T>
int init_t(mplb_batch* b) {
    // initial density into the snapshot buffer, then equilibrium + first collision
    std::vector<T> h(b->P, T(0));
    for (size_t c = 0; c < b->cases.size(); c++)
        for (int64_t i = 0; i < b->cases[c].n; i++) h[b->cases[c].off + i] = (T)b->rho_init[c][i];
    BCHK(b, hipMemcpyAsync(b->rho, h.data(), sizeof(T) * b->P, hipMemcpyHostToDevice, b->stream));
    for (auto& c : b->cases) c.state = 0;
    int rc = rebuild_work(b);
    if (rc) return rc;
    BArgs a = bargs(b);
    const dim3 grid((unsigned)b->nwork), block((unsigned)b->bsize);
    a.fout = b->fa;
    a.rout = b->ra;
    k_batch_psi<S, T, FROM_INIT><<<grid, block, 0, b->stream>>>(a);
    k_batch_collide<S, T, FROM_INIT><<<grid, block, 0, b->stream>>>(a);
    BCHK(b, hipGetLastError());
    BCHK(b, hipStreamSynchronize(b->stream));
    b->steps = 1;
    std::fill(b->case_steps.begin(), b->case_steps.end(), 1);
    b->initialized = true;
    b->fields_valid = false;
    return 0;
}

template <typename S, typename T>
int fields_t(mplb_batch* b) {
    // rho, psi and p of every case at its current level
    int rc = rebuild_work(b, true);
    if (rc) return rc;
    BArgs a = bargs(b);
    a.fin = b->fa;
    a.rin = b->ra;
    const dim3 grid((unsigned)b->nwork), block((unsigned)b->bsize);
    k_batch_rho_psi<S, T><<<grid, block, 0, b->stream>>>(a);
    k_batch_pressure<T><<<grid, block, 0, b->stream>>>(a);
    BCHK(b, hipGetLastError());
    BCHK(b, hipStreamSynchronize(b->stream));
    b->fields_valid = true;
    return rebuild_work(b);
}

template <typename S, typename T>
int steps_plain(mplb_batch* b, int n) {
    if (b->persist && b->mrows > 0) return launch_persist<S, T>(b, n);
    launch_steps<S, T>(b, n);
    BCHK(b, hipGetLastError());
    return 0;
}

template <typename S, typename T>
int measure_t(mplb_batch* b, bool compare) {
    const int nc = (int)b->cases.size();
    if (compare) k_fill_u32<<<(nc + 255) / 256, 256, 0, b->stream>>>(b->d_delta, nc, 0u);
    BArgs a = bargs(b);
    a.fin = b->fa;
    a.rin = b->ra;
    k_batch_measure<S, T><<<(unsigned)b->nwork, (unsigned)b->bsize, 0, b->stream>>>(a, compare ? 1 : 0);
    BCHK(b, hipGetLastError());
    return 0;
}

template <typename Fn>
int bdispatch(mplb_batch* b, Fn fn) {
    if (b->prec == MPLB_FP64) return fn(double{}, double{});
    if (b->prec == MPLB_FP16S) return fn(_Float16{}, float{});
    if (b->prec == MPLB_FP16N) return fn(h16n{}, float{});
    return fn(float{}, float{});
}

#define BATCH_CALL(b, fn, ...) \
    bdispatch((b), [&](auto s_, auto t_) { return fn<decltype(s_), decltype(t_)>(__VA_ARGS__); })

}  // namespace

extern "C" {

int mplb_batch_create(mplb_case* const* cases, int n, int precision, int device, mplb_batch** out) {
    if (!out) return MPLB_E_ARG;
    *out = nullptr;
    if (!cases || n <= 0) return MPLB_E_ARG;
    if (precision < MPLB_FP64 || precision > MPLB_FP16N) return MPLB_E_ARG;
    mplb_batch* b = new mplb_batch();
    *out = b;
    b->prec = precision;
    b->dev = device;
    b->sbytes = precision == MPLB_FP64 ? 8 : ((precision == MPLB_FP16S || precision == MPLB_FP16N) ? 2 : 4);
    b->cbytes = precision == MPLB_FP64 ? 8 : 4;
    int64_t off = 0;
    for (int i = 0; i < n; i++) {
        const mplb_case* c = cases[i];
        if (!c) return bfail(b, MPLB_E_ARG, "case " + std::to_string(i) + " is null");
        const Mesh& m = c->mesh;
        const std::string who = "case " + std::to_string(i) + ": ";
        if (m.lbmodel != 0 || m.btype != BOX_PERIODIC_XY || m.nx < 3 || m.ny < 3)
            return bfail(b, MPLB_E_MODEL, who + "batched cases are D2Q9 periodicXY boxes");
        if (m.N >= (int64_t(1) << 31)) return bfail(b, MPLB_E_ARG, who + "too many nodes");
        const ModelParams& mp = c->fc[0].mp;
        if (mp.model != M_LIMRT || mp.surf)
            return bfail(b, MPLB_E_MODEL, who + "batched cases use liMRT without liSurfTen");
        if (c->rho0 > 0) return bfail(b, MPLB_E_MODEL, who + "batched cases need EOS/rho_0 <= 0");
        if (c->opts.heat_transfer || c->opts.frozen)
            return bfail(b, MPLB_E_MODEL, who + "batched cases run with heat transfer off and f updated");
        if (i == 0) {
            b->eos = c->eos;
            for (int q = 0; q < 9; q++) b->lam[q] = mp.Lambda[q];
            b->sigma = mp.sigma;
        } else if (std::memcmp(&b->eos, &c->eos, sizeof(EosParams)) != 0 ||
                   std::memcmp(b->lam, mp.Lambda, sizeof(b->lam)) != 0 || b->sigma != mp.sigma) {
            return bfail(b, MPLB_E_MODEL, who + "batched cases share the EOS block and the f model constants");
        }
        if (m.nx > kMarchLanes) b->march_ok = false;
        BCase bc{};
        bc.nx = m.nx;
        bc.ny = m.ny;
        bc.off = off;
        bc.n = m.N;
        b->cases.push_back(bc);
        b->rho_init.emplace_back();
        off += ((m.N + 63) / 64) * 64;
    }
    b->case_steps.assign(n, 0);
    b->P = off;
    if (!b->march_ok) b->mrows = 0;
    BCHK(b, hipSetDevice(device));
    BCHK(b, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    BCHK(b, hipMalloc(&b->fa, b->sbytes * 9 * b->P));
    BCHK(b, hipMalloc(&b->fb, b->sbytes * 9 * b->P));
    BCHK(b, hipMalloc(&b->psi, b->cbytes * b->P));
    BCHK(b, hipMalloc(&b->rho, b->cbytes * b->P));
    BCHK(b, hipMalloc(&b->p, b->cbytes * b->P));
    BCHK(b, hipMemset(b->fa, 0, b->sbytes * 9 * b->P));
    BCHK(b, hipMemset(b->fb, 0, b->sbytes * 9 * b->P));
    if (b->prec == MPLB_FP16N) {
        BCHK(b, hipMalloc(&b->ra, sizeof(_Float16) * b->P));
        BCHK(b, hipMalloc(&b->rb, sizeof(_Float16) * b->P));
        BCHK(b, hipMemset(b->ra, 0, sizeof(_Float16) * b->P));
        BCHK(b, hipMemset(b->rb, 0, sizeof(_Float16) * b->P));
    }
    BCHK(b, hipMalloc(&b->d_cases, sizeof(BCase) * n));
    BCHK(b, hipMalloc(&b->d_delta, sizeof(unsigned) * n));
    return 0;
}

const char* mplb_batch_last_error(const mplb_batch* b) { return b ? b->err.c_str() : "null batch handle"; }

int mplb_batch_set_case(mplb_batch* b, int i, const double* rho, size_t n, double T) {
    if (!b) return MPLB_E_ARG;
    if (i < 0 || i >= (int)b->cases.size()) return bfail(b, MPLB_E_ARG, "case index out of range");
    if (!rho || n != (size_t)b->cases[i].n)
        return bfail(b, MPLB_E_ARG, "density array size differs from the case's nodes");
    b->rho_init[i].assign(rho, rho + n);
    b->cases[i].T = T;
    double s = 0;
    for (size_t k = 0; k < n; k++) s += rho[k];
    b->cases[i].rho_ref = (float)(s / (double)n);
    b->initialized = false;
    return 0;
}

int mplb_batch_init(mplb_batch* b) {
    if (!b) return MPLB_E_ARG;
    for (size_t c = 0; c < b->cases.size(); c++)
        if ((int64_t)b->rho_init[c].size() != b->cases[c].n)
            return bfail(b, MPLB_E_STATE,
                         "case " + std::to_string(c) + " has no initial density (mplb_batch_set_case)");
    BCHK(b, hipSetDevice(b->dev));
    return BATCH_CALL(b, init_t, b);
}

int mplb_batch_step(mplb_batch* b, int n) {
    if (!b || n < 0) return MPLB_E_ARG;
    if (!b->initialized) return bfail(b, MPLB_E_STATE, "mplb_batch_init first");
    BCHK(b, hipSetDevice(b->dev));
    bool any = false;
    for (auto& c : b->cases) {
        any |= c.state != 0;
        c.state = 0;
    }
    if (any) {
        int rc = rebuild_work(b);
        if (rc) return rc;
    }
    int rc = BATCH_CALL(b, steps_plain, b, n);
    if (rc) return rc;
    BCHK(b, hipStreamSynchronize(b->stream));
    b->steps += n;
    for (auto& s : b->case_steps) s += n;
    b->fields_valid = false;
    return 0;
}

int mplb_batch_tune(mplb_batch* b, int steps_per_trial, int* chosen_block, int* chosen_xcd_order) {
    if (!b) return MPLB_E_ARG;
    if (!b->initialized) return bfail(b, MPLB_E_STATE, "mplb_batch_init first");
    BCHK(b, hipSetDevice(b->dev));
    for (auto& c : b->cases) c.state = 0;
    const int cand[] = {64, 128, 256, 512, 1024};
    float best = 1e30f;
    int best_b = b->bsize, best_x = b->xcd_order;
    hipEvent_t e0, e1;
    BCHK(b, hipEventCreate(&e0));
    BCHK(b, hipEventCreate(&e1));
    const int trial = std::max(2, steps_per_trial & ~1);    // even: the buffers end where they started
    // candidates: rows per single-pass item (when every case fits a workgroup row), then
    // nodes per workgroup of the two-pass kernels
    const int rows_cand[] = {16, 32, 64};
    int best_r = b->mrows;
    const int bs0 = b->bsize;
    for (int xo = 0; xo < 2; xo++) {
        for (int ci = b->march_ok ? 0 : 3; ci < 8; ci++) {
            const int cr = ci < 3 ? rows_cand[ci] : 0;
            const int cb = ci < 3 ? bs0 : cand[ci - 3];
            b->mrows = cr;
            b->bsize = cb;
            b->xcd_order = xo;
            int rc = rebuild_work(b);
            if (rc) return rc;
            rc = BATCH_CALL(b, steps_plain, b, 2);
            if (rc) return rc;
            BCHK(b, hipEventRecord(e0, b->stream));
            rc = BATCH_CALL(b, steps_plain, b, trial);
            if (rc) return rc;
            BCHK(b, hipEventRecord(e1, b->stream));
            BCHK(b, hipEventSynchronize(e1));
            float ms = 0;
            BCHK(b, hipEventElapsedTime(&ms, e0, e1));
            b->steps += 2 + trial;
            for (auto& s : b->case_steps) s += 2 + trial;
            if (ms < best) {
                best = ms;
                best_b = cb;
                best_x = xo;
                best_r = cr;
            }
        }
    }
    BCHK(b, hipEventDestroy(e0));
    BCHK(b, hipEventDestroy(e1));
    b->bsize = best_b;
    b->xcd_order = best_x;
    b->mrows = best_r;
    b->fields_valid = false;
    if (chosen_block) *chosen_block = best_b;
    if (chosen_xcd_order) *chosen_xcd_order = best_x;
    return rebuild_work(b);
}

int mplb_batch_set_launch(mplb_batch* b, int block, int xcd_order) {
    if (!b) return MPLB_E_ARG;
    if (block < 64 || block > 1024 || (block & 63))
        return bfail(b, MPLB_E_ARG, "block must be a multiple of 64 in [64, 1024]");
    b->bsize = block;
    b->xcd_order = xcd_order ? 1 : 0;
    b->mrows = 0;
    if (!b->initialized) return 0;
    BCHK(b, hipSetDevice(b->dev));
    return rebuild_work(b);
}

int mplb_batch_set_march(mplb_batch* b, int rows) {
    if (!b) return MPLB_E_ARG;
    if (rows < 0) return bfail(b, MPLB_E_ARG, "rows must be >= 0");
    if (rows > 0 && !b->march_ok)
        return bfail(b, MPLB_E_MODEL, "the single-pass kernel needs every case at most 256 nodes wide");
    b->mrows = rows;
    if (!b->initialized) return 0;
    BCHK(b, hipSetDevice(b->dev));
    return rebuild_work(b);
}

int mplb_batch_march_rows(const mplb_batch* b) { return b ? b->mrows : MPLB_E_ARG; }

int mplb_batch_set_persistent(mplb_batch* b, int on) {
    if (!b) return MPLB_E_ARG;
    if (on && !b->march_ok)
        return bfail(b, MPLB_E_MODEL, "the persistent kernel is the single pass: every case at most 256 nodes wide");
    b->persist = on != 0;
    return 0;
}

int mplb_batch_run_to_steady(mplb_batch* b, double threshold, int check_every, int64_t max_steps,
                             int64_t* case_steps, double* case_delta, int* case_state) {
    if (!b) return MPLB_E_ARG;
    if (!b->initialized) return bfail(b, MPLB_E_STATE, "mplb_batch_init first");
    if (check_every < 2 || (check_every & 1)) return bfail(b, MPLB_E_ARG, "check_every must be even and >= 2");
    BCHK(b, hipSetDevice(b->dev));
    const int nc = (int)b->cases.size();
    // steady criterion scale: the largest initial density of the case (its liquid density)
    std::vector<double> rl(nc);
    for (int c = 0; c < nc; c++) rl[c] = *std::max_element(b->rho_init[c].begin(), b->rho_init[c].end());
    std::vector<unsigned> dbits(nc);
    std::vector<double> last(nc, NAN);
    int rc = 0;
    if (std::none_of(b->cases.begin(), b->cases.end(), [](const BCase& c) { return c.state == 0; })) {
        for (auto& c : b->cases) c.state = 0;
        rc = rebuild_work(b);
        if (rc) return rc;
    }
    rc = BATCH_CALL(b, measure_t, b, false);
    if (rc) return rc;
    int64_t done = 0;
    while (done < max_steps && b->nwork > 0) {
        rc = BATCH_CALL(b, run_graph, b, check_every);
        if (rc) return rc;
        rc = BATCH_CALL(b, measure_t, b, true);
        if (rc) return rc;
        BCHK(b, hipMemcpyAsync(dbits.data(), b->d_delta, sizeof(unsigned) * nc, hipMemcpyDeviceToHost, b->stream));
        BCHK(b, hipStreamSynchronize(b->stream));
        done += check_every;
        b->steps += check_every;
        bool changed = false;
        for (int c = 0; c < nc; c++) {
            if (b->cases[c].state != 0) continue;
            float d;
            std::memcpy(&d, &dbits[c], 4);
            last[c] = d / rl[c];
            b->case_steps[c] += check_every;
            if (std::isnan(d)) {
                b->cases[c].state = 2;
                changed = true;
            } else if (last[c] < threshold) {
                b->cases[c].state = 1;
                changed = true;
            }
        }
        if (changed) {
            rc = rebuild_work(b);
            if (rc) return rc;
        }
    }
    b->fields_valid = false;
    int nconv = 0;
    for (int c = 0; c < nc; c++) {
        nconv += b->cases[c].state == 1;
        if (case_steps) case_steps[c] = b->case_steps[c];
        if (case_delta) case_delta[c] = last[c];
        if (case_state) case_state[c] = b->cases[c].state;
    }
    return nconv;
}

int mplb_batch_get_field(mplb_batch* b, int i, const char* name, double* out, size_t n) {
    if (!b || !name) return MPLB_E_ARG;
    if (!b->initialized) return bfail(b, MPLB_E_STATE, "mplb_batch_init first");
    if (i < 0 || i >= (int)b->cases.size()) return bfail(b, MPLB_E_ARG, "case index out of range");
    if (!out || n != (size_t)b->cases[i].n) return bfail(b, MPLB_E_ARG, "output size differs from the case's nodes");
    void* src;
    if (!std::strcmp(name, "rho")) src = b->rho;
    else if (!std::strcmp(name, "p")) src = b->p;
    else return bfail(b, MPLB_E_ARG, std::string("unknown batch field ") + name);
    BCHK(b, hipSetDevice(b->dev));
    if (!b->fields_valid) {
        int rc = BATCH_CALL(b, fields_t, b);
        if (rc) return rc;
    }
    const int64_t off = b->cases[i].off;
    if (b->prec == MPLB_FP64) {
        BCHK(b, hipMemcpy(out, (double*)src + off, sizeof(double) * n, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> h(n);
        BCHK(b, hipMemcpy(h.data(), (float*)src + off, sizeof(float) * n, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < n; k++) out[k] = h[k];
    }
    return 0;
}

int64_t mplb_batch_case_steps(const mplb_batch* b, int i) {
    if (!b || i < 0 || i >= (int)b->cases.size()) return -1;
    return b->case_steps[i];
}

int mplb_batch_launch_info(const mplb_batch* b, int* block, int* xcd_order, int* workgroups, int64_t* nodes) {
    if (!b) return MPLB_E_ARG;
    if (block) *block = b->bsize;
    if (xcd_order) *xcd_order = b->xcd_order;
    if (workgroups) *workgroups = b->nwork;
    if (nodes) {
        int64_t s = 0;
        for (const auto& c : b->cases) s += c.n;
        *nodes = s;
    }
    return 0;
}

int mplb_batch_close(mplb_batch* b) {
    if (!b) return MPLB_E_ARG;
    // teardown: failures here have nobody left to report to
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    void* ptrs[] = {b->fa, b->fb, b->ra, b->rb, b->psi, b->rho, b->p, b->d_cases, b->d_work, b->d_mwork, b->d_delta};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return 0;
}

}  // extern "C"
