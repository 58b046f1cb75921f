// Per-routine HIP kernels of the lattice update, one thread per lattice node, for any
// mesh the reference's tables describe (neighbour table gather, boundary lists).  These
// are the FP64 parity kernels (this file is compiled with -ffp-contract=off and every
// expression keeps the reference's association order) and the general FP32 path; the
// fused structured-box kernels live in kernels_fused.hip.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace mplb {

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

template <class L, typename T>
__device__ inline void load_pops(const T* __restrict__ f, int64_t N, int64_t id, T* v) {
#pragma unroll
    for (int k = 0; k < L::Q; k++) v[k] = f[k * N + id];
}

template <class L, typename T>
__device__ inline void store_pops(T* __restrict__ f, int64_t N, int64_t id, const T* v) {
#pragma unroll
    for (int k = 0; k < L::Q; k++) f[k * N + id] = v[k];
}

// ---------------------------------------------------------------------------
// psi = potential(rho, T): potential.c:7-13.  A pure function of (rho, T), so storing
// it once per step is bit-identical to the reference's repeated evaluation.
template <typename T>
__global__ void k_psi(int64_t N, EosParams e, const T* __restrict__ rho, const T* __restrict__ temp,
                      T* __restrict__ psi) {
    const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id >= N) return;
    psi[id] = potential<T>(e, rho[id], temp[id]);
}

// interForce.c:16-39 with interactionForce.c:47-79
template <class L, typename T>
__global__ void k_inter_force(DevState<T> st, EosParams e, T rho0) {
    const int64_t N = st.N;
    const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id >= N) return;
    T F[3] = {T(0.0), T(0.0), T(0.0)};
    T Ff[3] = {T(0), T(0), T(0)};
    const bool interior = st.onbnd[id] == 0;
    if (interior) {
#pragma unroll
        for (int k = 1; k < L::Q; k++) {
            const int nid = st.nb[L::t.rev[k] * N + id];
            const T alpha = T(L::t.w[k]) * st.psi[nid];
#pragma unroll
            for (int i = 0; i < 3; i++) {
                const int c = L::t.c[k][i];
                if (c == 1) F[i] += alpha;
                else if (c == -1) F[i] += -alpha;
            }
        }
        const T beta = -T(e.G) * st.psi[id];
#pragma unroll
        for (int i = 0; i < 3; i++) Ff[i] = F[i] * beta;
    }
    const T r = st.rho[id];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        st.Fi[j * N + id] = Ff[j];
        T ft = Ff[j] + (r - rho0) * T(e.g[j]);
        if (interior) ft += T(e.fext[j]);
        st.Ft[j * N + id] = ft;
    }
}

// liMRTSurfaceTension.c:11-132 (D2Q9) / xuMRTSurfaceTension.c:57-122 (D3Q15)
template <class L, typename T>
__device__ inline void surface_tension(const DevState<T>& st, const EosParams& e, const ModelParams& mp,
                                       int64_t id, T* C) {
    const int64_t N = st.N;
    const T psi = st.psi[id];
    int noneigh = 0;
#pragma unroll
    for (int k = 1; k < L::Q; k++)
        if (st.nb[k * N + id] == -1) noneigh++;
#pragma unroll
    for (int k = 0; k < L::Q; k++) C[k] = T(0);
    if (noneigh != 0) return;
    T Qt[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
    for (int k = 1; k < L::Q; k++) {
        const int nid = st.nb[L::t.rev[k] * N + id];
        const T alpha = T(L::t.w[k]) * (st.psi[nid] - psi);
#pragma unroll
        for (int i = 0; i < L::D; i++)
#pragma unroll
            for (int j = (L::D == 2 ? 0 : i); j < L::D; j++) {
                const int cc = L::t.c[k][i] * L::t.c[k][j];
                if (cc == 1) Qt[i][j] += alpha;
                else if (cc == -1) Qt[i][j] += -alpha;
            }
    }
    const T G = T(e.G), kap = T(mp.kappa);
#pragma unroll
    for (int i = 0; i < L::D; i++)
#pragma unroll
        for (int j = (L::D == 2 ? 0 : i); j < L::D; j++) Qt[i][j] = Qt[i][j] * T(0.5) * psi * G * kap;
    if (L::D == 2) {
        C[1] = T(1.5) * T(mp.Lambda[1]) * (Qt[0][0] + Qt[1][1]);
        C[2] = T(-1.5) * T(mp.Lambda[2]) * (Qt[0][0] + Qt[1][1]);
        C[7] = -T(mp.Lambda[8]) * (Qt[0][0] - Qt[1][1]);
        C[8] = -T(mp.Lambda[8]) * Qt[0][1];
    } else {
        C[1] = T(4.0) * T(mp.Lambda[1]) * (Qt[0][0] + Qt[1][1] + Qt[2][2]) / T(5.0);
        const T l10 = -T(mp.Lambda[10]);
        C[9] = l10 * (T(2) * Qt[0][0] - Qt[1][1] - Qt[2][2]);
        C[10] = l10 * (Qt[1][1] - Qt[2][2]);
        C[11] = l10 * Qt[0][1];
        C[12] = l10 * Qt[1][2];
        C[13] = l10 * Qt[0][2];
    }
}

// liMRTCollision.c:33-106 with liMRTForce.c:13-42 (D2Q9)
template <class L, typename T>
__global__ void k_collide_limrt(DevState<T> st, EosParams e, ModelParams mp, T* __restrict__ f) {
    const int64_t N = st.N;
    const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id >= N) return;
    T fv[L::Q], m[L::Q], meq[L::Q], S[L::Q], C[L::Q];
    load_pops<L, T>(f, N, id, fv);
    const T rho = st.rho[id];
    const T U[3] = {st.U[id], st.U[N + id], st.U[2 * N + id]};
    T umag = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) umag += U[k] * U[k];
    meq[0] = rho;
    meq[1] = rho * (T(-2) + T(3) * umag);
    meq[2] = rho * (T(1) - T(3) * umag);
    meq[3] = rho * U[0];
    meq[4] = rho * (-U[0]);
    meq[5] = rho * U[1];
    meq[6] = rho * (-U[1]);
    meq[7] = rho * (U[0] * U[0] - U[1] * U[1]);
    meq[8] = rho * U[0] * U[1];
    moments<L, T>(fv, m);
    const T F[3] = {st.Ft[id], st.Ft[N + id], st.Ft[2 * N + id]};
    const T Fi0 = st.Fi[id], Fi1 = st.Fi[N + id];
    const T psi = st.psi[id];
    const T sigma = T(mp.sigma);
    const T L1 = T(mp.Lambda[1]), L2 = T(mp.Lambda[2]);
    S[0] = 0;
    S[1] = T(6) * (U[0] * F[0] + U[1] * F[1]) +
           T(12) * sigma * (Fi0 * Fi0 + Fi1 * Fi1) / (psi * psi * ((T(1) / L1) - T(0.5)));
    S[2] = T(-6) * (U[0] * F[0] + U[1] * F[1]) -
           T(12) * sigma * (Fi0 * Fi0 + Fi1 * Fi1) / (psi * psi * ((T(1) / L2) - T(0.5)));
    S[3] = F[0];
    S[4] = -F[0];
    S[5] = F[1];
    S[6] = -F[1];
    S[7] = T(2) * (U[0] * F[0] - U[1] * F[1]);
    S[8] = U[0] * F[1] + U[1] * F[0];
    if (mp.surf) surface_tension<L, T>(st, e, mp, id, C);
#pragma unroll
    for (int k = 0; k < L::Q; k++) {
        const T lam = T(mp.Lambda[k]);
        m[k] = m[k] - lam * (m[k] - meq[k]) + (T(1) - T(0.5) * lam) * S[k];
        if (mp.surf) m[k] = m[k] + C[k];
    }
    populations<L, T>(m, fv);
    store_pops<L, T>(f, N, id, fv);
}

// xuMRTCollision.c:42-92 with xuMRTEquilibriumMS.c:23-51 and xuMRTForce.c:18-69 (D3Q15)
template <class L, typename T>
__global__ void k_collide_xumrt(DevState<T> st, EosParams e, ModelParams mp, T* __restrict__ f, T rho0) {
    const int64_t N = st.N;
    const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id >= N) return;
    T fv[L::Q], m[L::Q], meq[L::Q], S[L::Q], C[L::Q];
    load_pops<L, T>(f, N, id, fv);
    const T U[3] = {st.U[id], st.U[N + id], st.U[2 * N + id]};
    const T rho = st.rho[id];
    T umag = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) umag += U[k] * U[k];
    meq[0] = rho;
    meq[1] = rho * (T(-1.0) + umag);
    meq[2] = rho * (T(1.0) - T(5.0) * umag);
    meq[3] = rho * U[0];
    meq[4] = rho * T(-7.0) * U[0] / T(3.0);
    meq[5] = rho * U[1];
    meq[6] = rho * T(-7.0) * U[1] / T(3.0);
    meq[7] = rho * U[2];
    meq[8] = rho * T(-7.0) * U[2] / T(3.0);
    meq[9] = rho * (T(2.0) * U[0] * U[0] - U[1] * U[1] - U[2] * U[2]);
    meq[10] = rho * (U[1] * U[1] - U[2] * U[2]);
    meq[11] = rho * U[0] * U[1];
    meq[12] = rho * U[1] * U[2];
    meq[13] = rho * U[0] * U[2];
    meq[14] = 0;
    moments<L, T>(fv, m);
    T F[3];
#pragma unroll
    for (int j = 0; j < 3; j++) F[j] = st.Fi[j * N + id] + (rho - rho0) * T(e.g[j]);
    const T psi = st.psi[id];
    const T uDotF = U[0] * F[0] + U[1] * F[1] + U[2] * F[2];
    T fmag = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) fmag += F[k] * F[k];
    const T sigma = T(mp.sigma), L1 = T(mp.Lambda[1]);
    S[0] = 0;
    S[1] = T(2.0) * uDotF + (T(6.0) * sigma * fmag) / (psi * psi * ((T(1) / L1) - T(0.5)));
    S[2] = T(-10.0) * uDotF;
    S[3] = F[0];
    S[4] = T(-7.0) * F[0] / T(3.0);
    S[5] = F[1];
    S[6] = T(-7.0) * F[1] / T(3.0);
    S[7] = F[2];
    S[8] = T(-7.0) * F[2] / T(3.0);
    S[9] = T(4.0) * U[0] * F[0] - T(2.0) * U[1] * F[1] - T(2.0) * U[2] * F[2];
    S[10] = T(2.0) * U[1] * F[1] - T(2.0) * U[2] * F[2];
    S[11] = U[0] * F[1] + U[1] * F[0];
    S[12] = U[1] * F[2] + U[2] * F[1];
    S[13] = U[0] * F[2] + U[2] * F[0];
    S[14] = 0;
    if (mp.surf) surface_tension<L, T>(st, e, mp, id, C);
#pragma unroll
    for (int k = 0; k < L::Q; k++) {
        const T lam = T(mp.Lambda[k]);
        m[k] = m[k] - lam * (m[k] - meq[k]) + (T(1) - T(0.5) * lam) * S[k];
        if (mp.surf) m[k] = m[k] + C[k];
    }
    populations<L, T>(m, fv);
    store_pops<L, T>(f, N, id, fv);
}

// liTempMRTCollision.c:31-84 with liTempMRTEquilibriumMS.c:20-41, liTempMRTSource.c:17-24
template <class L, typename T>
__global__ void k_collide_litemp(DevState<T> st, ModelParams mp, T* __restrict__ g) {
    const int64_t N = st.N;
    const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id >= N) return;
    T gv[L::Q], n[L::Q], neq[L::Q];
    load_pops<L, T>(g, N, id, gv);
    const T temp = st.temp[id];
    const T U[3] = {st.U[id], st.U[N + id], st.U[2 * N + id]};
    neq[0] = temp;
    neq[1] = T(-2.0) * temp;
    neq[2] = T(2.0) * temp;
    neq[3] = temp * U[0];
    neq[4] = temp * (-U[0]);
    neq[5] = temp * U[1];
    neq[6] = temp * (-U[1]);
    neq[7] = 0;
    neq[8] = 0;
    moments<L, T>(gv, n);
    const T n4 = n[4], n6 = n[6];
#pragma unroll
    for (int k = 0; k < L::Q; k++) {
        const T lam = T(mp.Lambda[k]);
        const T gam = (k == 0) ? st.ss[id] : T(0);
        n[k] = n[k] - lam * (n[k] - neq[k]) + gam;
    }
    n[3] = n[3] + (T(1) - T(mp.Lambda[3]) / T(2)) * T(mp.Lambda[4]) * n4;
    n[5] = n[5] + (T(1) - T(mp.Lambda[5]) / T(2)) * T(mp.Lambda[6]) * n6;
    populations<L, T>(n, gv);
    store_pops<L, T>(g, N, id, gv);
}

// myMRTCollision.c:35-212 with myMRTEquilibriumMS.c:20-77 and sparseMatVecMult.c:11-28:
// Q = diag(Lambda) + Q34, Q56 (, Q78), Q_aux = I - Q/2; n <- n - Q (n - n_eq) + Q_aux Gamma.
template <class L, typename T>
__global__ void k_collide_my(DevState<T> st, ModelParams mp, T* __restrict__ g) {
    const int64_t N = st.N;
    const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id >= N) return;
    T gv[L::Q], n[L::Q], neq[L::Q], aux1[L::Q], aux2[L::Q];
    load_pops<L, T>(g, N, id, gv);
    const T temp = st.temp[id];
    const T U[3] = {st.U[id], st.U[N + id], st.U[2 * N + id]};
#pragma unroll
    for (int k = 0; k < L::Q; k++) neq[k] = T(0);
    neq[0] = temp;
    neq[1] = T(mp.alpha1) * temp;
    neq[2] = T(mp.alpha2) * temp;
    if (L::D == 2) {
        neq[3] = temp * U[0];
        neq[4] = temp * (-U[0]);
        neq[5] = temp * U[1];
        neq[6] = temp * (-U[1]);
    } else {
        neq[3] = temp * U[0];
        neq[4] = -(temp * U[0]);
        neq[5] = temp * U[1];
        neq[6] = -(temp * U[1]);
        neq[7] = temp * U[2];
        neq[8] = -(temp * U[2]);
    }
    moments<L, T>(gv, n);
    // relaxation matrix entries (myMRTCollision.c:39-127), computed as the reference does
    T Qd[L::Q], Qa[L::Q];
#pragma unroll
    for (int k = 0; k < L::Q; k++) {
        Qd[k] = T(mp.Lambda[k]);
        Qa[k] = T(1.0) + T(-0.5) * Qd[k];
    }
    const int npair = L::D == 2 ? 2 : 3;
    T Qo[3], Qao[3];
#pragma unroll
    for (int pr = 0; pr < 3; pr++) {
        const int a = 3 + 2 * pr, b = 4 + 2 * pr;
        Qo[pr] = Qd[b] * (Qd[a] / T(2.0) - T(1.0));
        Qao[pr] = T(-0.5) * Qo[pr];
    }
#pragma unroll
    for (int k = 0; k < L::Q; k++) aux2[k] = n[k] - neq[k];
#pragma unroll
    for (int k = 0; k < L::Q; k++) aux1[k] = Qd[k] * aux2[k];
#pragma unroll
    for (int pr = 0; pr < npair; pr++) aux1[3 + 2 * pr] = aux1[3 + 2 * pr] + Qo[pr] * aux2[4 + 2 * pr];
    const T gam0 = st.ss[id];
#pragma unroll
    for (int k = 0; k < L::Q; k++) aux2[k] = Qa[k] * (k == 0 ? gam0 : T(0));
#pragma unroll
    for (int pr = 0; pr < npair; pr++) aux2[3 + 2 * pr] = aux2[3 + 2 * pr] + Qao[pr] * T(0);
#pragma unroll
    for (int k = 0; k < L::Q; k++) n[k] = n[k] - aux1[k] + aux2[k];
    populations<L, T>(n, gv);
    store_pops<L, T>(g, N, id, gv);
}

// lbstream.c:16-62: f[x][k] <- f[nb[x][k]][k] where nb != -1, else unchanged.
template <typename T>
__global__ void k_stream(int64_t N, int Q, const int* __restrict__ nb, const T* __restrict__ in,
                         T* __restrict__ out) {
    const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id >= N) return;
    for (int k = 0; k < Q; k++) {
        const int src = nb[k * N + id];
        out[k * N + id] = in[k * N + (src != -1 ? (int64_t)src : id)];
    }
}

// ---------------------------------------------------------------------------
// Boundary conditions, one list node at a time (sequential in k, in place).

// bounceBack.c:14-75
template <class L, typename T>
__device__ inline void bc_bounceback(const DevState<T>& st, T* f, int64_t id) {
    const int64_t N = st.N;
    T v[L::Q];
    int miss[L::Q];
    load_pops<L, T>(f, N, id, v);
#pragma unroll
    for (int k = 0; k < L::Q; k++) miss[k] = st.nb[k * N + id] == -1;
#pragma unroll
    for (int k = 0; k < L::Q; k++) {
        const int r = L::t.rev[k];
        if (miss[k]) {
            v[k] = v[r];
        } else if (!miss[r] && k >= 1 && k <= L::naxis) {
            v[k] = T(0.5) * v[k] + T(0.5) * v[r];
            v[r] = v[k];
        }
    }
    store_pops<L, T>(f, N, id, v);
}

// fixedT.c:45-122 and fixedRandomT.c:48-112: g_k = geq_k(Tw, U_nb) + (g_k[nb] - geq_k(T_nb, U_nb)).
template <class L, typename T>
__device__ inline void bc_fixed_temp(const DevState<T>& st, const ModelParams& mp, T* g, int64_t id,
                                     const BdList& b, int list_index) {
    const int64_t N = st.N;
    int draw = 0;
    const bool random = b.type == BD_FRDT;
    T minT = 0, maxT = 0;
    if (random) {
        minT = T(b.value[0]) * (T(100.0) - T(b.pert)) / T(100.0);
        maxT = T(b.value[0]) * (T(100.0) + T(b.pert)) / T(100.0);
        draw = b.rnd_off[list_index];
    }
    for (int k = 0; k < L::Q; k++) {
        if (st.nb[k * N + id] != -1) continue;
        const int nbid = st.nb[L::t.rev[k] * N + id];
        if (nbid == -1) continue;
        const T Unb[3] = {st.U[nbid], st.U[N + nbid], st.U[2 * N + nbid]};
        T geq[L::Q];
        if (mp.model == M_MYMRT) my_geq<L, T>(st.temp[nbid], Unb, T(mp.alpha1), T(mp.alpha2), geq);
        else litemp_geq<L, T>(st.temp[nbid], Unb, geq);
        const T eq_nb = geq[k];
        T Tw = T(b.value[0]);
        if (random && b.draw) {
            const T r = T(b.rnd[draw++]);
            Tw = minT + (maxT - minT) * r;
        }
        if (mp.model == M_MYMRT) my_geq<L, T>(Tw, Unb, T(mp.alpha1), T(mp.alpha2), geq);
        else litemp_geq<L, T>(Tw, Unb, geq);
        const T eq_bnd = geq[k];
        g[k * N + id] = eq_bnd + (g[k * N + nbid] - eq_nb);
    }
}

// fixedU.c:39-131 (liMRT): the neighbour's density and velocity are recomputed from its
// post-stream populations with the force Fi + (rho - rho0) g.
template <class L, typename T>
__device__ inline void bc_fixed_u(const DevState<T>& st, const EosParams& e, T* f, int64_t id, const BdList& b,
                                  T rho0) {
    const int64_t N = st.N;
    const T Uw[3] = {T(b.value[0]), T(b.value[1]), T(b.value[2])};
    for (int k = 0; k < L::Q; k++) {
        if (st.nb[k * N + id] != -1) continue;
        const int nbid = st.nb[L::t.rev[k] * N + id];
        if (nbid == -1) continue;
        T lv[3] = {0, 0, 0};
        T nbrho = 0;
        T F[3];
#pragma unroll
        for (int jj = 0; jj < 3; jj++) F[jj] = st.Fi[jj * N + nbid] + (st.rho[nbid] - rho0) * T(e.g[jj]);
        T fn[L::Q];
        load_pops<L, T>(f, N, nbid, fn);
#pragma unroll
        for (int jj = 0; jj < 3; jj++)
#pragma unroll
            for (int kk = 0; kk < L::Q; kk++) {
                const int c = L::t.c[kk][jj];
                if (c == 1) lv[jj] += fn[kk];
                else if (c == -1) lv[jj] += -fn[kk];
            }
#pragma unroll
        for (int kk = 0; kk < L::Q; kk++) nbrho += fn[kk];
#pragma unroll
        for (int jj = 0; jj < 3; jj++) lv[jj] = (lv[jj] + F[jj] * T(0.5)) / nbrho;
        const T feq_nb = ns_feq<L, T>(k, nbrho, lv);
        const T feq_bnd = ns_feq<L, T>(k, nbrho, Uw);
        f[k * N + id] = feq_bnd + (fn[k] - feq_nb);
    }
}

// outflow.c:26-195 (normal accumulated and normalised in float, as the reference does)
template <class L, typename T>
__device__ inline void bc_outflow(const DevState<T>& st, const EosParams& e, T* f, int64_t id, T rho0) {
    const int64_t N = st.N;
    float normal[3] = {0, 0, 0};
    for (int k = 0; k < L::Q; k++)
        if (st.nb[L::t.rev[k] * N + id] == -1)
            for (int j = 0; j < 3; j++) normal[j] += (float)L::t.c[k][j];
    float nmag = 0;
    for (int j = 0; j < 3; j++) nmag += normal[j] * normal[j];
    if (nmag == 0) return;   // rejected on the host before launch
    nmag = (float)sqrt((double)nmag);
    for (int j = 0; j < 3; j++) normal[j] = normal[j] / nmag;
    int ln = 0;
    for (int k = 0; k < L::Q; k++)
        if (L::t.c[k][0] == (int)normal[0] && L::t.c[k][1] == (int)normal[1] && L::t.c[k][2] == (int)normal[2]) ln = k;
    const int neigh = st.nb[ln * N + id];
    if (neigh == -1) return;
    T nbv[3] = {0, 0, 0};
    T nbrho = 0;
    T F[3];
    for (int jj = 0; jj < 3; jj++) F[jj] = st.Fi[jj * N + neigh] + (st.rho[neigh] - rho0) * T(e.g[jj]);
    T fn[L::Q];
    load_pops<L, T>(f, N, neigh, fn);
#pragma unroll
    for (int jj = 0; jj < 3; jj++)
#pragma unroll
        for (int kk = 0; kk < L::Q; kk++) {
            const int c = L::t.c[kk][jj];
            if (c == 1) nbv[jj] += fn[kk];
            else if (c == -1) nbv[jj] += -fn[kk];
        }
#pragma unroll
    for (int kk = 0; kk < L::Q; kk++) nbrho += fn[kk];
    for (int jj = 0; jj < 3; jj++) nbv[jj] = (nbv[jj] + F[jj] * T(0.5)) / nbrho;
    T uAdv = 0;
    for (int jj = 0; jj < 3; jj++) uAdv += T(normal[jj]) * nbv[jj];
    for (int k = 0; k < L::Q; k++)
        if (st.nb[k * N + id] == -1) f[k * N + id] = (f[k * N + id] + uAdv * fn[k]) / (T(1) + uAdv);
}

template <class L, typename T>
__device__ inline void apply_bc(const DevState<T>& st, const EosParams& e, const ModelParams& mp, T* f,
                                const BdList& b, int i, T rho0) {
    const int64_t id = b.nodes[i];
    switch (b.type) {
    case BD_BBACK: bc_bounceback<L, T>(st, f, id); break;
    case BD_FT:
    case BD_FRDT: bc_fixed_temp<L, T>(st, mp, f, id, b, i); break;
    case BD_FU: bc_fixed_u<L, T>(st, e, f, id, b, rho0); break;
    case BD_OF: bc_outflow<L, T>(st, e, f, id, rho0); break;
    default: break;
    }
}

template <class L, typename T>
__global__ void k_boundary_parallel(DevState<T> st, EosParams e, ModelParams mp, T* f, BdList b, T rho0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.n) return;
    apply_bc<L, T>(st, e, mp, f, b, i, rho0);
}

// Lists whose nodes read populations that other nodes of the same list write are
// applied in list order by one thread, exactly as the reference's loop.
template <class L, typename T>
__global__ void k_boundary_sequential(DevState<T> st, EosParams e, ModelParams mp, T* f, BdList b, T rho0) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int i = 0; i < b.n; i++) apply_bc<L, T>(st, e, mp, f, b, i, rho0);
}

// ---------------------------------------------------------------------------
// macroDensity.c:19-60
template <class L, typename T>
__global__ void k_density(DevState<T> st, const T* __restrict__ f) {
    const int64_t N = st.N;
    const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id >= N) return;
    T r = 0;
#pragma unroll
    for (int k = 0; k < L::Q; k++) r += f[k * N + id];
    st.rho[id] = r;
}

// macroVelocity.c:17-50 with liMRTVelocity.c:11-50 / xuMRTVelocity.c:19-50
template <class L, typename T>
__global__ void k_velocity(DevState<T> st, EosParams e, int fmodel, const T* __restrict__ f, T rho0) {
    const int64_t N = st.N;
    const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id >= N) return;
    T fv[L::Q];
    load_pops<L, T>(f, N, id, fv);
    const T rho = st.rho[id];
    T F[3];
#pragma unroll
    for (int j = 0; j < 3; j++)
        F[j] = fmodel == M_LIMRT ? st.Ft[j * N + id] : st.Fi[j * N + id] + (rho - rho0) * T(e.g[j]);
#pragma unroll
    for (int j = 0; j < 3; j++) {
        T lv = 0;
#pragma unroll
        for (int k = 0; k < L::Q; k++) {
            const int c = L::t.c[k][j];
            if (c == 1) lv += fv[k];
            else if (c == -1) lv += -fv[k];
        }
        st.U[j * N + id] = (lv + F[j] * T(0.5)) / rho;
    }
}

// macroTemperature.c:15-63 with myMRTTemperature.c:16-30 / liTempMRTTemperature.c:16-20
template <class L, typename T>
__global__ void k_temperature(DevState<T> st, int gmodel, const T* __restrict__ g) {
    const int64_t N = st.N;
    const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id >= N) return;
    T t = 0;
#pragma unroll
    for (int k = 0; k < L::Q; k++) t += g[k * N + id];
    if (gmodel == M_MYMRT) t = t + T(0.5) * st.ss[id];
    st.temp[id] = t;
}

// finiteDifference/{scalarGradient,invScalarGradient,vectorDivergence}.c along one axis.
template <typename T>
__device__ inline T fd_axis(const T* q, int a, int b, int64_t id, bool inv, bool* has) {
    *has = true;
    if (a != -1 && b != -1) {
        const T qa = inv ? T(1.0) / q[a] : q[a];
        const T qb = inv ? T(1.0) / q[b] : q[b];
        return T(0.5) * (qa - qb);
    }
    if (a == -1 && b != -1) {
        const T qi = inv ? T(1.0) / q[id] : q[id];
        const T qb = inv ? T(1.0) / q[b] : q[b];
        return qi - qb;
    }
    if (a != -1 && b == -1) {
        const T qa = inv ? T(1.0) / q[a] : q[a];
        const T qi = inv ? T(1.0) / q[id] : q[id];
        return qa - qi;
    }
    *has = false;
    return T(0);
}

// heatSource.c:26-170: S = -lambda (grad T . grad 1/rho)/Cv + div U T (1 - dp/dT/(rho Cv)),
// with old T and U and new rho; liTempMRT adds 0.5 (S - S_old).
template <class L, typename T>
__global__ void k_heat_source(DevState<T> st, EosParams e, ModelParams mp, int first_call) {
    const int64_t N = st.N;
    const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id >= N) return;
    const T rho = st.rho[id], temp = st.temp[id];
    const T Cv = T(e.Cv);
    const T L3 = T(mp.Lambda[3]);
    T lambda;
    if (mp.model == M_MYMRT) {
        if (L::D == 2)
            lambda = rho * Cv * (T(1) / L3 - T(0.5)) * (T(4.0) + T(3.0) * T(mp.alpha1) + T(2.0) * T(mp.alpha2)) / T(6.0);
        else
            lambda = rho * Cv * (T(1) / L3 - T(0.5)) * (T(6.0) + T(11.0) * T(mp.alpha1) + T(mp.alpha2)) / T(9.0);
    } else {
        lambda = rho * Cv * (T(1) / L3 - T(0.5));
    }
    T gT[3] = {0, 0, 0}, gR[3] = {0, 0, 0};
    T divU = 0;
#pragma unroll
    for (int ax = 0; ax < L::D; ax++) {
        const int a = st.nb[L::fd_plus[ax] * N + id];
        const int b = st.nb[L::fd_minus[ax] * N + id];
        bool has;
        gT[ax] = fd_axis<T>(st.temp, a, b, id, false, &has);
        gR[ax] = fd_axis<T>(st.rho, a, b, id, true, &has);
        const T dv = fd_axis<T>(st.U + ax * N, a, b, id, false, &has);
        if (has) divU += dv;
    }
    T dot = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) dot += gT[k] * gR[k];
    T dpdT = p_eos<T>(e, rho, T(1.1) * temp) - p_eos<T>(e, rho, T(0.9) * temp);
    dpdT = dpdT / (T(0.2) * temp);
    const T s = -lambda * dot / Cv + divU * temp * (T(1.0) - dpdT / (rho * Cv));
    if (mp.model == M_MYMRT) {
        st.ss[id] = s;
    } else {
        const T oldss = first_call ? T(0) : st.ss[id];
        st.ss[id] = s + T(0.5) * (s - oldss);
    }
}

// equilibrium.c:17-85
template <class L, typename T>
__global__ void k_equilibrium(DevState<T> st, ModelParams mp, T* __restrict__ f) {
    const int64_t N = st.N;
    const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id >= N) return;
    const T U[3] = {st.U[id], st.U[N + id], st.U[2 * N + id]};
    T fv[L::Q];
    if (mp.model == M_LIMRT || mp.model == M_XUMRT) {
        const T rho = st.rho[id];
#pragma unroll
        for (int k = 0; k < L::Q; k++) fv[k] = ns_feq<L, T>(k, rho, U);
    } else if (mp.model == M_MYMRT) {
        my_geq<L, T>(st.temp[id], U, T(mp.alpha1), T(mp.alpha2), fv);
    } else {
        if (L::D != 2) return;   // liTempMRTEquilibrium.c:8: D2Q9 only, populations kept
        litemp_geq<L, T>(st.temp[id], U, fv);
    }
    store_pops<L, T>(f, N, id, fv);
}

// macroPressure.c:20-65 with liMRTPressure.c:6-253 (D2Q9 stencil) / xuMRTPressure.c:12
template <class L, typename T>
__global__ void k_pressure(DevState<T> st, EosParams e, int fmodel, T sigma) {
    const int64_t N = st.N;
    const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id >= N) return;
    const T G = T(e.G);
    const T psi = st.psi[id];
    T p = st.rho[id] * T(kCs2) + T(0.5) * G * psi * psi;
    if (fmodel == M_LIMRT) {
        T lap = 0;
        T grad[3] = {0, 0, 0};
        if (L::D == 2) {
#pragma unroll
            for (int ax = 0; ax < 2; ax++) {
                const int ka = ax == 0 ? 3 : 4, kb = ax == 0 ? 1 : 2;
                const int a = st.nb[ka * N + id], b = st.nb[kb * N + id];
                // x: lap = ...; y: lap = lap + ... (left to right, liMRTPressure.c:40-117)
                if (a != -1 && b != -1) {
                    lap = ax == 0 ? st.psi[a] + st.psi[b] - T(2) * psi : lap + st.psi[a] + st.psi[b] - T(2) * psi;
                } else if (a == -1 && b != -1) {
                    const int c = st.nb[kb * N + b];
                    lap = ax == 0 ? st.psi[c] - T(2) * st.psi[b] + psi : lap + st.psi[c] - T(2) * st.psi[b] + psi;
                } else if (a != -1 && b == -1) {
                    const int c = st.nb[ka * N + a];
                    lap = ax == 0 ? st.psi[c] - T(2) * st.psi[a] + psi : lap + st.psi[c] - T(2) * st.psi[a] + psi;
                } else if (ax == 0) {
                    lap = 0;
                }
                if (a != -1 && b != -1) grad[ax] = T(0.5) * (st.psi[a] - st.psi[b]);
                else if (a == -1 && b != -1) grad[ax] = psi - st.psi[b];
                else if (a != -1 && b == -1) grad[ax] = st.psi[a] - psi;
                else grad[ax] = 0;
            }
        }
        p = p + G * psi * lap / T(12);
        T g2 = 0;
#pragma unroll
        for (int k = 0; k < 3; k++) g2 += grad[k] * grad[k];
        p = p + T(2.0) * G * G * sigma * g2;
    }
    st.p[id] = p;
}

// Serial-order sum (averageScalarField.c:28-32): chunks staged through LDS, one lane
// accumulates in index order.
template <typename T>
__global__ void k_sum_ordered(const T* __restrict__ x, int64_t n, double* out) {
    __shared__ T buf[1024];
    double s = 0;
    for (int64_t base = 0; base < n; base += 1024) {
        const int64_t m = (n - base) < 1024 ? (n - base) : 1024;
        for (int i = threadIdx.x; i < m; i += blockDim.x) buf[i] = x[base + i];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int i = 0; i < m; i++) s += (double)buf[i];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s;
}

// Tree sum with FP64 partials: wavefront shuffles, then one partial per block.
template <typename T>
__global__ void k_sum_tree_partial(const T* __restrict__ x, int64_t n, double* part) {
    __shared__ double ws[kBlock / 64];
    double s = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        s += (double)x[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < kBlock / 64; w++) t += ws[w];
        part[blockIdx.x] = t;
    }
}

__global__ void k_sum_tree_final(const double* __restrict__ part, int n, double* out) {
    __shared__ double ws[kBlock / 64];
    double s = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += part[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < kBlock / 64; w++) t += ws[w];
        *out = t;
    }
}

template <typename T>
__global__ void k_nan_count(const T* __restrict__ x, int64_t n, unsigned int* out) {
    unsigned int c = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        c += (x[i] != x[i]) ? 1u : 0u;
    if (c) atomicAdd(out, c);
}

template <class L, typename T>
void launch_collide(hipStream_t s, const EosParams& e, const ModelParams& mp, DevState<T> st, T* f, double rho0) {
    const unsigned g = grid_for(st.N);
    switch (mp.model) {
    case M_LIMRT: k_collide_limrt<L, T><<<g, kBlock, 0, s>>>(st, e, mp, f); break;
    case M_XUMRT:
        // D3Q15 only: case set-up refuses xuMRT on D2Q9 (the reference's own error)
        if constexpr (L::Q == 15) k_collide_xumrt<L, T><<<g, kBlock, 0, s>>>(st, e, mp, f, T(rho0));
        break;
    case M_LITEMPMRT: k_collide_litemp<L, T><<<g, kBlock, 0, s>>>(st, mp, f); break;
    case M_MYMRT: k_collide_my<L, T><<<g, kBlock, 0, s>>>(st, mp, f); break;
    default: break;
    }
}

}  // namespace

#define MPLB_DISPATCH(model, ...)                      \
    do {                                               \
        if ((model) == 0) {                            \
            using L = D2Q9;                            \
            __VA_ARGS__;                               \
        } else {                                       \
            using L = D3Q15;                           \
            __VA_ARGS__;                               \
        }                                              \
    } while (0)

template <typename T>
void GenericOps<T>::psi(int model, hipStream_t s, const EosParams& e, DevState<T> st) {
    (void)model;
    k_psi<T><<<grid_for(st.N), kBlock, 0, s>>>(st.N, e, st.rho, st.temp, st.psi);
}

template <typename T>
void GenericOps<T>::inter_force(int model, hipStream_t s, const EosParams& e, DevState<T> st, double rho0) {
    MPLB_DISPATCH(model, (k_inter_force<L, T><<<grid_for(st.N), kBlock, 0, s>>>(st, e, T(rho0))));
}

template <typename T>
void GenericOps<T>::collide(int model, hipStream_t s, const EosParams& e, const ModelParams& mp, DevState<T> st,
                            T* f, double rho0) {
    MPLB_DISPATCH(model, (launch_collide<L, T>(s, e, mp, st, f, rho0)));
}

template <typename T>
void GenericOps<T>::stream(int model, hipStream_t s, int64_t N, const int* nb, const T* in, T* out) {
    const int Q = model == 0 ? 9 : 15;
    k_stream<T><<<grid_for(N), kBlock, 0, s>>>(N, Q, nb, in, out);
}

template <typename T>
void GenericOps<T>::boundary(int model, hipStream_t s, const EosParams& e, const ModelParams& mp, DevState<T> st,
                             T* f, const BdList& b, double rho0) {
    if (b.n <= 0) return;
    if (b.sequential) {
        MPLB_DISPATCH(model, (k_boundary_sequential<L, T><<<1, 64, 0, s>>>(st, e, mp, f, b, T(rho0))));
    } else {
        MPLB_DISPATCH(model, (k_boundary_parallel<L, T><<<grid_for(b.n), kBlock, 0, s>>>(st, e, mp, f, b, T(rho0))));
    }
}

template <typename T>
void GenericOps<T>::macro_density(int model, hipStream_t s, DevState<T> st, const T* f) {
    MPLB_DISPATCH(model, (k_density<L, T><<<grid_for(st.N), kBlock, 0, s>>>(st, f)));
}

template <typename T>
void GenericOps<T>::macro_velocity(int model, hipStream_t s, const EosParams& e, int fmodel, DevState<T> st,
                                   const T* f, double rho0) {
    MPLB_DISPATCH(model, (k_velocity<L, T><<<grid_for(st.N), kBlock, 0, s>>>(st, e, fmodel, f, T(rho0))));
}

template <typename T>
void GenericOps<T>::macro_temperature(int model, hipStream_t s, int gmodel, DevState<T> st, const T* g) {
    MPLB_DISPATCH(model, (k_temperature<L, T><<<grid_for(st.N), kBlock, 0, s>>>(st, gmodel, g)));
}

template <typename T>
void GenericOps<T>::heat_source(int model, hipStream_t s, const EosParams& e, const ModelParams& mp,
                                DevState<T> st, int first_call) {
    MPLB_DISPATCH(model, (k_heat_source<L, T><<<grid_for(st.N), kBlock, 0, s>>>(st, e, mp, first_call)));
}

template <typename T>
void GenericOps<T>::equilibrium(int model, hipStream_t s, const ModelParams& mp, DevState<T> st, T* f) {
    MPLB_DISPATCH(model, (k_equilibrium<L, T><<<grid_for(st.N), kBlock, 0, s>>>(st, mp, f)));
}

template <typename T>
void GenericOps<T>::pressure(int model, hipStream_t s, const EosParams& e, int fmodel, double sigma,
                             DevState<T> st) {
    MPLB_DISPATCH(model, (k_pressure<L, T><<<grid_for(st.N), kBlock, 0, s>>>(st, e, fmodel, T(sigma))));
}

template <typename T>
void GenericOps<T>::sum(hipStream_t s, const T* x, int64_t n, double* out_dev, int ordered, double* scratch) {
    if (ordered) {
        k_sum_ordered<T><<<1, kBlock, 0, s>>>(x, n, out_dev);
    } else {
        int blocks = (int)((n + kBlock - 1) / kBlock);
        if (blocks > 2048) blocks = 2048;
        if (blocks < 1) blocks = 1;
        k_sum_tree_partial<T><<<blocks, kBlock, 0, s>>>(x, n, scratch);
        k_sum_tree_final<<<1, kBlock, 0, s>>>(scratch, blocks, out_dev);
    }
}

template <typename T>
void GenericOps<T>::nan_count(hipStream_t s, const T* x, int64_t n, unsigned int* out_dev) {
    int blocks = (int)((n + kBlock - 1) / kBlock);
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    k_nan_count<T><<<blocks, kBlock, 0, s>>>(x, n, out_dev);
}

template struct GenericOps<double>;
template struct GenericOps<float>;

}  // namespace mplb
