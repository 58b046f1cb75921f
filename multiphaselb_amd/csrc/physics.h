// Per-node physics shared by the HIP kernels: equation of state, pseudopotential,
// equilibria and the MRT moment transforms.  Every expression keeps the reference's
// association order (plain IEEE arithmetic; the FP64 translation unit is built with
// -ffp-contract=off so no multiply-add is fused), which is what makes the FP64 mode
// reproduce the reference results bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include "lattice.h"

namespace mplb {

enum EosType : int { EOS_IDEAL = 0, EOS_CARSTAR = 1, EOS_PENGROB = 2, EOS_VDW = 3 };

// Scalars of the EOS block (EOS/EOSInfo.h:80-102).
struct EosParams {
    int type;
    double a, b, w_pr, R, G, M, Cv;
    double g[3], fext[3];
};

// EOS/p_eos.c:7-64
template <typename T>
__device__ __host__ inline T p_eos(const EosParams& e, T rho, T temp) {
    T p = 0, alpha, beta;
    const T R = T(e.R), a = T(e.a), b = T(e.b), Mm = T(e.M);
    switch (e.type) {
    case EOS_IDEAL:
        p = rho * R * temp;
        break;
    case EOS_CARSTAR:
        alpha = b * rho / T(4.0);
        beta = T(1) - alpha;
        p = rho * R * temp * ((T(1) + alpha + alpha * alpha - alpha * alpha * alpha) / (beta * beta * beta)) -
            a * rho * rho;
        break;
    case EOS_PENGROB: {
        const T w = T(e.w_pr);
        alpha = T(0.170151343) * a / (b * R);
        beta = (T(1) + (T(0.37464) + T(1.5422) * w - T(0.26992) * w * w) * (T(1) - sqrt(temp / alpha)));
        p = rho * R * temp / (T(1) - b * rho) -
            a * beta * beta * rho * rho / (T(1) + T(2) * b * rho - b * rho * b * rho);
        break;
    }
    case EOS_VDW:
        p = rho * R * temp / (Mm - rho * b) - a * rho * rho / (Mm * Mm);
        break;
    default:
        break;
    }
    return p;
}

// lbEquation/pseudoPot/potential.c:7-13
template <typename T>
__device__ __host__ inline T potential(const EosParams& e, T rho, T temp) {
    return sqrt(T(2) * (p_eos<T>(e, rho, temp) - rho * T(kCs2)) / T(e.G));
}

// res = M v, j ascending from a +0 start; zero coefficients are skipped (adding the
// reference's +-0 products never changes a running sum) and the power-of-two
// coefficients multiply exactly, so the rounding sequence equals the reference's dense
// matVecMult (basic/matVecMult.c:4-20).
template <class L, typename T>
__device__ __host__ inline void moments(const T* v, T* res) {
#pragma unroll
    for (int i = 0; i < L::Q; i++) {
        T r = T(0);
#pragma unroll
        for (int j = 0; j < L::Q; j++) {
            const int m = L::t.M[i][j];
            if (m == 0) continue;
            r += (m == 1) ? v[j] : (m == -1 ? -v[j] : T(m) * v[j]);
        }
        res[i] = r;
    }
}

// res = invM v, j ascending, zero coefficients skipped.
template <class L, typename T>
__device__ __host__ inline void populations(const T* v, T* res) {
#pragma unroll
    for (int i = 0; i < L::Q; i++) {
        T r = T(0);
#pragma unroll
        for (int j = 0; j < L::Q; j++) {
            const double m = L::t.invM[i][j];
            if (m == 0.0) continue;
            r += T(m) * v[j];
        }
        res[i] = r;
    }
}

// liMRTEquilibrium.c:3-29 / xuMRTEquilibrium.c:3-29, population k only.
template <class L, typename T>
__device__ __host__ inline T ns_feq(int k, T rho, const T v[3]) {
    T alpha = 0, beta = 0;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int c = L::t.c[k][j];
        if (c == 1) alpha += v[j];
        else if (c == -1) alpha += -v[j];
        beta += v[j] * v[j];
    }
    const T cs2 = T(kCs2);
    return rho * T(L::t.omega[k]) *
           (T(1) + alpha / cs2 + T(0.5) * alpha * alpha / (cs2 * cs2) - T(0.5) * beta / cs2);
}

// myMRTEquilibrium.c:5-95: g_eq = invM n_eq.
template <class L, typename T>
__device__ __host__ inline void my_geq(T temp, const T U[3], T a1, T a2, T* geq) {
    T n[L::Q];
#pragma unroll
    for (int k = 0; k < L::Q; k++) n[k] = T(0);
    n[0] = temp;
    n[1] = a1 * temp;
    n[2] = a2 * temp;
    if (L::D == 2) {
        n[3] = temp * U[0];
        n[4] = temp * (-U[0]);
        n[5] = temp * U[1];
        n[6] = temp * (-U[1]);
    } else {
        n[3] = temp * U[0];
        n[4] = -(temp * U[0]);
        n[5] = temp * U[1];
        n[6] = -(temp * U[1]);
        n[7] = temp * U[2];
        n[8] = -(temp * U[2]);
    }
    populations<L, T>(n, geq);
}

// liTempMRTEquilibrium.c:5-44 (D2Q9 only).
template <class L, typename T>
__device__ __host__ inline void litemp_geq(T temp, const T U[3], T* geq) {
    T n[L::Q];
#pragma unroll
    for (int k = 0; k < L::Q; k++) n[k] = T(0);
    n[0] = temp;
    n[1] = T(-2.0) * temp;
    n[2] = T(2.0) * temp;
    n[3] = temp * U[0];
    n[4] = temp * (-U[0]);
    n[5] = temp * U[1];
    n[6] = temp * (-U[1]);
    populations<L, T>(n, geq);
}

// ---------------------------------------------------------------------------
// Inversion-symmetric forms for the reduced-precision kernels.  Summing populations
// in index order is not an odd function of the state under x -> -x (which swaps f_q and
// f_rev(q)): fl(fl(a - b) + c) and fl(fl(b - a) - c) need not be exact negatives, so a
// symmetric droplet picks up a bias of a fraction of an ulp per node per step and its
// total momentum drifts linearly in FP32 (in FP64 the same bias is ~1e-16 and harmless).
// Here every sum is formed from pair sums s = f_q + f_rev and differences
// d = f_q - f_rev: even moments from s, odd moments from d, and populations as
// f_q = E + O, f_rev = E - O, so the arithmetic is exactly equivariant under inversion
// and the rounding errors of mirror nodes cancel instead of adding up.
template <class L>
struct Pairs {
    static constexpr int n = (L::Q - 1) / 2;
    int q[n], r[n];
};

template <class L>
constexpr Pairs<L> make_pairs() {
    Pairs<L> p{};
    int k = 0;
    for (int q = 1; q < L::Q; q++)
        if (q < L::t.rev[q]) {
            p.q[k] = q;
            p.r[k] = L::t.rev[q];
            k++;
        }
    return p;
}

// +1: moment row j is even under inversion (M[j][q] == M[j][rev q]), -1: odd
template <class L>
constexpr int moment_parity(int j) {
    const Pairs<L> P = make_pairs<L>();
    for (int k = 0; k < Pairs<L>::n; k++) {
        const int a = L::t.M[j][P.q[k]];
        if (a != 0) return a == L::t.M[j][P.r[k]] ? 1 : -1;
    }
    return 1;
}

// every row is either even or odd, and odd rows do not see the rest population
template <class L>
constexpr bool parity_consistent() {
    const Pairs<L> P = make_pairs<L>();
    for (int j = 0; j < L::Q; j++) {
        const int s = moment_parity<L>(j);
        if (s == -1 && L::t.M[j][0] != 0) return false;
        for (int k = 0; k < Pairs<L>::n; k++)
            if (L::t.M[j][P.q[k]] != s * L::t.M[j][P.r[k]]) return false;
    }
    return true;
}

template <typename T>
__device__ __host__ inline T imul(int c, T x) {
    return c == 1 ? x : (c == -1 ? -x : T(c) * x);
}

// rho = f_0 + sum of pair sums
template <class L, typename T>
__device__ __host__ inline T rho_sym(const T* v) {
    constexpr Pairs<L> P = make_pairs<L>();
    T r = v[0];
#pragma unroll
    for (int k = 0; k < Pairs<L>::n; k++) r += v[P.q[k]] + v[P.r[k]];
    return r;
}

// m = M v from pair sums (even rows) and pair differences (odd rows)
template <class L, typename T>
__device__ __host__ inline void moments_sym(const T* v, T* m) {
    constexpr Pairs<L> P = make_pairs<L>();
    constexpr int NP = Pairs<L>::n;
    T sp[NP], dp[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) {
        sp[k] = v[P.q[k]] + v[P.r[k]];
        dp[k] = v[P.q[k]] - v[P.r[k]];
    }
#pragma unroll
    for (int j = 0; j < L::Q; j++) {
        T r = T(0);
        if (moment_parity<L>(j) == 1) {
            if (L::t.M[j][0] != 0) r = imul<T>(L::t.M[j][0], v[0]);
#pragma unroll
            for (int k = 0; k < NP; k++)
                if (L::t.M[j][P.q[k]] != 0) r += imul<T>(L::t.M[j][P.q[k]], sp[k]);
        } else {
#pragma unroll
            for (int k = 0; k < NP; k++)
                if (L::t.M[j][P.q[k]] != 0) r += imul<T>(L::t.M[j][P.q[k]], dp[k]);
        }
        m[j] = r;
    }
}

// v = M^-1 m as even part E and odd part O per pair: f_q = E + O, f_rev = E - O
template <class L, typename T>
__device__ __host__ inline void populations_sym(const T* m, T* v) {
    constexpr Pairs<L> P = make_pairs<L>();
    T r0 = T(0);
#pragma unroll
    for (int j = 0; j < L::Q; j++)
        if (moment_parity<L>(j) == 1 && L::t.invM[0][j] != 0.0) r0 += T(L::t.invM[0][j]) * m[j];
    v[0] = r0;
#pragma unroll
    for (int k = 0; k < Pairs<L>::n; k++) {
        const int q = P.q[k];
        T E = T(0), O = T(0);
#pragma unroll
        for (int j = 0; j < L::Q; j++) {
            const double c = L::t.invM[q][j];
            if (c == 0.0) continue;
            if (moment_parity<L>(j) == 1) E += T(c) * m[j];
            else O += T(c) * m[j];
        }
        v[q] = E + O;
        v[P.r[k]] = E - O;
    }
}

// v += M^-1 dm in the same even/odd form.  Collisions add the moment change to the
// pulled populations instead of rebuilding them from the post-collision moments: a
// rebuild rounds every population afresh each step (about an ulp of f, with a bias that
// does not average out near equilibrium, so the total mass drifts by ~5e-8 per step in
// FP32), while f + df rounds only the small change and leaves f exactly where df is
// below half an ulp.  The rest row of dm is zero for the conserved moments, so the mass
// changes only by the rounding of the additions.
template <class L, typename T>
__device__ __host__ inline void populations_add_sym(const T* dm, T* v) {
    constexpr Pairs<L> P = make_pairs<L>();
    T r0 = T(0);
#pragma unroll
    for (int j = 0; j < L::Q; j++)
        if (moment_parity<L>(j) == 1 && L::t.invM[0][j] != 0.0) r0 += T(L::t.invM[0][j]) * dm[j];
    v[0] += r0;
#pragma unroll
    for (int k = 0; k < Pairs<L>::n; k++) {
        const int q = P.q[k];
        T E = T(0), O = T(0);
#pragma unroll
        for (int j = 0; j < L::Q; j++) {
            const double c = L::t.invM[q][j];
            if (c == 0.0) continue;
            if (moment_parity<L>(j) == 1) E += T(c) * dm[j];
            else O += T(c) * dm[j];
        }
        v[q] += E + O;
        v[P.r[k]] += E - O;
    }
}

static_assert(parity_consistent<D2Q9>() && parity_consistent<D3Q15>(), "moment rows must be even or odd");

}  // namespace mplb
