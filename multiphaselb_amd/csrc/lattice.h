// Lattice descriptors D2Q9 and D3Q15, usable from host and device code.
//
// The velocity sets, reverse indices, weights and pseudopotential weights are the
// reference's (latticeModel/D2Q9/D2Q9_{Vel,Reverse,omega,ppw,cs2}.c,
// latticeModel/D3Q15/D3Q15_*.c).  The MRT matrices are not tabulated: M is built here
// from polynomial moments of the velocities, and because its rows are mutually
// orthogonal the inverse is M^T diag(1/|row|^2).  Each inverse entry is therefore a
// single correctly rounded division of two small integers, which reproduces the
// reference's decimal/rational constants (D2Q9_MRTInv.c:20-115, D3Q15_MRTInv.c:21-259)
// bit for bit; tests compare both matrices with the reference's compiled tables.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define MPLB_HD __host__ __device__
#else
#define MPLB_HD
#endif

namespace mplb {

template <int Q_>
struct Tables {
    int c[Q_][3];
    int rev[Q_];
    double omega[Q_];
    double w[Q_];
    int M[Q_][Q_];
    double invM[Q_][Q_];
};

// D2Q9 moment basis (rho, e, eps, jx, qx, jy, qy, pxx, pxy).
constexpr int d2q9_moment(int row, int cx, int cy) {
    const int c2 = cx * cx + cy * cy;
    switch (row) {
    case 0: return 1;
    case 1: return 3 * c2 - 4;
    case 2: return (9 * c2 * c2 - 21 * c2 + 8) / 2;
    case 3: return cx;
    case 4: return (3 * c2 - 5) * cx;
    case 5: return cy;
    case 6: return (3 * c2 - 5) * cy;
    case 7: return cx * cx - cy * cy;
    default: return cx * cy;
    }
}

// D3Q15 moment basis (rho, e, eps, jx, qx, jy, qy, jz, qz, 3pxx, pww, pxy, pyz, pxz, mxyz).
constexpr int d3q15_moment(int row, int cx, int cy, int cz) {
    const int c2 = cx * cx + cy * cy + cz * cz;
    switch (row) {
    case 0: return 1;
    case 1: return c2 - 2;
    case 2: return (15 * c2 * c2 - 55 * c2 + 32) / 2;
    case 3: return cx;
    case 4: return cx * (5 * c2 - 13) / 2;
    case 5: return cy;
    case 6: return cy * (5 * c2 - 13) / 2;
    case 7: return cz;
    case 8: return cz * (5 * c2 - 13) / 2;
    case 9: return 3 * cx * cx - c2;
    case 10: return cy * cy - cz * cz;
    case 11: return cx * cy;
    case 12: return cy * cz;
    case 13: return cx * cz;
    default: return cx * cy * cz;
    }
}

template <int Q_>
constexpr void finish_inverse(Tables<Q_>& t) {
    for (int i = 0; i < Q_; i++) {
        int n2 = 0;
        for (int k = 0; k < Q_; k++) n2 += t.M[i][k] * t.M[i][k];
        for (int k = 0; k < Q_; k++) t.invM[k][i] = double(t.M[i][k]) / double(n2);
    }
}

constexpr Tables<9> make_d2q9() {
    Tables<9> t{};
    const int c[9][2] = {{0, 0}, {1, 0}, {0, 1}, {-1, 0}, {0, -1}, {1, 1}, {-1, 1}, {-1, -1}, {1, -1}};
    for (int k = 0; k < 9; k++) {
        t.c[k][0] = c[k][0];
        t.c[k][1] = c[k][1];
        t.c[k][2] = 0;
    }
    for (int k = 0; k < 9; k++)
        for (int r = 0; r < 9; r++)
            if (t.c[r][0] == -t.c[k][0] && t.c[r][1] == -t.c[k][1]) t.rev[k] = r;
    for (int k = 0; k < 9; k++) {
        const int c2 = c[k][0] * c[k][0] + c[k][1] * c[k][1];
        t.omega[k] = c2 == 0 ? 4.0 / 9.0 : (c2 == 1 ? 1.0 / 9.0 : 1.0 / 36.0);
        t.w[k] = c2 == 0 ? 0.0 : (c2 == 1 ? 1.0 / 3 : 1.0 / 12);
    }
    for (int r = 0; r < 9; r++)
        for (int k = 0; k < 9; k++) t.M[r][k] = d2q9_moment(r, c[k][0], c[k][1]);
    finish_inverse(t);
    return t;
}

constexpr Tables<15> make_d3q15() {
    Tables<15> t{};
    // rest, the six axis links (+x,-x,+y,-y,+z,-z), then the eight corners with x
    // varying fastest and z = +1 first.
    const int c[15][3] = {{0, 0, 0},   {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},  {0, -1, 0},
                          {0, 0, 1},   {0, 0, -1},  {1, 1, 1},   {-1, 1, 1}, {1, -1, 1},
                          {-1, -1, 1}, {1, 1, -1},  {-1, 1, -1}, {1, -1, -1}, {-1, -1, -1}};
    for (int k = 0; k < 15; k++)
        for (int j = 0; j < 3; j++) t.c[k][j] = c[k][j];
    for (int k = 0; k < 15; k++)
        for (int r = 0; r < 15; r++)
            if (c[r][0] == -c[k][0] && c[r][1] == -c[k][1] && c[r][2] == -c[k][2]) t.rev[k] = r;
    for (int k = 0; k < 15; k++) {
        const int c2 = c[k][0] * c[k][0] + c[k][1] * c[k][1] + c[k][2] * c[k][2];
        t.omega[k] = c2 == 0 ? 2.0 / 9.0 : (c2 == 1 ? 1.0 / 9.0 : 1.0 / 72.0);
        t.w[k] = c2 == 0 ? 0.0 : (c2 == 1 ? 1.0 / 3.0 : 1.0 / 24.0);
    }
    for (int r = 0; r < 15; r++)
        for (int k = 0; k < 15; k++) t.M[r][k] = d3q15_moment(r, c[k][0], c[k][1], c[k][2]);
    finish_inverse(t);
    return t;
}

struct D2Q9 {
    static constexpr int Q = 9;
    static constexpr int D = 2;
    static constexpr int model = 0;
    static constexpr Tables<9> t = make_d2q9();
    // axis links used by the finite-difference stencils (finiteDifference/*.c):
    // derivative along axis a is 0.5 (q[nb[fd_plus[a]]] - q[nb[fd_minus[a]]]).
    static constexpr int fd_plus[3] = {3, 4, 0};
    static constexpr int fd_minus[3] = {1, 2, 0};
    static constexpr int naxis = 4;   // bounceBack tangential-average links 1..4
};

struct D3Q15 {
    static constexpr int Q = 15;
    static constexpr int D = 3;
    static constexpr int model = 1;
    static constexpr Tables<15> t = make_d3q15();
    static constexpr int fd_plus[3] = {2, 4, 6};
    static constexpr int fd_minus[3] = {1, 3, 5};
    static constexpr int naxis = 6;
};

constexpr double kCs2 = 1.0 / 3.0;   // D2Q9_cs2.c / D3Q15_cs2.c

}  // namespace mplb
