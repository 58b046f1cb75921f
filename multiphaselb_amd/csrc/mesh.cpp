// Lattice topology of structured boxes.
//
// The neighbour of node x along link k is the UPSTREAM node x - c_k
// (latticeBox2D.c:174, latticeBox3D.c:215); on the box surface a candidate is kept
// only if it lies within one lattice unit in every coordinate.  Periodic directions
// are then closed by copying each surface node's missing entries from the node on the
// opposite face, axis by axis in the reference's pass order, so the two end planes
// are duplicates of one site and the period is N-1 (periodicXY.c, periodicX.c,
// periodicY.c, periodicXYZ3D.c, periodicXY3D.c, periodicZ3D.c, periodicXY3DSpot.c).
// The pass order decides the corner and edge entries, so it is kept as is; the
// boundary lists keep the reference's node order and sizes, including its indexing
// slips (Y1 of the generic 2D box and the periodicY copy use (nx-1)*ny; periodicZ3D's
// x-edge pass uses j*ny), which only coincide with the intended nodes on square boxes.
#include "mesh.h"

#include <cmath>
#include <cstdlib>

namespace mplb {

namespace {

const int kC2[9][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {-1, 0, 0}, {0, -1, 0}, {1, 1, 0}, {-1, 1, 0}, {-1, -1, 0}, {1, -1, 0}};
const int kR2[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
const int kC3[15][3] = {{0, 0, 0},  {1, 0, 0},  {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},  {0, 0, 1},   {0, 0, -1}, {1, 1, 1},
                        {-1, 1, 1}, {1, -1, 1}, {-1, -1, 1}, {1, 1, -1},  {-1, 1, -1}, {1, -1, -1}, {-1, -1, -1}};
const int kR3[15] = {0, 2, 1, 4, 3, 6, 5, 14, 13, 12, 11, 10, 9, 8, 7};

struct Box {
    Mesh& m;
    int32_t& at(int64_t id, int k) { return m.nb[id * m.Q + k]; }
    // mutual completion of two nodes' missing entries (in link order)
    void pair(int64_t a, int64_t b, int kfirst = 0) {
        for (int k = kfirst; k < m.Q; k++) {
            if (at(a, k) == -1) at(a, k) = at(b, k);
            if (at(b, k) == -1) at(b, k) = at(a, k);
        }
    }
    // four-corner completion of one plane (periodicXY.c:67-96)
    void corners(const int64_t cid[4]) {
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++)
                if (i != j)
                    for (int k = 1; k < m.Q; k++)
                        if (at(cid[i], k) == -1 && at(cid[j], k) != -1) at(cid[i], k) = at(cid[j], k);
    }
    std::vector<int32_t>& list(const char* name, size_t n) {
        m.bdnames.push_back(name);
        m.bd.emplace_back(n, 0);
        return m.bd.back();
    }
};

void neighbours_2d(Mesh& m) {
    const int nx = m.nx, ny = m.ny, Q = 9;
    auto coord_ok = [&](int64_t id, int64_t nid) {
        return std::abs(int(id % nx) - int(nid % nx)) <= 1 && std::abs(int(id / nx) - int(nid / nx)) <= 1;
    };
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            const int64_t id = i + (int64_t)j * nx;
            const bool surface = (j == 0 || j == ny - 1 || i == 0 || i == nx - 1);
            for (int v = 0; v < Q; v++) {
                const int* c = kC2[kR2[v]];
                const int64_t nid = id + c[0] + (int64_t)c[1] * nx;
                if (!surface) {
                    m.nb[id * Q + v] = (int32_t)nid;
                } else if (nid >= 0 && nid <= (int64_t)nx * ny - 1 && coord_ok(id, nid)) {
                    m.nb[id * Q + v] = (int32_t)nid;
                }
            }
        }
}

void neighbours_3d(Mesh& m) {
    const int nx = m.nx, ny = m.ny, nz = m.nz, Q = 15;
    const int64_t nxy = (int64_t)nx * ny;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const int64_t id = i + j * (int64_t)nx + k * nxy;
                const bool surface = i == 0 || i == nx - 1 || j == 0 || j == ny - 1 || k == 0 || k == nz - 1;
                for (int v = 0; v < Q; v++) {
                    const int* c = kC3[kR3[v]];
                    const int64_t nid = id + c[0] + (int64_t)c[1] * nx + (int64_t)c[2] * nxy;
                    if (!surface) {
                        m.nb[id * Q + v] = (int32_t)nid;
                        continue;
                    }
                    if (nid < 0 || nid > m.N - 1) continue;
                    const int qi = int(nid % nx), qj = int((nid / nx) % ny), qk = int(nid / nxy);
                    if (std::abs(i - qi) <= 1 && std::abs(j - qj) <= 1 && std::abs(k - qk) <= 1)
                        m.nb[id * Q + v] = (int32_t)nid;
                }
            }
}

}  // namespace

std::string build_box(Mesh& m, double spot_radius, const double spot_centre[2]) {
    m.d = m.lbmodel == 0 ? 2 : 3;
    m.Q = m.lbmodel == 0 ? 9 : 15;
    if (m.d == 2) m.nz = 1;
    if (m.nx < 3 || m.ny < 3 || (m.d == 3 && m.nz < 3)) return "Lattice box must have at least 3 nodes per axis";
    m.N = (int64_t)m.nx * m.ny * m.nz;
    if (m.N > (int64_t)INT32_MAX) return "Lattice too large for 32-bit node indices";
    m.nb.assign(m.N * m.Q, -1);
    m.bd.clear();
    m.bdnames.clear();
    m.bd.reserve(8);   // Box::list hands out references into this vector
    Box b{m};
    const int nx = m.nx, ny = m.ny, nz = m.nz;
    const int64_t nxy = (int64_t)nx * ny;
    if (m.d == 2) {
        neighbours_2d(m);
        switch (m.btype) {
        case BOX_GENERIC: {                                  // genericBoundary.c:6-49
            auto& x0 = b.list("X0", ny);
            auto& x1 = b.list("X1", ny);
            auto& y0 = b.list("Y0", nx - 2);
            auto& y1 = b.list("Y1", nx - 2);
            for (int j = 0; j < ny; j++) { x0[j] = j * nx; x1[j] = (nx - 1) + j * nx; }
            for (int j = 1; j < nx - 1; j++) { y0[j - 1] = j; y1[j - 1] = j + (nx - 1) * ny; }
            break;
        }
        case BOX_PERIODIC_X: {                               // periodicX.c:6-73
            auto& x0 = b.list("X0", ny - 2);
            auto& x1 = b.list("X1", ny - 2);
            auto& y0 = b.list("Y0", nx);
            auto& y1 = b.list("Y1", nx);
            for (int j = 0; j < ny; j++) {
                if (j >= 1 && j <= ny - 2) { x0[j - 1] = j * nx; x1[j - 1] = (nx - 1) + j * nx; }
                b.pair((int64_t)j * nx, nx - 1 + (int64_t)j * nx);
            }
            for (int j = 0; j < nx; j++) { y0[j] = j; y1[j] = j + (ny - 1) * nx; }
            break;
        }
        case BOX_PERIODIC_Y: {                               // periodicY.c:6-71
            auto& x0 = b.list("X0", ny - 2);
            auto& x1 = b.list("X1", ny - 2);
            auto& y0 = b.list("Y0", nx);
            auto& y1 = b.list("Y1", nx);
            for (int j = 1; j < ny - 1; j++) { x0[j - 1] = j * nx; x1[j - 1] = (nx - 1) + j * nx; }
            for (int j = 0; j < nx; j++) {
                y0[j] = j;
                y1[j] = j + (ny - 1) * nx;
                const int64_t other = j + (int64_t)(nx - 1) * ny;
                if (other >= m.N) return "periodicY needs Nx == Ny";
                b.pair(j, other);
            }
            break;
        }
        case BOX_PERIODIC_XY: {                              // periodicXY.c:6-101
            for (int j = 1; j < ny - 1; j++) b.pair((int64_t)j * nx, nx - 1 + (int64_t)j * nx);
            for (int j = 1; j < nx - 1; j++) b.pair(j, j + (int64_t)(ny - 1) * nx);
            const int64_t cid[4] = {0, nx - 1, (int64_t)nx * (ny - 1), nxy - 1};
            b.corners(cid);
            break;
        }
        default:
            return "Unrecognized boundary type";
        }
    } else {
        neighbours_3d(m);
        switch (m.btype) {
        case BOX_GENERIC: {                                  // genericBoundary3D.c:6-93
            auto& x0 = b.list("X0", (size_t)ny * nz);
            auto& x1 = b.list("X1", (size_t)ny * nz);
            auto& y0 = b.list("Y0", (size_t)(nx - 2) * nz);
            auto& y1 = b.list("Y1", (size_t)(nx - 2) * nz);
            auto& z0 = b.list("Z0", (size_t)(nx - 2) * (ny - 2));
            auto& z1 = b.list("Z1", (size_t)(nx - 2) * (ny - 2));
            size_t id = 0;
            for (int k = 0; k < nz; k++)
                for (int j = 0; j < ny; j++, id++) {
                    x0[id] = int32_t(j * nx + k * nxy);
                    x1[id] = int32_t((nx - 1) + j * nx + k * nxy);
                }
            id = 0;
            for (int k = 0; k < nz; k++)
                for (int i = 1; i < nx - 1; i++, id++) {
                    y0[id] = int32_t(i + k * nxy);
                    y1[id] = int32_t(i + (ny - 1) * nx + k * nxy);
                }
            id = 0;
            for (int j = 1; j < ny - 1; j++)
                for (int i = 1; i < nx - 1; i++, id++) {
                    z0[id] = int32_t(i + j * nx);
                    z1[id] = int32_t(i + j * nx + (nz - 1) * nxy);
                }
            break;
        }
        case BOX_PERIODIC_XY:                               // periodicXY3D.c:6-226
        case BOX_PERIODIC_XY_SPOT: {                        // periodicXY3DSpot.c:7-313
            if (m.btype == BOX_PERIODIC_XY) {
                auto& z0 = b.list("Z0", (size_t)nxy);
                auto& z1 = b.list("Z1", (size_t)nxy);
                size_t k = 0;
                for (int j = 0; j < ny; j++)
                    for (int i = 0; i < nx; i++, k++) {
                        z0[k] = int32_t(i + j * nx);
                        z1[k] = int32_t(i + j * nx + (nz - 1) * nxy);
                    }
            } else {
                size_t count = 0;
                const double rr = spot_radius * spot_radius;
                auto inside = [&](int i, int j) {
                    const int x0 = (int)(spot_centre[0] - (int)i), y0 = (int)(spot_centre[1] - (int)j);
                    return (double)(x0 * x0 + y0 * y0) <= rr;
                };
                for (int i = 0; i < nx; i++)
                    for (int j = 0; j < ny; j++)
                        if (inside(i, j)) count++;
                auto& z0 = b.list("Z0", (size_t)nxy - count);
                auto& z1 = b.list("Z1", (size_t)nxy);
                auto& sp = b.list("Spot", count);
                size_t k = 0, k0 = 0, ks = 0;
                for (int j = 0; j < ny; j++)
                    for (int i = 0; i < nx; i++, k++) {
                        z1[k] = int32_t(i + j * nx + (nz - 1) * nxy);
                        if (inside(i, j)) sp[ks++] = i + j * nx;
                        else z0[k0++] = i + j * nx;
                    }
            }
            for (int k = 1; k < nz - 1; k++)
                for (int j = 0; j < ny; j++) b.pair(j * nx + k * nxy, j * nx + k * nxy + nx - 1);
            for (int k = 1; k < nz - 1; k++)
                for (int i = 0; i < nx; i++) b.pair(i + k * nxy, i + k * nxy + (int64_t)(ny - 1) * nx);
            const int64_t c0[4] = {0, nx - 1, (int64_t)nx * (ny - 1), nxy - 1};
            b.corners(c0);
            const int64_t top = (nz - 1) * nxy;
            const int64_t c1[4] = {top, nx - 1 + top, (int64_t)nx * (ny - 1) + top, nxy - 1 + top};
            b.corners(c1);
            break;
        }
        case BOX_PERIODIC_Z: {                              // periodicZ3D.c:6-251
            if (nz != nx) return "periodicZ boundary lists are only defined for Nz == Nx";
            // periodicZ3D.c:86-93 fills nz*(nx-2) of the (nx-2)*ny Y-list entries (the rest
            // stay unset for ny > nz) and :221-236 completes node j*ny + (nz-1)*nx*ny, which
            // leaves the lattice for ny >= nx + 2
            if (ny > nz) return "periodicZ boundary lists are only defined for Ny <= Nz";
            auto& x0 = b.list("X0", (size_t)nxy);
            auto& x1 = b.list("X1", (size_t)nxy);
            auto& y0 = b.list("Y0", (size_t)(nx - 2) * ny);
            auto& y1 = b.list("Y1", (size_t)(nx - 2) * ny);
            size_t i = 0;
            for (int k = 0; k < nz; k++)
                for (int j = 0; j < ny; j++, i++) {
                    x0[i] = int32_t(j * nx + k * nxy);
                    x1[i] = int32_t((nx - 1) + j * nx + k * nxy);
                }
            size_t jj = 0;
            for (int k = 0; k < nz; k++)
                for (int ii = 1; ii < nx - 1; ii++, jj++)
                    if (jj < y0.size()) {
                        y0[jj] = int32_t(ii + k * nxy);
                        y1[jj] = int32_t(ii + (ny - 1) * nx + k * nxy);
                    }
            const int64_t top = (nz - 1) * nxy;
            for (int j = 1; j < ny - 1; j++)
                for (int ii = 1; ii < nx - 1; ii++) b.pair(ii + j * nx, ii + j * nx + top);
            for (int ii = 0; ii < nx; ii++) {
                b.pair(ii, ii + top);
                b.pair(ii + (ny - 1) * nx, ii + (ny - 1) * nx + top);
            }
            for (int j = 1; j < ny - 1; j++) {
                b.pair((int64_t)j * ny, (int64_t)j * ny + top);
                b.pair((nx - 1) + (int64_t)j * nx, (nx - 1) + (int64_t)j * nx + top);
            }
            break;
        }
        case BOX_PERIODIC_XYZ: {                            // periodicXYZ3D.c:6-159
            for (int k = 0; k < nz; k++)
                for (int j = 0; j < ny; j++) b.pair(j * nx + k * nxy, j * nx + k * nxy + nx - 1);
            for (int k = 0; k < nz; k++)
                for (int i = 0; i < nx; i++) b.pair(i + k * nxy, i + k * nxy + (int64_t)(ny - 1) * nx);
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) b.pair(i + (int64_t)j * nx, i + (int64_t)j * nx + (nz - 1) * nxy);
            break;
        }
        default:
            return "Unrecognized boundary type";
        }
    }
    m.onbnd.assign(m.N, 0);
    for (int64_t id = 0; id < m.N; id++)
        for (int k = 0; k < m.Q; k++)
            if (m.nb[id * m.Q + k] == -1) {
                m.onbnd[id] = 1;
                break;
            }
    return "";
}

}  // namespace mplb
