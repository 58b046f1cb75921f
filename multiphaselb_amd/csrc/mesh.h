// Structured-box lattice tables with the reference's conventions.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace mplb {

enum BoxType : int { BOX_GENERIC = 0, BOX_PERIODIC_X = 1, BOX_PERIODIC_Y = 2, BOX_PERIODIC_XY = 3,
                     BOX_PERIODIC_Z = 4, BOX_PERIODIC_XYZ = 5, BOX_PERIODIC_XY_SPOT = 6 };

struct Mesh {
    int lbmodel = 0;      // 0 D2Q9, 1 D3Q15
    int Q = 9, d = 2;
    int nx = 0, ny = 0, nz = 1;
    int64_t N = 0;
    int btype = BOX_GENERIC;
    std::vector<int32_t> nb;                 // [N][Q] node-major, upstream node or -1
    std::vector<uint8_t> onbnd;              // isOnBnd
    std::vector<std::string> bdnames;
    std::vector<std::vector<int32_t>> bd;    // boundary lists in file order
    std::vector<int32_t> points;             // [N][3] when read from lattice files (else implied)
    bool structured = true;                  // node id = i + j nx (+ k nx ny)
};

// Builds the neighbour table and boundary lists of an nx x ny (x nz) box the way
// latticeBox2D/latticeBox3D followed by a one-rank latticeMeshPartition do.  Returns
// an empty string on success or the reference's error text.
std::string build_box(Mesh& m, double spot_radius, const double spot_centre[2]);

// processor0/lattice/* of a one-rank case (fieldio.hip)
bool lattice_files_present(const std::string& dir);
std::string read_lattice(Mesh& m, const std::string& dir);

// The wrap-around index the fused kernels use for a periodic axis with duplicate end
// nodes: coordinate u in [-1, n] maps to u in [0, n-1] with period n-1.
inline int wrap_dup(int u, int n) { return u < 0 ? u + (n - 1) : (u > n - 1 ? u - (n - 1) : u); }

}  // namespace mplb
