// Files in the reference's on-disk formats (host code):
//   * lattice tables of one rank, processor0/lattice/{points, ghosts, neighbours,
//     recvGhosts, sendGhosts, boundary, vtkCells} (IO/writeLatticeMesh.c:20-226 writes
//     them, IO/readLatticeMesh.c reads them);
//   * macroscopic and population fields of a time level, processor0/<time>/<name>, in
//     the asciiRaw (one node per line, "%g" values) and binaryRaw (node-major float32)
//     formats (IO/asciiRaw/*, IO/binaryRaw/*); values read back pass through float as
//     in readScalarFromAsciiRaw.c / readArrayFromAsciiRaw.c.
#include <hip/hip_runtime.h>

#include <sys/stat.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mplb.h"
#include "case.h"
#include "mesh.h"

using namespace mplb;

namespace {

int ffail(mplb_case* c, int code, const std::string& m) {
    if (c) c->err = m;
    return code;
}

bool mkdirs(const std::string& path) {
    std::string cur;
    size_t i = 0;
    while (i <= path.size()) {
        const size_t j = path.find('/', i);
        cur = path.substr(0, j == std::string::npos ? path.size() : j);
        if (!cur.empty() && mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST) return false;
        if (j == std::string::npos) break;
        i = j + 1;
    }
    return true;
}

struct File {
    FILE* f = nullptr;
    File(const std::string& p, const char* mode) : f(std::fopen(p.c_str(), mode)) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

// node coordinates of a structured box (latticeBox2D.c:110-130 / latticeBox3D.c:115-130)
void coords(const Mesh& m, int64_t id, int* p) {
    const int64_t nxy = (int64_t)m.nx * m.ny;
    p[2] = m.d == 3 ? (int)(id / nxy) : 0;
    const int64_t r = id - (int64_t)p[2] * nxy;
    p[1] = (int)(r / m.nx);
    p[0] = (int)(r - (int64_t)p[1] * m.nx);
}

}  // namespace

namespace mplb {

bool lattice_files_present(const std::string& dir) {
    struct stat st;
    return stat((dir + "/processor0/lattice/neighbours").c_str(), &st) == 0;
}

// readLatticeMesh.c: points, neighbours (header nlocal d Q), boundary lists in file
// order, isOnBnd = any missing link.  One rank: no ghosts.
std::string read_lattice(Mesh& m, const std::string& dir) {
    const std::string base = dir + "/processor0/lattice/";
    int64_t n = 0;
    {
        File f(base + "points", "r");
        if (!f.f) return "Unable to open processor0/lattice/points";
        long long nl = 0;
        if (std::fscanf(f.f, "%lld", &nl) != 1 || nl <= 0) return "Ill defined processor0/lattice/points";
        n = nl;
        m.points.resize(3 * n);
        int mx[3] = {0, 0, 0};
        for (int64_t i = 0; i < n; i++)
            for (int d = 0; d < 3; d++) {
                if (std::fscanf(f.f, "%d", &m.points[3 * i + d]) != 1) return "Ill defined processor0/lattice/points";
                mx[d] = std::max(mx[d], m.points[3 * i + d]);
            }
        m.nx = mx[0] + 1;
        m.ny = mx[1] + 1;
        m.nz = m.d == 3 ? mx[2] + 1 : 1;
    }
    {
        File f(base + "ghosts", "r");
        long long ng = 0;
        if (f.f && std::fscanf(f.f, "%lld", &ng) == 1 && ng != 0)
            return "processor0/lattice has ghost nodes: the case was partitioned for more than one rank";
    }
    m.N = n;
    m.nb.assign(n * m.Q, -1);
    {
        File f(base + "neighbours", "r");
        if (!f.f) return "Unable to open processor0/lattice/neighbours";
        long long hdr[3];
        if (std::fscanf(f.f, "%lld %lld %lld", &hdr[0], &hdr[1], &hdr[2]) != 3 || hdr[0] != n || hdr[2] != m.Q)
            return "processor0/lattice/neighbours does not match the points and the lattice model";
        for (int64_t i = 0; i < n * m.Q; i++)
            if (std::fscanf(f.f, "%d", &m.nb[i]) != 1) return "Ill defined processor0/lattice/neighbours";
        for (int32_t v : m.nb)
            if (v < -1 || v >= n) return "processor0/lattice/neighbours points outside the lattice";
    }
    m.onbnd.assign(n, 0);
    for (int64_t i = 0; i < n; i++)
        for (int q = 0; q < m.Q; q++)
            if (m.nb[i * m.Q + q] == -1) m.onbnd[i] = 1;
    m.bd.clear();
    m.bdnames.clear();
    {
        File f(base + "boundary", "r");
        if (!f.f) return "Unable to open processor0/lattice/boundary";
        int nbd = 0;
        if (std::fscanf(f.f, "%d", &nbd) != 1 || nbd < 0) return "Ill defined processor0/lattice/boundary";
        for (int b = 0; b < nbd; b++) {
            char name[256];
            long long cnt = 0;
            if (std::fscanf(f.f, "%255s %lld", name, &cnt) != 2 || cnt < 0) return "Ill defined processor0/lattice/boundary";
            m.bdnames.push_back(name);
            std::vector<int32_t> l(cnt);
            for (long long j = 0; j < cnt; j++)
                if (std::fscanf(f.f, "%d", &l[j]) != 1 || l[j] < 0 || l[j] >= n)
                    return "Ill defined processor0/lattice/boundary";
            m.bd.push_back(std::move(l));
        }
    }
    // structured order: node id = i + j nx (+ k nx ny)
    m.structured = (int64_t)m.nx * m.ny * m.nz == n;
    for (int64_t i = 0; i < n && m.structured; i++) {
        int p[3];
        coords(m, i, p);
        m.structured = p[0] == m.points[3 * i] && p[1] == m.points[3 * i + 1] && p[2] == m.points[3 * i + 2];
    }
    m.btype = -1;
    return "";
}

}  // namespace mplb

extern "C" {

int mplb_write_lattice(const mplb_case* cc, const char* dir) {
    mplb_case* c = const_cast<mplb_case*>(cc);
    if (!c || !dir) return MPLB_E_ARG;
    const Mesh& m = c->mesh;
    const std::string base = std::string(dir) + "/processor0/lattice/";
    if (!mkdirs(base)) return ffail(c, MPLB_E_IO, "Unable to create " + base);
    const int64_t N = m.N;
    {
        File f(base + "points", "w");
        if (!f.f) return ffail(c, MPLB_E_IO, "Unable to write points");
        std::fprintf(f.f, "%lld\n", (long long)N);
        for (int64_t i = 0; i < N; i++) {
            int p[3];
            if (!m.points.empty()) {
                p[0] = m.points[3 * i];
                p[1] = m.points[3 * i + 1];
                p[2] = m.points[3 * i + 2];
            } else {
                coords(m, i, p);
            }
            std::fprintf(f.f, "%d %d %d\n", p[0], p[1], p[2]);
        }
    }
    {
        File f(base + "ghosts", "w");
        if (!f.f) return ffail(c, MPLB_E_IO, "Unable to write ghosts");
        std::fprintf(f.f, "%d\n", 0);
    }
    {
        File f(base + "neighbours", "w");
        if (!f.f) return ffail(c, MPLB_E_IO, "Unable to write neighbours");
        std::fprintf(f.f, "%lld %d %d\n", (long long)N, m.d, m.Q);
        for (int64_t i = 0; i < N; i++) {
            for (int q = 0; q < m.Q; q++) std::fprintf(f.f, "%d ", m.nb[i * m.Q + q]);
            std::fprintf(f.f, "\n");
        }
    }
    for (const char* g : {"recvGhosts", "sendGhosts"}) {
        File f(base + g, "w");
        if (!f.f) return ffail(c, MPLB_E_IO, std::string("Unable to write ") + g);
        std::fprintf(f.f, "%d\n\n%d\n%d\n\n", 1, 0, 0);
    }
    {
        File f(base + "boundary", "w");
        if (!f.f) return ffail(c, MPLB_E_IO, "Unable to write boundary");
        std::fprintf(f.f, "%d\n\n", (int)m.bd.size());
        for (size_t b = 0; b < m.bd.size(); b++) {
            std::fprintf(f.f, "%s\n%d\n", m.bdnames[b].c_str(), (int)m.bd[b].size());
            for (int32_t id : m.bd[b]) std::fprintf(f.f, "%d\n", id);
            std::fprintf(f.f, "\n");
        }
    }
    {
        // VTK cells of a structured box (latticeBox2D.c:336-356 / latticeBox3D.c:360-392)
        File f(base + "vtkCells", "w");
        if (!f.f) return ffail(c, MPLB_E_IO, "Unable to write vtkCells");
        if (!m.points.empty() && !m.structured) {
            std::fprintf(f.f, "%d %d\n", 0, m.d == 3 ? 8 : 4);
        } else if (m.d == 3) {
            const long long nx = m.nx, ny = m.ny, nz = m.nz;
            std::fprintf(f.f, "%lld %d\n", (nx - 1) * (ny - 1) * (nz - 1), 8);
            for (long long k = 0; k < nz - 1; k++)
                for (long long j = 0; j < ny - 1; j++)
                    for (long long i = 0; i < nx - 1; i++) {
                        const long long b = i + j * nx + k * nx * ny, up = nx * ny;
                        std::fprintf(f.f, "%lld %lld %lld %lld %lld %lld %lld %lld \n", b, b + 1, b + nx, b + nx + 1,
                                     b + up, b + up + 1, b + up + nx, b + up + nx + 1);
                    }
        } else {
            const long long nx = m.nx, ny = m.ny;
            std::fprintf(f.f, "%lld %d\n", (nx - 1) * (ny - 1), 4);
            for (long long j = 0; j < ny - 1; j++)
                for (long long i = 0; i < nx - 1; i++) {
                    const long long b = i + j * nx;
                    std::fprintf(f.f, "%lld %lld %lld %lld \n", b, b + 1, b + nx, b + nx + 1);
                }
        }
    }
    return 0;
}

// writeScalarField.c / writeVectorField.c for rho, T, p, U, f, g (multiPhasePP.c:291-301)
int mplb_write_fields(mplb_case* c, const char* dir) {
    if (!c || !dir) return MPLB_E_ARG;
    const bool binary = c->data_format == "binaryRaw";
    if (!binary && c->data_format != "asciiRaw")
        return ffail(c, MPLB_E_MODEL, "dataFormat " + c->data_format + " is not written by this library");
    const std::string base = std::string(dir) + "/processor0/" + std::to_string(c->current) + "/";
    if (!mkdirs(base)) return ffail(c, MPLB_E_IO, "Unable to create time folder");
    const int64_t N = c->mesh.N;
    struct Item {
        const char* name;
        int comps;
        bool vector;
    };
    const Item items[] = {{"rho", 1, false}, {"T", 1, false}, {"p", 1, false},
                          {"U", 3, true},    {"f", c->mesh.Q, true}, {"g", c->mesh.Q, true}};
    std::vector<double> h;
    for (const Item& it : items) {
        h.assign((size_t)N * it.comps, 0.0);
        int rc = mplb_get_field(c, it.name, h.data(), h.size());
        if (rc) return rc;
        for (double v : h)
            if (std::isnan(v)) return ffail(c, MPLB_E_NAN, "Floating point exception. NaN solution");
        File f(base + it.name, binary ? "wb" : "w");
        if (!f.f) return ffail(c, MPLB_E_IO, std::string("Unable to write ") + it.name);
        if (binary) {
            std::vector<float> fv(h.begin(), h.end());
            if (std::fwrite(fv.data(), sizeof(float), fv.size(), f.f) != fv.size())
                return ffail(c, MPLB_E_IO, std::string("Unable to write ") + it.name);
        } else if (!it.vector) {
            for (int64_t i = 0; i < N; i++) std::fprintf(f.f, "%g\n", h[i]);
        } else {
            for (int64_t i = 0; i < N; i++) {
                for (int j = 0; j < it.comps; j++) std::fprintf(f.f, "%g ", h[i * it.comps + j]);
                std::fprintf(f.f, "\n");
            }
        }
    }
    return 0;
}

// readScalarFromAsciiRaw.c / readArrayFromAsciiRaw.c / *BinaryRaw.c: values through float
int mplb_read_field(const mplb_case* cc, const char* dir, int time, const char* name, int comps, double* out,
                    size_t n) {
    mplb_case* c = const_cast<mplb_case*>(cc);
    if (!c || !dir || !name || !out || comps < 1) return MPLB_E_ARG;
    if ((int64_t)n != c->mesh.N * comps) return ffail(c, MPLB_E_ARG, "output size differs from nodes x components");
    const bool binary = c->data_format == "binaryRaw";
    if (!binary && c->data_format != "asciiRaw")
        return ffail(c, MPLB_E_MODEL, "dataFormat " + c->data_format + " is not read by this library");
    const std::string path = std::string(dir) + "/processor0/" + std::to_string(time) + "/" + name;
    File f(path, binary ? "rb" : "r");
    if (!f.f) return ffail(c, MPLB_E_IO, "Unable to open " + path);
    if (binary) {
        std::vector<float> fv(n);
        if (std::fread(fv.data(), sizeof(float), n, f.f) != n) return ffail(c, MPLB_E_IO, "Short file " + path);
        for (size_t i = 0; i < n; i++) out[i] = (double)fv[i];
    } else {
        for (size_t i = 0; i < n; i++) {
            float v;
            if (std::fscanf(f.f, "%f", &v) != 1) return ffail(c, MPLB_E_IO, "Ill defined field " + path);
            out[i] = (double)v;
        }
    }
    return 0;
}

}  // extern "C"
