// Files in the reference's on-disk formats (host code):
//   * lattice tables of one rank, processor0/lattice/{points, ghosts, neighbours,
//     recvGhosts, sendGhosts, boundary, vtkCells} (IO/writeLatticeMesh.c:20-226 writes
//     them, IO/readLatticeMesh.c reads them);
//   * macroscopic and population fields of a time level, processor0/<time>/<name>, in
//     the asciiRaw (one node per line, "%g" values) and binaryRaw (node-major float32)
//     formats (IO/asciiRaw/*, IO/binaryRaw/*); values read back pass through float as
//     in readScalarFromAsciiRaw.c / readArrayFromAsciiRaw.c;
//   * the pvtu format (IO/pvtu/*): processor0/<time>/fields.vtu with the points, the
//     VTK cells and every field as an ascii Float32 DataArray, <time>/fields.pvtu
//     naming it, and case.pvd listing the time folders present.
#include <hip/hip_runtime.h>

#include <dirent.h>
#include <sys/stat.h>

#include <algorithm>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mplb.h"
#include "case.h"
#include "dict.h"
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

// VTK cells of the lattice, ctype node ids each: generated for a structured box
// (latticeBox2D.c:336-356 / latticeBox3D.c:360-392: the quad or hexahedron whose lowest
// node is (i, j, k)), otherwise those of the vtkCells file the lattice was read with
// (none when that file is absent).  false: a vtkCells file that does not parse.
bool vtk_cells(const mplb_case* c, std::vector<long long>* conn, int* ctype) {
    const Mesh& m = c->mesh;
    *ctype = m.d == 3 ? 8 : 4;
    conn->clear();
    if (!m.points.empty() && !m.structured) {
        File f(c->dir + "/processor0/lattice/vtkCells", "r");
        if (!f.f) return true;
        long long n = 0;
        int t = 0;
        if (std::fscanf(f.f, "%lld %d", &n, &t) != 2 || n < 0 || t <= 0 || t > 8) return false;
        *ctype = t;
        conn->resize((size_t)(n * t));
        for (long long& v : *conn)
            if (std::fscanf(f.f, "%lld", &v) != 1 || v < 0 || v >= m.N) return false;
        return true;
    }
    const long long nx = m.nx, ny = m.ny, nz = m.d == 3 ? m.nz : 1;
    const long long up = nx * ny;
    for (long long k = 0; k < (m.d == 3 ? nz - 1 : 1); k++)
        for (long long j = 0; j < ny - 1; j++)
            for (long long i = 0; i < nx - 1; i++) {
                const long long b = i + j * nx + k * up;
                for (long long v : {b, b + 1, b + nx, b + nx + 1}) conn->push_back(v);
                if (m.d == 3)
                    for (long long v : {b + up, b + up + 1, b + up + nx, b + up + nx + 1}) conn->push_back(v);
            }
    return true;
}

// field arrays of a write step (multiPhasePP.c:291-301): rho, T, p, U, f, g
struct FieldItem {
    const char* name;
    int comps;
    bool vector;
};

// one field from the device, refused with the reference's message when it holds a NaN
// (writeScalarField.c / writeVectorField.c)
int fetch_checked(mplb_case* c, const FieldItem& it, std::vector<double>* h) {
    h->assign((size_t)c->mesh.N * it.comps, 0.0);
    const int rc = mplb_get_field(c, it.name, h->data(), h->size());
    if (rc) return rc;
    for (double v : *h)
        if (std::isnan(v)) return ffail(c, MPLB_E_NAN, "Floating point exception. NaN solution");
    return 0;
}

// "a b c" with single spaces (writeMeshToPvtu.c prints the scalar and pdf names that way
// and every vector name followed by a space)
std::string joined(const std::vector<std::string>& v, bool trailing) {
    std::string s;
    for (size_t i = 0; i < v.size(); i++) s += v[i] + (trailing || i + 1 < v.size() ? " " : "");
    return s;
}

// IO/writeMeshToVTK.c + writeScalarToPvtu.c + writeArrayToPvtu.c + writePvtuExtra.c +
// writeMainPvd.c for one rank
int write_pvtu(mplb_case* c, const std::string& dir, const FieldItem* items, int nitems) {
    const Mesh& m = c->mesh;
    const int64_t N = m.N;
    // readVTKInfo.c: the names listed in start/initialFields
    std::vector<std::string> sc, vc, pc;
    Dict init;
    if (init.open(c->dir + "/start/initialFields")) {
        init.string_list("scalarFields", &sc);
        init.string_list("vectorFields", &vc);
        init.string_list("pdfFields", &pc);
    }
    std::vector<long long> conn;
    int ctype = 0;
    if (!vtk_cells(c, &conn, &ctype)) return ffail(c, MPLB_E_IO, "Ill defined vtkCells file");
    const long long ncells = (long long)(conn.size() / (size_t)ctype);
    std::vector<std::vector<double>> data(nitems);
    for (int k = 0; k < nitems; k++) {
        const int rc = fetch_checked(c, items[k], &data[k]);
        if (rc) return rc;
    }
    const std::string t = std::to_string(c->current);
    const std::string pdir = dir + "/processor0/" + t + "/";
    if (!mkdirs(pdir) || !mkdirs(dir + "/" + t)) return ffail(c, MPLB_E_IO, "Unable to create time folder");
    const std::string pd = joined(sc, false), vd = joined(vc, true) + joined(pc, false);
    {
        File f(pdir + "fields.vtu", "w");
        if (!f.f) return ffail(c, MPLB_E_IO, "Unable to write fields.vtu");
        std::fprintf(f.f, "<?xml version=\"1.0\"?>\n");
        std::fprintf(f.f, "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\" "
                          "compressor=\"vtkZLibDataCompressor\">\n");
        std::fprintf(f.f, "  <UnstructuredGrid>\n");
        std::fprintf(f.f, "    <Piece NumberOfPoints=\"%lld\" NumberOfCells=\"%lld\">\n", (long long)N, ncells);
        std::fprintf(f.f, "      <Points>\n");
        std::fprintf(f.f, "        <DataArray type=\"Float32\" NumberOfComponents=\"3\" format=\"ascii\">\n");
        for (int64_t i = 0; i < N; i++) {
            int p[3];
            if (!m.points.empty()) {
                p[0] = m.points[3 * i];
                p[1] = m.points[3 * i + 1];
                p[2] = m.points[3 * i + 2];
            } else {
                coords(m, i, p);
            }
            std::fprintf(f.f, "          %.1f %.1f %.1f\n", (double)(float)p[0], (double)(float)p[1], (double)(float)p[2]);
        }
        std::fprintf(f.f, "        </DataArray>\n      </Points>\n      <Cells>\n");
        std::fprintf(f.f, "        <DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n");
        for (long long i = 0; i < ncells; i++) {
            std::fprintf(f.f, "          ");
            for (int j = 0; j < ctype; j++) std::fprintf(f.f, "%lld ", conn[(size_t)(i * ctype + j)]);
            std::fprintf(f.f, "\n");
        }
        std::fprintf(f.f, "        </DataArray>\n");
        std::fprintf(f.f, "        <DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n");
        for (long long i = 0; i < ncells; i++) std::fprintf(f.f, "          %lld\n", (i + 1) * ctype);
        std::fprintf(f.f, "        </DataArray>\n");
        // the reference labels every cell type 8 (writeMeshToPvtu.c), 2D or 3D
        std::fprintf(f.f, "        <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n");
        for (long long i = 0; i < ncells; i++) std::fprintf(f.f, "          8\n");
        std::fprintf(f.f, "        </DataArray>\n      </Cells>\n");
        std::fprintf(f.f, "      <PointData Scalars=\"%s\" Vectors=\"%s\">\n", pd.c_str(), vd.c_str());
        for (int k = 0; k < nitems; k++) {
            const FieldItem& it = items[k];
            const std::vector<double>& h = data[k];
            if (!it.vector) {
                std::fprintf(f.f, "        <DataArray type=\"Float32\" Name=\"%s\" format=\"ascii\">\n", it.name);
                for (int64_t i = 0; i < N; i++) std::fprintf(f.f, "          %g\n", h[i]);
            } else {
                std::fprintf(f.f, "        <DataArray type=\"Float32\" Name=\"%s\"  NumberOfComponents=\"%d\" format=\"ascii\">\n",
                             it.name, it.comps);
                for (int64_t i = 0; i < N; i++) {
                    std::fprintf(f.f, "          ");
                    for (int j = 0; j < it.comps; j++) std::fprintf(f.f, "%g ", h[(size_t)i * it.comps + j]);
                    std::fprintf(f.f, "\n");
                }
            }
            std::fprintf(f.f, "        </DataArray>\n");
        }
        std::fprintf(f.f, "      </PointData>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n");
        if (std::ferror(f.f)) return ffail(c, MPLB_E_IO, "Unable to write fields.vtu");
    }
    {
        File f(dir + "/" + t + "/fields.pvtu", "w");
        if (!f.f) return ffail(c, MPLB_E_IO, "Unable to write fields.pvtu");
        std::fprintf(f.f, "<?xml version=\"1.0\"?>\n");
        std::fprintf(f.f, "<VTKFile type=\"PUnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
        std::fprintf(f.f, "  <PUnstructuredGrid GhostLevel=\"0\">\n");
        std::fprintf(f.f, "    <PPointData Scalars=\"%s\" Vectors=\"%s\">\n", pd.c_str(), vd.c_str());
        for (const std::string& n : sc) std::fprintf(f.f, "      <PDataArray type=\"Float32\" Name=\"%s\"/>\n", n.c_str());
        for (const std::string& n : vc)
            std::fprintf(f.f, "      <PDataArray type=\"Float32\" Name=\"%s\" NumberOfComponents=\"3\"/>\n", n.c_str());
        for (const std::string& n : pc)
            std::fprintf(f.f, "      <PDataArray type=\"Float32\" Name=\"%s\" NumberOfComponents=\"%d\"/>\n", n.c_str(), m.Q);
        std::fprintf(f.f, "    </PPointData>\n    <PPoints>\n");
        std::fprintf(f.f, "      <PDataArray type=\"Float32\" NumberOfComponents=\"3\"/>\n    </PPoints>\n");
        std::fprintf(f.f, "    <Piece Source=\"../processor0/%s/fields.vtu\"/>\n", t.c_str());
        std::fprintf(f.f, "  </PUnstructuredGrid>\n</VTKFile>\n");
    }
    // writeMainPvd.c: every entry of the case folder whose name starts with an integer,
    // in increasing order
    std::vector<int> times;
    if (DIR* d = opendir(dir.c_str())) {
        while (struct dirent* e = readdir(d)) {
            int v;
            if (std::sscanf(e->d_name, "%d", &v) == 1) times.push_back(v);
        }
        closedir(d);
    }
    std::sort(times.begin(), times.end());
    File f(dir + "/case.pvd", "w");
    if (!f.f) return ffail(c, MPLB_E_IO, "Unable to write case.pvd");
    std::fprintf(f.f, "<VTKFile type=\"Collection\" version=\"0.1\" byte_order=\"LittleEndian\">\n  <Collection>\n");
    for (int v : times) std::fprintf(f.f, "    <DataSet part=\"0\"  timestep=\"%d\" file=\"%d/fields.pvtu\" />\n", v, v);
    std::fprintf(f.f, "  </Collection>\n</VTKFile>\n");
    return 0;
}

// readScalarFromPvtu.c / readArrayFromPvtu.c: after <PointData, the DataArray whose
// second word is Name="<name>"; its values follow the header's remaining words
int read_pvtu(mplb_case* c, const std::string& path, const char* name, int comps, double* out, size_t n) {
    File f(path, "r");
    if (!f.f) return ffail(c, MPLB_E_IO, "Unable to open " + path);
    char w[4096];
    auto word = [&]() { return std::fscanf(f.f, "%4095s", w) == 1; };
    bool in_points = false;
    while (word())
        if (std::strcmp(w, "<PointData") == 0) {
            in_points = true;
            break;
        }
    if (!in_points) return ffail(c, MPLB_E_IO, "No PointData in " + path);
    const std::string want = std::string("Name=\"") + name + "\"";
    for (;;) {
        bool found = false;
        while (word())
            if (std::strcmp(w, "<DataArray") == 0) {
                found = true;
                break;
            }
        if (!found) return ffail(c, MPLB_E_IO, std::string("Field ") + name + " not found in " + path);
        if (!word() || !word()) return ffail(c, MPLB_E_IO, "Ill defined " + path);
        if (want != w) continue;
        for (int k = 0; k < (comps > 1 ? 2 : 1); k++)
            if (!word()) return ffail(c, MPLB_E_IO, "Ill defined " + path);
        for (size_t i = 0; i < n; i++)
            if (std::fscanf(f.f, "%lf", &out[i]) != 1) return ffail(c, MPLB_E_IO, "Ill defined field in " + path);
        return 0;
    }
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
        File f(base + "vtkCells", "w");
        if (!f.f) return ffail(c, MPLB_E_IO, "Unable to write vtkCells");
        std::vector<long long> conn;
        int ctype = 0;
        if (!vtk_cells(c, &conn, &ctype)) return ffail(c, MPLB_E_IO, "Ill defined vtkCells file");
        const long long ncells = (long long)(conn.size() / (size_t)ctype);
        std::fprintf(f.f, "%lld %d\n", ncells, ctype);
        for (long long i = 0; i < ncells; i++) {
            for (int j = 0; j < ctype; j++) std::fprintf(f.f, "%lld ", conn[(size_t)(i * ctype + j)]);
            std::fprintf(f.f, "\n");
        }
    }
    return 0;
}

// writeScalarField.c / writeVectorField.c for rho, T, p, U, f, g (multiPhasePP.c:291-301)
int mplb_write_fields(mplb_case* c, const char* dir) {
    if (!c || !dir) return MPLB_E_ARG;
    const FieldItem items[] = {{"rho", 1, false}, {"T", 1, false}, {"p", 1, false},
                               {"U", 3, true},    {"f", c->mesh.Q, true}, {"g", c->mesh.Q, true}};
    if (c->data_format == "pvtu") return write_pvtu(c, dir, items, 6);
    const bool binary = c->data_format == "binaryRaw";
    if (!binary && c->data_format != "asciiRaw")
        return ffail(c, MPLB_E_MODEL, "dataFormat " + c->data_format + " is not written by this library");
    const std::string base = std::string(dir) + "/processor0/" + std::to_string(c->current) + "/";
    if (!mkdirs(base)) return ffail(c, MPLB_E_IO, "Unable to create time folder");
    const int64_t N = c->mesh.N;
    std::vector<double> h;
    for (const FieldItem& it : items) {
        const int rc = fetch_checked(c, it, &h);
        if (rc) return rc;
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
    if (c->data_format == "pvtu")
        return read_pvtu(c, std::string(dir) + "/processor0/" + std::to_string(time) + "/fields.vtu", name, comps, out, n);
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
