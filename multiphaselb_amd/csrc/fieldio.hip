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
//     naming it, and case.pvd listing the time folders present;
//   * binary EnSight Gold (IO/ensight/*): lattice.geo, the lbm.case time list and one
//     float32 file per variable and time index in the case folder.
#include <hip/hip_runtime.h>

#include <dirent.h>
#include <sys/stat.h>

#include <algorithm>

#include <cerrno>
#include <climits>
#include <cstdint>
#include <cmath>
#include <cstdio>
#include <cstdlib>
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

// readVTKInfo.c: the scalar, vector and pdf field names listed in start/initialFields
// (empty lists when the file is absent)
void field_names(const mplb_case* c, std::vector<std::string>* sc, std::vector<std::string>* vc,
                 std::vector<std::string>* pc) {
    Dict init;
    if (init.open(c->dir + "/start/initialFields")) {
        init.string_list("scalarFields", sc);
        init.string_list("vectorFields", vc);
        init.string_list("pdfFields", pc);
    }
}

// IO/writeMeshToVTK.c + writeScalarToPvtu.c + writeArrayToPvtu.c + writePvtuExtra.c +
// writeMainPvd.c for one rank
int write_pvtu(mplb_case* c, const std::string& dir, const FieldItem* items, int nitems) {
    const Mesh& m = c->mesh;
    const int64_t N = m.N;
    std::vector<std::string> sc, vc, pc;
    field_names(c, &sc, &vc, &pc);
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

// ---- EnSight Gold, binary (IO/ensight/*, one rank: part 1) ------------------------
// Every file sits in the case folder: lattice.geo, lbm.case and one file per variable
// and time index, lattice.<name>_<index> (population k of an lbe field: lattice.<name>k_
// <index>).  Text records are 80-byte NUL-padded blocks, numbers int32 / float32.

bool put80(FILE* f, const char* s) {
    char b[80];
    std::memset(b, 0, sizeof b);
    std::strncpy(b, s, sizeof b - 1);
    return std::fwrite(b, 1, sizeof b, f) == sizeof b;
}

bool put_i32(FILE* f, int32_t v) { return std::fwrite(&v, sizeof v, 1, f) == 1; }

// the time list of <dir>/lbm.case as timeToIndex.c and updateCaseFile.c scan it: the
// integer after the words "number of steps:", then that many integers after "time
// values:".  false when the file is absent.
bool case_times(const std::string& dir, std::vector<long long>* t) {
    t->clear();
    File f(dir + "/lbm.case", "r");
    if (!f.f) return false;
    std::vector<std::string> w;
    char b[256];
    while (std::fscanf(f.f, "%255s", b) == 1) w.push_back(b);
    long long ns = -1;
    for (size_t i = 0; i + 3 < w.size() && ns < 0; i++)
        if (w[i] == "number" && w[i + 1] == "of" && w[i + 2] == "steps:") ns = std::atoll(w[i + 3].c_str());
    for (size_t i = 0; i + 1 < w.size() && ns > 0; i++)
        if (w[i] == "time" && w[i + 1] == "values:") {
            for (long long k = 0; k < ns && i + 2 + (size_t)k < w.size(); k++)
                t->push_back(std::atoll(w[i + 2 + (size_t)k].c_str()));
            break;
        }
    return true;
}

// timeToIndex.c: position of `time` in the case file's list (the last match), 0 when it is
// not listed or there is no case file
long long ensight_index(const std::string& dir, long long time) {
    std::vector<long long> t;
    case_times(dir, &t);
    long long idx = 0;
    for (size_t k = 0; k < t.size(); k++)
        if (t[k] == time) idx = (long long)k;
    return idx;
}

// updateCaseFile.c: append the current time to the case file's list unless it is already
// the last entry; a case folder without lbm.case starts the list at time 0 (whatever the
// current time: the field files of that write then carry index 0).  The VARIABLE block
// names the fields of start/initialFields, pdf fields one scalar per population.
int update_case_file(mplb_case* c, const std::string& dir) {
    std::vector<long long> t;
    if (case_times(dir, &t)) {
        if (t.empty() || t.back() != c->current) t.push_back(c->current);
    } else {
        t.assign(1, 0);
    }
    std::vector<std::string> sc, vc, pc;
    field_names(c, &sc, &vc, &pc);
    File f(dir + "/lbm.case", "w");
    if (!f.f) return ffail(c, MPLB_E_IO, "Unable to write lbm.case");
    // the value column starts at character 27 on every keyword line
    std::fprintf(f.f, "#\n# EnSight 7.4.1 ((n))\n# Case File: lattice.case\n\nFORMAT\n\n");
    std::fprintf(f.f, "type:  ensight gold\n\nGEOMETRY\n\n%-27s%s\n\nVARIABLE\n\n", "model:", "lattice.geo");
    for (const std::string& n : sc) std::fprintf(f.f, "%-27s%s lattice.%s_*\n", "scalar per node:", n.c_str(), n.c_str());
    for (const std::string& n : pc)
        for (int k = 0; k < c->mesh.Q; k++)
            std::fprintf(f.f, "%-27s%s%d lattice.%s%d_*\n", "scalar per node:", n.c_str(), k, n.c_str(), k);
    for (const std::string& n : vc) std::fprintf(f.f, "%-27s%s lattice.%s_*\n", "vector per node:", n.c_str(), n.c_str());
    std::fprintf(f.f, "\nTIME\n%-27s%d\n%-27s%d\n%-27s%d\n%-27s%d\n", "time set:", 1, "number of steps:", (int)t.size(),
                 "filename start number:", 0, "filename increment:", 1);
    for (size_t k = 0; k < t.size(); k++) std::fprintf(f.f, "%-27s%lld\n", k ? "" : "time values:", t[k]);
    if (std::ferror(f.f)) return ffail(c, MPLB_E_IO, "Unable to write lbm.case");
    return 0;
}

// writeMeshToEnsight.c: lattice.geo with the node coordinates (x block, y block, z block)
// and the VTK cells as quad4 (D2Q9) or hexa8 (D3Q15), 1-based, each face's last two
// nodes swapped (VTK's lexicographic voxel order -> EnSight's cyclic order)
int write_ensight_geo(mplb_case* c, const std::string& dir) {
    const Mesh& m = c->mesh;
    std::vector<long long> conn;
    int ctype = 0;
    if (!vtk_cells(c, &conn, &ctype)) return ffail(c, MPLB_E_IO, "Ill defined vtkCells file");
    if (ctype != (m.d == 3 ? 8 : 4)) return ffail(c, MPLB_E_IO, "vtkCells do not hold one voxel per lattice cell");
    if (m.N > INT32_MAX || (int64_t)(conn.size() / (size_t)ctype) > INT32_MAX)
        return ffail(c, MPLB_E_IO, "lattice.geo counts nodes and cells in 32-bit integers: lattice too large");
    {
        File f(dir + "/lattice.geo", "wb");
        if (!f.f) return ffail(c, MPLB_E_IO, "Unable to write lattice.geo");
        for (const char* h : {"C Binary", "EnSight Model Geometry File", "EnSight 7.4.1", "node id assign",
                              "element id assign"})
            put80(f.f, h);
    }
    const int rc = update_case_file(c, dir);
    if (rc) return rc;
    File f(dir + "/lattice.geo", "ab");
    if (!f.f) return ffail(c, MPLB_E_IO, "Unable to write lattice.geo");
    bool ok = put80(f.f, "part") && put_i32(f.f, 1) && put80(f.f, "Point field") && put80(f.f, "coordinates") &&
              put_i32(f.f, (int32_t)m.N);
    std::vector<float> xyz((size_t)m.N);
    for (int d = 0; d < 3 && ok; d++) {
        for (int64_t i = 0; i < m.N; i++) {
            int p[3];
            if (!m.points.empty()) {
                p[0] = m.points[3 * i];
                p[1] = m.points[3 * i + 1];
                p[2] = m.points[3 * i + 2];
            } else {
                coords(m, i, p);
            }
            xyz[(size_t)i] = (float)p[d];
        }
        ok = std::fwrite(xyz.data(), sizeof(float), xyz.size(), f.f) == xyz.size();
    }
    const long long ncells = (long long)(conn.size() / (size_t)ctype);
    ok = ok && put80(f.f, m.d == 3 ? "hexa8" : "quad4") && put_i32(f.f, (int32_t)ncells);
    static const int order[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    std::vector<int32_t> ids((size_t)conn.size());
    for (long long i = 0; i < ncells; i++)
        for (int j = 0; j < ctype; j++) ids[(size_t)(i * ctype + j)] = (int32_t)(conn[(size_t)(i * ctype + order[j])] + 1);
    ok = ok && std::fwrite(ids.data(), sizeof(int32_t), ids.size(), f.f) == ids.size();
    if (!ok) return ffail(c, MPLB_E_IO, "Unable to write lattice.geo");
    return 0;
}

// writeScalarToEnsight.c / writeVectorToEnsight.c / writeLbeFieldToEnsight.c body: the
// variable's name, "part", 1, "coordinates", then the values as float32
bool ensight_var(const std::string& path, const char* header, const float* v, size_t n) {
    File f(path, "wb");
    return f.f && put80(f.f, header) && put80(f.f, "part") && put_i32(f.f, 1) && put80(f.f, "coordinates") &&
           std::fwrite(v, sizeof(float), n, f.f) == n;
}

// multiPhasePP.c:283-301 with dataFormat ensight: the case file first, then each field
// (NaN check, then its file): scalars, U as three component blocks, f and g one file per
// population, each headed by the field's own name
int write_ensight(mplb_case* c, const std::string& dir, const FieldItem* items, int nitems) {
    if (!mkdirs(dir)) return ffail(c, MPLB_E_IO, "Unable to create " + dir);
    int rc = update_case_file(c, dir);
    if (rc) return rc;
    const long long idx = ensight_index(dir, c->current);
    const int64_t N = c->mesh.N;
    std::vector<double> h;
    std::vector<float> v((size_t)N * 3);
    for (int k = 0; k < nitems; k++) {
        const FieldItem& it = items[k];
        rc = fetch_checked(c, it, &h);
        if (rc) return rc;
        const std::string tail = "_" + std::to_string(idx);
        if (it.comps == 1 || it.comps == 3) {
            for (int j = 0; j < it.comps; j++)
                for (int64_t i = 0; i < N; i++) v[(size_t)(j * N + i)] = (float)h[(size_t)(i * it.comps + j)];
            if (!ensight_var(dir + "/lattice." + it.name + tail, it.name, v.data(), (size_t)(N * it.comps)))
                return ffail(c, MPLB_E_IO, std::string("Unable to write lattice.") + it.name + tail);
        } else {
            for (int q = 0; q < it.comps; q++) {
                for (int64_t i = 0; i < N; i++) v[(size_t)i] = (float)h[(size_t)(i * it.comps + q)];
                const std::string path = dir + "/lattice." + it.name + std::to_string(q) + tail;
                if (!ensight_var(path, it.name, v.data(), (size_t)N)) return ffail(c, MPLB_E_IO, "Unable to write " + path);
            }
        }
    }
    return 0;
}

// readScalarFromEnsight.c / readVectorFromEnsight.c / readLbeFieldFromEnsight.c: the
// file of the case-file index of `time`, values after the 244-byte header (three text
// records and the part number), through float
int read_ensight(mplb_case* c, const std::string& dir, int time, const char* name, int comps, double* out) {
    const int64_t N = c->mesh.N;
    const std::string tail = "_" + std::to_string(ensight_index(dir, time));
    const long header = 3 * 80 + (long)sizeof(int32_t);
    std::vector<float> v((size_t)N * (comps == 3 ? 3 : 1));
    auto load = [&](const std::string& path) -> int {
        File f(path, "rb");
        if (!f.f) return ffail(c, MPLB_E_IO, "Unable to open " + path);
        if (std::fseek(f.f, header, SEEK_SET) != 0 || std::fread(v.data(), sizeof(float), v.size(), f.f) != v.size())
            return ffail(c, MPLB_E_IO, "Short file " + path);
        return 0;
    };
    if (comps == 1 || comps == 3) {
        const int rc = load(dir + "/lattice." + name + tail);
        if (rc) return rc;
        for (int j = 0; j < comps; j++)
            for (int64_t i = 0; i < N; i++) out[(size_t)(i * comps + j)] = (double)v[(size_t)(j * N + i)];
        return 0;
    }
    for (int q = 0; q < comps; q++) {
        const int rc = load(dir + "/lattice." + name + std::to_string(q) + tail);
        if (rc) return rc;
        for (int64_t i = 0; i < N; i++) out[(size_t)(i * comps + q)] = (double)v[(size_t)i];
    }
    return 0;
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
    // writeLatticeMesh.c:17-22: with dataFormat ensight the geometry and case file first
    if (c->data_format == "ensight") {
        const int rc = write_ensight_geo(c, dir);
        if (rc) return rc;
    }
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
    if (c->data_format == "ensight") return write_ensight(c, dir, items, 6);
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
    if (c->data_format == "ensight") return read_ensight(c, dir, time, name, comps, out);
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
