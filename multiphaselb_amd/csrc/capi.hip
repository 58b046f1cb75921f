// C ABI of libmplb: case parsing (reference dictionaries), device state, and the
// per-step driver that replays multiPhasePP's time loop on the GPU.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <set>
#include <string>
#include <unordered_set>
#include <vector>

#include "../../include/mplb.h"
#include "case.h"
#include "fused.h"
#include "glibc_rand.h"

using namespace mplb;

namespace {

int fail(mplb_case* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(c, call)                                                                      \
    do {                                                                                     \
        hipError_t _e = (call);                                                              \
        if (_e != hipSuccess)                                                                \
            return fail((c), MPLB_E_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

int model_id(const std::string& s) {
    if (s == "liMRT") return M_LIMRT;
    if (s == "myMRT") return M_MYMRT;
    if (s == "xuMRT") return M_XUMRT;
    if (s == "liTempMRT") return M_LITEMPMRT;
    return -1;
}

int bd_id(const std::string& s) {
    if (s == "bounceBack") return BD_BBACK;
    if (s == "periodic") return BD_PERIODIC;
    if (s == "fixedT") return BD_FT;
    if (s == "fixedRandomT") return BD_FRDT;
    if (s == "fixedU") return BD_FU;
    if (s == "outflow") return BD_OF;
    if (s == "fixedGradT") return BD_FGT;
    return -1;
}

// readLiMRTConstants.c / readXuMRTConstants.c / readMyMRTConstants.c / readLiTempMRTConstants.c
int read_model(mplb_case* c, const Dict& mp, const char* fname, FieldCfg& fc) {
    std::string col;
    if (!mp.string(std::string(fname) + "/Collision", "", &col)) return fail(c, MPLB_E_IO, "Collision model not found");
    fc.mp.model = model_id(col);
    if (fc.mp.model < 0) return fail(c, MPLB_E_MODEL, std::string("Collision model is not yet implemented: ") + col);
    std::vector<double> lam;
    if (!mp.vector(std::string(fname) + "/Lambda", &lam))
        return fail(c, MPLB_E_IO, std::string("Ill defined entry ") + fname + "/Lambda");
    for (int k = 0; k < 15; k++) fc.mp.Lambda[k] = k < (int)lam.size() ? lam[k] : 0.0;
    if ((int)lam.size() < c->mesh.Q)
        return fail(c, MPLB_E_IO, std::string("Ill defined entry ") + fname + "/Lambda (fewer than Q values)");
    if (fc.mp.model == M_LIMRT || fc.mp.model == M_XUMRT) {
        if (!mp.scalar(std::string(fname) + "/sigma", 0, &fc.mp.sigma))
            return fail(c, MPLB_E_IO, std::string("Ill defined entry ") + fname + "/sigma");
        std::string st;
        mp.string(std::string(fname) + "/SurfaceTension", "none", &st);
        fc.mp.surf = 0;
        const char* want = fc.mp.model == M_LIMRT ? "liSurfTen" : "xuSurfTen";
        if (st == want) {
            fc.mp.surf = 1;
            if (!mp.scalar(std::string(fname) + "/kappa_st", -1, &fc.mp.kappa))
                return fail(c, MPLB_E_IO, std::string("Unable to find ") + fname + "/kappa_st");
        }
    }
    if (fc.mp.model == M_MYMRT) {
        if (!mp.scalar(std::string(fname) + "/alpha_1", 0, &fc.mp.alpha1))
            return fail(c, MPLB_E_IO, std::string("Ill defined entry ") + fname + "/alpha_1");
        if (!mp.scalar(std::string(fname) + "/alpha_2", 0, &fc.mp.alpha2))
            return fail(c, MPLB_E_IO, std::string("Ill defined entry ") + fname + "/alpha_2");
    }
    return 0;
}

// readEOSInfo.c:10-111 (g and extForce default to zero where the reference would leave
// the pointer unset)
int read_eos(mplb_case* c, const Dict& mp) {
    EosParams& e = c->eos;
    std::memset(&e, 0, sizeof(e));
    std::string model;
    e.type = -1;
    if (mp.string("EOS/model", "", &model)) {
        if (model == "idealGas") e.type = EOS_IDEAL;
        if (model == "Carnahan-Starling") {
            e.type = EOS_CARSTAR;
            if (!mp.scalar("EOS/a_cs", 0, &e.a)) return fail(c, MPLB_E_IO, "Unable to find constant a_cs");
            if (!mp.scalar("EOS/b_cs", 0, &e.b)) return fail(c, MPLB_E_IO, "Unable to find constant b_cs");
        }
        if (model == "Peng-Robinson") {
            e.type = EOS_PENGROB;
            if (!mp.scalar("EOS/a_pr", 0, &e.a)) return fail(c, MPLB_E_IO, "Unable to find constant a_pr");
            if (!mp.scalar("EOS/b_pr", 0, &e.b)) return fail(c, MPLB_E_IO, "Unable to find constant b_pr");
            if (!mp.scalar("EOS/w_pr", 0, &e.w_pr)) return fail(c, MPLB_E_IO, "Unable to find constant w_pr");
        }
        if (model == "vanDerWaals") {
            e.type = EOS_VDW;
            if (!mp.scalar("EOS/a_vdw", 0, &e.a)) return fail(c, MPLB_E_IO, "Unable to find constant a_vdw");
            if (!mp.scalar("EOS/b_vdw", 0, &e.b)) return fail(c, MPLB_E_IO, "Unable to find constant b_vdw");
        }
    }
    if (e.type < 0) return fail(c, MPLB_E_MODEL, "EOS not defined");
    mp.scalar("EOS/R", 0, &e.R);
    mp.scalar("EOS/G", 0, &e.G);
    mp.scalar("EOS/M", 0, &e.M);
    mp.scalar("EOS/Cv", 0, &e.Cv);
    mp.scalar("EOS/rho_0", 0, &c->rho0);
    c->rho0_dict = c->rho0;
    std::vector<double> v;
    if (mp.vector("EOS/g", &v))
        for (int j = 0; j < 3 && j < (int)v.size(); j++) e.g[j] = v[j];
    if (mp.vector("EOS/extForce", &v))
        for (int j = 0; j < 3 && j < (int)v.size(); j++) e.fext[j] = v[j];
    return 0;
}

// readLbeField.c:139-221 with readFixedTParam.c, readFixedRandomTParam.c:17-73, readFixedUParam.c
int read_boundaries(mplb_case* c, const Dict& bd, const char* fname, FieldCfg& fc) {
    const int nbd = (int)c->mesh.bd.size();
    fc.bctype.assign(nbd, BD_PERIODIC);
    fc.bcval.assign(nbd, {0, 0, 0});
    fc.pert.assign(nbd, 0);
    fc.psteps.assign(nbd, 0);
    fc.localT.assign(nbd, 0);
    fc.bcname.assign(nbd, "");
    for (int i = 0; i < nbd; i++) {
        const std::string base = std::string(fname) + "/" + c->mesh.bdnames[i];
        std::string type;
        if (!bd.string(base + "/type", "", &type)) return fail(c, MPLB_E_IO, "Unable to find entry " + base + "/type");
        fc.bcname[i] = type;
        fc.bctype[i] = bd_id(type);
        if (fc.bctype[i] == BD_FT || fc.bctype[i] == BD_FRDT) {
            if (!bd.scalar(base + "/value", 0, &fc.bcval[i][0])) return fail(c, MPLB_E_IO, "Unable to find " + base + "/value");
        }
        if (fc.bctype[i] == BD_FRDT) {
            if (!bd.scalar(base + "/pert", 0, &fc.pert[i])) {
                fc.pert[i] = 0;
                fc.psteps[i] = 0;
                fc.localT[i] = 1;
            } else {
                double steps = 0;
                if (!bd.scalar(base + "/duration", 0, &steps)) return fail(c, MPLB_E_IO, "Unable to find " + base + "/duration");
                fc.psteps[i] = (int)steps;
                fc.localT[i] = 0;
            }
        }
        if (fc.bctype[i] == BD_FU) {
            std::vector<double> u;
            if (!bd.vector(base + "/value", &u)) return fail(c, MPLB_E_IO, "Unable to find " + base + "/value");
            if (u.size() != 3) return fail(c, MPLB_E_IO, "Ill defined vector " + base + "/value value");
            for (int j = 0; j < 3; j++) fc.bcval[i][j] = u[j];
        }
    }
    return 0;
}

int parse_case(mplb_case* c) {
    Dict lp, mp, sim, bd;
    const std::string d = c->dir + "/";
    if (!lp.open(d + "properties/latticeProperties")) return fail(c, MPLB_E_IO, "Unable to open properties/latticeProperties");
    if (!mp.open(d + "properties/macroProperties")) return fail(c, MPLB_E_IO, "Unable to open properties/macroProperties");
    if (!sim.open(d + "properties/simulation")) return fail(c, MPLB_E_IO, "Unable to open properties/simulation");
    // setLatticeInfo.c:16-90
    std::string lbm;
    lp.string("LBModel", "D2Q9", &lbm);
    if (lbm == "D2Q9") c->mesh.lbmodel = 0;
    else if (lbm == "D3Q15") c->mesh.lbmodel = 1;
    else return fail(c, MPLB_E_MODEL, "Unrecognized model " + lbm);
    c->mesh.Q = c->mesh.lbmodel == 0 ? 9 : 15;
    c->mesh.d = c->mesh.lbmodel == 0 ? 2 : 3;
    // latticeBox2D.c:41-45 / latticeBox3D.c:49-56
    double dn;
    c->mesh.nx = lp.scalar("Nx", -1, &dn) ? (int)dn : 0;
    c->mesh.ny = lp.scalar("Ny", -1, &dn) ? (int)dn : 0;
    c->mesh.nz = c->mesh.d == 3 ? (lp.scalar("Nz", -1, &dn) ? (int)dn : 0) : 1;
    std::string bt;
    lp.string("boundaryType", "generic", &bt);
    static const char* names[] = {"generic", "periodicX", "periodicY", "periodicXY", "periodicZ", "periodicXYZ", "periodicXYSpot"};
    c->mesh.btype = -1;
    for (int i = 0; i < 7; i++)
        if (bt == names[i]) c->mesh.btype = i;
    if (c->mesh.btype < 0) return fail(c, MPLB_E_MODEL, "Unrecognized boundary type " + bt);
    if (c->mesh.d == 2 && c->mesh.btype > BOX_PERIODIC_XY) return fail(c, MPLB_E_MODEL, "Unrecognized boundary type " + bt);
    if (c->mesh.d == 3 && (c->mesh.btype == BOX_PERIODIC_X || c->mesh.btype == BOX_PERIODIC_Y))
        return fail(c, MPLB_E_MODEL, "Unrecognized boundary type " + bt);
    double spot_r = 0, spot_c[2] = {0, 0};
    if (c->mesh.btype == BOX_PERIODIC_XY_SPOT) {
        lp.scalar("spot/radius", 0, &spot_r);
        std::vector<double> v;
        if (lp.vector("spot/centre", &v) && v.size() >= 2) { spot_c[0] = v[0]; spot_c[1] = v[1]; }
    }
    // the reference solver reads the lattice written by latticeBox + latticeMeshPartition
    // (readLatticeMesh.c); without those files the box is built from latticeProperties
    const std::string merr =
        lattice_files_present(c->dir) ? read_lattice(c->mesh, c->dir) : build_box(c->mesh, spot_r, spot_c);
    if (!merr.empty()) return fail(c, MPLB_E_MODEL, merr);
    for (auto& l : c->mesh.bd)
        for (int32_t id : l)
            if (id < 0 || id >= c->mesh.N) return fail(c, MPLB_E_MODEL, "Boundary list leaves the lattice for this box shape");
    // readTimeInfo.c
    double aux;
    sim.scalar("startTime", 0, &aux);
    c->start = (int)aux;
    sim.scalar("endTime", 0, &aux);
    c->end = (int)aux;
    sim.scalar("writeInterval", 1, &aux);
    c->write_interval = (int)aux;
    sim.string("dataFormat", "asciiRaw", &c->data_format);
    c->current = c->start;
    c->stp = 0;
    int rc = read_eos(c, mp);
    if (rc) return rc;
    // f and g (readLbeField.c); a missing g block disables the energy field
    const char* fn[2] = {"f", "g"};
    const bool has_bd = bd.open(d + "start/boundaries");
    for (int i = 0; i < 2; i++) {
        std::string col;
        if (!mp.string(std::string(fn[i]) + "/Collision", "", &col)) {
            if (i == 0) return fail(c, MPLB_E_IO, "Collision model not found");
            c->fc[i].present = false;
            continue;
        }
        c->fc[i].present = true;
        rc = read_model(c, mp, fn[i], c->fc[i]);
        if (rc) return rc;
        if (!c->mesh.bd.empty() && !has_bd) return fail(c, MPLB_E_IO, "Unable to open start/boundaries");
        rc = read_boundaries(c, bd, fn[i], c->fc[i]);
        if (rc) return rc;
    }
    return 0;
}

// Model / lattice combinations the reference cannot run.
int check_models(mplb_case* c) {
    const int Q = c->mesh.Q;
    FieldCfg& f = c->fc[0];
    if (f.mp.model == M_LIMRT && Q == 15) {
        if (c->opts.extensions & MPLB_EXT_LIMRT_D3Q15) f.mp.model = M_XUMRT;
        else return fail(c, MPLB_E_MODEL, "MRT force squeme not implemented for D3Q15");
    }
    if (f.mp.model == M_XUMRT && Q == 9) return fail(c, MPLB_E_MODEL, "MRT force squeme not implemented for D2Q9");
    if (f.mp.model != M_LIMRT && f.mp.model != M_XUMRT) return fail(c, MPLB_E_MODEL, "Density update not available");
    if (c->fc[1].present && c->update[1]) {
        const int gm = c->fc[1].mp.model;
        if (gm != M_MYMRT && gm != M_LITEMPMRT) return fail(c, MPLB_E_MODEL, "Temperature update not available");
        if (gm == M_LITEMPMRT && Q == 15) return fail(c, MPLB_E_MODEL, "Lattice model not available for liTempMRTModel");
    }
    for (int i = 0; i < 2; i++) {
        if (!c->fc[i].present || !c->update[i]) continue;
        const int m = c->fc[i].mp.model;
        for (size_t b = 0; b < c->fc[i].bctype.size(); b++) {
            const int t = c->fc[i].bctype[b];
            if (t < 0) return fail(c, MPLB_E_MODEL, "Unrecognized boundary condition");
            if (t == BD_FGT && !(c->opts.extensions & MPLB_EXT_NO_FIXEDGRADT_ERROR))
                return fail(c, MPLB_E_MODEL, "Unrecognized boundary condition");
            if (t == BD_FT && m != M_MYMRT && m != M_LITEMPMRT) return fail(c, MPLB_E_MODEL, "Unable to apply boundary condition on temperature");
            if (t == BD_FRDT && m != M_MYMRT) return fail(c, MPLB_E_MODEL, "Unable to apply boundary condition on temperature");
            if (t == BD_OF) {
                // outflow.c:58-74: every node needs a detectable normal
                const Mesh& ms = c->mesh;
                for (int32_t id : ms.bd[b]) {
                    float nrm[3] = {0, 0, 0};
                    for (int k = 0; k < ms.Q; k++) {
                        const int rk = ms.lbmodel == 0 ? D2Q9::t.rev[k] : D3Q15::t.rev[k];
                        if (ms.nb[(int64_t)id * ms.Q + rk] == -1)
                            for (int j = 0; j < 3; j++) nrm[j] += ms.lbmodel == 0 ? D2Q9::t.c[k][j] : D3Q15::t.c[k][j];
                    }
                    if (nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2] == 0)
                        return fail(c, MPLB_E_MODEL, "Unable to detect normal in outflow boundary");
                }
            }
        }
    }
    return 0;
}

int reverse(const mplb_case* c, int k) { return c->mesh.lbmodel == 0 ? D2Q9::t.rev[k] : D3Q15::t.rev[k]; }

// A list must be applied in order when one of its nodes reads a population that another
// node of the same list writes (or a node repeats).
bool list_has_hazard(const mplb_case* c, int b, int type) {
    const Mesh& m = c->mesh;
    const auto& L = m.bd[b];
    std::unordered_set<int32_t> in(L.begin(), L.end());
    if (in.size() != L.size()) return true;
    auto missing = [&](int64_t id, int k) { return m.nb[id * m.Q + k] == -1; };
    auto any_missing = [&](int64_t id) {
        for (int k = 0; k < m.Q; k++)
            if (missing(id, k)) return true;
        return false;
    };
    for (int32_t id : L) {
        for (int k = 0; k < m.Q; k++) {
            if (!missing(id, k)) continue;
            const int32_t nbid = m.nb[(int64_t)id * m.Q + reverse(c, k)];
            if (nbid == -1 || !in.count(nbid)) continue;
            if (type == BD_FT || type == BD_FRDT) {
                if (missing(nbid, k)) return true;
            } else if (type == BD_FU) {
                if (any_missing(nbid)) return true;
            }
        }
        if (type == BD_OF) return true;   // neighbour along the normal: checked conservatively
    }
    return false;
}

int setup_device(mplb_case* c) {
    const Mesh& m = c->mesh;
    const int64_t N = m.N;
    const int Q = m.Q;
    HIPCHK(c, hipSetDevice(c->opts.device));
    HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    {
        const int rc = warm_scratch_setup(c);
        if (rc) return rc;
    }
    c->scalar_bytes = c->opts.precision == MPLB_FP64 ? 8 : 4;
    const size_t sb = c->scalar_bytes;
    for (int i = 0; i < 2; i++) {
        HIPCHK(c, hipMalloc(&c->f[i], sb * Q * N));
        HIPCHK(c, hipMemsetAsync(c->f[i], 0, sb * Q * N, c->stream));
        if (!c->fused) {
            HIPCHK(c, hipMalloc(&c->f_alt[i], sb * Q * N));
        }
    }
    void** scal[] = {&c->rho, &c->temp, &c->p, &c->ss, &c->psi};
    for (void** pp : scal) {
        HIPCHK(c, hipMalloc(pp, sb * N));
        HIPCHK(c, hipMemsetAsync(*pp, 0, sb * N, c->stream));
    }
    void** vec[] = {&c->U, &c->Fi, &c->Ft};
    for (void** pp : vec) {
        HIPCHK(c, hipMalloc(pp, sb * 3 * N));
        HIPCHK(c, hipMemsetAsync(*pp, 0, sb * 3 * N, c->stream));
    }
    // neighbour table [Q][N]
    std::vector<int32_t> nbt((size_t)Q * N);
    for (int64_t id = 0; id < N; id++)
        for (int k = 0; k < Q; k++) nbt[(size_t)k * N + id] = m.nb[id * Q + k];
    HIPCHK(c, hipMalloc(&c->nb, sizeof(int32_t) * Q * N));
    HIPCHK(c, hipMemcpy(c->nb, nbt.data(), sizeof(int32_t) * Q * N, hipMemcpyHostToDevice));
    HIPCHK(c, hipMalloc(&c->onbnd, N));
    HIPCHK(c, hipMemcpy(c->onbnd, m.onbnd.data(), N, hipMemcpyHostToDevice));
    // boundary lists, hazard flags and fixedRandomT draw tables
    c->bd.resize(m.bd.size());
    int max_draws = 0;
    for (size_t b = 0; b < m.bd.size(); b++) {
        DevBd& db = c->bd[b];
        db.n = (int)m.bd[b].size();
        if (db.n) {
            HIPCHK(c, hipMalloc(&db.nodes, sizeof(int) * db.n));
            HIPCHK(c, hipMemcpy(db.nodes, m.bd[b].data(), sizeof(int) * db.n, hipMemcpyHostToDevice));
        }
        for (int f = 0; f < 2; f++) {
            if (!c->fc[f].present) continue;
            const int t = c->fc[f].bctype[b];
            db.sequential[f] = list_has_hazard(c, (int)b, t) ? 1 : 0;
            if (t == BD_FRDT && db.n) {
                std::vector<int> off(db.n);
                int cnt = 0;
                for (int i = 0; i < db.n; i++) {
                    off[i] = cnt;
                    const int64_t id = m.bd[b][i];
                    for (int k = 0; k < Q; k++)
                        if (m.nb[id * Q + k] == -1 && m.nb[id * Q + reverse(c, k)] != -1) cnt++;
                }
                if (cnt > max_draws) max_draws = cnt;
                HIPCHK(c, hipMalloc(&db.rnd_off[f], sizeof(int) * db.n));
                HIPCHK(c, hipMemcpy(db.rnd_off[f], off.data(), sizeof(int) * db.n, hipMemcpyHostToDevice));
            }
        }
    }
    if (max_draws > 0) {
        // fixedRandomT.c:37,81: srand(pid) on every call, r = rand() / RAND_MAX per link
        std::vector<int32_t> rv(max_draws);
        glibc_rand_fill(0, max_draws, rv.data());
        std::vector<double> r(max_draws);
        for (int i = 0; i < max_draws; i++) r[i] = (double)rv[i] / 2147483647;
        HIPCHK(c, hipMalloc(&c->rnd, sizeof(double) * max_draws));
        HIPCHK(c, hipMemcpy(c->rnd, r.data(), sizeof(double) * max_draws, hipMemcpyHostToDevice));
    }
    HIPCHK(c, hipMalloc(&c->red, sizeof(double) * 4096));
    HIPCHK(c, hipMalloc(&c->nanc, sizeof(unsigned int) * 8));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

template <typename T>
DevState<T> state(mplb_case* c) {
    DevState<T> s;
    s.N = c->mesh.N;
    s.nb = c->nb;
    s.onbnd = c->onbnd;
    s.rho = (T*)c->rho;
    s.temp = (T*)c->temp;
    s.p = (T*)c->p;
    s.U = (T*)c->U;
    s.Fi = (T*)c->Fi;
    s.Ft = (T*)c->Ft;
    s.ss = (T*)c->ss;
    s.psi = (T*)c->psi;
    return s;
}

// ---------------------------------------------------------------------------
// timing
struct TimedRegion {
    mplb_case* c;
    int idx;
    hipEvent_t a = nullptr, b = nullptr;
    TimedRegion(mplb_case* cc, int i) : c(cc), idx(i) {
        if (!c->timing) return;
        // timing only: a failed event leaves this region out of the per-kernel times
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
        (void)hipEventRecord(a, c->stream);
    }
    ~TimedRegion() {
        if (!c->timing) return;
        (void)hipEventRecord(b, c->stream);
        c->ev_pool.push_back(a);
        c->ev_pool.push_back(b);
        c->ev_pool.push_back((hipEvent_t)(intptr_t)idx);
    }
};

const char* kTimerNames[MPLB_NTIMERS] = {"collide", "stream", "boundary", "macro", "force", "heat_source", "fused", "fused_aux"};
enum { TM_COLLIDE = 0, TM_STREAM, TM_BOUNDARY, TM_MACRO, TM_FORCE, TM_HEAT, TM_FUSED, TM_FUSED_AUX };

void flush_timers(mplb_case* c) {
    if (c->timers.empty()) {
        for (int i = 0; i < MPLB_NTIMERS; i++) c->timers.push_back(Timer{kTimerNames[i], 0, 0});
    }
    (void)hipStreamSynchronize(c->stream);
    for (size_t i = 0; i + 2 < c->ev_pool.size(); i += 3) {
        float ms = 0;
        const int idx = (int)(intptr_t)c->ev_pool[i + 2];
        if (hipEventElapsedTime(&ms, c->ev_pool[i], c->ev_pool[i + 1]) == hipSuccess) {
            c->timers[idx].ms += ms;
            c->timers[idx].launches++;
        }
        (void)hipEventDestroy(c->ev_pool[i]);
        (void)hipEventDestroy(c->ev_pool[i + 1]);
    }
    c->ev_pool.clear();
}

// ---------------------------------------------------------------------------
// generic (table-driven) step pieces
template <typename T>
void ensure_psi(mplb_case* c) {
    if (c->psi_valid) return;
    GenericOps<T>::psi(c->mesh.lbmodel, c->stream, c->eos, state<T>(c));
    c->psi_valid = true;
}

template <typename T>
int collision_t(mplb_case* c, int field) {
    if (!c->fc[field].present || !c->update[field]) return 0;
    if (field == 0) ensure_psi<T>(c);
    TimedRegion tr(c, TM_COLLIDE);
    GenericOps<T>::collide(c->mesh.lbmodel, c->stream, c->eos, c->fc[field].mp, state<T>(c), (T*)c->f[field], c->rho0);
    return 0;
}

template <typename T>
int stream_t(mplb_case* c, int field) {
    if (!c->fc[field].present || !c->update[field]) return 0;
    TimedRegion tr(c, TM_STREAM);
    GenericOps<T>::stream(c->mesh.lbmodel, c->stream, c->mesh.N, c->nb, (const T*)c->f[field], (T*)c->f_alt[field]);
    std::swap(c->f[field], c->f_alt[field]);
    return 0;
}

template <typename T>
int boundaries_t(mplb_case* c, int field) {
    if (!c->fc[field].present || !c->update[field]) return 0;
    FieldCfg& fc = c->fc[field];
    TimedRegion tr(c, TM_BOUNDARY);
    for (size_t b = 0; b < c->bd.size(); b++) {
        const int t = fc.bctype[b];
        if (t == BD_PERIODIC || t == BD_FEP || t == BD_UHF || t == BD_FGT) continue;
        if (t == BD_FU && fc.mp.model != M_LIMRT) continue;   // fixedU.c:10: liMRT only
        BdList bl{};
        bl.nodes = c->bd[b].nodes;
        bl.n = c->bd[b].n;
        bl.type = t;
        bl.sequential = c->bd[b].sequential[field];
        for (int j = 0; j < 3; j++) bl.value[j] = fc.bcval[b][j];
        bl.pert = fc.pert[b];
        bl.rnd = c->rnd;
        bl.rnd_off = c->bd[b].rnd_off[field];
        bl.draw = (t == BD_FRDT) ? (fc.localT[b] < fc.psteps[b] ? 1 : 0) : 0;
        GenericOps<T>::boundary(c->mesh.lbmodel, c->stream, c->eos, fc.mp, state<T>(c), (T*)c->f[field], bl, c->rho0);
        if (t == BD_FRDT) fc.localT[b]++;
    }
    return 0;
}

template <typename T>
int macro_density_t(mplb_case* c) {
    if (!c->update[0]) return 0;
    TimedRegion tr(c, TM_MACRO);
    GenericOps<T>::macro_density(c->mesh.lbmodel, c->stream, state<T>(c), (const T*)c->f[0]);
    c->psi_valid = false;
    return 0;
}

template <typename T>
int macro_velocity_t(mplb_case* c) {
    if (!c->update[0]) return 0;
    TimedRegion tr(c, TM_MACRO);
    GenericOps<T>::macro_velocity(c->mesh.lbmodel, c->stream, c->eos, c->fc[0].mp.model, state<T>(c), (const T*)c->f[0],
                                  c->rho0);
    return 0;
}

template <typename T>
int macro_temperature_t(mplb_case* c) {
    if (!c->fc[1].present || !c->update[1]) return 0;
    TimedRegion tr(c, TM_MACRO);
    GenericOps<T>::macro_temperature(c->mesh.lbmodel, c->stream, c->fc[1].mp.model, state<T>(c), (const T*)c->f[1]);
    c->psi_valid = false;
    return 0;
}

template <typename T>
int inter_force_t(mplb_case* c) {
    TimedRegion tr(c, TM_FORCE);
    ensure_psi<T>(c);
    GenericOps<T>::inter_force(c->mesh.lbmodel, c->stream, c->eos, state<T>(c), c->rho0);
    return 0;
}

template <typename T>
int heat_source_t(mplb_case* c) {
    if (!c->fc[1].present || !c->update[1]) return 0;
    TimedRegion tr(c, TM_HEAT);
    GenericOps<T>::heat_source(c->mesh.lbmodel, c->stream, c->eos, c->fc[1].mp, state<T>(c), c->current == 0 ? 1 : 0);
    return 0;
}

template <typename T>
int equilibrium_t(mplb_case* c, int field) {
    if (!c->fc[field].present || !c->update[field]) return 0;
    GenericOps<T>::equilibrium(c->mesh.lbmodel, c->stream, c->fc[field].mp, state<T>(c), (T*)c->f[field]);
    return 0;
}

template <typename T>
int pressure_t(mplb_case* c) {
    if (!c->update[0]) return 0;
    c->psi_valid = false;
    ensure_psi<T>(c);
    GenericOps<T>::pressure(c->mesh.lbmodel, c->stream, c->eos, c->fc[0].mp.model, c->fc[0].mp.sigma, state<T>(c));
    return 0;
}

template <typename T>
int average_t(mplb_case* c, double* out) {
    // averageScalarField.c:10-72
    if (c->rho0 <= 0) {
        *out = 0;
        return 0;
    }
    GenericOps<T>::sum(c->stream, (const T*)c->rho, c->mesh.N, c->red, c->opts.precision == MPLB_FP64 ? 1 : 0, c->red + 1);
    double s = 0;
    HIPCHK(c, hipMemcpyAsync(&s, c->red, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *out = s / (double)(int)c->mesh.N;
    return 0;
}

template <typename T>
int nan_check_t(mplb_case* c) {
    HIPCHK(c, hipMemsetAsync(c->nanc, 0, sizeof(unsigned int), c->stream));
    const int64_t N = c->mesh.N, Q = c->mesh.Q;
    GenericOps<T>::nan_count(c->stream, (const T*)c->rho, N, c->nanc);
    GenericOps<T>::nan_count(c->stream, (const T*)c->temp, N, c->nanc);
    GenericOps<T>::nan_count(c->stream, (const T*)c->p, N, c->nanc);
    GenericOps<T>::nan_count(c->stream, (const T*)c->U, 3 * N, c->nanc);
    GenericOps<T>::nan_count(c->stream, (const T*)c->f[0], Q * N, c->nanc);
    GenericOps<T>::nan_count(c->stream, (const T*)c->f[1], Q * N, c->nanc);
    unsigned int n = 0;
    HIPCHK(c, hipMemcpyAsync(&n, c->nanc, sizeof(unsigned int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (n) return fail(c, MPLB_E_NAN, "Floating point exception. NaN solution");
    return 0;
}

// IO/time/timeOptions.c:35-64
bool update_time(mplb_case* c) {
    bool upd = true;
    ++c->current;
    ++c->stp;
    if (c->current > c->end) upd = false;
    if (c->stp > c->write_interval) c->stp = 1;
    return upd;
}

template <typename T>
int write_step_t(mplb_case* c) {
    // multiPhasePP.c:249-315 (file output is left to the caller through mplb_get_field)
    int rc = pressure_t<T>(c);
    if (rc) return rc;
    double avg = 0;
    rc = average_t<T>(c, &avg);
    if (rc) return rc;
    c->rho0 = avg;
    return nan_check_t<T>(c);
}

// multiPhasePP.c:157-241
template <typename T>
int one_step_t(mplb_case* c) {
    if (c->fused) {
        int rc = fused_step(c);
        if (rc) return rc;
    } else {
        collision_t<T>(c, 0);
        stream_t<T>(c, 0);
        boundaries_t<T>(c, 0);
        collision_t<T>(c, 1);
        stream_t<T>(c, 1);
        boundaries_t<T>(c, 1);
        macro_density_t<T>(c);
        if (c->opts.heat_transfer) heat_source_t<T>(c);
        macro_temperature_t<T>(c);
        if (!c->opts.frozen) inter_force_t<T>(c);
        macro_velocity_t<T>(c);
    }
    if (c->stp == c->write_interval) {
        if (c->fused) {
            int rc = fused_sync_to_generic(c);
            if (rc) return rc;
        }
        int rc = write_step_t<T>(c);
        if (rc) return rc;
        if (c->fused) {
            rc = fused_after_write(c);
            if (rc) return rc;
        }
    }
    return 0;
}

template <typename T>
int steps_t(mplb_case* c, int nsteps, int* done) {
    int n = 0;
    while (n < nsteps && update_time(c)) {
        int rc = one_step_t<T>(c);
        n++;
        if (rc) {
            if (done) *done = n;
            return rc;
        }
    }
    if (done) *done = n;
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(c, MPLB_E_HIP, std::string("step: ") + hipGetErrorString(e));
    e = hipGetLastError();
    if (e != hipSuccess) return fail(c, MPLB_E_HIP, std::string("launch: ") + hipGetErrorString(e));
    return 0;
}

template <typename T>
int init_t(mplb_case* c) {
    double avg;
    int rc = average_t<T>(c, &avg);   // multiPhasePP.c:74
    if (rc) return rc;
    c->rho0 = avg;
    equilibrium_t<T>(c, 0);
    equilibrium_t<T>(c, 1);
    c->psi_valid = false;
    inter_force_t<T>(c);
    heat_source_t<T>(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->initialized = true;
    c->snap_valid = false;
    if (c->fused) return fused_after_init(c);
    return 0;
}

int field_layout(const mplb_case* c, const std::string& name, void** dptr, int* comps) {
    const int Q = c->mesh.Q;
    void* const* fp = c->f;
    if (name == "rho") { *dptr = c->rho; *comps = 1; return 0; }
    if (name == "T") { *dptr = c->temp; *comps = 1; return 0; }
    if (name == "p") { *dptr = c->p; *comps = 1; return 0; }
    if (name == "ss") { *dptr = c->ss; *comps = 1; return 0; }
    if (name == "U") { *dptr = c->U; *comps = 3; return 0; }
    if (name == "Fi") { *dptr = c->Fi; *comps = 3; return 0; }
    if (name == "Ft") { *dptr = c->Ft; *comps = 3; return 0; }
    if (name == "f") { *dptr = fp[0]; *comps = Q; return 0; }
    if (name == "g") { *dptr = fp[1]; *comps = Q; return 0; }
    return -1;
}

template <typename T>
int copy_field(mplb_case* c, const std::string& name, double* host, size_t n, bool to_device) {
    void* d;
    int comps;
    if (field_layout(c, name, &d, &comps)) return fail(c, MPLB_E_ARG, "Unknown field " + name);
    const int64_t N = c->mesh.N;
    if (n != (size_t)(N * comps)) return fail(c, MPLB_E_ARG, "Field " + name + ": wrong element count");
    std::vector<T> buf((size_t)N * comps);
    if (to_device) {
        for (int64_t i = 0; i < N; i++)
            for (int j = 0; j < comps; j++) buf[(size_t)j * N + i] = (T)host[(size_t)i * comps + j];
        HIPCHK(c, hipMemcpyAsync(d, buf.data(), sizeof(T) * buf.size(), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (name == "rho" || name == "T") c->psi_valid = false;
    } else {
        HIPCHK(c, hipMemcpyAsync(buf.data(), d, sizeof(T) * buf.size(), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int64_t i = 0; i < N; i++)
            for (int j = 0; j < comps; j++) host[(size_t)i * comps + j] = (double)buf[(size_t)j * N + i];
    }
    return 0;
}

#define DISPATCH_T(c, expr)                                  \
    ((c)->opts.precision == MPLB_FP64 ? [&]() { using T = double; return (expr); }() \
                                      : [&]() { using T = float; return (expr); }())

int check_open(mplb_case* c) {
    if (!c) return MPLB_E_ARG;
    c->err.clear();
    return 0;
}

}  // namespace

// ===========================================================================
extern "C" {

void mplb_default_opts(mplb_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->precision = MPLB_FP64;
    o->heat_transfer = 1;
    o->frozen = 0;
    o->fused = 0;
    o->extensions = 0;
    o->device = 0;
}

int mplb_case_open(const char* case_dir, const mplb_opts* opts, mplb_case** out) {
    if (!case_dir || !out) return MPLB_E_ARG;
    *out = nullptr;
    mplb_case* c = new mplb_case();
    if (opts) c->opts = *opts;
    else mplb_default_opts(&c->opts);
    c->dir = case_dir;
    *out = c;
    if (c->opts.precision < MPLB_FP64 || c->opts.precision > MPLB_FP16N)
        return fail(c, MPLB_E_ARG, "unknown precision " + std::to_string(c->opts.precision));
    int rc = parse_case(c);
    if (rc) return rc;
    // multiPhasePP.c:106,116 (--frozen / --nht disable the f / g updates)
    c->update[0] = c->opts.frozen ? 0 : 1;
    c->update[1] = (c->opts.heat_transfer && c->fc[1].present) ? 1 : 0;
    rc = check_models(c);
    if (rc) return rc;
    if (c->opts.device < 0) return 0;   // host-only: dictionaries and lattice tables, no GPU
    const bool half = c->opts.precision == MPLB_FP16S || c->opts.precision == MPLB_FP16N;
    if (half && !c->opts.fused)
        return fail(c, MPLB_E_ARG, "FP16 storage is only available with the fused kernels");
    if (c->opts.fused) {
        std::string why;
        if (!fused_supported(c, &why)) {
            if (half) return fail(c, MPLB_E_ARG, "fused kernels unavailable: " + why);
            c->fused = false;
        } else {
            c->fused = true;
        }
    }
    rc = setup_device(c);
    if (rc) return rc;
    if (c->fused) {
        rc = fused_create(c);
        if (rc) return rc;
    }
    return 0;
}

int mplb_case_close(mplb_case* c) {
    if (!c) return MPLB_E_ARG;
    // teardown: failures here have nobody left to report to
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->fused) fused_destroy(c);
    void* ptrs[] = {c->f[0], c->f[1], c->f_alt[0], c->f_alt[1], c->rho, c->temp, c->p, c->U, c->Fi, c->Ft, c->ss, c->psi,
                    c->nb, c->onbnd, c->rnd, c->red, c->nanc, c->rho_snap};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (auto& b : c->bd) {
        if (b.nodes) (void)hipFree(b.nodes);
        for (int f = 0; f < 2; f++)
            if (b.rnd_off[f]) (void)hipFree(b.rnd_off[f]);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

const char* mplb_last_error(const mplb_case* c) { return c ? c->err.c_str() : "null case"; }

int mplb_case_info(const mplb_case* c, mplb_info* o) {
    if (!c || !o) return MPLB_E_ARG;
    std::memset(o, 0, sizeof(*o));
    o->lbmodel = c->mesh.lbmodel;
    o->Q = c->mesh.Q;
    o->d = c->mesh.d;
    o->nx = c->mesh.nx;
    o->ny = c->mesh.ny;
    o->nz = c->mesh.nz;
    o->npoints = c->mesh.N;
    o->nbd = (int)c->mesh.bd.size();
    o->f_model = c->fc[0].present ? c->fc[0].mp.model : -1;
    o->g_model = c->fc[1].present ? c->fc[1].mp.model : -1;
    o->start_time = c->start;
    o->end_time = c->end;
    o->write_interval = c->write_interval;
    o->current_time = c->current;
    o->precision = c->opts.precision;
    o->fused_active = c->fused ? 1 : 0;
    o->rho0 = c->rho0;
    return 0;
}

int mplb_case_param(const mplb_case* c, const char* key, double* out, size_t n) {
    if (!c || !key) return MPLB_E_ARG;
    std::vector<double> v;
    const std::string k = key;
    const EosParams& e = c->eos;
    auto field = [&](const std::string& name) -> const FieldCfg* {
        if (name == "f") return &c->fc[0];
        if (name == "g") return c->fc[1].present ? &c->fc[1] : nullptr;
        return nullptr;
    };
    if (k == "EOS/model") v = {(double)e.type};
    else if (k == "EOS/a") v = {e.a};
    else if (k == "EOS/b") v = {e.b};
    else if (k == "EOS/w") v = {e.w_pr};
    else if (k == "EOS/R") v = {e.R};
    else if (k == "EOS/G") v = {e.G};
    else if (k == "EOS/M") v = {e.M};
    else if (k == "EOS/Cv") v = {e.Cv};
    else if (k == "EOS/rho_0") v = {c->rho0_dict};
    else if (k == "EOS/g") v = {e.g[0], e.g[1], e.g[2]};
    else if (k == "EOS/extForce") v = {e.fext[0], e.fext[1], e.fext[2]};
    else if (k == "time/start") v = {(double)c->start};
    else if (k == "time/end") v = {(double)c->end};
    else if (k == "time/writeInterval") v = {(double)c->write_interval};
    else {
        const size_t sl = k.find('/');
        const FieldCfg* fc = sl == std::string::npos ? nullptr : field(k.substr(0, sl));
        if (!fc) return MPLB_E_ARG;
        const std::string rest = k.substr(sl + 1);
        if (rest == "Collision") v = {(double)fc->mp.model};
        else if (rest == "Lambda") v.assign(fc->mp.Lambda, fc->mp.Lambda + c->mesh.Q);
        else if (rest == "sigma") v = {fc->mp.sigma};
        else if (rest == "SurfaceTension") v = {(double)fc->mp.surf};
        else if (rest == "kappa_st") v = {fc->mp.kappa};
        else if (rest == "alpha_1") v = {fc->mp.alpha1};
        else if (rest == "alpha_2") v = {fc->mp.alpha2};
        else if (rest.compare(0, 3, "bc/") == 0) {
            // bc/<list index>/<type|value|pert|duration>
            const size_t s2 = rest.find('/', 3);
            if (s2 == std::string::npos) return MPLB_E_ARG;
            const int b = std::atoi(rest.substr(3, s2 - 3).c_str());
            if (b < 0 || b >= (int)fc->bctype.size()) return MPLB_E_ARG;
            const std::string what = rest.substr(s2 + 1);
            if (what == "type") v = {(double)fc->bctype[b]};
            else if (what == "value") v = {fc->bcval[b][0], fc->bcval[b][1], fc->bcval[b][2]};
            else if (what == "pert") v = {fc->pert[b]};
            else if (what == "duration") v = {(double)fc->psteps[b]};
            else return MPLB_E_ARG;
        } else {
            return MPLB_E_ARG;
        }
    }
    for (size_t i = 0; i < v.size() && i < n && out; i++) out[i] = v[i];
    return (int)v.size();
}

int mplb_set_field(mplb_case* c, const char* name, const double* host, size_t n) {
    if (check_open(c) || !name || !host) return MPLB_E_ARG;
    if (!c->stream) return fail(c, MPLB_E_STATE, "case was opened without a device");
    if (c->fused) {
        int rc = fused_set_field(c, name, host, n);
        if (rc != 1) return rc;
    }
    return DISPATCH_T(c, copy_field<T>(c, name, const_cast<double*>(host), n, true));
}

int mplb_get_field(mplb_case* c, const char* name, double* host, size_t n) {
    if (check_open(c) || !name || !host) return MPLB_E_ARG;
    if (!c->stream) return fail(c, MPLB_E_STATE, "case was opened without a device");
    if (c->fused) {
        int rc = fused_get_field(c, name, host, n);
        if (rc != 1) return rc;
    }
    return DISPATCH_T(c, copy_field<T>(c, name, host, n, false));
}

int mplb_get_neighbours(const mplb_case* c, int32_t* nb, size_t n) {
    if (!c || !nb || n != (size_t)(c->mesh.N * c->mesh.Q)) return MPLB_E_ARG;
    std::memcpy(nb, c->mesh.nb.data(), sizeof(int32_t) * n);
    return 0;
}

int mplb_get_boundary(const mplb_case* c, int bid, int32_t* nodes, size_t n, char* name, size_t name_len) {
    if (!c || bid < 0 || bid >= (int)c->mesh.bd.size()) return MPLB_E_ARG;
    const auto& l = c->mesh.bd[bid];
    if (nodes) {
        if (n != l.size()) return MPLB_E_ARG;
        std::memcpy(nodes, l.data(), sizeof(int32_t) * n);
    }
    if (name && name_len) std::snprintf(name, name_len, "%s", c->mesh.bdnames[bid].c_str());
    return (int)l.size();
}

int mplb_init(mplb_case* c) {
    if (check_open(c)) return MPLB_E_ARG;
    if (!c->stream) return fail(c, MPLB_E_STATE, "case was opened without a device");
    return DISPATCH_T(c, init_t<T>(c));
}

int mplb_step(mplb_case* c, int nsteps, int* steps_done) {
    if (check_open(c)) return MPLB_E_ARG;
    if (!c->initialized) return fail(c, MPLB_E_STATE, "mplb_init must run before mplb_step");
    return DISPATCH_T(c, steps_t<T>(c, nsteps, steps_done));
}

#define FINE(c, expr)                                                                   \
    if (check_open(c)) return MPLB_E_ARG;                                               \
    if (!(c)->stream) return fail((c), MPLB_E_STATE, "case was opened without a device"); \
    if ((c)->fused) return fail((c), MPLB_E_STATE, "fine-grained calls need fused = 0"); \
    {                                                                                   \
        int _rc = DISPATCH_T(c, expr);                                                  \
        if (_rc) return _rc;                                                            \
        hipError_t _e = hipStreamSynchronize((c)->stream);                              \
        if (_e != hipSuccess) return fail((c), MPLB_E_HIP, hipGetErrorString(_e));      \
        return 0;                                                                       \
    }

int mplb_collision(mplb_case* c, int field) { FINE(c, collision_t<T>(c, field & 1)); }
int mplb_stream(mplb_case* c, int field) { FINE(c, stream_t<T>(c, field & 1)); }
int mplb_update_boundaries(mplb_case* c, int field) { FINE(c, boundaries_t<T>(c, field & 1)); }
int mplb_macro_density(mplb_case* c) { FINE(c, macro_density_t<T>(c)); }
int mplb_macro_velocity(mplb_case* c) { FINE(c, macro_velocity_t<T>(c)); }
int mplb_macro_temperature(mplb_case* c) { FINE(c, macro_temperature_t<T>(c)); }
int mplb_inter_force(mplb_case* c) { FINE(c, inter_force_t<T>(c)); }
int mplb_heat_source(mplb_case* c) { FINE(c, heat_source_t<T>(c)); }
int mplb_equilibrium(mplb_case* c, int field) { FINE(c, equilibrium_t<T>(c, field & 1)); }
int mplb_macro_pressure(mplb_case* c) { FINE(c, pressure_t<T>(c)); }

int mplb_average_density(mplb_case* c, double* out) {
    if (check_open(c) || !out) return MPLB_E_ARG;
    if (!c->stream) return fail(c, MPLB_E_STATE, "case was opened without a device");
    if (c->fused) {
        int rc = fused_sync_to_generic(c);
        if (rc) return rc;
    }
    return DISPATCH_T(c, average_t<T>(c, out));
}

int mplb_time_steps(mplb_case* c, int nsteps, double* total_ms, double* kernel_ms, int* kernel_launches) {
    if (check_open(c)) return MPLB_E_ARG;
    if (!c->stream) return fail(c, MPLB_E_STATE, "case was opened without a device");
    if (!c->initialized) return fail(c, MPLB_E_STATE, "mplb_init must run before mplb_time_steps");
    c->timers.clear();
    for (int i = 0; i < MPLB_NTIMERS; i++) c->timers.push_back(Timer{kTimerNames[i], 0, 0});
    hipEvent_t a, b;
    HIPCHK(c, hipEventCreate(&a));
    HIPCHK(c, hipEventCreate(&b));
    const bool per_kernel = kernel_ms != nullptr;
    c->timing = per_kernel;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipEventRecord(a, c->stream));
    int done = 0;
    int rc = DISPATCH_T(c, steps_t<T>(c, nsteps, &done));
    hipError_t he = hipEventRecord(b, c->stream);
    if (he == hipSuccess) he = hipEventSynchronize(b);
    float ms = 0;
    if (he == hipSuccess) he = hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (per_kernel) flush_timers(c);
    c->timing = false;
    if (rc) return rc;
    if (he != hipSuccess) return fail(c, MPLB_E_HIP, std::string("timing events: ") + hipGetErrorString(he));
    if (total_ms) *total_ms = ms;
    for (int i = 0; i < MPLB_NTIMERS; i++) {
        if (kernel_ms) kernel_ms[i] = c->timers[i].launches ? c->timers[i].ms / c->timers[i].launches : 0.0;
        if (kernel_launches) kernel_launches[i] = c->timers[i].launches;
    }
    return done;
}

const char* mplb_timer_name(int i) { return (i >= 0 && i < MPLB_NTIMERS) ? kTimerNames[i] : ""; }

int mplb_lattice_tables(int lbmodel, int* vel, int* rev, double* omega, double* w, double* M, double* invM) {
    auto fill = [&](auto L) {
        using LL = decltype(L);
        const int Q = LL::Q;
        for (int k = 0; k < Q; k++) {
            for (int j = 0; j < 3; j++)
                if (vel) vel[k * 3 + j] = LL::t.c[k][j];
            if (rev) rev[k] = LL::t.rev[k];
            if (omega) omega[k] = LL::t.omega[k];
            if (w) w[k] = LL::t.w[k];
            for (int j = 0; j < Q; j++) {
                if (M) M[k * Q + j] = LL::t.M[k][j];
                if (invM) invM[k * Q + j] = LL::t.invM[k][j];
            }
        }
        return Q;
    };
    if (lbmodel == 0) return fill(D2Q9{});
    if (lbmodel == 1) return fill(D3Q15{});
    return MPLB_E_ARG;
}

int mplb_glibc_rand(unsigned seed, int n, int32_t* out) {
    if (n < 0 || (!out && n)) return MPLB_E_ARG;
    glibc_rand_fill(seed, n, out);
    return 0;
}

}  // extern "C"
