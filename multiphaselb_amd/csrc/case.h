// Host-side state of one case: parsed dictionaries, lattice tables, device buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <string>
#include <vector>

#include "../../include/mplb.h"
#include "dict.h"
#include "kernels.h"
#include "mesh.h"

namespace mplb {

struct FieldCfg {
    bool present = false;
    ModelParams mp{};
    std::vector<int> bctype;                    // bdType per boundary list
    std::vector<std::array<double, 3>> bcval;   // value entry
    std::vector<double> pert;                   // fixedRandomT pert (percent)
    std::vector<int> psteps;                    // fixedRandomT duration
    std::vector<int> localT;                    // fixedRandomT localTStep
    std::vector<std::string> bcname;            // type string as read
};

struct DevBd {
    int* nodes = nullptr;
    int n = 0;
    int sequential[2] = {0, 0};                 // per field
    int* rnd_off[2] = {nullptr, nullptr};       // fixedRandomT draw offsets per field
};

struct Timer {
    const char* name;
    double ms = 0;
    int launches = 0;
};

}  // namespace mplb

struct mplb_case {
    mplb_opts opts{};
    std::string dir;
    std::string err;
    mplb::Mesh mesh;
    mplb::EosParams eos{};
    double rho0 = 0;
    double rho0_dict = 0;                       // EOS/rho_0 as the dictionary gives it
    mplb::FieldCfg fc[2];
    int update[2] = {1, 1};
    int start = 0, end = 0, write_interval = 1, current = 0, stp = 0;
    std::string data_format = "asciiRaw";
    bool initialized = false;
    bool psi_valid = false;
    int ss_first = 1;

    // device
    int dev = 0;
    hipStream_t stream = nullptr;
    size_t scalar_bytes = 8;
    void* f[2] = {nullptr, nullptr};
    void* f_alt[2] = {nullptr, nullptr};
    void* rho = nullptr;
    void* temp = nullptr;
    void* p = nullptr;
    void* U = nullptr;
    void* Fi = nullptr;
    void* Ft = nullptr;
    void* ss = nullptr;
    void* psi = nullptr;
    int* nb = nullptr;
    uint8_t* onbnd = nullptr;
    std::vector<mplb::DevBd> bd;
    double* rnd = nullptr;                      // fixedRandomT draws r = rand()/RAND_MAX
    double* red = nullptr;                      // reduction result + partials
    unsigned int* nanc = nullptr;
    void* rho_snap = nullptr;                   // density snapshot of mplb_density_change
    bool snap_valid = false;

    // fused structured-box path
    bool fused = false;
    void* fz = nullptr;                         // fused-state handle (kernels_fused.hip)

    // timers
    bool timing = false;
    std::vector<mplb::Timer> timers;
    std::vector<hipEvent_t> ev_pool;
};

namespace mplb {
// resolve the device address of warm.hip's allocation-free scratch array for the
// case's device (once per device; the device must be current)
int warm_scratch_setup(mplb_case* c);
}  // namespace mplb
