// Case-level helpers of the learned warm start and of runs to steady state, on the
// case's own device fields:
//   mplb_set_radial_density  rho(x) = profile(|x - centre| - R) (linear interpolation,
//                            clamped ends) written straight into the case's density,
//                            returning the mass of the distinct lattice sites (a
//                            periodic axis's last node duplicates its first and is not
//                            counted), so the host can pick R so the proposal carries
//                            the mass of the cold initial field;
//   mplb_density_change      max_x |rho(x) - rho_prev(x)| of the current time level
//                            against the previous call's snapshot (device reduction).
//   mplb_set_radial_density_noalloc
//                            the radial raster without any device allocation: the
//                            profile table, the per-block partial masses and the total
//                            live in this file's `scratch` array (one copy per GPU,
//                            held until the process ends), so callers that may not
//                            allocate (graph capture, allocation-free loops) can use it.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/mplb.h"
#include "case.h"
#include "fused.h"

using namespace mplb;

namespace {

constexpr int kBlock = 256;

// Scratch of the allocation-free raster: [0, kMaxProf) profile table, [kMaxProf,
// kMaxProf + kMaxPart) one partial mass per block, then the ordered total.  Limits:
// nprof <= kMaxProf and N <= kMaxPart * kBlock (2^24 nodes, a 256^3 box).  552 KiB per
// GPU.  Every call on a GPU shares that GPU's copy and leaves its contents behind;
// calls that overlap on one GPU (several host threads or streams) race.
constexpr int kMaxProf = 4096;
constexpr int64_t kMaxPart = 65536;
constexpr int64_t kScratchDoubles = kMaxProf + kMaxPart + 32;
static __device__ __align__(256) double scratch[kScratchDoubles];

// device address of `scratch` per device ordinal (the host-side name is only a stand-in)
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

int wfail(mplb_case* c, int code, const std::string& m) {
    if (c) c->err = m;
    return code;
}

#define WCHK(c, call)                                                                                           \
    do {                                                                                                        \
        hipError_t _e = (call);                                                                                 \
        if (_e != hipSuccess) return wfail((c), MPLB_E_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

struct Radial {
    int64_t N;
    int nx, ny, nz;
    double cx, cy, cz, R, d0, dd;
    const double* prof;
    int nprof;
    int dup;          // bit a set: the last node along axis a duplicates the first (periodic)
};

// one partial mass per block (FP64, fixed order); the host adds the partials in order
template <typename T>
__global__ void k_radial(Radial a, T* __restrict__ rho, double* __restrict__ part) {
    __shared__ double s[kBlock];
    const int64_t id = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    double v = 0.0;
    if (id < a.N) {
        const int64_t nxy = (int64_t)a.nx * a.ny;
        const int64_t k = id / nxy, r = id - k * nxy;
        const int64_t j = r / a.nx, i = r - j * a.nx;
        const double dx = (double)i - a.cx, dy = (double)j - a.cy, dz = (double)k - a.cz;
        const double t = (sqrt(dx * dx + dy * dy + dz * dz) - a.R - a.d0) / a.dd;
        double x;
        if (t <= 0.0) {
            x = a.prof[0];
        } else if (t >= (double)(a.nprof - 1)) {
            x = a.prof[a.nprof - 1];
        } else {
            const int q = (int)t;
            const double w = t - (double)q;
            x = a.prof[q] + w * (a.prof[q + 1] - a.prof[q]);
        }
        const T xv = (T)x;
        rho[id] = xv;
        const bool copy = ((a.dup & 1) && i == a.nx - 1) || ((a.dup & 2) && j == a.ny - 1) || ((a.dup & 4) && k == a.nz - 1);
        v = copy ? 0.0 : (double)xv;
    }
    s[threadIdx.x] = v;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// total of the per-block partials in block order, as the host loop of
// mplb_set_radial_density adds them (same FP64 additions, same result)
__global__ void k_sum_in_order(const double* __restrict__ part, unsigned n, double* __restrict__ out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double s = 0;
    for (unsigned b = 0; b < n; b++) s += part[b];
    *out = s;
}

int radial_dup(const mplb_case* c) {
    // periodic axes of the built boxes (latticeBox2D/3D boundaryType); none for lattices
    // read from files
    switch (c->mesh.btype) {
    case BOX_PERIODIC_X: return 1;
    case BOX_PERIODIC_Y: return 2;
    case BOX_PERIODIC_XY: return 3;
    case BOX_PERIODIC_Z: return 4;
    case BOX_PERIODIC_XYZ: return 7;
    case BOX_PERIODIC_XY_SPOT: return 3;
    default: return 0;
    }
}

// per-block max of |x - snap| as the bit pattern of a non-negative double (a NaN
// difference orders above every number), then snap := x
template <typename T>
__global__ void k_change(const T* __restrict__ x, T* __restrict__ snap, int64_t n, unsigned long long* part) {
    __shared__ unsigned long long s[kBlock];
    const int64_t id = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    unsigned long long u = 0;
    if (id < n) {
        const double d = fabs((double)x[id] - (double)snap[id]);
        u = (unsigned long long)__double_as_longlong(d);
        snap[id] = x[id];
    }
    s[threadIdx.x] = u;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) s[threadIdx.x] = s[threadIdx.x] > s[threadIdx.x + h] ? s[threadIdx.x] : s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

}  // namespace

namespace mplb {

int warm_scratch_setup(mplb_case* c) {
    const int dev = c->opts.device;
    if (dev < 0 || dev >= kMaxDevices) return wfail(c, MPLB_E_ARG, "device ordinal out of range for the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[dev]) return 0;
    void* p = nullptr;
    WCHK(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return wfail(c, MPLB_E_HIP, "hipGetSymbolAddress(scratch) returned no address");
    g_scratch_dev[dev] = (double*)p;
    return 0;
}

}  // namespace mplb

extern "C" {

int mplb_set_radial_density(mplb_case* c, const double* prof, int nprof, double d0, double dd, const double* centre,
                            double R, double* mass) {
    if (!c || !prof || nprof < 2 || !(dd > 0) || !centre) return MPLB_E_ARG;
    if (c->opts.device < 0 || !c->rho) return wfail(c, MPLB_E_STATE, "the case has no device fields (device < 0)");
    if (!c->mesh.structured) return wfail(c, MPLB_E_MODEL, "radial fields need a lattice in structured box order");
    WCHK(c, hipSetDevice(c->opts.device));
    if (c->fused) {
        // bring the generic fields to the current level first, as mplb_set_field does, so
        // that later field transfers do not overwrite the density written here
        const int rc = fused_set_field(c, "rho", nullptr, 0);
        if (rc < 0) return rc;
    }
    const int64_t N = c->mesh.N;
    const unsigned nb = (unsigned)((N + kBlock - 1) / kBlock);
    double* dprof = nullptr;
    double* part = nullptr;
    WCHK(c, hipMalloc(&dprof, sizeof(double) * nprof));
    WCHK(c, hipMalloc(&part, sizeof(double) * nb));
    WCHK(c, hipMemcpyAsync(dprof, prof, sizeof(double) * nprof, hipMemcpyHostToDevice, c->stream));
    Radial a{N, c->mesh.nx, c->mesh.ny, c->mesh.nz, centre[0], centre[1], c->mesh.d == 3 ? centre[2] : 0.0, R, d0, dd,
             dprof, nprof, radial_dup(c)};
    if (c->scalar_bytes == 8) k_radial<double><<<nb, kBlock, 0, c->stream>>>(a, (double*)c->rho, part);
    else k_radial<float><<<nb, kBlock, 0, c->stream>>>(a, (float*)c->rho, part);
    WCHK(c, hipGetLastError());
    std::string hostbuf(sizeof(double) * nb, '\0');
    double* hp = (double*)&hostbuf[0];
    WCHK(c, hipMemcpyAsync(hp, part, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
    WCHK(c, hipStreamSynchronize(c->stream));
    WCHK(c, hipFree(dprof));
    WCHK(c, hipFree(part));
    double s = 0;
    for (unsigned b = 0; b < nb; b++) s += hp[b];
    if (mass) *mass = s;
    c->initialized = false;       // the new density takes effect at mplb_init
    return 0;
}

int mplb_set_radial_density_noalloc(mplb_case* c, const double* prof, int nprof, double d0, double dd,
                                    const double* centre, double R, double* mass) {
    if (!c || !prof || nprof < 2 || !(dd > 0) || !centre) return MPLB_E_ARG;
    if (c->opts.device < 0 || !c->rho) return wfail(c, MPLB_E_STATE, "the case has no device fields (device < 0)");
    if (!c->mesh.structured) return wfail(c, MPLB_E_MODEL, "radial fields need a lattice in structured box order");
    if (nprof > kMaxProf)
        return wfail(c, MPLB_E_ARG, "allocation-free raster: nprof " + std::to_string(nprof) + " exceeds the scratch limit " +
                                        std::to_string(kMaxProf));
    const int64_t N = c->mesh.N;
    if (N > kMaxPart * kBlock)
        return wfail(c, MPLB_E_ARG, "allocation-free raster: " + std::to_string(N) + " nodes exceed the scratch limit " +
                                        std::to_string(kMaxPart * kBlock));
    double* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (c->opts.device < kMaxDevices) base = g_scratch_dev[c->opts.device];
    }
    if (!base) return wfail(c, MPLB_E_STATE, "scratch address not resolved for this device");
    WCHK(c, hipSetDevice(c->opts.device));
    if (c->fused) {
        const int rc = fused_set_field(c, "rho", nullptr, 0);   // launches only, no allocation
        if (rc < 0) return rc;
    }
    const unsigned nb = (unsigned)((N + kBlock - 1) / kBlock);
    double* dprof = base;
    double* part = base + kMaxProf;
    double* total = base + kMaxProf + kMaxPart;
    WCHK(c, hipMemcpyAsync(dprof, prof, sizeof(double) * nprof, hipMemcpyHostToDevice, c->stream));
    Radial a{N, c->mesh.nx, c->mesh.ny, c->mesh.nz, centre[0], centre[1], c->mesh.d == 3 ? centre[2] : 0.0, R, d0, dd,
             dprof, nprof, radial_dup(c)};
    if (c->scalar_bytes == 8) k_radial<double><<<nb, kBlock, 0, c->stream>>>(a, (double*)c->rho, part);
    else k_radial<float><<<nb, kBlock, 0, c->stream>>>(a, (float*)c->rho, part);
    WCHK(c, hipGetLastError());
    k_sum_in_order<<<1, 64, 0, c->stream>>>(part, nb, total);
    WCHK(c, hipGetLastError());
    double s = 0;
    WCHK(c, hipMemcpyAsync(&s, total, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    WCHK(c, hipStreamSynchronize(c->stream));
    if (mass) *mass = s;
    c->initialized = false;       // the new density takes effect at mplb_init
    return 0;
}

int mplb_density_change(mplb_case* c, double* max_abs) {
    if (!c || !max_abs) return MPLB_E_ARG;
    if (c->opts.device < 0 || !c->rho) return wfail(c, MPLB_E_STATE, "the case has no device fields (device < 0)");
    if (!c->initialized) return wfail(c, MPLB_E_STATE, "mplb_init first");
    WCHK(c, hipSetDevice(c->opts.device));
    if (c->fused) {
        const int rc = fused_sync_to_generic(c);
        if (rc) return rc;
    }
    const int64_t N = c->mesh.N;
    const unsigned nb = (unsigned)((N + kBlock - 1) / kBlock);
    if (!c->rho_snap) WCHK(c, hipMalloc(&c->rho_snap, c->scalar_bytes * N));
    if (!c->snap_valid) {
        // first call after mplb_init: take the snapshot, no change to report yet
        WCHK(c, hipMemcpyAsync(c->rho_snap, c->rho, c->scalar_bytes * N, hipMemcpyDeviceToDevice, c->stream));
        WCHK(c, hipStreamSynchronize(c->stream));
        c->snap_valid = true;
        *max_abs = NAN;
        return 0;
    }
    unsigned long long* part = nullptr;
    WCHK(c, hipMalloc(&part, sizeof(unsigned long long) * nb));
    if (c->scalar_bytes == 8) k_change<double><<<nb, kBlock, 0, c->stream>>>((double*)c->rho, (double*)c->rho_snap, N, part);
    else k_change<float><<<nb, kBlock, 0, c->stream>>>((float*)c->rho, (float*)c->rho_snap, N, part);
    WCHK(c, hipGetLastError());
    std::string hostbuf(sizeof(unsigned long long) * nb, '\0');
    unsigned long long* hp = (unsigned long long*)&hostbuf[0];
    WCHK(c, hipMemcpyAsync(hp, part, sizeof(unsigned long long) * nb, hipMemcpyDeviceToHost, c->stream));
    WCHK(c, hipStreamSynchronize(c->stream));
    WCHK(c, hipFree(part));
    unsigned long long m = 0;
    for (unsigned b = 0; b < nb; b++) m = hp[b] > m ? hp[b] : m;
    double d;
    std::memcpy(&d, &m, sizeof(d));
    *max_abs = d;
    return 0;
}

}  // extern "C"
