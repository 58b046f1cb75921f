// Memory-pattern ceilings for the D3Q15 single pass (a measurement program, not part of
// libmplb.so).  On an n^3 lattice of 15 FP32 population planes it times, with HIP
// events over 20 launches each:
//   copy     a flat float4 copy of the 15 planes (the plain streaming rate),
//   pull     the single pass's memory pattern: every site of a 64 x 4 column pulls its
//            15 populations from x - c_q (periodic) and stores them at x, marching up a
//            z chunk, work items handed to XCDs in contiguous ranges,
//   pull+ring the same plus the pulls of the one-site ring around the column,
// with no collision arithmetic.  Bytes counted: 120 per node (15 x 4 read + 15 x 4
// written), the single pass's algorithmic count.  One JSON object per line on stdout.
// "sweep" times streaming variants of the same planes instead: the copy at other grid
// sizes, one float4 per work item, four float4 per work item with the loads issued
// first, plain stores, non-temporal loads, and a read-only and a write-only pass
// (60 B per node counted for those two).
// "shapes" times the single pass's access shape without its march: one workgroup per
// (column, plane), each site copying (or pulling) its 15 populations, for 64 x 4,
// 128 x 2 and 256 x 1 columns, and the marching pull (no ring) with those columns.
// Usage: stream_probe [n [copy|sweep|shapes]] ("copy": the flat copy only).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <string>

#define PCHK(x)                                                                      \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            std::fprintf(stderr, "stream_probe: %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            std::exit(1);                                                            \
        }                                                                            \
    } while (0)

namespace {

constexpr int Q = 15, TX = 64, TY = 4, HX = TX + 2, NRING = 2 * HX + 2 * TY;
constexpr double PEAK_GBS = 8000.0;

// D3Q15 velocities (any order: the pattern, not the physics, is measured)
__constant__ int kC[Q][3] = {{0, 0, 0},  {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},
                             {0, 0, 1},  {0, 0, -1},  {1, 1, 1},   {-1, -1, -1}, {1, 1, -1},
                             {-1, -1, 1}, {1, -1, 1}, {-1, 1, -1}, {-1, 1, 1},  {1, -1, -1}};

struct Grid {
    const float* fin;
    float* fout;
    int n;
    long long plane, P;          // n^2, n^3
    int tiles, items, per, zc;   // tiles per plane, work items, items per XCD, planes per chunk
};

__device__ inline int wrap(int u, int n) { return u < 0 ? u + n : (u >= n ? u - n : u); }

typedef float f4 __attribute__((ext_vector_type(4)));

__global__ void k_copy(const f4* in, f4* out, long long n4) {
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n4; i += (long long)gridDim.x * 256)
        __builtin_nontemporal_store(in[i], out + i);
}

// grid-stride copy; NT_LOAD: non-temporal loads, NT_STORE: non-temporal stores
template <bool NT_LOAD, bool NT_STORE>
__global__ void k_copy_v(const f4* in, f4* out, long long n4) {
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n4; i += (long long)gridDim.x * 256) {
        const f4 v = NT_LOAD ? __builtin_nontemporal_load(in + i) : in[i];
        if (NT_STORE) __builtin_nontemporal_store(v, out + i);
        else out[i] = v;
    }
}

// one block per 256 * U float4: U loads issued before the U stores
template <int U>
__global__ void k_copy_blk(const f4* in, f4* out, long long n4) {
    const long long base = blockIdx.x * (256ll * U) + threadIdx.x;
    f4 v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const long long i = base + 256ll * u;
        if (i < n4) v[u] = in[i];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        const long long i = base + 256ll * u;
        if (i < n4) __builtin_nontemporal_store(v[u], out + i);
    }
}

// read-only pass: the sum is stored only if it equals a value the zero-filled planes
// never give, so the loads stay and nothing is written
__global__ void k_read(const f4* in, float* out, long long n4) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n4; i += (long long)gridDim.x * 256) acc += in[i];
    const float s = acc.x + acc.y + acc.z + acc.w;
    if (s == 12345.f) out[blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ void k_write(f4* out, long long n4, float v) {
    const f4 w = {v, v, v, v};
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n4; i += (long long)gridDim.x * 256)
        __builtin_nontemporal_store(w, out + i);
}

// one workgroup per (column of CX x CY sites, plane); PULL: from x - c_q, else from x
template <int CX, int CY, bool PULL>
__global__ void __launch_bounds__(CX * CY) k_site(const float* __restrict__ fin, float* __restrict__ fout, int n) {
    const int mtx = n / CX, tiles = mtx * (n / CY);
    const int z = blockIdx.x / tiles, tile = blockIdx.x % tiles;
    const int x = (tile % mtx) * CX + threadIdx.x % CX, y = (tile / mtx) * CY + threadIdx.x / CX;
    const long long plane = (long long)n * n, P = plane * n;
    const long long id = x + (long long)y * n + z * plane;
    float v[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const long long s = PULL ? wrap(x - kC[q][0], n) + (long long)wrap(y - kC[q][1], n) * n +
                                       (long long)wrap(z - kC[q][2], n) * plane
                                 : id;
        v[q] = fin[q * P + s];
    }
#pragma unroll
    for (int q = 0; q < Q; q++) __builtin_nontemporal_store(v[q], fout + q * P + id);
}

// the marching pull without ring for a CX x CY column (shape sweep)
template <int CX, int CY>
__global__ void __launch_bounds__(CX * CY) k_march_shape(Grid g) {
    const unsigned wi = (blockIdx.x & 7u) * (unsigned)g.per + (blockIdx.x >> 3);
    if (wi >= (unsigned)g.items) return;
    const int chunk = (int)(wi / (unsigned)g.tiles), tile = (int)(wi % (unsigned)g.tiles);
    const int mtx = g.n / CX;
    const int x = (tile % mtx) * CX + threadIdx.x % CX, y = (tile / mtx) * CY + threadIdx.x / CX;
    const int z0 = chunk * g.zc, z1 = min(z0 + g.zc, g.n);
    for (int z = z0; z < z1; z++) {
        float v[Q];
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const long long s = wrap(x - kC[q][0], g.n) + (long long)wrap(y - kC[q][1], g.n) * g.n +
                                (long long)wrap(z - kC[q][2], g.n) * g.plane;
            v[q] = g.fin[q * g.P + s];
        }
        const long long id = x + (long long)y * g.n + (long long)z * g.plane;
#pragma unroll
        for (int q = 0; q < Q; q++) __builtin_nontemporal_store(v[q], g.fout + q * g.P + id);
    }
}

template <bool RING>
__global__ void __launch_bounds__(TX * TY) k_pull(Grid g) {
    const unsigned wi = (blockIdx.x & 7u) * (unsigned)g.per + (blockIdx.x >> 3);
    if (wi >= (unsigned)g.items) return;
    const int chunk = (int)(wi / (unsigned)g.tiles), tile = (int)(wi % (unsigned)g.tiles);
    const int mtx = g.n / TX;
    const int x0 = (tile % mtx) * TX, y0 = (tile / mtx) * TY;
    const int z0 = chunk * g.zc, z1 = min(z0 + g.zc, g.n);
    const int t = threadIdx.x, x = x0 + t % TX, y = y0 + t / TX;
    // ring site of thread t < NRING: the rows below and above the column, then its sides
    int rhx = 0, rhy = 0;
    if (t < HX) {
        rhx = t;
    } else if (t < 2 * HX) {
        rhx = t - HX;
        rhy = TY + 1;
    } else if (t < 2 * HX + TY) {
        rhy = 1 + (t - 2 * HX);
    } else {
        rhx = HX - 1;
        rhy = 1 + (t - 2 * HX - TY);
    }
    const int rx = wrap(x0 - 1 + rhx, g.n), ry = wrap(y0 - 1 + rhy, g.n);
    float acc = 0.f;
    for (int z = z0; z < z1; z++) {
        float v[Q];
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const long long s = wrap(x - kC[q][0], g.n) + (long long)wrap(y - kC[q][1], g.n) * g.n +
                                (long long)wrap(z - kC[q][2], g.n) * g.plane;
            v[q] = g.fin[q * g.P + s];
        }
        if (RING && t < NRING) {
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const long long s = wrap(rx - kC[q][0], g.n) + (long long)wrap(ry - kC[q][1], g.n) * g.n +
                                    (long long)wrap(wrap(z + 1, g.n) - kC[q][2], g.n) * g.plane;
                acc += g.fin[q * g.P + s];
            }
        }
        const long long id = x + (long long)y * g.n + (long long)z * g.plane;
#pragma unroll
        for (int q = 0; q < Q; q++) __builtin_nontemporal_store(v[q] + acc * 1e-30f, g.fout + q * g.P + id);
    }
}

void report(const char* what, int n, int chunks, int items, float ms, int reps, double bytes_per_node = 120.0) {
    const double nodes = (double)n * n * n, sec = ms / reps * 1e-3;
    const double gbs = bytes_per_node * nodes / sec / 1e9;
    std::printf("{\"pattern\": \"%s\", \"n\": %d, \"chunks\": %d, \"items\": %d, \"ms\": %.5f, \"GBs\": %.1f, "
                "\"frac_of_8TBs\": %.4f}\n",
                what, n, chunks, items, ms / reps, gbs, gbs / PEAK_GBS);
    std::fflush(stdout);
}

}  // namespace

int main(int argc, char** argv) {
    const int n = argc > 1 ? std::atoi(argv[1]) : 256;
    const int reps = 20;
    if (n < TX || n % TX != 0 || n % TY != 0) {
        std::fprintf(stderr, "stream_probe: n must be a multiple of %d\n", TX);
        return 2;
    }
    const long long P = (long long)n * n * n;
    float *fin = nullptr, *fout = nullptr;
    PCHK(hipMalloc(&fin, sizeof(float) * Q * P));
    PCHK(hipMalloc(&fout, sizeof(float) * Q * P));
    PCHK(hipMemset(fin, 0, sizeof(float) * Q * P));
    int cus = 0;
    PCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
    hipEvent_t e0, e1;
    PCHK(hipEventCreate(&e0));
    PCHK(hipEventCreate(&e1));
    float ms = 0.f;
    // times `launch` (one launch per call) over `reps` launches after a warm-up pass;
    // `blocks` is reported as the work-item count
    auto timed_b = [&](const char* what, int chunks, long long blocks, double bpn, auto launch) {
        for (int pass = 0; pass < 2; pass++) {
            PCHK(hipEventRecord(e0));
            for (int r = 0; r < reps; r++) launch();
            PCHK(hipGetLastError());
            PCHK(hipEventRecord(e1));
            PCHK(hipEventSynchronize(e1));
            PCHK(hipEventElapsedTime(&ms, e0, e1));
        }
        report(what, n, chunks, (int)blocks, ms, reps, bpn);
    };
    const long long n4 = Q * P / 4;
    // the two copies every mode reports first: grid-stride over 8 workgroups per CU, and
    // one float4 per work item over the whole grid (the faster of the two on MI355X)
    timed_b("copy", 0, 0, 120.0, [&] { k_copy<<<cus * 8, 256>>>((const f4*)fin, (f4*)fout, n4); });
    const long long g1 = (n4 + 255) / 256;
    timed_b("copy 1 float4 per item", 0, g1, 120.0,
            [&] { k_copy_blk<1><<<g1, 256>>>((const f4*)fin, (f4*)fout, n4); });
    const bool copy_only = argc > 2 && std::string(argv[2]) == "copy";
    if (argc > 2 && std::string(argv[2]) == "shapes") {
        auto timed = [&](const char* what, int chunks, long long blocks, auto launch) {
            timed_b(what, chunks, blocks, 120.0, launch);
        };
        const long long b = (long long)n * n * n / 256;      // 256 sites per workgroup
        timed("site copy 64x4", 0, b, [&] { k_site<64, 4, false><<<b, 256>>>(fin, fout, n); });
        timed("site copy 256x1", 0, b, [&] { k_site<256, 1, false><<<b, 256>>>(fin, fout, n); });
        timed("site pull 64x4", 0, b, [&] { k_site<64, 4, true><<<b, 256>>>(fin, fout, n); });
        timed("site pull 128x2", 0, b, [&] { k_site<128, 2, true><<<b, 256>>>(fin, fout, n); });
        timed("site pull 256x1", 0, b, [&] { k_site<256, 1, true><<<b, 256>>>(fin, fout, n); });
        for (int chunks : {1, 3, 8}) {
            Grid m{fin, fout, n, (long long)n * n, P, 0, 0, 0, 0};
            m.zc = (n + chunks - 1) / chunks;
            const int nch = (n + m.zc - 1) / m.zc;
            m.tiles = (n / 64) * (n / 4);
            m.items = m.tiles * nch;
            m.per = (m.items + 7) / 8;
            timed("march pull 64x4", chunks, m.items, [&] { k_march_shape<64, 4><<<8 * m.per, 256>>>(m); });
            m.tiles = (n / 128) * (n / 2);
            m.items = m.tiles * nch;
            m.per = (m.items + 7) / 8;
            timed("march pull 128x2", chunks, m.items, [&] { k_march_shape<128, 2><<<8 * m.per, 256>>>(m); });
            m.tiles = (n / 256) * n;
            m.items = m.tiles * nch;
            m.per = (m.items + 7) / 8;
            if (n % 256 == 0)
                timed("march pull 256x1", chunks, m.items, [&] { k_march_shape<256, 1><<<8 * m.per, 256>>>(m); });
        }
        PCHK(hipFree(fin));
        PCHK(hipFree(fout));
        return 0;
    }
    if (argc > 2 && std::string(argv[2]) == "sweep") {
        auto timed = [&](const char* what, long long blocks, double bpn, auto launch) {
            timed_b(what, 0, blocks, bpn, launch);
        };
        for (int per_cu : {2, 4, 16, 32}) {
            const int b = cus * per_cu;
            timed("copy", b, 120.0, [&] { k_copy<<<b, 256>>>((const f4*)fin, (f4*)fout, n4); });
        }
        const int b8 = cus * 8;
        timed("copy plain stores", b8, 120.0, [&] { k_copy_v<false, false><<<b8, 256>>>((const f4*)fin, (f4*)fout, n4); });
        timed("copy nt loads", b8, 120.0, [&] { k_copy_v<true, true><<<b8, 256>>>((const f4*)fin, (f4*)fout, n4); });
        const long long g4 = (n4 + 1023) / 1024;
        timed("copy 4 float4 per item", g4, 120.0, [&] { k_copy_blk<4><<<g4, 256>>>((const f4*)fin, (f4*)fout, n4); });
        for (int per_cu : {8, 32}) {
            const int b = cus * per_cu;
            timed("read only", b, 60.0, [&] { k_read<<<b, 256>>>((const f4*)fin, fout, n4); });
            timed("write only", b, 60.0, [&] { k_write<<<b, 256>>>((f4*)fout, n4, 0.f); });
        }
        PCHK(hipFree(fin));
        PCHK(hipFree(fout));
        return 0;
    }
    Grid g{fin, fout, n, (long long)n * n, P, (n / TX) * (n / TY), 0, 0, 0};
    for (int ring = 0; ring < 2 && !copy_only; ring++) {
        for (int chunks : {1, 2, 3}) {
            g.zc = (n + chunks - 1) / chunks;
            g.items = g.tiles * ((n + g.zc - 1) / g.zc);
            g.per = (g.items + 7) / 8;
            for (int pass = 0; pass < 2; pass++) {
                PCHK(hipEventRecord(e0));
                for (int r = 0; r < reps; r++) {
                    if (ring) k_pull<true><<<8 * g.per, TX * TY>>>(g);
                    else k_pull<false><<<8 * g.per, TX * TY>>>(g);
                }
                PCHK(hipGetLastError());
                PCHK(hipEventRecord(e1));
                PCHK(hipEventSynchronize(e1));
                PCHK(hipEventElapsedTime(&ms, e0, e1));
            }
            report(ring ? "pull+ring" : "pull", n, chunks, g.items, ms, reps);
        }
    }
    PCHK(hipFree(fin));
    PCHK(hipFree(fout));
    return 0;
}
