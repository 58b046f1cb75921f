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
// Usage: stream_probe [n [copy]] ("copy": the flat copy only).
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

void report(const char* what, int n, int chunks, int items, float ms, int reps) {
    const double nodes = (double)n * n * n, sec = ms / reps * 1e-3;
    const double gbs = 120.0 * nodes / sec / 1e9;
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
    for (int pass = 0; pass < 2; pass++) {     // the first pass warms up
        PCHK(hipEventRecord(e0));
        for (int r = 0; r < reps; r++) k_copy<<<cus * 8, 256>>>((const f4*)fin, (f4*)fout, Q * P / 4);
        PCHK(hipGetLastError());
        PCHK(hipEventRecord(e1));
        PCHK(hipEventSynchronize(e1));
        PCHK(hipEventElapsedTime(&ms, e0, e1));
    }
    report("copy", n, 0, 0, ms, reps);
    const bool copy_only = argc > 2 && std::string(argv[2]) == "copy";
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
