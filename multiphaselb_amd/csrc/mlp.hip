// Multilayer perceptron of the learned warm start: tanh hidden layers, linear output,
// mean-squared-error loss, Adam.  Matrix products run on the gfx950 matrix cores
// (v_mfma_f32_16x16x32_bf16): weights and activations enter as BF16, products
// accumulate in FP32; the master parameters, the Adam moments, the biases, the loss
// and the activation derivatives stay FP32.
//
// Every product is C[M][N] = A[M][K] * Bt[N][K]^T with both operands row-major in K,
// so each lane loads its 8 consecutive K values of a row with one 16-byte load.  The
// three products of a layer map onto that form by keeping each operand in the layout
// it needs:
//   forward  Z  = X W        A = X   [B][in]   Bt = W^T [out][in]
//   input    dX = dZ W^T     A = dZ  [B][out]  Bt = W   [in][out]
//   weight   dW = X^T dZ     A = X^T [in][B]   Bt = dZ^T [out][B]
// so the forward epilogue writes each activation as [B][w] and [w][B], the backward
// epilogue writes dZ both ways, and the BF16 weight copies exist in both layouts.
// A training iteration (gather, forward, loss, backward, Adam, BF16 refresh) has no
// host-side state: the minibatch index and the Adam step count come from a device
// counter, so one captured graph is replayed for every iteration.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mplb.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 64;         // workgroup tile (4 waves of 32 x 32)
constexpr int kKStep = 32;        // K per MFMA
constexpr int kInPad = 32;        // input width rounded up to the K step

enum Epi : int { E_HIDDEN = 0, E_OUT = 1, E_DX = 2, E_DW = 3 };

struct Gemm {
    const __bf16* A;
    int lda;
    const __bf16* Bt;
    int ldb;
    int M, N, K;
    const float* bias;        // E_HIDDEN, E_OUT
    float* out32;             // a (E_HIDDEN), y (E_OUT), dz (E_DX), dW (E_DW)
    int ld32;
    __bf16* outb;             // [M][N] BF16 copy (E_HIDDEN, E_DX)
    int ldob;
    __bf16* outbt;            // [N][M] BF16 copy (E_HIDDEN, E_DX)
    int ldobt;
    const float* act;         // E_DX: activation of the layer's input level, [M][N]
    int ldact;
    int mvalid;               // E_DW: rows m >= mvalid are input padding (not written)
};

__device__ inline __bf16 bf(float x) { return (__bf16)x; }

// Wave w of the workgroup computes the 32 x 32 block (m0, n0) as 2 x 2 MFMA tiles of
// 16 x 16; lane l supplies row l % 16 of each A tile and row l % 16 of each Bt tile,
// K values 8 (l / 16) .. 8 (l / 16) + 7 of the current K step, and receives
// C[4 (l / 16) + e][l % 16], e = 0..3, of each tile.
template <int EPI>
__global__ void __launch_bounds__(256) k_gemm(Gemm g) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m0 = blockIdx.y * kTile + (wave >> 1) * 32;
    const int n0 = blockIdx.x * kTile + (wave & 1) * 32;
    if (m0 >= g.M || n0 >= g.N) return;
    const int r = lane & 15, kq = (lane >> 4) * 8;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const __bf16* pa0 = g.A + (size_t)(m0 + r) * g.lda + kq;
    const __bf16* pa1 = pa0 + (size_t)16 * g.lda;
    const __bf16* pb0 = g.Bt + (size_t)(n0 + r) * g.ldb + kq;
    const __bf16* pb1 = pb0 + (size_t)16 * g.ldb;
    for (int k = 0; k < g.K; k += kKStep) {
        const bf16x8 a0 = *(const bf16x8*)(pa0 + k), a1 = *(const bf16x8*)(pa1 + k);
        const bf16x8 b0 = *(const bf16x8*)(pb0 + k), b1 = *(const bf16x8*)(pb1 + k);
        acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc[1][1], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 2; i++) {
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int n = n0 + 16 * j + r;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int m = m0 + 16 * i + (lane >> 4) * 4 + e;
                const float v = acc[i][j][e];
                if (EPI == E_HIDDEN) {
                    const float a = tanhf(v + g.bias[n]);
                    g.out32[(size_t)m * g.ld32 + n] = a;
                    g.outb[(size_t)m * g.ldob + n] = bf(a);
                    g.outbt[(size_t)n * g.ldobt + m] = bf(a);
                } else if (EPI == E_OUT) {
                    g.out32[(size_t)m * g.ld32 + n] = v + g.bias[n];
                } else if (EPI == E_DX) {
                    const float a = g.act[(size_t)m * g.ldact + n];
                    const float dz = v * (1.0f - a * a);
                    g.out32[(size_t)m * g.ld32 + n] = dz;
                    g.outb[(size_t)m * g.ldob + n] = bf(dz);
                    g.outbt[(size_t)n * g.ldobt + m] = bf(dz);
                } else {
                    if (m < g.mvalid) g.out32[(size_t)m * g.ld32 + n] = v;
                }
            }
        }
    }
}

// minibatch rows idx[it * B + b] of X [n][din] / T [n][dout] -> BF16 input in both
// layouts (zero padding columns) and the FP32 targets; it = *counter (or 0 with no idx:
// rows 0..B-1 in order)
__global__ void k_gather(const float* __restrict__ X, const float* __restrict__ T, const int* __restrict__ idx,
                         const int* __restrict__ counter, int B, int din, int dout, __bf16* ab, __bf16* abt,
                         float* tgt) {
    const int b = blockIdx.x;
    const int row = idx ? idx[(size_t)(*counter) * B + b] : b;
    for (int c = threadIdx.x; c < kInPad; c += blockDim.x) {
        const float x = c < din ? X[(size_t)row * din + c] : 0.f;
        ab[(size_t)b * kInPad + c] = bf(x);
        abt[(size_t)c * B + b] = bf(x);
    }
    if (T)
        for (int c = threadIdx.x; c < dout; c += blockDim.x) tgt[(size_t)b * dout + c] = T[(size_t)row * dout + c];
}

// dL/dy of L = sum (y - t)^2 / (B n) and the row sums of squared errors
__global__ void k_loss(const float* __restrict__ y, const float* __restrict__ t, int B, int n, float* dz32,
                       __bf16* dzb, __bf16* dzt, float* rowsum) {
    const int b = blockIdx.x;
    const float scale = 2.0f / ((float)B * (float)n);
    __shared__ float part[256];
    float s = 0.f;
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        const float d = y[(size_t)b * n + c] - t[(size_t)b * n + c];
        s += d * d;
        const float g = scale * d;
        dz32[(size_t)b * n + c] = g;
        dzb[(size_t)b * n + c] = bf(g);
        dzt[(size_t)c * B + b] = bf(g);
    }
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = 0.f;
        for (unsigned i = 0; i < blockDim.x; i++) r += part[i];
        rowsum[b] = r;
    }
}

__global__ void k_loss_final(const float* __restrict__ rowsum, int B, int n, const int* __restrict__ counter,
                             float* hist, float* last) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0;
    for (int b = 0; b < B; b++) s += rowsum[b];
    const float L = (float)(s / ((double)B * (double)n));
    *last = L;
    if (hist) hist[*counter] = L;
}

// db[c] = sum over the batch of dz[b][c].  A 256-lane workgroup takes 64 columns: lane
// (c, k) sums rows k, k + 4, k + 8, ... with four independent accumulators (loads in
// flight instead of one dependent chain), and the four slices of a column are added in
// slice order through LDS, so the result does not depend on scheduling.
__global__ void __launch_bounds__(256) k_bias_grad(const float* __restrict__ dz, int B, int n, float* db) {
    __shared__ float part[4][64];
    const int cl = threadIdx.x & 63, k = threadIdx.x >> 6;
    const int c = blockIdx.x * 64 + cl;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
    if (c < n) {
        int b = k;
        for (; b + 12 < B; b += 16) {
            s0 += dz[(size_t)b * n + c];
            s1 += dz[(size_t)(b + 4) * n + c];
            s2 += dz[(size_t)(b + 8) * n + c];
            s3 += dz[(size_t)(b + 12) * n + c];
        }
        for (; b < B; b += 4) s0 += dz[(size_t)b * n + c];
    }
    part[k][cl] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    if (k == 0 && c < n) db[c] = ((part[0][cl] + part[1][cl]) + part[2][cl]) + part[3][cl];
}

// Adam with bias correction; t = *counter + 1
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                       const int* __restrict__ counter) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float t = (float)(*counter + 1);
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.f - powf(b1, t));
    const float vh = vi / (1.f - powf(b2, t));
    p[i] -= lr * mh / (sqrtf(vh) + eps);
}

// BF16 copies of layer weights W [din][dout] (FP32 master): Wb [dinp][dout] and
// Wbt [dout][dinp], zero rows for the input padding
__global__ void k_refresh(const float* __restrict__ W, int din, int dinp, int dout, __bf16* Wb, __bf16* Wbt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= dinp * dout) return;
    const int r = i / dout, c = i - r * dout;
    const __bf16 v = bf(r < din ? W[(size_t)r * dout + c] : 0.f);
    Wb[(size_t)r * dout + c] = v;
    Wbt[(size_t)c * dinp + r] = v;
}

// all layers' copies in one launch (training replays this once per iteration, where a
// launch per layer costs more than the copy): element i belongs to the layer l with
// start[l] <= i < start[l + 1]
constexpr int kRefreshLayers = 8;
struct RefreshAll {
    int L;
    int64_t start[kRefreshLayers + 1];
    const float* W[kRefreshLayers];
    int din[kRefreshLayers], dinp[kRefreshLayers], dout[kRefreshLayers];
    __bf16* Wb[kRefreshLayers];
    __bf16* Wbt[kRefreshLayers];
};

__global__ void k_refresh_all(RefreshAll a) {
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g >= a.start[a.L]) return;
    int l = 0;
    while (g >= a.start[l + 1]) l++;
    const int i = (int)(g - a.start[l]), dout = a.dout[l], dinp = a.dinp[l];
    const int r = i / dout, c = i - r * dout;
    const __bf16 v = bf(r < a.din[l] ? a.W[l][(size_t)r * dout + c] : 0.f);
    a.Wb[l][(size_t)r * dout + c] = v;
    a.Wbt[l][(size_t)c * dinp + r] = v;
}

__global__ void k_advance(int* counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *counter += 1;
}

__global__ void k_set_int(int* p, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

}  // namespace

struct mplb_mlp {
    std::string err;
    int dev = 0;
    hipStream_t stream = nullptr;
    int L = 0;                                  // layers
    std::vector<int> w;                         // level widths (w[0] = raw input width)
    std::vector<int> wp;                        // padded widths used on the device
    int B = 0;                                  // rows per product
    std::vector<int64_t> woff, boff;            // parameter offsets
    int64_t P = 0;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    std::vector<__bf16*> Wb, Wbt;               // per layer
    std::vector<__bf16*> ab, abt;               // per level 0..L-1
    std::vector<float*> a32;                    // per level 1..L-1 (index by level)
    float* y = nullptr;                         // [B][w[L]]
    float* tgt = nullptr;
    std::vector<float*> dz32;                   // per level 1..L
    std::vector<__bf16*> dzb, dzt;
    float* rowsum = nullptr;
    float* loss = nullptr;
    int* counter = nullptr;
    std::vector<void*> owned;
    // training data resident on the device
    float *X = nullptr, *T = nullptr;
    int* idx = nullptr;
    size_t X_cap = 0, T_cap = 0, idx_cap = 0;
    float* hist = nullptr;
    size_t hist_cap = 0;
};

namespace {

int mfail(mplb_mlp* n, int code, const std::string& m) {
    if (n) n->err = m;
    return code;
}

#define MCHK(n, call)                                                                                           \
    do {                                                                                                        \
        hipError_t _e = (call);                                                                                 \
        if (_e != hipSuccess) return mfail((n), MPLB_E_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

template <typename T>
int alloc(mplb_mlp* n, T** ptr, size_t count) {
    void* q = nullptr;
    MCHK(n, hipMalloc(&q, sizeof(T) * (count ? count : 1)));
    MCHK(n, hipMemset(q, 0, sizeof(T) * (count ? count : 1)));
    n->owned.push_back(q);
    *ptr = (T*)q;
    return 0;
}

template <typename T>
int ensure(mplb_mlp* n, T** ptr, size_t* cap, size_t count) {
    if (count <= *cap) return 0;
    if (*ptr) MCHK(n, hipFree(*ptr));
    *ptr = nullptr;
    MCHK(n, hipMalloc((void**)ptr, sizeof(T) * count));
    *cap = count;
    return 0;
}

dim3 gemm_grid(int M, int N) { return dim3((unsigned)((N + kTile - 1) / kTile), (unsigned)((M + kTile - 1) / kTile)); }

void refresh(mplb_mlp* n) {
    if (n->L <= kRefreshLayers) {
        RefreshAll a{};
        a.L = n->L;
        for (int l = 0; l < n->L; l++) {
            a.start[l + 1] = a.start[l] + (int64_t)n->wp[l] * n->w[l + 1];
            a.W[l] = n->p + n->woff[l];
            a.din[l] = n->w[l];
            a.dinp[l] = n->wp[l];
            a.dout[l] = n->w[l + 1];
            a.Wb[l] = n->Wb[l];
            a.Wbt[l] = n->Wbt[l];
        }
        k_refresh_all<<<(unsigned)((a.start[n->L] + 255) / 256), 256, 0, n->stream>>>(a);
        return;
    }
    for (int l = 0; l < n->L; l++) {
        const int total = n->wp[l] * n->w[l + 1];
        k_refresh<<<(total + 255) / 256, 256, 0, n->stream>>>(n->p + n->woff[l], n->w[l], n->wp[l], n->w[l + 1],
                                                              n->Wb[l], n->Wbt[l]);
    }
}

void forward(mplb_mlp* n) {
    const int B = n->B;
    for (int l = 0; l < n->L; l++) {
        Gemm g{};
        g.A = n->ab[l];
        g.lda = n->wp[l];
        g.Bt = n->Wbt[l];
        g.ldb = n->wp[l];
        g.M = B;
        g.N = n->w[l + 1];
        g.K = n->wp[l];
        g.bias = n->p + n->boff[l];
        if (l < n->L - 1) {
            g.out32 = n->a32[l + 1];
            g.ld32 = g.N;
            g.outb = n->ab[l + 1];
            g.ldob = g.N;
            g.outbt = n->abt[l + 1];
            g.ldobt = B;
            k_gemm<E_HIDDEN><<<gemm_grid(g.M, g.N), 256, 0, n->stream>>>(g);
        } else {
            g.out32 = n->y;
            g.ld32 = g.N;
            k_gemm<E_OUT><<<gemm_grid(g.M, g.N), 256, 0, n->stream>>>(g);
        }
    }
}

void backward(mplb_mlp* n) {
    const int B = n->B, L = n->L;
    const int nout = n->w[L];
    k_loss<<<B, 256, 0, n->stream>>>(n->y, n->tgt, B, nout, n->dz32[L], n->dzb[L], n->dzt[L], n->rowsum);
    k_loss_final<<<1, 64, 0, n->stream>>>(n->rowsum, B, nout, n->counter, n->hist, n->loss);
    for (int l = L - 1; l >= 0; l--) {
        const int din = n->wp[l], dout = n->w[l + 1];
        Gemm gw{};
        gw.A = n->abt[l];
        gw.lda = B;
        gw.Bt = n->dzt[l + 1];
        gw.ldb = B;
        gw.M = din;
        gw.N = dout;
        gw.K = B;
        gw.out32 = n->g + n->woff[l];
        gw.ld32 = dout;
        gw.mvalid = n->w[l];
        k_gemm<E_DW><<<gemm_grid(gw.M, gw.N), 256, 0, n->stream>>>(gw);
        k_bias_grad<<<(dout + 63) / 64, 256, 0, n->stream>>>(n->dz32[l + 1], B, dout, n->g + n->boff[l]);
        if (l > 0) {
            Gemm gx{};
            gx.A = n->dzb[l + 1];
            gx.lda = dout;
            gx.Bt = n->Wb[l];
            gx.ldb = dout;
            gx.M = B;
            gx.N = din;
            gx.K = dout;
            gx.out32 = n->dz32[l];
            gx.ld32 = din;
            gx.outb = n->dzb[l];
            gx.ldob = din;
            gx.outbt = n->dzt[l];
            gx.ldobt = B;
            gx.act = n->a32[l];
            gx.ldact = din;
            k_gemm<E_DX><<<gemm_grid(gx.M, gx.N), 256, 0, n->stream>>>(gx);
        }
    }
}

}  // namespace

extern "C" {

int mplb_mlp_create(int nlayers, const int* widths, int batch, int device, mplb_mlp** out) {
    if (!out) return MPLB_E_ARG;
    *out = nullptr;
    if (nlayers < 1 || !widths) return MPLB_E_ARG;
    mplb_mlp* n = new mplb_mlp();
    *out = n;
    n->L = nlayers;
    n->w.assign(widths, widths + nlayers + 1);
    if (n->w[0] < 1 || n->w[0] > kInPad) return mfail(n, MPLB_E_ARG, "input width must be in [1, 32]");
    for (int l = 1; l <= nlayers; l++)
        if (n->w[l] < 32 || n->w[l] % 32) return mfail(n, MPLB_E_ARG, "layer widths must be multiples of 32");
    if (batch < 32 || batch % 32) return mfail(n, MPLB_E_ARG, "batch must be a multiple of 32");
    n->B = batch;
    n->wp = n->w;
    n->wp[0] = kInPad;
    for (int l = 0; l < nlayers; l++) {
        n->woff.push_back(n->P);
        n->P += (int64_t)n->w[l] * n->w[l + 1];
        n->boff.push_back(n->P);
        n->P += n->w[l + 1];
    }
    n->dev = device;
    MCHK(n, hipSetDevice(device));
    MCHK(n, hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
    int rc = 0;
    rc |= alloc(n, &n->p, n->P);
    rc |= alloc(n, &n->g, n->P);
    rc |= alloc(n, &n->m, n->P);
    rc |= alloc(n, &n->v, n->P);
    n->Wb.resize(nlayers);
    n->Wbt.resize(nlayers);
    for (int l = 0; l < nlayers; l++) {
        rc |= alloc(n, &n->Wb[l], (size_t)n->wp[l] * n->w[l + 1]);
        rc |= alloc(n, &n->Wbt[l], (size_t)n->wp[l] * n->w[l + 1]);
    }
    n->ab.assign(nlayers, nullptr);
    n->abt.assign(nlayers, nullptr);
    n->a32.assign(nlayers + 1, nullptr);
    n->dz32.assign(nlayers + 1, nullptr);
    n->dzb.assign(nlayers + 1, nullptr);
    n->dzt.assign(nlayers + 1, nullptr);
    for (int j = 0; j < nlayers; j++) {
        rc |= alloc(n, &n->ab[j], (size_t)batch * n->wp[j]);
        rc |= alloc(n, &n->abt[j], (size_t)batch * n->wp[j]);
        if (j > 0) rc |= alloc(n, &n->a32[j], (size_t)batch * n->wp[j]);
    }
    for (int j = 1; j <= nlayers; j++) {
        rc |= alloc(n, &n->dz32[j], (size_t)batch * n->wp[j]);
        rc |= alloc(n, &n->dzb[j], (size_t)batch * n->wp[j]);
        rc |= alloc(n, &n->dzt[j], (size_t)batch * n->wp[j]);
    }
    rc |= alloc(n, &n->y, (size_t)batch * n->w[nlayers]);
    rc |= alloc(n, &n->tgt, (size_t)batch * n->w[nlayers]);
    rc |= alloc(n, &n->rowsum, (size_t)batch);
    rc |= alloc(n, &n->loss, 1);
    rc |= alloc(n, &n->counter, 1);
    return rc ? MPLB_E_HIP : 0;
}

int mplb_mlp_close(mplb_mlp* n) {
    if (!n) return MPLB_E_ARG;
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    for (void* q : n->owned) (void)hipFree(q);
    void* extra[] = {n->X, n->T, n->idx, n->hist};
    for (void* q : extra)
        if (q) (void)hipFree(q);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
    return 0;
}

const char* mplb_mlp_last_error(const mplb_mlp* n) { return n ? n->err.c_str() : "null network handle"; }

int64_t mplb_mlp_num_params(const mplb_mlp* n) { return n ? n->P : -1; }

int mplb_mlp_set_params(mplb_mlp* n, const float* params, size_t count) {
    if (!n || !params) return MPLB_E_ARG;
    if ((int64_t)count != n->P) return mfail(n, MPLB_E_ARG, "parameter count differs from the network's");
    MCHK(n, hipSetDevice(n->dev));
    MCHK(n, hipMemcpyAsync(n->p, params, sizeof(float) * count, hipMemcpyHostToDevice, n->stream));
    hipLaunchKernelGGL(k_set_int, 1, 1, 0, n->stream, n->counter, 0);
    const size_t bytes = sizeof(float) * count;
    MCHK(n, hipMemsetAsync(n->m, 0, bytes, n->stream));
    MCHK(n, hipMemsetAsync(n->v, 0, bytes, n->stream));
    refresh(n);
    MCHK(n, hipGetLastError());
    MCHK(n, hipStreamSynchronize(n->stream));
    return 0;
}

int mplb_mlp_get_params(mplb_mlp* n, float* params, size_t count) {
    if (!n || !params) return MPLB_E_ARG;
    if ((int64_t)count != n->P) return mfail(n, MPLB_E_ARG, "parameter count differs from the network's");
    MCHK(n, hipSetDevice(n->dev));
    MCHK(n, hipMemcpyAsync(params, n->p, sizeof(float) * count, hipMemcpyDeviceToHost, n->stream));
    MCHK(n, hipStreamSynchronize(n->stream));
    return 0;
}

int mplb_mlp_forward(mplb_mlp* n, const float* x, int rows, float* y) {
    if (!n || !x || !y || rows < 0) return MPLB_E_ARG;
    MCHK(n, hipSetDevice(n->dev));
    const int din = n->w[0], dout = n->w[n->L], B = n->B;
    int rc = ensure(n, &n->X, &n->X_cap, (size_t)B * din);
    if (rc) return rc;
    std::vector<float> xb((size_t)B * din), yb((size_t)B * dout);
    for (int r0 = 0; r0 < rows; r0 += B) {
        const int cnt = std::min(B, rows - r0);
        std::fill(xb.begin(), xb.end(), 0.f);
        std::memcpy(xb.data(), x + (size_t)r0 * din, sizeof(float) * cnt * din);
        MCHK(n, hipMemcpyAsync(n->X, xb.data(), sizeof(float) * xb.size(), hipMemcpyHostToDevice, n->stream));
        k_gather<<<B, 64, 0, n->stream>>>(n->X, nullptr, nullptr, nullptr, B, din, dout, n->ab[0], n->abt[0],
                                          nullptr);
        forward(n);
        MCHK(n, hipGetLastError());
        MCHK(n, hipMemcpyAsync(yb.data(), n->y, sizeof(float) * yb.size(), hipMemcpyDeviceToHost, n->stream));
        MCHK(n, hipStreamSynchronize(n->stream));
        std::memcpy(y + (size_t)r0 * dout, yb.data(), sizeof(float) * cnt * dout);
    }
    return 0;
}

int mplb_mlp_gradient(mplb_mlp* n, const float* x, const float* t, float* loss, float* grad, size_t ngrad) {
    if (!n || !x || !t) return MPLB_E_ARG;
    if (grad && (int64_t)ngrad != n->P) return mfail(n, MPLB_E_ARG, "gradient size differs from the parameter count");
    MCHK(n, hipSetDevice(n->dev));
    const int din = n->w[0], dout = n->w[n->L], B = n->B;
    int rc = ensure(n, &n->X, &n->X_cap, (size_t)B * din);
    if (!rc) rc = ensure(n, &n->T, &n->T_cap, (size_t)B * dout);
    if (rc) return rc;
    MCHK(n, hipMemcpyAsync(n->X, x, sizeof(float) * B * din, hipMemcpyHostToDevice, n->stream));
    MCHK(n, hipMemcpyAsync(n->T, t, sizeof(float) * B * dout, hipMemcpyHostToDevice, n->stream));
    k_gather<<<B, 64, 0, n->stream>>>(n->X, n->T, nullptr, nullptr, B, din, dout, n->ab[0], n->abt[0], n->tgt);
    float* keep = n->hist;
    n->hist = nullptr;
    forward(n);
    backward(n);
    n->hist = keep;
    MCHK(n, hipGetLastError());
    if (loss) MCHK(n, hipMemcpyAsync(loss, n->loss, sizeof(float), hipMemcpyDeviceToHost, n->stream));
    if (grad) MCHK(n, hipMemcpyAsync(grad, n->g, sizeof(float) * ngrad, hipMemcpyDeviceToHost, n->stream));
    MCHK(n, hipStreamSynchronize(n->stream));
    return 0;
}

int mplb_mlp_train(mplb_mlp* n, const float* X, const float* T, int nsamples, const int32_t* idx, int iters,
                   float lr, float beta1, float beta2, float eps, float* loss_hist) {
    if (!n || !X || !T || !idx || nsamples <= 0 || iters <= 0) return MPLB_E_ARG;
    MCHK(n, hipSetDevice(n->dev));
    const int din = n->w[0], dout = n->w[n->L], B = n->B;
    for (int64_t i = 0; i < (int64_t)iters * B; i++)
        if (idx[i] < 0 || idx[i] >= nsamples) return mfail(n, MPLB_E_ARG, "minibatch index out of range");
    int rc = ensure(n, &n->X, &n->X_cap, (size_t)nsamples * din);
    if (!rc) rc = ensure(n, &n->T, &n->T_cap, (size_t)nsamples * dout);
    if (!rc) rc = ensure(n, &n->idx, &n->idx_cap, (size_t)iters * B);
    if (!rc) rc = ensure(n, &n->hist, &n->hist_cap, (size_t)iters);
    if (rc) return rc;
    MCHK(n, hipMemcpyAsync(n->X, X, sizeof(float) * nsamples * din, hipMemcpyHostToDevice, n->stream));
    MCHK(n, hipMemcpyAsync(n->T, T, sizeof(float) * nsamples * dout, hipMemcpyHostToDevice, n->stream));
    MCHK(n, hipMemcpyAsync(n->idx, idx, sizeof(int32_t) * iters * B, hipMemcpyHostToDevice, n->stream));
    // every call is a fresh Adam run: the iteration counter (index table, loss history,
    // bias-correction step count) restarts and the first and second moments are zeroed
    hipLaunchKernelGGL(k_set_int, 1, 1, 0, n->stream, n->counter, 0);
    MCHK(n, hipMemsetAsync(n->m, 0, sizeof(float) * n->P, n->stream));
    MCHK(n, hipMemsetAsync(n->v, 0, sizeof(float) * n->P, n->stream));
    MCHK(n, hipStreamSynchronize(n->stream));
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    MCHK(n, hipStreamBeginCapture(n->stream, hipStreamCaptureModeThreadLocal));
    k_gather<<<B, 64, 0, n->stream>>>(n->X, n->T, n->idx, n->counter, B, din, dout, n->ab[0], n->abt[0], n->tgt);
    forward(n);
    backward(n);
    k_adam<<<(unsigned)((n->P + 255) / 256), 256, 0, n->stream>>>(n->p, n->g, n->m, n->v, n->P, lr, beta1, beta2, eps,
                                                                   n->counter);
    refresh(n);
    k_advance<<<1, 64, 0, n->stream>>>(n->counter);
    MCHK(n, hipStreamEndCapture(n->stream, &graph));
    MCHK(n, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    for (int it = 0; it < iters; it++) MCHK(n, hipGraphLaunch(exec, n->stream));
    MCHK(n, hipStreamSynchronize(n->stream));
    MCHK(n, hipGraphExecDestroy(exec));
    MCHK(n, hipGraphDestroy(graph));
    if (loss_hist) MCHK(n, hipMemcpy(loss_hist, n->hist, sizeof(float) * iters, hipMemcpyDeviceToHost));
    return 0;
}

// Intermediate tensors of the last forward/backward, as FP32 (BF16 ones widened):
// "x" level-0 BF16 input [B][32]; "a" level j BF16 activation [B][w_j] (j = 0..L-1);
// "a32" FP32 activation (1..L-1); "y" output [B][w_L]; "dz" FP32 dL/dz of level j
// (1..L); "dzb" its BF16 copy; "wb" BF16 weights of layer j [32 or w_j][w_j+1].
int mplb_mlp_tensor(mplb_mlp* n, const char* name, int j, float* out, size_t count) {
    if (!n || !name || !out) return MPLB_E_ARG;
    MCHK(n, hipSetDevice(n->dev));
    const int B = n->B, L = n->L;
    const void* src = nullptr;
    bool is_bf = false;
    size_t want = 0;
    if (!std::strcmp(name, "a") && j >= 0 && j < L) {
        src = n->ab[j];
        is_bf = true;
        want = (size_t)B * n->wp[j];
    } else if (!std::strcmp(name, "a32") && j >= 1 && j < L) {
        src = n->a32[j];
        want = (size_t)B * n->wp[j];
    } else if (!std::strcmp(name, "y")) {
        src = n->y;
        want = (size_t)B * n->w[L];
    } else if (!std::strcmp(name, "dz") && j >= 1 && j <= L) {
        src = n->dz32[j];
        want = (size_t)B * n->wp[j];
    } else if (!std::strcmp(name, "dzb") && j >= 1 && j <= L) {
        src = n->dzb[j];
        is_bf = true;
        want = (size_t)B * n->wp[j];
    } else if (!std::strcmp(name, "wb") && j >= 0 && j < L) {
        src = n->Wb[j];
        is_bf = true;
        want = (size_t)n->wp[j] * n->w[j + 1];
    } else {
        return mfail(n, MPLB_E_ARG, std::string("unknown tensor ") + name);
    }
    if (count != want) return mfail(n, MPLB_E_ARG, "tensor size mismatch");
    if (is_bf) {
        std::vector<uint16_t> h(count);
        MCHK(n, hipMemcpy(h.data(), src, 2 * count, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < count; i++) {
            const uint32_t u = (uint32_t)h[i] << 16;
            std::memcpy(&out[i], &u, 4);
        }
    } else {
        MCHK(n, hipMemcpy(out, src, 4 * count, hipMemcpyDeviceToHost));
    }
    return 0;
}

}  // extern "C"
