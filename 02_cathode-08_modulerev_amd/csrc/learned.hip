// learned.hip -- learned initial guess for the Poisson correction solve.
//
// A 6-layer fully convolutional network, 3x3 kernels, 32 hidden channels:
//   in  3 feature planes (padded to 16 channels)
//   16 -> 32 -> 32 -> 32 -> 32 -> 32 -> 16 (ReLU after the first five; only
//   output channel 0 is used)
// Weights and activations are BF16, every product is accumulated in FP32 on
// the matrix cores (v_mfma_f32_16x16x16_bf16): a layer is an implicit GEMM
// with M = 16 pixels of one row, N = 16 output channels, K = 9 taps x 16
// input channels per step.  Training runs on the device with the library's
// own kernels: the backward data pass is the same MFMA convolution with the
// weights transposed and flipped, the weight gradient is a GEMM over pixels
// on the matrix cores (per-band partials summed in a fixed order, so
// gradients are reproducible), and the optimiser is Adam on FP32 master
// weights.
//
// Activations are NHWC over the whole (ni+2) x (nj+2) grid including the
// ghost ring, which stays zero (the convolution's zero padding).
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "common.hpp"
#include "learned.hpp"

namespace cathode {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16_t;

constexpr int CNN_L = 6;
constexpr int CNN_CIN[CNN_L] = {16, 32, 32, 32, 32, 32};
constexpr int CNN_COUT[CNN_L] = {32, 32, 32, 32, 32, 16};
constexpr int CNN_FEAT = 3;  // real input channels
constexpr int CNN_MAXC = 32;

// float -> bf16, round to nearest even (NaN kept quiet)
__device__ inline bf16_t f2bf(float f)
{
    unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__device__ inline float bf2f(bf16_t h) { return __uint_as_float(((unsigned)h) << 16); }

// parameter layout (FP32): per layer W[cout][9][cin] then b[cout]
struct CnnLayout {
    size_t w_off[CNN_L], b_off[CNN_L], total;
    CnnLayout()
    {
        size_t o = 0;
        for (int l = 0; l < CNN_L; l++) {
            w_off[l] = o;
            o += (size_t)CNN_COUT[l] * 9 * CNN_CIN[l];
            b_off[l] = o;
            o += CNN_COUT[l];
        }
        total = o;
    }
};
static const CnnLayout g_lay;

// ------------------------------------------------------------------ kernels
// feature planes (FP64, pitch nj+2, interior only) -> BF16 NHWC, 16 channels
__global__ void k_cnn_features(int ni, int nj, const double* __restrict__ feat, bf16_t* __restrict__ act)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    const int W2 = nj + 2;
    if (j >= W2 || i >= ni + 2) return;
    const size_t p = (size_t)i * W2 + j, plane = (size_t)(ni + 2) * W2;
    const bool in = i >= 1 && i <= ni && j >= 1 && j <= nj;
    for (int c = 0; c < 16; c++) {
        const float v = (in && c < CNN_FEAT) ? (float)feat[c * plane + p] : 0.f;
        act[p * 16 + c] = f2bf(v);
    }
}

// One convolution layer on MFMA.  Block = 4 wavefronts, one image row each,
// 16 consecutive pixels per wavefront; every wavefront computes all COUT
// output channels of its 16 pixels.
// MODE 0: + bias, ReLU, BF16 NHWC out
// MODE 1: + bias, output channel 0 to an FP32 plane (last layer)
// MODE 2: backward data: no bias, times 1[mask > 0] (the forward input of
//         the layer the gradient flows into), BF16 NHWC out
template <int CIN, int COUT, int MODE>
__global__ void __launch_bounds__(256) k_cnn_conv(int ni, int nj, const bf16_t* __restrict__ X,
                                                  const bf16x4* __restrict__ Wp, const float* __restrict__ bias,
                                                  bf16_t* __restrict__ Y, float* __restrict__ yplane,
                                                  const bf16_t* __restrict__ mask)
{
    constexpr int NB = COUT / 16, KC = CIN / 16;
    const int lane = threadIdx.x;  // 0..63
    const int i = blockIdx.y * 4 + threadIdx.y + 1;
    const int j0 = blockIdx.x * 16 + 1;
    const int W2 = nj + 2;
    if (i > ni) return;  // whole wavefront (no barriers below)
    const int m = lane & 15, g = lane >> 4;
    f32x4 acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
    const bool pix_ok = j0 + m <= nj;
#pragma unroll
    for (int tap = 0; tap < 9; tap++) {
        const int ii = i + tap / 3 - 1, jj = j0 + m + tap % 3 - 1;
        const size_t base = ((size_t)ii * W2 + jj) * CIN + 4 * g;
#pragma unroll
        for (int kc = 0; kc < KC; kc++) {
            bf16x4 a = bf16x4{0, 0, 0, 0};
            if (pix_ok) a = *reinterpret_cast<const bf16x4*>(X + base + kc * 16);
#pragma unroll
            for (int nb = 0; nb < NB; nb++) {
                const bf16x4 b = Wp[((tap * KC + kc) * NB + nb) * 64 + lane];
                acc[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[nb], 0, 0, 0);
            }
        }
    }
    // D[pixel 4g + r][channel nb*16 + m]
#pragma unroll
    for (int nb = 0; nb < NB; nb++) {
        const int co = nb * 16 + m;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int j = j0 + 4 * g + r;
            if (j > nj) continue;
            const size_t p = (size_t)i * W2 + j;
            float v = acc[nb][r];
            if (MODE == 0) {
                v += bias[co];
                Y[p * COUT + co] = f2bf(v > 0.f ? v : 0.f);
            } else if (MODE == 1) {
                if (co == 0) yplane[p] = v + bias[0];
            } else {
                Y[p * COUT + co] = bf2f(mask[p * COUT + co]) > 0.f ? f2bf(v) : (bf16_t)0;
            }
        }
    }
}

// forward weights -> MFMA B fragments: Wp[tap][kc][nb][lane] = 4 x bf16 of
// W[co = nb*16 + lane%16][tap][ci = kc*16 + 4*(lane/16) + t]
// transposed = 1: fragments of the backward-data convolution (CIN and COUT
// swapped, taps flipped): W[co = ci_b][8 - tap][ci = co_b]
__global__ void k_cnn_pack(const float* __restrict__ W, int cin, int cout, int transposed, bf16_t* __restrict__ Wp)
{
    const int ci_n = transposed ? cout : cin, co_n = transposed ? cin : cout;
    const int KC = ci_n / 16, NB = co_n / 16;
    const int total = 9 * KC * NB * 64;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < total; q += gridDim.x * blockDim.x) {
        const int lane = q % 64, nb = (q / 64) % NB, kc = (q / (64 * NB)) % KC, tap = q / (64 * NB * KC);
        const int n = nb * 16 + lane % 16;
        for (int t = 0; t < 4; t++) {
            const int k = kc * 16 + 4 * (lane / 16) + t;
            float v;
            if (!transposed) v = W[((size_t)n * 9 + tap) * cin + k];
            else v = W[((size_t)k * 9 + (8 - tap)) * cin + n];
            Wp[(size_t)q * 4 + t] = f2bf(v);
        }
    }
}

// loss = mean over interior cells of (y - target)^2; dY (16-channel NHWC,
// channel 0) = 2 (y - target) / N; block partial sums of the squared error
__global__ void k_cnn_loss(int ni, int nj, const float* __restrict__ y, const double* __restrict__ target,
                           bf16_t* __restrict__ dY, double* __restrict__ part)
{
    __shared__ double lds[8];
    const int W2 = nj + 2;
    const double N = (double)ni * nj;
    double s = 0.0;
    const int n = ni * nj;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const int i = 1 + q / nj, j = 1 + q % nj;
        const size_t p = (size_t)i * W2 + j;
        const double d = (double)y[p] - target[p];
        s += d * d;
        dY[p * 16] = f2bf((float)(2.0 * d / N));
    }
    const double t = block_sum<256>(s, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void k_cnn_sum(const double* __restrict__ part, int nb, double scale, double* __restrict__ out)
{
    __shared__ double lds[8];
    double s = 0.0;
    for (int k = threadIdx.x; k < nb; k += blockDim.x) s += part[k];
    s = block_sum<256>(s, lds);
    if (threadIdx.x == 0) *out = s * scale;
}

// Weight and bias gradient of one layer, partial over a band of rows:
// dW[co][tap][ci] = sum_p dY[p][co] X[p + off(tap)][ci], db[co] = sum_p dY[p][co].
// Each block stages 16 pixels at a time (dY and the 3x3 neighbourhood of X)
// in LDS.  Thread t < 9*CIN owns the pair (tap, ci) and accumulates it for
// every co; the next COUT threads own the biases.
template <int CIN, int COUT>
constexpr int wgrad_threads()
{
    return ((9 * CIN + COUT) + 63) / 64 * 64;
}

template <int CIN, int COUT>
__global__ void __launch_bounds__((wgrad_threads<CIN, COUT>()))
    k_cnn_wgrad(int ni, int nj, int rows_per_block, const bf16_t* __restrict__ X, const bf16_t* __restrict__ dY,
                float* __restrict__ part)
{
    constexpr int NT = wgrad_threads<CIN, COUT>(), NW = 9 * CIN * COUT, NO = NW + COUT, PB = 16;
    __shared__ float sdy[PB][COUT];
    __shared__ float sx[PB][9][CIN];
    const int W2 = nj + 2;
    const int t = threadIdx.x;
    const int i0 = 1 + blockIdx.x * rows_per_block, i1 = min(ni, i0 + rows_per_block - 1);
    const int tap = t / CIN, ci = t % CIN;  // valid for t < 9*CIN
    float acc[COUT];
#pragma unroll
    for (int k = 0; k < COUT; k++) acc[k] = 0.f;
    float bacc = 0.f;
    const int npix = (i1 >= i0) ? (i1 - i0 + 1) * nj : 0;
    for (int p0 = 0; p0 < npix; p0 += PB) {
        __syncthreads();
        for (int q = t; q < PB * COUT; q += NT) {
            const int px = q / COUT, cc = q % COUT, pp = p0 + px;
            float v = 0.f;
            if (pp < npix) {
                const int i = i0 + pp / nj, j = 1 + pp % nj;
                v = bf2f(dY[((size_t)i * W2 + j) * COUT + cc]);
            }
            sdy[px][cc] = v;
        }
        for (int q = t; q < PB * 9 * CIN; q += NT) {
            const int px = q / (9 * CIN), tp = (q / CIN) % 9, cc = q % CIN, pp = p0 + px;
            float v = 0.f;
            if (pp < npix) {
                const int i = i0 + pp / nj + tp / 3 - 1, j = 1 + pp % nj + tp % 3 - 1;
                v = bf2f(X[((size_t)i * W2 + j) * CIN + cc]);
            }
            sx[px][tp][cc] = v;
        }
        __syncthreads();
        if (t < 9 * CIN) {
            for (int px = 0; px < PB; px++) {
                const float xv = sx[px][tap][ci];
#pragma unroll
                for (int co = 0; co < COUT; co++) acc[co] += sdy[px][co] * xv;
            }
        } else if (t < 9 * CIN + COUT) {
            const int co = t - 9 * CIN;
            for (int px = 0; px < PB; px++) bacc += sdy[px][co];
        }
    }
    float* out = part + (size_t)blockIdx.x * NO;
    if (t < 9 * CIN) {
#pragma unroll
        for (int co = 0; co < COUT; co++) out[(size_t)co * 9 * CIN + tap * CIN + ci] = acc[co];
    } else if (t < 9 * CIN + COUT) {
        out[NW + (t - 9 * CIN)] = bacc;
    }
}

// The same partials on the matrix cores: dW is a GEMM with M = COUT output
// channels, N = 9 CIN (tap, input channel) columns and K = the band's pixels.
// A chunk of 64 consecutive pixels of one row is staged in LDS with the
// pixel index contiguous: sA[co][px] = dY, sB[tap * CIN + ci][px] = X at the
// tap's neighbour (zero beyond the row's end).  Each of the 4 wavefronts owns
// every 4th 16 x 16 output tile and runs v_mfma_f32_16x16x16_bf16 over the
// chunk's 4 steps of 16 pixels; the products of BF16 values are exact in
// FP32, so only the summation order differs from k_cnn_wgrad.  The biases
// are FP32 sums of sA.  Same partial layout as k_cnn_wgrad.
constexpr int WG_PX = 64;
template <int CIN, int COUT>
__global__ void __launch_bounds__(256)
    k_cnn_wgrad_mfma(int ni, int nj, int rows_per_block, const bf16_t* __restrict__ X, const bf16_t* __restrict__ dY,
                     float* __restrict__ part)
{
    constexpr int NCOL = 9 * CIN, MT = COUT / 16, NTL = NCOL / 16, T = MT * NTL, TPW = (T + 3) / 4;
    constexpr int NW = 9 * CIN * COUT, NO = NW + COUT;
    constexpr int LD = WG_PX + 4;  // row pitch of the staged tiles (bf16), 8-byte aligned rows
    __shared__ __align__(16) bf16_t sA[COUT * LD];
    __shared__ __align__(16) bf16_t sB[NCOL * LD];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int W2 = nj + 2;
    const int i0 = 1 + blockIdx.x * rows_per_block, i1 = min(ni, i0 + rows_per_block - 1);
    const int m = lane & 15, g = lane >> 4;
    f32x4 acc[TPW];
#pragma unroll
    for (int q = 0; q < TPW; q++) acc[q] = f32x4{0.f, 0.f, 0.f, 0.f};
    float bacc = 0.f;
    const int chunks_per_row = (nj + WG_PX - 1) / WG_PX;
    const int nchunks = (i1 >= i0) ? (i1 - i0 + 1) * chunks_per_row : 0;
    for (int c = 0; c < nchunks; c++) {
        const int i = i0 + c / chunks_per_row, jb = 1 + (c % chunks_per_row) * WG_PX;
        __syncthreads();
        // dY: thread per (pixel, 4 channels)
        for (int q = t; q < WG_PX * (COUT / 4); q += 256) {
            const int px = q / (COUT / 4), c4 = (q % (COUT / 4)) * 4, j = jb + px;
            bf16x4 v = bf16x4{0, 0, 0, 0};
            if (j <= nj) v = *reinterpret_cast<const bf16x4*>(dY + ((size_t)i * W2 + j) * COUT + c4);
#pragma unroll
            for (int e = 0; e < 4; e++) sA[(c4 + e) * LD + px] = (bf16_t)v[e];
        }
        // X neighbourhoods: thread per (pixel, tap, 4 channels)
        for (int q = t; q < WG_PX * 9 * (CIN / 4); q += 256) {
            const int px = q / (9 * (CIN / 4)), rem = q % (9 * (CIN / 4));
            const int tap = rem / (CIN / 4), c4 = (rem % (CIN / 4)) * 4, j = jb + px;
            bf16x4 v = bf16x4{0, 0, 0, 0};
            if (j <= nj) {
                const int ii = i + tap / 3 - 1, jj = j + tap % 3 - 1;
                v = *reinterpret_cast<const bf16x4*>(X + ((size_t)ii * W2 + jj) * CIN + c4);
            }
#pragma unroll
            for (int e = 0; e < 4; e++) sB[(tap * CIN + c4 + e) * LD + px] = (bf16_t)v[e];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < TPW; q++) {
            const int tile = wv + 4 * q;
            if (tile >= T) break;
            const int mt = tile / NTL, nt = tile % NTL;
#pragma unroll
            for (int ks = 0; ks < WG_PX / 16; ks++) {
                // A[co][pixel], B[pixel][column]: lane supplies 4 consecutive pixels
                const bf16x4 a = *reinterpret_cast<const bf16x4*>(&sA[(mt * 16 + m) * LD + ks * 16 + 4 * g]);
                const bf16x4 b = *reinterpret_cast<const bf16x4*>(&sB[(nt * 16 + m) * LD + ks * 16 + 4 * g]);
                acc[q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[q], 0, 0, 0);
            }
        }
        if (t < COUT)
            for (int px = 0; px < WG_PX; px++) bacc += bf2f(sA[t * LD + px]);
    }
    float* out = part + (size_t)blockIdx.x * NO;
#pragma unroll
    for (int q = 0; q < TPW; q++) {
        const int tile = wv + 4 * q;
        if (tile >= T) break;
        const int mt = tile / NTL, nt = tile % NTL;
        // D[row 4g + r][column m]: row = output channel, column = (tap, ci)
#pragma unroll
        for (int r = 0; r < 4; r++) out[(size_t)(mt * 16 + 4 * g + r) * NCOL + nt * 16 + m] = acc[q][r];
    }
    if (t < COUT) out[NW + t] = bacc;
}

// grad[o] (+)= sum over blocks of part[b][o], fixed order; o < NW are the
// weights (same order as the parameter layout), then the biases
__global__ void k_cnn_wgrad_reduce(const float* __restrict__ part, int nblocks, int no, float* __restrict__ grad,
                                   int accumulate)
{
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= no) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; b++) s += part[(size_t)b * no + o];
    grad[o] = accumulate ? grad[o] + (float)s : (float)s;
}

__global__ void k_cnn_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                           float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float bc1, float bc2,
                           float gscale)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k] * gscale;
    const float mk = b1 * m[k] + (1.f - b1) * gk;
    const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    w[k] -= lr * (mk / bc1) / (sqrtf(vk / bc2) + eps);
}

__global__ void k_cnn_out_plane(int ni, int nj, const float* __restrict__ y, double* __restrict__ out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    const int W2 = nj + 2;
    if (j >= W2 || i >= ni + 2) return;
    const size_t p = (size_t)i * W2 + j;
    out[p] = (i >= 1 && i <= ni && j >= 1 && j <= nj) ? (double)y[p] : 0.0;
}

// ------------------------------------------------------------------ network object
struct CnnNet {
    DevBuf<float> w, g, m, v, part;       // master weights, gradients, Adam moments, wgrad partials
    DevBuf<bf16_t> wf[CNN_L], wb[CNN_L];  // MFMA fragments: forward and backward data
    DevBuf<bf16_t> act[CNN_L];            // layer inputs (act[0] = features)
    DevBuf<bf16_t> ga, gb, g6;            // gradient activations
    DevBuf<float> y;                      // output plane
    DevBuf<double> dpart, scal;
    int ni = -1, nj = -1;
    int adam_t = 0;
};

static size_t frag_elems(int cin, int cout) { return (size_t)9 * (cin / 16) * (cout / 16) * 64 * 4; }

static void cnn_pack(CnnNet& N, hipStream_t s)
{
    for (int l = 0; l < CNN_L; l++) {
        hipLaunchKernelGGL(k_cnn_pack, dim3(64), dim3(256), 0, s, N.w.p + g_lay.w_off[l], CNN_CIN[l], CNN_COUT[l], 0,
                           N.wf[l].p);
        if (l > 0)
            hipLaunchKernelGGL(k_cnn_pack, dim3(64), dim3(256), 0, s, N.w.p + g_lay.w_off[l], CNN_CIN[l], CNN_COUT[l],
                               1, N.wb[l].p);
    }
    CATHODE_LAUNCH_CHECK();
}

static void cnn_shape(CnnNet& N, int ni, int nj)
{
    if (N.ni == ni && N.nj == nj) return;
    const size_t px = (size_t)(ni + 2) * (nj + 2);
    for (int l = 0; l < CNN_L; l++) {
        N.act[l].alloc(px * CNN_CIN[l]);
        CATHODE_HIP(hipMemset(N.act[l].p, 0, px * CNN_CIN[l] * sizeof(bf16_t)));
    }
    N.ga.alloc(px * CNN_MAXC);
    N.gb.alloc(px * CNN_MAXC);
    N.g6.alloc(px * 16);
    CATHODE_HIP(hipMemset(N.ga.p, 0, px * CNN_MAXC * sizeof(bf16_t)));
    CATHODE_HIP(hipMemset(N.gb.p, 0, px * CNN_MAXC * sizeof(bf16_t)));
    CATHODE_HIP(hipMemset(N.g6.p, 0, px * 16 * sizeof(bf16_t)));
    N.y.alloc(px);
    CATHODE_HIP(hipMemset(N.y.p, 0, px * sizeof(float)));
    CATHODE_HIP(hipDeviceSynchronize());  // null-stream memsets before any stream use
    // weight-gradient partials: 2-row bands, the largest layer's outputs
    N.part.alloc((size_t)cdiv(ni, 2) * (9 * CNN_MAXC * CNN_MAXC + CNN_MAXC));
    N.dpart.alloc(256);
    N.ni = ni;
    N.nj = nj;
}

static dim3 conv_grid(int ni, int nj) { return dim3(cdiv(nj, 16), cdiv(ni, 4)); }

// forward pass into the activation planes act[0..CNN_L-1] and the output y
// (zero ghosts); the weight fragments are packed whenever the weights change
static void cnn_forward(const CnnNet& N, int ni, int nj, const double* feat, bf16_t* const* act, float* y,
                        hipStream_t s)
{
    const dim3 gp(cdiv(nj + 2, 256), ni + 2), gc = conv_grid(ni, nj), blk(64, 4);
    hipLaunchKernelGGL(k_cnn_features, gp, dim3(256), 0, s, ni, nj, feat, act[0]);
    auto W = [&](int l) { return reinterpret_cast<const bf16x4*>(N.wf[l].p); };
    auto B = [&](int l) { return (const float*)(N.w.p + g_lay.b_off[l]); };
    hipLaunchKernelGGL((k_cnn_conv<16, 32, 0>), gc, blk, 0, s, ni, nj, act[0], W(0), B(0), act[1], (float*)nullptr,
                       (const bf16_t*)nullptr);
    for (int l = 1; l < CNN_L - 1; l++)
        hipLaunchKernelGGL((k_cnn_conv<32, 32, 0>), gc, blk, 0, s, ni, nj, act[l], W(l), B(l), act[l + 1],
                           (float*)nullptr, (const bf16_t*)nullptr);
    hipLaunchKernelGGL((k_cnn_conv<32, 16, 1>), gc, blk, 0, s, ni, nj, act[CNN_L - 1], W(CNN_L - 1), B(CNN_L - 1),
                       (bf16_t*)nullptr, y, (const bf16_t*)nullptr);
    CATHODE_LAUNCH_CHECK();
}

static void cnn_forward(CnnNet& N, int ni, int nj, const double* feat, hipStream_t s)
{
    bf16_t* act[CNN_L];
    for (int l = 0; l < CNN_L; l++) act[l] = N.act[l].p;
    cnn_forward(N, ni, nj, feat, act, N.y.p, s);
}

// the weight gradient on the matrix cores (default) or on the VALU
// (cathode_cnn_set_wgrad_valu, for comparison)
static bool g_wgrad_valu = false;

template <int CIN, int COUT>
static void wgrad(CnnNet& N, int l, int ni, int nj, const bf16_t* X, const bf16_t* dY, int accumulate, hipStream_t s)
{
    constexpr int NO = 9 * CIN * COUT + COUT;
    const int rows = 2;  // N.part holds cdiv(ni, 2) bands (cnn_shape)
    const int nblocks = cdiv(ni, rows);
    if (g_wgrad_valu)
        hipLaunchKernelGGL((k_cnn_wgrad<CIN, COUT>), dim3(nblocks), dim3(wgrad_threads<CIN, COUT>()), 0, s, ni, nj,
                           rows, X, dY, N.part.p);
    else
        hipLaunchKernelGGL((k_cnn_wgrad_mfma<CIN, COUT>), dim3(nblocks), dim3(256), 0, s, ni, nj, rows, X, dY,
                           N.part.p);
    // weights and biases are contiguous in the layout (W then b)
    hipLaunchKernelGGL(k_cnn_wgrad_reduce, dim3(cdiv(NO, 256)), dim3(256), 0, s, (const float*)N.part.p, nblocks, NO,
                       N.g.p + g_lay.w_off[l], accumulate);
}

// forward + loss + backward for one sample; gradients (accumulate = 1: added)
static void cnn_fwd_bwd(CnnNet& N, int ni, int nj, const double* feat, const double* target, int accumulate,
                        double* loss_dev, hipStream_t s)
{
    cnn_forward(N, ni, nj, feat, s);
    const int nbl = 256;  // N.dpart size (cnn_shape)
    hipLaunchKernelGGL(k_cnn_loss, dim3(nbl), dim3(256), 0, s, ni, nj, (const float*)N.y.p, target, N.g6.p,
                       N.dpart.p);
    hipLaunchKernelGGL(k_cnn_sum, dim3(1), dim3(256), 0, s, (const double*)N.dpart.p, nbl, 1.0 / ((double)ni * nj),
                       loss_dev);
    const dim3 gc = conv_grid(ni, nj), blk(64, 4);
    auto WB = [&](int l) { return reinterpret_cast<const bf16x4*>(N.wb[l].p); };
    // last layer (32 -> 16)
    wgrad<32, 16>(N, CNN_L - 1, ni, nj, N.act[CNN_L - 1].p, N.g6.p, accumulate, s);
    hipLaunchKernelGGL((k_cnn_conv<16, 32, 2>), gc, blk, 0, s, ni, nj, N.g6.p, WB(CNN_L - 1), (const float*)nullptr,
                       N.ga.p, (float*)nullptr, N.act[CNN_L - 1].p);
    bf16_t* cur = N.ga.p;
    bf16_t* nxt = N.gb.p;
    for (int l = CNN_L - 2; l >= 1; l--) {
        wgrad<32, 32>(N, l, ni, nj, N.act[l].p, cur, accumulate, s);
        hipLaunchKernelGGL((k_cnn_conv<32, 32, 2>), gc, blk, 0, s, ni, nj, cur, WB(l), (const float*)nullptr, nxt,
                           (float*)nullptr, N.act[l].p);
        std::swap(cur, nxt);
    }
    wgrad<16, 32>(N, 0, ni, nj, N.act[0].p, cur, accumulate, s);
    CATHODE_LAUNCH_CHECK();
}

}  // namespace cathode

using namespace cathode;

// ------------------------------------------------------------------ internal API for the NDD step
namespace cathode {

struct CnnScratch {
    DevBuf<bf16_t> act[CNN_L];
    DevBuf<float> y;
    int ni = -1, nj = -1;
};

CnnScratch* cnn_scratch_new() { return new CnnScratch(); }
void cnn_scratch_free(CnnScratch* sc) { delete sc; }

void cnn_infer(void* net, int ni, int nj, const double* feat, double* out, CnnScratch* sc, hipStream_t s)
{
    const CnnNet& N = *(const CnnNet*)net;
    if (sc->ni != ni || sc->nj != nj) {
        // zero ghosts and padding, ordered on the caller's stream
        const size_t px = (size_t)(ni + 2) * (nj + 2);
        for (int l = 0; l < CNN_L; l++) {
            sc->act[l].alloc(px * CNN_CIN[l]);
            CATHODE_HIP(hipMemsetAsync(sc->act[l].p, 0, px * CNN_CIN[l] * sizeof(bf16_t), s));
        }
        sc->y.alloc(px);
        CATHODE_HIP(hipMemsetAsync(sc->y.p, 0, px * sizeof(float), s));
        sc->ni = ni;
        sc->nj = nj;
    }
    bf16_t* act[CNN_L];
    for (int l = 0; l < CNN_L; l++) act[l] = sc->act[l].p;
    cnn_forward(N, ni, nj, feat, act, sc->y.p, s);
    hipLaunchKernelGGL(k_cnn_out_plane, dim3(cdiv(nj + 2, 256), ni + 2), dim3(256), 0, s, ni, nj,
                       (const float*)sc->y.p, out);
    CATHODE_LAUNCH_CHECK();
}

}  // namespace cathode

extern "C" {

int cathode_cnn_num_params(void) { return (int)g_lay.total; }

int cathode_cnn_set_wgrad_valu(int on)
{
    g_wgrad_valu = on != 0;
    return CATHODE_OK;
}

int cathode_cnn_create(int seed, void** net)
{
    CATHODE_GUARD_BEGIN
    if (!require_init()) return CATHODE_ERR_NOINIT;
    if (!net) {
        set_error("cnn_create: null argument");
        return CATHODE_ERR_ARG;
    }
    std::unique_ptr<CnnNet> N(new CnnNet());
    const size_t n = g_lay.total;
    std::vector<float> w(n, 0.f);
    std::mt19937_64 rng((unsigned long long)seed);
    for (int l = 0; l < CNN_L; l++) {
        const int cin_real = l == 0 ? CNN_FEAT : CNN_CIN[l];
        const int cout_real = l == CNN_L - 1 ? 1 : CNN_COUT[l];
        const double sd = (l == CNN_L - 1 ? 0.5 : 1.0) * std::sqrt(2.0 / (9.0 * cin_real));
        std::normal_distribution<double> nd(0.0, sd);
        for (int co = 0; co < cout_real; co++)
            for (int t = 0; t < 9; t++)
                for (int ci = 0; ci < cin_real; ci++)
                    w[g_lay.w_off[l] + ((size_t)co * 9 + t) * CNN_CIN[l] + ci] = (float)nd(rng);
    }
    N->w.alloc(n);
    N->g.alloc(n);
    N->m.alloc(n);
    N->v.alloc(n);
    CATHODE_HIP(hipMemcpy(N->w.p, w.data(), n * sizeof(float), hipMemcpyHostToDevice));
    CATHODE_HIP(hipMemset(N->g.p, 0, n * sizeof(float)));
    CATHODE_HIP(hipMemset(N->m.p, 0, n * sizeof(float)));
    CATHODE_HIP(hipMemset(N->v.p, 0, n * sizeof(float)));
    CATHODE_HIP(hipDeviceSynchronize());  // null-stream memsets before any stream use
    for (int l = 0; l < CNN_L; l++) {
        N->wf[l].alloc(frag_elems(CNN_CIN[l], CNN_COUT[l]));
        N->wb[l].alloc(frag_elems(CNN_COUT[l], CNN_CIN[l]));
    }
    N->scal.alloc(4);
    cnn_pack(*N, ctx().stream);
    CATHODE_HIP(hipStreamSynchronize(ctx().stream));
    *net = N.release();
    return CATHODE_OK;
    CATHODE_GUARD_END
}

int cathode_cnn_destroy(void* net)
{
    delete (CnnNet*)net;
    return CATHODE_OK;
}

int cathode_cnn_get_params(void* net, float* host)
{
    CATHODE_GUARD_BEGIN
    if (!net || !host) {
        set_error("cnn_get_params: null argument");
        return CATHODE_ERR_ARG;
    }
    CATHODE_HIP(hipMemcpy(host, ((CnnNet*)net)->w.p, g_lay.total * sizeof(float), hipMemcpyDeviceToHost));
    return CATHODE_OK;
    CATHODE_GUARD_END
}

int cathode_cnn_set_params(void* net, const float* host)
{
    CATHODE_GUARD_BEGIN
    if (!net || !host) {
        set_error("cnn_set_params: null argument");
        return CATHODE_ERR_ARG;
    }
    CnnNet& N = *(CnnNet*)net;
    CATHODE_HIP(hipMemcpy(N.w.p, host, g_lay.total * sizeof(float), hipMemcpyHostToDevice));
    // the fragments the passes read are packed here, so inference (which may
    // run for several handles at once) only reads the network
    cnn_pack(N, ctx().stream);
    CATHODE_HIP(hipStreamSynchronize(ctx().stream));
    return CATHODE_OK;
    CATHODE_GUARD_END
}

int cathode_cnn_forward_dev(void* net, int ni, int nj, const double* feat, double* out)
{
    CATHODE_GUARD_BEGIN
    if (!require_init()) return CATHODE_ERR_NOINIT;
    if (!net || !feat || !out || ni < 1 || nj < 1) {
        set_error("cnn_forward: invalid arguments");
        return CATHODE_ERR_ARG;
    }
    hipStream_t s = ctx().stream;
    CnnNet& N = *(CnnNet*)net;
    cnn_shape(N, ni, nj);
    cnn_forward(N, ni, nj, feat, s);
    hipLaunchKernelGGL(k_cnn_out_plane, dim3(cdiv(nj + 2, 256), ni + 2), dim3(256), 0, s, ni, nj,
                       (const float*)N.y.p, out);
    CATHODE_LAUNCH_CHECK();
    CATHODE_HIP(hipStreamSynchronize(s));
    return CATHODE_OK;
    CATHODE_GUARD_END
}

int cathode_cnn_grad_dev(void* net, int ni, int nj, const double* feat, const double* target, double* loss,
                         float* grad_host)
{
    CATHODE_GUARD_BEGIN
    if (!require_init()) return CATHODE_ERR_NOINIT;
    if (!net || !feat || !target || ni < 1 || nj < 1) {
        set_error("cnn_grad: invalid arguments");
        return CATHODE_ERR_ARG;
    }
    CnnNet& N = *(CnnNet*)net;
    hipStream_t s = ctx().stream;
    cnn_shape(N, ni, nj);
    cnn_fwd_bwd(N, ni, nj, feat, target, 0, N.scal.p, s);
    double l = 0.0;
    CATHODE_HIP(hipMemcpyAsync(&l, N.scal.p, sizeof(double), hipMemcpyDeviceToHost, s));
    if (grad_host)
        CATHODE_HIP(hipMemcpyAsync(grad_host, N.g.p, g_lay.total * sizeof(float), hipMemcpyDeviceToHost, s));
    CATHODE_HIP(hipStreamSynchronize(s));
    if (loss) *loss = l;
    return CATHODE_OK;
    CATHODE_GUARD_END
}

int cathode_cnn_train_step_dev(void* net, int ni, int nj, int batch, const double* feat, const double* target,
                               float lr, double* loss)
{
    CATHODE_GUARD_BEGIN
    if (!require_init()) return CATHODE_ERR_NOINIT;
    if (!net || !feat || !target || ni < 1 || nj < 1 || batch < 1 || !(lr > 0.f)) {
        set_error("cnn_train_step: invalid arguments");
        return CATHODE_ERR_ARG;
    }
    CnnNet& N = *(CnnNet*)net;
    hipStream_t s = ctx().stream;
    cnn_shape(N, ni, nj);
    const size_t plane = (size_t)(ni + 2) * (nj + 2);
    N.scal.alloc(4 + batch);
    for (int b = 0; b < batch; b++)
        cnn_fwd_bwd(N, ni, nj, feat + (size_t)b * CNN_FEAT * plane, target + (size_t)b * plane, b > 0,
                    N.scal.p + 4 + b, s);
    N.adam_t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.f - std::pow(b1, (float)N.adam_t), bc2 = 1.f - std::pow(b2, (float)N.adam_t);
    const int n = (int)g_lay.total;
    hipLaunchKernelGGL(k_cnn_adam, dim3(cdiv(n, 256)), dim3(256), 0, s, N.w.p, (const float*)N.g.p, N.m.p, N.v.p, n,
                       lr, b1, b2, 1e-8f, bc1, bc2, 1.f / batch);
    CATHODE_LAUNCH_CHECK();
    cnn_pack(N, s);
    std::vector<double> ls(batch);
    CATHODE_HIP(hipMemcpyAsync(ls.data(), N.scal.p + 4, batch * sizeof(double), hipMemcpyDeviceToHost, s));
    CATHODE_HIP(hipStreamSynchronize(s));
    double tot = 0.0;
    for (double v : ls) tot += v;
    if (loss) *loss = tot / batch;
    return CATHODE_OK;
    CATHODE_GUARD_END
}

}  // extern "C"
