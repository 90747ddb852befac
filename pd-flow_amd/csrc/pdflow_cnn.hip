// Kernels of the learned warm start: a small BF16 U-Net evaluated and trained on gfx950 matrix
// cores (v_mfma_f32_16x16x16_bf16, FP32 accumulation).  No counterpart in the reference.
//
// Tensors are NHWC (channels fastest), activations and their gradients in BF16, weights kept in
// FP32 (master copy, Adam state, gradients) and cast to BF16 for the products.  A 3x3 "same"
// convolution is an implicit GEMM: M = pixels, N = output channels, K = 9 * Cin ordered
// (tap, channel) with tap = 3*(dy+1) + (dx+1), padded with zero weights to a multiple of 16.
// Weights are stored transposed, wt[co][k], so one lane's four consecutive k are one 8-byte
// load.  The data gradient is the same convolution with the flipped, channel-transposed
// weights; the weight gradient is a GEMM over pixels staged through LDS.
//
// MFMA 16x16x16 operand layout (lane l of 64): A[i][k] with i = l % 16, k = 4*(l/16) + 0..3;
// B[k][j] with j = l % 16, k = 4*(l/16) + 0..3; D[i][j] with j = l % 16, i = 4*(l/16) + 0..3.
// pdf_cnn_mfma_probe exposes one product so a test checks this layout.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "pdflow_cnn_internal.hpp"

namespace pdf {
namespace cnn {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ uint16_t f2bf(float f)
{
    // round to nearest even; NaN stays NaN
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// ------------------------------------------------------------------ layout probe
__global__ void k_probe(const uint16_t *A, const uint16_t *B, float *C)
{
    const int l = threadIdx.x;
    bf16x4 a, b;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        a[j] = (short)A[(l % 16) * 16 + 4 * (l / 16) + j];      // A row-major 16x16
        b[j] = (short)B[(4 * (l / 16) + j) * 16 + l % 16];      // B row-major 16x16
    }
    f32x4 c = {0.f, 0.f, 0.f, 0.f};
    c = mfma(a, b, c);
#pragma unroll
    for (int j = 0; j < 4; j++) C[(4 * (l / 16) + j) * 16 + l % 16] = c[j];
}

// ------------------------------------------------------------------ 3x3 convolution
// WG: 256 threads (4 waves) -> 64 pixels x 32 output channels; wave w owns pixels
// 16w..16w+15 of the tile and both 16-channel blocks.
constexpr int CT_P = 64, CT_C = 32;

__global__ void __launch_bounds__(256) k_conv3x3(ConvArgs a)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int P = a.nb * a.h * a.w;
    const int p0 = blockIdx.x * CT_P + wave * 16;
    const int co0 = blockIdx.y * CT_C;
    const int i = lane % 16, kq = 4 * (lane / 16);
    // this lane's A row: one output pixel
    const int pa = p0 + i;
    const bool pa_ok = pa < P;
    const int xa = pa_ok ? pa % a.w : 0, ya = pa_ok ? (pa / a.w) % a.h : 0, na = pa_ok ? pa / (a.w * a.h) : 0;
    const int kreal = 9 * a.cin;
    f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
    const int cb0 = co0 + i, cb1 = co0 + 16 + i;          // B columns of this lane
    const bool c0_ok = cb0 < a.coutp, c1_ok = cb1 < a.coutp;
    for (int kc = 0; kc < a.kp; kc += 16) {
        const int k = kc + kq;
        bf16x4 av = {0, 0, 0, 0};
        if (pa_ok && k < kreal) {
            const int tap = k / a.cin, ci = k - tap * a.cin;
            const int yy = ya + tap / 3 - 1, xx = xa + tap % 3 - 1;
            if (yy >= 0 && yy < a.h && xx >= 0 && xx < a.w)
                av = *reinterpret_cast<const bf16x4 *>(a.x + ((size_t)(na * a.h + yy) * a.w + xx) * a.cin + ci);
        }
        bf16x4 b0 = {0, 0, 0, 0}, b1 = {0, 0, 0, 0};
        if (c0_ok) b0 = *reinterpret_cast<const bf16x4 *>(a.wt + (size_t)cb0 * a.kp + k);
        if (c1_ok) b1 = *reinterpret_cast<const bf16x4 *>(a.wt + (size_t)cb1 * a.kp + k);
        acc0 = mfma(av, b0, acc0);
        acc1 = mfma(av, b1, acc1);
    }
    // epilogue: lane holds channel co0 + 16b + i for pixels p0 + kq + j
#pragma unroll
    for (int bi = 0; bi < 2; bi++) {
        const int co = co0 + 16 * bi + i;
        if (co >= a.cout) continue;
        const float bias = a.bias ? a.bias[co] : 0.f;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int p = p0 + kq + j;
            if (p >= P) continue;
            float v = (bi == 0 ? acc0[j] : acc1[j]) + bias;
            if (a.relu) v = v > 0.f ? v : 0.f;
            if (a.y16) a.y16[(size_t)p * a.ystride + a.yoff + co] = f2bf(v);
            if (a.y32) a.y32[(size_t)p * a.ystride + a.yoff + co] = v;
        }
    }
}

// LDS-staged variant: a workgroup (4 waves) computes 128 pixels x WN output channels; each wave
// 32 pixels (two 16-row MFMA tiles) x WN channels (WN/16 tiles), so every A fragment feeds WN/16
// products and every B fragment two.  Per K chunk of 32 the im2col rows and the weight rows are
// staged through LDS once for the whole workgroup.  Each thread's staging pixels and channels
// are fixed across chunks, so their coordinates are computed once and the (tap, channel) split
// of k advances incrementally.
constexpr int CS_P = 128, CS_K = 32;

template <int WN>
__global__ void __launch_bounds__(256) k_conv3x3_lds(ConvArgs a)
{
    constexpr int NT = WN / 16;
    constexpr int AG = CS_P * CS_K / 4 / 256;      // A groups of 4 k per thread per chunk (4)
    constexpr int BG = (WN * CS_K / 4 + 255) / 256;
    __shared__ uint16_t sA[CS_P][CS_K + 4];
    __shared__ uint16_t sB[WN][CS_K + 4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int P = a.nb * a.h * a.w;
    const int p0 = blockIdx.x * CS_P, co0 = blockIdx.y * WN;
    const int kreal = 9 * a.cin;
    const int kg = 4 * (tid % (CS_K / 4));          // this thread's k offset within a chunk
    // staging pixels: rows tid/8 + 32 r of the tile
    int sx[AG], sy[AG], sn[AG];
    bool sok[AG];
#pragma unroll
    for (int r = 0; r < AG; r++) {
        const int p = p0 + tid / (CS_K / 4) + 32 * r;
        sok[r] = p < P;
        const int pp = sok[r] ? p : 0;
        sx[r] = pp % a.w;
        sy[r] = (pp / a.w) % a.h;
        sn[r] = pp / (a.w * a.h);
    }
    int tap = kg / a.cin, ci = kg - tap * a.cin;   // (tap, channel) of k = kc + kg
    f32x4 acc[2][NT];
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int t = 0; t < NT; t++) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    // global -> registers for the chunk at kc (the next chunk's loads are in flight while the
    // current chunk's products run), registers -> LDS at the top of each chunk
    bf16x4 ra[AG], rb[BG];
    auto gload = [&](int kc) {
        const int k = kc + kg;
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
#pragma unroll
        for (int r = 0; r < AG; r++) {
            ra[r] = bf16x4{0, 0, 0, 0};
            const int yy = sy[r] + dy, xx = sx[r] + dx;
            if (sok[r] && k < kreal && yy >= 0 && yy < a.h && xx >= 0 && xx < a.w)
                ra[r] = *reinterpret_cast<const bf16x4 *>(a.x + ((size_t)(sn[r] * a.h + yy) * a.w + xx) * a.cin + ci);
        }
#pragma unroll
        for (int r = 0; r < BG; r++) {
            const int e = tid + 256 * r, co = co0 + e / (CS_K / 4);
            rb[r] = bf16x4{0, 0, 0, 0};
            if (e < WN * CS_K / 4 && co < a.coutp && k < a.kp)
                rb[r] = *reinterpret_cast<const bf16x4 *>(a.wt + (size_t)co * a.kp + k);
        }
        ci += CS_K;                          // (tap, channel) of the next chunk
        while (ci >= a.cin) {
            ci -= a.cin;
            tap++;
        }
    };
    gload(0);
    for (int kc = 0; kc < a.kp; kc += CS_K) {
#pragma unroll
        for (int r = 0; r < AG; r++) *reinterpret_cast<bf16x4 *>(&sA[tid / (CS_K / 4) + 32 * r][kg]) = ra[r];
#pragma unroll
        for (int r = 0; r < BG; r++) {
            const int e = tid + 256 * r;
            if (e < WN * CS_K / 4) *reinterpret_cast<bf16x4 *>(&sB[e / (CS_K / 4)][kg]) = rb[r];
        }
        __syncthreads();
        if (kc + CS_K < a.kp) gload(kc + CS_K);
#pragma unroll
        for (int ks = 0; ks < CS_K; ks += 16) {
            const int kq = ks + 4 * (lane / 16);
            bf16x4 af[2], bfr[NT];
#pragma unroll
            for (int m = 0; m < 2; m++) af[m] = *reinterpret_cast<const bf16x4 *>(&sA[wave * 32 + m * 16 + lane % 16][kq]);
#pragma unroll
            for (int t = 0; t < NT; t++) bfr[t] = *reinterpret_cast<const bf16x4 *>(&sB[t * 16 + lane % 16][kq]);
#pragma unroll
            for (int m = 0; m < 2; m++)
#pragma unroll
                for (int t = 0; t < NT; t++) acc[m][t] = mfma(af[m], bfr[t], acc[m][t]);
        }
        __syncthreads();
    }
    // epilogue: lane holds channel co0 + 16 t + lane % 16 for pixels (tile m) 4 * (lane / 16) + j
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int co = co0 + 16 * t + lane % 16;
        if (co >= a.cout) continue;
        const float bias = a.bias ? a.bias[co] : 0.f;
#pragma unroll
        for (int m = 0; m < 2; m++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int p = p0 + wave * 32 + m * 16 + 4 * (lane / 16) + j;
                if (p >= P) continue;
                float v = acc[m][t][j] + bias;
                if (a.relu) v = v > 0.f ? v : 0.f;
                if (a.y16) a.y16[(size_t)p * a.ystride + a.yoff + co] = f2bf(v);
                if (a.y32) a.y32[(size_t)p * a.ystride + a.yoff + co] = v;
            }
    }
}

// ------------------------------------------------------------------ weight gradient
// dwt[co][k] += sum over pixels of dz[p][co] * xcol[p][k].  WG: 256 threads -> 32 channels x
// 64 k over a chunk of pixels, staged 64 pixels at a time through LDS in transposed form so
// that each lane's four consecutive pixels are one 8-byte LDS read.
constexpr int WG_C = 32, WG_K = 64, WG_STEP = 64, WG_CHUNK = 1024;

__global__ void __launch_bounds__(256) k_conv3x3_wgrad(WGradArgs a)
{
    __shared__ uint16_t sZ[WG_C][WG_STEP + 4];     // dz transposed: [co][pixel]
    __shared__ uint16_t sX[WG_K][WG_STEP + 4];     // im2col transposed: [k][pixel]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int P = a.nb * a.h * a.w;
    const int k0 = blockIdx.x * WG_K, co0 = blockIdx.y * WG_C;
    const int pbeg = blockIdx.z * WG_CHUNK, pend = min(P, pbeg + WG_CHUNK);
    const int kreal = 9 * a.cin;
    // wave w: channel block (w & 1), k blocks 2*(w >> 1) and 2*(w >> 1) + 1
    const int cblk = wave & 1, kblk = 2 * (wave >> 1);
    f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
    for (int ps = pbeg; ps < pend; ps += WG_STEP) {
        // stage dz: 64 pixels x 32 channels, 8 per thread (two 4-channel groups)
        for (int e = tid; e < WG_STEP * (WG_C / 4); e += 256) {
            const int pl = e / (WG_C / 4), cg = 4 * (e % (WG_C / 4));
            const int p = ps + pl;
            bf16x4 v = {0, 0, 0, 0};
            if (p < pend && co0 + cg < a.coutp)
                v = *reinterpret_cast<const bf16x4 *>(a.dz + (size_t)p * a.coutp + co0 + cg);
#pragma unroll
            for (int j = 0; j < 4; j++) sZ[cg + j][pl] = (uint16_t)v[j];
        }
        // stage im2col: 64 pixels x 64 k (groups of 4 channels within one tap)
        for (int e = tid; e < WG_STEP * (WG_K / 4); e += 256) {
            const int pl = e / (WG_K / 4), kg = 4 * (e % (WG_K / 4));
            const int p = ps + pl, k = k0 + kg;
            bf16x4 v = {0, 0, 0, 0};
            if (p < pend && k < kreal) {
                const int x = p % a.w, y = (p / a.w) % a.h, n = p / (a.w * a.h);
                const int tap = k / a.cin, ci = k - tap * a.cin;
                const int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
                if (yy >= 0 && yy < a.h && xx >= 0 && xx < a.w)
                    v = *reinterpret_cast<const bf16x4 *>(a.x + ((size_t)(n * a.h + yy) * a.w + xx) * a.cin + ci);
            }
#pragma unroll
            for (int j = 0; j < 4; j++) sX[kg + j][pl] = (uint16_t)v[j];
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < WG_STEP; s += 16) {
            const int pq = s + 4 * (lane / 16);
            bf16x4 av, b0, b1;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                av[j] = (short)sZ[16 * cblk + lane % 16][pq + j];
                b0[j] = (short)sX[16 * kblk + lane % 16][pq + j];
                b1[j] = (short)sX[16 * (kblk + 1) + lane % 16][pq + j];
            }
            acc0 = mfma(av, b0, acc0);
            acc1 = mfma(av, b1, acc1);
        }
        __syncthreads();
    }
    // lane holds k column (block kblk / kblk+1, index lane % 16) for channels 4*(lane/16)+j
#pragma unroll
    for (int bi = 0; bi < 2; bi++) {
        const int k = k0 + 16 * (kblk + bi) + lane % 16;
        if (k >= a.kp) continue;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int co = co0 + 16 * cblk + 4 * (lane / 16) + j;
            if (co >= a.coutp) continue;
            atomicAdd(a.dwt + (size_t)co * a.kp + k, bi == 0 ? acc0[j] : acc1[j]);
        }
    }
}

// bias gradient: db[co] += sum over a chunk of pixels of dz[p][co]
__global__ void __launch_bounds__(256) k_bias_grad(const uint16_t *dz, float *db, int P, int coutp, int cout)
{
    const int co = threadIdx.x % 64 + 64 * blockIdx.y;
    const int lanep = threadIdx.x / 64;
    if (co >= cout) return;
    float s = 0.f;
    for (int p = blockIdx.x * 1024 + lanep; p < min(P, (int)(blockIdx.x + 1) * 1024); p += 4)
        s += bf2f(dz[(size_t)p * coutp + co]);
    atomicAdd(db + co, s);
}

// ------------------------------------------------------------------ elementwise
// dz = dy * (y > 0) over channel slices: reads dy (bf16, row stride dys), y (bf16)
__global__ void k_relu_grad(const uint16_t *dy, const uint16_t *y, uint16_t *dz, size_t n)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    dz[i] = bf2f(y[i]) > 0.f ? dy[i] : (uint16_t)0;
}

// 2x2 average pool (H, W even), NHWC
__global__ void k_pool(const uint16_t *x, uint16_t *y, int nb, int h, int w, int c)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const int ho = h / 2, wo = w / 2;
    if (i >= (size_t)nb * ho * wo * c) return;
    const int ch = i % c;
    const size_t q = i / c;
    const int xo = q % wo, yo = (q / wo) % ho, n = q / (wo * ho);
    float s = 0.f;
#pragma unroll
    for (int dy = 0; dy < 2; dy++)
#pragma unroll
        for (int dx = 0; dx < 2; dx++)
            s += bf2f(x[((size_t)(n * h + 2 * yo + dy) * w + 2 * xo + dx) * c + ch]);
    y[i] = f2bf(0.25f * s);
}

// gradient of the pool, accumulated into dx (f32 accumulators are not needed: one writer per
// element), optionally added to what dx already holds
__global__ void k_pool_grad(const uint16_t *dy, uint16_t *dx, int nb, int h, int w, int c, int accumulate)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)nb * h * w * c) return;
    const int ch = i % c;
    const size_t q = i / c;
    const int x = q % w, y = (q / w) % h, n = q / (w * h);
    float g = 0.25f * bf2f(dy[((size_t)(n * (h / 2) + y / 2) * (w / 2) + x / 2) * c + ch]);
    if (accumulate) g += bf2f(dx[i]);
    dx[i] = f2bf(g);
}

// nearest 2x upsample of x (nb, h, w, c) into channels [off, off + c) of y (row stride ys)
__global__ void k_upsample(const uint16_t *x, uint16_t *y, int nb, int h, int w, int c, int ys, int off)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const int H = 2 * h, W = 2 * w;
    if (i >= (size_t)nb * H * W * c) return;
    const int ch = i % c;
    const size_t q = i / c;
    const int X = q % W, Y = (q / W) % H, n = q / (W * H);
    y[q * ys + off + ch] = x[((size_t)(n * h + Y / 2) * w + X / 2) * c + ch];
}

// k_upsample with 8 channels (16 bytes) per thread: c, ys and off multiples of 8, both tensors
// 16-byte aligned (c8 = c / 8)
__global__ void k_upsample8(const uint16_t *x, uint16_t *y, int nb, int h, int w, int c8, int ys, int off)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const int H = 2 * h, W = 2 * w;
    if (i >= (size_t)nb * H * W * c8) return;
    const int ch = 8 * (int)(i % c8);
    const size_t q = i / c8;
    const int X = q % W, Y = (q / W) % H, n = q / (W * H);
    const uint4 v = *reinterpret_cast<const uint4 *>(x + ((size_t)(n * h + Y / 2) * w + X / 2) * (8 * c8) + ch);
    *reinterpret_cast<uint4 *>(y + q * ys + off + ch) = v;
}

// gradient of the upsample: dx[y][x] = sum of the four dy children (channel slice of dy)
__global__ void k_upsample_grad(const uint16_t *dy, uint16_t *dx, int nb, int h, int w, int c, int ys, int off)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)nb * h * w * c) return;
    const int ch = i % c;
    const size_t q = i / c;
    const int x = q % w, y = (q / w) % h, n = q / (w * h);
    float s = 0.f;
#pragma unroll
    for (int dy2 = 0; dy2 < 2; dy2++)
#pragma unroll
        for (int dx2 = 0; dx2 < 2; dx2++)
            s += bf2f(dy[((size_t)(n * 2 * h + 2 * y + dy2) * (2 * w) + 2 * x + dx2) * ys + off + ch]);
    dx[i] = f2bf(s);
}

// copy a tensor (n pixels, c channels) into / out of a channel slice [off, off+c) of a wider
// tensor (row stride ys); `accumulate` adds to the destination
__global__ void k_slice(const uint16_t *src, int ss, int soff, uint16_t *dst, int ds, int doff, size_t npix, int c,
                        int accumulate)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= npix * c) return;
    const size_t p = i / c;
    const int ch = i % c;
    const uint16_t v = src[p * ss + soff + ch];
    uint16_t *d = dst + p * ds + doff + ch;
    *d = accumulate ? f2bf(bf2f(*d) + bf2f(v)) : v;
}

// k_slice without accumulation, 8 channels (16 bytes) per thread: c, strides and offsets
// multiples of 8, both tensors 16-byte aligned (c8 = c / 8)
__global__ void k_slice8(const uint16_t *src, int ss, int soff, uint16_t *dst, int ds, int doff, size_t npix, int c8)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= npix * c8) return;
    const size_t p = i / c8;
    const int ch = 8 * (int)(i % c8);
    const uint4 v = *reinterpret_cast<const uint4 *>(src + p * ss + soff + ch);
    *reinterpret_cast<uint4 *>(dst + p * ds + doff + ch) = v;
}

__global__ void k_f32_to_bf16(const float *x, uint16_t *y, size_t n)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// weights: master wm[co][k] (f32, [coutp][kp], zero where co >= cout or k >= 9*cin) ->
// wt[co][k] bf16 and the data-gradient weights wd[ci][kd] (bf16, [cin][kdp]) with
// kd = tap' * coutp + co, tap' = 8 - tap (the flipped, transposed kernel)
__global__ void k_prepare_weights(const float *wm, uint16_t *wt, uint16_t *wd, int cin, int cout, int coutp, int kp,
                                  int kdp)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= coutp * kp) return;
    const int co = i / kp, k = i % kp;
    const float v = (co < cout && k < 9 * cin) ? wm[i] : 0.f;
    wt[i] = f2bf(v);
    if (k < 9 * cin) {
        const int tap = k / cin, ci = k % cin;
        wd[(size_t)ci * kdp + (8 - tap) * coutp + co] = f2bf(v);
    }
}

// L1 loss on three output channels: out (f32 [p][3]), target (f32 [p][3]), mask [p];
// loss partials (sum |d|, count) per block; gradient sign(d) * scale (bf16, row stride 16)
__global__ void k_l1(const float *out, const float *target, const float *mask, uint16_t *grad, float *partial,
                     int P, float scale)
{
    __shared__ float sl[256], sc[256];
    const int p = blockIdx.x * 256 + threadIdx.x;
    float l = 0.f, cnt = 0.f;
    if (p < P) {
        const float m = mask[p];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float d = out[(size_t)p * 3 + c] - target[(size_t)p * 3 + c];
            l += m * fabsf(d);
            grad[(size_t)p * 16 + c] = f2bf(m * scale * (d > 0.f ? 1.f : (d < 0.f ? -1.f : 0.f)));
        }
#pragma unroll
        for (int c = 3; c < 16; c++) grad[(size_t)p * 16 + c] = 0;
        cnt = m;
    }
    sl[threadIdx.x] = l;
    sc[threadIdx.x] = cnt;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sl[threadIdx.x] += sl[threadIdx.x + s];
            sc[threadIdx.x] += sc[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = sl[0];
        partial[2 * blockIdx.x + 1] = sc[0];
    }
}

// Adam (Kingma & Ba 2015) on f32 parameters
__global__ void k_adam(float *w, const float *g, float *m, float *v, size_t n, float lr, float b1, float b2,
                       float eps, float bc1, float bc2, float gscale)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] * gscale;
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// dst row r = src row idx[r] (rows of row4 float4s)
__global__ void k_gather_rows(const float4 *src, const unsigned *idx, float4 *dst, unsigned count, size_t row4)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= count * row4) return;
    const size_t r = i / row4, c = i % row4;
    dst[i] = src[(size_t)idx[r] * row4 + c];
}

}  // namespace

inline dim3 g1(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

hipError_t probe(const uint16_t *A, const uint16_t *B, float *C, hipStream_t s)
{
    hipLaunchKernelGGL(k_probe, dim3(1), dim3(64), 0, s, A, B, C);
    return hipGetLastError();
}

hipError_t conv3x3(const ConvArgs &a, hipStream_t s)
{
    const int P = a.nb * a.h * a.w;
    if (a.cin % 4 || a.kp % 16) return hipErrorInvalidValue;
    const unsigned gx = (unsigned)((P + CS_P - 1) / CS_P);
    if (a.coutp >= 64)
        hipLaunchKernelGGL(k_conv3x3_lds<64>, dim3(gx, (a.coutp + 63) / 64), dim3(256), 0, s, a);
    else if (a.coutp >= 32)
        hipLaunchKernelGGL(k_conv3x3_lds<32>, dim3(gx, (a.coutp + 31) / 32), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(k_conv3x3_lds<16>, dim3(gx, (a.coutp + 15) / 16), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t conv3x3_direct(const ConvArgs &a, hipStream_t s)
{
    const int P = a.nb * a.h * a.w;
    hipLaunchKernelGGL(k_conv3x3, dim3((P + CT_P - 1) / CT_P, (a.coutp + CT_C - 1) / CT_C), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t conv3x3_wgrad(const WGradArgs &a, hipStream_t s)
{
    const int P = a.nb * a.h * a.w;
    hipLaunchKernelGGL(k_conv3x3_wgrad, dim3((a.kp + WG_K - 1) / WG_K, (a.coutp + WG_C - 1) / WG_C, (P + WG_CHUNK - 1) / WG_CHUNK),
                       dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t bias_grad(const uint16_t *dz, float *db, int P, int coutp, int cout, hipStream_t s)
{
    hipLaunchKernelGGL(k_bias_grad, dim3((P + 1023) / 1024, (cout + 63) / 64), dim3(256), 0, s, dz, db, P, coutp, cout);
    return hipGetLastError();
}

hipError_t relu_grad(const uint16_t *dy, const uint16_t *y, uint16_t *dz, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_relu_grad, g1(n), dim3(256), 0, s, dy, y, dz, n);
    return hipGetLastError();
}

hipError_t pool(const uint16_t *x, uint16_t *y, int nb, int h, int w, int c, hipStream_t s)
{
    hipLaunchKernelGGL(k_pool, g1((size_t)nb * (h / 2) * (w / 2) * c), dim3(256), 0, s, x, y, nb, h, w, c);
    return hipGetLastError();
}

hipError_t pool_grad(const uint16_t *dy, uint16_t *dx, int nb, int h, int w, int c, bool acc, hipStream_t s)
{
    hipLaunchKernelGGL(k_pool_grad, g1((size_t)nb * h * w * c), dim3(256), 0, s, dy, dx, nb, h, w, c, acc ? 1 : 0);
    return hipGetLastError();
}

// 16-byte accesses are possible when every channel run starts on a 16-byte boundary
static bool by8(const void *a, const void *b, int c, int s1, int o1, int s2, int o2)
{
    return c % 8 == 0 && s1 % 8 == 0 && o1 % 8 == 0 && s2 % 8 == 0 && o2 % 8 == 0 &&
           reinterpret_cast<uintptr_t>(a) % 16 == 0 && reinterpret_cast<uintptr_t>(b) % 16 == 0;
}

hipError_t upsample(const uint16_t *x, uint16_t *y, int nb, int h, int w, int c, int ys, int off, hipStream_t s)
{
    if (by8(x, y, c, c, 0, ys, off))
        hipLaunchKernelGGL(k_upsample8, g1((size_t)nb * 4 * h * w * (c / 8)), dim3(256), 0, s, x, y, nb, h, w, c / 8,
                           ys, off);
    else
        hipLaunchKernelGGL(k_upsample, g1((size_t)nb * 4 * h * w * c), dim3(256), 0, s, x, y, nb, h, w, c, ys, off);
    return hipGetLastError();
}

hipError_t upsample_grad(const uint16_t *dy, uint16_t *dx, int nb, int h, int w, int c, int ys, int off, hipStream_t s)
{
    hipLaunchKernelGGL(k_upsample_grad, g1((size_t)nb * h * w * c), dim3(256), 0, s, dy, dx, nb, h, w, c, ys, off);
    return hipGetLastError();
}

hipError_t slice(const uint16_t *src, int ss, int soff, uint16_t *dst, int ds, int doff, size_t npix, int c, bool acc,
                 hipStream_t s)
{
    if (!acc && by8(src, dst, c, ss, soff, ds, doff))
        hipLaunchKernelGGL(k_slice8, g1(npix * (c / 8)), dim3(256), 0, s, src, ss, soff, dst, ds, doff, npix, c / 8);
    else
        hipLaunchKernelGGL(k_slice, g1(npix * c), dim3(256), 0, s, src, ss, soff, dst, ds, doff, npix, c, acc ? 1 : 0);
    return hipGetLastError();
}

hipError_t f32_to_bf16(const float *x, uint16_t *y, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_f32_to_bf16, g1(n), dim3(256), 0, s, x, y, n);
    return hipGetLastError();
}

hipError_t prepare_weights(const float *wm, uint16_t *wt, uint16_t *wd, int cin, int cout, int coutp, int kp, int kdp,
                           hipStream_t s)
{
    hipLaunchKernelGGL(k_prepare_weights, g1((size_t)coutp * kp), dim3(256), 0, s, wm, wt, wd, cin, cout, coutp, kp, kdp);
    return hipGetLastError();
}

hipError_t l1_loss(const float *out, const float *target, const float *mask, uint16_t *grad, float *partial, int P,
                   float scale, hipStream_t s)
{
    hipLaunchKernelGGL(k_l1, g1(P), dim3(256), 0, s, out, target, mask, grad, partial, P, scale);
    return hipGetLastError();
}

hipError_t gather_rows(const float *src, const unsigned *idx, float *dst, unsigned count, size_t row_floats,
                       hipStream_t s)
{
    const size_t row4 = row_floats / 4;
    hipLaunchKernelGGL(k_gather_rows, g1(count * row4), dim3(256), 0, s, (const float4 *)src, idx, (float4 *)dst, count,
                       row4);
    return hipGetLastError();
}

hipError_t adam(float *w, const float *g, float *m, float *v, size_t n, float lr, float b1, float b2, float eps,
                float bc1, float bc2, float gscale, hipStream_t s)
{
    hipLaunchKernelGGL(k_adam, g1(n), dim3(256), 0, s, w, g, m, v, n, lr, b1, b2, eps, bc1, bc2, gscale);
    return hipGetLastError();
}

}  // namespace cnn
}  // namespace pdf
