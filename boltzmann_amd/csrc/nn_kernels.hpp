// Kernels of the warm-start network: a bf16 MFMA GEMM with LDS staging that also reads 3x3
// convolution operands implicitly, and the element-wise layers.  Activations are
// channels-last bf16 ([batch][nx][ny][C], C contiguous); gradients and master weights are
// float.
//
// The GEMM computes C(m, n) = sum_k A(m, k) B(n, k) with
//   A in "MK" form (A(m,k) at a[m*lda + k]), "KM" form (a[k*lda + m]) or "MKconv" form
//     (the 3x3 neighbourhood of pixel m, k = tap*C + c, gathered from a channels-last image),
//   B in "NK" form (B(n,k) at b[n*ldb + k]), "KN" form (b[k*ldb + n]) or "KNconv" form
//     (n = tap*C + c, k = pixel),
// which covers the forward convolution (MKconv(X) x W^T), the input gradient (MKconv(dZ) x
// Wt^T with the reflected, transposed kernel Wt) and the weight gradient (dZ^T x KNconv(X)),
// with no im2col buffer.  Work-group tile 64 x 64, K step 64, four waves each owning a 32 x 32
// block computed with v_mfma_f32_32x32x16_bf16, float accumulation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nnk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint16_t u16;

// bfloat16 bits of v, rounded to nearest even (NaN kept quiet): gfx950's conversion
// instruction, the same bits as the integer form u + 0x7fff + ((u >> 16) & 1) >> 16 for every
// float
__device__ inline u16 bf16_bits(float v) { return __builtin_bit_cast(u16, (__bf16)v); }
__device__ inline float bf16_float(u16 b) { return __uint_as_float(((uint32_t)b) << 16); }

enum : int { kMK = 0, kKM = 1, kMKconv = 2 };   // A forms
enum : int { kNK = 0, kKN = 1, kKNconv = 2 };   // B forms
enum : int { kStoreF32 = 0, kStoreBF16 = 1, kAtomicF32 = 2 };

// Unsigned 32-bit division by a divisor fixed for a launch, as a multiply-high (Granlund and
// Montgomery, "Division by invariant integers using multiplication", PLDI 1994): with
// s = ceil(log2 d) and m = floor(2^32 (2^s - d) / d) + 1, n / d = (mulhi(n, m) + n) >> s for
// every 32-bit n (the sum is formed in 64 bits).
struct Divisor {
    uint32_t m;
    int s;
};

inline Divisor make_divisor(uint32_t d) {
    int s = 0;
    while ((1ull << s) < d) ++s;
    return {(uint32_t)((((1ull << 32) * ((1ull << s) - d)) / d) + 1), s};
}

__host__ __device__ inline uint32_t divide(uint32_t n, Divisor v) {
    const uint64_t hi = ((uint64_t)n * v.m) >> 32;  // mulhi: v_mul_hi_u32 on the device
    return (uint32_t)((hi + n) >> v.s);
}

// Geometry of an implicit 3x3 (padding 1) convolution operand: the element for pixel p and
// column j = tap * C + c is X[b, x + tap/3 - 1, y + tap%3 - 1, c] (zero off the image), with
// p = (b*nx + x)*ny + y and X channels-last with channel stride ld (the operand's lda/ldb).
// Pixel counts stay below 2^31.
struct ConvGeom {
    int nx, ny, C;
    Divisor dx, dy, dc;  // division by nx, ny and C
};

inline ConvGeom conv_geom(int nx, int ny, int C) {
    return {nx, ny, C, make_divisor((uint32_t)nx), make_divisor((uint32_t)ny), make_divisor((uint32_t)C)};
}

struct GemmArgs {
    const u16* a;
    const u16* b;
    void* c;
    const float* bias;   // per column n, or null
    int M, N, K;
    int lda, ldb, ldc;
    int c_off;           // column offset into c (channel slice of a concatenation)
    int relu;            // apply max(0, .) in the epilogue
    int k_split;         // K range per z-slice (multiple of 32)
    ConvGeom conv;       // for kMKconv / kKNconv operands
};

// address of the 8 channels [c, c+8) of pixel p's neighbour through tap, or null off the image
__device__ inline const u16* conv_src(const u16* X, int ld, const ConvGeom& g, uint32_t p, uint32_t j) {
    const uint32_t tap = divide(j, g.dc), c = j - tap * (uint32_t)g.C;
    const uint32_t bx = divide(p, g.dy), b = divide(bx, g.dx);
    const int y = (int)(p - bx * (uint32_t)g.ny), x = (int)(bx - b * (uint32_t)g.nx);
    const int xx = x + (int)(tap / 3) - 1, yy = y + (int)(tap % 3) - 1;
    if (xx < 0 || xx >= g.nx || yy < 0 || yy >= g.ny) return nullptr;
    return X + (size_t)((b * (uint32_t)g.nx + (uint32_t)xx) * (uint32_t)g.ny + (uint32_t)yy) * ld + c;
}

// element e (0..7) of 8 packed bf16 values, by shifts: a pointer cast of the vector would
// place it in scratch memory
__device__ inline uint32_t half16(const uint4& q, int e) {
    const uint32_t w = e < 2 ? q.x : e < 4 ? q.y : e < 6 ? q.z : q.w;
    return (e & 1) ? (w >> 16) : (w & 0xffffu);
}

// kTK: the K width one tile load fills; kStage: the K width staged in LDS per barrier pair
constexpr int kTM = 64, kTN = 64, kTK = 32, kStage = 64, kPad = 8, kLdsRow = kStage + kPad;

// Loads a 64 x 32 tile of an operand into LDS as [row][k] (k contiguous), zero-filled
// outside the matrix.  kform: 0 -> element (row, k) at p[row*ld + k]; 1 -> at p[k*ld + row].
// FORM 2 (kMKconv): row = pixel, k = tap*C + c gathered from the image (rows x K = P x 9C).
// FORM 3 (kKNconv): row = tap*C + c, k = pixel (rows x K = 9C x P).  C % 8 == 0, so every
// 8-element chunk lies in one tap; K and row counts of conv operands are multiples of 8.
template <int FORM>
__device__ inline void load_tile(u16* lds, const u16* __restrict__ p, int ld, int row0, int rows,
                                 int k0, int kend, const ConvGeom& cg) {
    const int t = threadIdx.x;  // 256 threads, 8 elements each
    if (FORM == 2) {
        const int r = t >> 2, kc = (t & 3) * 8;
        const int gr = row0 + r, gk = k0 + kc;
        uint4 q = {0, 0, 0, 0};
        if (gr < rows && gk < kend) {
            const u16* src = conv_src(p, ld, cg, gr, gk);
            if (src) q = *(const uint4*)src;
        }
        *(uint4*)(lds + r * kLdsRow + kc) = q;
        return;
    }
    if (FORM == 3) {
        const int k = t >> 3, rc = (t & 7) * 8;
        const int gk = k0 + k, gr = row0 + rc;
        uint4 q = {0, 0, 0, 0};
        if (gk < kend && gr < rows) {
            const u16* src = conv_src(p, ld, cg, gk, gr);
            if (src) q = *(const uint4*)src;
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) lds[(rc + e) * kLdsRow + k] = (u16)half16(q, e);
        return;
    }
    if (FORM == 0) {
        const int r = t >> 2, kc = (t & 3) * 8;
        const int gr = row0 + r, gk = k0 + kc;
        u16 v[8];
        if (gr < rows && gk + 8 <= kend && ((ld & 7) == 0)) {
            const uint4 q = *(const uint4*)(p + (size_t)gr * ld + gk);
            *(uint4*)v = q;
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e)
                v[e] = (gr < rows && gk + e < kend) ? p[(size_t)gr * ld + gk + e] : (u16)0;
        }
        *(uint4*)(lds + r * kLdsRow + kc) = *(uint4*)v;
    } else {
        const int k = t >> 3, rc = (t & 7) * 8;
        const int gk = k0 + k, gr = row0 + rc;
        u16 v[8];
        if (gk < kend && gr + 8 <= rows && ((ld & 7) == 0)) {
            const uint4 q = *(const uint4*)(p + (size_t)gk * ld + gr);
            *(uint4*)v = q;
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e)
                v[e] = (gk < kend && gr + e < rows) ? p[(size_t)gk * ld + gr + e] : (u16)0;
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) lds[(rc + e) * kLdsRow + k] = v[e];
    }
}

// 8 consecutive rows [r, r+8) at one k of a k-major operand (FORM 1: p[k*ld + row];
// FORM 3: the gathered image, k = pixel, row = tap*C + c), zero outside the matrix
template <int FORM>
__device__ inline uint4 load8_kmajor(const u16* __restrict__ p, int ld, int r, int rows, int k, int kend,
                                     const ConvGeom& cg) {
    uint4 q = {0, 0, 0, 0};
    if (k >= kend || r >= rows) return q;
    if (FORM == 3) {
        const u16* src = conv_src(p, ld, cg, k, r);
        if (src) q = *(const uint4*)src;
        return q;
    }
    if (r + 8 <= rows && (ld & 7) == 0) return *(const uint4*)(p + (size_t)k * ld + r);
    uint32_t w[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t lo = (r + 2 * i < rows) ? p[(size_t)k * ld + r + 2 * i] : 0u;
        const uint32_t hi = (r + 2 * i + 1 < rows) ? p[(size_t)k * ld + r + 2 * i + 1] : 0u;
        w[i] = lo | (hi << 16);
    }
    return {w[0], w[1], w[2], w[3]};
}

// Fills one 64 x 32 LDS tile ([row][k]) of a k-major operand with 128 threads: thread t
// loads k = 2*(t>>3) and k+1 for rows 8*(t&7) .. +7 and writes the 8 (k, k+1) pairs as
// 4-byte stores (half the LDS store instructions of one k per thread)
template <int FORM>
__device__ inline void load_tile_kpairs(u16* lds, const u16* __restrict__ p, int ld, int row0, int rows,
                                        int k0, int kend, const ConvGeom& cg, int t) {
    const int kp = 2 * (t >> 3), rc = (t & 7) * 8;
    const uint4 q0 = load8_kmajor<FORM>(p, ld, row0 + rc, rows, k0 + kp, kend, cg);
    const uint4 q1 = load8_kmajor<FORM>(p, ld, row0 + rc, rows, k0 + kp + 1, kend, cg);
#pragma unroll
    for (int e = 0; e < 8; ++e)
        *(uint32_t*)(lds + (rc + e) * kLdsRow + kp) = half16(q0, e) | (half16(q1, e) << 16);
}

template <int AF, int BF, int OUT>
__global__ void __launch_bounds__(256) gemm_bf16(const GemmArgs g) {
    __shared__ __attribute__((aligned(16))) u16 sa[kTM * kLdsRow];
    __shared__ __attribute__((aligned(16))) u16 sb[kTN * kLdsRow];
    const int m0 = blockIdx.x * kTM, n0 = blockIdx.y * kTN;
    const int kb = blockIdx.z * g.k_split;
    const int ke = min(g.K, kb + g.k_split);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x16 acc = {0};
    for (int k0 = kb; k0 < ke; k0 += kStage) {
        bool pairs = false;
        if constexpr (AF == kKM && (BF == kKN || BF == kKNconv)) pairs = g.M >= 16;
#pragma unroll
        for (int h = 0; h < kStage; h += kTK) {
            // columns past ke (the slice end) load as zeros
            if (pairs) {
                // both operands k-major (the weight gradients of the 3x3 layers): half the
                // threads fill each tile; a skinny A (the 1x1 output layer, M = 3) keeps one k
                // per thread
                if (threadIdx.x < 128)
                    load_tile_kpairs<1>(sa + h, g.a, g.lda, m0, g.M, k0 + h, ke, g.conv, threadIdx.x);
                else
                    load_tile_kpairs<BF == kKNconv ? 3 : 1>(sb + h, g.b, g.ldb, n0, g.N, k0 + h, ke,
                                                            g.conv, threadIdx.x - 128);
            } else {
                load_tile<AF>(sa + h, g.a, g.lda, m0, g.M, k0 + h, ke, g.conv);
                load_tile<BF == kKNconv ? 3 : BF>(sb + h, g.b, g.ldb, n0, g.N, k0 + h, ke, g.conv);
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kStage; kk += 16) {
            // A(i, k): lane holds row i = lane % 32, k = 8 * (lane / 32) .. + 7
            const bf16x8 a = *(const bf16x8*)(sa + (wm + (lane & 31)) * kLdsRow + kk + 8 * (lane >> 5));
            const bf16x8 b = *(const bf16x8*)(sb + (wn + (lane & 31)) * kLdsRow + kk + 8 * (lane >> 5));
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // D(i, j): lane holds column j = lane % 32, rows (r/4)*8 + (lane/32)*4 + r%4
    const int j = n0 + wn + (lane & 31);
    if (j >= g.N) return;
    const float bj = g.bias ? g.bias[j] : 0.f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int i = m0 + wm + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
        if (i >= g.M) continue;
        float v = acc[r] + bj;
        if (g.relu) v = fmaxf(v, 0.f);
        const size_t o = (size_t)i * g.ldc + g.c_off + j;
        if (OUT == kStoreF32) ((float*)g.c)[o] = v;
        else if (OUT == kStoreBF16) ((u16*)g.c)[o] = bf16_bits(v);
        else atomicAdd((float*)g.c + o, v);
    }
}

// Wt[ci][t * cout + co] = W[co][(8 - t) * cin + ci]: the kernel of the input gradient of a
// 3x3 convolution (its taps reflected and its channels transposed), bf16
__global__ void flip_weights(const u16* __restrict__ W, int cout, int cin, u16* __restrict__ Wt) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)cout * 9 * cin) return;
    const int ci = (int)(t % cin);
    const int tap = (int)((t / cin) % 9);
    const int co = (int)(t / (9 * cin));
    Wt[((size_t)ci * 9 + (8 - tap)) * cout + co] = W[t];
}

// dZ = dY * (Y > 0) as bf16, dY float with channel stride ld / offset off
// Backward of a layer's bias + ReLU, in one pass over its P x C pre-activation gradient:
//   dZ[p][c] = bf16(dY[p][c] masked by Y[p][c] > 0 when relu)  and  dbias[c] += sum_p dZ[p][c].
// C <= 256 and 256 threads: with R = 256 / C rows per pass, thread t handles channel t % C of
// rows p0 + t / C, + R, + 2R, ... of its work-group's chunk [p0, p0 + rows), then the group
// adds its channel sums with one atomic per channel.
__global__ void relu_bias_grad(const float* __restrict__ dY, int ldy, int yoff, const u16* __restrict__ Y,
                               int ldY, int Yoff, u16* __restrict__ dZ, long long P, int C, int relu,
                               long long rows, float* __restrict__ db) {
    const int R = (int)blockDim.x / C;
    const int c = (int)threadIdx.x % C, r0 = (int)threadIdx.x / C;
    const long long p0 = (long long)blockIdx.x * rows;
    const long long p1 = p0 + rows < P ? p0 + rows : P;
    float s = 0.f;
    if (r0 < R)
        for (long long p = p0 + r0; p < p1; p += R) {
            float g = dY[p * ldy + yoff + c];
            if (relu && !(bf16_float(Y[p * ldY + Yoff + c]) > 0.f)) g = 0.f;
            const u16 b = bf16_bits(g);
            dZ[p * C + c] = b;
            s += bf16_float(b);
        }
    __shared__ float sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    if ((int)threadIdx.x < C) {
        float t = 0.f;
        for (int j = threadIdx.x; j < R * C; j += C) t += sh[j];
        atomicAdd(&db[threadIdx.x], t);
    }
}

// The pooling and upsampling kernels walk a channels-last image 8 channels per thread:
// thread t takes pixel p = t / (C/8) and channels [c0, c0 + 8), c0 = 8 (t % (C/8)), with
// p = (b*nx + x)*ny + y.  C, the leading dimensions and the channel offsets are multiples
// of 8, so each thread moves 16 bytes of bf16 or 32 of float per pixel; the thread count
// (pixels x C/8 <= 2^26 x 32) fits in 32 bits.
struct Grid8 {
    int nx, ny, C;
    Divisor dx, dy, dq;  // division by nx, ny and C / 8
};

inline Grid8 grid8(int nx, int ny, int C) {
    return {nx, ny, C, make_divisor((uint32_t)nx), make_divisor((uint32_t)ny), make_divisor((uint32_t)(C / 8))};
}

struct Pix8 {
    uint32_t p;
    int b, x, y, c0;
};

__device__ inline Pix8 pix8(const Grid8& g, uint32_t t) {
    Pix8 r;
    r.p = divide(t, g.dq);
    r.c0 = (int)(t - r.p * (uint32_t)(g.C / 8)) * 8;
    const uint32_t bx = divide(r.p, g.dy), b = divide(bx, g.dx);
    r.y = (int)(r.p - bx * (uint32_t)g.ny);
    r.x = (int)(bx - b * (uint32_t)g.nx);
    r.b = (int)b;
    return r;
}

__device__ inline void load8f(const float* p, float (&v)[8]) {
    const float4 a = *(const float4*)p, b = *(const float4*)(p + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}

__device__ inline void store8f(float* p, const float (&v)[8]) {
    *(float4*)p = make_float4(v[0], v[1], v[2], v[3]);
    *(float4*)(p + 4) = make_float4(v[4], v[5], v[6], v[7]);
}

// 2x2 average pooling (channels-last) into Y [b][nx/2][ny/2][C]; ld/off select a channel
// slice of X.  g describes the pooled image (nx/2, ny/2, C).
__global__ void avgpool2(const u16* __restrict__ X, int ldx, int xoff, u16* __restrict__ Y, uint32_t n,
                         Grid8 g) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const Pix8 q = pix8(g, t);
    const int nx = 2 * g.nx, ny = 2 * g.ny;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const size_t fp = ((size_t)q.b * nx + 2 * q.x + i) * ny + 2 * q.y + j;
            const uint4 v = *(const uint4*)(X + fp * ldx + xoff + q.c0);
#pragma unroll
            for (int e = 0; e < 8; ++e) s[e] += bf16_float((u16)half16(v, e));
        }
    uint32_t w[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        w[k] = (uint32_t)bf16_bits(0.25f * s[2 * k]) | ((uint32_t)bf16_bits(0.25f * s[2 * k + 1]) << 16);
    *(uint4*)(Y + (size_t)q.p * g.C + q.c0) = {w[0], w[1], w[2], w[3]};
}

// backward of avgpool2: dX[b, x, y, c] (+)= dY[b, x/2, y/2, c] / 4.  g describes the
// full-resolution image (nx, ny, C).
__global__ void avgpool2_grad(const float* __restrict__ dY, float* __restrict__ dX, int lddx, int dxoff,
                              uint32_t n, Grid8 g) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const Pix8 q = pix8(g, t);
    const size_t cp = ((size_t)q.b * (g.nx / 2) + q.x / 2) * (g.ny / 2) + q.y / 2;
    float v[8], d[8];
    load8f(dY + cp * g.C + q.c0, v);
    float* out = dX + (size_t)q.p * lddx + dxoff + q.c0;
    load8f(out, d);
#pragma unroll
    for (int e = 0; e < 8; ++e) d[e] += 0.25f * v[e];
    store8f(out, d);
}

// nearest 2x upsampling into a channel slice: Y[b, x, y, off + c] = X[b, x/2, y/2, c].  g
// describes the upsampled image (nx, ny, C).
__global__ void upsample2(const u16* __restrict__ X, u16* __restrict__ Y, int ldy, int yoff, uint32_t n,
                          Grid8 g) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const Pix8 q = pix8(g, t);
    const size_t cp = ((size_t)q.b * (g.nx / 2) + q.x / 2) * (g.ny / 2) + q.y / 2;
    *(uint4*)(Y + (size_t)q.p * ldy + yoff + q.c0) = *(const uint4*)(X + cp * g.C + q.c0);
}

// backward of upsample2: dX[b, x, y, c] = sum of the four dY[.., off + c].  g describes the
// coarse image (nx/2, ny/2, C).
__global__ void upsample2_grad(const float* __restrict__ dY, int ldy, int yoff, float* __restrict__ dX,
                               uint32_t n, Grid8 g) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const Pix8 q = pix8(g, t);
    const int nx = 2 * g.nx, ny = 2 * g.ny;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            float v[8];
            load8f(dY + (((size_t)q.b * nx + 2 * q.x + i) * ny + 2 * q.y + j) * ldy + yoff + q.c0, v);
#pragma unroll
            for (int e = 0; e < 8; ++e) s[e] += v[e];
        }
    store8f(dX + (size_t)q.p * g.C + q.c0, s);
}

// mean-squared error over all elements; dP = 2 (P - T) / n; partial sums per block
__global__ void mse_loss(const float* __restrict__ P, const float* __restrict__ T, float* __restrict__ dP,
                         long long n, double* __restrict__ partial) {
    double s = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const float d = P[i] - T[i];
        s += (double)d * d;
        dP[i] = 2.0f * d / (float)n;
    }
    __shared__ double sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// Adam with bias correction; refreshes the bf16 copy of the weights
__global__ void adam_step(float* __restrict__ w, u16* __restrict__ wb, const float* __restrict__ g,
                          float* __restrict__ m, float* __restrict__ v, long long n, float lr,
                          float b1, float b2, float eps, float c1, float c2) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = bf16_bits(wi);
}

// Input features [pixels][FEAT] float -> [pixels][INC] bf16 with the channels from FEAT and
// the pixels from `valid` on zero
// features [pixels][feat] float -> [pixels][INC] bf16, channels feat.. and pixels valid..
// zeroed (feat <= INC)
template <int INC>
__global__ void pad_features(const float* __restrict__ x, long long pixels, long long valid,
                             int feat, u16* __restrict__ out) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= pixels * INC) return;
    const long long p = t / INC;
    const int c = (int)(t - p * INC);
    out[t] = (c < feat && p < valid) ? bf16_bits(x[p * feat + c]) : (u16)0;
}

__global__ void to_bf16(const float* __restrict__ x, u16* __restrict__ y, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = bf16_bits(x[i]);
}

__global__ void fill_f32(float* __restrict__ x, long long n, float v) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

}  // namespace nnk
