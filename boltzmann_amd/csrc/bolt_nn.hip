// The warm-start network: a small U-Net trained and run with this library's own kernels
// (nn_kernels.hpp: bf16 MFMA GEMM, im2col/col2im, element-wise layers).
//
// Input  [batch][nx][ny][16]  bf16 (the caller's 4 feature channels, zero-padded to 16)
// Output [batch][nx][ny][3]   float
//
//   level 0 (nx, ny):     e1a = conv(x, 16->w)          e1b = conv(e1a, w->w)
//   level 1 (nx/2, ny/2): e2a = conv(pool e1b, w->2w)   e2b = conv(e2a, 2w->2w)
//   level 2:              e3a = conv(pool e2b, 2w->4w)  e3b = conv(e3a, 4w->4w)
//   level 3:              ba  = conv(pool e3b, 4w->4w)  bb  = conv(ba, 4w->4w)
//   level 2:  u3 = conv([up bb | e3b], 8w->2w)
//   level 1:  u2 = conv([up u3 | e2b], 4w->w)
//   level 0:  u1 = conv([up u2 | e1b], 2w->w)          out = conv1x1(u1, w->3)
// 3x3 convolutions with padding 1 and ReLU, 2x2 average pooling, nearest upsampling.
// Parameters are float master weights (layer order above, each W[cout][taps*cin] then
// b[cout], tap = kx*3 + ky), with a bf16 copy for the GEMMs; Adam updates both.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <numeric>
#include <random>
#include <vector>

#include "nn_kernels.hpp"
#include "runtime_internal.hpp"

using namespace nnk;

namespace {

constexpr int kInC = 16;   // padded input channels
constexpr int kFeat = 4;   // caller's feature channels
constexpr int kOutC = 3;

inline unsigned nblk(long long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

struct Conv {
    int cin, cout, taps;   // taps: 9 (3x3) or 1 (1x1)
    bool relu;
    size_t w_off, b_off;   // offsets into the parameter vector
};

struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct bolt_nn {
    bolt_ctx* ctx = nullptr;
    int nx = 0, ny = 0, batch = 0, w = 32;
    std::vector<Conv> conv;
    size_t nparam = 0;
    float* W = nullptr;    // float master parameters
    u16* Wb = nullptr;     // bf16 copy
    float* G = nullptr;    // gradients
    float* M1 = nullptr;   // Adam moments
    float* M2 = nullptr;
    int adam_t = 0;
    std::vector<Buf> bufs;
    // activations (bf16)
    u16 *x0, *e1a, *cat1, *p1, *e2a, *cat2, *p2, *e3a, *cat3, *p3, *ba, *bb, *u3, *u2, *u1;
    float* out;
    // gradients (float)
    float *dout, *du1, *dcat1, *du2, *dcat2, *du3, *dcat3, *dbb, *dba, *dp3, *de3a, *dp2, *de2a,
        *dp1, *de1a;
    u16* dz;       // bf16 pre-activation gradient of the layer being processed
    u16* col;      // im2col scratch
    float* dcol;   // input-gradient columns
    float* target;
    double* partial;
    // dataset on the device
    u16* data_x = nullptr;
    float* data_y = nullptr;
    int data_n = 0;
};

static int nn_alloc(bolt_nn* N, void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(BOLT_ENOMEM, "network: hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    N->bufs.push_back({*p, bytes});
    return BOLT_OK;
}

// ---- launch helpers ---------------------------------------------------------------------
static void gemm(bolt_nn* N, int af, int bf, int out, const u16* a, int lda, const u16* b, int ldb,
                 void* c, int ldc, int c_off, int M, int Nn, int K, const float* bias, bool relu,
                 int k_split = 0) {
    GemmArgs g;
    g.a = a;
    g.b = b;
    g.c = c;
    g.bias = bias;
    g.M = M;
    g.N = Nn;
    g.K = K;
    g.lda = lda;
    g.ldb = ldb;
    g.ldc = ldc;
    g.c_off = c_off;
    g.relu = relu ? 1 : 0;
    g.k_split = k_split > 0 ? k_split : ((K + kTK - 1) / kTK) * kTK;
    const dim3 grid(nblk(M, kTM), nblk(Nn, kTN), (unsigned)((K + g.k_split - 1) / g.k_split));
    hipStream_t s = N->ctx->stream;
#define BOLT_GEMM(A, B, O) hipLaunchKernelGGL((gemm_bf16<A, B, O>), grid, dim3(256), 0, s, g)
#define BOLT_GEMM_OUT(A, B)                                  \
    do {                                                     \
        if (out == kStoreF32) BOLT_GEMM(A, B, kStoreF32);    \
        else if (out == kStoreBF16) BOLT_GEMM(A, B, kStoreBF16); \
        else BOLT_GEMM(A, B, kAtomicF32);                    \
    } while (0)
    if (af == kMK && bf == kNK) BOLT_GEMM_OUT(kMK, kNK);
    else if (af == kMK && bf == kKN) BOLT_GEMM_OUT(kMK, kKN);
    else if (af == kKM && bf == kNK) BOLT_GEMM_OUT(kKM, kNK);
    else BOLT_GEMM_OUT(kKM, kKN);
#undef BOLT_GEMM_OUT
#undef BOLT_GEMM
}

// forward 3x3 conv of X (channel count cin, all channels) into Y (ldy, channel offset yoff)
static void conv_fwd(bolt_nn* N, const Conv& L, const u16* X, int nx, int ny, u16* Y, int ldy, int yoff) {
    hipStream_t s = N->ctx->stream;
    const long long P = (long long)N->batch * nx * ny;
    if (L.taps == 9) {
        hipLaunchKernelGGL(im2col3x3, dim3(nblk(P * 9 * (L.cin / 8))), dim3(256), 0, s, X, L.cin, 0,
                           N->col, N->batch, nx, ny, L.cin);
        gemm(N, kMK, kNK, kStoreBF16, N->col, 9 * L.cin, N->Wb + L.w_off, 9 * L.cin, Y, ldy, yoff,
             (int)P, L.cout, 9 * L.cin, N->W + L.b_off, L.relu);
    }
}

// backward of a 3x3 conv whose output was written at (Y, ldy, yoff) with gradient (dY, lddy,
// dyoff); accumulates parameter gradients and writes the input gradient into dX (if any)
static void conv_bwd(bolt_nn* N, const Conv& L, const u16* X, int nx, int ny, const u16* Y, int ldy,
                     int yoff, const float* dY, int lddy, int dyoff, float* dX) {
    hipStream_t s = N->ctx->stream;
    const long long P = (long long)N->batch * nx * ny;
    hipLaunchKernelGGL(relu_grad, dim3(nblk(P * L.cout)), dim3(256), 0, s, dY, lddy, dyoff, Y, ldy, yoff,
                       N->dz, P, L.cout, L.relu ? 1 : 0);
    hipLaunchKernelGGL(bias_grad, dim3(L.cout), dim3(256), 0, s, N->dz, P, L.cout, N->G + L.b_off);
    hipLaunchKernelGGL(im2col3x3, dim3(nblk(P * 9 * (L.cin / 8))), dim3(256), 0, s, X, L.cin, 0, N->col,
                       N->batch, nx, ny, L.cin);
    // dW[co][j] += sum_p dz[p][co] col[p][j]
    gemm(N, kKM, kKN, kAtomicF32, N->dz, L.cout, N->col, 9 * L.cin, N->G + L.w_off, 9 * L.cin, 0,
         L.cout, 9 * L.cin, (int)P, nullptr, false, 4096);
    if (dX) {
        // dcol[p][j] = sum_co dz[p][co] W[co][j]; dX = col2im(dcol)
        gemm(N, kMK, kKN, kStoreF32, N->dz, L.cout, N->Wb + L.w_off, 9 * L.cin, N->dcol, 9 * L.cin, 0,
             (int)P, 9 * L.cin, L.cout, nullptr, false);
        hipLaunchKernelGGL(col2im3x3, dim3(nblk(P * L.cin)), dim3(256), 0, s, N->dcol, dX, L.cin, 0, 0,
                           N->batch, nx, ny, L.cin);
    }
}

static void pool(bolt_nn* N, const u16* X, int ldx, int xoff, u16* Y, int nx, int ny, int C) {
    hipLaunchKernelGGL(avgpool2, dim3(nblk((long long)N->batch * (nx / 2) * (ny / 2) * C)), dim3(256), 0,
                       N->ctx->stream, X, ldx, xoff, Y, N->batch, nx, ny, C);
}
static void pool_bwd(bolt_nn* N, const float* dY, float* dX, int lddx, int dxoff, int nx, int ny, int C) {
    hipLaunchKernelGGL(avgpool2_grad, dim3(nblk((long long)N->batch * nx * ny * C)), dim3(256), 0,
                       N->ctx->stream, dY, dX, lddx, dxoff, N->batch, nx, ny, C);
}
static void up(bolt_nn* N, const u16* X, u16* Y, int ldy, int yoff, int nx, int ny, int C) {
    hipLaunchKernelGGL(upsample2, dim3(nblk((long long)N->batch * nx * ny * C)), dim3(256), 0,
                       N->ctx->stream, X, Y, ldy, yoff, N->batch, nx, ny, C);
}
static void up_bwd(bolt_nn* N, const float* dY, int ldy, int yoff, float* dX, int nx, int ny, int C) {
    hipLaunchKernelGGL(upsample2_grad, dim3(nblk((long long)N->batch * (nx / 2) * (ny / 2) * C)), dim3(256),
                       0, N->ctx->stream, dY, ldy, yoff, dX, N->batch, nx, ny, C);
}

static void forward(bolt_nn* N) {
    const int w = N->w, X0 = N->nx, Y0 = N->ny;
    const int X1 = X0 / 2, Y1 = Y0 / 2, X2 = X1 / 2, Y2 = Y1 / 2, X3 = X2 / 2, Y3 = Y2 / 2;
    auto& C = N->conv;
    conv_fwd(N, C[0], N->x0, X0, Y0, N->e1a, w, 0);
    conv_fwd(N, C[1], N->e1a, X0, Y0, N->cat1, 2 * w, w);
    pool(N, N->cat1, 2 * w, w, N->p1, X0, Y0, w);
    conv_fwd(N, C[2], N->p1, X1, Y1, N->e2a, 2 * w, 0);
    conv_fwd(N, C[3], N->e2a, X1, Y1, N->cat2, 4 * w, 2 * w);
    pool(N, N->cat2, 4 * w, 2 * w, N->p2, X1, Y1, 2 * w);
    conv_fwd(N, C[4], N->p2, X2, Y2, N->e3a, 4 * w, 0);
    conv_fwd(N, C[5], N->e3a, X2, Y2, N->cat3, 8 * w, 4 * w);
    pool(N, N->cat3, 8 * w, 4 * w, N->p3, X2, Y2, 4 * w);
    conv_fwd(N, C[6], N->p3, X3, Y3, N->ba, 4 * w, 0);
    conv_fwd(N, C[7], N->ba, X3, Y3, N->bb, 4 * w, 0);
    up(N, N->bb, N->cat3, 8 * w, 0, X2, Y2, 4 * w);
    conv_fwd(N, C[8], N->cat3, X2, Y2, N->u3, 2 * w, 0);
    up(N, N->u3, N->cat2, 4 * w, 0, X1, Y1, 2 * w);
    conv_fwd(N, C[9], N->cat2, X1, Y1, N->u2, w, 0);
    up(N, N->u2, N->cat1, 2 * w, 0, X0, Y0, w);
    conv_fwd(N, C[10], N->cat1, X0, Y0, N->u1, w, 0);
    const Conv& O = C[11];
    const long long P = (long long)N->batch * X0 * Y0;
    gemm(N, kMK, kNK, kStoreF32, N->u1, w, N->Wb + O.w_off, w, N->out, kOutC, 0, (int)P, kOutC, w,
         N->W + O.b_off, false);
}

static void backward(bolt_nn* N) {
    const int w = N->w, X0 = N->nx, Y0 = N->ny;
    const int X1 = X0 / 2, Y1 = Y0 / 2, X2 = X1 / 2, Y2 = Y1 / 2, X3 = X2 / 2, Y3 = Y2 / 2;
    auto& C = N->conv;
    hipStream_t s = N->ctx->stream;
    const long long P0 = (long long)N->batch * X0 * Y0;
    // output 1x1 conv
    const Conv& O = C[11];
    hipLaunchKernelGGL(relu_grad, dim3(nblk(P0 * kOutC)), dim3(256), 0, s, N->dout, kOutC, 0, (const u16*)nullptr,
                       0, 0, N->dz, P0, kOutC, 0);
    hipLaunchKernelGGL(bias_grad, dim3(kOutC), dim3(256), 0, s, N->dz, P0, kOutC, N->G + O.b_off);
    gemm(N, kKM, kKN, kAtomicF32, N->dz, kOutC, N->u1, w, N->G + O.w_off, w, 0, kOutC, w, (int)P0, nullptr,
         false, 4096);
    gemm(N, kMK, kKN, kStoreF32, N->dz, kOutC, N->Wb + O.w_off, w, N->du1, w, 0, (int)P0, w, kOutC, nullptr,
         false);
    conv_bwd(N, C[10], N->cat1, X0, Y0, N->u1, w, 0, N->du1, w, 0, N->dcat1);
    up_bwd(N, N->dcat1, 2 * w, 0, N->du2, X0, Y0, w);
    conv_bwd(N, C[9], N->cat2, X1, Y1, N->u2, w, 0, N->du2, w, 0, N->dcat2);
    up_bwd(N, N->dcat2, 4 * w, 0, N->du3, X1, Y1, 2 * w);
    conv_bwd(N, C[8], N->cat3, X2, Y2, N->u3, 2 * w, 0, N->du3, 2 * w, 0, N->dcat3);
    up_bwd(N, N->dcat3, 8 * w, 0, N->dbb, X2, Y2, 4 * w);
    conv_bwd(N, C[7], N->ba, X3, Y3, N->bb, 4 * w, 0, N->dbb, 4 * w, 0, N->dba);
    conv_bwd(N, C[6], N->p3, X3, Y3, N->ba, 4 * w, 0, N->dba, 4 * w, 0, N->dp3);
    pool_bwd(N, N->dp3, N->dcat3, 8 * w, 4 * w, X2, Y2, 4 * w);
    conv_bwd(N, C[5], N->e3a, X2, Y2, N->cat3, 8 * w, 4 * w, N->dcat3, 8 * w, 4 * w, N->de3a);
    conv_bwd(N, C[4], N->p2, X2, Y2, N->e3a, 4 * w, 0, N->de3a, 4 * w, 0, N->dp2);
    pool_bwd(N, N->dp2, N->dcat2, 4 * w, 2 * w, X1, Y1, 2 * w);
    conv_bwd(N, C[3], N->e2a, X1, Y1, N->cat2, 4 * w, 2 * w, N->dcat2, 4 * w, 2 * w, N->de2a);
    conv_bwd(N, C[2], N->p1, X1, Y1, N->e2a, 2 * w, 0, N->de2a, 2 * w, 0, N->dp1);
    pool_bwd(N, N->dp1, N->dcat1, 2 * w, w, X0, Y0, w);
    conv_bwd(N, C[1], N->e1a, X0, Y0, N->cat1, 2 * w, w, N->dcat1, 2 * w, w, N->de1a);
    conv_bwd(N, C[0], N->x0, X0, Y0, N->e1a, w, 0, N->de1a, w, 0, nullptr);
}

static int zero_grads(bolt_nn* N) {
    hipLaunchKernelGGL(fill_f32, dim3(nblk((long long)N->nparam)), dim3(256), 0, N->ctx->stream, N->G,
                       (long long)N->nparam, 0.f);
    HIP_TRY(hipGetLastError());
    return BOLT_OK;
}

static int loss(bolt_nn* N, double* out) {
    const long long n = (long long)N->batch * N->nx * N->ny * kOutC;
    hipStream_t s = N->ctx->stream;
    hipLaunchKernelGGL(mse_loss, dim3(256), dim3(256), 0, s, N->out, N->target, N->dout, n, N->partial);
    HIP_TRY(hipGetLastError());
    double part[256];
    HIP_TRY(hipMemcpyAsync(part, N->partial, sizeof part, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    double t = 0;
    for (double v : part) t += v;
    if (out) *out = t / (double)n;
    return BOLT_OK;
}

static int adam(bolt_nn* N, float lr) {
    N->adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)N->adam_t), c2 = 1.f - powf(b2, (float)N->adam_t);
    hipLaunchKernelGGL(adam_step, dim3(nblk((long long)N->nparam)), dim3(256), 0, N->ctx->stream, N->W, N->Wb,
                       N->G, N->M1, N->M2, (long long)N->nparam, lr, b1, b2, 1e-8f, c1, c2);
    HIP_TRY(hipGetLastError());
    return BOLT_OK;
}

// host features [n][nx][ny][4] (float) -> device [n][nx][ny][16] bf16 (zero padded)
static void pad_features(const float* x, size_t pixels, std::vector<uint16_t>& out) {
    out.assign(pixels * kInC, 0);
    for (size_t p = 0; p < pixels; ++p)
        for (int c = 0; c < kFeat; ++c) {
            float v = x[p * kFeat + c];
            uint32_t u;
            memcpy(&u, &v, 4);
            u += 0x7fffu + ((u >> 16) & 1u);
            out[p * kInC + c] = (uint16_t)(u >> 16);
        }
}

static int upload_batch_inputs(bolt_nn* N, const float* x, const float* y) {
    const size_t pix = (size_t)N->batch * N->nx * N->ny;
    std::vector<uint16_t> xb;
    pad_features(x, pix, xb);
    hipStream_t s = N->ctx->stream;
    HIP_TRY(hipMemcpyAsync(N->x0, xb.data(), xb.size() * 2, hipMemcpyHostToDevice, s));
    if (y) HIP_TRY(hipMemcpyAsync(N->target, y, pix * kOutC * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return BOLT_OK;
}

int bolt_nn_destroy(bolt_nn* N) {
    if (!N) return BOLT_OK;
    hipSetDevice(N->ctx->device);
    hipStreamSynchronize(N->ctx->stream);
    for (auto& b : N->bufs) hipFree(b.p);
    delete N;
    return BOLT_OK;
}

int bolt_nn_create(bolt_ctx* ctx, int nx, int ny, int batch, int width, uint64_t seed, bolt_nn** out) {
    if (!ctx || !out) return fail(BOLT_EINVAL, "null argument");
    *out = nullptr;
    if (nx % 8 || ny % 8 || nx < 8 || ny < 8) return fail(BOLT_EINVAL, "nx and ny must be multiples of 8");
    if (width < 16 || width % 16) return fail(BOLT_EINVAL, "width must be a multiple of 16");
    if (batch < 1 || batch > 64) return fail(BOLT_EINVAL, "batch must be 1..64");
    if ((long long)batch * nx * ny > (1ll << 26)) return fail(BOLT_EINVAL, "batch * nx * ny too large");
    HIP_TRY(hipSetDevice(ctx->device));
    auto N = std::make_unique<bolt_nn>();
    N->ctx = ctx;
    N->nx = nx;
    N->ny = ny;
    N->batch = batch;
    N->w = width;
    const int w = width;
    const int spec[12][3] = {{kInC, w, 9},  {w, w, 9},         {w, 2 * w, 9},     {2 * w, 2 * w, 9},
                             {2 * w, 4 * w, 9}, {4 * w, 4 * w, 9}, {4 * w, 4 * w, 9}, {4 * w, 4 * w, 9},
                             {8 * w, 2 * w, 9}, {4 * w, w, 9},     {2 * w, w, 9},     {w, kOutC, 1}};
    size_t off = 0;
    for (int i = 0; i < 12; ++i) {
        Conv c{spec[i][0], spec[i][1], spec[i][2], i != 11, 0, 0};
        c.w_off = off;
        off += (size_t)c.cout * c.taps * c.cin;
        c.b_off = off;
        off += (size_t)c.cout;
        N->conv.push_back(c);
    }
    N->nparam = off;
    bolt_nn* n = N.get();
    int rc = 0;
    auto A = [&](void** p, size_t bytes) {
        if (!rc) rc = nn_alloc(n, p, bytes);
    };
    A((void**)&n->W, off * 4);
    A((void**)&n->Wb, off * 2);
    A((void**)&n->G, off * 4);
    A((void**)&n->M1, off * 4);
    A((void**)&n->M2, off * 4);
    const size_t P0 = (size_t)batch * nx * ny, P1 = P0 / 4, P2 = P0 / 16, P3 = P0 / 64;
    A((void**)&n->x0, P0 * kInC * 2);
    A((void**)&n->e1a, P0 * w * 2);
    A((void**)&n->cat1, P0 * 2 * w * 2);
    A((void**)&n->p1, P1 * w * 2);
    A((void**)&n->e2a, P1 * 2 * w * 2);
    A((void**)&n->cat2, P1 * 4 * w * 2);
    A((void**)&n->p2, P2 * 2 * w * 2);
    A((void**)&n->e3a, P2 * 4 * w * 2);
    A((void**)&n->cat3, P2 * 8 * w * 2);
    A((void**)&n->p3, P3 * 4 * w * 2);
    A((void**)&n->ba, P3 * 4 * w * 2);
    A((void**)&n->bb, P3 * 4 * w * 2);
    A((void**)&n->u3, P2 * 2 * w * 2);
    A((void**)&n->u2, P1 * w * 2);
    A((void**)&n->u1, P0 * w * 2);
    A((void**)&n->out, P0 * kOutC * 4);
    A((void**)&n->target, P0 * kOutC * 4);
    A((void**)&n->dout, P0 * kOutC * 4);
    A((void**)&n->du1, P0 * w * 4);
    A((void**)&n->dcat1, P0 * 2 * w * 4);
    A((void**)&n->du2, P1 * w * 4);
    A((void**)&n->dcat2, P1 * 4 * w * 4);
    A((void**)&n->du3, P2 * 2 * w * 4);
    A((void**)&n->dcat3, P2 * 8 * w * 4);
    A((void**)&n->dbb, P3 * 4 * w * 4);
    A((void**)&n->dba, P3 * 4 * w * 4);
    A((void**)&n->dp3, P3 * 4 * w * 4);
    A((void**)&n->de3a, P2 * 4 * w * 4);
    A((void**)&n->dp2, P2 * 2 * w * 4);
    A((void**)&n->de2a, P1 * 2 * w * 4);
    A((void**)&n->dp1, P1 * w * 4);
    A((void**)&n->de1a, P0 * w * 4);
    // scratch sized for the widest layer at each resolution: max over layers of P * cout / cin
    size_t zmax = 0, colmax = 0;
    const size_t Ps[12] = {P0, P0, P1, P1, P2, P2, P3, P3, P2, P1, P0, P0};
    for (int i = 0; i < 12; ++i) {
        zmax = std::max(zmax, Ps[i] * (size_t)N->conv[i].cout);
        colmax = std::max(colmax, Ps[i] * 9 * (size_t)N->conv[i].cin);
    }
    A((void**)&n->dz, zmax * 2);
    A((void**)&n->col, colmax * 2);
    A((void**)&n->dcol, colmax * 4);
    A((void**)&n->partial, 256 * sizeof(double));
    if (rc) return rc;
    // He-normal initialisation, zero biases
    std::vector<float> h(off, 0.f);
    std::mt19937_64 rng(seed);
    for (auto& c : N->conv) {
        std::normal_distribution<float> nd(0.f, sqrtf(2.f / (float)(c.cin * c.taps)));
        for (size_t i = 0; i < (size_t)c.cout * c.taps * c.cin; ++i) h[c.w_off + i] = nd(rng);
        if (!c.relu)
            for (size_t i = 0; i < (size_t)c.cout * c.taps * c.cin; ++i) h[c.w_off + i] *= 0.1f;
    }
    HIP_TRY(hipMemcpy(n->W, h.data(), off * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(n->M1, 0, off * 4));
    HIP_TRY(hipMemset(n->M2, 0, off * 4));
    hipLaunchKernelGGL(to_bf16, dim3(nblk((long long)off)), dim3(256), 0, ctx->stream, n->W, n->Wb, (long long)off);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    *out = N.release();
    return BOLT_OK;
}

int bolt_nn_param_count(const bolt_nn* N, uint64_t* out) {
    if (!N || !out) return fail(BOLT_EINVAL, "null argument");
    *out = N->nparam;
    return BOLT_OK;
}

int bolt_nn_get_params(bolt_nn* N, float* out, uint64_t n) {
    if (!N || !out || n != N->nparam) return fail(BOLT_EINVAL, "expected %zu parameters", N ? N->nparam : 0);
    HIP_TRY(hipSetDevice(N->ctx->device));
    HIP_TRY(hipMemcpy(out, N->W, n * 4, hipMemcpyDeviceToHost));
    return BOLT_OK;
}

int bolt_nn_set_params(bolt_nn* N, const float* in, uint64_t n) {
    if (!N || !in || n != N->nparam) return fail(BOLT_EINVAL, "expected %zu parameters", N ? N->nparam : 0);
    HIP_TRY(hipSetDevice(N->ctx->device));
    HIP_TRY(hipMemcpy(N->W, in, n * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(to_bf16, dim3(nblk((long long)n)), dim3(256), 0, N->ctx->stream, N->W, N->Wb, (long long)n);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(N->ctx->stream));
    return BOLT_OK;
}

int bolt_nn_loss_grad(bolt_nn* N, const float* x, const float* y, double* loss_out, float* grads, uint64_t ng) {
    if (!N || !x || !y) return fail(BOLT_EINVAL, "null argument");
    if (grads && ng != N->nparam) return fail(BOLT_EINVAL, "expected %zu gradients", N->nparam);
    HIP_TRY(hipSetDevice(N->ctx->device));
    int rc = upload_batch_inputs(N, x, y);
    if (!rc) rc = zero_grads(N);
    if (rc) return rc;
    forward(N);
    rc = loss(N, loss_out);
    if (rc) return rc;
    backward(N);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(N->ctx->stream));
    if (grads) HIP_TRY(hipMemcpy(grads, N->G, ng * 4, hipMemcpyDeviceToHost));
    return BOLT_OK;
}

int bolt_nn_set_data(bolt_nn* N, int n, const float* x, const float* y) {
    if (!N || !x || !y || n < 1) return fail(BOLT_EINVAL, "bad arguments");
    HIP_TRY(hipSetDevice(N->ctx->device));
    if (N->data_x) {
        hipFree(N->data_x);
        hipFree(N->data_y);
        N->data_x = nullptr;
        N->data_y = nullptr;
    }
    const size_t pix = (size_t)N->nx * N->ny;
    HIP_TRY(hipMalloc(&N->data_x, (size_t)n * pix * kInC * 2));
    HIP_TRY(hipMalloc(&N->data_y, (size_t)n * pix * kOutC * 4));
    std::vector<uint16_t> xb;
    for (int i = 0; i < n; ++i) {
        pad_features(x + (size_t)i * pix * kFeat, pix, xb);
        HIP_TRY(hipMemcpy(N->data_x + (size_t)i * pix * kInC, xb.data(), xb.size() * 2, hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMemcpy(N->data_y, y, (size_t)n * pix * kOutC * 4, hipMemcpyHostToDevice));
    N->data_n = n;
    return BOLT_OK;
}

// Minibatch Adam over the uploaded dataset; losses[e] receives epoch e's mean batch loss.
int bolt_nn_train(bolt_nn* N, int epochs, float lr, uint64_t seed, float* losses) {
    return bolt_nn_train_part(N, 0, epochs, epochs, lr, seed, losses);
}

// Epochs [first, first + epochs) of a run of total epochs: the same shuffles and cosine
// learning rates as those epochs of one bolt_nn_train(total) call.
int bolt_nn_train_part(bolt_nn* N, int first, int epochs, int total, float lr, uint64_t seed,
                       float* losses) {
    if (!N || epochs < 1 || first < 0 || first + epochs > total)
        return fail(BOLT_EINVAL, "bad arguments");
    if (N->data_n < N->batch) return fail(BOLT_ESTATE, "dataset smaller than one batch (bolt_nn_set_data)");
    HIP_TRY(hipSetDevice(N->ctx->device));
    hipStream_t s = N->ctx->stream;
    const size_t pix = (size_t)N->nx * N->ny;
    std::vector<int> order(N->data_n);
    std::iota(order.begin(), order.end(), 0);
    std::mt19937_64 rng(seed);
    const int per_epoch = N->data_n / N->batch;
    for (int e = 0; e < first; ++e) std::shuffle(order.begin(), order.end(), rng);
    for (int e = first; e < first + epochs; ++e) {
        std::shuffle(order.begin(), order.end(), rng);
        // cosine decay of the learning rate over the run
        double tot = 0;
        for (int b = 0; b < per_epoch; ++b) {
            const float frac = (float)((long)e * per_epoch + b) / (float)((long)total * per_epoch);
            const float lr_t = lr * 0.5f * (1.f + cosf(3.14159265f * frac));
            for (int k = 0; k < N->batch; ++k) {
                const int idx = order[b * N->batch + k];
                HIP_TRY(hipMemcpyAsync(N->x0 + (size_t)k * pix * kInC, N->data_x + (size_t)idx * pix * kInC,
                                       pix * kInC * 2, hipMemcpyDeviceToDevice, s));
                HIP_TRY(hipMemcpyAsync(N->target + (size_t)k * pix * kOutC, N->data_y + (size_t)idx * pix * kOutC,
                                       pix * kOutC * 4, hipMemcpyDeviceToDevice, s));
            }
            int rc = zero_grads(N);
            if (rc) return rc;
            forward(N);
            double l = 0;
            rc = loss(N, &l);
            if (rc) return rc;
            backward(N);
            rc = adam(N, lr_t);
            if (rc) return rc;
            tot += l;
        }
        if (losses) losses[e - first] = (float)(tot / per_epoch);
    }
    HIP_TRY(hipStreamSynchronize(s));
    return BOLT_OK;
}

int bolt_nn_predict(bolt_nn* N, int n, const float* x, float* y) {
    if (!N || !x || !y || n < 1) return fail(BOLT_EINVAL, "bad arguments");
    HIP_TRY(hipSetDevice(N->ctx->device));
    const size_t pix = (size_t)N->nx * N->ny;
    std::vector<float> xin((size_t)N->batch * pix * kFeat, 0.f);
    for (int i0 = 0; i0 < n; i0 += N->batch) {
        const int cnt = std::min(N->batch, n - i0);
        std::fill(xin.begin(), xin.end(), 0.f);
        memcpy(xin.data(), x + (size_t)i0 * pix * kFeat, (size_t)cnt * pix * kFeat * 4);
        int rc = upload_batch_inputs(N, xin.data(), nullptr);
        if (rc) return rc;
        forward(N);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(y + (size_t)i0 * pix * kOutC, N->out, (size_t)cnt * pix * kOutC * 4,
                               hipMemcpyDeviceToHost, N->ctx->stream));
        HIP_TRY(hipStreamSynchronize(N->ctx->stream));
    }
    return BOLT_OK;
}

// Test hook: C(m, n) = sum_k A(m, k) B(n, k) through the network's GEMM kernel.  a_form:
// 0 = A[m][k], 1 = A[k][m]; b_form: 0 = B[n][k], 1 = B[k][n]; inputs rounded to bf16.
int bolt_nn_gemm_check(bolt_ctx* ctx, int a_form, int b_form, int M, int Nn, int K, const float* a,
                       const float* b, float* c, int k_split) {
    if (!ctx || !a || !b || !c) return fail(BOLT_EINVAL, "null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    bolt_nn tmp;
    tmp.ctx = ctx;
    std::vector<uint16_t> ha((size_t)M * K), hb((size_t)Nn * K);
    auto cv = [](float v) {
        uint32_t u;
        memcpy(&u, &v, 4);
        u += 0x7fffu + ((u >> 16) & 1u);
        return (uint16_t)(u >> 16);
    };
    for (size_t i = 0; i < ha.size(); ++i) ha[i] = cv(a[i]);
    for (size_t i = 0; i < hb.size(); ++i) hb[i] = cv(b[i]);
    u16 *da = nullptr, *db = nullptr;
    float* dc = nullptr;
    int rc = BOLT_OK;
    if (hipMalloc(&da, ha.size() * 2) || hipMalloc(&db, hb.size() * 2) || hipMalloc(&dc, (size_t)M * Nn * 4)) {
        rc = fail(BOLT_ENOMEM, "gemm check allocation failed");
    } else {
        hipMemcpy(da, ha.data(), ha.size() * 2, hipMemcpyHostToDevice);
        hipMemcpy(db, hb.data(), hb.size() * 2, hipMemcpyHostToDevice);
        hipMemset(dc, 0, (size_t)M * Nn * 4);
        const int lda = a_form == kMK ? K : M, ldb = b_form == kNK ? K : Nn;
        gemm(&tmp, a_form, b_form, k_split > 0 ? kAtomicF32 : kStoreF32, da, lda, db, ldb, dc, Nn, 0, M, Nn, K,
             nullptr, false, k_split);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess) e = hipMemcpy(c, dc, (size_t)M * Nn * 4, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(BOLT_EDEVICE, "gemm check: %s", hipGetErrorString(e));
    }
    hipFree(da);
    hipFree(db);
    hipFree(dc);
    return rc;
}
