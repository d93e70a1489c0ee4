// The warm-start network: a small U-Net trained and run with this library's own kernels
// (nn_kernels.hpp: bf16 MFMA GEMM with implicit 3x3 convolution operands, element-wise layers).
//
// Input  [batch][nx][ny][16]  bf16 (the caller's 8 feature channels, zero-padded to 16)
// Output [batch][nx][ny][3]   float
//
// A U-Net of depth D (2x2 poolings; D = 3 below), channels ch[l] = w * min(2^l, 4) at
// level l and decoder channels dch[l] = ch[max(l - 1, 0)]:
//   level 0 (nx, ny):     e1a = conv(x, 16->w)          e1b = conv(e1a, w->w)
//   level 1 (nx/2, ny/2): e2a = conv(pool e1b, w->2w)   e2b = conv(e2a, 2w->2w)
//   level 2:              e3a = conv(pool e2b, 2w->4w)  e3b = conv(e3a, 4w->4w)
//   level 3 (bottleneck): ba  = conv(pool e3b, 4w->4w)  bb  = conv(ba, 4w->4w)
//   level 2:  u3 = conv([up bb | e3b], 8w->2w)
//   level 1:  u2 = conv([up u3 | e2b], 4w->w)
//   level 0:  u1 = conv([up u2 | e1b], 2w->w)          out = conv1x1(u1, w->3)
// 3x3 convolutions with padding 1 and ReLU, 2x2 average pooling, nearest upsampling.
// Parameters are float master weights (layer order: the encoder pairs level by level, the
// bottleneck pair, the decoder convolutions from the deepest level up, the 1x1 output;
// each W[cout][taps*cin] then b[cout], tap = kx*3 + ky), with a bf16 copy for the GEMMs;
// Adam updates both.
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
constexpr int kFeat = 8;   // caller's feature channels unless bolt_nn_set_features changes it
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
    int nx = 0, ny = 0, batch = 0, w = 32, depth = 3;
    int nfeat = kFeat;     // feature channels of the host inputs (<= kInC)
    std::vector<Conv> conv;
    std::vector<int> ch, dch;   // encoder / decoder channels per level (0..depth)
    size_t nparam = 0;
    float* W = nullptr;    // float master parameters
    u16* Wb = nullptr;     // bf16 copy
    float* G = nullptr;    // gradients
    float* M1 = nullptr;   // Adam moments
    float* M2 = nullptr;
    int adam_t = 0;
    std::vector<Buf> bufs;
    // activations (bf16), per level: ea = first conv of the level, cat = [up | second conv]
    // (levels < depth), pin = pooled input (levels >= 1; pin[0] = x0), u = decoder output
    u16 *x0 = nullptr, *bb = nullptr;
    float* xf = nullptr;   // float features of one batch as uploaded, before padding to bf16
    std::vector<u16*> ea, cat, pin, u;
    float* out;
    // gradients (float)
    float *dout = nullptr, *dbb = nullptr;
    std::vector<float*> du, dcat, dpin, dea;
    u16* dz;       // bf16 pre-activation gradient of the layer being processed
    u16* wt;       // reflected, transposed kernel of the layer being processed (input gradient)
    float* target;
    double* partial;
    // dataset on the device
    u16* data_x = nullptr;
    float* data_y = nullptr;
    int data_n = 0;

    // layer indices: level l's pair (the bottleneck is level depth), decoder, output
    int la(int l) const { return 2 * l; }
    int lb(int l) const { return 2 * l + 1; }
    int lu(int l) const { return 2 * depth + 2 + (depth - 1 - l); }
    int lout() const { return 3 * depth + 2; }
    int catc(int l) const { return dch[l + 1] + ch[l]; }
};

static int nn_alloc(bolt_nn* N, void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(BOLT_ENOMEM, "network: hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    N->bufs.push_back({*p, bytes});
    return BOLT_OK;
}

// Pixels per split-K slice of the weight gradients.  Small slices give many work-groups,
// which hide the operand loads these GEMMs are bound by, at the cost of more float atomics.
// Training samples/s at 512 x 256, depth 6, width 32, batch 8 (one session): slices of 4096
// -> 1,343; 2048 -> 1,485; 1024 -> 1,551; 512 -> 1,574; 256 -> 1,479.
constexpr int kWgradSlice = 512;

// ---- launch helpers ---------------------------------------------------------------------
static void gemm(bolt_nn* N, int af, int bf, int out, const u16* a, int lda, const u16* b, int ldb,
                 void* c, int ldc, int c_off, int M, int Nn, int K, const float* bias, bool relu,
                 int k_split = 0, ConvGeom cg = conv_geom(1, 1, 8)) {
    GemmArgs g;
    g.conv = cg;
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
    g.k_split = k_split > 0 ? k_split : ((K + kStage - 1) / kStage) * kStage;
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
    else if (af == kKM && bf == kKN) BOLT_GEMM_OUT(kKM, kKN);
    else if (af == kMKconv && bf == kNK) BOLT_GEMM_OUT(kMKconv, kNK);
    else BOLT_GEMM_OUT(kKM, kKNconv);
#undef BOLT_GEMM_OUT
#undef BOLT_GEMM
}

// dz = the layer's pre-activation gradient (bf16) and dbias += its column sums: up to 1024
// work-groups of contiguous row chunks
static void relu_bias_grad_launch(bolt_nn* N, const float* dY, int lddy, int dyoff, const u16* Y, int ldy,
                                  int yoff, long long P, int C, bool relu, float* db) {
    const long long groups = std::min<long long>(1024, std::max<long long>(1, P / 64));
    const long long rows = (P + groups - 1) / groups;
    hipLaunchKernelGGL(relu_bias_grad, dim3((unsigned)((P + rows - 1) / rows)), dim3(256), 0, N->ctx->stream,
                       dY, lddy, dyoff, Y, ldy, yoff, N->dz, P, C, relu ? 1 : 0, rows, db);
}

// forward 3x3 conv of X (channel count cin, all channels) into Y (ldy, channel offset yoff):
// an implicit GEMM whose A operand gathers the 3x3 neighbourhoods of X
static void conv_fwd(bolt_nn* N, const Conv& L, const u16* X, int nx, int ny, u16* Y, int ldy, int yoff) {
    const long long P = (long long)N->batch * nx * ny;
    if (L.taps == 9)
        gemm(N, kMKconv, kNK, kStoreBF16, X, L.cin, N->Wb + L.w_off, 9 * L.cin, Y, ldy, yoff, (int)P, L.cout,
             9 * L.cin, N->W + L.b_off, L.relu, 0, conv_geom(nx, ny, L.cin));
}

// backward of a 3x3 conv whose output was written at (Y, ldy, yoff) with gradient (dY, lddy,
// dyoff); accumulates parameter gradients and writes the input gradient into dX (if any)
static void conv_bwd(bolt_nn* N, const Conv& L, const u16* X, int nx, int ny, const u16* Y, int ldy,
                     int yoff, const float* dY, int lddy, int dyoff, float* dX) {
    hipStream_t s = N->ctx->stream;
    const long long P = (long long)N->batch * nx * ny;
    relu_bias_grad_launch(N, dY, lddy, dyoff, Y, ldy, yoff, P, L.cout, L.relu, N->G + L.b_off);
    // dW[co][tap*cin + ci] += sum_p dz[p][co] X[p + shift(tap)][ci]  (B gathers X)
    gemm(N, kKM, kKNconv, kAtomicF32, N->dz, L.cout, X, L.cin, N->G + L.w_off, 9 * L.cin, 0, L.cout,
         9 * L.cin, (int)P, nullptr, false, kWgradSlice, conv_geom(nx, ny, L.cin));
    if (dX) {
        // dX[p][ci] = sum_tap sum_co dz[p + shift(tap)][co] W[co][(8 - tap)*cin + ci]: the
        // forward form on dz with the reflected, transposed kernel Wt
        const long long nw = (long long)L.cout * 9 * L.cin;
        hipLaunchKernelGGL(flip_weights, dim3(nblk(nw)), dim3(256), 0, s, N->Wb + L.w_off, L.cout, L.cin, N->wt);
        gemm(N, kMKconv, kNK, kStoreF32, N->dz, L.cout, N->wt, 9 * L.cout, dX, L.cin, 0, (int)P, L.cin,
             9 * L.cout, nullptr, false, 0, conv_geom(nx, ny, L.cout));
    }
}

// pooling and upsampling, 8 channels per thread (every channel count, leading dimension and
// channel offset here is a multiple of 16)
static void pool(bolt_nn* N, const u16* X, int ldx, int xoff, u16* Y, int nx, int ny, int C) {
    const uint32_t n = (uint32_t)((long long)N->batch * (nx / 2) * (ny / 2) * (C / 8));
    hipLaunchKernelGGL(avgpool2, dim3(nblk(n)), dim3(256), 0, N->ctx->stream, X, ldx, xoff, Y, n,
                       grid8(nx / 2, ny / 2, C));
}
static void pool_bwd(bolt_nn* N, const float* dY, float* dX, int lddx, int dxoff, int nx, int ny, int C) {
    const uint32_t n = (uint32_t)((long long)N->batch * nx * ny * (C / 8));
    hipLaunchKernelGGL(avgpool2_grad, dim3(nblk(n)), dim3(256), 0, N->ctx->stream, dY, dX, lddx, dxoff, n,
                       grid8(nx, ny, C));
}
static void up(bolt_nn* N, const u16* X, u16* Y, int ldy, int yoff, int nx, int ny, int C) {
    const uint32_t n = (uint32_t)((long long)N->batch * nx * ny * (C / 8));
    hipLaunchKernelGGL(upsample2, dim3(nblk(n)), dim3(256), 0, N->ctx->stream, X, Y, ldy, yoff, n,
                       grid8(nx, ny, C));
}
static void up_bwd(bolt_nn* N, const float* dY, int ldy, int yoff, float* dX, int nx, int ny, int C) {
    const uint32_t n = (uint32_t)((long long)N->batch * (nx / 2) * (ny / 2) * (C / 8));
    hipLaunchKernelGGL(upsample2_grad, dim3(nblk(n)), dim3(256), 0, N->ctx->stream, dY, ldy, yoff, dX, n,
                       grid8(nx / 2, ny / 2, C));
}

static void forward(bolt_nn* N) {
    const int D = N->depth;
    auto& C = N->conv;
    for (int l = 0; l < D; ++l) {
        const int X = N->nx >> l, Y = N->ny >> l, cc = N->catc(l);
        conv_fwd(N, C[N->la(l)], N->pin[l], X, Y, N->ea[l], N->ch[l], 0);
        conv_fwd(N, C[N->lb(l)], N->ea[l], X, Y, N->cat[l], cc, N->dch[l + 1]);
        pool(N, N->cat[l], cc, N->dch[l + 1], N->pin[l + 1], X, Y, N->ch[l]);
    }
    const int XD = N->nx >> D, YD = N->ny >> D;
    conv_fwd(N, C[N->la(D)], N->pin[D], XD, YD, N->ea[D], N->ch[D], 0);
    conv_fwd(N, C[N->lb(D)], N->ea[D], XD, YD, N->bb, N->ch[D], 0);
    const u16* prev = N->bb;
    for (int l = D - 1; l >= 0; --l) {
        const int X = N->nx >> l, Y = N->ny >> l;
        up(N, prev, N->cat[l], N->catc(l), 0, X, Y, N->dch[l + 1]);
        conv_fwd(N, C[N->lu(l)], N->cat[l], X, Y, N->u[l], N->dch[l], 0);
        prev = N->u[l];
    }
    const Conv& O = C[N->lout()];
    const long long P = (long long)N->batch * N->nx * N->ny;
    gemm(N, kMK, kNK, kStoreF32, N->u[0], O.cin, N->Wb + O.w_off, O.cin, N->out, kOutC, 0, (int)P, kOutC,
         O.cin, N->W + O.b_off, false);
}

static void backward(bolt_nn* N) {
    const int D = N->depth;
    auto& C = N->conv;
    const long long P0 = (long long)N->batch * N->nx * N->ny;
    // output 1x1 conv
    const Conv& O = C[N->lout()];
    const int w0 = O.cin;
    relu_bias_grad_launch(N, N->dout, kOutC, 0, nullptr, 0, 0, P0, kOutC, false, N->G + O.b_off);
    gemm(N, kKM, kKN, kAtomicF32, N->dz, kOutC, N->u[0], w0, N->G + O.w_off, w0, 0, kOutC, w0, (int)P0, nullptr,
         false, kWgradSlice);
    gemm(N, kMK, kKN, kStoreF32, N->dz, kOutC, N->Wb + O.w_off, w0, N->du[0], w0, 0, (int)P0, w0, kOutC, nullptr,
         false);
    // decoder, from level 0 down to the bottleneck
    for (int l = 0; l < D; ++l) {
        const int X = N->nx >> l, Y = N->ny >> l, cc = N->catc(l);
        conv_bwd(N, C[N->lu(l)], N->cat[l], X, Y, N->u[l], N->dch[l], 0, N->du[l], N->dch[l], 0, N->dcat[l]);
        up_bwd(N, N->dcat[l], cc, 0, l + 1 < D ? N->du[l + 1] : N->dbb, X, Y, N->dch[l + 1]);
    }
    const int XD = N->nx >> D, YD = N->ny >> D;
    conv_bwd(N, C[N->lb(D)], N->ea[D], XD, YD, N->bb, N->ch[D], 0, N->dbb, N->ch[D], 0, N->dea[D]);
    conv_bwd(N, C[N->la(D)], N->pin[D], XD, YD, N->ea[D], N->ch[D], 0, N->dea[D], N->ch[D], 0, N->dpin[D]);
    // encoder; the pooling gradient adds to the skip gradient of the level's second conv
    for (int l = D - 1; l >= 0; --l) {
        const int X = N->nx >> l, Y = N->ny >> l, cc = N->catc(l);
        pool_bwd(N, N->dpin[l + 1], N->dcat[l], cc, N->dch[l + 1], X, Y, N->ch[l]);
        conv_bwd(N, C[N->lb(l)], N->ea[l], X, Y, N->cat[l], cc, N->dch[l + 1], N->dcat[l], cc, N->dch[l + 1],
                 N->dea[l]);
        conv_bwd(N, C[N->la(l)], N->pin[l], X, Y, N->ea[l], N->ch[l], 0, N->dea[l], N->ch[l], 0,
                 l > 0 ? N->dpin[l] : nullptr);
    }
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

// host features [n][nx][ny][nfeat] (float) -> device [n][nx][ny][16] bf16 (zero padded)
// Uploads the float features of `cnt` samples and pads them to bf16 [pixels][kInC] at dst
// on the device; pixels of the samples after `cnt` (up to `slots` samples) are zeroed.
static int upload_features(bolt_nn* N, const float* x, int cnt, int slots, u16* dst) {
    const size_t pix = (size_t)N->nx * N->ny;
    hipStream_t s = N->ctx->stream;
    HIP_TRY(hipMemcpyAsync(N->xf, x, (size_t)cnt * pix * N->nfeat * 4, hipMemcpyHostToDevice, s));
    const long long n = (long long)slots * pix;
    hipLaunchKernelGGL((pad_features<kInC>), dim3(nblk(n * kInC)), dim3(256), 0, s, N->xf, n,
                       (long long)cnt * pix, N->nfeat, dst);
    HIP_TRY(hipGetLastError());
    return BOLT_OK;
}

static int upload_batch_inputs(bolt_nn* N, const float* x, int cnt, const float* y) {
    const size_t pix = (size_t)N->batch * N->nx * N->ny;
    int rc = upload_features(N, x, cnt, N->batch, N->x0);
    if (rc) return rc;
    hipStream_t s = N->ctx->stream;
    if (y) HIP_TRY(hipMemcpyAsync(N->target, y, pix * kOutC * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return BOLT_OK;
}

int bolt_nn_destroy(bolt_nn* N) {
    if (!N) return BOLT_OK;
    // teardown: errors here have no caller to act on them
    (void)hipSetDevice(N->ctx->device);
    (void)hipStreamSynchronize(N->ctx->stream);
    for (auto& b : N->bufs) (void)hipFree(b.p);
    delete N;
    return BOLT_OK;
}

int bolt_nn_create(bolt_ctx* ctx, int nx, int ny, int batch, int width, int depth, uint64_t seed,
                   bolt_nn** out) {
    if (!ctx || !out) return fail(BOLT_EINVAL, "null argument");
    *out = nullptr;
    if (depth < 1 || depth > 8) return fail(BOLT_EINVAL, "depth must be 1..8");
    const int m = 1 << depth;
    if (nx % m || ny % m) return fail(BOLT_EINVAL, "nx and ny must be multiples of 2^depth = %d", m);
    if (width < 16 || width % 16 || width > 64)
        return fail(BOLT_EINVAL, "width must be a multiple of 16 from 16 to 64");
    if (batch < 1 || batch > 64) return fail(BOLT_EINVAL, "batch must be 1..64");
    if ((long long)batch * nx * ny > (1ll << 26)) return fail(BOLT_EINVAL, "batch * nx * ny too large");
    HIP_TRY(hipSetDevice(ctx->device));
    auto N = std::make_unique<bolt_nn>();
    N->ctx = ctx;
    N->nx = nx;
    N->ny = ny;
    N->batch = batch;
    N->w = width;
    N->depth = depth;
    const int D = depth, w = width;
    for (int l = 0; l <= D; ++l) N->ch.push_back(w * std::min(1 << l, 4));
    for (int l = 0; l <= D; ++l) N->dch.push_back(l == D ? N->ch[D] : N->ch[std::max(l - 1, 0)]);
    // layers in parameter order, with the pixel count each runs over
    std::vector<size_t> Ps;
    const size_t P0 = (size_t)batch * nx * ny;
    auto add = [&](int cin, int cout, int taps, bool relu, int level) {
        N->conv.push_back(Conv{cin, cout, taps, relu, 0, 0});
        Ps.push_back(P0 >> (2 * level));
    };
    for (int l = 0; l <= D; ++l) {
        add(l == 0 ? kInC : N->ch[l - 1], N->ch[l], 9, true, l);
        add(N->ch[l], N->ch[l], 9, true, l);
    }
    for (int l = D - 1; l >= 0; --l) add(N->catc(l), N->dch[l], 9, true, l);
    add(N->dch[0], kOutC, 1, false, 0);
    size_t off = 0;
    for (auto& c : N->conv) {
        c.w_off = off;
        off += (size_t)c.cout * c.taps * c.cin;
        c.b_off = off;
        off += (size_t)c.cout;
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
    n->ea.assign(D + 1, nullptr);
    n->cat.assign(D, nullptr);
    n->pin.assign(D + 1, nullptr);
    n->u.assign(D, nullptr);
    n->dea.assign(D + 1, nullptr);
    n->dcat.assign(D, nullptr);
    n->dpin.assign(D + 1, nullptr);
    n->du.assign(D, nullptr);
    A((void**)&n->x0, P0 * kInC * 2);
    A((void**)&n->xf, P0 * kInC * 4);
    n->pin[0] = n->x0;
    for (int l = 0; l <= D; ++l) {
        const size_t P = P0 >> (2 * l);
        A((void**)&n->ea[l], P * n->ch[l] * 2);
        A((void**)&n->dea[l], P * n->ch[l] * 4);
        if (l >= 1) {
            A((void**)&n->pin[l], P * n->ch[l - 1] * 2);
            A((void**)&n->dpin[l], P * n->ch[l - 1] * 4);
        }
        if (l < D) {
            A((void**)&n->cat[l], P * n->catc(l) * 2);
            A((void**)&n->dcat[l], P * n->catc(l) * 4);
            A((void**)&n->u[l], P * n->dch[l] * 2);
            A((void**)&n->du[l], P * n->dch[l] * 4);
        }
    }
    const size_t PD = P0 >> (2 * D);
    A((void**)&n->bb, PD * n->ch[D] * 2);
    A((void**)&n->dbb, PD * n->ch[D] * 4);
    A((void**)&n->out, P0 * kOutC * 4);
    A((void**)&n->target, P0 * kOutC * 4);
    A((void**)&n->dout, P0 * kOutC * 4);
    // scratch sized for the widest layer: max over layers of P * cout and of the kernel size
    size_t zmax = 0, wmax = 0;
    for (size_t i = 0; i < N->conv.size(); ++i) {
        zmax = std::max(zmax, Ps[i] * (size_t)N->conv[i].cout);
        wmax = std::max(wmax, (size_t)N->conv[i].cout * N->conv[i].taps * N->conv[i].cin);
    }
    A((void**)&n->dz, zmax * 2);
    A((void**)&n->wt, wmax * 2);
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
    // everything on the context's non-blocking stream: the null stream orders nothing here
    HIP_TRY(hipMemcpyAsync(n->W, h.data(), off * 4, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemsetAsync(n->M1, 0, off * 4, ctx->stream));
    HIP_TRY(hipMemsetAsync(n->M2, 0, off * 4, ctx->stream));
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
    HIP_TRY(hipMemcpyAsync(out, N->W, n * 4, hipMemcpyDeviceToHost, N->ctx->stream));
    HIP_TRY(hipStreamSynchronize(N->ctx->stream));
    return BOLT_OK;
}

int bolt_nn_set_params(bolt_nn* N, const float* in, uint64_t n) {
    if (!N || !in || n != N->nparam) return fail(BOLT_EINVAL, "expected %zu parameters", N ? N->nparam : 0);
    HIP_TRY(hipSetDevice(N->ctx->device));
    HIP_TRY(hipMemcpyAsync(N->W, in, n * 4, hipMemcpyHostToDevice, N->ctx->stream));
    hipLaunchKernelGGL(to_bf16, dim3(nblk((long long)n)), dim3(256), 0, N->ctx->stream, N->W, N->Wb, (long long)n);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(N->ctx->stream));
    return BOLT_OK;
}

int bolt_nn_set_features(bolt_nn* N, int nfeat) {
    if (!N) return fail(BOLT_EINVAL, "null network");
    if (nfeat < 1 || nfeat > kInC) return fail(BOLT_EINVAL, "feature channels must be 1..%d", kInC);
    N->nfeat = nfeat;
    return BOLT_OK;
}

int bolt_nn_features(const bolt_nn* N, int* out) {
    if (!N || !out) return fail(BOLT_EINVAL, "null argument");
    *out = N->nfeat;
    return BOLT_OK;
}

int bolt_nn_loss_grad(bolt_nn* N, const float* x, const float* y, double* loss_out, float* grads, uint64_t ng) {
    if (!N || !x || !y) return fail(BOLT_EINVAL, "null argument");
    if (grads && ng != N->nparam) return fail(BOLT_EINVAL, "expected %zu gradients", N->nparam);
    HIP_TRY(hipSetDevice(N->ctx->device));
    int rc = upload_batch_inputs(N, x, N->batch, y);
    if (!rc) rc = zero_grads(N);
    if (rc) return rc;
    forward(N);
    rc = loss(N, loss_out);
    if (rc) return rc;
    backward(N);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(N->ctx->stream));
    if (grads) {
        HIP_TRY(hipMemcpyAsync(grads, N->G, ng * 4, hipMemcpyDeviceToHost, N->ctx->stream));
        HIP_TRY(hipStreamSynchronize(N->ctx->stream));
    }
    return BOLT_OK;
}

int bolt_nn_set_data(bolt_nn* N, int n, const float* x, const float* y) {
    if (!N || !x || !y || n < 1) return fail(BOLT_EINVAL, "bad arguments");
    HIP_TRY(hipSetDevice(N->ctx->device));
    if (N->data_x) {
        (void)hipFree(N->data_x);
        (void)hipFree(N->data_y);
        N->data_x = nullptr;
        N->data_y = nullptr;
    }
    const size_t pix = (size_t)N->nx * N->ny;
    HIP_TRY(hipMalloc(&N->data_x, (size_t)n * pix * kInC * 2));
    HIP_TRY(hipMalloc(&N->data_y, (size_t)n * pix * kOutC * 4));
    // a batch of samples at a time through the float staging buffer (stream-ordered reuse)
    for (int i0 = 0; i0 < n; i0 += N->batch) {
        const int cnt = std::min(N->batch, n - i0);
        int rc = upload_features(N, x + (size_t)i0 * pix * N->nfeat, cnt, cnt, N->data_x + (size_t)i0 * pix * kInC);
        if (rc) return rc;
    }
    HIP_TRY(hipMemcpyAsync(N->data_y, y, (size_t)n * pix * kOutC * 4, hipMemcpyHostToDevice, N->ctx->stream));
    HIP_TRY(hipStreamSynchronize(N->ctx->stream));
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
    for (int i0 = 0; i0 < n; i0 += N->batch) {
        const int cnt = std::min(N->batch, n - i0);
        int rc = upload_batch_inputs(N, x + (size_t)i0 * pix * N->nfeat, cnt, nullptr);
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
        // stream-ordered before the GEMM (the context stream does not wait for the null
        // stream; a late zero-fill would wipe split-K partial sums)
        hipStream_t s = ctx->stream;
        hipError_t e = hipMemcpyAsync(da, ha.data(), ha.size() * 2, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(db, hb.data(), hb.size() * 2, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemsetAsync(dc, 0, (size_t)M * Nn * 4, s);
        const int lda = a_form == kMK ? K : M, ldb = b_form == kNK ? K : Nn;
        if (e == hipSuccess) {
            gemm(&tmp, a_form, b_form, k_split > 0 ? kAtomicF32 : kStoreF32, da, lda, db, ldb, dc, Nn, 0, M,
                 Nn, K, nullptr, false, k_split);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(c, dc, (size_t)M * Nn * 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = fail(BOLT_EDEVICE, "gemm check: %s", hipGetErrorString(e));
    }
    (void)hipFree(da);
    (void)hipFree(db);
    (void)hipFree(dc);
    return rc;
}
