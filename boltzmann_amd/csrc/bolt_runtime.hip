// libbolt_hip.so: the C ABI of include/bolt.h on HIP for gfx950.
//
// Replaces the reference's OpenCL device layer (src/opencl.rs: device selection, context,
// in-order queue, buffers, kernel registry) and the per-step kernel dispatch of
// Simulation::iterate (src/simulation.rs:297-433).
//
// Device layout (private to this file):
//   populations  SoA planes, plane i at f + i * stride, stride = cells rounded up to 64
//   flags        uint8 per cell (the reference's int32 flags use bits 0..3 only)
//   rho, C       one value per cell (float, or double for double storage)
//   vel          SoA, vel + a * cells
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <memory>
#include <string>
#include <vector>
#include <algorithm>
#include <mutex>

#include "../../include/bolt.h"
#include "update.hpp"
#include "update_wide.hpp"  // WideGeom and the launch geometry; kernels in bolt_wide.hip

using namespace bolt;

#include "runtime_internal.hpp"

// ---------------------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------------------
static thread_local std::string g_last_error;

int bolt_fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

const char* bolt_error_text() { return g_last_error.c_str(); }

struct Tracer {
    int q = 0;
    float omega = 0.f;
    void* g = nullptr;
    void* C = nullptr;
};

struct Coupling {
    int tracer = 0;
    float alpha = 0.f, c0 = 0.f;
    float g[3] = {0, 0, 0};
};

struct bolt_lattice {
    bolt_ctx* ctx = nullptr;
    int ndim = 0;
    int n[3] = {1, 1, 1};
    long long ncell = 0;
    long long stride = 0;
    int q = 0;
    int collision = 0;
    float om_p = 0.f, om_n = 0.f;
    int storage = 0;
    void* f = nullptr;
    void* rho = nullptr;
    void* vel = nullptr;
    uint8_t* flags = nullptr;
    std::vector<Tracer> tracers;
    bool has_gravity = false;
    float gravity[3] = {0, 0, 0};
    std::vector<Coupling> couplings;
    int block = 0;
    int vw = 0;                  // cells per lane of the fluid step (0: chosen per lattice)
    int row_mode = BOLT_ROW_AUTO;  // even steps row-staged (lbm_step_row): BOLT_ROW_*
    int cache_policy = BOLT_CACHE_AUTO;  // population load form: BOLT_CACHE_*
    double* partials = nullptr;  // reduction scratch
    double* result = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint64_t device_bytes = 0;
    double* vel_snap = nullptr;  // velocity snapshot for the steady-state residual
};

static size_t store_size(int s) {
    switch (s) {
        case kF32:
        case kF32F: return 4;
        case kF16S:
        case kBF16S: return 2;
        case kF64: return 8;
    }
    return 0;
}
static bool is_f64(const bolt_lattice* L) { return L->storage == kF64; }
static size_t macro_size(const bolt_lattice* L) { return is_f64(L) ? 8 : 4; }
static int tracer_storage(const bolt_lattice* L) { return is_f64(L) ? kF64 : kF32; }

// The population load form.  MI355X's 256 MB Infinity Cache sits in front of HBM.  A step's
// working set that fits it is loaded with ordinary loads, so the lines stay for the next step;
// measured 1.1-1.4x the streaming form on working sets of 150-240 MB (D2Q9 and D3Q27, float
// and FP16-shifted).  A larger one streams: at 268-319 MB ordinary loads gained up to 4 % or
// lost up to 8 %, and at 1.8 GB they lost 4 %.
static constexpr uint64_t kCacheFitBytes = 240ull << 20;

static bool loads_stream(int policy, uint64_t working_set) {
    if (policy == BOLT_CACHE_STREAM) return true;
    if (policy == BOLT_CACHE_KEEP) return false;
    return working_set > kCacheFitBytes;
}

// Bytes one step of the lattice touches: its population planes (fluid and tracers), the
// flags, every tracer's value and, when a tracer runs as a launch of its own, the density
// and velocity the fluid step then writes every step.
static uint64_t step_working_set(const bolt_lattice* L) {
    const uint64_t nc = (uint64_t)L->ncell;
    const uint64_t ms = macro_size(L);
    uint64_t b = nc * ((uint64_t)L->q * store_size(L->storage) + 1);
    for (const Tracer& t : L->tracers) b += nc * ((uint64_t)t.q * store_size(tracer_storage(L)) + ms);
    const bool can_fuse = (L->storage == kF32 || L->storage == kF64) && !L->tracers.empty();
    if (L->tracers.size() > (can_fuse ? 1u : 0u)) b += nc * (uint64_t)(1 + L->ndim) * ms;
    return b;
}

// ---------------------------------------------------------------------------------------
// kernel table
// ---------------------------------------------------------------------------------------
using StepFn = void (*)(StepParams);

// Every kernel comes in two population load forms (update.hpp pop_load): streaming
// (`stream` true) or ordinary loads.
template <int D, int Q, int C, int S, int TQ, bool ST>
static StepFn pick_even(bool even) {
    return even ? (StepFn)lbm_step<D, Q, C, S, true, TQ, ST> : (StepFn)lbm_step<D, Q, C, S, false, TQ, ST>;
}

template <int D, int Q, int C, int S, bool ST>
static StepFn step_fn(bool even, int tq) {
    if (tq == 0) return pick_even<D, Q, C, S, 0, ST>(even);
    if constexpr (S == kF32 || S == kF64) {
        constexpr int TQ = D == 2 ? 5 : 7;
        if (tq == TQ) return pick_even<D, Q, C, S, TQ, ST>(even);
    }
    return nullptr;
}

template <int D, int Q, int C, bool ST>
static StepFn step_fn_s(int s, bool even, int tq) {
    switch (s) {
        case kF32: return step_fn<D, Q, C, kF32, ST>(even, tq);
        case kF16S: return step_fn<D, Q, C, kF16S, ST>(even, tq);
        case kBF16S: return step_fn<D, Q, C, kBF16S, ST>(even, tq);
        case kF64: return step_fn<D, Q, C, kF64, ST>(even, tq);
        case kF32F: return step_fn<D, Q, C, kF32F, ST>(even, tq);
    }
    return nullptr;
}

template <bool ST>
static StepFn step_table(int d, int q, int coll, int s, bool even, int tq) {
    if (d == 2 && q == 9)
        return coll == 0 ? step_fn_s<2, 9, 0, ST>(s, even, tq) : step_fn_s<2, 9, 1, ST>(s, even, tq);
    if (d == 3 && q == 27)
        return coll == 0 ? step_fn_s<3, 27, 0, ST>(s, even, tq) : step_fn_s<3, 27, 1, ST>(s, even, tq);
    return nullptr;
}

static StepFn lookup_step(int d, int q, int coll, int s, bool even, int tq, bool stream) {
    return stream ? step_table<true>(d, q, coll, s, even, tq) : step_table<false>(d, q, coll, s, even, tq);
}

// wide fluid steps (update_wide.hpp, instantiated in bolt_wide.hip and bolt_wide_keep.hip)
using WideFn = void (*)(StepParams, WideGeom);
WideFn bolt_lookup_wide(int d, int q, int coll, int s, int vw, bool even, bool stream);

// row-staged even steps (update.hpp lbm_step_row, instantiated in bolt_row.hip)
StepFn bolt_lookup_row(int d, int q, int coll, int s, int tq, int block, bool stream);
int bolt_row_slots(int d, int q);

template <int D, int TQ, int S, bool ST>
static StepFn ad_fn(bool even) {
    return even ? (StepFn)ad_step<D, TQ, S, true, ST> : (StepFn)ad_step<D, TQ, S, false, ST>;
}

template <bool ST>
static StepFn ad_table(int d, int tq, int s, bool even) {
    if (d == 2 && tq == 5) return s == kF64 ? ad_fn<2, 5, kF64, ST>(even) : ad_fn<2, 5, kF32, ST>(even);
    if (d == 3 && tq == 7) return s == kF64 ? ad_fn<3, 7, kF64, ST>(even) : ad_fn<3, 7, kF32, ST>(even);
    return nullptr;
}

static StepFn lookup_ad(int d, int tq, int s, bool even, bool stream) {
    return stream ? ad_table<true>(d, tq, s, even) : ad_table<false>(d, tq, s, even);
}

// ---------------------------------------------------------------------------------------
// layout / conversion kernels (host AoS <-> device SoA)
// ---------------------------------------------------------------------------------------
// weights of slot i for the shift of the reduced-precision formats
template <int D, int Q>
__device__ float weight_of(int i) {
    float w = 0.f;
#pragma unroll
    for (int k = 0; k < Q; ++k)
        if (k == i) w = VSet<D, Q>::w(k);
    return w;
}

__device__ float weight_any(int d, int q, int i) {
    if (d == 2 && q == 9) return weight_of<2, 9>(i);
    if (d == 3 && q == 27) return weight_of<3, 27>(i);
    if (d == 2 && q == 5) return weight_of<2, 5>(i);
    if (d == 3 && q == 7) return weight_of<3, 7>(i);
    return 0.f;
}

template <typename H, int S>
__global__ void k_pop_to_device(const H* __restrict__ src, long long c0, long long nc, int d,
                                int q, void* dst, long long stride) {
    using T = typename Store<S>::T;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nc * q) return;
    const long long cl = k / q;
    const int i = (int)(k - cl * q);
    const long long cell = c0 + cl;
    T* out = (T*)dst + (long long)i * stride + cell;
    const double v = (double)src[k];
    if constexpr (S == kF32 || S == kF32F) *out = (float)v;
    else if constexpr (S == kF64) *out = v;
    else if constexpr (S == kF16S) *out = (_Float16)((float)v - weight_any(d, q, i));
    else *out = f32_to_bf16_rne((float)v - weight_any(d, q, i));
}

template <typename H, int S>
__global__ void k_pop_to_host(H* __restrict__ dst, long long c0, long long nc, int d, int q,
                              const void* src, long long stride) {
    using T = typename Store<S>::T;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nc * q) return;
    const long long cl = k / q;
    const int i = (int)(k - cl * q);
    const long long cell = c0 + cl;
    const T* in = (const T*)src + (long long)i * stride + cell;
    double v;
    if constexpr (S == kF32 || S == kF32F || S == kF64) v = (double)*in;
    else if constexpr (S == kF16S) v = (double)((float)*in + weight_any(d, q, i));
    else v = (double)(__uint_as_float(((uint32_t)*in) << 16) + weight_any(d, q, i));
    dst[k] = (H)v;
}

// per-cell vectors [cells, m] host <-> SoA device (m = 1 for rho/C)
template <typename H, typename R>
__global__ void k_vec_to_device(const H* __restrict__ src, long long c0, long long nc, int m,
                                R* dst, long long ncell) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nc * m) return;
    const long long cl = k / m;
    const int a = (int)(k - cl * m);
    dst[(long long)a * ncell + c0 + cl] = (R)src[k];
}

template <typename H, typename R>
__global__ void k_vec_to_host(H* __restrict__ dst, long long c0, long long nc, int m,
                              const R* src, long long ncell) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nc * m) return;
    const long long cl = k / m;
    const int a = (int)(k - cl * m);
    dst[k] = (H)src[(long long)a * ncell + c0 + cl];
}

__global__ void k_flags_to_device(const int32_t* __restrict__ src, long long c0, long long nc,
                                  uint8_t* dst) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nc) return;
    dst[c0 + k] = (uint8_t)(src[k] & 0xff);
}

__global__ void k_flags_to_host(int32_t* __restrict__ dst, long long c0, long long nc,
                                const uint8_t* src) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nc) return;
    dst[k] = (int32_t)src[c0 + k];
}

template <int S>
__global__ void k_fill_zero_pops(void* f, long long stride, long long ncell, int d, int q) {
    using T = typename Store<S>::T;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncell * q) return;
    const int i = (int)(k / ncell);
    const long long cell = k - (long long)i * ncell;
    T* out = (T*)f + (long long)i * stride + cell;
    if constexpr (S == kF16S) *out = (_Float16)(-weight_any(d, q, i));
    else *out = f32_to_bf16_rne(-weight_any(d, q, i));
}

template <typename T>
__global__ void k_fill(T* p, long long n, T v) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

// ---------------------------------------------------------------------------------------
// equilibrium initialisation: velocities.rs:43-52 in float,
//   f_q = rho * w_q * (1 + 3 vq + 4.5 vq vq - 1.5 vv), vq and vv summed from 0.
// ---------------------------------------------------------------------------------------
template <int D, int Q, int S>
__global__ void k_equilibrate(long long ncell, const void* rho_, const void* vel_, void* f_,
                              long long stride) {
#pragma clang fp contract(off)
    using V = VSet<D, Q>;
    using T = typename Store<S>::T;
    const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= ncell) return;
    T* f = (T*)f_;
    if constexpr (S == kF64) {
        const double* rho = (const double*)rho_;
        const double* vel = (const double*)vel_;
        double u[3] = {0, 0, 0};
#pragma unroll
        for (int a = 0; a < D; ++a) u[a] = vel[a * ncell + cell];
        double vv = 0.0;
#pragma unroll
        for (int a = 0; a < D; ++a) vv += u[a] * u[a];
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            double cu = 0.0;
#pragma unroll
            for (int a = 0; a < D; ++a) cu += (double)V::c(i, a) * u[a];
            const double w = (double)V::wn(i) / (double)V::wd(i);
            f[i * stride + cell] = w * rho[cell] * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * vv);
        }
    } else {
        const float* rho = (const float*)rho_;
        const float* vel = (const float*)vel_;
        float u[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int a = 0; a < D; ++a) u[a] = vel[a * ncell + cell];
        float vv = 0.0f;
#pragma unroll
        for (int a = 0; a < D; ++a) vv = vv + u[a] * u[a];
        const float r = rho[cell];
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            float vq = 0.0f;
#pragma unroll
            for (int a = 0; a < D; ++a) vq = vq + u[a] * (float)V::c(i, a);
            const float w = V::w(i);
            if constexpr (S == kF32 || S == kF32F) {
                f[i * stride + cell] = r * w * (1.0f + 3.0f * vq + 4.5f * vq * vq - 1.5f * vv);
            } else {
                // deviation w (rho - 1 + rho (3 vq + 4.5 vq^2 - 1.5 vv)), never forming f
                const float x = 3.0f * vq + 4.5f * vq * vq - 1.5f * vv;
                const float dv = w * ((r - 1.0f) + r * x);
                if constexpr (S == kF16S) f[i * stride + cell] = (_Float16)dv;
                else f[i * stride + cell] = f32_to_bf16_rne(dv);
            }
        }
    }
}

template <int D, int Q, int S>
static void launch_equilibrate(bolt_lattice* L, const void* rho, void* f, long long stride) {
    const long long nb = (L->ncell + 255) / 256;
    hipLaunchKernelGGL((k_equilibrate<D, Q, S>), dim3((unsigned)nb), dim3(256), 0, L->ctx->stream,
                       L->ncell, rho, L->vel, f, stride);
}

// ---------------------------------------------------------------------------------------
// reductions (double, deterministic: per-block partials then one block in fixed order)
// ---------------------------------------------------------------------------------------
constexpr int kRedBlocks = 1024;

__device__ inline double block_sum(double v) {
    __shared__ double sh[256];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    return sh[0];
}

template <int S>
__global__ void k_sum_pops(const void* f_, long long stride, long long ncell, int d, int q,
                           double* partials) {
    using T = typename Store<S>::T;
    const T* f = (const T*)f_;
    double acc = 0.0;
    const long long total = ncell * q;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < total;
         k += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(k / ncell);
        const long long cell = k - (long long)i * ncell;
        const T v = f[(long long)i * stride + cell];
        if constexpr (S == kF16S) acc += (double)((float)v) + (double)weight_any(d, q, i);
        else if constexpr (S == kBF16S)
            acc += (double)__uint_as_float(((uint32_t)v) << 16) + (double)weight_any(d, q, i);
        else acc += (double)v;
    }
    const double s = block_sum(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

template <typename R>
__global__ void k_sum_momentum(const R* rho, const R* vel, const uint8_t* flags, long long ncell,
                               int a, double* partials) {
    double acc = 0.0;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
         c += (long long)gridDim.x * blockDim.x)
        if (!(flags[c] & 1)) acc += (double)rho[c] * (double)vel[(long long)a * ncell + c];
    const double s = block_sum(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// count of the non-finite values of x[0, n) per block, as a double (exact below 2^53)
template <typename R>
__global__ void k_count_nonfinite(const R* x, long long n, double* partials) {
    double acc = 0.0;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x)
        acc += __builtin_isfinite((double)x[k]) ? 0.0 : 1.0;
    const double s = block_sum(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void k_sum_partials(const double* partials, int n, double* out) {
    double acc = 0.0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) acc += partials[k];
    const double s = block_sum(acc);
    if (threadIdx.x == 0) *out = s;
}

// ---------------------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------------------
static int dev_alloc(bolt_lattice* L, void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess)
        return fail(BOLT_ENOMEM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    L->device_bytes += bytes;
    return BOLT_OK;
}

template <typename T>
static int dev_fill(bolt_lattice* L, void* p, long long n, T v) {
    if (n <= 0) return BOLT_OK;
    const long long nb = (n + 255) / 256;
    hipLaunchKernelGGL((k_fill<T>), dim3((unsigned)nb), dim3(256), 0, L->ctx->stream, (T*)p, n, v);
    HIP_TRY(hipGetLastError());
    return BOLT_OK;
}

static long long round_up(long long x, long long m) { return (x + m - 1) / m * m; }

static int pick_block(const bolt_lattice* L) {
    if (L->block > 0) return L->block;
    const int len = L->n[L->ndim - 1];
    const long long rows = L->ncell / len;
    for (int b : {256, 128}) {
        const long long nb = rows * ((len + b - 1) / b);
        if (nb >= 2048 && len >= b) return b;
    }
    return 64;
}

// Cells per lane of the fluid step on even or odd steps.  set_vector_width forces one choice
// for both.  The built-in choice, from the per-launch kernel trace on MI355X (D3Q27 TRT
// 256^3): the odd (own-cell) step of the 2-byte formats runs 2 cells per lane, one 4-byte
// access per plane (f16s 353.5 us against 375.0 us at one cell per lane, bf16s 356.5
// against 378.6); everything else runs one cell per lane, whose even (streaming) step and
// float/double odd step were the fastest (f32 730.1 / 675.6 us against 938.9 / 702.9 at 2
// cells and 802.8 / 706.8 at 4).  A fused tracer always runs one cell per lane.
// Cells per lane of the fluid step.  The built-in choice (vw 0): 2 on the odd steps of the
// 2-byte formats, except for D2Q9 lattices whose population loads keep their lines in the
// Infinity Cache (`stream` false), which run one cell per lane (measured 0.4-8.7 % faster on
// 1024x1024 to 8192x1536 lattices; the D3Q27 ones gained or lost up to 4 % and keep 2);
// one cell per lane otherwise.
static int pick_vw(const bolt_lattice* L, bool fused_tracer, bool even, bool stream) {
    if (fused_tracer) return 1;
    const int len = L->n[L->ndim - 1];
    if (L->vw == 0) {
        const bool half = L->storage == kF16S || L->storage == kBF16S;
        if (half && L->ndim == 2 && !stream) return 1;
        return (half && !even && len % 2 == 0 && len >= 64) ? 2 : 1;
    }
    if (L->vw <= 1) return 1;
    return len % L->vw == 0 ? L->vw : 1;
}

static void fill_geometry(const bolt_lattice* L, StepParams& p, int block) {
    p.n0 = L->n[0];
    p.n1 = L->n[1];
    p.n2 = L->ndim == 3 ? L->n[2] : 1;
    p.len = L->n[L->ndim - 1];
    p.cpr = (p.len + block - 1) / block;
    const long long rows = L->ncell / p.len;
    p.nblocks = (int)(rows * p.cpr);
    p.flags = L->flags;
}

// plane i of a population array starts i * stride elements in
static void fill_planes(const bolt_lattice* L, void* base, int q, int storage, void** planes) {
    const size_t es = store_size(storage);
    for (int i = 0; i < q; ++i) planes[i] = (char*)base + (size_t)i * L->stride * es;
}

static void fill_vel(const bolt_lattice* L, StepParams& p) {
    for (int a = 0; a < 3; ++a)
        p.vel[a] = a < L->ndim ? (char*)L->vel + (size_t)a * L->ncell * macro_size(L) : nullptr;
}

// ---------------------------------------------------------------------------------------
// C ABI (C linkage comes from the declarations in bolt.h)
// ---------------------------------------------------------------------------------------

const char* bolt_last_error(void) { return g_last_error.c_str(); }

const char* bolt_version(void) { return "bolt-hip 0.1 gfx950"; }

static int scratch_setup(int device);

int bolt_ctx_create(int device_kind, int device_index, bolt_ctx** out) {
    if (!out) return fail(BOLT_EINVAL, "out is null");
    *out = nullptr;
    if (device_kind != BOLT_DEVICE_GPU)
        return fail(BOLT_ENOTFOUND,
                    "no CPU device: this library runs the lattice update on HIP GPUs only");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device_index < 0 || device_index >= n)
        return fail(BOLT_ENOTFOUND, "no HIP device %d (found %d)", device_index, n);
    auto c = std::make_unique<bolt_ctx>();
    c->device = device_index;
    HIP_TRY(hipSetDevice(device_index));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device_index));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(BOLT_ENOTFOUND, "device %d is %s; this library is built for gfx950",
                    device_index, prop.gcnArchName);
    HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    int rc = scratch_setup(device_index);  // the read-back staging area on this GPU
    if (rc) {
        (void)hipStreamDestroy(c->stream);
        return rc;
    }
    c->info = std::string("AMD GPU ") + prop.name + " (" + prop.gcnArchName + ")";
    *out = c.release();
    return BOLT_OK;
}

int bolt_ctx_destroy(bolt_ctx* ctx) {
    if (!ctx) return BOLT_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
    return BOLT_OK;
}

int bolt_ctx_device_info(const bolt_ctx* ctx, char* buf, size_t len) {
    if (!ctx || !buf || len == 0) return fail(BOLT_EINVAL, "bad arguments");
    snprintf(buf, len, "%s", ctx->info.c_str());
    return BOLT_OK;
}

int bolt_lattice_destroy(bolt_lattice* L) {
    if (!L) return BOLT_OK;
    (void)hipSetDevice(L->ctx->device);
    (void)hipStreamSynchronize(L->ctx->stream);
    (void)hipFree(L->f);
    (void)hipFree(L->rho);
    (void)hipFree(L->vel);
    (void)hipFree(L->flags);
    for (auto& t : L->tracers) {
        (void)hipFree(t.g);
        (void)hipFree(t.C);
    }
    (void)hipFree(L->partials);
    (void)hipFree(L->result);
    (void)hipFree(L->vel_snap);
    if (L->ev0) (void)hipEventDestroy(L->ev0);
    if (L->ev1) (void)hipEventDestroy(L->ev1);
    delete L;
    return BOLT_OK;
}

int bolt_lattice_create(bolt_ctx* ctx, int ndim, const int64_t* counts, int q, int collision,
                        float omega_pos, float omega_neg, int storage, bolt_lattice** out) {
    if (!ctx || !counts || !out) return fail(BOLT_EINVAL, "null argument");
    *out = nullptr;
    if (ndim != 2 && ndim != 3)
        return fail(BOLT_ENOTFOUND, "Kernel not found: D%d (only 2D and 3D lattices)", ndim);
    if (!((ndim == 2 && q == 9) || (ndim == 3 && q == 27)))
        return fail(BOLT_ENOTFOUND, "Kernel not found: D%dQ%d NavierStokes", ndim, q);
    if (collision != BOLT_BGK && collision != BOLT_TRT)
        return fail(BOLT_EINVAL, "unknown collision %d", collision);
    if (storage < 0 || storage > BOLT_STORE_F32F)
        return fail(BOLT_EINVAL, "unknown storage %d", storage);
    long long ncell = 1;
    for (int a = 0; a < ndim; ++a) {
        if (counts[a] < 1 || counts[a] > (1ll << 30))
            return fail(BOLT_EINVAL, "count[%d] = %lld out of range", a, (long long)counts[a]);
        ncell *= counts[a];
    }
    if (ncell / counts[ndim - 1] > (1ll << 31) / 8)
        return fail(BOLT_EINVAL, "too many rows for one launch");
    // kernels address every array with 32-bit byte offsets per cell
    const long long widest = (storage == BOLT_STORE_F64) ? 8 : 4;
    if (ncell * widest >= (1ll << 32))
        return fail(BOLT_EINVAL,
                    "%lld cells exceed one lattice (at most 2^32 bytes per population plane: "
                    "%lld cells in this storage)", ncell, ((1ll << 32) - 1) / widest);
    auto L = std::make_unique<bolt_lattice>();
    L->ctx = ctx;
    L->ndim = ndim;
    for (int a = 0; a < ndim; ++a) L->n[a] = (int)counts[a];
    L->ncell = ncell;
    L->stride = round_up(ncell, 64);
    L->q = q;
    L->collision = collision;
    L->om_p = omega_pos;
    L->om_n = collision == BOLT_TRT ? omega_neg : omega_pos;
    L->storage = storage;
    HIP_TRY(hipSetDevice(ctx->device));
    int rc;
    if ((rc = dev_alloc(L.get(), &L->f, (size_t)L->stride * q * store_size(storage)))) return rc;
    if ((rc = dev_alloc(L.get(), &L->rho, (size_t)ncell * macro_size(L.get())))) return rc;
    if ((rc = dev_alloc(L.get(), &L->vel, (size_t)ncell * ndim * macro_size(L.get())))) return rc;
    if ((rc = dev_alloc(L.get(), (void**)&L->flags, (size_t)ncell))) return rc;
    if ((rc = dev_alloc(L.get(), (void**)&L->partials, kRedBlocks * sizeof(double)))) return rc;
    if ((rc = dev_alloc(L.get(), (void**)&L->result, 8 * sizeof(double)))) return rc;
    HIP_TRY(hipMemsetAsync(L->f, 0, (size_t)L->stride * q * store_size(storage), ctx->stream));
    if (storage == kF16S || storage == kBF16S) {
        // the reference's default f = 0 is the deviation -w_i in the shifted formats
        const long long n = ncell * q;
        const dim3 grid((unsigned)((n + 255) / 256)), blk(256);
        if (storage == kF16S)
            hipLaunchKernelGGL(k_fill_zero_pops<kF16S>, grid, blk, 0, ctx->stream, L->f, L->stride, ncell, ndim, q);
        else
            hipLaunchKernelGGL(k_fill_zero_pops<kBF16S>, grid, blk, 0, ctx->stream, L->f, L->stride, ncell, ndim, q);
        HIP_TRY(hipGetLastError());
    }
    if (is_f64(L.get())) {
        if ((rc = dev_fill<double>(L.get(), L->rho, ncell, 1.0))) return rc;
    } else {
        if ((rc = dev_fill<float>(L.get(), L->rho, ncell, 1.0f))) return rc;
    }
    HIP_TRY(hipMemsetAsync(L->vel, 0, (size_t)ncell * ndim * macro_size(L.get()), ctx->stream));
    HIP_TRY(hipMemsetAsync(L->flags, 0, (size_t)ncell, ctx->stream));
    HIP_TRY(hipEventCreate(&L->ev0));
    HIP_TRY(hipEventCreate(&L->ev1));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    *out = L.release();
    return BOLT_OK;
}

int bolt_lattice_add_tracer(bolt_lattice* L, int q, float omega, int* tracer_id) {
    if (!L) return fail(BOLT_EINVAL, "null lattice");
    if (!((L->ndim == 2 && q == 5) || (L->ndim == 3 && q == 7)))
        return fail(BOLT_ENOTFOUND, "Kernel not found: D%dQ%d AdvectionDiffusion BGK", L->ndim, q);
    HIP_TRY(hipSetDevice(L->ctx->device));
    Tracer t;
    t.q = q;
    t.omega = omega;
    const size_t es = store_size(tracer_storage(L));
    int rc;
    if ((rc = dev_alloc(L, &t.g, (size_t)L->stride * q * es))) return rc;
    if ((rc = dev_alloc(L, &t.C, (size_t)L->ncell * macro_size(L)))) return rc;
    HIP_TRY(hipMemsetAsync(t.g, 0, (size_t)L->stride * q * es, L->ctx->stream));
    HIP_TRY(hipMemsetAsync(t.C, 0, (size_t)L->ncell * macro_size(L), L->ctx->stream));
    L->tracers.push_back(t);
    if (tracer_id) *tracer_id = (int)L->tracers.size() - 1;
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    return BOLT_OK;
}

int bolt_lattice_set_gravity(bolt_lattice* L, const float* g) {
    if (!L || !g) return fail(BOLT_EINVAL, "null argument");
    L->has_gravity = true;
    for (int a = 0; a < L->ndim; ++a) L->gravity[a] = g[a];
    return BOLT_OK;
}

int bolt_lattice_add_coupling(bolt_lattice* L, int tracer_id, float alpha, float c0,
                              const float* g) {
    if (!L || !g) return fail(BOLT_EINVAL, "null argument");
    if (tracer_id < 0 || tracer_id >= (int)L->tracers.size())
        return fail(BOLT_EINVAL, "unknown tracer %d", tracer_id);
    Coupling c;
    c.tracer = tracer_id;
    c.alpha = alpha;
    c.c0 = c0;
    for (int a = 0; a < L->ndim; ++a) c.g[a] = g[a];
    L->couplings.push_back(c);
    return BOLT_OK;
}

// shape of a field: elements per cell and the device pointer
struct FieldDesc {
    int per_cell;
    void* ptr;
    int kind;  // 0 populations, 1 cell vector, 2 flags
    int q;
    int storage;
};

static int describe(bolt_lattice* L, int field, int tracer_id, FieldDesc& fd) {
    switch (field) {
        case BOLT_FIELD_FLAGS: fd = {1, L->flags, 2, 0, 0}; return BOLT_OK;
        case BOLT_FIELD_F: fd = {L->q, L->f, 0, L->q, L->storage}; return BOLT_OK;
        case BOLT_FIELD_RHO: fd = {1, L->rho, 1, 0, 0}; return BOLT_OK;
        case BOLT_FIELD_VEL: fd = {L->ndim, L->vel, 1, 0, 0}; return BOLT_OK;
        case BOLT_FIELD_G:
        case BOLT_FIELD_C: {
            if (tracer_id < 0 || tracer_id >= (int)L->tracers.size())
                return fail(BOLT_EINVAL, "unknown tracer %d", tracer_id);
            Tracer& t = L->tracers[tracer_id];
            if (field == BOLT_FIELD_G) fd = {t.q, t.g, 0, t.q, tracer_storage(L)};
            else fd = {1, t.C, 1, 0, 0};
            return BOLT_OK;
        }
    }
    return fail(BOLT_EINVAL, "unknown field %d", field);
}

template <typename H>
static void launch_pop_to_device(bolt_lattice* L, const H* src, long long c0, long long nc, int q,
                                 void* dst, int storage) {
    const long long n = nc * q;
    const dim3 grid((unsigned)((n + 255) / 256)), blk(256);
    hipStream_t s = L->ctx->stream;
    switch (storage) {
        case kF32: hipLaunchKernelGGL((k_pop_to_device<H, kF32>), grid, blk, 0, s, src, c0, nc, L->ndim, q, dst, L->stride); break;
        case kF32F: hipLaunchKernelGGL((k_pop_to_device<H, kF32F>), grid, blk, 0, s, src, c0, nc, L->ndim, q, dst, L->stride); break;
        case kF16S: hipLaunchKernelGGL((k_pop_to_device<H, kF16S>), grid, blk, 0, s, src, c0, nc, L->ndim, q, dst, L->stride); break;
        case kBF16S: hipLaunchKernelGGL((k_pop_to_device<H, kBF16S>), grid, blk, 0, s, src, c0, nc, L->ndim, q, dst, L->stride); break;
        case kF64: hipLaunchKernelGGL((k_pop_to_device<H, kF64>), grid, blk, 0, s, src, c0, nc, L->ndim, q, dst, L->stride); break;
    }
}

template <typename H>
static void launch_pop_to_host(bolt_lattice* L, H* dst, long long c0, long long nc, int q,
                               const void* src, int storage) {
    const long long n = nc * q;
    const dim3 grid((unsigned)((n + 255) / 256)), blk(256);
    hipStream_t s = L->ctx->stream;
    switch (storage) {
        case kF32: hipLaunchKernelGGL((k_pop_to_host<H, kF32>), grid, blk, 0, s, dst, c0, nc, L->ndim, q, src, L->stride); break;
        case kF32F: hipLaunchKernelGGL((k_pop_to_host<H, kF32F>), grid, blk, 0, s, dst, c0, nc, L->ndim, q, src, L->stride); break;
        case kF16S: hipLaunchKernelGGL((k_pop_to_host<H, kF16S>), grid, blk, 0, s, dst, c0, nc, L->ndim, q, src, L->stride); break;
        case kBF16S: hipLaunchKernelGGL((k_pop_to_host<H, kBF16S>), grid, blk, 0, s, dst, c0, nc, L->ndim, q, src, L->stride); break;
        case kF64: hipLaunchKernelGGL((k_pop_to_host<H, kF64>), grid, blk, 0, s, dst, c0, nc, L->ndim, q, src, L->stride); break;
    }
}

static size_t host_size(int dt) { return dt == BOLT_DT_F64 ? 8 : 4; }

// Moves a field in chunks of cells through a device staging buffer: one it allocates
// (64 MiB at most) and frees, or, when `fixed_stage` is given, that buffer of
// `fixed_stage_bytes` (nothing is allocated then).
static int transfer(bolt_lattice* L, int field, int tracer_id, int host_dtype, void* host,
                    size_t nbytes, bool to_device, void* fixed_stage = nullptr,
                    size_t fixed_stage_bytes = 0) {
    if (!L || !host) return fail(BOLT_EINVAL, "null argument");
    FieldDesc fd;
    int rc = describe(L, field, tracer_id, fd);
    if (rc) return rc;
    if (fd.kind == 2 && host_dtype != BOLT_DT_I32)
        return fail(BOLT_EINVAL, "flags are int32 on the host");
    if (fd.kind != 2 && host_dtype != BOLT_DT_F32 && host_dtype != BOLT_DT_F64)
        return fail(BOLT_EINVAL, "host dtype must be f32 or f64 for field %d", field);
    const size_t es = host_size(host_dtype);
    const size_t want = (size_t)L->ncell * fd.per_cell * es;
    if (nbytes != want)
        return fail(BOLT_EINVAL, "field %d needs %zu bytes, got %zu", field, want, nbytes);
    HIP_TRY(hipSetDevice(L->ctx->device));
    const size_t stage_bytes = fixed_stage ? fixed_stage_bytes : (64u << 20);
    long long chunk = (long long)(stage_bytes / (fd.per_cell * es));
    if (chunk < 1) return fail(BOLT_EINVAL, "staging area of %zu bytes holds no cell", stage_bytes);
    if (chunk > L->ncell) chunk = L->ncell;
    void* stage = fixed_stage;
    if (!stage) HIP_TRY(hipMalloc(&stage, (size_t)chunk * fd.per_cell * es));
    hipStream_t s = L->ctx->stream;
    const bool f64m = is_f64(L);
    int status = BOLT_OK;
    for (long long c0 = 0; c0 < L->ncell && status == BOLT_OK; c0 += chunk) {
        const long long nc = (L->ncell - c0 < chunk) ? L->ncell - c0 : chunk;
        const size_t bytes = (size_t)nc * fd.per_cell * es;
        char* hp = (char*)host + (size_t)c0 * fd.per_cell * es;
        const dim3 grid((unsigned)((nc * fd.per_cell + 255) / 256)), blk(256);
        hipError_t e = hipSuccess;
        if (to_device) {
            e = hipMemcpyAsync(stage, hp, bytes, hipMemcpyHostToDevice, s);
            if (e == hipSuccess) {
                if (fd.kind == 2) {
                    hipLaunchKernelGGL(k_flags_to_device, grid, blk, 0, s, (const int32_t*)stage, c0, nc, L->flags);
                } else if (fd.kind == 0) {
                    if (host_dtype == BOLT_DT_F32) launch_pop_to_device<float>(L, (const float*)stage, c0, nc, fd.q, fd.ptr, fd.storage);
                    else launch_pop_to_device<double>(L, (const double*)stage, c0, nc, fd.q, fd.ptr, fd.storage);
                } else {
                    if (host_dtype == BOLT_DT_F32) {
                        if (f64m) hipLaunchKernelGGL((k_vec_to_device<float, double>), grid, blk, 0, s, (const float*)stage, c0, nc, fd.per_cell, (double*)fd.ptr, L->ncell);
                        else hipLaunchKernelGGL((k_vec_to_device<float, float>), grid, blk, 0, s, (const float*)stage, c0, nc, fd.per_cell, (float*)fd.ptr, L->ncell);
                    } else {
                        if (f64m) hipLaunchKernelGGL((k_vec_to_device<double, double>), grid, blk, 0, s, (const double*)stage, c0, nc, fd.per_cell, (double*)fd.ptr, L->ncell);
                        else hipLaunchKernelGGL((k_vec_to_device<double, float>), grid, blk, 0, s, (const double*)stage, c0, nc, fd.per_cell, (float*)fd.ptr, L->ncell);
                    }
                }
                e = hipGetLastError();
            }
            if (e == hipSuccess) e = hipStreamSynchronize(s);
        } else {
            if (fd.kind == 2) {
                hipLaunchKernelGGL(k_flags_to_host, grid, blk, 0, s, (int32_t*)stage, c0, nc, L->flags);
            } else if (fd.kind == 0) {
                if (host_dtype == BOLT_DT_F32) launch_pop_to_host<float>(L, (float*)stage, c0, nc, fd.q, fd.ptr, fd.storage);
                else launch_pop_to_host<double>(L, (double*)stage, c0, nc, fd.q, fd.ptr, fd.storage);
            } else {
                if (host_dtype == BOLT_DT_F32) {
                    if (f64m) hipLaunchKernelGGL((k_vec_to_host<float, double>), grid, blk, 0, s, (float*)stage, c0, nc, fd.per_cell, (const double*)fd.ptr, L->ncell);
                    else hipLaunchKernelGGL((k_vec_to_host<float, float>), grid, blk, 0, s, (float*)stage, c0, nc, fd.per_cell, (const float*)fd.ptr, L->ncell);
                } else {
                    if (f64m) hipLaunchKernelGGL((k_vec_to_host<double, double>), grid, blk, 0, s, (double*)stage, c0, nc, fd.per_cell, (const double*)fd.ptr, L->ncell);
                    else hipLaunchKernelGGL((k_vec_to_host<double, float>), grid, blk, 0, s, (double*)stage, c0, nc, fd.per_cell, (const float*)fd.ptr, L->ncell);
                }
            }
            e = hipGetLastError();
            if (e == hipSuccess) e = hipMemcpyAsync(hp, stage, bytes, hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
        }
        if (e != hipSuccess)
            status = fail(BOLT_EDEVICE, "field transfer failed: %s", hipGetErrorString(e));
    }
    (void)hipStreamSynchronize(s);
    if (!fixed_stage) (void)hipFree(stage);
    return status;
}

int bolt_lattice_write(bolt_lattice* L, int field, int tracer_id, int host_dtype,
                       const void* host, size_t nbytes) {
    return transfer(L, field, tracer_id, host_dtype, (void*)host, nbytes, true);
}

int bolt_lattice_read(bolt_lattice* L, int field, int tracer_id, int host_dtype, void* host,
                      size_t nbytes) {
    return transfer(L, field, tracer_id, host_dtype, host, nbytes, false);
}

// ---------------------------------------------------------------------------------------
// Allocation-free read-back.  `scratch` is the staging area of bolt_lattice_read_noalloc:
// every GPU the library runs on holds its own copy for the life of the process.  In host
// code the array's name is only a stand-in; its device address on each GPU comes from
// hipGetSymbolAddress, taken once per device when a context is created on it
// (bolt_ctx_create) and kept in g_scratch_dev.  All calls on one GPU share that GPU's
// array: calls that overlap on one GPU (two threads, two contexts) race on it.
// ---------------------------------------------------------------------------------------
static constexpr size_t kNoAllocScratchBytes = 16u << 20;  // 16 MiB per GPU
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static constexpr int kMaxDevices = 64;
static void* g_scratch_dev[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// Resolves the device address of `scratch` on `device`, which must be the current device.
static int scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(BOLT_EINVAL, "device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[device]) return BOLT_OK;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(BOLT_EDEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                    hipGetErrorString(e));
    g_scratch_dev[device] = ptr;
    return BOLT_OK;
}

int bolt_lattice_read_noalloc(bolt_lattice* L, int field, int tracer_id, int host_dtype,
                              void* host, size_t nbytes) {
    if (!L || !host) return fail(BOLT_EINVAL, "null argument");
    const int dev = L->ctx->device;
    void* ptr = (dev >= 0 && dev < kMaxDevices) ? g_scratch_dev[dev] : nullptr;
    if (!ptr) return fail(BOLT_ESTATE, "no staging area on device %d (contexts set it up)", dev);
    return transfer(L, field, tracer_id, host_dtype, host, nbytes, false, ptr, kNoAllocScratchBytes);
}

int bolt_lattice_equilibrate(bolt_lattice* L) {
    if (!L) return fail(BOLT_EINVAL, "null lattice");
    HIP_TRY(hipSetDevice(L->ctx->device));
    const int s = L->storage;
#define EQ(D, Q, S) launch_equilibrate<D, Q, S>(L, L->rho, L->f, L->stride)
    if (L->ndim == 2) {
        if (s == kF32 || s == kF32F) EQ(2, 9, kF32);
        else if (s == kF16S) EQ(2, 9, kF16S);
        else if (s == kBF16S) EQ(2, 9, kBF16S);
        else EQ(2, 9, kF64);
    } else {
        if (s == kF32 || s == kF32F) EQ(3, 27, kF32);
        else if (s == kF16S) EQ(3, 27, kF16S);
        else if (s == kBF16S) EQ(3, 27, kBF16S);
        else EQ(3, 27, kF64);
    }
#undef EQ
    for (auto& t : L->tracers) {
        if (L->ndim == 2) {
            if (is_f64(L)) launch_equilibrate<2, 5, kF64>(L, t.C, t.g, L->stride);
            else launch_equilibrate<2, 5, kF32>(L, t.C, t.g, L->stride);
        } else {
            if (is_f64(L)) launch_equilibrate<3, 7, kF64>(L, t.C, t.g, L->stride);
            else launch_equilibrate<3, 7, kF32>(L, t.C, t.g, L->stride);
        }
    }
    HIP_TRY(hipGetLastError());
    return BOLT_OK;
}

static int enqueue_steps(bolt_lattice* L, uint64_t n) {
    if (n % 2 != 0) return fail(BOLT_EINVAL, "iters must be even");
    HIP_TRY(hipSetDevice(L->ctx->device));
    const int block = pick_block(L);
    StepParams p{};
    fill_geometry(L, p, block);
    fill_planes(L, L->f, L->q, L->storage, p.fpl);
    p.rho = L->rho;
    fill_vel(L, p);
    p.om_p = L->om_p;
    p.om_n = L->om_n;
    // acceleration: the last coupling wins and overrides gravity (simulation.rs:326-373)
    p.acc_mode = kAccNone;
    if (!L->couplings.empty()) {
        const Coupling& c = L->couplings.back();
        p.acc_mode = kAccBoussinesq;
        for (int a = 0; a < 3; ++a) p.g[a] = c.g[a];
        p.alpha = c.alpha;
        p.c0 = c.c0;
        p.bq_C = L->tracers[c.tracer].C;
    } else if (L->has_gravity) {
        p.acc_mode = kAccConst;
        for (int a = 0; a < 3; ++a) p.g[a] = L->gravity[a];
    }
    // the first tracer is fused into the fluid step when the storage allows it
    const bool can_fuse = (L->storage == kF32 || L->storage == kF64) && !L->tracers.empty();
    const int tq = can_fuse ? L->tracers[0].q : 0;
    if (can_fuse) {
        fill_planes(L, L->tracers[0].g, L->tracers[0].q, tracer_storage(L), p.gpl);
        p.tC = L->tracers[0].C;
        p.t_om = L->tracers[0].omega;
    }
    const size_t first_sep = can_fuse ? 1 : 0;
    const bool sep_tracers = L->tracers.size() > first_sep;
    const bool st = loads_stream(L->cache_policy, step_working_set(L));
    StepFn fe = lookup_step(L->ndim, L->q, L->collision, L->storage, true, tq, st);
    StepFn fo = lookup_step(L->ndim, L->q, L->collision, L->storage, false, tq, st);
    if (!fe || !fo) return fail(BOLT_ENOTFOUND, "Kernel not found: D%dQ%d collision %d storage %d",
                                L->ndim, L->q, L->collision, L->storage);
    const int vwe = pick_vw(L, can_fuse, true, st), vwo = pick_vw(L, can_fuse, false, st);
    const WideGeom wge = make_wide_geom(L->ncell, p.len, vwe), wgo = make_wide_geom(L->ncell, p.len, vwo);
    WideFn we = nullptr, wo = nullptr;
    if (vwe > 1 && !(we = bolt_lookup_wide(L->ndim, L->q, L->collision, L->storage, vwe, true, st)))
        return fail(BOLT_ENOTFOUND, "Kernel not found: %d cells per lane, storage %d", vwe, L->storage);
    if (vwo > 1 && !(wo = bolt_lookup_wide(L->ndim, L->q, L->collision, L->storage, vwo, false, st)))
        return fail(BOLT_ENOTFOUND, "Kernel not found: %d cells per lane, storage %d", vwo, L->storage);
    // the even step row-staged: one work-group per row of the contiguous axis, which must
    // fit one work-group (1024 cells) and its shifted planes' rows (fluid and fused tracer)
    // 64 KiB of LDS
    StepFn fr = nullptr;
    const int row_block = (p.len + 63) / 64 * 64;
    const size_t row_lds = (size_t)(bolt_row_slots(L->ndim, L->q) + bolt_row_slots(L->ndim, tq)) * row_block *
                           store_size(L->storage);
    // BOLT_ROW_AUTO takes it where it measured faster on MI355X (D2Q9 and D3Q27 TRT): rows
    // that fill whole 128-byte cache lines, so that no line is shared between work-groups
    // (+0.2 to +10.6 %; rows of 1000 float cells lost 2 %), except D3Q27 rows over 512 cells
    // in the 2-byte formats, whose 1024-work-item groups fit one per CU (-11 %)
    const bool two_byte = L->storage == kF16S || L->storage == kBF16S;
    // nor D2Q9 lattices that fit the Infinity Cache, whose direct even step measured as fast
    // in float storage and 0.2-6.4 % faster in the 2-byte formats and with a fused tracer
    const bool row_pays = (size_t)p.len * store_size(L->storage) % 128 == 0 &&
                          !(L->ndim == 3 && two_byte && row_block > 512) && !(L->ndim == 2 && !st);
    const bool row_wanted = L->row_mode == BOLT_ROW_ON || (L->row_mode == BOLT_ROW_AUTO && row_pays);
    if (row_wanted && vwe == 1 && row_block <= 1024 && row_lds <= 65536)
        fr = bolt_lookup_row(L->ndim, L->q, L->collision, L->storage, tq, row_block, st);
    const long long rows = L->ncell / p.len;
    const dim3 grid((unsigned)p.nblocks), blk((unsigned)block);
    const dim3 wgride((unsigned)wge.nblocks), wgrido((unsigned)wgo.nblocks), wblk(256);
    const dim3 rgrid((unsigned)rows), rblk((unsigned)row_block);
    hipStream_t s = L->ctx->stream;
    for (uint64_t it = 0; it < n; ++it) {
        const bool even = (it % 2) == 0;
        p.write_macro = (it + 1 == n) || sep_tracers;
        // a lattice larger than the Infinity Cache walks its odd steps' tiles last to first,
        // so that each step starts on the lines the previous one left in that cache
        // (measured: +0.6 to +1.9 % f32 and +3.8 to +6 % f16s on D3Q27 256^3); a lattice that
        // fits it keeps one order (reversing cost the cache-resident D2Q9 lattices 1-4 %)
        p.reverse = (st && !even) ? 1 : 0;
        if (even && fr) {
            StepParams pr = p;
            pr.nblocks = (int)rows;
            hipLaunchKernelGGL(fr, rgrid, rblk, row_lds, s, pr);
        } else if (even && vwe > 1) hipLaunchKernelGGL(we, wgride, wblk, 0, s, p, wge);
        else if (!even && vwo > 1) hipLaunchKernelGGL(wo, wgrido, wblk, 0, s, p, wgo);
        else hipLaunchKernelGGL(even ? fe : fo, grid, blk, 0, s, p);
        for (size_t k = first_sep; k < L->tracers.size(); ++k) {
            const Tracer& t = L->tracers[k];
            StepParams pt = p;
            fill_planes(L, t.g, t.q, tracer_storage(L), pt.gpl);
            pt.tC = t.C;
            pt.t_om = t.omega;
            StepFn fa = lookup_ad(L->ndim, t.q, tracer_storage(L), even, st);
            if (!fa) return fail(BOLT_ENOTFOUND, "Kernel not found: D%dQ%d AdvectionDiffusion", L->ndim, t.q);
            hipLaunchKernelGGL(fa, grid, blk, 0, s, pt);
        }
    }
    HIP_TRY(hipGetLastError());
    return BOLT_OK;
}

int bolt_lattice_step(bolt_lattice* L, uint64_t n) {
    if (!L) return fail(BOLT_EINVAL, "null lattice");
    return enqueue_steps(L, n);
}

int bolt_lattice_step_timed(bolt_lattice* L, uint64_t n, double* ms) {
    if (!L) return fail(BOLT_EINVAL, "null lattice");
    HIP_TRY(hipSetDevice(L->ctx->device));
    HIP_TRY(hipEventRecord(L->ev0, L->ctx->stream));
    int rc = enqueue_steps(L, n);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(L->ev1, L->ctx->stream));
    HIP_TRY(hipEventSynchronize(L->ev1));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, L->ev0, L->ev1));
    if (ms) *ms = t;
    return BOLT_OK;
}

int bolt_lattice_sync(bolt_lattice* L) {
    if (!L) return fail(BOLT_EINVAL, "null lattice");
    HIP_TRY(hipSetDevice(L->ctx->device));
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    return BOLT_OK;
}

static int finish_reduce(bolt_lattice* L, double* out) {
    hipStream_t s = L->ctx->stream;
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, L->partials, kRedBlocks, L->result);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, L->result, sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return BOLT_OK;
}

int bolt_host_alloc(size_t bytes, void** out) {
    if (!out) return fail(BOLT_EINVAL, "null argument");
    *out = nullptr;
    if (bytes == 0) return BOLT_OK;
    HIP_TRY(hipHostMalloc(out, bytes, hipHostMallocDefault));
    return BOLT_OK;
}

int bolt_host_free(void* p) {
    if (p) HIP_TRY(hipHostFree(p));
    return BOLT_OK;
}

int bolt_lattice_count_nonfinite(bolt_lattice* L, int field, uint64_t* count) {
    if (!L || !count) return fail(BOLT_EINVAL, "null argument");
    if (field != BOLT_FIELD_RHO && field != BOLT_FIELD_VEL)
        return fail(BOLT_EINVAL, "count_nonfinite reads BOLT_FIELD_RHO or BOLT_FIELD_VEL");
    HIP_TRY(hipSetDevice(L->ctx->device));
    const long long n = (long long)L->ncell * (field == BOLT_FIELD_VEL ? L->ndim : 1);
    const void* x = field == BOLT_FIELD_VEL ? L->vel : L->rho;
    hipStream_t s = L->ctx->stream;
    if (is_f64(L))
        hipLaunchKernelGGL(k_count_nonfinite<double>, dim3(kRedBlocks), dim3(256), 0, s, (const double*)x, n, L->partials);
    else
        hipLaunchKernelGGL(k_count_nonfinite<float>, dim3(kRedBlocks), dim3(256), 0, s, (const float*)x, n, L->partials);
    double r = 0.0;
    const int rc = finish_reduce(L, &r);
    if (rc) return rc;
    *count = (uint64_t)r;
    return BOLT_OK;
}

int bolt_lattice_reduce(bolt_lattice* L, int what, int tracer_id, double* out) {
    if (!L || !out) return fail(BOLT_EINVAL, "null argument");
    HIP_TRY(hipSetDevice(L->ctx->device));
    hipStream_t s = L->ctx->stream;
    const dim3 grid(kRedBlocks), blk(256);
    if (what == BOLT_REDUCE_MASS || what == BOLT_REDUCE_TRACER) {
        const void* f = L->f;
        int q = L->q, st = L->storage;
        if (what == BOLT_REDUCE_TRACER) {
            if (tracer_id < 0 || tracer_id >= (int)L->tracers.size())
                return fail(BOLT_EINVAL, "unknown tracer %d", tracer_id);
            f = L->tracers[tracer_id].g;
            q = L->tracers[tracer_id].q;
            st = tracer_storage(L);
        }
        switch (st) {
            case kF32:
            case kF32F: hipLaunchKernelGGL(k_sum_pops<kF32>, grid, blk, 0, s, f, L->stride, L->ncell, L->ndim, q, L->partials); break;
            case kF16S: hipLaunchKernelGGL(k_sum_pops<kF16S>, grid, blk, 0, s, f, L->stride, L->ncell, L->ndim, q, L->partials); break;
            case kBF16S: hipLaunchKernelGGL(k_sum_pops<kBF16S>, grid, blk, 0, s, f, L->stride, L->ncell, L->ndim, q, L->partials); break;
            case kF64: hipLaunchKernelGGL(k_sum_pops<kF64>, grid, blk, 0, s, f, L->stride, L->ncell, L->ndim, q, L->partials); break;
        }
        HIP_TRY(hipGetLastError());
        return finish_reduce(L, out);
    }
    if (what == BOLT_REDUCE_MOMENTUM) {
        for (int a = 0; a < L->ndim; ++a) {
            if (is_f64(L))
                hipLaunchKernelGGL(k_sum_momentum<double>, grid, blk, 0, s, (const double*)L->rho, (const double*)L->vel, L->flags, L->ncell, a, L->partials);
            else
                hipLaunchKernelGGL(k_sum_momentum<float>, grid, blk, 0, s, (const float*)L->rho, (const float*)L->vel, L->flags, L->ncell, a, L->partials);
            HIP_TRY(hipGetLastError());
            int rc = finish_reduce(L, out + a);
            if (rc) return rc;
        }
        return BOLT_OK;
    }
    return fail(BOLT_EINVAL, "unknown reduction %d", what);
}

int bolt_lattice_device_bytes(const bolt_lattice* L, uint64_t* out) {
    if (!L || !out) return fail(BOLT_EINVAL, "null argument");
    *out = L->device_bytes;
    return BOLT_OK;
}

int bolt_lattice_set_vector_width(bolt_lattice* L, int vw) {
    if (!L) return fail(BOLT_EINVAL, "null lattice");
    if (vw != 0 && vw != 1 && vw != 2 && vw != 4)
        return fail(BOLT_EINVAL, "cells per lane must be 0 (chosen per lattice), 1, 2 or 4");
    if (vw == 4 && L->storage == kF64) return fail(BOLT_EINVAL, "double storage takes 1 or 2 cells per lane");
    if (vw > 1 && L->n[L->ndim - 1] % vw != 0)
        return fail(BOLT_EINVAL, "the contiguous axis (%d cells) is not a multiple of %d", L->n[L->ndim - 1], vw);
    L->vw = vw;
    return BOLT_OK;
}

int bolt_lattice_set_row_staging(bolt_lattice* L, int mode) {
    if (!L) return fail(BOLT_EINVAL, "null lattice");
    if (mode != BOLT_ROW_OFF && mode != BOLT_ROW_ON && mode != BOLT_ROW_AUTO)
        return fail(BOLT_EINVAL, "row staging mode must be BOLT_ROW_OFF, _ON or _AUTO");
    L->row_mode = mode;
    return BOLT_OK;
}

static bool valid_cache_policy(int policy) {
    return policy == BOLT_CACHE_STREAM || policy == BOLT_CACHE_KEEP || policy == BOLT_CACHE_AUTO;
}

int bolt_lattice_set_cache_policy(bolt_lattice* L, int policy) {
    if (!L) return fail(BOLT_EINVAL, "null lattice");
    if (!valid_cache_policy(policy))
        return fail(BOLT_EINVAL, "cache policy must be BOLT_CACHE_STREAM, _KEEP or _AUTO");
    L->cache_policy = policy;
    return BOLT_OK;
}

int bolt_lattice_set_block(bolt_lattice* L, int block) {
    if (!L) return fail(BOLT_EINVAL, "null lattice");
    if (block != 0 && (block < 64 || block > 256 || block % 64 != 0))
        return fail(BOLT_EINVAL, "block must be 64, 128, 192 or 256 (or 0)");
    L->block = block;
    return BOLT_OK;
}

// ---------------------------------------------------------------------------------------
// Steady-state residual, double precision:
//   r = max_x |u - u_snap| / (k * max_x |u|)   over non-wall cells, |.| the Euclidean norm,
// where u_snap is the velocity saved by the previous call and k the steps between them.
// ---------------------------------------------------------------------------------------
__device__ inline double block_max(double v) {
    __shared__ double sh[256];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    return sh[0];
}

template <typename R>
__global__ void k_residual(const R* vel, double* snap, const uint8_t* flags, long long ncell,
                           int d, int have_snap, double* partials) {
    double dmax = 0.0, umax = 0.0;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
         c += (long long)gridDim.x * blockDim.x) {
        double du2 = 0.0, u2 = 0.0;
        for (int a = 0; a < d; ++a) {
            const double u = (double)vel[(long long)a * ncell + c];
            const double w = u - snap[(long long)a * ncell + c];
            du2 += w * w;
            u2 += u * u;
            snap[(long long)a * ncell + c] = u;
        }
        if (!(flags[c] & 1)) {
            dmax = fmax(dmax, du2);
            umax = fmax(umax, u2);
        }
    }
    const double a = block_max(have_snap ? dmax : 0.0);
    const double b = block_max(umax);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = a;
        partials[2 * blockIdx.x + 1] = b;
    }
}

__global__ void k_max_pairs(const double* partials, int n, double* out) {
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        a = fmax(a, partials[2 * k]);
        b = fmax(b, partials[2 * k + 1]);
    }
    a = block_max(a);
    b = block_max(b);
    if (threadIdx.x == 0) {
        out[0] = sqrt(a);
        out[1] = sqrt(b);
    }
}

int bolt_lattice_residual(bolt_lattice* L, uint64_t steps_between, double* out) {
    if (!L || !out) return fail(BOLT_EINVAL, "null argument");
    HIP_TRY(hipSetDevice(L->ctx->device));
    hipStream_t s = L->ctx->stream;
    int have = L->vel_snap != nullptr;
    if (!have) {
        int rc = dev_alloc(L, (void**)&L->vel_snap, (size_t)L->ncell * L->ndim * sizeof(double));
        if (rc) return rc;
    }
    constexpr int kBlocks = kRedBlocks / 2;
    if (is_f64(L))
        hipLaunchKernelGGL(k_residual<double>, dim3(kBlocks), dim3(256), 0, s, (const double*)L->vel, L->vel_snap, L->flags, L->ncell, L->ndim, have, L->partials);
    else
        hipLaunchKernelGGL(k_residual<float>, dim3(kBlocks), dim3(256), 0, s, (const float*)L->vel, L->vel_snap, L->flags, L->ncell, L->ndim, have, L->partials);
    hipLaunchKernelGGL(k_max_pairs, dim3(1), dim3(256), 0, s, L->partials, kBlocks, L->result);
    HIP_TRY(hipGetLastError());
    double r[2];
    HIP_TRY(hipMemcpyAsync(r, L->result, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (!have || steps_between == 0 || r[1] == 0.0) *out = HUGE_VAL;
    else *out = r[0] / ((double)steps_between * r[1]);
    return BOLT_OK;
}

// ---------------------------------------------------------------------------------------
// Post-processing of the device velocity (utils/postproc.py:12-127), in float like the
// reference's numpy on the host's float32 arrays: np.gradient (central differences inside,
// one-sided first-order at the two ends of an axis) and sequential float32 cumsum.
// ---------------------------------------------------------------------------------------
template <typename R>
__device__ inline float velf(const R* vel, long long ncell, int a, long long c) {
    return (float)vel[(long long)a * ncell + c];
}

// d/dx_axis of component a at cell (i0, i1, i2) of counts n, strides st (np.gradient)
template <typename R>
__device__ inline float grad(const R* vel, long long ncell, int a, const int* n, const long long* st,
                             const int* idx, int axis, long long c) {
    const int i = idx[axis];
    const long long s = st[axis];
    if (i == 0) return velf(vel, ncell, a, c + s) - velf(vel, ncell, a, c);
    if (i == n[axis] - 1) return velf(vel, ncell, a, c) - velf(vel, ncell, a, c - s);
    return (velf(vel, ncell, a, c + s) - velf(vel, ncell, a, c - s)) / 2.0f;
}

template <typename R>
__global__ void k_vmag(const R* vel, long long ncell, int d, float* out) {
#pragma clang fp contract(off)  // numpy rounds u*u before the sum: no fused multiply-add
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    float s = 0.f;
    for (int a = 0; a < d; ++a) {
        const float u = velf(vel, ncell, a, c);
        s = s + u * u;
    }
    // correctly rounded float sqrt, as numpy's: the double square root rounded to float
    out[c] = (float)sqrt((double)s);
}

template <typename R>
__global__ void k_curl(const R* vel, long long ncell, int d, int n0, int n1, int n2, float* out) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int n[3] = {n0, n1, n2};
    const long long st[3] = {(long long)n1 * n2, (long long)n2, 1};
    const int idx[3] = {(int)(c / st[0]), (int)((c / n2) % n1), (int)(c % n2)};
    if (d == 2) {
        // du/dy - dv/dx (postproc.py: grad[1][..., 0] - grad[0][..., 1])
        out[c] = grad(vel, ncell, 0, n, st, idx, 1, c) - grad(vel, ncell, 1, n, st, idx, 0, c);
    } else {
        out[3 * c + 0] = grad(vel, ncell, 1, n, st, idx, 2, c) - grad(vel, ncell, 2, n, st, idx, 1, c);
        out[3 * c + 1] = grad(vel, ncell, 2, n, st, idx, 0, c) - grad(vel, ncell, 0, n, st, idx, 2, c);
        out[3 * c + 2] = grad(vel, ncell, 0, n, st, idx, 1, c) - grad(vel, ncell, 1, n, st, idx, 0, c);
    }
}

// ivdx = cumsum(v, axis=0): one thread per y walks x (coalesced across threads)
template <typename R>
__global__ void k_cumsum_v_x(const R* vel, long long ncell, int nx, int ny, float* ivdx) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= ny) return;
    float acc = 0.f;
    for (int x = 0; x < nx; ++x) {
        const long long c = (long long)x * ny + y;
        acc = acc + velf(vel, ncell, 1, c);
        ivdx[c] = acc;
    }
}

// iudy[0, :] = cumsum(u[0, :]): one work-item walks row x = 0
template <typename R>
__global__ void k_cumsum_u_row0(const R* vel, long long ncell, int ny, float* iudy0) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    float acc = 0.f;
    for (int y = 0; y < ny; ++y) {
        acc = acc + velf(vel, ncell, 0, y);
        iudy0[y] = acc;
    }
}

// phi = iudy[0, :] - ivdx, then phi -= phi[origin]
__global__ void k_stream_combine(const float* ivdx, const float* iudy0, long long ncell, int ny,
                                 int ox, int oy, float* phi) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const long long o = (long long)ox * ny + oy;
    const float p0 = iudy0[oy] - ivdx[o];
    const float p = iudy0[c % ny] - ivdx[c];
    phi[c] = p - p0;
}

int bolt_lattice_postproc(bolt_lattice* L, int kind, const int* origin, float* host, size_t nbytes) {
    if (!L || !host) return fail(BOLT_EINVAL, "null argument");
    const int d = L->ndim;
    size_t per = 1;
    if (kind == BOLT_POST_CURL && d == 3) per = 3;
    if (kind == BOLT_POST_STREAM && d != 2)
        return fail(BOLT_EINVAL, "the stream function is defined for 2D lattices only");
    if (kind != BOLT_POST_VMAG && kind != BOLT_POST_CURL && kind != BOLT_POST_STREAM)
        return fail(BOLT_EINVAL, "unknown post-processing kind %d", kind);
    if (kind != BOLT_POST_VMAG)
        for (int a = 0; a < d; ++a)
            if (L->n[a] < 2) return fail(BOLT_EINVAL, "gradients need at least 2 cells along every axis");
    const size_t want = (size_t)L->ncell * per * sizeof(float);
    if (nbytes != want) return fail(BOLT_EINVAL, "expected %zu bytes, got %zu", want, nbytes);
    int ox = 0, oy = 0;
    if (kind == BOLT_POST_STREAM && origin) {
        ox = origin[0];
        oy = origin[1];
        if (ox < 0 || ox >= L->n[0] || oy < 0 || oy >= L->n[1])
            return fail(BOLT_EINVAL, "origin (%d, %d) outside the lattice", ox, oy);
    }
    HIP_TRY(hipSetDevice(L->ctx->device));
    hipStream_t s = L->ctx->stream;
    // scratch: the result (ncell * per floats) and, for the stream function, ivdx and row 0
    const size_t need = (size_t)L->ncell * per + (kind == BOLT_POST_STREAM ? (size_t)L->ncell + L->n[1] : 0);
    float* buf = nullptr;
    HIP_TRY(hipMalloc(&buf, need * sizeof(float)));
    const unsigned blocks = (unsigned)((L->ncell + 255) / 256);
    const bool f64 = is_f64(L);
    const void* v = L->vel;
    if (kind == BOLT_POST_VMAG) {
        if (f64) hipLaunchKernelGGL(k_vmag<double>, dim3(blocks), dim3(256), 0, s, (const double*)v, L->ncell, d, buf);
        else hipLaunchKernelGGL(k_vmag<float>, dim3(blocks), dim3(256), 0, s, (const float*)v, L->ncell, d, buf);
    } else if (kind == BOLT_POST_CURL) {
        if (f64)
            hipLaunchKernelGGL(k_curl<double>, dim3(blocks), dim3(256), 0, s, (const double*)v, L->ncell, d, L->n[0],
                               L->n[1], L->n[2], buf);
        else
            hipLaunchKernelGGL(k_curl<float>, dim3(blocks), dim3(256), 0, s, (const float*)v, L->ncell, d, L->n[0],
                               L->n[1], L->n[2], buf);
    } else {
        float* phi = buf;
        float* ivdx = buf + L->ncell;
        float* iudy0 = ivdx + L->ncell;
        const unsigned yb = (unsigned)((L->n[1] + 255) / 256);
        if (f64) {
            hipLaunchKernelGGL(k_cumsum_v_x<double>, dim3(yb), dim3(256), 0, s, (const double*)v, L->ncell, L->n[0],
                               L->n[1], ivdx);
            hipLaunchKernelGGL(k_cumsum_u_row0<double>, dim3(1), dim3(64), 0, s, (const double*)v, L->ncell, L->n[1],
                               iudy0);
        } else {
            hipLaunchKernelGGL(k_cumsum_v_x<float>, dim3(yb), dim3(256), 0, s, (const float*)v, L->ncell, L->n[0],
                               L->n[1], ivdx);
            hipLaunchKernelGGL(k_cumsum_u_row0<float>, dim3(1), dim3(64), 0, s, (const float*)v, L->ncell, L->n[1],
                               iudy0);
        }
        hipLaunchKernelGGL(k_stream_combine, dim3(blocks), dim3(256), 0, s, ivdx, iudy0, L->ncell, L->n[1], ox, oy,
                           phi);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host, buf, want, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(buf);
    if (e != hipSuccess) return fail(BOLT_EDEVICE, "postproc: %s", hipGetErrorString(e));
    return BOLT_OK;
}

// =======================================================================================
// Batches: many independent small D2Q9 + D2Q5 Boussinesq lattices advanced together.
//
// Each instance is an ordinary bolt_lattice (one fluid, one tracer, one coupling), so
// transfers, the equilibrium initialisation and the reductions are the lattice ones.  A
// step is ONE launch over all instances: every work-group reads its tile (instance, row,
// chunk of the contiguous axis) and the instance's descriptor, and runs the same
// update_cell as the single-lattice kernel, so every instance evolves bit-identically to
// running it alone.  Steps are captured into a HIP graph and replayed; kernel boundaries
// order the in-place AA steps.  bolt_batch_tune times the tile width and the graph length
// on the batch itself and keeps the fastest.
// =======================================================================================
struct InstDesc {
    char* f;            // plane 0 of the fluid populations
    char* g;            // plane 0 of the tracer populations (null: fluid only)
    void* rho;
    void* vel0;
    void* vel1;
    void* C;
    const uint8_t* flags;
    uint64_t plane_bytes;
    int nx, ny;
    int acc_mode;
    float om_p, om_n, om_g;
    float alpha, c0, gx, gy;
};

struct Tile {
    uint16_t inst, row, chunk, pad;
};

struct BatchParams {
    const InstDesc* inst;
    const Tile* tiles;
    int ntiles;
    int write_macro;
};

// A pointer read from a descriptor in device memory, as a pointer into global memory: the
// compiler then addresses it with the global (scalar-base) memory instructions instead of
// the flat ones it must use for a pointer of unknown address space -- which also count
// against the LDS wait counter and would order every LDS access behind them.
template <class T>
__device__ inline T* global_ptr(T* p) {
    // through an integer: a cast of the generic pointer to the global space and back would
    // fold away
    return (T*)(__attribute__((address_space(1))) T*)(uintptr_t)p;
}

// The step parameters of one instance (TQ: its fused tracer's velocity set, or 0).
template <int TQ>
__device__ inline StepParams batch_step_params(const InstDesc& d, int write_macro) {
    StepParams p;
    p.n0 = d.nx;
    p.n1 = d.ny;
    p.n2 = 1;
    p.len = d.ny;
    char* const f = global_ptr(d.f);
#pragma unroll
    for (int i = 0; i < 9; ++i) p.fpl[i] = f + (size_t)i * d.plane_bytes;
    if constexpr (TQ != 0) {
        char* const g = global_ptr(d.g);
#pragma unroll
        for (int i = 0; i < TQ; ++i) p.gpl[i] = g + (size_t)i * d.plane_bytes;
    }
    p.rho = global_ptr(d.rho);
    p.vel[0] = global_ptr(d.vel0);
    p.vel[1] = global_ptr(d.vel1);
    p.flags = global_ptr(d.flags);
    p.om_p = d.om_p;
    p.om_n = d.om_n;
    p.acc_mode = d.acc_mode;
    p.g[0] = d.gx;
    p.g[1] = d.gy;
    p.alpha = d.alpha;
    p.c0 = d.c0;
    p.bq_C = global_ptr(d.C);
    p.write_macro = write_macro;
    p.tC = global_ptr(d.C);
    p.t_om = d.om_g;
    p.reverse = 0;  // the batch kernels map work-groups to tiles themselves
    return p;
}

template <int COLL, bool EVEN, int TQ, bool STREAM>
__global__ void __launch_bounds__(256) rb_batch_step(const BatchParams bp) {
    const int b = xcd_swizzle((int)blockIdx.x, bp.ntiles);
    const Tile t = bp.tiles[b];
    const InstDesc& d = bp.inst[t.inst];
    const int iy = (int)t.chunk * (int)blockDim.x + (int)threadIdx.x;
    if (iy >= d.ny) return;
    const StepParams p = batch_step_params<TQ>(d, bp.write_macro);
    update_cell<2, 9, COLL, kF32, EVEN, TQ, STREAM>(p, t.row, iy, 0, (uint32_t)t.row * (uint32_t)d.ny + iy);
}

// The even step row-staged (update.hpp row_step): one work-group per (instance, row) tile,
// launched once per group of tiles whose rows round up to the same multiple of 64 (the
// group's blockDim.x).  Same per-cell arithmetic as rb_batch_step and the single lattices.
template <int COLL, int TQ, bool STREAM>
__global__ void __launch_bounds__(1024) rb_batch_row(const BatchParams bp) {
    extern __shared__ __align__(16) unsigned char rb_row_lds_bytes[];
    const int b = xcd_swizzle((int)blockIdx.x, bp.ntiles);
    const Tile t = bp.tiles[b];
    const InstDesc& d = bp.inst[t.inst];
    const StepParams p = batch_step_params<TQ>(d, bp.write_macro);
    row_step<2, 9, COLL, kF32, TQ, STREAM>(p, (int)t.row, rb_row_lds_bytes);
}

struct bolt_batch {
    bolt_ctx* ctx = nullptr;
    int collision = 0;
    std::vector<bolt_lattice*> lat;
    InstDesc* d_inst = nullptr;
    Tile* d_tiles = nullptr;  // ntiles (instance, row, chunk) tiles, then nrtiles row tiles
    int ntiles = 0;
    int block = 128;     // tile width (cells of the contiguous axis per work-group)
    int nrtiles = 0;     // (instance, row) tiles of the row-staged even step, grouped by
                         // row length rounded up to 64: one launch per group
    struct RowGroup { int first, count, block; };
    std::vector<RowGroup> rgroups;
    int row_mode = BOLT_ROW_AUTO;  // even steps row-staged: BOLT_ROW_* (the autotuner tries both)
    bool rows_whole_lines = false; // every active row fills whole 128-byte lines
    int cache_policy = BOLT_CACHE_AUTO;  // population load form: BOLT_CACHE_* (tuner: both)
    uint64_t working_set = 0;      // bytes one step of the active instances touches
    int graph_steps = 100;
    hipGraphExec_t graph = nullptr;
    int graph_block = 0, graph_len = 0;
    bool graph_row = false;
    bool graph_stream = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint64_t cells = 0;
    bool tracers = true;
    std::vector<uint8_t> active;   // instances a step advances (empty: all)
    // batched residual: per-instance descriptors, partial maxima and results on the device
    struct ResDesc* d_res = nullptr;
    double* d_res_part = nullptr;
    double* d_res_out = nullptr;
    std::vector<uint8_t> res_have; // instance holds a velocity snapshot
};

// Destroys the captured graph.  The stream is drained first, so no replay of the graph is
// still queued or running when its executable (and the kernel arguments it owns) goes away.
static void batch_drop_graph(bolt_batch* B) {
    if (B->graph) {
        (void)hipStreamSynchronize(B->ctx->stream);
        (void)hipGraphExecDestroy(B->graph);
    }
    B->graph = nullptr;
    B->graph_block = B->graph_len = 0;
}

static int batch_build_tiles(bolt_batch* B, int block) {
    // The captured graph holds the current tile table in its kernel arguments: it goes
    // before the table does (and waits for the stream).
    batch_drop_graph(B);
    std::vector<Tile> tiles;
    std::vector<std::vector<Tile>> rt(17);  // row tiles by work-group size / 64 (64 .. 1024)
    for (size_t k = 0; k < B->lat.size(); ++k) {
        if (!B->active.empty() && !B->active[k]) continue;
        const bolt_lattice* L = B->lat[k];
        const int cpr = (L->n[1] + block - 1) / block;
        const int g = (L->n[1] + 63) / 64;  // rows longer than 1024 cells: no row staging
        for (int r = 0; r < L->n[0]; ++r) {
            for (int c = 0; c < cpr; ++c) tiles.push_back({(uint16_t)k, (uint16_t)r, (uint16_t)c, 0});
            if (g <= 16) rt[g].push_back({(uint16_t)k, (uint16_t)r, 0, 0});
            else rt[0].push_back({(uint16_t)k, (uint16_t)r, 0, 0});
        }
    }
    const size_t nchunk = tiles.size();
    B->rows_whole_lines = true;
    B->working_set = 0;
    for (size_t k = 0; k < B->lat.size(); ++k) {
        if (!B->active.empty() && !B->active[k]) continue;
        if ((size_t)B->lat[k]->n[1] * sizeof(float) % 128 != 0) B->rows_whole_lines = false;
        B->working_set += step_working_set(B->lat[k]);
    }
    B->rgroups.clear();
    const bool too_long = !rt[0].empty();
    for (int g = 1; g <= 16 && !too_long; ++g) {
        if (rt[g].empty()) continue;
        B->rgroups.push_back({(int)(tiles.size() - nchunk), (int)rt[g].size(), 64 * g});
        tiles.insert(tiles.end(), rt[g].begin(), rt[g].end());
    }
    HIP_TRY(hipStreamSynchronize(B->ctx->stream));  // no launch still reads the old table
    (void)hipFree(B->d_tiles);
    B->d_tiles = nullptr;
    if (!tiles.empty()) {
        HIP_TRY(hipMalloc(&B->d_tiles, tiles.size() * sizeof(Tile)));
        // on the batch's (non-blocking) stream, so the launches that read the table follow it
        HIP_TRY(hipMemcpyAsync(B->d_tiles, tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice,
                               B->ctx->stream));
        HIP_TRY(hipStreamSynchronize(B->ctx->stream));
    }
    B->ntiles = (int)nchunk;
    B->nrtiles = (int)(tiles.size() - nchunk);
    B->block = block;
    return BOLT_OK;
}

// LDS bytes of one row-staged work-group of `block` work-items (fluid and tracer rows
// shifted along the row; at most 8 rows of 1024 floats: 32 KiB).
static size_t batch_row_lds(const bolt_batch* B, int block) {
    return (size_t)(bolt_row_slots(2, 9) + (B->tracers ? bolt_row_slots(2, 5) : 0)) * block * sizeof(float);
}

// Whether the even step runs row-staged: every active row fits one work-group, and the mode
// asks for it (BOLT_ROW_AUTO: where the rows fill whole cache lines, as for single lattices;
// measured faster on the RB sweep and the warm-start channels).
static bool batch_row(const bolt_batch* B) {
    if (B->rgroups.empty()) return false;
    return B->row_mode == BOLT_ROW_ON || (B->row_mode == BOLT_ROW_AUTO && B->rows_whole_lines);
}

// Whether the batch's population loads stream (the working set of the active instances
// against the Infinity Cache, as for single lattices).
static bool batch_stream(const bolt_batch* B) { return loads_stream(B->cache_policy, B->working_set); }

template <int C, int TQ, bool ST>
static void batch_launch_t(bolt_batch* B, bool even, int write_macro, hipStream_t s) {
    if (even && batch_row(B)) {
        for (const auto& g : B->rgroups) {
            BatchParams bp{B->d_inst, B->d_tiles + B->ntiles + g.first, g.count, write_macro};
            hipLaunchKernelGGL((rb_batch_row<C, TQ, ST>), dim3((unsigned)g.count), dim3((unsigned)g.block),
                               batch_row_lds(B, g.block), s, bp);
        }
        return;
    }
    BatchParams bp{B->d_inst, B->d_tiles, B->ntiles, write_macro};
    const dim3 grid((unsigned)B->ntiles), blk((unsigned)B->block);
    if (even) hipLaunchKernelGGL((rb_batch_step<C, true, TQ, ST>), grid, blk, 0, s, bp);
    else hipLaunchKernelGGL((rb_batch_step<C, false, TQ, ST>), grid, blk, 0, s, bp);
}

template <int C, int TQ>
static void batch_launch_c(bolt_batch* B, bool even, int write_macro, hipStream_t s) {
    if (batch_stream(B)) batch_launch_t<C, TQ, true>(B, even, write_macro, s);
    else batch_launch_t<C, TQ, false>(B, even, write_macro, s);
}

static void batch_launch(bolt_batch* B, bool even, bool write_macro, hipStream_t s) {
    const int wm = write_macro ? 1 : 0;
    if (B->tracers) {
        if (B->collision == BOLT_BGK) batch_launch_c<0, 5>(B, even, wm, s);
        else batch_launch_c<1, 5>(B, even, wm, s);
    } else {
        if (B->collision == BOLT_BGK) batch_launch_c<0, 0>(B, even, wm, s);
        else batch_launch_c<1, 0>(B, even, wm, s);
    }
}

static int batch_capture(bolt_batch* B, int len) {
    batch_drop_graph(B);
    hipStream_t s = B->ctx->stream;
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < len; ++k) batch_launch(B, (k % 2) == 0, false, s);
    HIP_TRY(hipStreamEndCapture(s, &g));
    hipError_t e = hipGraphInstantiate(&B->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(BOLT_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
    B->graph_block = B->block;
    B->graph_len = len;
    B->graph_row = batch_row(B);
    B->graph_stream = batch_stream(B);
    return BOLT_OK;
}

static int batch_enqueue(bolt_batch* B, uint64_t n) {
    if (n % 2 != 0) return fail(BOLT_EINVAL, "iters must be even");
    if (B->ntiles == 0) return BOLT_OK;  // every instance is inactive
    HIP_TRY(hipSetDevice(B->ctx->device));
    hipStream_t s = B->ctx->stream;
    const uint64_t len = (uint64_t)B->graph_steps;
    uint64_t done = 0;
    if (len > 0 && n > len + 2) {  // graph_steps 0: every step is a direct launch
        if (!B->graph || B->graph_block != B->block || B->graph_len != (int)len || B->graph_row != batch_row(B) ||
            B->graph_stream != batch_stream(B)) {
            int rc = batch_capture(B, (int)len);
            if (rc) return rc;
        }
        const uint64_t replays = (n - 2) / len;
        for (uint64_t r = 0; r < replays; ++r) HIP_TRY(hipGraphLaunch(B->graph, s));
        done = replays * len;
    }
    for (uint64_t k = done; k < n; ++k) batch_launch(B, (k % 2) == 0, k + 1 == n, s);
    HIP_TRY(hipGetLastError());
    return BOLT_OK;
}

static int batch_upload_desc(bolt_batch* B) {
    std::vector<InstDesc> h(B->lat.size());
    for (size_t k = 0; k < B->lat.size(); ++k) {
        bolt_lattice* L = B->lat[k];
        InstDesc& d = h[k];
        d.f = (char*)L->f;
        d.g = L->tracers.empty() ? nullptr : (char*)L->tracers[0].g;
        d.rho = L->rho;
        d.vel0 = L->vel;
        d.vel1 = (char*)L->vel + (size_t)L->ncell * 4;
        d.C = L->tracers.empty() ? nullptr : L->tracers[0].C;
        d.flags = L->flags;
        d.plane_bytes = (uint64_t)L->stride * 4;
        d.nx = L->n[0];
        d.ny = L->n[1];
        d.om_p = L->om_p;
        d.om_n = L->om_n;
        d.om_g = L->tracers.empty() ? 0.f : L->tracers[0].omega;
        if (!L->couplings.empty()) {
            const Coupling& c = L->couplings.back();
            d.acc_mode = kAccBoussinesq;
            d.alpha = c.alpha;
            d.c0 = c.c0;
            d.gx = c.g[0];
            d.gy = c.g[1];
        } else {
            d.acc_mode = L->has_gravity ? kAccConst : kAccNone;
            d.gx = L->gravity[0];
            d.gy = L->gravity[1];
        }
    }
    (void)hipFree(B->d_inst);
    B->d_inst = nullptr;
    HIP_TRY(hipMalloc(&B->d_inst, h.size() * sizeof(InstDesc)));
    HIP_TRY(hipMemcpyAsync(B->d_inst, h.data(), h.size() * sizeof(InstDesc), hipMemcpyHostToDevice,
                           B->ctx->stream));
    HIP_TRY(hipStreamSynchronize(B->ctx->stream));
    return BOLT_OK;
}

int bolt_batch_create(bolt_ctx* ctx, int n, const int64_t* counts, int collision,
                      const float* omega_f, const float* omega_g, const float* alpha,
                      const float* c0, const float* g, bolt_batch** out) {
    if (!ctx || !counts || !omega_f || !out) return fail(BOLT_EINVAL, "null argument");
    if (omega_g && (!alpha || !c0 || !g)) return fail(BOLT_EINVAL, "null coupling argument");
    *out = nullptr;
    if (n < 1 || n > 65535) return fail(BOLT_EINVAL, "instance count %d out of range", n);
    auto B = std::make_unique<bolt_batch>();
    B->ctx = ctx;
    B->collision = collision;
    B->tracers = omega_g != nullptr;
    for (int k = 0; k < n; ++k) {
        // tiles hold the instance, the row and the 64-cell chunk of the row in 16 bits each
        if (counts[2 * k] > 65535 || counts[2 * k + 1] > 64ll * 65535) {
            for (auto* L : B->lat) bolt_lattice_destroy(L);
            return fail(BOLT_EINVAL, "instance %d: nx above 65535 or ny above 64*65535", k);
        }
        bolt_lattice* L = nullptr;
        int rc = bolt_lattice_create(ctx, 2, counts + 2 * k, 9, collision, omega_f[2 * k],
                                     omega_f[2 * k + 1], BOLT_STORE_F32, &L);
        int tid = 0;
        if (!rc) B->lat.push_back(L);
        if (!rc && omega_g) rc = bolt_lattice_add_tracer(L, 5, omega_g[k], &tid);
        if (!rc && omega_g) rc = bolt_lattice_add_coupling(L, tid, alpha[k], c0[k], g + 2 * k);
        if (rc) {
            std::string msg = std::string("instance ") + std::to_string(k) + ": " + bolt_error_text();
            for (auto* l : B->lat) bolt_lattice_destroy(l);
            return fail(rc, "%s", msg.c_str());
        }
        B->cells += (uint64_t)L->ncell;
    }
    HIP_TRY(hipEventCreate(&B->ev0));
    HIP_TRY(hipEventCreate(&B->ev1));
    int rc = batch_upload_desc(B.get());
    if (!rc) rc = batch_build_tiles(B.get(), 128);
    if (rc) {
        for (auto* l : B->lat) bolt_lattice_destroy(l);
        return rc;
    }
    *out = B.release();
    return BOLT_OK;
}

int bolt_batch_destroy(bolt_batch* B) {
    if (!B) return BOLT_OK;
    (void)hipSetDevice(B->ctx->device);
    (void)hipStreamSynchronize(B->ctx->stream);
    batch_drop_graph(B);
    for (auto* L : B->lat) bolt_lattice_destroy(L);
    (void)hipFree(B->d_inst);
    (void)hipFree(B->d_tiles);
    (void)hipFree(B->d_res);
    (void)hipFree(B->d_res_part);
    (void)hipFree(B->d_res_out);
    if (B->ev0) (void)hipEventDestroy(B->ev0);
    if (B->ev1) (void)hipEventDestroy(B->ev1);
    delete B;
    return BOLT_OK;
}

static int batch_lat(bolt_batch* B, int inst, bolt_lattice** L) {
    if (!B) return fail(BOLT_EINVAL, "null batch");
    if (inst < 0 || inst >= (int)B->lat.size()) return fail(BOLT_EINVAL, "unknown instance %d", inst);
    *L = B->lat[inst];
    return BOLT_OK;
}

int bolt_batch_write(bolt_batch* B, int inst, int field, int host_dtype, const void* host,
                     size_t nbytes) {
    bolt_lattice* L;
    int rc = batch_lat(B, inst, &L);
    return rc ? rc : bolt_lattice_write(L, field, 0, host_dtype, host, nbytes);
}

int bolt_batch_read(bolt_batch* B, int inst, int field, int host_dtype, void* host,
                    size_t nbytes) {
    bolt_lattice* L;
    int rc = batch_lat(B, inst, &L);
    return rc ? rc : bolt_lattice_read(L, field, 0, host_dtype, host, nbytes);
}

int bolt_batch_equilibrate(bolt_batch* B) {
    if (!B) return fail(BOLT_EINVAL, "null batch");
    for (auto* L : B->lat) {
        int rc = bolt_lattice_equilibrate(L);
        if (rc) return rc;
    }
    return BOLT_OK;
}

int bolt_batch_step(bolt_batch* B, uint64_t n) {
    if (!B) return fail(BOLT_EINVAL, "null batch");
    return batch_enqueue(B, n);
}

int bolt_batch_set_active(bolt_batch* B, const uint8_t* active, int n) {
    if (!B || !active) return fail(BOLT_EINVAL, "null argument");
    if (n != (int)B->lat.size()) return fail(BOLT_EINVAL, "expected %zu flags, got %d", B->lat.size(), n);
    HIP_TRY(hipSetDevice(B->ctx->device));
    B->active.assign(active, active + n);
    return batch_build_tiles(B, B->block);
}

int bolt_batch_step_timed(bolt_batch* B, uint64_t n, double* ms) {
    if (!B) return fail(BOLT_EINVAL, "null batch");
    HIP_TRY(hipSetDevice(B->ctx->device));
    HIP_TRY(hipEventRecord(B->ev0, B->ctx->stream));
    int rc = batch_enqueue(B, n);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(B->ev1, B->ctx->stream));
    HIP_TRY(hipEventSynchronize(B->ev1));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, B->ev0, B->ev1));
    if (ms) *ms = t;
    return BOLT_OK;
}

int bolt_batch_reduce(bolt_batch* B, int inst, int what, double* out) {
    bolt_lattice* L;
    int rc = batch_lat(B, inst, &L);
    return rc ? rc : bolt_lattice_reduce(L, what, 0, out);
}

int bolt_batch_cells(const bolt_batch* B, uint64_t* out) {
    if (!B || !out) return fail(BOLT_EINVAL, "null argument");
    *out = B->cells;
    return BOLT_OK;
}

int bolt_batch_set_row_staging(bolt_batch* B, int mode) {
    if (!B) return fail(BOLT_EINVAL, "null batch");
    if (mode != BOLT_ROW_OFF && mode != BOLT_ROW_ON && mode != BOLT_ROW_AUTO)
        return fail(BOLT_EINVAL, "row staging mode must be BOLT_ROW_OFF, _ON or _AUTO");
    B->row_mode = mode;
    batch_drop_graph(B);
    return BOLT_OK;
}

int bolt_batch_set_cache_policy(bolt_batch* B, int policy) {
    if (!B) return fail(BOLT_EINVAL, "null batch");
    if (!valid_cache_policy(policy))
        return fail(BOLT_EINVAL, "cache policy must be BOLT_CACHE_STREAM, _KEEP or _AUTO");
    B->cache_policy = policy;
    batch_drop_graph(B);
    return BOLT_OK;
}

int bolt_batch_set_shape(bolt_batch* B, int block, int graph_steps) {
    if (!B) return fail(BOLT_EINVAL, "null batch");
    if (block < 64 || block > 256 || block % 64 != 0)
        return fail(BOLT_EINVAL, "block must be 64, 128, 192 or 256");
    if (graph_steps < 0 || graph_steps % 2 != 0 || graph_steps > 4096)
        return fail(BOLT_EINVAL, "graph_steps must be even, 0..4096");
    B->graph_steps = graph_steps;
    if (block != B->block) return batch_build_tiles(B, block);
    return BOLT_OK;
}

// Autotuner: times every (population load form, even steps row-staged or direct, tile
// width, graph length) candidate over `steps` steps on the batch itself, restores the
// batch's state afterwards, and keeps the fastest.
int bolt_batch_tune(bolt_batch* B, int steps, char* report, size_t len) {
    if (!B) return fail(BOLT_EINVAL, "null batch");
    if (steps < 4 || steps % 2) return fail(BOLT_EINVAL, "steps must be even and >= 4");
    HIP_TRY(hipSetDevice(B->ctx->device));
    hipStream_t s = B->ctx->stream;
    // snapshot the state the tuning steps overwrite
    struct Saved { void* dst; void* src; size_t bytes; };
    std::vector<Saved> saved;
    auto keep = [&](void* p, size_t bytes) -> int {
        void* c = nullptr;
        HIP_TRY(hipMalloc(&c, bytes));
        HIP_TRY(hipMemcpyAsync(c, p, bytes, hipMemcpyDeviceToDevice, s));
        saved.push_back({p, c, bytes});
        return BOLT_OK;
    };
    int rc = BOLT_OK;
    for (auto* L : B->lat) {
        if (!rc) rc = keep(L->f, (size_t)L->stride * 9 * 4);
        if (!rc && B->tracers) rc = keep(L->tracers[0].g, (size_t)L->stride * 5 * 4);
        if (!rc && B->tracers) rc = keep(L->tracers[0].C, (size_t)L->ncell * 4);
        if (!rc) rc = keep(L->rho, (size_t)L->ncell * 4);
        if (!rc) rc = keep(L->vel, (size_t)L->ncell * 2 * 4);
    }
    std::string rep;
    int best_block = B->block, best_len = B->graph_steps;
    bool best_row = batch_row(B), best_stream = batch_stream(B);
    double best = 1e300;
    const int blocks[] = {64, 128, 256};
    const int lens[] = {20, 100, 400};
    for (int stream = 1; stream >= 0 && !rc; --stream) {
        rc = bolt_batch_set_cache_policy(B, stream ? BOLT_CACHE_STREAM : BOLT_CACHE_KEEP);
        for (int row = 1; row >= 0 && !rc; --row) {
            rc = bolt_batch_set_row_staging(B, row ? BOLT_ROW_ON : BOLT_ROW_OFF);
            for (int blk : blocks) {
                if (rc) break;
                for (int gl : lens) {
                    if (rc) break;
                    rc = bolt_batch_set_shape(B, blk, gl);
                    if (rc) break;
                    double ms = 0;
                    rc = bolt_batch_step_timed(B, (uint64_t)std::max(steps, gl + 4), &ms);  // warm + capture
                    if (!rc) rc = bolt_batch_step_timed(B, (uint64_t)std::max(steps, gl + 4), &ms);
                    if (rc) break;
                    const double per = ms / std::max(steps, gl + 4);
                    char line[112];
                    snprintf(line, sizeof line, "block %d graph %d rows %s loads %s: %.4f ms/step\n", blk, gl,
                             row ? "staged" : "direct", stream ? "streaming" : "cached", per);
                    rep += line;
                    if (per < best) {
                        best = per;
                        best_block = blk;
                        best_len = gl;
                        best_row = row != 0;
                        best_stream = stream != 0;
                    }
                }
            }
        }
    }
    // restore the state the candidates stepped; a failed restore is reported, since the
    // batch would otherwise continue from a tuning run's state
    hipError_t e = hipSuccess;
    for (auto& sv : saved)
        if (e == hipSuccess) e = hipMemcpyAsync(sv.dst, sv.src, sv.bytes, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    for (auto& sv : saved) (void)hipFree(sv.src);
    if (rc) return rc;
    if (e != hipSuccess) return fail(BOLT_EDEVICE, "tune: restoring the batch state failed: %s", hipGetErrorString(e));
    rc = bolt_batch_set_row_staging(B, best_row ? BOLT_ROW_ON : BOLT_ROW_OFF);
    if (!rc) rc = bolt_batch_set_cache_policy(B, best_stream ? BOLT_CACHE_STREAM : BOLT_CACHE_KEEP);
    if (!rc) rc = bolt_batch_set_shape(B, best_block, best_len);
    char line[112];
    snprintf(line, sizeof line, "chosen: block %d graph %d rows %s loads %s\n", best_block, best_len,
             best_row ? "staged" : "direct", best_stream ? "streaming" : "cached");
    rep += line;
    if (report && len) snprintf(report, len, "%s", rep.c_str());
    return rc;
}

int bolt_batch_residual(bolt_batch* B, int inst, uint64_t steps_between, double* out) {
    bolt_lattice* L;
    int rc = batch_lat(B, inst, &L);
    return rc ? rc : bolt_lattice_residual(L, steps_between, out);
}

// The residual of every active instance in one launch: work-group (x, k) takes slice x of
// instance k's cells (the arithmetic of k_residual), then one work-group per instance
// reduces its slices.
struct ResDesc {
    const float* vel;     // D2Q9 instances store float velocity planes
    double* snap;
    const uint8_t* flags;
    long long ncell;
    int active, have;
};

constexpr int kResSlices = 16;

__global__ void k_batch_residual(const ResDesc* desc, double* part) {
    const ResDesc r = desc[blockIdx.y];
    const float* const vel = global_ptr(r.vel);
    double* const snap = global_ptr(r.snap);
    const uint8_t* const flags = global_ptr(r.flags);
    double dmax = 0.0, umax = 0.0;
    if (r.active) {
        for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < r.ncell;
             c += (long long)gridDim.x * blockDim.x) {
            double du2 = 0.0, u2 = 0.0;
            for (int a = 0; a < 2; ++a) {
                const double u = (double)vel[(long long)a * r.ncell + c];
                const double w = u - snap[(long long)a * r.ncell + c];
                du2 += w * w;
                u2 += u * u;
                snap[(long long)a * r.ncell + c] = u;
            }
            if (!(flags[c] & 1)) {
                dmax = fmax(dmax, du2);
                umax = fmax(umax, u2);
            }
        }
    }
    const double a = block_max(r.have ? dmax : 0.0);
    const double b = block_max(umax);
    if (threadIdx.x == 0) {
        part[2 * ((long long)blockIdx.y * gridDim.x + blockIdx.x)] = a;
        part[2 * ((long long)blockIdx.y * gridDim.x + blockIdx.x) + 1] = b;
    }
}

__global__ void k_batch_residual_max(const double* part, int slices, double* out) {
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < slices; k += blockDim.x) {
        a = fmax(a, part[2 * ((long long)blockIdx.x * slices + k)]);
        b = fmax(b, part[2 * ((long long)blockIdx.x * slices + k) + 1]);
    }
    a = block_max(a);
    b = block_max(b);
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = sqrt(a);
        out[2 * blockIdx.x + 1] = sqrt(b);
    }
}

int bolt_batch_residuals(bolt_batch* B, uint64_t steps_between, double* out) {
    if (!B || !out) return fail(BOLT_EINVAL, "null argument");
    HIP_TRY(hipSetDevice(B->ctx->device));
    hipStream_t s = B->ctx->stream;
    const int n = (int)B->lat.size();
    if (!B->d_res) {
        for (auto* L : B->lat)
            if (!L->vel_snap) {
                int rc = dev_alloc(L, (void**)&L->vel_snap, (size_t)L->ncell * L->ndim * sizeof(double));
                if (rc) return rc;
            }
        HIP_TRY(hipMalloc(&B->d_res, n * sizeof(ResDesc)));
        HIP_TRY(hipMalloc(&B->d_res_part, (size_t)n * kResSlices * 2 * sizeof(double)));
        HIP_TRY(hipMalloc(&B->d_res_out, (size_t)n * 2 * sizeof(double)));
        B->res_have.assign(n, 0);
    }
    std::vector<ResDesc> h(n);
    for (int k = 0; k < n; ++k) {
        const bolt_lattice* L = B->lat[k];
        h[k] = {(const float*)L->vel, L->vel_snap, L->flags, L->ncell,
                B->active.empty() || B->active[k] ? 1 : 0, B->res_have[k]};
    }
    HIP_TRY(hipMemcpyAsync(B->d_res, h.data(), n * sizeof(ResDesc), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_batch_residual, dim3(kResSlices, n), dim3(256), 0, s, B->d_res, B->d_res_part);
    hipLaunchKernelGGL(k_batch_residual_max, dim3(n), dim3(256), 0, s, B->d_res_part, kResSlices, B->d_res_out);
    HIP_TRY(hipGetLastError());
    std::vector<double> r((size_t)n * 2);
    HIP_TRY(hipMemcpyAsync(r.data(), B->d_res_out, r.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int k = 0; k < n; ++k) {
        if (!h[k].active) {
            out[k] = NAN;
            continue;
        }
        out[k] = (!h[k].have || steps_between == 0 || r[2 * k + 1] == 0.0)
                     ? HUGE_VAL : r[2 * k] / ((double)steps_between * r[2 * k + 1]);
        B->res_have[k] = 1;
    }
    return BOLT_OK;
}
