// Memory ceilings for the lattice update's access pattern on the current GPU.
//
// Times, with HIP events, kernels that move the same bytes as one AA step of a 256^3 D3Q27
// lattice (27 planes of 16,777,216 elements) without the stencil shift or the collision:
//   copy4         float4 copy between two buffers (the usual "stream copy" figure)
//   read27        27 float planes read, nothing written
//   write27       27 float planes written, nothing read
//   inplace27     27 float planes read and written back in place (the update's mix)
//   shift27       the same, each population at its D3Q27 neighbour (periodic): the
//                 access mix of the step that streams, where 18 of the 27 accesses of a
//                 wavefront start one element off a 256-byte boundary
//   inplace27_h   the same as inplace27 with 2-byte elements (the f16s/bf16s storage)
//   shift27_h     the same as shift27 with 2-byte elements
//   inplace9      9 float planes in place over 3x the cells (the D2Q9 mix)
// (each in-place figure the better of plain and streaming loads) and prints one JSON
// object of GB/s (bytes read + written per second).  Every index is
// checked against the element count it addresses.
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#define CK(x)                                                                              \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            fprintf(stderr, "bw_probe: %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            exit(1);                                                                       \
        }                                                                                  \
    } while (0)

__global__ void copy4(const float4* __restrict__ a, float4* __restrict__ b, long long n4) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n4) b[i] = a[i];
}

template <int Q>
__global__ void read_planes(const float* __restrict__ f, float* __restrict__ sink, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
#pragma unroll
    for (int q = 0; q < Q; ++q) s += f[q * n + i];
    if (s == 1234.5f) sink[i] = s;  // keeps the loads; never true for the zeroed input
}

template <int Q>
__global__ void write_planes(float* __restrict__ f, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q * n + i] = (float)q;
}

// NT: streaming (non-temporal) loads, as the update kernels use
template <int Q, typename T, bool NT>
__global__ void inplace_planes(T* __restrict__ f, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T v[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        if constexpr (NT) v[q] = __builtin_nontemporal_load(f + q * n + i);
        else v[q] = f[q * n + i];
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q * n + i] = v[(q + 1) % Q];
}

// Each element (q, j) is read and written by exactly one thread: for a fixed q the map
// i -> (i + shift_q) mod n is a bijection.  n is a power of two, so the wrap is a mask.
template <typename T, bool NT>
__global__ void shifted_planes(T* __restrict__ f, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T v[27];
    long long a[27];
#pragma unroll
    for (int q = 0; q < 27; ++q) {
        const long long d = (q % 3 - 1) + 256 * ((q / 3) % 3 - 1) + 65536 * (q / 9 - 1);
        a[q] = q * n + ((i + d) & (n - 1));
        if constexpr (NT) v[q] = __builtin_nontemporal_load(f + a[q]);
        else v[q] = f[a[q]];
    }
#pragma unroll
    for (int q = 0; q < 27; ++q) f[a[q]] = v[(q + 1) % 27];
}

template <typename F>
static double gbs(F launch, double bytes, int reps) {
    hipEvent_t a, b;
    CK(hipEventCreate(&a));
    CK(hipEventCreate(&b));
    launch();
    launch();
    CK(hipGetLastError());
    CK(hipDeviceSynchronize());
    CK(hipEventRecord(a, 0));
    for (int r = 0; r < reps; ++r) launch();
    CK(hipEventRecord(b, 0));
    CK(hipEventSynchronize(b));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, a, b));
    CK(hipEventDestroy(a));
    CK(hipEventDestroy(b));
    return bytes / (ms * 1e-3 / reps) / 1e9;
}

int main(int argc, char** argv) {
    const long long n = 256ll * 256 * 256;  // cells of the 256^3 lattice
    const int reps = argc > 1 ? atoi(argv[1]) : 20;
    const long long total = 27 * n;         // floats in one lattice
    float* f = nullptr;
    float* sink = nullptr;
    CK(hipMalloc(&f, sizeof(float) * total * 2));  // two lattices: copy source and target
    CK(hipMalloc(&sink, sizeof(float) * n));
    CK(hipMemset(f, 0, sizeof(float) * total * 2));
    float* g = f + total;
    const unsigned b = 256;
    auto grid = [&](long long items) { return dim3((unsigned)((items + b - 1) / b)); };
    const long long n4 = total / 4;  // total is a multiple of 4
    const double lat = 2.0 * 4.0 * (double)total;

    double r_copy = gbs([&] { copy4<<<grid(n4), b>>>((const float4*)f, (float4*)g, n4); }, lat, reps);
    double r_read = gbs([&] { read_planes<27><<<grid(n), b>>>(f, sink, n); }, 4.0 * total, reps);
    double r_write = gbs([&] { write_planes<27><<<grid(n), b>>>(f, n); }, 4.0 * total, reps);
    // the in-place mixes, each the better of plain and streaming loads
    auto best = [](double a, double c) { return a > c ? a : c; };
    double r_in = best(gbs([&] { inplace_planes<27, float, false><<<grid(n), b>>>(f, n); }, lat, reps),
                       gbs([&] { inplace_planes<27, float, true><<<grid(n), b>>>(f, n); }, lat, reps));
    double r_sh = best(gbs([&] { shifted_planes<float, false><<<grid(n), b>>>(f, n); }, lat, reps),
                       gbs([&] { shifted_planes<float, true><<<grid(n), b>>>(f, n); }, lat, reps));
    // 2-byte elements: 27 planes of n halves occupy half of the first lattice
    const double half = 2.0 * 2.0 * (double)total;
    double r_inh = best(
        gbs([&] { inplace_planes<27, uint16_t, false><<<grid(n), b>>>((uint16_t*)f, n); }, half, reps),
        gbs([&] { inplace_planes<27, uint16_t, true><<<grid(n), b>>>((uint16_t*)f, n); }, half, reps));
    double r_shh = best(
        gbs([&] { shifted_planes<uint16_t, false><<<grid(n), b>>>((uint16_t*)f, n); }, half, reps),
        gbs([&] { shifted_planes<uint16_t, true><<<grid(n), b>>>((uint16_t*)f, n); }, half, reps));
    // 9 planes over 3n cells: the same 27n floats
    double r_in9 = best(
        gbs([&] { inplace_planes<9, float, false><<<grid(3 * n), b>>>(f, 3 * n); }, lat, reps),
        gbs([&] { inplace_planes<9, float, true><<<grid(3 * n), b>>>(f, 3 * n); }, lat, reps));

    hipDeviceProp_t p;
    CK(hipGetDeviceProperties(&p, 0));
    printf("{\"device\": \"%s\", \"bytes_per_pass\": %.0f, \"copy4\": %.1f, \"read27\": %.1f, "
           "\"write27\": %.1f, \"inplace27\": %.1f, \"shift27\": %.1f, \"inplace27_h\": %.1f, "
           "\"shift27_h\": %.1f, \"inplace9\": %.1f}\n",
           p.gcnArchName, lat, r_copy, r_read, r_write, r_in, r_sh, r_inh, r_shh, r_in9);
    CK(hipFree(f));
    CK(hipFree(sink));
    return 0;
}
