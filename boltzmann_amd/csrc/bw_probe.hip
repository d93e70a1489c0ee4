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
//   inplace27_w2, inplace27_w4
//                 inplace27 with 2 or 4 consecutive cells per lane (8- and 16-byte
//                 aligned accesses, dwordx2 / dwordx4)
//   shift27_w4    shift27 with 4 cells per lane: every access is an aligned 16-byte
//                 load or store, a wavefront covers one 256-cell row, and the +-1 shift
//                 along the row is a one-lane rotation across the wavefront (DPP)
//   shift27_w4h   the same with the general-row handling a row longer than one
//                 wavefront needs: lane 0 loads and stores the element before the
//                 wavefront's range itself and the last lane stores three elements
//   inplace27_hw4, shift27_hw4
//                 inplace27_h and shift27_h with 4 cells per lane (8-byte accesses)
//   shift27_w2, shift27_w2h
//                 shift27_w4 and shift27_w4h with 2 cells per lane (8-byte accesses) on
//                 rows of 128 cells
// (each in-place figure the better of plain and streaming loads) and prints one JSON
// object of GB/s (bytes read + written per second).
//
// `bw_probe REPS layout` measures instead how the plane layout moves the inplace27 and
// shift27 ceilings: planes 2^26 bytes apart (the unpadded 256^3 lattice) against planes
// padded by 256 B to 1 MiB, and the 27 planes interleaved in blocks of 64-1024 cells.  Every index is
// checked against the element count it addresses.
//
// `bw_probe REPS cache` times the inplace27 mix on lattices of 2^17 to 2^24 cells (14 MB to
// 1.8 GB of planes) with ordinary and with streaming loads: where the planes fit MI355X's
// 256 MB Infinity Cache, ordinary loads leave the lines there for the next pass.
//
// `bw_probe REPS march` times a step pair of the inplace27 mix on the 256^3 lattice as two
// full passes, and as a march over K = 8, 16, 32 slabs of consecutive cells: pass one over
// slab s, then pass two over slab s - 1, so that pass two finds its slab's lines fresh.
//
// `bw_probe REPS sync` prices the step boundary of a persistent kernel against launches: the
// time of one grid-wide barrier (cooperative launch, every work-group resident, cooperative
// groups' grid sync) against one launch of an empty kernel replayed from a HIP graph and
// one issued directly on a stream.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

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
// Plane q starts q * st elements in (st >= n: st = n is the unpadded layout).
template <int Q, typename T, bool NT>
__global__ void inplace_planes(T* __restrict__ f, long long n, long long st) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T v[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        if constexpr (NT) v[q] = __builtin_nontemporal_load(f + q * st + i);
        else v[q] = f[q * st + i];
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q * st + i] = v[(q + 1) % Q];
}

// The same mix with the planes interleaved in blocks of B cells: block b holds Q runs of B
// elements, population q of cell b * B + j at b * Q * B + q * B + j (n a multiple of B).
template <int Q, int B, bool NT>
__global__ void inplace_blocked(float* __restrict__ f, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float* base = f + (i / B) * (Q * B) + (i % B);
    float v[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        if constexpr (NT) v[q] = __builtin_nontemporal_load(base + q * B);
        else v[q] = base[q * B];
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) base[q * B] = v[(q + 1) % Q];
}

// Each element (q, j) is read and written by exactly one thread: for a fixed q the map
// i -> (i + shift_q) mod n is a bijection.  n is a power of two, so the wrap is a mask.
template <typename T, bool NT>
__global__ void shifted_planes(T* __restrict__ f, long long n, long long st) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T v[27];
    long long a[27];
#pragma unroll
    for (int q = 0; q < 27; ++q) {
        const long long d = (q % 3 - 1) + 256 * ((q / 3) % 3 - 1) + 65536 * (q / 9 - 1);
        a[q] = q * st + ((i + d) & (n - 1));
        if constexpr (NT) v[q] = __builtin_nontemporal_load(f + a[q]);
        else v[q] = f[a[q]];
    }
#pragma unroll
    for (int q = 0; q < 27; ++q) f[a[q]] = v[(q + 1) % 27];
}

// ---- wide accesses: VW consecutive 4-byte words of a plane per lane ----------------------
template <int VW>
struct WideT;
template <>
struct WideT<2> {
    typedef uint32_t T __attribute__((ext_vector_type(2)));
};
template <>
struct WideT<4> {
    typedef uint32_t T __attribute__((ext_vector_type(4)));
};

template <int Q, int VW, bool NT>
__global__ void __launch_bounds__(256) inplace_wide(uint32_t* __restrict__ f, long long nv) {
    using V = typename WideT<VW>::T;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nv) return;  // nv = vectors per plane
    V* p = (V*)f;
    V v[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        if constexpr (NT) v[q] = __builtin_nontemporal_load(p + q * nv + t);
        else v[q] = p[q * nv + t];
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) p[q * nv + t] = v[(q + 1) % Q];
}

// One-lane rotations across the 64-lane wavefront (DPP wave_ror:1 / wave_rol:1):
// ror1(x) on lane l returns x of lane (l - 1) mod 64, rol1(x) that of lane (l + 1) mod 64.
__device__ inline uint32_t ror1(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x13C, 0xF, 0xF, false);
}
__device__ inline uint32_t rol1(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x134, 0xF, 0xF, false);
}

// Shift of one 4-word vector per lane by one element along the row, for 4-byte elements
// (H = false) or 2-byte elements packed two to a word (H = true).
//   down: element e of the result is element e - 1 of the row (lane 0 takes lane 63's last)
//   up:   element e of the result is element e + 1 of the row (lane 63 takes lane 0's first)
template <int VW, bool H>
__device__ inline void shift_down(typename WideT<VW>::T& a) {
    const uint32_t prev = ror1(a[VW - 1]);
    if constexpr (!H) {
#pragma unroll
        for (int k = VW - 1; k > 0; --k) a[k] = a[k - 1];
        a[0] = prev;
    } else {
#pragma unroll
        for (int k = VW - 1; k > 0; --k) a[k] = __builtin_amdgcn_alignbyte(a[k], a[k - 1], 2);
        a[0] = __builtin_amdgcn_alignbyte(a[0], prev, 2);
    }
}
template <int VW, bool H>
__device__ inline void shift_up(typename WideT<VW>::T& a) {
    const uint32_t next = rol1(a[0]);
    if constexpr (!H) {
#pragma unroll
        for (int k = 0; k < VW - 1; ++k) a[k] = a[k + 1];
        a[VW - 1] = next;
    } else {
#pragma unroll
        for (int k = 0; k < VW - 1; ++k) a[k] = __builtin_amdgcn_alignbyte(a[k + 1], a[k], 2);
        a[VW - 1] = __builtin_amdgcn_alignbyte(next, a[VW - 1], 2);
    }
}

// The D3Q27 neighbour mix with VW words per lane on a lattice of (rows / 256) x 256 rows of
// 64*VW words each: one wavefront per row, row r = (block * 256 + thread) / 64.  Plane q's access
// of row r is the aligned vector of the row at (x + dx, y + dy), shifted by dz elements
// along the row by rotation.  HALO: the general handling a row longer than one wavefront
// needs, where the element across the wavefront's edge belongs to another wavefront's
// range.  For dz = -1 lane 0 loads and stores the element before the range itself and
// lane 63 stores VW - 1 elements; for dz = +1 lane 63 loads and stores the element after
// the range and lane 0 stores the last VW - 1.  (On these one-wavefront rows that element
// is the periodic wrap inside the same row.)  Every address is the plane base plus a row
// index below nv / 64 times 64 plus a lane (vector units), or one word of those vectors.
template <int VW, bool H, bool HALO>
__global__ void __launch_bounds__(256) shift_wide(uint32_t* __restrict__ f, long long nv) {
    using V = typename WideT<VW>::T;
    static_assert(!(H && HALO), "the halo probe is for 4-byte elements");
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nv) return;  // nv = rows * 64 lanes, rows = 256 * (a power of two)
    const int lane = (int)(t & 63);
    const int r = (int)(t >> 6);
    const int xmask = (int)(nv >> 14) - 1;  // rows / 256 - 1
    const int x = r >> 8, y = r & 255;
    V* p = (V*)f;
    V v[27];
    long long a[27];
#pragma unroll
    for (int q = 0; q < 27; ++q) {
        const int dz = q % 3 - 1, dy = (q / 3) % 3 - 1, dx = q / 9 - 1;
        const int rr = (((x + dx) & xmask) << 8) | ((y + dy) & 255);
        a[q] = q * nv + ((long long)rr << 6) + lane;
        v[q] = __builtin_nontemporal_load(p + a[q]);
        if (dz == -1) {
            shift_down<VW, H>(v[q]);
            if (HALO && lane == 0)  // last word of the row: lane 63's vector
                v[q][0] = __builtin_nontemporal_load((uint32_t*)(p + a[q] + 63) + VW - 1);
        } else if (dz == 1) {
            shift_up<VW, H>(v[q]);
            if (HALO && lane == 63)  // first word of the row: lane 0's vector
                v[q][VW - 1] = __builtin_nontemporal_load((uint32_t*)(p + a[q] - 63));
        }
    }
#pragma unroll
    for (int q = 0; q < 27; ++q) {
        const int dz = q % 3 - 1;
        V w = v[(q + 1) % 27];
        if (dz == -1) {
            const uint32_t edge = w[0];
            shift_up<VW, H>(w);
            if (HALO) {
                uint32_t* pw = (uint32_t*)(p + a[q]);
                if (lane == 0) *((uint32_t*)(p + a[q] + 63) + VW - 1) = edge;
                if (lane == 63) {
#pragma unroll
                    for (int k = 0; k < VW - 1; ++k) pw[k] = w[k];
                } else {
                    p[a[q]] = w;
                }
                continue;
            }
        } else if (dz == 1) {
            const uint32_t edge = w[VW - 1];
            shift_down<VW, H>(w);
            if (HALO) {
                uint32_t* pw = (uint32_t*)(p + a[q]);
                if (lane == 63) *(uint32_t*)(p + a[q] - 63) = edge;
                if (lane == 0) {
#pragma unroll
                    for (int k = 1; k < VW; ++k) pw[k] = w[k];
                } else {
                    p[a[q]] = w;
                }
                continue;
            }
        }
        p[a[q]] = w;
    }
}

// The neighbour mix with every access an aligned row: one 256-thread work-group per row of
// 256 cells (rows are (x, y) pairs of a 256^3 lattice, the row axis contiguous), one cell per
// thread.  Each plane is read from row (x + dx, y + dy) at the thread's own position; the 18
// planes with a shift along the row go through LDS, where thread z takes element
// (z + dz) mod 256 (the periodic neighbour inside the same row) and later puts its result
// back at that element before the row is stored aligned.  Thread z owns element (z + dz) mod
// 256 of each shifted plane, a bijection, so every element is read and written once.
template <typename T>
__global__ void __launch_bounds__(256) shift_lds(T* __restrict__ f, long long n) {
    __shared__ T buf[18][256];
    const int z = (int)threadIdx.x;
    const int r = (int)blockIdx.x;  // n / 256 rows: x = r / 256, y = r % 256
    const int x = r >> 8, y = r & 255;
    T v[27];
    long long a[27];
#pragma unroll
    for (int q = 0, k = 0; q < 27; ++q) {
        const int dz = q % 3 - 1, dy = (q / 3) % 3 - 1, dx = q / 9 - 1;
        const long long rr = (((x + dx) & 255) << 8) | ((y + dy) & 255);
        a[q] = q * n + rr * 256 + z;
        const T e = __builtin_nontemporal_load(f + a[q]);
        if (dz == 0) v[q] = e;
        else buf[k++][z] = e;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0, k = 0; q < 27; ++q) {
        const int dz = q % 3 - 1;
        if (dz != 0) v[q] = buf[k++][(z + dz) & 255];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0, k = 0; q < 27; ++q) {
        const int dz = q % 3 - 1;
        if (dz != 0) buf[k++][(z + dz) & 255] = v[(q + 1) % 27];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0, k = 0; q < 27; ++q) {
        const int dz = q % 3 - 1;
        f[a[q]] = dz == 0 ? v[(q + 1) % 27] : buf[k++][z];
    }
}

// inplace27 with the work-group's loads and stores in two phases: every thread of the
// work-group has loaded its 27 elements before any of them stores (a barrier between).
template <typename T>
__global__ void __launch_bounds__(256) inplace_phased(T* __restrict__ f, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;  // n % 256 == 0
    T v[27];
#pragma unroll
    for (int q = 0; q < 27; ++q) v[q] = __builtin_nontemporal_load(f + q * n + i);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 27; ++q) f[q * n + i] = v[(q + 1) % 27];
}

// iters grid-wide barriers (launched cooperatively: every work-group is resident)
__global__ void grid_barriers(int iters, int* sink) {
    cooperative_groups::grid_group g = cooperative_groups::this_grid();
    for (int i = 0; i < iters; ++i) g.sync();
    if (iters < 0) sink[0] = 1;  // never taken: keeps the argument used
}

__global__ void empty_kernel(int* sink) {
    if (sink == nullptr) return;
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
    const double half = 2.0 * 2.0 * (double)total;
    auto best = [](double a, double c) { return a > c ? a : c; };

    if (argc > 2 && !strcmp(argv[2], "sync")) {
        int dev = 0, cus = 0, per = 0;
        CK(hipGetDevice(&dev));
        CK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        CK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, grid_barriers, 256, 0));
        const int blocks = cus * std::min(per, 8);
        const int iters = 200 * reps;
        int* isink = (int*)sink;
        hipEvent_t e0, e1;
        CK(hipEventCreate(&e0));
        CK(hipEventCreate(&e1));
        auto coop = [&](int nit) {
            void* args[] = {(void*)&nit, (void*)&isink};
            CK(hipLaunchCooperativeKernel((const void*)grid_barriers, dim3(blocks), dim3(256), args, 0, 0));
        };
        coop(10);
        CK(hipDeviceSynchronize());
        float ms1 = 0.f, ms0 = 0.f;
        CK(hipEventRecord(e0, 0));
        coop(iters);
        CK(hipEventRecord(e1, 0));
        CK(hipEventSynchronize(e1));
        CK(hipEventElapsedTime(&ms1, e0, e1));
        CK(hipEventRecord(e0, 0));
        coop(0);
        CK(hipEventRecord(e1, 0));
        CK(hipEventSynchronize(e1));
        CK(hipEventElapsedTime(&ms0, e0, e1));
        // the same number of empty launches, from a captured graph and directly on a stream
        hipStream_t st;
        CK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        const int K = 400;
        hipGraph_t gr;
        hipGraphExec_t ge;
        CK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < K; ++k) empty_kernel<<<blocks, 256, 0, st>>>(isink);
        CK(hipStreamEndCapture(st, &gr));
        CK(hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0));
        CK(hipGraphLaunch(ge, st));
        CK(hipStreamSynchronize(st));
        float msg = 0.f, mss = 0.f;
        CK(hipEventRecord(e0, st));
        for (int r = 0; r < reps; ++r) CK(hipGraphLaunch(ge, st));
        CK(hipEventRecord(e1, st));
        CK(hipEventSynchronize(e1));
        CK(hipEventElapsedTime(&msg, e0, e1));
        CK(hipEventRecord(e0, st));
        for (int r = 0; r < reps * K; ++r) empty_kernel<<<blocks, 256, 0, st>>>(isink);
        CK(hipEventRecord(e1, st));
        CK(hipEventSynchronize(e1));
        CK(hipEventElapsedTime(&mss, e0, e1));
        CK(hipGetLastError());
        printf("{\"work_groups\": %d, \"grid_barrier_us\": %.2f, \"graph_launch_us\": %.2f, "
               "\"stream_launch_us\": %.2f}\n",
               blocks, (ms1 - ms0) * 1e3 / iters, msg * 1e3 / (reps * K), mss * 1e3 / (reps * K));
        CK(hipGraphExecDestroy(ge));
        CK(hipGraphDestroy(gr));
        CK(hipStreamDestroy(st));
        CK(hipFree(f));
        CK(hipFree(sink));
        return 0;
    }

    if (argc > 2 && !strcmp(argv[2], "march")) {
        // slab s of every plane: cells [s * m, (s + 1) * m), addressed as a lattice of m cells
        // whose planes lie n apart (inplace_planes with the base moved to the slab)
        const double pair = 2.0 * lat;
        for (int nt = 1; nt >= 0; --nt) {
            auto pass = [&](long long c0, long long m) {
                if (nt) inplace_planes<27, float, true><<<grid(m), b>>>(f + c0, m, n);
                else inplace_planes<27, float, false><<<grid(m), b>>>(f + c0, m, n);
            };
            printf("%s\"%s\": {\"full\": %.1f", nt ? "{" : ", ", nt ? "streaming" : "ordinary",
                   gbs([&] { pass(0, n); pass(0, n); }, pair, reps));
            for (int K : {8, 16, 32}) {
                const long long m = n / K;  // n is a multiple of 32
                printf(", \"march%d\": %.1f", K, gbs([&] {
                    for (int s = 0; s <= K; ++s) {
                        if (s < K) pass(s * m, m);
                        if (s >= 1) pass((s - 1) * m, m);
                    }
                }, pair, reps));
                fflush(stdout);
            }
            printf("}");
        }
        printf("}\n");
        CK(hipFree(f));
        CK(hipFree(sink));
        return 0;
    }

    if (argc > 2 && !strcmp(argv[2], "cache")) {
        // planes n apart, 27 * n <= total floats for every n below; more passes on the small
        // lattices, so that each figure times at least ~2^29 cells
        printf("{\"plane_mb\": [");
        for (int k = 17; k <= 24; ++k) printf("%s%.1f", k > 17 ? ", " : "", 27.0 * 4 * (1ll << k) / 1e6);
        for (int nt = 0; nt < 2; ++nt) {
            printf("], \"%s\": [", nt ? "inplace27_streaming" : "inplace27_ordinary");
            for (int k = 17; k <= 24; ++k) {
                const long long m = 1ll << k;
                const int r = (int)(reps * std::max(1ll, (1ll << 24) / m));
                const double by = 2.0 * 4.0 * 27 * (double)m;
                const double v = nt ? gbs([&] { inplace_planes<27, float, true><<<grid(m), b>>>(f, m, m); }, by, r)
                                    : gbs([&] { inplace_planes<27, float, false><<<grid(m), b>>>(f, m, m); }, by, r);
                printf("%s%.1f", k > 17 ? ", " : "", v);
                fflush(stdout);
            }
        }
        printf("]}\n");
        CK(hipFree(f));
        CK(hipFree(sink));
        return 0;
    }

    if (argc > 2 && !strcmp(argv[2], "layout")) {
        // inplace27 and shift27 with the planes st = n + pad floats apart (the unpadded
        // stride is 2^26 bytes), and inplace27 with the planes interleaved in blocks.
        // 26 * st + n <= 2 * total floats holds for every pad below.
        const long long pads[] = {0, 64, 576, 1088, 16448, 262208};
        printf("{\"pad_bytes\": [");
        for (int k = 0; k < 6; ++k) printf("%s%lld", k ? ", " : "", pads[k] * 4);
        printf("], \"inplace27\": [");
        for (int k = 0; k < 6; ++k) {
            const long long st = n + pads[k];
            double r = best(gbs([&] { inplace_planes<27, float, false><<<grid(n), b>>>(f, n, st); }, lat, reps),
                            gbs([&] { inplace_planes<27, float, true><<<grid(n), b>>>(f, n, st); }, lat, reps));
            printf("%s%.1f", k ? ", " : "", r);
            fflush(stdout);
        }
        printf("], \"shift27\": [");
        for (int k = 0; k < 6; ++k) {
            const long long st = n + pads[k];
            double r = best(gbs([&] { shifted_planes<float, false><<<grid(n), b>>>(f, n, st); }, lat, reps),
                            gbs([&] { shifted_planes<float, true><<<grid(n), b>>>(f, n, st); }, lat, reps));
            printf("%s%.1f", k ? ", " : "", r);
            fflush(stdout);
        }
        double r64 = best(gbs([&] { inplace_blocked<27, 64, false><<<grid(n), b>>>(f, n); }, lat, reps),
                          gbs([&] { inplace_blocked<27, 64, true><<<grid(n), b>>>(f, n); }, lat, reps));
        double r256 = best(gbs([&] { inplace_blocked<27, 256, false><<<grid(n), b>>>(f, n); }, lat, reps),
                           gbs([&] { inplace_blocked<27, 256, true><<<grid(n), b>>>(f, n); }, lat, reps));
        double r1k = best(gbs([&] { inplace_blocked<27, 1024, false><<<grid(n), b>>>(f, n); }, lat, reps),
                          gbs([&] { inplace_blocked<27, 1024, true><<<grid(n), b>>>(f, n); }, lat, reps));
        // the aligned-row neighbour mix through LDS, for 4-, 2- and 8-byte elements, beside
        // the plain in-place and neighbour mixes of 8-byte elements (27 n doubles fill the
        // buffer); 256 rows x 256 cells per row x 256 = n cells
        const unsigned rows = (unsigned)(n / 256);
        double rl = gbs([&] { shift_lds<float><<<rows, 256>>>(f, n); }, lat, reps);
        double rlh = gbs([&] { shift_lds<uint16_t><<<rows, 256>>>((uint16_t*)f, n); }, half, reps);
        const double dbl = 2.0 * 8.0 * (double)total;
        double rld = gbs([&] { shift_lds<double><<<rows, 256>>>((double*)f, n); }, dbl, reps);
        double rid = best(gbs([&] { inplace_planes<27, double, false><<<grid(n), b>>>((double*)f, n, n); }, dbl, reps),
                          gbs([&] { inplace_planes<27, double, true><<<grid(n), b>>>((double*)f, n, n); }, dbl, reps));
        double rsd = best(gbs([&] { shifted_planes<double, false><<<grid(n), b>>>((double*)f, n, n); }, dbl, reps),
                          gbs([&] { shifted_planes<double, true><<<grid(n), b>>>((double*)f, n, n); }, dbl, reps));
        double rp = gbs([&] { inplace_phased<float><<<grid(n), b>>>(f, n); }, lat, reps);
        double rph = gbs([&] { inplace_phased<uint16_t><<<grid(n), b>>>((uint16_t*)f, n); }, half, reps);
        double rpd = gbs([&] { inplace_phased<double><<<grid(n), b>>>((double*)f, n); }, dbl, reps);
        printf("], \"inplace27_blocked\": {\"64\": %.1f, \"256\": %.1f, \"1024\": %.1f}, "
               "\"shift27_lds\": %.1f, \"shift27_lds_h\": %.1f, \"shift27_lds_d\": %.1f, "
               "\"inplace27_d\": %.1f, \"shift27_d\": %.1f, \"inplace27_phased\": %.1f, "
               "\"inplace27_phased_h\": %.1f, \"inplace27_phased_d\": %.1f}\n",
               r64, r256, r1k, rl, rlh, rld, rid, rsd, rp, rph, rpd);
        CK(hipFree(f));
        CK(hipFree(sink));
        return 0;
    }

    double r_copy = gbs([&] { copy4<<<grid(n4), b>>>((const float4*)f, (float4*)g, n4); }, lat, reps);
    double r_read = gbs([&] { read_planes<27><<<grid(n), b>>>(f, sink, n); }, 4.0 * total, reps);
    double r_write = gbs([&] { write_planes<27><<<grid(n), b>>>(f, n); }, 4.0 * total, reps);
    // the in-place mixes, each the better of plain and streaming loads
    double r_in = best(gbs([&] { inplace_planes<27, float, false><<<grid(n), b>>>(f, n, n); }, lat, reps),
                       gbs([&] { inplace_planes<27, float, true><<<grid(n), b>>>(f, n, n); }, lat, reps));
    double r_sh = best(gbs([&] { shifted_planes<float, false><<<grid(n), b>>>(f, n, n); }, lat, reps),
                       gbs([&] { shifted_planes<float, true><<<grid(n), b>>>(f, n, n); }, lat, reps));
    // 2-byte elements: 27 planes of n halves occupy half of the first lattice (half: bytes)
    double r_inh = best(
        gbs([&] { inplace_planes<27, uint16_t, false><<<grid(n), b>>>((uint16_t*)f, n, n); }, half, reps),
        gbs([&] { inplace_planes<27, uint16_t, true><<<grid(n), b>>>((uint16_t*)f, n, n); }, half, reps));
    double r_shh = best(
        gbs([&] { shifted_planes<uint16_t, false><<<grid(n), b>>>((uint16_t*)f, n, n); }, half, reps),
        gbs([&] { shifted_planes<uint16_t, true><<<grid(n), b>>>((uint16_t*)f, n, n); }, half, reps));
    // 9 planes over 3n cells: the same 27n floats
    double r_in9 = best(
        gbs([&] { inplace_planes<9, float, false><<<grid(3 * n), b>>>(f, 3 * n, 3 * n); }, lat, reps),
        gbs([&] { inplace_planes<9, float, true><<<grid(3 * n), b>>>(f, 3 * n, 3 * n); }, lat, reps));

    // wide accesses (vectors of VW words per lane; the 2-byte lattice is half the words)
    const long long nw = n;  // words per float plane
    double r_w2 = best(
        gbs([&] { inplace_wide<27, 2, false><<<grid(nw / 2), b>>>((uint32_t*)f, nw / 2); }, lat, reps),
        gbs([&] { inplace_wide<27, 2, true><<<grid(nw / 2), b>>>((uint32_t*)f, nw / 2); }, lat, reps));
    double r_w4 = best(
        gbs([&] { inplace_wide<27, 4, false><<<grid(nw / 4), b>>>((uint32_t*)f, nw / 4); }, lat, reps),
        gbs([&] { inplace_wide<27, 4, true><<<grid(nw / 4), b>>>((uint32_t*)f, nw / 4); }, lat, reps));
    double r_sw4 = gbs([&] { shift_wide<4, false, false><<<grid(nw / 4), b>>>((uint32_t*)f, nw / 4); },
                       lat, reps);
    double r_sw4h = gbs([&] { shift_wide<4, false, true><<<grid(nw / 4), b>>>((uint32_t*)f, nw / 4); },
                        lat, reps);
    // 2 cells per lane: rows of 128 cells, 512 x 256 of them
    double r_sw2 = gbs([&] { shift_wide<2, false, false><<<grid(nw / 2), b>>>((uint32_t*)f, nw / 2); },
                       lat, reps);
    double r_sw2h = gbs([&] { shift_wide<2, false, true><<<grid(nw / 2), b>>>((uint32_t*)f, nw / 2); },
                        lat, reps);
    // 2-byte elements, 4 per lane: the 256-cell rows are 128 words, 2 words per lane
    double r_hw4 = best(
        gbs([&] { inplace_wide<27, 2, false><<<grid(nw / 4), b>>>((uint32_t*)f, nw / 4); }, half, reps),
        gbs([&] { inplace_wide<27, 2, true><<<grid(nw / 4), b>>>((uint32_t*)f, nw / 4); }, half, reps));
    double r_shw4 = gbs([&] { shift_wide<2, true, false><<<grid(nw / 4), b>>>((uint32_t*)f, nw / 4); },
                        half, reps);

    hipDeviceProp_t p;
    CK(hipGetDeviceProperties(&p, 0));
    printf("{\"device\": \"%s\", \"bytes_per_pass\": %.0f, \"copy4\": %.1f, \"read27\": %.1f, "
           "\"write27\": %.1f, \"inplace27\": %.1f, \"shift27\": %.1f, \"inplace27_h\": %.1f, "
           "\"shift27_h\": %.1f, \"inplace9\": %.1f, \"inplace27_w2\": %.1f, "
           "\"inplace27_w4\": %.1f, \"shift27_w4\": %.1f, \"shift27_w4h\": %.1f, "
           "\"inplace27_hw4\": %.1f, \"shift27_hw4\": %.1f, \"shift27_w2\": %.1f, "
           "\"shift27_w2h\": %.1f}\n",
           p.gcnArchName, lat, r_copy, r_read, r_write, r_in, r_sh, r_inh, r_shh, r_in9, r_w2, r_w4,
           r_sw4, r_sw4h, r_hw4, r_shw4, r_sw2, r_sw2h);
    CK(hipFree(f));
    CK(hipFree(sink));
    return 0;
}
