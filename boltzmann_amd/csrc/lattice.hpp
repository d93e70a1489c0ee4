// Velocity sets, storage formats and per-cell collision arithmetic shared by every
// lattice-update kernel.  Header-only, compiled for gfx950 by hipcc.
//
// Velocity ordering and weights are the reference's (velocities.rs:66-191, the qs tables
// of lib.cl): slot 0 is the rest population and slots (2k-1, 2k) are opposite pairs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bolt {

// ---------------------------------------------------------------------------------------
// Velocity sets
// ---------------------------------------------------------------------------------------
template <int D, int Q>
struct VSet;

#define BOLT_VSET_COMMON                                                                   \
    __host__ __device__ static constexpr int opp(int i) {                                  \
        return i == 0 ? 0 : ((i & 1) ? i + 1 : i - 1);                                     \
    }                                                                                      \
    __host__ __device__ static constexpr int nnz(int i) {                                  \
        return (c(i, 0) != 0) + (c(i, 1) != 0) + (c(i, 2) != 0);                           \
    }                                                                                      \
    __host__ __device__ static constexpr float w(int i) {                                  \
        return (float)((double)wn(i) / (double)wd(i));                                     \
    }

template <>
struct VSet<2, 9> {
    static constexpr int d = 2, q = 9;
    __host__ __device__ static constexpr int c(int i, int a) {
        constexpr int t[9][3] = {{0, 0, 0}, {1, 0, 0},   {-1, 0, 0}, {0, 1, 0}, {0, -1, 0},
                                 {1, 1, 0}, {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}};
        return t[i][a];
    }
    __host__ __device__ static constexpr int wn(int i) { return i == 0 ? 4 : 1; }
    __host__ __device__ static constexpr int wd(int i) { return i == 0 ? 9 : (i < 5 ? 9 : 36); }
    BOLT_VSET_COMMON
};

template <>
struct VSet<2, 5> {
    static constexpr int d = 2, q = 5;
    __host__ __device__ static constexpr int c(int i, int a) {
        constexpr int t[5][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}};
        return t[i][a];
    }
    __host__ __device__ static constexpr int wn(int i) { return 1; }
    __host__ __device__ static constexpr int wd(int i) { return i == 0 ? 3 : 6; }
    BOLT_VSET_COMMON
};

template <>
struct VSet<3, 7> {
    static constexpr int d = 3, q = 7;
    __host__ __device__ static constexpr int c(int i, int a) {
        constexpr int t[7][3] = {{0, 0, 0},  {1, 0, 0}, {-1, 0, 0}, {0, 1, 0},
                                 {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
        return t[i][a];
    }
    __host__ __device__ static constexpr int wn(int i) { return i == 0 ? 1 : 19; }
    __host__ __device__ static constexpr int wd(int i) { return i == 0 ? 20 : 120; }
    BOLT_VSET_COMMON
};

template <>
struct VSet<3, 27> {
    static constexpr int d = 3, q = 27;
    __host__ __device__ static constexpr int c(int i, int a) {
        constexpr int t[27][3] = {
            {0, 0, 0},   {1, 0, 0},   {-1, 0, 0},   {0, 1, 0},   {0, -1, 0},  {0, 0, 1},
            {0, 0, -1},  {1, 1, 0},   {-1, -1, 0},  {1, -1, 0},  {-1, 1, 0},  {1, 0, 1},
            {-1, 0, -1}, {0, 1, 1},   {0, -1, -1},  {1, 0, -1},  {-1, 0, 1},  {0, 1, -1},
            {0, -1, 1},  {1, 1, 1},   {-1, -1, -1}, {1, -1, 1},  {-1, 1, -1}, {1, 1, -1},
            {-1, -1, 1}, {1, -1, -1}, {-1, 1, 1}};
        return t[i][a];
    }
    __host__ __device__ static constexpr int wn(int i) {
        return i == 0 ? 8 : (i < 7 ? 2 : 1);
    }
    __host__ __device__ static constexpr int wd(int i) {
        return i == 0 ? 27 : (i < 7 ? 27 : (i < 19 ? 54 : 216));
    }
    BOLT_VSET_COMMON
};

#undef BOLT_VSET_COMMON

// Order in which the reference sums c_i f_i for the velocity moments: the lexicographic
// order of the names "f_[i]" (its generator printed sympy terms that way, lib.cl:509-511).
template <int Q>
__host__ __device__ constexpr int vel_order(int k) {
    if (Q == 9) return k;  // single digits: numeric order
    // "f_[0]" < "f_[10]" < ... < "f_[19]" < "f_[1]" < "f_[20]" < ... < "f_[2]" < "f_[3]" ...
    constexpr int t27[27] = {0,  10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 1, 20, 21,
                             22, 23, 24, 25, 26, 2,  3,  4,  5,  6,  7,  8, 9};
    return t27[k];
}

// ---------------------------------------------------------------------------------------
// Storage formats of the population arrays.
//   F32   : float, the reference's format
//   F16S  : binary16 holding f_i - w_i ("shifted"), arithmetic in float
//   BF16S : bfloat16 holding f_i - w_i, arithmetic in float
//   F64   : double, arithmetic in double (the comparison run)
//   F32F  : float storage with plain float arithmetic (kF32 evaluates the equilibrium the
//           way the reference does, in double, narrowed to float)
// ---------------------------------------------------------------------------------------
enum Storage : int { kF32 = 0, kF16S = 1, kBF16S = 2, kF64 = 3, kF32F = 4 };

template <int S>
struct Store;

template <>
struct Store<kF32> {
    using T = float;
    using R = float;
    __device__ static R load(const T* p, float) { return *p; }
    __device__ static void store(T* p, R v, float) { *p = v; }
};

template <>
struct Store<kF32F> : Store<kF32> {};

template <>
struct Store<kF64> {
    using T = double;
    using R = double;
    __device__ static R load(const T* p, float) { return *p; }
    __device__ static void store(T* p, R v, float) { *p = v; }
};

template <>
struct Store<kF16S> {
    using T = _Float16;
    using R = float;
    __device__ static R load(const T* p, float w) { return (float)(*p) + w; }
    __device__ static void store(T* p, R v, float w) { *p = (_Float16)(v - w); }
};

// bfloat16 bits with round-to-nearest-even on store (NaN kept quiet), by gfx950's conversion
// instruction (v_cvt_pk_bf16_f32).  On MI355X it gives the same bits as the integer form
// u + 0x7fff + ((u >> 16) & 1) >> 16 (with NaN -> (u >> 16) | 0x40) for all 2^32 float
// patterns, in one instruction where that form takes five; the bfloat16-shifted steps run
// 1.4-7.6 % faster with it.
__device__ inline uint16_t f32_to_bf16_rne(float v) {
    return __builtin_bit_cast(uint16_t, (__bf16)v);
}

template <>
struct Store<kBF16S> {
    using T = uint16_t;
    using R = float;
    __device__ static R load(const T* p, float w) {
        return __uint_as_float(((uint32_t)(*p)) << 16) + w;
    }
    __device__ static void store(T* p, R v, float w) { *p = f32_to_bf16_rne(v - w); }
};

}  // namespace bolt
