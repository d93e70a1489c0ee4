// Wide AA update: VW consecutive cells of the contiguous axis per lane, every population
// access an aligned 8- or 16-byte vector.
//
// Layout and physics are those of update.hpp (structure-of-arrays planes, the reference's
// row-major cell order, the same per-cell arithmetic, so results are bit-identical to the
// one-cell-per-lane kernel and to the oracle).  What changes is the access:
//
//   * a wavefront owns 64*VW consecutive cells of one row (row = every index but the
//     contiguous one); lane l owns cells zb + VW*l .. zb + VW*l + VW-1;
//   * odd step: every access is the lane's own aligned vector of a plane;
//   * even step: plane i is read and written at the neighbour x - c_i.  The row part of
//     that shift only changes the row.  The +-1 part along the row is done in registers:
//     the wavefront loads the aligned vectors of the neighbour row and moves each lane's
//     edge element to the next (or previous) lane with a one-lane rotation across the
//     wavefront (DPP wave_ror:1 / wave_rol:1).  Stores take the reverse path.
//   * when a row is longer than one wavefront (or shorter), the element just across the
//     wavefront's edge belongs to the wavefront itself in the AA sense (cell zb reads and
//     writes F[zb - 1]), so its first (last) active lane loads and stores that one element
//     directly and the edge lane stores VW - 1 elements of its vector.  When the row is
//     exactly one wavefront long (L = 64*VW) the rotation itself wraps periodically and
//     every access is a full vector.
//
// Wall cells are loaded and written back unchanged (the collision is replaced by the
// bounce-back permutation f*_i = f_opp(i), which stores back the values loaded from the
// same locations); no other cell touches those locations in that step.
#pragma once

#include "update.hpp"

namespace bolt {

// ---- launch geometry (host and device) --------------------------------------------------
struct WideGeom {
    int len;           // cells along the contiguous axis (a multiple of vw)
    int vw;            // cells per lane
    int wpr;           // wavefronts per row
    int fullrow;       // len == 64 * vw: one wavefront per row, the shift is a rotation
    long long rows;
    long long tasks;   // rows * wpr wavefronts
    int nblocks;       // work-groups of four wavefronts
};

__host__ __device__ inline WideGeom make_wide_geom(long long ncell, int len, int vw) {
    WideGeom g;
    g.len = len;
    g.vw = vw;
    g.wpr = (len + 64 * vw - 1) / (64 * vw);
    g.fullrow = len == 64 * vw;
    g.rows = ncell / len;
    g.tasks = g.rows * g.wpr;
    g.nblocks = (int)((g.tasks + 3) / 4);
    return g;
}

// the wavefront's row, the first cell of its range and its last active lane
struct WaveTask {
    uint32_t row;
    int zb;
    int last;
};

__host__ __device__ inline WaveTask wave_task(const WideGeom& g, long long task) {
    WaveTask t;
    t.row = (uint32_t)(task / g.wpr);
    const int w = (int)(task - (long long)t.row * g.wpr);
    t.zb = w * 64 * g.vw;
    const int n = (g.len - t.zb) / g.vw;
    t.last = (n < 64 ? n : 64) - 1;
    return t;
}

// element of the row just before / just after the wavefront's range (periodic)
__host__ __device__ inline int halo_prev(const WideGeom& g, const WaveTask& t) {
    return t.zb == 0 ? g.len - 1 : t.zb - 1;
}
__host__ __device__ inline int halo_next(const WideGeom& g, const WaveTask& t) {
    const int e = t.zb + g.vw * (t.last + 1);
    return e == g.len ? 0 : e;
}

// neighbour row index: row = ix (2D) or ix * n1 + iy (3D); (dx, dy) periodic
template <int D>
__host__ __device__ inline uint32_t row_shift(uint32_t row, int dx, int dy, int n0, int n1) {
    if (D == 2) {
        int x = (int)row + dx;
        x = x < 0 ? x + n0 : (x >= n0 ? x - n0 : x);
        return (uint32_t)x;
    } else {
        const int ix = (int)(row / (uint32_t)n1), iy = (int)(row - (uint32_t)ix * (uint32_t)n1);
        int x = ix + dx, y = iy + dy;
        x = x < 0 ? x + n0 : (x >= n0 ? x - n0 : x);
        y = y < 0 ? y + n1 : (y >= n1 ? y - n1 : y);
        return (uint32_t)x * (uint32_t)n1 + (uint32_t)y;
    }
}

// ---- per-lane vectors of raw words ------------------------------------------------------
template <int W>
struct RawV;
template <>
struct RawV<1> {
    typedef uint32_t T;
};
template <>
struct RawV<2> {
    typedef uint32_t T __attribute__((ext_vector_type(2)));
};
template <>
struct RawV<4> {
    typedef uint32_t T __attribute__((ext_vector_type(4)));
};

template <int W>
struct Words {
    uint32_t w[W];
};

// STREAM: streaming or ordinary load, as update.hpp's pop_load
template <int W, bool STREAM>
__device__ inline Words<W> vload(const void* plane, uint32_t elem_byte) {
    using R = typename RawV<W>::T;
    const R* const a = (const R*)((const char*)plane + elem_byte);
    R r;
    if constexpr (STREAM) r = __builtin_nontemporal_load(a);
    else r = *a;
    Words<W> o;
    if constexpr (W == 1) o.w[0] = r;
    else
#pragma unroll
        for (int j = 0; j < W; ++j) o.w[j] = r[j];
    return o;
}

template <int W>
__device__ inline void vstore(void* plane, uint32_t elem_byte, const Words<W>& v) {
    using R = typename RawV<W>::T;
    R r;
    if constexpr (W == 1) r = v.w[0];
    else
#pragma unroll
        for (int j = 0; j < W; ++j) r[j] = v.w[j];
    *(R*)((char*)plane + elem_byte) = r;
}

// One-lane rotations across the 64-lane wavefront: lane l receives x of lane l-1 (ror1) or
// of lane l+1 (rol1), lane 0 from lane 63 and lane 63 from lane 0.
__device__ inline uint32_t ror1(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x13C, 0xF, 0xF, false);
}
__device__ inline uint32_t rol1(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x134, 0xF, 0xF, false);
}

// Element k of ES-byte elements packed in words (little-endian: element 0 first).
template <int ES, int W>
__device__ inline uint64_t get_el(const Words<W>& v, int k) {
    if constexpr (ES == 2) return (v.w[k >> 1] >> ((k & 1) * 16)) & 0xffffu;
    else if constexpr (ES == 4) return v.w[k];
    else return (uint64_t)v.w[2 * k] | ((uint64_t)v.w[2 * k + 1] << 32);
}
template <int ES, int W>
__device__ inline void set_el(Words<W>& v, int k, uint64_t x) {
    if constexpr (ES == 2) {
        const int s = (k & 1) * 16;
        v.w[k >> 1] = (v.w[k >> 1] & ~(0xffffu << s)) | ((uint32_t)(x & 0xffffu) << s);
    } else if constexpr (ES == 4) {
        v.w[k] = (uint32_t)x;
    } else {
        v.w[2 * k] = (uint32_t)x;
        v.w[2 * k + 1] = (uint32_t)(x >> 32);
    }
}

// element k of the result = element k-1 of the lane's vector; element 0 = the previous
// lane's last element
template <int ES, int W>
__device__ inline void shift_prev(Words<W>& v) {
    if constexpr (ES == 2) {
        const uint32_t prev = ror1(v.w[W - 1]);
#pragma unroll
        for (int j = W - 1; j > 0; --j) v.w[j] = __builtin_amdgcn_alignbyte(v.w[j], v.w[j - 1], 2);
        v.w[0] = __builtin_amdgcn_alignbyte(v.w[0], prev, 2);
    } else {
        constexpr int K = ES / 4;  // words per element
        uint32_t prev[K];
#pragma unroll
        for (int j = 0; j < K; ++j) prev[j] = ror1(v.w[W - K + j]);
#pragma unroll
        for (int j = W - 1; j >= K; --j) v.w[j] = v.w[j - K];
#pragma unroll
        for (int j = 0; j < K; ++j) v.w[j] = prev[j];
    }
}

// element k of the result = element k+1; the last element = the next lane's element 0
template <int ES, int W>
__device__ inline void shift_next(Words<W>& v) {
    if constexpr (ES == 2) {
        const uint32_t next = rol1(v.w[0]);
#pragma unroll
        for (int j = 0; j < W - 1; ++j) v.w[j] = __builtin_amdgcn_alignbyte(v.w[j + 1], v.w[j], 2);
        v.w[W - 1] = __builtin_amdgcn_alignbyte(next, v.w[W - 1], 2);
    } else {
        constexpr int K = ES / 4;
        uint32_t next[K];
#pragma unroll
        for (int j = 0; j < K; ++j) next[j] = rol1(v.w[j]);
#pragma unroll
        for (int j = 0; j < W - K; ++j) v.w[j] = v.w[j + K];
#pragma unroll
        for (int j = 0; j < K; ++j) v.w[W - K + j] = next[j];
    }
}

template <int ES>
__device__ inline uint64_t load_el(const void* plane, uint32_t elem) {
    const char* p = (const char*)plane + elem * (uint32_t)ES;
    if constexpr (ES == 2) return *(const uint16_t*)p;
    else if constexpr (ES == 4) return *(const uint32_t*)p;
    else return *(const uint64_t*)p;
}
template <int ES>
__device__ inline void store_el(void* plane, uint32_t elem, uint64_t x) {
    char* p = (char*)plane + elem * (uint32_t)ES;
    if constexpr (ES == 2) *(uint16_t*)p = (uint16_t)x;
    else if constexpr (ES == 4) *(uint32_t*)p = (uint32_t)x;
    else *(uint64_t*)p = x;
}

// ---- conversions between stored bits and the arithmetic type ----------------------------
template <int S>
struct WideConv;
template <>
struct WideConv<kF32> {
    using A = float;
    __device__ static A in(uint64_t b) { return __uint_as_float((uint32_t)b); }
    __device__ static uint64_t out(A v) { return __float_as_uint(v); }
};
template <>
struct WideConv<kF32F> : WideConv<kF32> {};
template <>
struct WideConv<kF16S> {
    using A = float;
    __device__ static A in(uint64_t b) { return (float)__builtin_bit_cast(_Float16, (uint16_t)b); }
    __device__ static uint64_t out(A v) { return __builtin_bit_cast(uint16_t, (_Float16)v); }
};
template <>
struct WideConv<kBF16S> {
    using A = float;
    __device__ static A in(uint64_t b) { return __uint_as_float(((uint32_t)b) << 16); }
    __device__ static uint64_t out(A v) { return f32_to_bf16_rne(v); }
};
template <>
struct WideConv<kF64> {
    using A = double;
    __device__ static A in(uint64_t b) { return __builtin_bit_cast(double, b); }
    __device__ static uint64_t out(A v) { return __builtin_bit_cast(uint64_t, v); }
};

// Where the lane's accesses of one plane go in this step.
struct WideLane {
    int lane;
    bool active;       // the lane owns cells of the row
    uint32_t row;      // the lane's row
    int z0;            // first cell of the lane along the row
    uint32_t rb[3][3]; // first element of the row shifted by (dx, dy) = (a - 1, b - 1)
};

// Loads plane `pl` for this step as the lane's packed vector: element k is the value cell
// z0 + k reads -- even step, the plane's slot at x - c (row shift rdx, rdy and shift
// dz = -c along the row); odd step, dz = 0 on the own row.
template <int ES, int VW, bool STREAM>
__device__ inline Words<VW * ES / 4> wide_load(const void* pl, const WideGeom& g, const WaveTask& t,
                                               const WideLane& l, int rdx, int rdy, int dz) {
    constexpr int W = VW * ES / 4;
    const uint32_t rb = l.rb[rdx + 1][rdy + 1];
    // a lane past the end of the row loads the wavefront's first vector instead (its values
    // are never used), so the loads need no branch
    Words<W> v = vload<W, STREAM>(pl, (rb + (uint32_t)(l.active ? l.z0 : t.zb)) * (uint32_t)ES);
    if (dz < 0) {
        shift_prev<ES, W>(v);
        if (!g.fullrow && l.lane == 0) set_el<ES, W>(v, 0, load_el<ES>(pl, rb + (uint32_t)halo_prev(g, t)));
    } else if (dz > 0) {
        shift_next<ES, W>(v);
        if (!g.fullrow && l.lane == t.last)
            set_el<ES, W>(v, VW - 1, load_el<ES>(pl, rb + (uint32_t)halo_next(g, t)));
    }
    return v;
}

// Stores the packed vector v (element k: the value cell z0 + k sends to element z0 + k + dz
// of the shifted row).
template <int ES, int VW>
__device__ inline void wide_store(void* pl, const WideGeom& g, const WaveTask& t, const WideLane& l,
                                  int rdx, int rdy, int dz, Words<VW * ES / 4> v) {
    constexpr int W = VW * ES / 4;
    const uint32_t rb = l.rb[rdx + 1][rdy + 1];
    const uint32_t at = rb + (uint32_t)l.z0;
    if (dz == 0) {
        if (l.active) vstore<W>(pl, at * (uint32_t)ES, v);
        return;
    }
    if (dz < 0) {
        // element e of the vector receives cell e + 1's value: the lane's first value goes
        // to the element before its range, the last element comes from the next lane
        const uint64_t edge = get_el<ES, W>(v, 0);
        shift_next<ES, W>(v);
        if (g.fullrow) {
            vstore<W>(pl, at * (uint32_t)ES, v);
            return;
        }
        if (l.lane < t.last) vstore<W>(pl, at * (uint32_t)ES, v);
        if (l.lane == t.last)
#pragma unroll
            for (int k = 0; k < VW - 1; ++k) store_el<ES>(pl, at + (uint32_t)k, get_el<ES, W>(v, k));
        if (l.lane == 0) store_el<ES>(pl, rb + (uint32_t)halo_prev(g, t), edge);
    } else {
        // element e receives cell e - 1's value: the lane's last value goes to the element
        // after its range, the first element comes from the previous lane
        const uint64_t edge = get_el<ES, W>(v, VW - 1);
        shift_prev<ES, W>(v);
        if (g.fullrow) {
            vstore<W>(pl, at * (uint32_t)ES, v);
            return;
        }
        if (l.lane > 0 && l.lane <= t.last) vstore<W>(pl, at * (uint32_t)ES, v);
        if (l.lane == 0 && l.active)
#pragma unroll
            for (int k = 1; k < VW; ++k) store_el<ES>(pl, at + (uint32_t)k, get_el<ES, W>(v, k));
        if (l.lane == t.last) store_el<ES>(pl, rb + (uint32_t)halo_next(g, t), edge);
    }
}

// row shift (rdx, rdy) and shift along the row dz of plane i's location in this step
template <int D, class V, bool EVEN>
__device__ constexpr void plane_shift(int i, int& rdx, int& rdy, int& dz) {
    if (!EVEN) {
        rdx = rdy = dz = 0;
        return;
    }
    rdx = -V::c(i, 0);
    rdy = D == 3 ? -V::c(i, 1) : 0;
    dz = -V::c(i, D - 1);
}

// Deviation f_i - w_i of a loaded value (kF32F stores f_i; the shifted formats store the
// deviation itself) and back.
template <class V, int S>
__device__ inline float to_dev(float x, int i) {
#pragma clang fp contract(off)
    return S == kF32F ? x - V::w(i) : x;
}
template <class V, int S>
__device__ inline float from_dev(float d, int i) {
#pragma clang fp contract(off)
    return S == kF32F ? d + V::w(i) : d;
}

// The collision of every opposite pair (I, opp I), each followed by the stores of planes I
// and opp I.  The populations stay packed as loaded (fr[i], VW elements of the storage
// format) and are decoded and re-encoded per use, which keeps 2-byte formats at half the
// registers.  Per cell: r, u, vv and the two rates from the preparation pass (r holds
// rho - 1 on the deviation path); wall cells swap the raw elements of each pair (the
// bounce-back permutation of what they loaded).
template <int D, int Q, int COLL, int S, bool EVEN, int VW, int I = 0>
__device__ inline void collide_store_wide(Words<VW * sizeof(typename Store<S>::T) / 4> (&fr)[Q],
                                          const typename WideConv<S>::A (&r)[VW],
                                          const typename WideConv<S>::A (&u)[VW][3],
                                          const typename WideConv<S>::A (&vv)[VW],
                                          const typename WideConv<S>::A (&omp)[VW],
                                          const typename WideConv<S>::A (&omn)[VW],
                                          const bool (&wall)[VW], const StepParams& p,
                                          const WideGeom& g, const WaveTask& t, const WideLane& l) {
    using V = VSet<D, Q>;
    using A = typename WideConv<S>::A;
    constexpr int ES = sizeof(typename Store<S>::T), W = VW * ES / 4;
    if constexpr (I < Q) {
        constexpr int J = V::opp(I);
        if constexpr (J >= I) {
#pragma unroll
            for (int k = 0; k < VW; ++k) {
                const uint64_t bi = get_el<ES, W>(fr[I], k), bj = get_el<ES, W>(fr[J], k);
                if (wall[k]) {
                    if constexpr (J != I) {
                        set_el<ES, W>(fr[I], k, bj);
                        set_el<ES, W>(fr[J], k, bi);
                    }
                    continue;
                }
                A fk[Q] = {};
                fk[I] = WideConv<S>::in(bi);
                fk[J] = WideConv<S>::in(bj);
                if constexpr (S == kF32) {
                    ref_pair<V, COLL, I>(fk, u[k], vv[k], r[k], omp[k], omn[k]);
                } else if constexpr (S == kF64) {
                    f64_pair<V, COLL, I>(fk, r[k], u[k], vv[k], omp[k], omn[k]);
                } else {
#pragma clang fp contract(off)
                    fk[I] = to_dev<V, S>(fk[I], I);
                    if constexpr (J != I) fk[J] = to_dev<V, S>(fk[J], J);
                    dev_pair<V, COLL, I>(fk, u[k], -1.5f * vv[k], r[k], 1.0f + r[k], omp[k], omn[k]);
                    fk[I] = from_dev<V, S>(fk[I], I);
                    if constexpr (J != I) fk[J] = from_dev<V, S>(fk[J], J);
                }
                set_el<ES, W>(fr[I], k, WideConv<S>::out(fk[I]));
                set_el<ES, W>(fr[J], k, WideConv<S>::out(fk[J]));
            }
            int rdx, rdy, dz;
            // even: F[x - c_i, i] = f*_opp(i); odd: F[x, i] = f*_i
            plane_shift<D, V, EVEN>(I, rdx, rdy, dz);
            wide_store<ES, VW>(p.fpl[I], g, t, l, rdx, rdy, dz, EVEN ? fr[J] : fr[I]);
            if constexpr (J != I) {
                plane_shift<D, V, EVEN>(J, rdx, rdy, dz);
                wide_store<ES, VW>(p.fpl[J], g, t, l, rdx, rdy, dz, EVEN ? fr[I] : fr[J]);
            }
        }
        collide_store_wide<D, Q, COLL, S, EVEN, VW, I + 1>(fr, r, u, vv, omp, omn, wall, p, g, t, l);
    }
}

template <int Q, int S, int VW>
constexpr int wide_waves_target() {
#ifdef BOLT_WIDE_WAVES
    return BOLT_WIDE_WAVES;
#else
    return (Q * VW * (S == kF64 ? 2 : 1)) > 64 ? 2 : 4;
#endif
}

// The fluid step with VW cells per lane (no fused tracer); STREAM: the population load form.
template <int D, int Q, int COLL, int S, bool EVEN, int VW, bool STREAM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(wide_waves_target<Q, S, VW>())))
    lbm_step_wide(const StepParams p, const WideGeom g) {
    static_assert(VW == 2 || VW == 4, "2 or 4 cells per lane");
    using V = VSet<D, Q>;
    using A = typename WideConv<S>::A;
    // the wavefront's task, row and neighbour rows are the same on every lane: computed
    // once from a wavefront-uniform index (scalar registers)
    int task = __builtin_amdgcn_readfirstlane(
        xcd_swizzle((int)blockIdx.x, g.nblocks) * 4 + (int)(threadIdx.x >> 6));
    if (task >= g.tasks) return;  // a whole wavefront
    if (p.reverse) task = g.tasks - 1 - task;
    const WaveTask t = wave_task(g, task);
    WideLane l;
    l.lane = (int)(threadIdx.x & 63);
    l.active = l.lane <= t.last;
    l.row = t.row;
    l.z0 = t.zb + VW * l.lane;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
            l.rb[a][b] = (D == 3 || b == 1) ? row_shift<D>(t.row, a - 1, b - 1, p.n0, p.n1) * (uint32_t)g.len : 0u;
    const uint32_t cell0 = t.row * (uint32_t)g.len + (uint32_t)l.z0;

    uint8_t fl[VW];
    if (l.active) {
        Words<(VW + 3) / 4> fw;
        if constexpr (VW == 4) fw.w[0] = *(const uint32_t*)(p.flags + cell0);
        else if constexpr (VW == 2) fw.w[0] = *(const uint16_t*)(p.flags + cell0);
        else fw.w[0] = p.flags[cell0];
#pragma unroll
        for (int k = 0; k < VW; ++k) fl[k] = (uint8_t)(fw.w[0] >> (8 * k));
    } else {
#pragma unroll
        for (int k = 0; k < VW; ++k) fl[k] = 1;  // nothing to compute or write
    }
    bool wall[VW], fixed[VW];
#pragma unroll
    for (int k = 0; k < VW; ++k) {
        wall[k] = (fl[k] & 1) != 0;
        fixed[k] = !wall[k] && (fl[k] & 6) != 0;
    }

    constexpr int ES = sizeof(typename Store<S>::T), W = VW * ES / 4;
    Words<W> fr[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        // even: f_i = F[x - c_i, i]; odd: f_i = F[x, opp i]
        int rdx, rdy, dz;
        plane_shift<D, V, EVEN>(i, rdx, rdy, dz);
        fr[i] = wide_load<ES, VW, STREAM>(p.fpl[EVEN ? i : V::opp(i)], g, t, l, rdx, rdy, dz);
    }

    A r[VW], u[VW][3], vv[VW], omp[VW], omn[VW];
#pragma unroll
    for (int k = 0; k < VW; ++k) {
        const uint32_t cell = cell0 + (uint32_t)k;
        if (wall[k]) {  // bounce-back only (also every cell of a lane past the row's end)
            r[k] = vv[k] = omp[k] = omn[k] = 0;
            u[k][0] = u[k][1] = u[k][2] = 0;
            continue;
        }
        if constexpr (S == kF64) {
            double rfix = 0.0, ufix[3] = {0, 0, 0}, acc[3] = {0, 0, 0};
            if (fixed[k]) {
                rfix = elem<double>((const void*)p.rho, cell);
#pragma unroll
                for (int a = 0; a < D; ++a) ufix[a] = elem<double>((const void*)p.vel[a], cell);
            }
            if (p.acc_mode == kAccConst) {
#pragma unroll
                for (int a = 0; a < D; ++a) acc[a] = (double)p.g[a];
            } else if (p.acc_mode == kAccBoussinesq) {
                const double b = (double)p.alpha * (elem<double>(p.bq_C, cell) - (double)p.c0);
#pragma unroll
                for (int a = 0; a < D; ++a) acc[a] = (double)p.g[a] * b;
            }
            double fk[Q];
#pragma unroll
            for (int i = 0; i < Q; ++i) fk[i] = WideConv<S>::in(get_el<ES, W>(fr[i], k));
            f64_prepare<V>(fk, r[k], u[k], vv[k], omp[k], omn[k], fixed[k], p, rfix, ufix, acc,
                           p.acc_mode != kAccNone);
        } else {
            u[k][0] = u[k][1] = u[k][2] = 0.f;
            omp[k] = p.om_p;
            omn[k] = p.om_n;
            float fk[Q];
#pragma unroll
            for (int i = 0; i < Q; ++i) fk[i] = WideConv<S>::in(get_el<ES, W>(fr[i], k));
            if constexpr (S == kF32) {
                ref_moments<V>(fk, r[k], u[k]);
            } else {
#pragma unroll
                for (int i = 0; i < Q; ++i) fk[i] = to_dev<V, S>(fk[i], i);
                dev_moments<V>(fk, r[k], u[k]);  // r holds rho - 1
            }
            if (p.acc_mode != kAccNone) {
                float acc[3];
                if (p.acc_mode == kAccConst) {
#pragma unroll
                    for (int a = 0; a < D; ++a) acc[a] = p.g[a];
                } else {
                    const float b = p.alpha * (elem<float>(p.bq_C, cell) - p.c0);
#pragma unroll
                    for (int a = 0; a < D; ++a) acc[a] = p.g[a] * b;
                }
#pragma unroll
                for (int a = 0; a < D; ++a) u[k][a] = u[k][a] + acc[a] / omp[k];
            }
            if (fixed[k]) {
                omp[k] = 1.0f;
                omn[k] = 1.0f;
                r[k] = elem<float>((const void*)p.rho, cell);
                if constexpr (S != kF32) r[k] = r[k] - 1.0f;
#pragma unroll
                for (int a = 0; a < D; ++a) u[k][a] = elem<float>((const void*)p.vel[a], cell);
            }
            if constexpr (S == kF32) vv[k] = ref_vv<D>(u[k]);
            else vv[k] = dev_vv(u[k]);
        }
    }
    collide_store_wide<D, Q, COLL, S, EVEN, VW>(fr, r, u, vv, omp, omn, wall, p, g, t, l);
    if (p.write_macro) {
#pragma unroll
        for (int k = 0; k < VW; ++k) {
            if (!l.active || wall[k] || fixed[k]) continue;
            const uint32_t cell = cell0 + (uint32_t)k;
            elem<A>(p.rho, cell) = (S == kF32 || S == kF64) ? r[k] : (A)(r[k] + 1.0f);
#pragma unroll
            for (int a = 0; a < D; ++a) elem<A>(p.vel[a], cell) = u[k][a];
        }
    }
}

}  // namespace bolt
