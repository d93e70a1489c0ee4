// AA-pattern in-place stream + collide for gfx950.
//
// One work-item per cell.  Populations are structure-of-arrays: plane i holds slot i of
// every cell, f[i * stride + cell], with cell = (ix*ny + iy)*nz + iz (2D: ix*ny + iy), the
// reference's row-major cell order (lib.cl:204,426).  The last axis is contiguous, so the
// 64 lanes of a wavefront read and write 64 consecutive cells of one plane.
//
// AA addressing (lib.cl:226-246, 276-296):
//   even step  load  f_i  = F[x - c_i, i]          store F[x - c_i, i]  = f*_{opp i}
//   odd step   load  f_i  = F[x, opp i]            store F[x, i]        = f*_i
// Every work-item reads and writes the same set of locations and the sets of different
// cells are disjoint, so one launch updates the array in place with no ordering between
// work-items.  Wall cells (flag & 1) neither read nor write (full-way bounce-back).  The
// even step also has a row-staged form (lbm_step_row, below): one work-group per row, every
// access an aligned row, the shift along the row taken in LDS, with the same bits.
//
// Per-cell physics (lib.cl:249-300, 680-717; kernel_gen.py:402-420 for D3Q27-TRT):
//   rho = sum f, u = sum c f / rho, u += a / omega(+) when an acceleration is active,
//   fixed fluid cells (flag & 6) relax fully (omega = 1) towards feq(rho[x], vel[x]),
//   BGK:  f += omega (feq - f);   TRT: symmetric half with omega+, antisymmetric with omega-.
// Advection-diffusion tracer (lib.cl:102-189, 303-405): rho = sum g, u = the fluid's u of
// the same step, fixed scalar cells (flag & 8) relax fully towards geq(C[x], u).
//
// Arithmetic: storage kF32 evaluates each equilibrium polynomial, the relaxation and the
// population update in double and narrows to float exactly where the reference's kernels
// do (its unsuffixed literals are double); the operation order is the reference's and
// multiply-adds are not contracted, so results are bit-identical to the CPU restatement.
// Storage kF32F/kF16S/kBF16S use float arithmetic on the deviations f_i - w_i; kF64 uses
// double throughout.
#pragma once

#include "lattice.hpp"

namespace bolt {

// Kernel arguments.  Every array is addressed as a uniform base pointer plus a 32-bit
// per-cell byte offset (cell * sizeof(element) < 2^32, enforced at lattice creation), so
// loads and stores use the scalar-base form of the global memory instructions and keep no
// 64-bit address in vector registers.
struct StepParams {
    int n0, n1, n2;      // counts; 2D runs have n2 == 1 and the contiguous axis is n1
    int len;             // length of the contiguous axis
    int cpr;             // work-groups per row of the contiguous axis
    int nblocks;         // work-groups in the launch
    void* fpl[27];       // population planes of the fluid: slot i of cell c at fpl[i][c]
    void* rho;           // float, or double for kF64
    void* vel[3];        // velocity component planes
    const uint8_t* flags;
    float om_p, om_n;
    int acc_mode;        // 0: none, 1: constant g, 2: Boussinesq from bq_C
    float g[3];
    float alpha, c0;
    const void* bq_C;    // tracer values read for the buoyancy (previous step)
    int write_macro;     // write rho/vel of non-fixed cells this step
    // advection-diffusion tracer (fused into lbm_step, or the one ad_step updates)
    void* gpl[7];
    void* tC;
    float t_om;
    // Visit the launch's tiles last to first.  The odd steps of a lattice larger than the
    // Infinity Cache do, so that each step starts on the lines the previous one touched last,
    // which that cache still holds (an AA step's result does not depend on the order in
    // which cells are visited).
    int reverse;
};

template <class T>
__device__ inline T& elem(void* base, uint32_t cell) {
    return *(T*)((char*)base + cell * (uint32_t)sizeof(T));
}
template <class T>
__device__ inline const T& elem(const void* base, uint32_t cell) {
    return *(const T*)((const char*)base + cell * (uint32_t)sizeof(T));
}

enum AccMode : int { kAccNone = 0, kAccConst = 1, kAccBoussinesq = 2 };

// Work-groups are dealt to the 8 XCDs round-robin (XCD = id % 8).  Renumber them so that
// each XCD works on one contiguous range of rows: neighbouring work-groups then share an
// L2 for the cache lines straddling their boundary.
__device__ inline int xcd_swizzle(int p, int n) {
    const int q = n >> 3, r = n & 7;
    const int k = p & 7, j = p >> 3;
    return k * q + (k < r ? k : r) + j;
}

// Cell coordinates and periodic neighbour indices of the cell owned by this work-item.
template <int D>
struct Nbr;

template <>
struct Nbr<2> {
    uint32_t xb[3];  // (ix + dx mod nx) * ny for dx = -1, 0, 1
    uint32_t y[3];   // iy + dy mod ny
    __device__ Nbr(int ix, int iy, int, const StepParams& p) {
        const int xm = ix == 0 ? p.n0 - 1 : ix - 1, xp = ix == p.n0 - 1 ? 0 : ix + 1;
        xb[0] = (uint32_t)xm * (uint32_t)p.n1;
        xb[1] = (uint32_t)ix * (uint32_t)p.n1;
        xb[2] = (uint32_t)xp * (uint32_t)p.n1;
        y[0] = iy == 0 ? p.n1 - 1 : iy - 1;
        y[1] = iy;
        y[2] = iy == p.n1 - 1 ? 0 : iy + 1;
    }
    __device__ uint32_t at(int dx, int dy, int) const { return xb[dx + 1] + y[dy + 1]; }
    // first cell of the row at (ix + dx); the row is the contiguous axis
    __device__ uint32_t row_at(int dx, int) const { return xb[dx + 1]; }
};

template <>
struct Nbr<3> {
    uint32_t rb[3][3];  // ((ix+dx)*ny + (iy+dy)) * nz
    uint32_t z[3];
    __device__ Nbr(int ix, int iy, int iz, const StepParams& p) {
        const uint32_t xs[3] = {(uint32_t)(ix == 0 ? p.n0 - 1 : ix - 1), (uint32_t)ix,
                                (uint32_t)(ix == p.n0 - 1 ? 0 : ix + 1)};
        const uint32_t ys[3] = {(uint32_t)(iy == 0 ? p.n1 - 1 : iy - 1), (uint32_t)iy,
                                (uint32_t)(iy == p.n1 - 1 ? 0 : iy + 1)};
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) rb[a][b] = (xs[a] * (uint32_t)p.n1 + ys[b]) * (uint32_t)p.n2;
        z[0] = iz == 0 ? p.n2 - 1 : iz - 1;
        z[1] = iz;
        z[2] = iz == p.n2 - 1 ? 0 : iz + 1;
    }
    __device__ uint32_t at(int dx, int dy, int dz) const { return rb[dx + 1][dy + 1] + z[dz + 1]; }
    // first cell of the row at (ix + dx, iy + dy); the row is the contiguous axis
    __device__ uint32_t row_at(int dx, int dy) const { return rb[dx + 1][dy + 1]; }
};

// ---- AA loads and stores ---------------------------------------------------------------
// Populations are touched once per step.  STREAM picks the load form, per lattice:
//  - streaming (non-temporal) loads, for a working set larger than MI355X's 256 MB
//    Infinity Cache.  Measured on D3Q27 TRT 256^3, they lift FP16-shifted storage from
//    42.1k to 45.7k MLUPS and float storage by 2-4 %;
//  - ordinary loads, for a working set that fits that cache.  The lines a step loads and
//    stores then stay there for the next step: 1.1-1.4x the streaming form on working
//    sets of 150-240 MB, where the streaming hint keeps them out.
// Streaming stores measured slower and are off unless BOLT_NT_STORE is defined.
template <class T, bool STREAM>
__device__ inline T pop_load(const void* plane, uint32_t cell) {
    if constexpr (STREAM) return __builtin_nontemporal_load(&elem<T>(plane, cell));
    else return elem<T>(plane, cell);
}

template <class T>
__device__ inline void pop_store(void* plane, uint32_t cell, T v) {
#ifdef BOLT_NT_STORE
    __builtin_nontemporal_store(v, &elem<T>(plane, cell));
#else
    elem<T>(plane, cell) = v;
#endif
}

// An accessor reaches the AA locations of one cell by position k: slot k of plane k at
// x - c_k on the even step, at x on the odd step.  ld<V, T>(k) loads position k and
// st<V, T>(k, v) stores it.  CellAA addresses every position in memory (one cell per
// work-item); the row-staged even step (RowAA, lbm_step_row) keeps the positions shifted
// along the row in LDS.
template <int D, bool EVEN, bool STREAM>
struct CellAA {
    void* const* planes;
    const Nbr<D>& nb;
    uint32_t cell;
    template <class V>
    __device__ uint32_t at(int k) const {
        return EVEN ? nb.at(-V::c(k, 0), -V::c(k, 1), -V::c(k, 2)) : cell;
    }
    template <class V, class T>
    __device__ T ld(int k) const {
        return pop_load<T, STREAM>(planes[k], at<V>(k));
    }
    template <class V, class T>
    __device__ void st(int k, T v) const {
        pop_store<T>(planes[k], at<V>(k), v);
    }
};

// f_i of the cell: position i on the even step, position opp(i) on the odd step.
template <class V, bool EVEN, class T, class A>
__device__ inline void aa_load_raw(const A& a, T (&v)[V::q]) {
#pragma unroll
    for (int i = 0; i < V::q; ++i) v[i] = a.template ld<V, T>(EVEN ? i : V::opp(i));
}

// The post-collision populations: position i takes f*_opp(i) on the even step, f*_i on
// the odd step.
template <class V, bool EVEN, class T, class A>
__device__ inline void aa_store_raw(const A& a, const T (&v)[V::q]) {
#pragma unroll
    for (int i = 0; i < V::q; ++i) a.template st<V, T>(i, EVEN ? v[V::opp(i)] : v[i]);
}

// ---- reference arithmetic (storage kF32) -------------------------------------------------
// Equilibrium polynomial of slot i, double, in the reference's printed term order.
template <class V, int I>
__device__ inline double ref_feq_poly(const float (&u)[3], float vv) {
#pragma clang fp contract(off)
    constexpr double c_vv = (-3.0 * V::wn(I)) / (2.0 * V::wd(I));
    constexpr double c_lin = (3.0 * V::wn(I)) / (double)V::wd(I);
    constexpr double c_sq = (9.0 * V::wn(I)) / (2.0 * V::wd(I));
    constexpr double c_0 = (double)V::wn(I) / (double)V::wd(I);
    if constexpr (V::nnz(I) == 0) {
        return c_0 + c_vv * (double)vv;
    } else {
        float cu = 0.0f;
        bool first = true;
#pragma unroll
        for (int a = 0; a < V::d; ++a) {
            if (V::c(I, a) == 0) continue;
            const float t = V::c(I, a) > 0 ? u[a] : -u[a];
            cu = first ? t : cu + t;
            first = false;
        }
        const float sq = cu * cu;
        double p = c_vv * (double)vv;
        if constexpr (V::nnz(I) == 1) {
            p = p + c_sq * (double)sq;
#pragma unroll
            for (int a = 0; a < V::d; ++a)
                if (V::c(I, a) != 0) p = p + (V::c(I, a) > 0 ? c_lin : -c_lin) * (double)u[a];
        } else {
#pragma unroll
            for (int a = 0; a < V::d; ++a)
                if (V::c(I, a) != 0) p = p + (V::c(I, a) > 0 ? c_lin : -c_lin) * (double)u[a];
            p = p + c_sq * (double)sq;
        }
        return p + c_0;
    }
}

template <class V, int I>
__device__ inline double ref_feq_pos(const float (&u)[3], float vv) {
#pragma clang fp contract(off)
    constexpr double c_vv = (-3.0 * V::wn(I)) / (2.0 * V::wd(I));
    constexpr double c_sq = (9.0 * V::wn(I)) / (2.0 * V::wd(I));
    constexpr double c_0 = (double)V::wn(I) / (double)V::wd(I);
    if constexpr (V::nnz(I) == 0) {
        return c_0 + c_vv * (double)vv;
    } else {
        float cu = 0.0f;
        bool first = true;
#pragma unroll
        for (int a = 0; a < V::d; ++a) {
            if (V::c(I, a) == 0) continue;
            const float t = V::c(I, a) > 0 ? u[a] : -u[a];
            cu = first ? t : cu + t;
            first = false;
        }
        const float sq = cu * cu;
        double p = c_vv * (double)vv;
        p = p + c_sq * (double)sq;
        return p + c_0;
    }
}

template <class V, int I>
__device__ inline double ref_feq_neg(const float (&u)[3]) {
#pragma clang fp contract(off)
    constexpr double c_lin = (3.0 * V::wn(I)) / (double)V::wd(I);
    double p = 0.0;
    bool first = true;
#pragma unroll
    for (int a = 0; a < V::d; ++a) {
        if (V::c(I, a) == 0) continue;
        const double t = (V::c(I, a) > 0 ? c_lin : -c_lin) * (double)u[a];
        p = first ? t : p + t;
        first = false;
    }
    return p;
}

template <class V, int I = 0>
__device__ inline void ref_bgk(float (&f)[V::q], const float (&u)[3], float vv, float r,
                               float om) {
#pragma clang fp contract(off)
    if constexpr (I < V::q) {
        const double p = ref_feq_poly<V, I>(u, vv);
        const double upd = (double)om * ((double)r * p - (double)f[I]);
        f[I] = (float)((double)f[I] + upd);
        ref_bgk<V, I + 1>(f, u, vv, r, om);
    }
}

// The reference-arithmetic collision of one opposite pair (I, opp I) of one cell
// (lib.cl:266-274 BGK, 701-719 TRT: f+/- halves, omega+ on the even part, omega- on the odd
// part, double polynomials narrowed to float).
template <class V, int COLL, int I>
__device__ inline void ref_pair(float (&f)[V::q], const float (&u)[3], float vv, float r,
                                float omp, float omn) {
#pragma clang fp contract(off)
    constexpr int J = V::opp(I);
    if constexpr (COLL == 0) {
        {
            const double pI = ref_feq_poly<V, I>(u, vv);
            f[I] = (float)((double)f[I] + (double)omp * ((double)r * pI - (double)f[I]));
        }
        if constexpr (J != I) {
            const double pJ = ref_feq_poly<V, J>(u, vv);
            f[J] = (float)((double)f[J] + (double)omp * ((double)r * pJ - (double)f[J]));
        }
    } else {
        const double pp = ref_feq_pos<V, I>(u, vv);
        const double pn = ref_feq_neg<V, I>(u);
        const float fpos = (float)((double)omp * ((double)r * pp - (double)(f[I] + f[J]) * 0.5));
        const float fneg = (float)((double)omn * ((double)r * pn - (double)(f[I] - f[J]) * 0.5));
        f[I] = f[I] + (fpos + fneg);
        if constexpr (J != I) f[J] = f[J] + (fpos - fneg);
    }
}

// The reference-arithmetic collision fused with the AA stores, one opposite pair at a time,
// so each population's register is free as soon as it is stored.
template <class V, bool EVEN, int COLL, class A, int I = 0>
__device__ inline void ref_collide_store(float (&f)[V::q], const float (&u)[3], float vv, float r,
                                         float omp, float omn, const A& a) {
    if constexpr (I < V::q) {
        constexpr int J = V::opp(I);
        if constexpr (J >= I) {
            ref_pair<V, COLL, I>(f, u, vv, r, omp, omn);
            if constexpr (EVEN) {
                a.template st<V, float>(I, f[J]);
                if constexpr (J != I) a.template st<V, float>(J, f[I]);
            } else {
                a.template st<V, float>(I, f[I]);
                if constexpr (J != I) a.template st<V, float>(J, f[J]);
            }
        }
        ref_collide_store<V, EVEN, COLL, A, I + 1>(f, u, vv, r, omp, omn, a);
    }
}

// rho and u in the reference's order.
template <class V>
__device__ inline void ref_moments(const float (&f)[V::q], float& r, float (&u)[3]) {
#pragma clang fp contract(off)
    r = f[0];
#pragma unroll
    for (int i = 1; i < V::q; ++i) r = r + f[i];
#pragma unroll
    for (int a = 0; a < V::d; ++a) {
        float s = 0.0f;
        bool first = true;
#pragma unroll
        for (int k = 0; k < V::q; ++k) {
            const int i = vel_order<V::q>(k);
            if (V::c(i, a) == 0) continue;
            const float t = V::c(i, a) > 0 ? f[i] : -f[i];
            s = first ? t : s + t;
            first = false;
        }
        u[a] = s / r;
    }
}

template <int D>
__device__ inline float ref_vv(const float (&u)[3]) {
#pragma clang fp contract(off)
    float vv = u[0] * u[0];
    vv = vv + u[1] * u[1];
    if (D == 3) vv = vv + u[2] * u[2];
    return vv;
}

template <class V>
__device__ inline float cdotu(int i, const float (&u)[3]) {
    float cu = 0.f;
#pragma unroll
    for (int a = 0; a < V::d; ++a)
        if (V::c(i, a) != 0) cu += V::c(i, a) > 0 ? u[a] : -u[a];
    return cu;
}

// ---- float arithmetic on deviations d_i = f_i - w_i (kF32F and the shifted formats) ----
// rho - 1 = sum d, j = sum c d (the w_i cancel pairwise), and the update
//   d_i += omega (w_i (rho - 1) + w_i rho (3 cu + 9/2 cu^2 - 3/2 u.u) - d_i)
// never forms w_i + d_i, which keeps the small deviations exact to float rounding.  The
// multiply-adds are written out (fmaf) and nothing else is contracted, so every kernel that
// uses these functions produces the same bits.
template <class V>
__device__ inline void dev_moments(const float (&d)[V::q], float& dr, float (&u)[3]) {
#pragma clang fp contract(off)
    dr = 0.f;
    float j[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < V::q; ++i) {
        dr += d[i];
#pragma unroll
        for (int a = 0; a < V::d; ++a)
            if (V::c(i, a) != 0) j[a] += V::c(i, a) > 0 ? d[i] : -d[i];
    }
    const float inv = 1.0f / (1.0f + dr);
#pragma unroll
    for (int a = 0; a < V::d; ++a) u[a] = j[a] * inv;
}

__device__ inline float dev_vv(const float (&u)[3]) {
#pragma clang fp contract(off)
    return u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
}

// BGK of slot I
template <class V, int I>
__device__ inline void dev_bgk_slot(float (&d)[V::q], const float (&u)[3], float m15vv, float dr,
                                    float r, float om) {
#pragma clang fp contract(off)
    const float cu = cdotu<V>(I, u);
    const float t = __builtin_fmaf(cu, __builtin_fmaf(4.5f, cu, 3.0f), m15vv);
    const float deq = V::w(I) * __builtin_fmaf(r, t, dr);
    d[I] = __builtin_fmaf(om, deq - d[I], d[I]);
}

// TRT of the pair (I, opp I) (I odd), or of the rest slot (I = 0)
template <class V, int I>
__device__ inline void dev_trt_pair(float (&d)[V::q], const float (&u)[3], float m15vv, float dr,
                                    float r, float omp, float omn) {
#pragma clang fp contract(off)
    if constexpr (I == 0) {
        const float eq0 = V::w(0) * __builtin_fmaf(r, m15vv, dr);
        d[0] = __builtin_fmaf(omp, eq0 - d[0], d[0]);
    } else {
        constexpr int J = V::opp(I);
        const float cu = cdotu<V>(I, u);
        const float eqp = V::w(I) * __builtin_fmaf(r, __builtin_fmaf(4.5f * cu, cu, m15vv), dr);
        const float eqm = V::w(I) * (r * 3.0f * cu);
        const float dp = omp * (eqp - 0.5f * (d[I] + d[J]));
        const float dm = omn * (eqm - 0.5f * (d[I] - d[J]));
        d[I] += dp + dm;
        d[J] += dp - dm;
    }
}

// the collision of the opposite pair (I, opp I), I <= opp I, of one cell
template <class V, int COLL, int I>
__device__ inline void dev_pair(float (&d)[V::q], const float (&u)[3], float m15vv, float dr,
                                float r, float omp, float omn) {
    if constexpr (COLL == 0) {
        dev_bgk_slot<V, I>(d, u, m15vv, dr, r, omp);
        if constexpr (V::opp(I) != I) dev_bgk_slot<V, V::opp(I)>(d, u, m15vv, dr, r, omp);
    } else {
        dev_trt_pair<V, I>(d, u, m15vv, dr, r, omp, omn);
    }
}

template <class V, int COLL, int I = 0>
__device__ inline void dev_collide(float (&d)[V::q], const float (&u)[3], float vv, float dr,
                                   float omp, float omn) {
    if constexpr (I < V::q) {
        if constexpr (V::opp(I) >= I) {
#pragma clang fp contract(off)
            dev_pair<V, COLL, I>(d, u, -1.5f * vv, dr, 1.0f + dr, omp, omn);
        }
        dev_collide<V, COLL, I + 1>(d, u, vv, dr, omp, omn);
    }
}

// ---- double arithmetic (kF64) -------------------------------------------------------------
template <class V>
__device__ inline void f64_moments(const double (&f)[V::q], double& r, double (&u)[3]) {
#pragma clang fp contract(off)
    r = 0.0;
    u[0] = u[1] = u[2] = 0.0;
#pragma unroll
    for (int i = 0; i < V::q; ++i) {
        r += f[i];
#pragma unroll
        for (int a = 0; a < V::d; ++a) u[a] += (double)V::c(i, a) * f[i];
    }
#pragma unroll
    for (int a = 0; a < V::d; ++a) u[a] /= r;
}

template <class V, int I>
__device__ inline double f64_feq(double r, const double (&u)[3], double vv) {
#pragma clang fp contract(off)
    double cu = 0.0;
#pragma unroll
    for (int a = 0; a < V::d; ++a) cu += (double)V::c(I, a) * u[a];
    const double w = (double)V::wn(I) / (double)V::wd(I);
    return w * r * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * vv);
}

template <class V, int COLL, int I>
__device__ inline void f64_pair(double (&f)[V::q], double r, const double (&u)[3], double vv,
                                double op, double on) {
#pragma clang fp contract(off)
    constexpr int J = V::opp(I);
    const double ei = f64_feq<V, I>(r, u, vv);
    const double ej = f64_feq<V, J>(r, u, vv);
    if constexpr (COLL == 0) {
        f[I] += op * (ei - f[I]);
        if constexpr (J != I) f[J] += op * (ej - f[J]);
    } else {
        const double fpos = op * (0.5 * (ei + ej) - 0.5 * (f[I] + f[J]));
        const double fneg = on * (0.5 * (ei - ej) - 0.5 * (f[I] - f[J]));
        f[I] += fpos + fneg;
        if constexpr (J != I) f[J] += fpos - fneg;
    }
}

template <class V, int COLL, int I = 0>
__device__ inline void f64_collide(double (&f)[V::q], double r, const double (&u)[3], double vv,
                                   double op, double on) {
    if constexpr (I < V::q) {
        if constexpr (V::opp(I) >= I) f64_pair<V, COLL, I>(f, r, u, vv, op, on);
        f64_collide<V, COLL, I + 1>(f, r, u, vv, op, on);
    }
}

// moments, forcing, fixed-cell override and vv of one cell in double
template <class V>
__device__ inline void f64_prepare(const double (&f)[V::q], double& r, double (&u)[3], double& vv,
                                   double& op, double& on, bool fixed, const StepParams& p,
                                   double rfix, const double (&ufix)[3], const double (&acc)[3],
                                   bool use_acc) {
#pragma clang fp contract(off)
    f64_moments<V>(f, r, u);
    op = p.om_p;
    on = p.om_n;
    if (use_acc)
#pragma unroll
        for (int a = 0; a < V::d; ++a) u[a] += acc[a] / op;
    if (fixed) {
        op = 1.0;
        on = 1.0;
        r = rfix;
#pragma unroll
        for (int a = 0; a < V::d; ++a) u[a] = ufix[a];
    }
    vv = 0.0;
#pragma unroll
    for (int a = 0; a < V::d; ++a) vv += u[a] * u[a];
}

template <class V, int COLL>
__device__ inline void f64_step(double (&f)[V::q], double& r, double (&u)[3], bool fixed,
                                const StepParams& p, double rfix, const double (&ufix)[3],
                                const double (&acc)[3], bool use_acc) {
    double vv, op, on;
    f64_prepare<V>(f, r, u, vv, op, on, fixed, p, rfix, ufix, acc, use_acc);
    f64_collide<V, COLL>(f, r, u, vv, op, on);
}

template <class V>
__device__ inline void f64_ad(double (&g)[V::q], double& r, const double (&u)[3], bool fixed,
                              double om, double cfix) {
#pragma clang fp contract(off)
    r = 0.0;
#pragma unroll
    for (int i = 0; i < V::q; ++i) r += g[i];
    if (fixed) {
        om = 1.0;
        r = cfix;
    }
    double vv = 0.0;
#pragma unroll
    for (int a = 0; a < V::d; ++a) vv += u[a] * u[a];
#pragma unroll
    for (int i = 0; i < V::q; ++i) {
        double cu = 0.0;
#pragma unroll
        for (int a = 0; a < V::d; ++a) cu += (double)V::c(i, a) * u[a];
        const double w = (double)V::wn(i) / (double)V::wd(i);
        g[i] += om * (w * r * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * vv) - g[i]);
    }
}

// ---- advection-diffusion update of one cell, float storage, reference arithmetic --------
template <class V>
__device__ inline void ref_ad(float (&g)[V::q], float& r, const float (&u)[3], bool fixed,
                              float om, float cfix) {
#pragma clang fp contract(off)
    r = g[0];
#pragma unroll
    for (int i = 1; i < V::q; ++i) r = r + g[i];
    if (fixed) {
        om = 1.0f;
        r = cfix;
    }
    const float vv = ref_vv<V::d>(u);
    ref_bgk<V>(g, u, vv, r, om);
}

// Maps the work-item to its cell.  Returns false for work-items past the end of a row.
template <int D>
__device__ inline bool locate(const StepParams& p, int& ix, int& iy, int& iz, uint32_t& cell) {
    int b = xcd_swizzle((int)blockIdx.x, p.nblocks);
    if (p.reverse) b = p.nblocks - 1 - b;
    const int row = b / p.cpr;
    const int chunk = b - row * p.cpr;
    const int zc = chunk * (int)blockDim.x + (int)threadIdx.x;
    if (zc >= p.len) return false;
    if (D == 2) {
        ix = row;
        iy = zc;
        iz = 0;
    } else {
        ix = row / p.n1;
        iy = row - ix * p.n1;
        iz = zc;
    }
    cell = (uint32_t)row * (uint32_t)p.len + (uint32_t)zc;
    return true;
}

// ---------------------------------------------------------------------------------------
// The fluid step, optionally fused with one advection-diffusion tracer (TQ = 5 or 7).
// ---------------------------------------------------------------------------------------
// Waves per SIMD the register allocation must allow: more resident waves keep more
// loads in flight, which is what bounds this memory-streaming kernel.
template <int Q, int S, int TQ>
constexpr int waves_target() {
#ifdef BOLT_WAVES
    return BOLT_WAVES;
#else
    if (S == kF64 || TQ != 0) return 2;
    return Q == 27 ? 4 : 8;
#endif
}

// One non-wall cell of one time step: fluid update through the accessor `a` (its AA
// locations), optionally followed by one tracer (TQ != 0) through the accessor `ga`.
template <int D, int Q, int COLL, int S, bool EVEN, int TQ, class A, class GA>
__device__ inline void update_cell_with(const StepParams& p, const A& a, const GA& ga, int flag,
                                        uint32_t cell) {
    using V = VSet<D, Q>;
    using ST = Store<S>;
    using T = typename ST::T;
    const bool fixed = (flag & 6) != 0;

    if constexpr (S == kF64) {
        double f[Q];
        aa_load_raw<V, EVEN>(a, f);
        double rfix = 0.0, ufix[3] = {0, 0, 0}, acc[3] = {0, 0, 0};
        if (fixed) {
            rfix = elem<double>((const void*)p.rho, cell);
#pragma unroll
            for (int a = 0; a < D; ++a) ufix[a] = elem<double>((const void*)p.vel[a], cell);
        }
        const bool use_acc = p.acc_mode != kAccNone;
        if (p.acc_mode == kAccConst) {
#pragma unroll
            for (int a = 0; a < D; ++a) acc[a] = (double)p.g[a];
        } else if (p.acc_mode == kAccBoussinesq) {
            const double b = (double)p.alpha * (elem<double>(p.bq_C, cell) - (double)p.c0);
#pragma unroll
            for (int a = 0; a < D; ++a) acc[a] = (double)p.g[a] * b;
        }
        double r, u[3];
        f64_step<V, COLL>(f, r, u, fixed, p, rfix, ufix, acc, use_acc);
        aa_store_raw<V, EVEN>(a, f);
        if (p.write_macro && !fixed) {
            elem<double>(p.rho, cell) = r;
#pragma unroll
            for (int a = 0; a < D; ++a) elem<double>(p.vel[a], cell) = u[a];
        }
        if constexpr (TQ != 0) {
            using W = VSet<D, TQ>;
            double g[TQ];
            aa_load_raw<W, EVEN>(ga, g);
            const bool tfixed = (flag & 8) != 0;
            double rt;
            f64_ad<W>(g, rt, u, tfixed, (double)p.t_om,
                      tfixed ? elem<double>((const void*)p.tC, cell) : 0.0);
            aa_store_raw<W, EVEN>(ga, g);
            elem<double>(p.tC, cell) = rt;
        }
    } else {
        float f[Q];
        float r, u[3] = {0.f, 0.f, 0.f};
        float omp = p.om_p, omn = p.om_n;
        if constexpr (S == kF32) {
            aa_load_raw<V, EVEN>(a, f);
            ref_moments<V>(f, r, u);
        } else if constexpr (S == kF32F) {
            aa_load_raw<V, EVEN>(a, f);
#pragma unroll
            for (int i = 0; i < Q; ++i) f[i] -= V::w(i);
            dev_moments<V>(f, r, u);  // r holds rho - 1 here
        } else {
            T raw[Q];
            aa_load_raw<V, EVEN>(a, raw);
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                if constexpr (S == kF16S) f[i] = (float)raw[i];
                else f[i] = __uint_as_float(((uint32_t)raw[i]) << 16);
            }
            dev_moments<V>(f, r, u);  // r holds rho - 1 here
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
            for (int a = 0; a < D; ++a) u[a] = u[a] + acc[a] / omp;
        }
        if (fixed) {
            omp = 1.0f;
            omn = 1.0f;
            r = elem<float>((const void*)p.rho, cell);
            if constexpr (S != kF32) r = r - 1.0f;
#pragma unroll
            for (int a = 0; a < D; ++a) u[a] = elem<float>((const void*)p.vel[a], cell);
        }
        if constexpr (S == kF32) {
            const float vv = ref_vv<D>(u);
            ref_collide_store<V, EVEN, COLL>(f, u, vv, r, omp, omn, a);
        } else if constexpr (S == kF32F) {
            dev_collide<V, COLL>(f, u, dev_vv(u), r, omp, omn);
#pragma unroll
            for (int i = 0; i < Q; ++i) f[i] += V::w(i);
            aa_store_raw<V, EVEN>(a, f);
            r = r + 1.0f;
        } else {
            dev_collide<V, COLL>(f, u, dev_vv(u), r, omp, omn);
            T raw[Q];
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                if constexpr (S == kF16S) raw[i] = (_Float16)f[i];
                else raw[i] = f32_to_bf16_rne(f[i]);
            }
            aa_store_raw<V, EVEN>(a, raw);
            r = r + 1.0f;
        }
        if (p.write_macro && !fixed) {
            elem<float>(p.rho, cell) = r;
#pragma unroll
            for (int a = 0; a < D; ++a) elem<float>(p.vel[a], cell) = u[a];
        }
        if constexpr (TQ != 0) {
            static_assert(S == kF32, "fused tracers use float storage");
            using W = VSet<D, TQ>;
            float g[TQ];
            aa_load_raw<W, EVEN>(ga, g);
            const bool tfixed = (flag & 8) != 0;
            float rt;
            ref_ad<W>(g, rt, u, tfixed, p.t_om,
                      tfixed ? elem<float>((const void*)p.tC, cell) : 0.0f);
            aa_store_raw<W, EVEN>(ga, g);
            elem<float>(p.tC, cell) = rt;
        }
    }
}

// One cell of one time step, every AA location addressed in memory (one cell per work-item).
template <int D, int Q, int COLL, int S, bool EVEN, int TQ, bool STREAM>
__device__ inline void update_cell(const StepParams& p, int ix, int iy, int iz, uint32_t cell) {
    const int flag = elem<uint8_t>((const void*)p.flags, cell);
    if (flag & 1) return;
    const Nbr<D> nb(ix, iy, iz, p);
    update_cell_with<D, Q, COLL, S, EVEN, TQ>(p, CellAA<D, EVEN, STREAM>{p.fpl, nb, cell},
                                              CellAA<D, EVEN, STREAM>{p.gpl, nb, cell}, flag, cell);
}

template <int D, int Q, int COLL, int S, bool EVEN, int TQ, bool STREAM>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(waves_target<Q, S, TQ>())))
    lbm_step(const StepParams p) {
    int ix, iy, iz;
    uint32_t cell;
    if (!locate<D>(p, ix, iy, iz, cell)) return;
    update_cell<D, Q, COLL, S, EVEN, TQ, STREAM>(p, ix, iy, iz, cell);
}

// ---- row-staged even step ----------------------------------------------------------------
// The even (streaming) step with every global access an aligned row.  One work-group owns a
// whole row of the contiguous axis (len <= blockDim.x <= 1024 cells, one per work-item).
// Position k of cell z lies in row (x - c_k) of plane k, at element z - c_k,row of that row
// (c_k,row: the component along the contiguous axis; the lattice is periodic, so the
// element wraps within the row).  Work-item z loads element z of every plane's row: where
// c_k,row = 0 that is its own position (kept in registers); the other planes' rows go to
// LDS, where work-item z reads and later writes element (z - c_k,row) mod len -- a
// bijection, so each element has exactly one owner and needs no barrier between its read
// and its write.  After a barrier the shifted planes' rows are stored from LDS at element
// z.  Elements owned by wall cells, which neither read nor write, go back as loaded.  A
// fused tracer's planes are staged the same way.  The per-cell arithmetic is
// update_cell_with, shared with lbm_step (and the batch kernels), so the kernels give
// identical bits.

// LDS row of plane k (planes with no component along the contiguous axis have none).
template <class V>
__host__ __device__ constexpr int row_slot(int k) {
    int s = 0;
    for (int j = 0; j < k; ++j) s += V::c(j, V::d - 1) != 0;
    return s;
}
template <class V>
__host__ __device__ constexpr int row_slots() {
    return row_slot<V>(V::q);
}

template <int D, class T>
struct RowAA {
    void* const* planes;
    const Nbr<D>& nb;
    const T* own;  // element z of the rows of the planes with c_k,row = 0
    T* lds;        // rows of the other planes: row_slot(k) * stride + element
    int stride, z, len;
    template <class V>
    __device__ int lds_index(int k) const {
        const int e = z - V::c(k, D - 1);
        return row_slot<V>(k) * stride + (e < 0 ? e + len : (e >= len ? e - len : e));
    }
    template <class V, class U>
    __device__ U ld(int k) const {
        static_assert(sizeof(U) == sizeof(T), "storage element");
        if (V::c(k, D - 1) == 0) return own[k];
        return lds[lds_index<V>(k)];
    }
    template <class V, class U>
    __device__ void st(int k, U v) const {
        static_assert(sizeof(U) == sizeof(T), "storage element");
        if (V::c(k, D - 1) == 0)
            pop_store<U>(planes[k], nb.row_at(-V::c(k, 0), D == 3 ? -V::c(k, 1) : 0) + (uint32_t)z, v);
        else
            lds[lds_index<V>(k)] = v;
    }
};

// One row of the row-staged even step, fluid (Q) and, with TQ != 0, the fused tracer: the
// work-group owns row `row`, work-item z = threadIdx.x its cell z (z < p.len), and `lds`
// holds (row_slots<fluid> + row_slots<tracer>) * blockDim.x elements of the storage type.
// Every work-item of the work-group must call it (it holds two barriers).
template <int D, int Q, int COLL, int S, int TQ, bool STREAM>
__device__ inline void row_step(const StepParams& p, int row, unsigned char* lds_bytes) {
    using V = VSet<D, Q>;
    using W = VSet<D, TQ != 0 ? TQ : (D == 2 ? 5 : 7)>;  // the tracer's set (unused if TQ == 0)
    using T = typename Store<S>::T;
    static_assert(TQ == 0 || S == kF32 || S == kF64, "fused tracers use float or double storage");
    constexpr int GQ = TQ != 0 ? TQ : 1;
    const int z = (int)threadIdx.x;
    const int stride = (int)blockDim.x;
    const bool live = z < p.len;
    T* lds = (T*)lds_bytes;
    T* glds = lds + row_slots<V>() * stride;
    int ix, iy;
    if (D == 2) {
        ix = row;
        iy = z;
    } else {
        ix = row / p.n1;
        iy = row - ix * p.n1;
    }
    const Nbr<D> nb(ix, iy, z, p);
    const uint32_t cell = nb.row_at(0, 0) + (uint32_t)z;
    int flag = 1;
    T own[Q], gown[GQ];
    if (live) {
        flag = elem<uint8_t>((const void*)p.flags, cell);
#pragma unroll
        for (int k = 0; k < Q; ++k) {
            const T v = pop_load<T, STREAM>(p.fpl[k], nb.row_at(-V::c(k, 0), D == 3 ? -V::c(k, 1) : 0) + (uint32_t)z);
            if (V::c(k, D - 1) == 0) own[k] = v;
            else lds[row_slot<V>(k) * stride + z] = v;
        }
        if constexpr (TQ != 0) {
#pragma unroll
            for (int k = 0; k < TQ; ++k) {
                const T v = pop_load<T, STREAM>(p.gpl[k], nb.row_at(-W::c(k, 0), D == 3 ? -W::c(k, 1) : 0) + (uint32_t)z);
                if (W::c(k, D - 1) == 0) gown[k] = v;
                else glds[row_slot<W>(k) * stride + z] = v;
            }
        }
    }
    __syncthreads();
    if (live && !(flag & 1))
        update_cell_with<D, Q, COLL, S, true, TQ>(p, RowAA<D, T>{p.fpl, nb, own, lds, stride, z, p.len},
                                                  RowAA<D, T>{p.gpl, nb, gown, glds, stride, z, p.len}, flag,
                                                  cell);
    __syncthreads();
    if (live) {
#pragma unroll
        for (int k = 0; k < Q; ++k)
            if (V::c(k, D - 1) != 0)
                pop_store<T>(p.fpl[k], nb.row_at(-V::c(k, 0), D == 3 ? -V::c(k, 1) : 0) + (uint32_t)z,
                             lds[row_slot<V>(k) * stride + z]);
        if constexpr (TQ != 0) {
#pragma unroll
            for (int k = 0; k < TQ; ++k)
                if (W::c(k, D - 1) != 0)
                    pop_store<T>(p.gpl[k], nb.row_at(-W::c(k, 0), D == 3 ? -W::c(k, 1) : 0) + (uint32_t)z,
                                 glds[row_slot<W>(k) * stride + z]);
        }
    }
}

// BMAX: the largest work-group the instantiation serves (256 or 1024).  Launched with
// p.nblocks = rows, blockDim.x = len rounded up to 64 and (row_slots of the fluid + row_slots
// of the fused tracer) * blockDim.x elements of dynamic LDS.
template <int D, int Q, int COLL, int S, int BMAX, int TQ, bool STREAM>
__global__ void __launch_bounds__(BMAX)
    __attribute__((amdgpu_waves_per_eu(BMAX > 256 ? BMAX / 256 : waves_target<Q, S, TQ>())))
    lbm_step_row(const StepParams p) {
    extern __shared__ __align__(16) unsigned char row_lds_bytes[];
    int row = xcd_swizzle((int)blockIdx.x, p.nblocks);
    if (p.reverse) row = p.nblocks - 1 - row;
    row_step<D, Q, COLL, S, TQ, STREAM>(p, row, row_lds_bytes);
}

// Advection-diffusion tracer on its own, reading the fluid velocity written this step.
template <int D, int TQ, int S, bool EVEN, bool STREAM>
__global__ void __launch_bounds__(256) ad_step(const StepParams p) {
    using W = VSet<D, TQ>;
    int ix, iy, iz;
    uint32_t cell;
    if (!locate<D>(p, ix, iy, iz, cell)) return;
    const int flag = elem<uint8_t>((const void*)p.flags, cell);
    if (flag & 1) return;
    const bool tfixed = (flag & 8) != 0;
    const Nbr<D> nb(ix, iy, iz, p);
    if constexpr (S == kF64) {
        double g[TQ];
        aa_load_raw<W, EVEN>(CellAA<D, EVEN, STREAM>{p.gpl, nb, cell}, g);
        double u[3] = {0, 0, 0};
#pragma unroll
        for (int a = 0; a < D; ++a) u[a] = elem<double>((const void*)p.vel[a], cell);
        double rt;
        f64_ad<W>(g, rt, u, tfixed, (double)p.t_om,
                  tfixed ? elem<double>((const void*)p.tC, cell) : 0.0);
        aa_store_raw<W, EVEN>(CellAA<D, EVEN, STREAM>{p.gpl, nb, cell}, g);
        elem<double>(p.tC, cell) = rt;
    } else {
        float g[TQ];
        aa_load_raw<W, EVEN>(CellAA<D, EVEN, STREAM>{p.gpl, nb, cell}, g);
        float u[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int a = 0; a < D; ++a) u[a] = elem<float>((const void*)p.vel[a], cell);
        float rt;
        ref_ad<W>(g, rt, u, tfixed, p.t_om,
                  tfixed ? elem<float>((const void*)p.tC, cell) : 0.0f);
        aa_store_raw<W, EVEN>(CellAA<D, EVEN, STREAM>{p.gpl, nb, cell}, g);
        elem<float>(p.tC, cell) = rt;
    }
}

}  // namespace bolt
