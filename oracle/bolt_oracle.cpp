// TEST INFRASTRUCTURE -- NOT PART OF THE PRODUCT.
//
// CPU restatement of the reference's lattice update (djbarker/boltzmann, src/lib.cl),
// used only by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as the
// checker.  The product (libbolt_hip.so) never links, loads or calls this file.
//
// Layout follows the reference exactly (array-of-structures, x slowest):
//   cell index ii = (ix*ny + iy)*nz + iz        (2D: ix*ny + iy)      lib.cl:204,426
//   population slot f[ii*Q + i]                                        lib.cl:222,445
//   vel/acc       v[ii*D + d]                                         lib.cl:298-300
//   flags int32: bit0 wall, bits1|2 fixed fluid, bit3 fixed scalar     simulation.rs:80-89
//
// Arithmetic follows lib.cl's "FP32" kernels, which are mixed precision: macroscopic
// moments and the velocity shift are float; each equilibrium polynomial, the
// relaxation and the population update are evaluated in double (unsuffixed literals
// promote) and narrowed to float on store (lib.cl:249-274, 701-717).  Offsets use
// 64-bit indices (lib.cl:222,445 use int, which overflows past 2^31 elements).
// Build with -ffp-contract=off so no multiply-add is fused behind our back.
//
// Functions:
//   orc_update_ns   lib.cl:191-301 (d2q9 bgk), 620-748 (d2q9 trt), 407-618 (d3q27 bgk);
//                   D3Q27-TRT per kernel_gen.py:402-420 applied to kernel_gen.py:100-101
//   orc_update_ad   lib.cl:102-189 (d2q5), 303-405 (d3q7)
//   orc_set_constant_acc   lib.cl:14-44  (bounds test fixed: iz < nz, not iy < nz, lib.cl:36)
//   orc_set_boussinesq_acc lib.cl:60-92  (same fix, lib.cl:83)
//   orc_equilibrate        velocities.rs:43-52 via fields.rs:106-114,185-193 (host, f32)
//   orc_update_ns_f64 / orc_update_ad_f64: the same update carried out in double
//                   throughout (double storage); the FP64 comparison run.
//
// How each part is pinned (tests/test_oracle.py): the velocity tables against the
// fixture parsed from velocities.rs; the omega=0 two-step map against a particle model
// of free streaming with full-way bounce-back; the full update against a textbook
// push/pull stream-and-collide in numpy float64 (np.roll streaming, generic
// equilibrium, bounce-back walls, equilibrium-Dirichlet fixed cells, velocity-shift
// forcing) -- to 2e-6 for the float kernels and to 1e-13 for both double paths
// (orc_update_ns_f64, and orc_update_ad_f64 on its own with a moving fluid); equilibrium
// moments against their closed forms; TRT(w,w) == BGK(w); TRT linearity in (w+, w-); the
// reference's own integration asserts (test_core.py:57-127); orc_equilibrate against the
// product's host equilibrium (tests/test_host_sequence.py).

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <string>
#include <vector>

#include <omp.h>

namespace {

// Velocity sets: c_i ordering and weights of velocities.rs:66-191 (the same order as the
// qs tables in lib.cl).  Weights are kept as exact rationals so that every polynomial
// coefficient can be formed as one correctly rounded double division, which is what the
// reference's literal quotients (e.g. 1.0/18.0) evaluate to.
struct VSet {
    int d, q;
    int c[27][3];
    int wn[27];  // weight numerator
    int wd[27];  // weight denominator
};

const VSet kD2Q9 = {2, 9,
    {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0},
     {1, 1, 0}, {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}},
    {4, 1, 1, 1, 1, 1, 1, 1, 1},
    {9, 9, 9, 9, 9, 36, 36, 36, 36}};

const VSet kD2Q5 = {2, 5,
    {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}},
    {1, 1, 1, 1, 1},
    {3, 6, 6, 6, 6}};

const VSet kD3Q7 = {3, 7,
    {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}},
    {1, 19, 19, 19, 19, 19, 19},
    {20, 120, 120, 120, 120, 120, 120}};

const VSet kD3Q27 = {3, 27,
    {{0, 0, 0},   {1, 0, 0},   {-1, 0, 0},   {0, 1, 0},   {0, -1, 0},  {0, 0, 1},
     {0, 0, -1},  {1, 1, 0},   {-1, -1, 0},  {1, -1, 0},  {-1, 1, 0},  {1, 0, 1},
     {-1, 0, -1}, {0, 1, 1},   {0, -1, -1},  {1, 0, -1},  {-1, 0, 1},  {0, 1, -1},
     {0, -1, 1},  {1, 1, 1},   {-1, -1, -1}, {1, -1, 1},  {-1, 1, -1}, {1, 1, -1},
     {-1, -1, 1}, {1, -1, -1}, {-1, 1, 1}},
    {8, 2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
    {27, 27, 27, 27, 27, 27, 27, 54, 54, 54, 54, 54, 54, 54, 54, 54, 54, 54, 54,
     216, 216, 216, 216, 216, 216, 216, 216}};

const VSet* find_set(int d, int q) {
    if (d == 2 && q == 9) return &kD2Q9;
    if (d == 2 && q == 5) return &kD2Q5;
    if (d == 3 && q == 7) return &kD3Q7;
    if (d == 3 && q == 27) return &kD3Q27;
    return nullptr;
}

// Per-set derived data: opposite index, per-slot polynomial coefficients, the summation
// order of the velocity numerators.
struct Derived {
    const VSet* vs;
    int opp[27];
    int nnz[27];        // number of non-zero velocity components of slot i
    double c_vv[27];    // -3/2 w      (coefficient of |u|^2)
    double c_lin[27];   //  3 w        (coefficient of c.u, applied per component)
    double c_sq[27];    //  9/2 w      (coefficient of (c.u)^2)
    double c_0[27];     //  w
    // TRT pair coefficients: feq+ = w(1 + 9/2 (c.u)^2 - 3/2 |u|^2), feq- = 3 w c.u
    // Velocity numerator order: lib.cl sums sum_i c_i f_i in the order its generator
    // printed the terms, which is the lexicographic order of the names "f_[i]".
    int vel_order[27];
};

Derived make_derived(const VSet* vs) {
    Derived d{};
    d.vs = vs;
    for (int i = 0; i < vs->q; ++i) {
        d.opp[i] = -1;
        for (int j = 0; j < vs->q; ++j) {
            bool neg = true;
            for (int a = 0; a < 3; ++a) neg = neg && (vs->c[j][a] == -vs->c[i][a]);
            if (neg) d.opp[i] = j;
        }
        int n = 0;
        for (int a = 0; a < vs->d; ++a) n += vs->c[i][a] != 0;
        d.nnz[i] = n;
        const double wn = vs->wn[i], wd = vs->wd[i];
        d.c_vv[i] = (-3.0 * wn) / (2.0 * wd);
        d.c_lin[i] = (3.0 * wn) / wd;
        d.c_sq[i] = (9.0 * wn) / (2.0 * wd);
        d.c_0[i] = wn / wd;
    }
    std::vector<std::pair<std::string, int>> names;
    for (int i = 0; i < vs->q; ++i) names.push_back({"f_[" + std::to_string(i) + "]", i});
    std::sort(names.begin(), names.end());
    for (int i = 0; i < vs->q; ++i) d.vel_order[i] = names[i].second;
    return d;
}

const Derived& derived_for(const VSet* vs) {
    static const Derived d2q9 = make_derived(&kD2Q9);
    static const Derived d2q5 = make_derived(&kD2Q5);
    static const Derived d3q7 = make_derived(&kD3Q7);
    static const Derived d3q27 = make_derived(&kD3Q27);
    if (vs == &kD2Q9) return d2q9;
    if (vs == &kD2Q5) return d2q5;
    if (vs == &kD3Q7) return d3q7;
    return d3q27;
}

// Periodic neighbour offsets (lib.cl:218-223): off[i] = index of cell x + c_i.
struct Geometry {
    int d;
    int64_t n[3];
    int64_t ncell;
};

Geometry make_geom(int d, const int* s) {
    Geometry g{};
    g.d = d;
    g.n[0] = s[0];
    g.n[1] = s[1];
    g.n[2] = d == 3 ? s[2] : 1;
    g.ncell = g.n[0] * g.n[1] * g.n[2];
    return g;
}

inline int64_t neighbour(const Geometry& g, int64_t ix, int64_t iy, int64_t iz, const int* c) {
    const int64_t jx = (ix + c[0] + g.n[0]) % g.n[0];
    const int64_t jy = (iy + c[1] + g.n[1]) % g.n[1];
    const int64_t jz = (iz + c[2] + g.n[2]) % g.n[2];
    return (jx * g.n[1] + jy) * g.n[2] + jz;
}

// Equilibrium polynomial of slot i in double, in the reference's printed term order:
// rest slot "w0 + cvv*vv"; axis slots "cvv*vv + csq*sq + lin + w"; diagonal slots
// "cvv*vv + lin + csq*sq + w".  sq = (c.u)^2 is a float square of a float sum
// (lib.cl's pow(vx + vy, 2) is the float overload).
inline double feq_poly(const Derived& D, int i, const float* u, float vv) {
    const VSet* vs = D.vs;
    if (D.nnz[i] == 0) return D.c_0[i] + D.c_vv[i] * (double)vv;
    float cu = 0.0f;
    bool first = true;
    for (int a = 0; a < vs->d; ++a) {
        const int ca = vs->c[i][a];
        if (ca == 0) continue;
        const float t = ca > 0 ? u[a] : -u[a];
        cu = first ? t : cu + t;
        first = false;
    }
    const float sq = cu * cu;
    double lin = 0.0;
    bool lfirst = true;
    for (int a = 0; a < vs->d; ++a) {
        const int ca = vs->c[i][a];
        if (ca == 0) continue;
        const double t = (ca > 0 ? D.c_lin[i] : -D.c_lin[i]) * (double)u[a];
        lin = lfirst ? t : lin + t;
        lfirst = false;
    }
    double p = D.c_vv[i] * (double)vv;
    if (D.nnz[i] == 1) {
        p = p + D.c_sq[i] * (double)sq;
        p = p + lin;
    } else {
        // lin is a sum of per-component terms appended one by one in the reference.
        for (int a = 0; a < vs->d; ++a) {
            const int ca = vs->c[i][a];
            if (ca == 0) continue;
            p = p + (ca > 0 ? D.c_lin[i] : -D.c_lin[i]) * (double)u[a];
        }
        p = p + D.c_sq[i] * (double)sq;
    }
    return p + D.c_0[i];
}

// TRT symmetric / antisymmetric parts for the pair (i, opp i), kernel_gen.py:402-420.
inline double feq_pos_poly(const Derived& D, int i, const float* u, float vv) {
    const VSet* vs = D.vs;
    if (D.nnz[i] == 0) return D.c_0[i] + D.c_vv[i] * (double)vv;
    float cu = 0.0f;
    bool first = true;
    for (int a = 0; a < vs->d; ++a) {
        const int ca = vs->c[i][a];
        if (ca == 0) continue;
        const float t = ca > 0 ? u[a] : -u[a];
        cu = first ? t : cu + t;
        first = false;
    }
    const float sq = cu * cu;
    double p = D.c_vv[i] * (double)vv;
    p = p + D.c_sq[i] * (double)sq;
    return p + D.c_0[i];
}

inline double feq_neg_poly(const Derived& D, int i, const float* u) {
    const VSet* vs = D.vs;
    if (D.nnz[i] == 0) return 0.0;
    double p = 0.0;
    bool first = true;
    for (int a = 0; a < vs->d; ++a) {
        const int ca = vs->c[i][a];
        if (ca == 0) continue;
        const double t = (ca > 0 ? D.c_lin[i] : -D.c_lin[i]) * (double)u[a];
        p = first ? t : p + t;
        first = false;
    }
    return p;
}

// ---- AA addressing (lib.cl:226-246, 276-296) ----------------------------------------
// even: slot i is read from / written to cell x - c_i (= off[opp i]); the value written
//       there is the post-collision population opp(i).
// odd:  slot opp(i) of cell x is read as f_i; slot i of cell x is written with f_i.
template <typename T>
inline void aa_load(const Derived& D, const Geometry& g, const T* f, int64_t ii,
                    const int64_t* off, bool even, T* f_) {
    const int Q = D.vs->q;
    for (int i = 0; i < Q; ++i) {
        if (even)
            f_[i] = f[off[D.opp[i]] * Q + i];
        else
            f_[i] = f[ii * Q + D.opp[i]];
    }
    (void)g;
}

template <typename T>
inline void aa_store(const Derived& D, T* f, int64_t ii, const int64_t* off, bool even,
                     const T* f_) {
    const int Q = D.vs->q;
    for (int i = 0; i < Q; ++i) {
        if (even)
            f[off[D.opp[i]] * Q + i] = f_[D.opp[i]];
        else
            f[ii * Q + i] = f_[i];
    }
}

inline void cell_coords(const Geometry& g, int64_t ii, int64_t* ix, int64_t* iy, int64_t* iz) {
    *iz = ii % g.n[2];
    const int64_t r = ii / g.n[2];
    *iy = r % g.n[1];
    *ix = r / g.n[1];
}

// ---- Navier-Stokes update, mirroring the reference arithmetic ---------------------
void update_ns_mixed(const Derived& D, const Geometry& g, int collision, int even,
                     float omega_in, float omega_neg_in, float* f, float* rho, float* vel,
                     const float* acc, int use_acc, const int32_t* cell) {
    const int Q = D.vs->q, DIM = D.vs->d;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < g.ncell; ++ii) {
        const int c = cell[ii];
        if (c & 1) continue;  // wall: no reads, no writes (lib.cl:207-213)
        const bool fixed = (c & 6) != 0;
        int64_t ix, iy, iz;
        cell_coords(g, ii, &ix, &iy, &iz);
        int64_t off[27];
        for (int i = 0; i < Q; ++i) off[i] = neighbour(g, ix, iy, iz, D.vs->c[i]);
        float f_[27];
        aa_load(D, g, f, ii, off, even != 0, f_);

        float r = f_[0];
        for (int i = 1; i < Q; ++i) r = r + f_[i];
        float u[3] = {0.f, 0.f, 0.f};
        for (int a = 0; a < DIM; ++a) {
            float s = 0.0f;
            bool first = true;
            for (int k = 0; k < Q; ++k) {
                const int i = D.vel_order[k];
                const int ca = D.vs->c[i][a];
                if (ca == 0) continue;
                const float t = ca > 0 ? f_[i] : -f_[i];
                s = first ? t : s + t;
                first = false;
            }
            u[a] = s / r;
        }
        float omega = omega_in, omega_neg = omega_neg_in;
        if (use_acc) {
            for (int a = 0; a < DIM; ++a) u[a] = u[a] + acc[ii * DIM + a] / omega;
        }
        if (fixed) {
            omega = 1.0f;
            omega_neg = 1.0f;
            r = rho[ii];
            for (int a = 0; a < DIM; ++a) u[a] = vel[ii * DIM + a];
        }
        float vv = u[0] * u[0];
        for (int a = 1; a < DIM; ++a) vv = vv + u[a] * u[a];

        if (collision == 0) {  // BGK
            for (int i = 0; i < Q; ++i) {
                const double p = feq_poly(D, i, u, vv);
                const double upd = (double)omega * ((double)r * p - (double)f_[i]);
                f_[i] = (float)((double)f_[i] + upd);
            }
        } else {  // TRT: pairs (i, opp i) with opp i >= i, rest slot first
            for (int i = 0; i < Q; ++i) {
                const int j = D.opp[i];
                if (j < i) continue;
                const double pp = feq_pos_poly(D, i, u, vv);
                const double pn = feq_neg_poly(D, i, u);
                const float fpos = (float)((double)omega *
                                           ((double)r * pp - (double)(f_[i] + f_[j]) * 0.5));
                const float fneg = (float)((double)omega_neg *
                                           ((double)r * pn - (double)(f_[i] - f_[j]) * 0.5));
                f_[i] = f_[i] + (fpos + fneg);
                if (j != i) f_[j] = f_[j] + (fpos - fneg);
            }
        }
        aa_store(D, f, ii, off, even != 0, f_);
        rho[ii] = r;
        for (int a = 0; a < DIM; ++a) vel[ii * DIM + a] = u[a];
    }
}

// ---- Advection-diffusion update (lib.cl:102-189, 303-405) ----------------------------
void update_ad_mixed(const Derived& D, const Geometry& g, int even, float omega_in, float* f,
                     float* val, const float* vel, const int32_t* cell) {
    const int Q = D.vs->q, DIM = D.vs->d;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < g.ncell; ++ii) {
        const int c = cell[ii];
        if (c & 1) continue;
        const bool fixed = (c & 8) != 0;
        int64_t ix, iy, iz;
        cell_coords(g, ii, &ix, &iy, &iz);
        int64_t off[27];
        for (int i = 0; i < Q; ++i) off[i] = neighbour(g, ix, iy, iz, D.vs->c[i]);
        float f_[27];
        aa_load(D, g, f, ii, off, even != 0, f_);
        float r = f_[0];
        for (int i = 1; i < Q; ++i) r = r + f_[i];
        float u[3] = {0.f, 0.f, 0.f};
        for (int a = 0; a < DIM; ++a) u[a] = vel[ii * DIM + a];
        float omega = omega_in;
        if (fixed) {
            omega = 1.0f;
            r = val[ii];
        }
        float vv = u[0] * u[0];
        for (int a = 1; a < DIM; ++a) vv = vv + u[a] * u[a];
        for (int i = 0; i < Q; ++i) {
            const double p = feq_poly(D, i, u, vv);
            const double upd = (double)omega * ((double)r * p - (double)f_[i]);
            f_[i] = (float)((double)f_[i] + upd);
        }
        aa_store(D, f, ii, off, even != 0, f_);
        val[ii] = r;
    }
}

// ---- FP64 variants: same update, double storage and double arithmetic ----------------
inline double feq_d(const Derived& D, int i, const double* u, double vv, double r) {
    double cu = 0.0;
    for (int a = 0; a < D.vs->d; ++a) cu += D.vs->c[i][a] * u[a];
    const double w = (double)D.vs->wn[i] / (double)D.vs->wd[i];
    return w * r * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * vv);
}

void update_ns_f64(const Derived& D, const Geometry& g, int collision, int even, double om_p,
                   double om_n, double* f, double* rho, double* vel, const double* acc,
                   int use_acc, const int32_t* cell) {
    const int Q = D.vs->q, DIM = D.vs->d;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < g.ncell; ++ii) {
        const int c = cell[ii];
        if (c & 1) continue;
        const bool fixed = (c & 6) != 0;
        int64_t ix, iy, iz;
        cell_coords(g, ii, &ix, &iy, &iz);
        int64_t off[27];
        for (int i = 0; i < Q; ++i) off[i] = neighbour(g, ix, iy, iz, D.vs->c[i]);
        double f_[27];
        aa_load(D, g, f, ii, off, even != 0, f_);
        double r = 0.0, u[3] = {0, 0, 0};
        for (int i = 0; i < Q; ++i) {
            r += f_[i];
            for (int a = 0; a < DIM; ++a) u[a] += D.vs->c[i][a] * f_[i];
        }
        for (int a = 0; a < DIM; ++a) u[a] /= r;
        double op = om_p, on = om_n;
        if (use_acc)
            for (int a = 0; a < DIM; ++a) u[a] += acc[ii * DIM + a] / op;
        if (fixed) {
            op = 1.0;
            on = 1.0;
            r = rho[ii];
            for (int a = 0; a < DIM; ++a) u[a] = vel[ii * DIM + a];
        }
        double vv = 0.0;
        for (int a = 0; a < DIM; ++a) vv += u[a] * u[a];
        double feq[27];
        for (int i = 0; i < Q; ++i) feq[i] = feq_d(D, i, u, vv, r);
        if (collision == 0) {
            for (int i = 0; i < Q; ++i) f_[i] += op * (feq[i] - f_[i]);
        } else {
            for (int i = 0; i < Q; ++i) {
                const int j = D.opp[i];
                if (j < i) continue;
                const double fpos = op * (0.5 * (feq[i] + feq[j]) - 0.5 * (f_[i] + f_[j]));
                const double fneg = on * (0.5 * (feq[i] - feq[j]) - 0.5 * (f_[i] - f_[j]));
                f_[i] += fpos + fneg;
                if (j != i) f_[j] += fpos - fneg;
            }
        }
        aa_store(D, f, ii, off, even != 0, f_);
        rho[ii] = r;
        for (int a = 0; a < DIM; ++a) vel[ii * DIM + a] = u[a];
    }
}

void update_ad_f64(const Derived& D, const Geometry& g, int even, double omega_in, double* f,
                   double* val, const double* vel, const int32_t* cell) {
    const int Q = D.vs->q, DIM = D.vs->d;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < g.ncell; ++ii) {
        const int c = cell[ii];
        if (c & 1) continue;
        const bool fixed = (c & 8) != 0;
        int64_t ix, iy, iz;
        cell_coords(g, ii, &ix, &iy, &iz);
        int64_t off[27];
        for (int i = 0; i < Q; ++i) off[i] = neighbour(g, ix, iy, iz, D.vs->c[i]);
        double f_[27];
        aa_load(D, g, f, ii, off, even != 0, f_);
        double r = 0.0;
        for (int i = 0; i < Q; ++i) r += f_[i];
        double u[3] = {0, 0, 0};
        for (int a = 0; a < DIM; ++a) u[a] = vel[ii * DIM + a];
        double omega = omega_in;
        if (fixed) {
            omega = 1.0;
            r = val[ii];
        }
        double vv = 0.0;
        for (int a = 0; a < DIM; ++a) vv += u[a] * u[a];
        for (int i = 0; i < Q; ++i) f_[i] += omega * (feq_d(D, i, u, vv, r) - f_[i]);
        aa_store(D, f, ii, off, even != 0, f_);
        val[ii] = r;
    }
}

}  // namespace

extern "C" {

// Returns 0 on success, -1 for an unknown (D, Q).
int orc_update_ns(int ndim, const int* s, int q, int collision, int even, float omega_pos,
                  float omega_neg, float* f, float* rho, float* vel, const float* acc,
                  int use_acc, const int32_t* cell) {
    const VSet* vs = find_set(ndim, q);
    if (!vs || (q != 9 && q != 27)) return -1;
    update_ns_mixed(derived_for(vs), make_geom(ndim, s), collision, even, omega_pos,
                    omega_neg, f, rho, vel, acc, use_acc, cell);
    return 0;
}

int orc_update_ad(int ndim, const int* s, int q, int even, float omega, float* g, float* val,
                  const float* vel, const int32_t* cell) {
    const VSet* vs = find_set(ndim, q);
    if (!vs || (q != 5 && q != 7)) return -1;
    update_ad_mixed(derived_for(vs), make_geom(ndim, s), even, omega, g, val, vel, cell);
    return 0;
}

int orc_update_ns_f64(int ndim, const int* s, int q, int collision, int even, double omega_pos,
                      double omega_neg, double* f, double* rho, double* vel, const double* acc,
                      int use_acc, const int32_t* cell) {
    const VSet* vs = find_set(ndim, q);
    if (!vs || (q != 9 && q != 27)) return -1;
    update_ns_f64(derived_for(vs), make_geom(ndim, s), collision, even, omega_pos, omega_neg,
                  f, rho, vel, acc, use_acc, cell);
    return 0;
}

int orc_update_ad_f64(int ndim, const int* s, int q, int even, double omega, double* g,
                      double* val, const double* vel, const int32_t* cell) {
    const VSet* vs = find_set(ndim, q);
    if (!vs || (q != 5 && q != 7)) return -1;
    update_ad_f64(derived_for(vs), make_geom(ndim, s), even, omega, g, val, vel, cell);
    return 0;
}

// acc[ii*D + d] = g[d] for every cell (lib.cl:14-44).
void orc_set_constant_acc(int ndim, const int* s, float* acc, const float* g) {
    const Geometry geo = make_geom(ndim, s);
    for (int64_t ii = 0; ii < geo.ncell; ++ii)
        for (int a = 0; a < ndim; ++a) acc[ii * ndim + a] = g[a];
}

// acc[ii*D + d] = g[d] * (alpha * (C[ii] - c0)), float arithmetic (lib.cl:60-92).
void orc_set_boussinesq_acc(int ndim, const int* s, float* acc, const float* conc, float alpha,
                            float c0, const float* g) {
    const Geometry geo = make_geom(ndim, s);
    for (int64_t ii = 0; ii < geo.ncell; ++ii) {
        const float buoyancy = alpha * (conc[ii] - c0);
        for (int a = 0; a < ndim; ++a) acc[ii * ndim + a] = g[a] * buoyancy;
    }
}

// Host equilibrium initialisation in f32 (velocities.rs:43-52):
//   out[q] = rho * w_q * (1 + 3 vq + 4.5 vq vq - 1.5 vv),  vq = u.c_q, vv = u.u,
// both dot products summed sequentially from 0 (ndarray dot of length <= 3), evaluated
// left to right with f32 weights.
int orc_equilibrate(int ndim, int q, int64_t ncell, const float* rho, const float* vel,
                    float* f) {
    const VSet* vs = find_set(ndim, q);
    if (!vs) return -1;
    float w[27];
    for (int i = 0; i < q; ++i) w[i] = (float)((double)vs->wn[i] / (double)vs->wd[i]);
    for (int64_t ii = 0; ii < ncell; ++ii) {
        const float* u = vel + ii * ndim;
        float vv = 0.0f;
        for (int a = 0; a < ndim; ++a) vv = vv + u[a] * u[a];
        for (int i = 0; i < q; ++i) {
            float vq = 0.0f;
            for (int a = 0; a < ndim; ++a) vq = vq + u[a] * (float)vs->c[i][a];
            f[ii * q + i] =
                rho[ii] * w[i] * (1.0f + 3.0f * vq + 4.5f * vq * vq - 1.5f * vv);
        }
    }
    return 0;
}

// Table access for tests: copies c (q x 3 ints), weights (q floats, f32 of the rational)
// and opposite indices.
int orc_velocity_set(int ndim, int q, int32_t* c, float* w, int32_t* opp) {
    const VSet* vs = find_set(ndim, q);
    if (!vs) return -1;
    const Derived& D = derived_for(vs);
    for (int i = 0; i < q; ++i) {
        for (int a = 0; a < 3; ++a) c[i * 3 + a] = vs->c[i][a];
        w[i] = (float)((double)vs->wn[i] / (double)vs->wd[i]);
        opp[i] = D.opp[i];
    }
    return 0;
}

int orc_num_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
