// Host-side velocity tables and the first-iteration equilibrium, in f32.
//
// equilibrate() follows VelocitySet::feq (velocities.rs:43-52) as called by
// Fluid::equilibrate / Scalar::equilibrate (fields.rs:106-114, 185-193):
//   f_q = rho * w_q * (1 + 3 vq + 4.5 vq vq - 1.5 vv),  vq = u.c_q,  vv = u.u,
// with both dot products accumulated from 0 in component order and the expression
// evaluated left to right in float.  This file is compiled with -ffp-contract=off.
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace bolt_host {

struct Table {
    int d, q;
    int c[27][3];
    int wn[27], wd[27];
};

inline const Table* table(int d, int q) {
    static const Table d2q9 = {2, 9,
        {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0},
         {1, 1, 0}, {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}},
        {4, 1, 1, 1, 1, 1, 1, 1, 1}, {9, 9, 9, 9, 9, 36, 36, 36, 36}};
    static const Table d2q5 = {2, 5,
        {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}},
        {1, 1, 1, 1, 1}, {3, 6, 6, 6, 6}};
    static const Table d3q7 = {3, 7,
        {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}},
        {1, 19, 19, 19, 19, 19, 19}, {20, 120, 120, 120, 120, 120, 120}};
    static const Table d3q27 = {3, 27,
        {{0, 0, 0},   {1, 0, 0},   {-1, 0, 0},   {0, 1, 0},   {0, -1, 0},  {0, 0, 1},
         {0, 0, -1},  {1, 1, 0},   {-1, -1, 0},  {1, -1, 0},  {-1, 1, 0},  {1, 0, 1},
         {-1, 0, -1}, {0, 1, 1},   {0, -1, -1},  {1, 0, -1},  {-1, 0, 1},  {0, 1, -1},
         {0, -1, 1},  {1, 1, 1},   {-1, -1, -1}, {1, -1, 1},  {-1, 1, -1}, {1, 1, -1},
         {-1, -1, 1}, {1, -1, -1}, {-1, 1, 1}},
        {8, 2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
        {27, 27, 27, 27, 27, 27, 27, 54, 54, 54, 54, 54, 54, 54, 54, 54, 54, 54, 54,
         216, 216, 216, 216, 216, 216, 216, 216}};
    if (d == 2 && q == 9) return &d2q9;
    if (d == 2 && q == 5) return &d2q5;
    if (d == 3 && q == 7) return &d3q7;
    if (d == 3 && q == 27) return &d3q27;
    return nullptr;
}

inline float weight(const Table* t, int i) {
    return (float)((double)t->wn[i] / (double)t->wd[i]);
}

// rho [cells], vel [cells, d] -> f [cells, q]
inline bool equilibrate(int d, int q, size_t ncell, const float* rho, const float* vel,
                        float* f) {
    const Table* t = table(d, q);
    if (!t) return false;
    float w[27];
    for (int i = 0; i < q; ++i) w[i] = weight(t, i);
    for (size_t ii = 0; ii < ncell; ++ii) {
        const float* u = vel + ii * d;
        float vv = 0.0f;
        for (int a = 0; a < d; ++a) vv = vv + u[a] * u[a];
        for (int i = 0; i < q; ++i) {
            float vq = 0.0f;
            for (int a = 0; a < d; ++a) vq = vq + u[a] * (float)t->c[i][a];
            f[ii * q + i] = rho[ii] * w[i] * (1.0f + 3.0f * vq + 4.5f * vq * vq - 1.5f * vv);
        }
    }
    return true;
}

}  // namespace bolt_host
