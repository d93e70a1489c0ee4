// Host emulation of the wide AA kernel's access map (boltzmann_amd/csrc/update_wide.hpp):
// for every wavefront of a launch it replays, lane by lane, which elements of a plane the
// kernel loads and stores on the even (streaming) step -- the aligned vector of each active
// lane, the one-lane rotation across the wavefront (ror1 / rol1, emulated), the edge
// lanes' single-element loads and stores -- using the header's own geometry functions
// (make_wide_geom, wave_task, halo_prev, halo_next, row_shift).  For each (D, shape, cells
// per lane, shift) it checks that
//   * every address lies inside the plane (rows * len elements),
//   * after the load, cell z of a row holds element z + dz of the shifted row (periodic),
//     i.e. F[x - c_i, i] of the AA pattern,
//   * the stores write every element of the plane exactly once, and element e receives the
//     value of the cell whose AA target it is.
// Prints "ok" or the first failure.
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "../../boltzmann_amd/csrc/update_wide.hpp"

using namespace bolt;

static int fails = 0;
#define CHECK(c, ...)                                 \
    do {                                              \
        if (!(c)) {                                   \
            if (fails++ < 5) {                        \
                printf("FAIL: " __VA_ARGS__);         \
                printf("\n");                         \
            }                                         \
        }                                             \
    } while (0)

// One plane of one shape: the even step of a population with row shift (rdx, rdy) and
// shift dz along the row.  Values are element indices (loads) and cell indices (stores).
template <int D>
static void check(int n0, int n1, int n2, int vw, int rdx, int rdy, int dz) {
    const int len = D == 2 ? n1 : n2;
    const long long ncell = (long long)n0 * n1 * (D == 3 ? n2 : 1);
    const WideGeom g = make_wide_geom(ncell, len, vw);
    const long long nel = g.rows * (long long)len;
    std::vector<int> written(nel, 0);
    std::vector<long long> wval(nel, -1);
    for (long long task = 0; task < g.tasks; ++task) {
        const WaveTask t = wave_task(g, task);
        const uint32_t rb = row_shift<D>(t.row, rdx, rdy, n0, n1) * (uint32_t)len;
        CHECK(t.last >= 0 && t.last < 64 && t.zb < len, "task %lld: zb %d last %d", task, t.zb, t.last);
        // loads: v[lane][k] = element index loaded into slot k of the lane's vector
        std::vector<long long> v(64 * vw);
        for (int lane = 0; lane < 64; ++lane) {
            const bool active = lane <= t.last;
            const long long base = (long long)rb + (active ? t.zb + vw * lane : t.zb);
            CHECK(base >= 0 && base + vw <= nel, "load out of the plane");
            for (int k = 0; k < vw; ++k) v[lane * vw + k] = base + k;
        }
        std::vector<long long> s = v;
        if (dz != 0) {
            for (int lane = 0; lane < 64; ++lane)
                for (int k = 0; k < vw; ++k) {
                    int sl = lane, sk = k + dz;  // shift_prev (dz < 0) / shift_next (dz > 0)
                    if (sk < 0) { sl = (lane + 63) & 63; sk = vw - 1; }       // ror1
                    if (sk >= vw) { sl = (lane + 1) & 63; sk = 0; }            // rol1
                    s[lane * vw + k] = v[sl * vw + sk];
                }
            if (!g.fullrow) {
                if (dz < 0) s[0] = (long long)rb + halo_prev(g, t);
                else s[t.last * vw + vw - 1] = (long long)rb + halo_next(g, t);
            }
        }
        for (int lane = 0; lane <= t.last; ++lane)
            for (int k = 0; k < vw; ++k) {
                const int z = t.zb + vw * lane + k;
                const long long want = (long long)rb + ((z + dz) % len + len) % len;
                CHECK(s[lane * vw + k] == want, "D%d %dx%dx%d vw %d dz %d: cell %d of row %u loads %lld, want %lld",
                      D, n0, n1, n2, vw, dz, z, t.row, s[lane * vw + k], want);
            }
        // stores: cell z sends its value to element z + dz of the shifted row
        std::vector<long long> val(64 * vw, -1);
        for (int lane = 0; lane < 64; ++lane)
            for (int k = 0; k < vw; ++k) val[lane * vw + k] = (long long)t.row * len + t.zb + vw * lane + k;
        auto put = [&](long long e, long long x) {
            CHECK(e >= 0 && e < nel, "store out of the plane");
            if (e < 0 || e >= nel) return;
            written[e]++;
            wval[e] = x;
        };
        if (dz == 0) {
            for (int lane = 0; lane <= t.last; ++lane)
                for (int k = 0; k < vw; ++k) put((long long)rb + t.zb + vw * lane + k, val[lane * vw + k]);
            continue;
        }
        std::vector<long long> w(64 * vw);
        for (int lane = 0; lane < 64; ++lane)
            for (int k = 0; k < vw; ++k) {
                int sl = lane, sk = k - dz;  // dz < 0: shift_next; dz > 0: shift_prev
                if (sk < 0) { sl = (lane + 63) & 63; sk = vw - 1; }
                if (sk >= vw) { sl = (lane + 1) & 63; sk = 0; }
                w[lane * vw + k] = val[sl * vw + sk];
            }
        for (int lane = 0; lane <= t.last; ++lane) {
            const long long at = (long long)rb + t.zb + vw * lane;
            int k0 = 0, k1 = vw;  // elements of the vector this lane stores
            if (!g.fullrow) {
                if (dz < 0 && lane == t.last) k1 = vw - 1;
                if (dz > 0 && lane == 0) k0 = 1;
            }
            for (int k = k0; k < k1; ++k) put(at + k, w[lane * vw + k]);
        }
        if (!g.fullrow) {
            if (dz < 0) put((long long)rb + halo_prev(g, t), val[0]);
            else put((long long)rb + halo_next(g, t), val[t.last * vw + vw - 1]);
        }
    }
    // every element written once, with the value of the cell that targets it
    for (long long r = 0; r < g.rows; ++r) {
        for (int z = 0; z < len; ++z) {
            const long long e = r * len + z;
            CHECK(written[e] == 1, "D%d %dx%dx%d vw %d dz %d: element %lld written %d times", D, n0, n1,
                  n2, vw, dz, e, written[e]);
        }
    }
    // the value at element (row', z') came from cell (row, z' - dz) with row' = shift(row)
    for (long long r = 0; r < g.rows; ++r) {
        const uint32_t rs = row_shift<D>((uint32_t)r, rdx, rdy, n0, n1);
        for (int z = 0; z < len; ++z) {
            const long long e = (long long)rs * len + ((z + dz) % len + len) % len;
            CHECK(wval[e] == r * len + z, "D%d vw %d dz %d: element %lld holds cell %lld, want %lld", D, vw, dz,
                  e, wval[e], r * len + z);
        }
    }
}

int main() {
    const int vws[] = {2, 4};
    // 2D: rows = n0, contiguous n1
    const int n2d[][2] = {{5, 256}, {3, 512}, {4, 128}, {3, 200}, {2, 520}, {3, 1024}, {2, 4}, {3, 8}, {2, 132}, {1, 64}};
    for (auto& s : n2d)
        for (int vw : vws) {
            if (s[1] % vw) continue;
            for (int dx = -1; dx <= 1; ++dx)
                for (int dz = -1; dz <= 1; ++dz) check<2>(s[0], s[1], 1, vw, dx, 0, dz);
        }
    // 3D: rows = n0 * n1, contiguous n2
    const int n3d[][3] = {{3, 4, 256}, {2, 3, 128}, {2, 2, 520}, {3, 2, 68}, {2, 3, 512}, {1, 1, 4}, {2, 2, 300}};
    for (auto& s : n3d)
        for (int vw : vws) {
            if (s[2] % vw) continue;
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dz = -1; dz <= 1; ++dz) check<3>(s[0], s[1], s[2], vw, dx, dy, dz);
        }
    if (fails) {
        printf("%d failures\n", fails);
        return 1;
    }
    printf("ok\n");
    return 0;
}
