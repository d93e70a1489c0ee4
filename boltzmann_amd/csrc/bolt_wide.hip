// The wide fluid steps of libbolt_hip.so (update_wide.hpp: 2 or 4 cells per lane, no fused
// tracer), in translation units of their own so that they compile beside bolt_runtime.hip:
// this one holds the streaming-load kernels, bolt_wide_keep.hip the ordinary-load ones.
// bolt_lookup_wide returns the kernel for a (D, Q, collision, storage, cells per lane,
// parity, load form), or null where none is built (double storage takes 2 cells per lane
// at most).
#include <hip/hip_runtime.h>

#include "wide_table.hpp"

using namespace bolt;

WideFn bolt_lookup_wide_keep(int d, int q, int coll, int s, int vw, bool even);

WideFn bolt_lookup_wide(int d, int q, int coll, int s, int vw, bool even, bool stream) {
    return stream ? wide_table<true>(d, q, coll, s, vw, even) : bolt_lookup_wide_keep(d, q, coll, s, vw, even);
}
