// The wide fluid steps with ordinary population loads (for lattices whose working set fits
// the Infinity Cache), instantiated apart from bolt_wide.hip's streaming-load ones so that
// the two compile side by side.
#include <hip/hip_runtime.h>

#include "wide_table.hpp"

using namespace bolt;

WideFn bolt_lookup_wide_keep(int d, int q, int coll, int s, int vw, bool even) {
    return wide_table<false>(d, q, coll, s, vw, even);
}
