// The table of wide fluid steps (update_wide.hpp) for one population load form, shared by
// bolt_wide.hip (streaming loads) and bolt_wide_keep.hip (ordinary loads), which instantiate
// one form each so that the two compile side by side.
#pragma once

#include "update_wide.hpp"

namespace bolt {

using WideFn = void (*)(StepParams, WideGeom);

template <int D, int Q, int C, int S, int VW, bool STREAM>
static WideFn wide_fn(bool even) {
    return even ? (WideFn)lbm_step_wide<D, Q, C, S, true, VW, STREAM>
                : (WideFn)lbm_step_wide<D, Q, C, S, false, VW, STREAM>;
}

template <int D, int Q, int C, bool STREAM>
static WideFn wide_fn_s(int s, int vw, bool even) {
    switch (s) {
        case kF32: return vw == 4 ? wide_fn<D, Q, C, kF32, 4, STREAM>(even) : wide_fn<D, Q, C, kF32, 2, STREAM>(even);
        case kF32F: return vw == 4 ? wide_fn<D, Q, C, kF32F, 4, STREAM>(even) : wide_fn<D, Q, C, kF32F, 2, STREAM>(even);
        case kF16S: return vw == 4 ? wide_fn<D, Q, C, kF16S, 4, STREAM>(even) : wide_fn<D, Q, C, kF16S, 2, STREAM>(even);
        case kBF16S:
            return vw == 4 ? wide_fn<D, Q, C, kBF16S, 4, STREAM>(even) : wide_fn<D, Q, C, kBF16S, 2, STREAM>(even);
        case kF64: return vw == 2 ? wide_fn<D, Q, C, kF64, 2, STREAM>(even) : nullptr;
    }
    return nullptr;
}

// The kernel for a (D, Q, collision, storage, cells per lane, parity), or null where none is
// built (double storage takes 2 cells per lane at most).
template <bool STREAM>
static WideFn wide_table(int d, int q, int coll, int s, int vw, bool even) {
    if (vw != 2 && vw != 4) return nullptr;
    if (d == 2 && q == 9)
        return coll == 0 ? wide_fn_s<2, 9, 0, STREAM>(s, vw, even) : wide_fn_s<2, 9, 1, STREAM>(s, vw, even);
    if (d == 3 && q == 27)
        return coll == 0 ? wide_fn_s<3, 27, 0, STREAM>(s, vw, even) : wide_fn_s<3, 27, 1, STREAM>(s, vw, even);
    return nullptr;
}

}  // namespace bolt
