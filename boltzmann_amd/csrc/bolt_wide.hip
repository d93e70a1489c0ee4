// The wide fluid steps of libbolt_hip.so (update_wide.hpp: 2 or 4 cells per lane, no fused
// tracer), in a translation unit of their own so that they compile beside bolt_runtime.hip.
// bolt_lookup_wide returns the kernel for a (D, Q, collision, storage, cells per lane,
// parity), or null where none is built (double storage takes 2 cells per lane at most).
#include <hip/hip_runtime.h>

#include "update_wide.hpp"

using namespace bolt;

using WideFn = void (*)(StepParams, WideGeom);

template <int D, int Q, int C, int S, int VW>
static WideFn wide_fn(bool even) {
    return even ? (WideFn)lbm_step_wide<D, Q, C, S, true, VW> : (WideFn)lbm_step_wide<D, Q, C, S, false, VW>;
}

template <int D, int Q, int C>
static WideFn wide_fn_s(int s, int vw, bool even) {
    switch (s) {
        case kF32: return vw == 4 ? wide_fn<D, Q, C, kF32, 4>(even) : wide_fn<D, Q, C, kF32, 2>(even);
        case kF32F: return vw == 4 ? wide_fn<D, Q, C, kF32F, 4>(even) : wide_fn<D, Q, C, kF32F, 2>(even);
        case kF16S: return vw == 4 ? wide_fn<D, Q, C, kF16S, 4>(even) : wide_fn<D, Q, C, kF16S, 2>(even);
        case kBF16S: return vw == 4 ? wide_fn<D, Q, C, kBF16S, 4>(even) : wide_fn<D, Q, C, kBF16S, 2>(even);
        case kF64: return vw == 2 ? wide_fn<D, Q, C, kF64, 2>(even) : nullptr;
    }
    return nullptr;
}

WideFn bolt_lookup_wide(int d, int q, int coll, int s, int vw, bool even) {
    if (vw != 2 && vw != 4) return nullptr;
    if (d == 2 && q == 9) return coll == 0 ? wide_fn_s<2, 9, 0>(s, vw, even) : wide_fn_s<2, 9, 1>(s, vw, even);
    if (d == 3 && q == 27)
        return coll == 0 ? wide_fn_s<3, 27, 0>(s, vw, even) : wide_fn_s<3, 27, 1>(s, vw, even);
    return nullptr;
}

