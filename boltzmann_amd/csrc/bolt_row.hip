// The row-staged even steps of libbolt_hip.so (update.hpp lbm_step_row: one work-group per
// row of the contiguous axis, every population access an aligned row, the shift along the
// row taken in LDS; with or without a fused tracer), in a translation unit of their own so
// that they compile beside bolt_runtime.hip.  bolt_lookup_row returns the kernel for a (D,
// Q, collision, storage, fused tracer's Q or 0), a work-group of `block` threads (block
// <= 256, or <= 1024 where an instantiation for long rows exists) and a population load
// form (streaming or ordinary), or null.
#include <hip/hip_runtime.h>

#include "update.hpp"

using namespace bolt;

using StepFn = void (*)(StepParams);

template <int D, int Q, int C, int S, int TQ, bool STREAM>
static StepFn row_fn(int block) {
    if (block <= 256) return (StepFn)lbm_step_row<D, Q, C, S, 256, TQ, STREAM>;
    // long rows: D2Q9 in every storage, D3Q27 in the formats whose registers fit 1024
    // work-items (double storage and fused tracers do not)
    if constexpr (Q == 9 || (S != kF64 && TQ == 0)) return (StepFn)lbm_step_row<D, Q, C, S, 1024, TQ, STREAM>;
    return nullptr;
}

template <int D, int Q, int C, int S, bool STREAM>
static StepFn row_fn_t(int tq, int block) {
    if (tq == 0) return row_fn<D, Q, C, S, 0, STREAM>(block);
    if constexpr (S == kF32 || S == kF64) {
        constexpr int TQ = D == 2 ? 5 : 7;
        if (tq == TQ) return row_fn<D, Q, C, S, TQ, STREAM>(block);
    }
    return nullptr;
}

template <int D, int Q, int C, bool STREAM>
static StepFn row_fn_s(int s, int tq, int block) {
    switch (s) {
        case kF32: return row_fn_t<D, Q, C, kF32, STREAM>(tq, block);
        case kF32F: return row_fn_t<D, Q, C, kF32F, STREAM>(tq, block);
        case kF16S: return row_fn_t<D, Q, C, kF16S, STREAM>(tq, block);
        case kBF16S: return row_fn_t<D, Q, C, kBF16S, STREAM>(tq, block);
        case kF64: return row_fn_t<D, Q, C, kF64, STREAM>(tq, block);
    }
    return nullptr;
}

template <bool STREAM>
static StepFn row_table(int d, int q, int coll, int s, int tq, int block) {
    if (d == 2 && q == 9)
        return coll == 0 ? row_fn_s<2, 9, 0, STREAM>(s, tq, block) : row_fn_s<2, 9, 1, STREAM>(s, tq, block);
    if (d == 3 && q == 27)
        return coll == 0 ? row_fn_s<3, 27, 0, STREAM>(s, tq, block) : row_fn_s<3, 27, 1, STREAM>(s, tq, block);
    return nullptr;
}

StepFn bolt_lookup_row(int d, int q, int coll, int s, int tq, int block, bool stream) {
    if (block < 64 || block > 1024 || block % 64 != 0) return nullptr;
    return stream ? row_table<true>(d, q, coll, s, tq, block) : row_table<false>(d, q, coll, s, tq, block);
}

// LDS elements one work-group of the row-staged step needs per cell of its row, for a
// fluid (q = 9, 27) or a tracer (q = 5, 7) velocity set; 0 for q = 0.
int bolt_row_slots(int d, int q) {
    if (d == 2 && q == 9) return row_slots<VSet<2, 9>>();
    if (d == 3 && q == 27) return row_slots<VSet<3, 27>>();
    if (d == 2 && q == 5) return row_slots<VSet<2, 5>>();
    if (d == 3 && q == 7) return row_slots<VSet<3, 7>>();
    return 0;
}
