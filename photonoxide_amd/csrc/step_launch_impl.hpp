// The definitions of the step launches declared in step_launch.hpp. Included only by the
// step_*.hip units, each of which instantiates one precision and coefficient mode.
#pragma once
#include "step_launch.hpp"
#include "march64.hpp"
#include "march2d.hpp"
#include "march2s.hpp"

namespace pxfdtd {

template <typename T, int MODE, int TY>
static int launch_ty(const StepArgs<T> &A, hipStream_t s, size_t lds) {
    hipLaunchKernelGGL((fused_step<T, MODE, TY>), dim3(8 * A.per_xcd), dim3(TX, TY + 1), lds, s, A);
    return PXFDTD_OK;
}

template <typename T, int MODE, int RY, bool PF>
static int launch_march(const StepArgs<T> &A0, hipStream_t s, size_t lds) {
    // the units touching no slab, then those that do, each spread over the 8 XCDs
    StepArgs<T> A = A0;
    const int lean = A0.blocks, slab = A0.per_xcd;  // counts passed through blocks, per_xcd
    if (lean) {
        A.blocks = lean;
        A.per_xcd = ((lean + 3) / 4 + 7) / 8;
        hipLaunchKernelGGL((march_step<T, MODE, RY, false, PF>), dim3(8 * A.per_xcd), dim3(256), lds, s, A);
    }
    if (slab) {
        A.units = A0.units + lean;
        A.blocks = slab;
        A.per_xcd = ((slab + 3) / 4 + 7) / 8;
        hipLaunchKernelGGL((march_step<T, MODE, RY, true, false>), dim3(8 * A.per_xcd), dim3(256), lds, s, A);
    }
    return PXFDTD_OK;
}

// The units touching no CPML slab and those that do, as two launches: with a side stream the
// slab launch (the longer per unit) runs there beside the lean one, so neither's tail idles
// the chip; the main stream waits for both before the step's next work.
template <typename T, int MODE, int RY, int AX>
static void launch_units(const StepArgs<T> &A0, int first, int count, hipStream_t s, size_t lds) {
    if (!count) return;
    StepArgs<T> A = A0;
    A.units = A0.units + first;
    A.blocks = count;
    A.per_xcd = ((count + 3) / 4 + 7) / 8;
    hipLaunchKernelGGL((march64_step<T, MODE, RY, AX>), dim3(8 * A.per_xcd), dim3(256), lds, s, A);
}

// the slab units at one row each, in the kernel held to four waves per SIMD
template <typename T, int MODE>
static void launch_units_4w(const StepArgs<T> &A0, int first, int count, hipStream_t s, size_t lds) {
    if (!count) return;
    StepArgs<T> A = A0;
    A.units = A0.units + first;
    A.blocks = count;
    A.per_xcd = ((count + 3) / 4 + 7) / 8;
    hipLaunchKernelGGL((march64_step_4w<T, MODE, 1, 7>), dim3(8 * A.per_xcd), dim3(256), lds, s, A);
}

// The units touching no CPML slab, those touching only the x slabs and the others, each in the
// variant that compiles in just that psi work. With a side stream the slab launch runs there
// beside the others; the main stream waits for it before the step's next work.
template <typename T, int MODE, int RY>
static int launch_march64(const StepArgs<T> &A, hipStream_t s, size_t lds, const SideStream *side,
                          const UnitCounts &n) {
    const bool both = (n.lean || n.x) && n.slab && side;
    hipStream_t ss = both ? side->s : s;
    if (both) {
        HIPCHK(hipEventRecord(side->fork, s));
        HIPCHK(hipStreamWaitEvent(ss, side->fork, 0));
    }
    if (side && side->first) TRY(side->first(ss));
    if (n.slab_rows == 1)
        launch_units_4w<T, MODE>(A, n.lean + n.x, n.slab, ss, lds);
    else
        launch_units<T, MODE, RY, 7>(A, n.lean + n.x, n.slab, ss, lds);
    launch_units<T, MODE, RY, 0>(A, 0, n.lean, s, lds);
    launch_units<T, MODE, RY, 1>(A, n.lean, n.x, s, lds);
    if (both) {
        HIPCHK(hipEventRecord(side->join, ss));
        HIPCHK(hipStreamWaitEvent(s, side->join, 0));
    }
    return PXFDTD_OK;
}

template <typename T, int MODE, int AX, int SET>
static void launch_units_2d(const StepArgs<T> &A0, int first, int count, hipStream_t s, size_t lds) {
    if (!count) return;
    StepArgs<T> A = A0;
    A.units = A0.units + first;
    A.blocks = count;
    A.per_xcd = ((count + 3) / 4 + 7) / 8;
    hipLaunchKernelGGL((march2d_step<T, MODE, AX, SET>), dim3(8 * A.per_xcd), dim3(256), lds, s, A);
}

template <typename T, int MODE, int SET>
static int launch_march2d(const StepArgs<T> &A, hipStream_t s, size_t lds, const UnitCounts &n) {
    launch_units_2d<T, MODE, 3, SET>(A, n.lean, n.slab, s, lds);
    launch_units_2d<T, MODE, 0, SET>(A, 0, n.lean, s, lds);
    return PXFDTD_OK;
}

template <typename T, int MODE>
int launch_mode(const StepArgs<T> &A, int kernel, int ty, hipStream_t s, size_t lds,
                       const SideStream *side, const UnitCounts &n, int set) {
    if (kernel == 4) {
        switch (set) {
        case 1: return launch_march2d<T, MODE, 1>(A, s, lds, n);
        case 2: return launch_march2d<T, MODE, 2>(A, s, lds, n);
        default: return launch_march2d<T, MODE, 3>(A, s, lds, n);
        }
    }
    if (kernel == 3) {
        switch (ty) {
        case 1: return launch_march64<T, MODE, 1>(A, s, lds, side, n);
        case 2: return launch_march64<T, MODE, 2>(A, s, lds, side, n);
        default: return launch_march64<T, MODE, 4>(A, s, lds, side, n);
        }
    }
    if (kernel == 1) {
        switch (ty) {
        case 1: return launch_march<T, MODE, 1, true>(A, s, lds);
        case 2: return launch_march<T, MODE, 2, true>(A, s, lds);
        default: return launch_march<T, MODE, 4, true>(A, s, lds);
        }
    }
    if (kernel == 2) {
        switch (ty) {
        case 1: return launch_march<T, MODE, 1, false>(A, s, lds);
        case 2: return launch_march<T, MODE, 2, false>(A, s, lds);
        default: return launch_march<T, MODE, 4, false>(A, s, lds);
        }
    }
    switch (ty) {
    case 2: return launch_ty<T, MODE, 2>(A, s, lds);
    case 4: return launch_ty<T, MODE, 4>(A, s, lds);
    default: return launch_ty<T, MODE, 8>(A, s, lds);
    }
}

template <typename T, int MODE, int RY, int AX>
static void launch_pair_units(const StepArgs<T> &A0, const PairTaps<T> &TP, int first, int count, hipStream_t s,
                              size_t lds) {
    if (!count) return;
    StepArgs<T> A = A0;
    A.units = A0.units + first;
    A.blocks = count;
    A.per_xcd = ((count + 3) / 4 + 7) / 8;
    hipLaunchKernelGGL((march2s_step<T, MODE, RY, AX>), dim3(8 * A.per_xcd), dim3(256), lds, s, A, TP);
}

template <typename T, int MODE, int RY>
static int launch_pair_ry(const StepArgs<T> &A, const PairTaps<T> &TP, hipStream_t s, size_t lds,
                          const PairStreams *side, const PairCounts &n) {
    const int count[4] = {n.x, n.y, n.z, n.rest};
    const int first[4] = {n.lean, n.lean + n.x, n.lean + n.x + n.y, n.lean + n.x + n.y + n.z};
    if (side) HIPCHK(hipEventRecord(side->fork, s));
    for (int c = 0; c < 4; c++) {
        if (!count[c]) continue;
        hipStream_t sc = side ? side->s[c] : s;
        if (side) HIPCHK(hipStreamWaitEvent(sc, side->fork, 0));
        if (c == 0) launch_pair_units<T, MODE, RY, 1>(A, TP, first[c], count[c], sc, lds);
        if (c == 1) launch_pair_units<T, MODE, RY, 2>(A, TP, first[c], count[c], sc, lds);
        if (c == 2) launch_pair_units<T, MODE, RY, 4>(A, TP, first[c], count[c], sc, lds);
        if (c == 3) launch_pair_units<T, MODE, RY, 7>(A, TP, first[c], count[c], sc, lds);
        if (side) HIPCHK(hipEventRecord(side->join[c], sc));
    }
    launch_pair_units<T, MODE, RY, 0>(A, TP, 0, n.lean, s, lds);
    if (side)
        for (int c = 0; c < 4; c++)
            if (count[c]) HIPCHK(hipStreamWaitEvent(s, side->join[c], 0));
    return PXFDTD_OK;
}

template <typename T, int MODE>
int launch_pair(const StepArgs<T> &A, const PairTaps<T> &TP, int ry, hipStream_t s, size_t lds,
                const PairStreams *side, const PairCounts &n) {
    if (ry == 3) return launch_pair_ry<T, MODE, 3>(A, TP, s, lds, side, n);
    return launch_pair_ry<T, MODE, 2>(A, TP, s, lds, side, n);
}

}  // namespace pxfdtd
