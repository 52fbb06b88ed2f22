// Launch interface of the gfx950 kernels (host side of sis_kernels.hip).
#pragma once
#include "sis_internal.h"

namespace sis {

struct FinalizeJob {
    int n;
    int iters;                  // FP64 CG polish iterations
    double L;                   // box side the source configuration was feasible in
    const float4* src32;        // n (x, y, cos, sin) or null
    const double* src64;        // n*3 (x, y, angle) when src32 is null
    long long out_off;          // double index into q_out
};

struct FinalizeOut {
    double e64_before, e64_after, lambda, L_cert, span;
    int iters, pad;
};

hipError_t launch_sweep(hipStream_t st, int ntiles, size_t lds, const Tile* tiles,
                        InstDesc* insts, RepState* reps, const SlotState* slots, float4* state);
// list: ninst instance indices into insts, or null for insts[0..ninst).
hipError_t launch_ladder(hipStream_t st, int ninst, InstDesc* insts, RepState* reps,
                         SlotState* slots, int* rep_of_slot, float4* state, float4* arch,
                         const int* list = nullptr);
hipError_t launch_finalize(hipStream_t st, int njobs, const FinalizeJob* jobs, FinalizeOut* outs,
                           double* q_out);
hipError_t launch_depth32(hipStream_t st, int npairs, const float* in, float* depth, float* d2);
hipError_t launch_cellcoord(hipStream_t st, int count, const float* x, float shift, float w,
                            int lo, int hi, int* out);
hipError_t sweep_kernel_attrs(int* max_lds);

}  // namespace sis
