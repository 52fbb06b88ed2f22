// Launch interface of the gfx950 kernels (host side of sis_kernels.hip).
#pragma once
#include "sis_internal.h"

namespace sis {

struct FinalizeJob {
    int n;
    int iters;                  // FP64 CG polish iterations
    double L;                   // box side the polish runs in
    const float4* src32;        // n (x, y, cos, sin) or null
    const double* src64;        // n*3 (x, y, angle) when src32 is null
    long long out_off;          // double index into q_out
    // When set, src32 is replica 0 of R consecutive replicas and the kernel
    // takes the one with the lowest FP32 energy (the first on ties).
    const RepState* reps;
    int R, pad;
};

struct FinalizeOut {
    double e64_before, e64_after, lambda, L_cert, span;
    int iters;
    int best;                   // replica taken when the job named replicas
    double e32;                 // its FP32 energy
    int pre;                    // 1: the export is the configuration before the polish
    int pad;
};

// threads: kSweepThreads for instances below kBigN squares, kSweepThreadsBig
// for the others (sweep_threads(n)); every tile of one launch uses the same.
// near_pairs: also count every proposal's near pairs (SIS_FLAG_NEAR_PAIRS).
// alone: the tiles are planned for an idle device (Plan::wide); the launch
// takes the kernel compiled for one wavefront per SIMD.
hipError_t launch_sweep(hipStream_t st, int ntiles, size_t lds, const Tile* tiles,
                        InstDesc* insts, RepState* reps, const SlotState* slots, float4* state,
                        int threads, bool near_pairs, bool alone = false);
// Probe: the exchange kernel's zero-overlap test and the all-pairs FP64
// energy of one configuration.
hipError_t launch_zero64(int n, double L, const float4* q, int* zero, double* e64);
// list: ninst instance indices into insts, or null for insts[0..ninst).
// lds: the largest ladder_lds_bytes (n, initial L) over those instances.
// seq: null, or one word per instance in mapped host memory: each workgroup
// that runs a step writes its instance's new interval count to seq[k] (for
// insts[k]), so a host that polls seq[k] sees the step's end before the
// launch has ended.  An instance that has finished its sweeps writes nothing.
hipError_t launch_ladder(hipStream_t st, int ninst, size_t lds, InstDesc* insts, RepState* reps,
                         SlotState* slots, int* rep_of_slot, float4* state, float4* arch,
                         const int* list = nullptr, unsigned* seq = nullptr);
hipError_t launch_finalize(hipStream_t st, int njobs, const FinalizeJob* jobs, FinalizeOut* outs,
                           double* q_out);
// out[k] = insts[list[k]] for k < count (a kernel, so it sees every earlier
// kernel's writes to the descriptors).
hipError_t launch_gather_desc(hipStream_t st, int count, const InstDesc* insts, const int* list,
                              InstDesc* out);
hipError_t launch_depth32(hipStream_t st, int npairs, const float* in, float* depth, float* d2);
hipError_t launch_cellcoord(hipStream_t st, int count, const float* x, float shift, float w,
                            int lo, int hi, int* out);
hipError_t launch_accept_exp2(hipStream_t st, int count, const float* x, float* out);

}  // namespace sis
