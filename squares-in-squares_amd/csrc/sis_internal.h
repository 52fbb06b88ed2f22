// Internal layout shared by the host runtime and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sis_abi.h"

namespace sis {

constexpr int kMaxN = 1000;
constexpr int kRankCap = 4;              // occupancy ranks with their own group
constexpr int kGroups = 9 * kRankCap;    // colour x rank groups; one more = serial group
constexpr int kSweepThreads = 64;        // sweep workgroup = one wavefront
constexpr int kSweepThreadsBig = 256;    // four wavefronts for instances of kBigN squares or more
// Four-wavefront tiles from 96 squares: alone on the device they run a
// launch of 10 sweeps 24-44% faster than one wavefront (n = 100: 0.133 vs
// 0.174 ms; n = 200: 0.159 vs 0.272 ms) and under load they match it up to
// n = 128 and beat it above (n = 255: 14.0e9 vs 13.1e9 proposals/s).
constexpr int kBigN = 96;
__host__ __device__ inline int sweep_threads(int n) { return n >= kBigN ? kSweepThreadsBig : kSweepThreads; }
constexpr int kLadderThreads = 256;
constexpr int kPolishThreads = 256;
constexpr int kPolishKPT = (kMaxN + kPolishThreads - 1) / kPolishThreads;
constexpr float kMinCell = 0.7075f;      // 2w > sqrt(2) + 6e-4: same colour never interacts

// One resident packing instance.  Written by the host at admission, then
// owned by the device until the instance finishes.
struct InstDesc {
    int n, R, ladder, G;
    int state_off;              // float4 index of (replica 0, square 0)
    int rep_off;                // first global replica / temperature-slot index
    int M;                      // cell-grid side (row stride), sized for the initial L
    int spe;                    // sweeps per exchange interval
    int spsh;                   // sweeps per random shift of the cell grid
    long long sweeps_total;
    long long sweeps_done;
    double L;                   // improve 0: current box side; 1: best certified side so far
    double L_arch;              // improve 0: box side of the archived zero-overlap configuration;
                                // 1: certified side of the archived configuration
    float Lf;                   // L rounded to float: the cell grid covers [0, Lf]; every slot box is <= Lf
    float scale;                // centre scale applied when the next sweep stages the state
    float w;                    // cell width
    float shrink;
    float target_acc;
    int arch_off;               // float4 index of the archive
    uint32_t key0, key1;        // Philox key (request seed)
    uint32_t interval;          // exchange steps done
    int arch_valid;
    int shrinks;
    int feasible;
    int fp64_rejects;           // FP32-zero configurations the FP64 test found overlapping
    int collective;             // collective proposals per replica after each launch's sweeps
    int kick_every;             // intervals between kicks (0: none)
    int kicks;                  // kicks applied (counted once per replica)
    int improve;                // 0: shrink on a zero-overlap replica; 1: certified-side improvement path
    int cool_period;            // reheating cycle in intervals (0: none)
    int certify_every;          // improve = 1: launches between certified-side passes
    double L_box_arch;          // box side the archived configuration was in
    int seeded;                 // replica 0 started from the caller's configuration (the initial archive)
    float E_arch;               // FP32 energy of the archived configuration when it was archived
    double Lmin;                // sqrt(n): no box goes below it
    unsigned long long proposals, accepted, cand, near, ex_tried, ex_acc;
    unsigned long long coll_tried, coll_acc;
    unsigned long long certs;   // improve = 1: replica checkpoints with a certified side (dilation <= 1.008)
};

// Per-replica scratch of the collective moves and the kick (LDS).
struct CollWork {
    float cx, cy, cc, cs;       // anchor square
    float mv, rt, tc, ts;       // translation, rotation cos/sin
    float th, uacc, eold, enew;
    int mode, anchor, accept, worst;
};

struct Tile {
    int inst, r0, g;
    int wide;                   // stage proposals up to kMaxN squares (sweep_prop_in_lds)
};

// Per global replica and per global temperature slot: the records of the C
// ABI's exchange-step probe, used as they are.
using RepState = sis_rep_state;
using SlotState = sis_slot_state;
static_assert(sizeof(RepState) == 80 && sizeof(SlotState) == 32, "record layouts");

// The reference's retarget multipliers by temperature class (cuda/packing.cu:503, 841).
constexpr double kTargetMults[8] = {0.02, 0.1, 0.4, 1.0, 2.5, 5.0, 10.0, 20.0};
// Certified side (cuda/packing.cu:254-297): pairs nearer than sqrt(2.05), margin
// tau, dilations above 1.008 rejected (side 1e9).
constexpr double kCertTau = 2.6e-8;
constexpr double kCertNear2 = 2.05;
constexpr double kCertMaxLam = 1.008;
constexpr double kCertNone = 1e9;
// 2^x in FP64 operations that each round once or not at all (round to
// nearest integer, an exact subtraction, a Horner chain of fused
// multiply-adds over the Taylor coefficients (ln 2)^j / j!, j <= 13, an exact
// power-of-two scaling), so the host, the device and the oracle compute the
// same bits.  Relative error below 1e-16 for |x| < 1000.
__host__ __device__ inline double exp2_fixed64(double x) {
    const double k = rint(x);
    const double f = x - k;
    double p = 1.3691488853904128e-12;
    p = fma(p, f, 2.5678435993488206e-11);
    p = fma(p, f, 4.4455382718708116e-10);
    p = fma(p, f, 7.054911620801123e-09);
    p = fma(p, f, 1.01780860092397e-07);
    p = fma(p, f, 1.321548679014431e-06);
    p = fma(p, f, 1.5252733804059841e-05);
    p = fma(p, f, 0.0001540353039338161);
    p = fma(p, f, 0.0013333558146428443);
    p = fma(p, f, 0.009618129107628477);
    p = fma(p, f, 0.05550410866482158);
    p = fma(p, f, 0.24022650695910072);
    p = fma(p, f, 0.6931471805599453);
    p = fma(p, f, 1.0);
    return ldexp(p, (int)k);
}
// Reheating factor of interval k (cuda/packing.cu:516): 0.002^((k mod P) /
// (P - 1)), and its square root for the step sizes (:523); 1 without a cycle.
__host__ __device__ inline void cool_factor(uint32_t k, int P, double& cool, double& root) {
    if (P < 2) {
        cool = 1.0;
        root = 1.0;
        return;
    }
    const double x = (double)(k % (uint32_t)P) / (double)(P - 1) * -8.965784284662087;  // log2(0.002)
    cool = exp2_fixed64(x);
    root = exp2_fixed64(0.5 * x);
}
// Effective slot values of an interval: 1/T in log2 units and the step
// half-widths, from the ladder temperature and the adapted bases.
__host__ __device__ inline void slot_effective(SlotState& s, double cool, double root) {
    s.inv_t_log2e = (float)(1.4426950408889634 / ((double)s.T * cool));
    s.step_xy = (float)((double)s.base_xy * root);
    s.step_th = (float)((double)s.base_th * root);
}

// LDS of one sweep tile (G replicas of n squares, M x M cells).
// Cell starts per replica: M*M + 2 u16, padded to 16 bytes.
__host__ __device__ inline int grid_words(int M) { return (M * M + 2 + 7) & ~7; }
// Small tiles stage every square's proposal of the sweep (variates, proposed
// position and orientation, wall-energy change) into LDS in one pass over all
// lanes before the first level, so the sparse levels only evaluate the
// stencil; large tiles move one square per lane and propose inline, keeping
// their LDS, and with it the tiles resident per compute unit, smaller.  When
// the launch leaves compute units to spare (`wide`, one replica per tile) the
// residency does not matter and tiles up to kMaxN squares stage too (n = 1000
// alone: 0.351 vs 0.382 ms per 10-sweep launch; under load staging them costs
// 10-23 % of the throughput).
constexpr int kPropLdsMaxSquares = 512;
__host__ __device__ inline bool sweep_prop_in_lds(int gn, bool wide) {
    return gn <= (wide ? kMaxN : kPropLdsMaxSquares);
}
// Tiles up to this size also keep their positions in cell-list order (one
// LDS read per candidate instead of an index read and a position read).
// Larger tiles skip the copy so that more of them fit a compute unit
// (measured at 2048 replicas: n=1000 15.4e9 proposals/s with it, 17.0e9
// without; n=256..600 unchanged).
constexpr int kSposLdsMaxSquares = 768;
__host__ __device__ inline bool sweep_spos_in_lds(int gn) { return gn <= kSposLdsMaxSquares; }
// One-wavefront tiles that stage their proposals write them as 80-byte mover
// records in level order (position, move, wall change, uniform, 1/T, indices
// and the stencil's row offsets), so that a pass reads everything but the
// candidates in one LDS round trip.  Under load only tiles of up to 64
// squares use them: at 80 squares the extra LDS costs residency (n = 80,
// 4096 replicas: 6.3e9 proposals/s with records, 8.9e9 without; n = 64:
// 8.2e9 with, 8.0e9 without).
constexpr int kMoverRecBytes = 80;
constexpr int kRecLoadedMaxSquares = 64;
__host__ __device__ inline bool sweep_records(int n, int gn, bool wide) {
    return n < kBigN && sweep_prop_in_lds(gn, wide) && sweep_spos_in_lds(gn) &&
           (wide || gn <= kRecLoadedMaxSquares);
}
__host__ __device__ inline size_t sweep_lds_bytes(int n, int M, int G, bool wide) {
    size_t gn = (size_t)G * n;
    size_t b = 16 * gn;                         // float4 x, y, cos, sin
    if (sweep_records(n, (int)gn, wide))
        b += kMoverRecBytes * gn;               // mover records
    else if (sweep_prop_in_lds((int)gn, wide))
        b += 32 * gn;                           // staged proposals by level item
    if (sweep_spos_in_lds((int)gn)) b += 16 * gn;  // float4 positions in cell-list order
    b += 2 * (size_t)G * grid_words(M);         // cell starts (u16)
    b += 2 * gn * 5;                            // sorted, cell coords, rank/key, items, list slot
    b = (b + 15) & ~(size_t)15;
    b += (size_t)G * 8;                         // per-replica occupied-group mask
    b += (size_t)G * sizeof(CollWork);          // collective-move / kick scratch
    b += (size_t)G * 11 * 4;                    // shift x/y (this and next sweep), steps, 1/T, box, accepts, energy,
                                                // largest overlap
    b += 2 * (kGroups + 2) * 4;                 // level counts / cursors, level starts
    return b;
}

// Dynamic LDS of the exchange kernel for an instance of n squares, box side
// L and R replicas: first the ladder by slot, then, reusing it, FP64 x, y,
// cos, sin of the replica under the zero-overlap test and that test's uniform
// grid (cells of width kZeroCell > sqrt(2.000001)): cell starts, scan
// partials, a flag and the square indices in cell order.
constexpr double kZeroCell = 1.5;
constexpr int kZeroGridMax = 45;         // cells per side for L <= 64
__host__ __device__ inline int zero_grid_side(double L) {
    const int m = (int)(L / kZeroCell) + 2;
    return m < kZeroGridMax ? m : kZeroGridMax;
}
__host__ __device__ inline size_t ladder_lds_bytes(int n, double L, int R) {
    const size_t mc = (size_t)zero_grid_side(L);
    size_t b = 32 * (size_t)n;               // x, y, cos, sin (FP64)
    b += 4 * (mc * mc + 1);                  // cell starts
    b += 4 * (size_t)kLadderThreads + 4;     // scan partials, flag
    b += 2 * (size_t)n;                      // square indices in cell order
    const size_t ladder = 16 * (size_t)R;    // before the test: energy, 1/T, replica, acceptance per slot
    return ((b > ladder ? b : ladder) + 15) & ~(size_t)15;
}

__host__ __device__ inline int grid_cells(float L, float w) {
    return (int)floorf((L + w) / w) + 1;
}

}  // namespace sis
