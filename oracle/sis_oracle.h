/*
 * CPU oracle for the PT-SA square-packing path.  TEST INFRASTRUCTURE ONLY:
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as the checker.  Nothing in the product links it.
 *
 * Every function restates the reference's algorithm (koboshchan/squares-in-
 * squares, file:line cited per function in sis_oracle.cpp) in plain C++17
 * compiled with g++ -ffp-contract=off.  Parity status of each group is stated
 * in the header comment of sis_oracle.cpp.
 */
#ifndef SIS_ORACLE_H
#define SIS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter-based RNG (Salmon et al., SC'11) ---- */
void o_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float o_u01f(uint32_t r);
double o_u01d(uint32_t hi, uint32_t lo);

/* ---- reference FP32 geometry (cuda/packing.cu:22-47) ---- */
int o_ref_cell(float x, float y);
float o_ref_depth32(const float p[3], const float q[3]);
float o_ref_wall32(const float p[3], float L);
void o_ref_local32(int n, const float* q, float L, float* local_out, float* full_out);

/* ---- the product's FP32 formulation (cached cos/sin, explicit fma) ---- */
int o_cellcoord(float x, float shift, float w, int lo, int hi);
float o_depth32_cs(float dx, float dy, float ci, float si, float cj, float sj);
float o_depth2_32_cs(float dx, float dy, float ci, float si, float cj, float sj);
float o_wall32_cs(float x, float y, float c, float s, float L);
void o_rot32(float c, float s, float t, float* c_out, float* s_out);
float o_accept_exp2(float x);

/* ---- reference FP64 geometry (cuda/packing.cu:163-297, cuda/seeded.cu:40-113) ---- */
double o_pair64(const double p[3], const double q[3], double grad[3]);
double o_pair64_smooth_cs(const double p[3], const double q[3], double c, double s,
                          double d, double t, double grad[3]);
double o_boundary64(const double p[3], double L, double grad[3]);
double o_sat_oracle64(const double p[3], const double q[3]);
double o_sat_oracle_depth_f(const float p[3], const float q[3]);
double o_total64(int n, const double* q, double L);
double o_total64_cs(int n, const double* x, const double* y, const double* c,
                    const double* s, double L);
double o_certified_side(int n, const double* q, double current_best, double* out,
                        double* lam_out);
int o_polish64(int n, double* q, double L, int iters, double* energy_inout);
double o_wrap_angle(double a);

/* ---- CPU PT-SA engine (FP64), the reported CPU baseline ---- */
typedef struct o_ptsa_cfg {
    int n;
    double L0;
    const double* init;      /* replicas*n*3 (x, y, angle) or NULL: uniform random */
    int replicas;
    int ladder;              /* temperatures per ladder (replicas % ladder == 0) */
    double tmin, tmax;
    double move_sigma, rot_sigma, target_accept;
    int64_t sweeps;
    int swap_every;          /* sweeps between exchange attempts */
    double shrink;           /* L decrement on a zero-overlap configuration */
    uint64_t seed;
    int threads;
} o_ptsa_cfg;

typedef struct o_ptsa_out {
    double L_final;          /* box side at the end */
    double L_feasible;       /* smallest box side with a zero-energy replica (or L0) */
    uint64_t proposals;
    uint64_t accepted;
    uint64_t exchanges_tried, exchanges_accepted;
    int shrinks;
    double seconds;
    double* best_q;          /* caller-owned n*3: best zero-energy configuration */
    double e_min;            /* lowest replica energy at the end */
} o_ptsa_out;

int o_ptsa_run(const o_ptsa_cfg* cfg, o_ptsa_out* out);

/* ---- CPU model of the product's FP32 colour-checkerboard sweep ---- */
typedef struct o_sweep_cfg {
    int n;
    float L;
    float w;                 /* cell width, >= sqrt(2)/2 */
    float step_xy, step_th;  /* half-widths of the uniform proposal */
    float inv_t_log2e;       /* log2(e) / T */
    uint32_t key0, key1;     /* Philox key (request seed) */
    uint32_t replica;        /* replica index within the request */
    uint32_t sweep;          /* global sweep counter of this replica */
    uint32_t sweeps_per_shift; /* the grid shift of sweep s is drawn at counter
                                  (s / k) k (0 or 1: every sweep) */
    float* dE32;             /* optional, n: the FP32 energy change of square i's proposal this
                                sweep (NaN when rejected before the stencil: it leaves the box
                                or its cell) */
    double* dE64;            /* optional, n: the same change recomputed in FP64 over every pair
                                and the wall, from the same FP32 inputs */
    double* depth64;         /* optional, n: the summed overlap depths of the pairs and wall
                                terms of that change (old and new pose), in FP64 */
} o_sweep_cfg;

/* x, y, c, s: n floats each, updated in place.  Returns accepted moves. */
int o_sweep32_model(const o_sweep_cfg* cfg, float* x, float* y, float* c, float* s,
                    uint64_t counts[3]);
/* Canonical FP32 replica energy, same summation order as the device. */
float o_energy32_model(int n, float L, float w, const float* x, const float* y,
                       const float* c, const float* s);
/* `moves` collective proposals after a launch whose last sweep counter is
 * cfg->sweep - 1 (draws use cfg->sweep = the sweep count after the launch);
 * step_xy / step_th are the slot's half-widths.  Returns accepts; *energy =
 * canonical energy at the end. */
int o_collective32_model(const o_sweep_cfg* cfg, int moves, float* x, float* y, float* c, float* s,
                         float* energy);
/* Kick after interval `interval` (relocating when `relocate`).  Returns 1 when
 * the replica had energy (and was kicked); *energy = canonical energy after. */
int o_kick32_model(const o_sweep_cfg* cfg, uint32_t interval, int relocate, float* x, float* y, float* c,
                   float* s, float* energy);

/* ---- certified side of an FP32 state (cuda/packing.cu:254-297) ---- */
/* xycs: n float4 (x, y, cos, sin); the orientation is scaled to unit length
 * in FP64.  out: side (1e9 when the dilation exceeds 1.008), dilation, lower
 * corner x, y of the dilated configuration. */
void o_certify_xycs(int n, const float* xycs, double out[4]);

/* ---- the exchange step (cuda/packing.cu:141-152, 750-765, 823-853) ---- */
/* Same layouts as the engine's sis_rep_state / sis_slot_state / sis_ladder_desc. */
typedef struct o_rep_state {
    float E; int slot; unsigned acc; unsigned pad;
    double Lc, lam, minx, miny;
    uint64_t cand, near_pairs;           /* sweep-workgroup counters, cleared by the exchange step */
    uint32_t coll_tried, coll_acc, kicks, pad2;
} o_rep_state;
typedef struct o_slot_state {
    float inv_t_log2e, step_xy, step_th, L, T, base_xy, base_th, pad;
} o_slot_state;
typedef struct o_ladder_desc {
    int n, R, ladder, improve, cool_period, spe;
    int64_t sweeps_total, sweeps_done;
    uint32_t key0, key1, interval;
    float shrink, target_acc, scale, Lf;
    double L, L_arch, L_box_arch;
    int arch_valid, shrinks, feasible, fp64_rejects;
    uint64_t ex_tried, ex_acc, certs;
} o_ladder_desc;
/* One exchange step of one instance, in place (xycs: R*n*4, arch: n*4). */
int o_ladder_model(o_ladder_desc* d, o_rep_state* reps, o_slot_state* slots, int* rep_of_slot, float* xycs,
                   float* arch);
/* 2^x as the fixed FP64 sequence of the exchange step, and the reheating
 * factor 0.002^((k mod P) / (P - 1)) with its square root. */
double o_exp2_fixed64(double x);
void o_cool_factor(uint32_t k, int P, double* cool, double* root);

/* ---- analytic fixtures (cuda/scheduler.py:25-74, src/squares/core.py:27-35) ---- */
int o_grid_packing(int n, double* q);

#ifdef __cplusplus
}
#endif
#endif
