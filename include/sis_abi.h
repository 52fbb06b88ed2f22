/*
 * squares-in-squares on MI355X: C ABI of the PT-SA square-packing engine.
 *
 * The reference (koboshchan/squares-in-squares) exposes this path only as a
 * process: `packing <seed.txt> <epochs> <steps> <chains> <rng_seed> <out.json>`
 * (cuda/packing.cu:768-868), driven through subprocess by cuda/cpp_loop.py:62-63,
 * cuda/full_run.py:50-51, cuda/scheduler.py:86-87, cuda/batch.py:31,
 * cuda/demo.py:84-86 and cuda/autotune.py:18.  Each entry point below replaces
 * one piece of that process contract; the `packing` binary built from this
 * library keeps the argv/env/JSON contract itself.
 *
 * Conventions: plain pointers and sizes; inputs are caller-owned and only read
 * during the call; outputs go to caller-allocated buffers.  Every function
 * returns an int status: SIS_OK (0), SIS_EINVAL (2, the CLI's exit code for bad
 * input, cuda/packing.cu:769-775), SIS_EDEVICE (3, a HIP error), SIS_PENDING (1,
 * sis_poll only).  sis_last_error() gives the message of the calling thread's
 * last failure.  No C++ exception crosses this boundary.
 */
#ifndef SIS_ABI_H
#define SIS_ABI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SIS_OK 0
#define SIS_PENDING 1
#define SIS_EINVAL 2
#define SIS_EDEVICE 3
/* A served request that reached its deadline (sis_submit_timeout,
 * sis_server_cfg.request_timeout_s; the reference gives each case's `packing`
 * process 900 s, cuda/cpp_loop.py:63, 78).  As sis_result.status: stopped at
 * a round boundary, the result is the best configuration found so far.  As
 * sis_poll's return: the deadline passed before admission, no result. */
#define SIS_ETIMEOUT 4

/* sis_params.flags: count the near pairs (centres closer than sqrt(2), before
 * and after the move) of every proposal into sis_result.near_pairs, the P of
 * the roofline's flop model.  A diagnostic: it costs about 3 % of sweep
 * throughput, and without it near_pairs is 0. */
#define SIS_FLAG_NEAR_PAIRS 1

/* One packing instance: n unit squares in a box of side L.
 * Replaces the seed file `n L` + n lines `x y angle` (cuda/packing.cu:772-775). */
typedef struct sis_instance {
    int n;              /* 1..1000 */
    double L;           /* (0, 64] */
    const double* q;    /* n*3 (x, y, angle in radians) or NULL for uniform random init */
} sis_instance;

/* Search parameters.  Replaces argv epochs/steps/chains/rng_seed and the
 * PACK_* environment knobs (cuda/packing.cu:769-786). */
typedef struct sis_params {
    double tmin, tmax;          /* geometric temperature ladder (PACK_TMIN/PACK_TMAX) */
    double move_sigma;          /* translation step sigma (PACK_MOVE) */
    double rot_sigma;           /* rotation step sigma, radians (PACK_ROTATION) */
    double target_accept;       /* acceptance the step sizes adapt to */
    double shrink;              /* L decrement on a zero-overlap configuration (PACK_SHRINK) */
    double cell;                /* cell width of the colour checkerboard, >= 0.7075 */
    double init_jitter;         /* uniform jitter applied to seed configurations */
    int64_t sweeps;             /* sweeps per replica; one sweep = n proposals */
    int replicas;               /* replicas (chains) per instance, 1..4096 */
    int ladder;                 /* temperatures per ladder; 0 = one ladder over all replicas */
    int sweeps_per_exchange;    /* sweeps between replica-exchange steps */
    int replicas_per_group;     /* replicas per workgroup tile; 0 = autotuned */
    int polish_iters;           /* FP64 conjugate-gradient polish iterations at the end */
    int init_mode;              /* 0: seed configuration (+ init_jitter) in every replica;
                                   1: uniform random; 2: portfolio (sis_portfolio) */
    int flags;                  /* 0 or SIS_FLAG_NEAR_PAIRS */
    uint64_t rng_seed;          /* Philox key */
    int collective_moves;       /* collective proposals per replica after each interval's sweeps
                                   (row / column / cluster / orientation / diagonal / tilted-split
                                   moves of cuda/packing.cu:564-616), 0..64 */
    int kick_every;             /* intervals between kicks (worst-square relocation + jitter,
                                   cuda/packing.cu:696-745); 0 = none */
    int sweeps_per_shift;       /* sweeps between random shifts of the cell grid, 1..100000: a
                                   move never leaves its cell, so the cell lists of a shift stay
                                   exact and the sweeps after the first reuse them */
    int improve;                /* how the box side moves after an exchange interval:
                                   0: when the coldest replica with FP32 energy exactly zero passes
                                      the FP64 zero-overlap test, the box shrinks by `shrink` and
                                      every centre is rescaled with it;
                                   1: the reference's improvement path (cuda/packing.cu:254-297,
                                      503-504, 831-848): every replica's FP64 certified side is
                                      computed after each sweep launch; when the smallest beats the
                                      best side by 1e-11, the best side moves to it, every
                                      temperature slot gets the box best - shrink * mults[class]
                                      (mults 0.02 .. 20 over eight classes by ladder position), the
                                      coldest slot of every ladder (every slot of the first ladder
                                      when the gain exceeds 5e-4) restarts from the best
                                      configuration scaled to its box, and that interval does no
                                      exchange */
    int cool_period;            /* reheating cycle of cuda/packing.cu:516-517,756-758 in exchange
                                   intervals: temperatures times 0.002^((k % P) / (P - 1)) and step
                                   sizes times its square root in interval k; 0 = none (P >= 2) */
    int certify_every;          /* improve = 1: the certified sides are computed at the end of every
                                   k-th sweep launch (interval k - 1, 2k - 1, ...) and of the last one,
                                   1..1e6, default 1; the reference computes them once per epoch
                                   (cuda/packing.cu:506, 689) */
} sis_params;

/* Result of one instance.  Replaces raw{n}.json (cuda/packing.cu:859-865). */
typedef struct sis_result {
    double* q_out;              /* caller-owned n*3 (x, y, angle): translated, dilated export */
    double target_L;            /* box side at which the exported configuration had zero overlap */
    double L_cert;              /* bounding side of q_out (FP64 certified-side test) */
    double lambda;              /* centre dilation applied by the certified-side test */
    double e32;                 /* FP32 energy of the source configuration at the end of the sweep
                                   launch it came from (0 for a seed that was never improved on) */
    double e64_before;          /* FP64 energy of the same configuration before polish */
    double e64_after;           /* FP64 energy after polish */
    double e_gap;               /* |e32 - e64_before|; 0 for a seed that was never improved on */
    uint64_t proposals;         /* attempted single-square moves */
    uint64_t accepted;
    uint64_t sweeps;            /* sweeps per replica completed */
    uint64_t exchanges_tried, exchanges_accepted;
    uint64_t candidates;        /* stencil neighbours examined (C) */
    uint64_t near_pairs;        /* neighbours with d^2 < 2, old + new position (P) */
    int shrinks;                /* number of box shrinks */
    int feasible;               /* 1 if a zero-overlap configuration was found */
    int replicas;               /* replicas actually used */
    int status;                 /* SIS_OK or an error code */
    double seconds;             /* host wall time of the solve */
    double device_seconds;      /* time of the sweep + exchange kernels (HIP events) */
    int fp64_rejects;           /* zero-overlap configurations in FP32 that the FP64 test rejected */
    int kicks;                  /* replica kicks applied */
    uint64_t collective_tried, collective_accepted;
    uint64_t feasible_checkpoints; /* improve = 1: replica checkpoints (end of a sweep launch) whose
                                   configuration certified (dilation <= 1.008), the count the
                                   reference reports as feasible_checkpoints (cuda/packing.cu:506-507,
                                   689-693, 857); improve = 0: zero-overlap configurations archived */
} sis_result;

/* Server configuration for the continuous-batching request loop. */
typedef struct sis_server_cfg {
    double hbm_budget_bytes;    /* admission cap on resident replica state */
    int max_instances;          /* resident instances at once */
    int intervals_per_round;    /* exchange intervals per device round */
    int certifier_threads;      /* host threads that certify finished packings */
    int certify;                /* 1: certify every result (dilation ladder + exact certifier) */
    double request_timeout_s;   /* deadline of every request, seconds after sis_submit; 0: none */
} sis_server_cfg;

typedef struct sis_server sis_server;

/* Library set-up on a device (no reference counterpart: the reference binds
 * device 0 implicitly, cuda/packing.cu:787). */
int sis_init(int device);
const char* sis_last_error(void);
/* Status code of the calling thread's last failure: the code of an entry point
 * that returns a handle (sis_server_create, sis_batch_create) and NULL on
 * failure.  The reference reports failures only as a process exit code
 * (cuda/packing.cu:13,769-775). */
int sis_last_error_code(void);
/* Compute units and marketing name of the device in use (the reference prints
 * prop.multiProcessorCount and prop.name into raw.json, cuda/packing.cu:862-863). */
int sis_device_info(int* compute_units, char* name, int name_len);

/* Sweep workgroups of this instance that one compute unit holds at once, by
 * the HIP occupancy query for the tile shape the engine picks when the
 * instance runs alone (replicas per workgroup, LDS), and those replicas per
 * workgroup.  Replaces cudaOccupancyMaxActiveBlocksPerMultiprocessor for the
 * search kernel (cuda/packing.cu:790-794), whose result the reference reports
 * as blocks_per_sm and uses to size its chains. */
int sis_occupancy(const sis_instance* inst, const sis_params* params, int* blocks_per_cu,
                  int* replicas_per_block);

/* Defaults of cuda/packing.cu:777-786 mapped to this engine. */
void sis_default_params(sis_params* p);

/* Starting configurations of init_mode 2 (cuda/packing.cu:304-452, the portfolio
 * built on the host before the first epoch, cuda/packing.cu:806, and the
 * structured jitter of cuda/packing.cu:454-475): replica r at temperature
 * position r % ladder of ladder r / ladder gets, by that position scaled to
 * eight classes, the seed or a Goebel-strip topology (classes 0-1, the seed
 * on the first two ladders), a Goebel-strip topology (2-4), a wall-hugging grid
 * with tilted rows (5-6) or corner squares with a tilted grid (7); then,
 * when `jitter` is 1, class 0 stays exact, class 1 moves by 1e-4 x 100^u and
 * the rest by 0.03.  out: replicas*n*3 (x, y, angle in radians).  Host only:
 * needs no device. */
int sis_portfolio(const sis_instance* inst, const sis_params* params, int jitter, double* out);

/* Plain form: one instance per launch sequence with a host synchronisation
 * every exchange interval, as the reference's epoch loop
 * (cuda/packing.cu:823-853) driven once per n by cuda/cpp_loop.py:47-80. */
int sis_solve(const sis_instance* inst, const sis_params* params, sis_result* out);

/* The plain form (sis_solve, same results bit for bit) without any device
 * allocation: no counterpart in the reference.  For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Every device buffer comes from one
 * 32 MiB array per device that the library holds until the process ends;
 * requests whose state needs more (about n x replicas > 2 million squares)
 * are rejected with SIS_EINVAL.  Calls on one device share the array: the
 * library serialises them.  Work runs on the null stream with synchronous
 * copies; device_seconds is host time around the launches. */
int sis_solve_noalloc(const sis_instance* inst, const sis_params* params, sis_result* out);

/* Batched form: `count` instances of any n, all replica tiles in one launch per
 * interval, the whole schedule captured into a HIP graph.  Replaces the
 * reference's one-process-per-n pool (cuda/cpp_loop.py:82-106). */
int sis_solve_batch(const sis_instance* insts, int count, const sis_params* params,
                    sis_result* out);

/* Step-wise control of resident instances: the reference's epoch loop
 * (cuda/packing.cu:823-853) opened up so a caller can run, time and finish
 * it in pieces.  mode 0 = plain (the host reads every interval's status,
 * with the next interval already queued), mode 1 = HIP graph (no host
 * synchronisation).  intervals <= 0 runs what is
 * left of the budget; a plain run stops early once every instance has run its
 * sweeps.  sweep_ms is the HIP-event time of the sweep launches that were
 * timed and sweep_launches their number (plain mode: every eighth launch of
 * the call, starting with the first; 0 in graph mode), total_ms the time of
 * the whole run. */
typedef struct sis_batch sis_batch;
sis_batch* sis_batch_create(const sis_instance* insts, int count, const sis_params* params);
int sis_batch_run(sis_batch* b, int64_t intervals, int mode, double* sweep_ms, int64_t* sweep_launches,
                  double* total_ms);
int sis_batch_finalize(sis_batch* b, sis_result* out);
int sis_batch_destroy(sis_batch* b);
/* The status of instance k as the host last read it, without a device copy:
 * every run copies the descriptors to a host mirror once at its end (the
 * reference's per-epoch blocking read of the chains' best sides,
 * cuda/packing.cu:831, reduced to what the host needs between intervals:
 * in mode 0 each exchange step writes its instance's interval count into
 * mapped host memory, which the plain loop waits on). */
struct sis_ladder_desc;  /* declared with the parity probes below */
int sis_batch_status(sis_batch* b, int k, struct sis_ladder_desc* d);

/* Autotuning (no reference counterpart: the reference picks chains from an
 * occupancy query, cuda/packing.cu:788-799, and its autotune.py runs a fixed
 * launch matrix, cuda/autotune.py:8-20, ranked by polished L first and
 * chain-steps/s second, :37; here speed ranks, guarded by a quality
 * probe where the trajectory changes).  Runs `probe_sweeps` sweeps of the
 * given batch for every candidate of replicas per workgroup (0 = the automatic
 * size of each instance by its n and the load, 1, 2, 4, 8), sweeps per
 * exchange (5, 10, 20) and cell width (0.7075, 0.8), by HIP-event time.
 * Replicas per workgroup never change results; the other two change the
 * trajectory, so a faster candidate that changes them is put through a
 * quality probe first: the batch runs 50 x probe_sweeps sweeps (polish off)
 * under the candidate and under the caller's own settings, to the same sweep
 * of every instance, and the candidate is taken when the mean paired loss of
 * relative side reduction 1 - min(L_cert, L) / L plus two standard errors is
 * at most 0.001 (non-inferiority).  At most three such candidates are
 * checked, fastest first; otherwise the fastest candidate with the caller's
 * sweeps per exchange and cell wins (the caller's own values are added as
 * candidates when they are not on the grid).  sweeps_per_exchange and cell
 * report instance 0's values. */
typedef struct sis_tuning {
    int replicas_per_group;
    int sweeps_per_exchange;
    double cell;
    double proposals_per_s;     /* of the winner during its probe */
    int candidates_tried;
    int trajectory_kept;        /* 1: the winner keeps the caller's sweeps per exchange and cell */
    int quality_checked;        /* candidates put through the quality probe */
    int quality_sweeps;         /* sweeps of the quality probe (0: none ran) */
    /* The last candidate put through the quality probe (the winner when
     * trajectory_kept = 0; zeros when none was): its sweeps per exchange and
     * cell, and its mean paired loss of relative side reduction with the
     * standard error of that mean. */
    int checked_sweeps_per_exchange;
    double checked_cell;
    double quality_delta;
    double quality_se;
} sis_tuning;
int sis_autotune(const sis_instance* insts, int count, const sis_params* params, int probe_sweeps,
                 sis_tuning* out);

/* Serving: requests of any n admitted while others run (continuous batching).
 * Replaces cuda/cpp_loop.py:82-106 (ThreadPoolExecutor + VRAM admission). */
sis_server* sis_server_create(const sis_server_cfg* cfg);
int sis_submit(sis_server* s, const sis_instance* inst, const sis_params* params,
               uint64_t* ticket);
/* sis_submit with a deadline of its own, timeout_s seconds after the call
 * (0: none), in place of the configuration's request_timeout_s
 * (cuda/cpp_loop.py:63: `timeout=900` on each case). */
int sis_submit_timeout(sis_server* s, const sis_instance* inst, const sis_params* params, double timeout_s,
                       uint64_t* ticket);
/* SIS_OK when finished (out filled; out->q_out must hold n*3 doubles;
 * out->status SIS_OK, SIS_EINVAL when the certifier rejected it, SIS_ETIMEOUT
 * when it stopped at its deadline), SIS_PENDING while queued or running,
 * SIS_ETIMEOUT when its deadline passed before admission and SIS_EDEVICE when
 * a device round that held it failed (the message in sis_last_error; requests
 * in other lanes are not affected).  The reference records such cases as
 * "rejected" with the exception text (cuda/cpp_loop.py:77-78). */
int sis_poll(sis_server* s, uint64_t ticket, sis_result* out);
/* Seconds from submit to a certified result, and the certified L (50 decimals). */
int sis_ticket_times(sis_server* s, uint64_t ticket, double* queued_s, double* done_s,
                     double* certified_s, char* L50, int L50_len);
/* Forget a finished ticket (its result, packing and times): a long-running
 * server keeps every ticket until released or destroyed.  SIS_PENDING (and the
 * ticket kept) while it is queued or running.  No counterpart in the reference,
 * whose request loop ends with its process pool. */
int sis_ticket_release(sis_server* s, uint64_t ticket);
int sis_server_destroy(sis_server* s);
/* Admission state: the device arena's capacity in bytes (hbm_budget_bytes,
 * at most 90 % of the device memory free at creation), the bytes of replica
 * state resident now and at most so far, and the requests resident now.  The
 * reference's admission reads free VRAM and RAM (cuda/cpp_loop.py:35-45). */
int sis_server_stats(sis_server* s, double* arena_bytes, double* resident_bytes, double* peak_resident_bytes,
                     int* resident_requests);
/* Test hook, no counterpart in the reference: the next round of lane `lane`
 * (0: n < 64, 1: n < 256, 2 and 3: the rest) fails at its first sweep launch,
 * before anything is enqueued, as a failed launch would.  For tests of the
 * lane-failure path, which need no faulting kernel. */
int sis_server_inject_lane_fault(sis_server* s, int lane);

/* Exact certificate of one packing: decimal strings in, 50-decimal L out.
 * Replaces the certifier process boundary (cuda/certifier.cpp:933-1077,
 * called by cuda/cpp_loop.py:17-28); semantics of
 * cuda/certifier.cpp:738-912 (accept contact, reject any interior overlap).
 * Returns SIS_OK and writes L (>= 52 bytes) or SIS_EINVAL with the reason. */
int sis_certify_strings(int n, const char* const* xs, const char* const* ys,
                        const char* const* angles, char* L50, int L50_len);
/* Dilation-ladder export of a raw configuration then certification
 * (cuda/demo.py:25-40): tries centre factors 1.000002, 1.00001, 1.0001.
 * On success writes the accepted factor and the 50-decimal L. */
int sis_export_certify(int n, const double* q, double* dilation, char* L50, int L50_len);

/* ---- device parity entry points (hot-path kernels on caller data) ---- */

/* SAT depth of pairs: in[8k..8k+5] = dx, dy, ci, si, cj, sj (dx = x_j - x_i).
 * The reference predicate is cuda/packing.cu:26-34 (depth), restated with
 * cached cos/sin as in cuda/packing.cu:177-189.  Writes depth and depth^2. */
int sis_dev_depth32(int npairs, const float* in, float* depth, float* depth2);
/* Cell coordinate clamp(floor((x + shift) / w), lo, hi); with w = 2, shift = 0,
 * lo = 0, hi = 15 this is the reference's binning (cuda/packing.cu:22-25). */
int sis_dev_cellcoord(int count, const float* x, float shift, float w, int lo, int hi,
                      int* out);
/* The Metropolis test's 2^x (x <= 0), out[k] = 2^x[k] as the sweep computes it
 * for `u < 2^(-dE/T log2 e)`; replaces the exp of the acceptance in
 * cuda/packing.cu:80,123 (walk, step).  Same bits as the oracle's FP32 form. */
int sis_dev_accept_exp2(int count, const float* x, float* out);
/* Run `sweeps` sweeps of the FP32 sweep kernel on R replicas of one instance,
 * in place: xycs = R*n*4 floats (x, y, cos, sin).  temps/steps per replica;
 * the cell grid gets a new random shift every `sweeps_per_shift` sweeps.
 * Returns per-replica accepted moves and FP32 energies after the sweeps.
 * Semantics: cuda/packing.cu:52-95 (walk). */
int sis_dev_sweeps(int n, int R, float L, float cell, int group, const float* inv_t_log2e,
                   const float* step_xy, const float* step_th, uint32_t key0, uint32_t key1,
                   uint32_t sweep0, int sweeps, int sweeps_per_shift, float* xycs, int* accepted,
                   float* energy, uint64_t counts[3]);
/* One sweep launch as above followed by `collective` collective proposals per
 * replica (cuda/packing.cu:564-616) and, when (interval + 1) is a multiple of
 * kick_every > 0, the kick (cuda/packing.cu:696-745; relocating on odd kick
 * numbers).  counts: proposals, candidates, near pairs, collective proposals,
 * collective accepts, kicked replicas. */
int sis_dev_epoch(int n, int R, float L, float cell, int group, const float* inv_t_log2e,
                  const float* step_xy, const float* step_th, uint32_t key0, uint32_t key1,
                  uint32_t sweep0, int sweeps, int sweeps_per_shift, int collective, int kick_every,
                  uint32_t interval, float* xycs, int* accepted, float* energy, uint64_t counts[6]);
/* The FP64 zero-overlap test that gates archiving and shrinking after each
 * exchange step (cuda/packing.cu:625-632, where every pair is summed): the
 * grid test used by the exchange kernel (zero = 1 when the configuration has
 * no overlap) beside the all-pairs FP64 energy of the same configuration.
 * xycs: n float4 (x, y, cos, sin). */
int sis_dev_zero64(int n, double L, const float* xycs, int* zero, double* e64);
/* FP64 polish + certified side of one configuration (cuda/packing.cu:624-697):
 * q in/out n*3; writes the exported (translated, dilated) configuration: the
 * certified side before and after the polish, whichever is smaller, as the
 * reference keeps the smaller of init_cert and post_cert (:506-507, :689-693). */
int sis_dev_polish(int n, double L, int iters, double* q, double* energy_before,
                   double* energy_after, double* lam, double* L_cert, double* q_export);

/* ---- probes of the exchange step and of a resident batch's state ---- */

/* Per-replica record of the engine (the reference keeps energies[b] and
 * best_Ls[b] per chain, cuda/packing.cu:479, 506). */
typedef struct sis_rep_state {
    float E;                    /* FP32 energy at the end of the last sweep launch */
    int slot;                   /* temperature slot held */
    unsigned acc;               /* accepted moves in the last interval */
    unsigned pad;
    double Lc;                  /* improve = 1: FP64 certified side of the configuration at the end of
                                   the last launch (1e9 when the dilation would exceed 1.008) */
    double lam;                 /* its centre dilation */
    double minx, miny;          /* lower corner of the dilated configuration's bounding box */
    /* Counters of the last sweep launch's workgroup whose first replica
     * this is (zero in the other replicas), moved into the instance's totals
     * and cleared by the exchange step that follows every launch. */
    uint64_t cand;              /* stencil candidates examined */
    uint64_t near_pairs;        /* candidate pairs with centres closer than sqrt(2), before and after */
    uint32_t coll_tried, coll_acc;  /* collective proposals tried / accepted */
    uint32_t kicks;             /* replicas kicked */
    uint32_t pad2;
} sis_rep_state;

/* Per temperature slot.  The sweep reads the first four fields. */
typedef struct sis_slot_state {
    float inv_t_log2e;          /* log2(e) / (T cool) of the current interval */
    float step_xy, step_th;     /* step half-widths of the current interval: base x sqrt(cool) */
    float L;                    /* box side of the slot */
    float T;                    /* ladder temperature */
    float base_xy, base_th;     /* adapted step half-widths */
    float pad;
} sis_slot_state;

/* Instance fields the exchange step reads and writes. */
typedef struct sis_ladder_desc {
    int n, R, ladder, improve, cool_period, spe;
    int64_t sweeps_total, sweeps_done;
    uint32_t key0, key1, interval;
    float shrink, target_acc, scale, Lf;
    double L, L_arch, L_box_arch;
    int arch_valid, shrinks, feasible, fp64_rejects;
    uint64_t ex_tried, ex_acc, certs;
} sis_ladder_desc;

/* One exchange step (the kernel run after every sweep launch) on caller data:
 * step adaptation, replica exchange (cuda/packing.cu:141-152, 750-765) or the
 * improvement path (improve = 1, :831-848), the FP64 zero-overlap shrink
 * (improve = 0), the cooling cycle.  reps, slots, rep_of_slot: R each;
 * xycs: R*n*4 floats (x, y, cos, sin), arch: n*4; all updated in place. */
int sis_dev_ladder(sis_ladder_desc* d, sis_rep_state* reps, sis_slot_state* slots, int* rep_of_slot,
                   float* xycs, float* arch);
/* The device state of instance k of a resident batch (sis_batch_create):
 * descriptor, replica and slot records, slot owners, replica states (R*n*4
 * floats) and archive (n*4).  Any output may be NULL. */
int sis_dev_batch_state(sis_batch* b, int k, sis_ladder_desc* d, sis_rep_state* reps, sis_slot_state* slots,
                        int* rep_of_slot, float* xycs, float* arch);
/* Sweep launch of R replicas of one instance as sis_dev_epoch, with a box per
 * replica (box[r] <= Lgrid, the side the cell grid is sized for) and, when
 * certify is 1, every replica's FP64 certified side at the end (Lc, lam,
 * minx, miny of sis_rep_state; pass reps, R records, or NULL). */
int sis_dev_launch(int n, int R, float Lgrid, const float* box, float cell, int group, const float* inv_t_log2e,
                   const float* step_xy, const float* step_th, uint32_t key0, uint32_t key1, uint32_t sweep0,
                   int sweeps, int sweeps_per_shift, int collective, int kick_every, uint32_t interval,
                   int certify, float* xycs, int* accepted, float* energy, uint64_t counts[6],
                   sis_rep_state* reps);

#ifdef __cplusplus
}
#endif
#endif
