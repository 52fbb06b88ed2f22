// Host runtime: resident instances in HBM, replica tiles, the plain and the
// batched (HIP graph) launch schedules, FP64 finalisation, the C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sis_abi.h"
#include "sis_host.h"
#include "sis_internal.h"
#include "sis_kernels.h"

namespace sis {

thread_local std::string t_err;
thread_local int t_code = 0;

int fail(int code, const std::string& msg) {
    t_err = msg;
    t_code = code;
    return code;
}

#define HIPCHECK(x)                                                                        \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            return fail(SIS_EDEVICE, std::string(#x) + ": " + hipGetErrorString(e_));     \
    } while (0)

// ---------------------------------------------------------------------------
// Device set-up.
// ---------------------------------------------------------------------------
static std::mutex g_init_mu;
static DeviceInfo g_dev;

int ensure_device(int device) {
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_dev.id >= 0 && (device < 0 || device == g_dev.id)) {
        HIPCHECK(hipSetDevice(g_dev.id));
        return SIS_OK;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1)
        return fail(SIS_EDEVICE, "no HIP device available");
    const int id = device < 0 ? 0 : device;
    if (id >= count) return fail(SIS_EINVAL, "device index out of range");
    HIPCHECK(hipSetDevice(id));
    hipDeviceProp_t prop;
    HIPCHECK(hipGetDeviceProperties(&prop, id));
    // a sweep tile may use all the LDS a workgroup can hold (160 KB on gfx950)
    for (int threads : {kSweepThreads, kSweepThreadsBig})
        for (bool near_pairs : {false, true})
            for (bool alone : {false, true})
                HIPCHECK(hipFuncSetAttribute(sweep_kernel_ptr(threads, near_pairs, alone),
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)prop.sharedMemPerBlock));
    // the allocation-free solve's device buffer: its address on this device
    if (int rc = noalloc_setup(id)) return rc;
    g_dev.cus = prop.multiProcessorCount;
    g_dev.lds_per_block = (size_t)prop.sharedMemPerBlock;
    g_dev.name = prop.name[0] ? std::string(prop.name) : std::string("AMD ") + prop.gcnArchName;
    g_dev.arch = prop.gcnArchName;
    g_dev.id = id;  // last: a failed set-up is retried on the next call
    return SIS_OK;
}

const DeviceInfo& device_info() { return g_dev; }

// ---------------------------------------------------------------------------
// Parameters.
// ---------------------------------------------------------------------------
void default_params(sis_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->tmin = 1e-5;
    p->tmax = 5e-2;
    p->move_sigma = 0.05;
    p->rot_sigma = 0.05;
    p->target_accept = 0.3;
    p->shrink = 1e-4;
    p->cell = 0.7075;
    p->init_jitter = 0.0;
    p->sweeps = 1000;
    p->replicas = 64;
    p->ladder = 0;
    p->sweeps_per_exchange = 10;
    p->sweeps_per_shift = 4;
    p->replicas_per_group = 0;
    p->polish_iters = 200;
    p->init_mode = 0;
    p->flags = 0;
    p->rng_seed = 7;
    p->improve = 0;
    p->cool_period = 0;
    p->certify_every = 1;
}

int validate(const sis_instance* in, const sis_params* p) {
    if (!in || !p) return fail(SIS_EINVAL, "null instance or params");
    if (in->n < 1 || in->n > kMaxN) return fail(SIS_EINVAL, "n must be 1..1000");
    if (!std::isfinite(in->L) || in->L <= 0 || in->L > 64) return fail(SIS_EINVAL, "L must be in (0, 64]");
    if (in->q)
        for (int k = 0; k < 3 * in->n; ++k)
            if (!std::isfinite(in->q[k])) return fail(SIS_EINVAL, "non-finite seed coordinate");
    if (!(p->tmin > 0) || !(p->tmax >= p->tmin) || !std::isfinite(p->tmax))
        return fail(SIS_EINVAL, "need 0 < tmin <= tmax");
    if (!(p->move_sigma > 0) || !(p->rot_sigma > 0)) return fail(SIS_EINVAL, "step sigmas must be > 0");
    if (!(p->target_accept > 0 && p->target_accept < 1)) return fail(SIS_EINVAL, "target_accept in (0,1)");
    if (!(p->shrink >= 0 && p->shrink <= 0.1)) return fail(SIS_EINVAL, "shrink in [0, 0.1]");
    if (!((float)p->cell >= kMinCell && p->cell <= 4.0)) return fail(SIS_EINVAL, "cell width in [0.7075, 4]");
    if (p->replicas < 1 || p->replicas > 4096) return fail(SIS_EINVAL, "replicas must be 1..4096");
    if (p->ladder < 0 || (p->ladder > 0 && p->replicas % p->ladder != 0))
        return fail(SIS_EINVAL, "ladder must divide replicas");
    if (p->sweeps < 0 || p->sweeps > (1ll << 31)) return fail(SIS_EINVAL, "sweeps out of range");
    if (p->sweeps_per_exchange < 1 || p->sweeps_per_exchange > 100000)
        return fail(SIS_EINVAL, "sweeps_per_exchange out of range");
    if (p->sweeps_per_shift < 1 || p->sweeps_per_shift > 100000)
        return fail(SIS_EINVAL, "sweeps_per_shift out of range");
    if (p->replicas_per_group < 0 || p->replicas_per_group > 256)
        return fail(SIS_EINVAL, "replicas_per_group out of range");
    if (p->polish_iters < 0 || p->polish_iters > 100000) return fail(SIS_EINVAL, "polish_iters out of range");
    if (!(p->init_jitter >= 0 && p->init_jitter < 1)) return fail(SIS_EINVAL, "init_jitter in [0,1)");
    if (p->init_mode < 0 || p->init_mode > 2) return fail(SIS_EINVAL, "init_mode must be 0, 1 or 2");
    if (p->flags & ~SIS_FLAG_NEAR_PAIRS) return fail(SIS_EINVAL, "unknown flags");
    if (p->collective_moves < 0 || p->collective_moves > 64) return fail(SIS_EINVAL, "collective_moves in 0..64");
    if (p->kick_every < 0 || p->kick_every > 1000000) return fail(SIS_EINVAL, "kick_every in 0..1e6");
    if (p->improve < 0 || p->improve > 1) return fail(SIS_EINVAL, "improve must be 0 or 1");
    if (p->cool_period < 0 || p->cool_period == 1 || p->cool_period > 1000000)
        return fail(SIS_EINVAL, "cool_period must be 0 or 2..1e6");
    if (p->certify_every < 1 || p->certify_every > 1000000) return fail(SIS_EINVAL, "certify_every must be 1..1e6");
    const float w = (float)p->cell;
    const int M = grid_cells((float)in->L, w) + 4;
    if (M > 255) return fail(SIS_EINVAL, "cell grid wider than 255 cells: raise the cell width");
    return SIS_OK;
}

// ---------------------------------------------------------------------------
// Host-side initial states (Philox purpose 3; same generator as the kernels).
// ---------------------------------------------------------------------------
static inline void philox_host(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                               uint32_t k1, uint32_t out[4]) {
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
static inline double u01d(uint32_t hi, uint32_t lo) {
    const uint64_t k = (((uint64_t)hi << 21) ^ (uint64_t)(lo >> 11)) & ((1ull << 52) - 1);
    return ((double)k + 0.5) * (1.0 / 4503599627370496.0);
}

// Portfolio starts (init_mode 2): the structured families of every replica by
// its ladder position, then, when `jitter`, the per-class jitter of
// cuda/packing.cu:454-475 drawn from Philox (amplitude: counter (0, r, 2, 3);
// square i: (i, r, 0, 3) for x and y, (i, r, 1, 3) for the angle), centres
// clamped to [1/2, L - 1/2].  Without a seed configuration the seed is a
// uniform random one (replica 0's draws of init_mode 1).
void portfolio_state(const InstSpec& s, bool jitter, std::vector<double>& q) {
    const int n = s.n, R = s.p.replicas, lad = s.p.ladder > 0 ? s.p.ladder : R;
    const uint32_t k0 = (uint32_t)s.p.rng_seed, k1 = (uint32_t)(s.p.rng_seed >> 32);
    const double L = s.L;
    std::vector<double> seed(s.q);
    if (seed.empty()) {
        seed.resize(3 * (size_t)n);
        const double lo = std::min(0.5, 0.5 * L), span = std::max(0.0, L - 2 * lo);
        for (int i = 0; i < n; ++i) {
            uint32_t o[4], o2[4];
            philox_host((uint32_t)i, 0u, 0u, 3u, k0, k1, o);
            philox_host((uint32_t)i, 0u, 1u, 3u, k0, k1, o2);
            seed[3 * i] = lo + span * u01d(o[0], o[1]);
            seed[3 * i + 1] = lo + span * u01d(o[2], o[3]);
            seed[3 * i + 2] = 1.5707963267948966 * u01d(o2[0], o2[1]);
        }
    }
    portfolio_layouts(n, R, lad, L, seed.data(), q);
    if (!jitter) return;
    const double lo = std::min(0.5, 0.5 * L), hi = std::max(L - 0.5, 0.5 * L);
    for (int r = 0; r < R; ++r) {
        uint32_t o[4], o2[4];
        philox_host(0u, (uint32_t)r, 2u, 3u, k0, k1, o);
        const double j = portfolio_jitter((r % lad) * 8 / lad, u01d(o[0], o[1]));
        if (j == 0.0) continue;
        for (int i = 0; i < n; ++i) {
            philox_host((uint32_t)i, (uint32_t)r, 0u, 3u, k0, k1, o);
            philox_host((uint32_t)i, (uint32_t)r, 1u, 3u, k0, k1, o2);
            double* p = &q[((size_t)r * n + i) * 3];
            p[0] = std::min(hi, std::max(lo, p[0] + j * (2 * u01d(o[0], o[1]) - 1)));
            p[1] = std::min(hi, std::max(lo, p[1] + j * (2 * u01d(o[2], o[3]) - 1)));
            p[2] += j * (2 * u01d(o2[0], o2[1]) - 1);
        }
    }
}

void initial_state(const InstSpec& s, std::vector<float4>& out) {
    const int n = s.n, R = s.p.replicas;
    const uint32_t k0 = (uint32_t)s.p.rng_seed, k1 = (uint32_t)(s.p.rng_seed >> 32);
    out.resize((size_t)n * R);
    const double L = s.L;
    if (s.p.init_mode == 2) {
        std::vector<double> q;
        portfolio_state(s, true, q);
        for (size_t k = 0; k < (size_t)n * R; ++k) {
            const double x = std::min(std::max(q[3 * k], 0.0), L), y = std::min(std::max(q[3 * k + 1], 0.0), L);
            out[k] = make_float4((float)x, (float)y, (float)std::cos(q[3 * k + 2]), (float)std::sin(q[3 * k + 2]));
        }
        return;
    }
    for (int r = 0; r < R; ++r)
        for (int i = 0; i < n; ++i) {
            double x, y, a;
            uint32_t o[4], o2[4];
            philox_host((uint32_t)i, (uint32_t)r, 0u, 3u, k0, k1, o);
            philox_host((uint32_t)i, (uint32_t)r, 1u, 3u, k0, k1, o2);
            if (!s.q.empty() && s.p.init_mode == 0) {
                const double j = s.p.init_jitter;
                x = s.q[3 * i] + j * (2 * u01d(o[0], o[1]) - 1);
                y = s.q[3 * i + 1] + j * (2 * u01d(o[2], o[3]) - 1);
                a = s.q[3 * i + 2] + j * (2 * u01d(o2[0], o2[1]) - 1);
            } else {
                const double lo = std::min(0.5, 0.5 * L), span = std::max(0.0, L - 2 * lo);
                x = lo + span * u01d(o[0], o[1]);
                y = lo + span * u01d(o[2], o[3]);
                a = 1.5707963267948966 * u01d(o2[0], o2[1]);
            }
            x = std::min(std::max(x, 0.0), L);
            y = std::min(std::max(y, 0.0), L);
            out[(size_t)r * n + i] = make_float4((float)x, (float)y, (float)std::cos(a), (float)std::sin(a));
        }
}

// Box side of temperature slot s (position s % ladder) with best side Lb:
// the reference's per-chain target max(sqrt n + 1e-6, best - shrink *
// mults[b % 8]) (cuda/packing.cu:503-504, 841), the eight classes spread over
// a ladder of any length.  improve = 0: every slot has the instance's box.
double slot_box(const InstSpec& s, int ladder, int slot, double Lb) {
    if (s.p.improve == 0) return Lb;
    return std::max(std::sqrt((double)s.n) + 1e-6, Lb - s.p.shrink * kTargetMults[((slot % ladder) * 8) / ladder]);
}

void instance_records(const InstSpec& s, const Plan& pl, InstDesc& D, std::vector<RepState>& reps,
                      std::vector<SlotState>& slots, std::vector<int>& ros) {
    const int n = s.n, R = s.p.replicas;
    D.n = n; D.R = R; D.ladder = pl.ladder; D.G = pl.G;
    D.M = pl.M; D.spe = pl.spe; D.spsh = s.p.sweeps_per_shift;
    D.sweeps_total = s.p.sweeps; D.sweeps_done = 0;
    D.L = s.L; D.L_arch = s.L; D.L_box_arch = s.L; D.Lf = (float)s.L; D.scale = 1.0f;
    D.Lmin = std::sqrt((double)n);
    D.w = (float)s.p.cell; D.shrink = (float)s.p.shrink; D.target_acc = (float)s.p.target_accept;
    D.key0 = (uint32_t)s.p.rng_seed; D.key1 = (uint32_t)(s.p.rng_seed >> 32);
    D.collective = s.p.collective_moves; D.kick_every = s.p.kick_every;
    D.improve = s.p.improve; D.cool_period = s.p.cool_period; D.certify_every = s.p.certify_every;
    // replica 0 starts from the seed exactly (init_mode 0 without jitter, or
    // the portfolio's unjittered class 0), and the archive starts as replica 0
    D.seeded = !s.q.empty() && ((s.p.init_mode == 0 && s.p.init_jitter == 0.0) || s.p.init_mode == 2);
    reps.resize(R);
    slots.resize(R);
    ros.resize(R);
    const int lad = pl.ladder;
    double cool, root;
    cool_factor(0u, s.p.cool_period, cool, root);
    for (int k = 0; k < R; ++k) {
        const int pos = k % lad;
        const double T = lad == 1 ? s.p.tmin : s.p.tmin * std::pow(s.p.tmax / s.p.tmin, (double)pos / (lad - 1));
        SlotState& sl = slots[k];
        sl = SlotState{};
        sl.T = (float)T;
        sl.base_xy = (float)(std::sqrt(3.0) * s.p.move_sigma);
        sl.base_th = (float)std::min(0.7794229, std::sqrt(3.0) * s.p.rot_sigma);
        sl.L = (float)slot_box(s, lad, k, s.L);
        slot_effective(sl, cool, root);
        reps[k] = RepState{0.0f, k, 0u, 0u, kCertNone, 1.0, 0.0, 0.0};
        ros[k] = k;
    }
}

// The FP64 finalisation of an instance: the archive when there is one
// (polished in the box it was found in); under the improvement path without
// an improvement, the seed itself, which the archive holds from the start
// (the reference keeps the seed and its side then, cuda/packing.cu:820,
// 859-865); else the best replica by FP32 energy in the instance's box.
// Returns whether the archive is the source.
bool finalize_source(const InstDesc& D, double& box, double& target_L) {
    if (D.arch_valid) {
        box = D.improve ? D.L_box_arch : D.L_arch;
        target_L = D.L_arch;
        return true;
    }
    box = D.L;
    target_L = D.L;
    return D.improve && D.seeded;
}

// ---------------------------------------------------------------------------
// Tile planning.
// ---------------------------------------------------------------------------
bool device_loaded(long long replicas) { return replicas >= 8ll * std::max(1, device_info().cus); }

int plan_instance(const InstSpec& s, int batch_replicas, Plan& pl) {
    const DeviceInfo& dev = device_info();
    const float w = (float)s.p.cell;
    pl.M = grid_cells((float)s.L, w) + 4;
    pl.ladder = s.p.ladder > 0 ? s.p.ladder : s.p.replicas;
    pl.spe = s.p.sweeps_per_exchange;
    const size_t hard = dev.lds_per_block ? dev.lds_per_block : 65536;
    int G = s.p.replicas_per_group;
    const bool big = s.n >= kBigN;
    // Automatic tiles aim at 20 KB of LDS per wavefront-sized tile (eight
    // tiles share a 160 KB compute unit) and 40 KB per four-wavefront tile;
    // an explicit group size may use all the LDS of a workgroup.
    const size_t budget = G > 0 ? hard : std::min<size_t>(hard, big ? 40 * 1024 : 20 * 1024);
    if (G <= 0) {
        // With few replicas every replica gets a workgroup of its own, spread
        // over the compute units.  Under load (several tiles per compute unit)
        // the tile that keeps the most lanes busy per pass wins: measured at
        // 2048 replicas of one instance (1e9 proposals/s), one-wavefront tiles
        // of about 64 squares (n=10: 1 replica 1.04, 2: 1.20, 4: 1.30, 8: 1.21;
        // n=30: 2.71, 3.07, 2.65; n=64: 5.83, 5.54) and four-wavefront tiles
        // of 400-1000 squares (replicas 1, 2, 3, 4: n=100: 4.36, 6.70, 7.20,
        // 7.53; n=150: 6.00, 8.93, 9.99, 8.61; n=200: 7.69, 11.53, 9.77, 9.89;
        // n=300: 9.90, 12.38, 12.09, 11.26; n=500: 14.86, 15.85, 13.63, 14.80;
        // n=700: 14.71, 13.29; n=1000: 17.23, 15.43).
        G = 1;
        pl.wide = !device_loaded(batch_replicas);
        if (!pl.wide) {
            if (big)
                G = s.n < 600 ? std::min(4, std::max(2, 512 / s.n)) : 1;
            else
                while (2 * G * s.n <= 64) G *= 2;
        }
    }
    G = std::min(G, s.p.replicas);
    while (G > 1 && (sweep_lds_bytes(s.n, pl.M, G, pl.wide) > budget || G * s.n > 65535)) --G;
    if (sweep_lds_bytes(s.n, pl.M, G, pl.wide) > hard)
        return fail(SIS_EINVAL, "one replica does not fit the LDS of a workgroup");
    pl.G = G;
    pl.lds = sweep_lds_bytes(s.n, pl.M, G, pl.wide);
    pl.ntiles = (s.p.replicas + G - 1) / G;
    return SIS_OK;
}

// ---------------------------------------------------------------------------
// Batch: a set of instances resident in device memory.
// ---------------------------------------------------------------------------
Batch::~Batch() { release(); }

void Batch::release() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
    for (void* p : {(void*)d_inst, (void*)d_tiles, (void*)d_reps, (void*)d_slots, (void*)d_ros,
                    (void*)d_state, (void*)d_arch, (void*)d_jobs, (void*)d_fout, (void*)d_qout})
        if (p) (void)hipFree(p);
    if (h_pin) (void)hipHostFree(h_pin);
    if (h_seq) (void)hipHostFree(h_seq);
    h_pin = nullptr;
    h_seq = nullptr;
    d_seq = nullptr;
    for (Chain& c : chains) {
        if (c.done) (void)hipEventDestroy(c.done);
        if (c.st) (void)hipStreamDestroy(c.st);
    }
    chains.clear();
    for (void* p : {(void*)d_ctiles, (void*)d_clist})
        if (p) (void)hipFree(p);
    d_ctiles = nullptr;
    d_clist = nullptr;
    if (ev_cfork) (void)hipEventDestroy(ev_cfork);
    ev_cfork = nullptr;
    d_inst = nullptr; d_tiles = nullptr; d_reps = nullptr; d_slots = nullptr; d_ros = nullptr;
    d_state = nullptr; d_arch = nullptr; d_jobs = nullptr; d_fout = nullptr; d_qout = nullptr;
    for (hipEvent_t e : {ev0, ev1, evs0, evs1, ev_fork, ev_join})
        if (e) (void)hipEventDestroy(e);
    ev0 = ev1 = evs0 = evs1 = ev_fork = ev_join = nullptr;
    if (stream) (void)hipStreamDestroy(stream);
    if (stream2) (void)hipStreamDestroy(stream2);
    stream = stream2 = nullptr;
}

int Batch::build(std::vector<InstSpec>&& in) {
    specs = std::move(in);
    const int ni = (int)specs.size();
    if (ni < 1) return fail(SIS_EINVAL, "empty batch");
    long long total_reps = 0;
    for (auto& s : specs) total_reps += s.p.replicas;
    plans.resize(ni);
    h_inst.assign(ni, InstDesc{});
    h_tiles.clear();
    size_t nsq = 0, nrep = 0, narch = 0;
    lds_big = lds_small = lds_ladder = 0;
    max_intervals = 0;
    near_pairs = false;
    for (int k = 0; k < ni; ++k) {
        const InstSpec& s = specs[k];
        near_pairs = near_pairs || (s.p.flags & SIS_FLAG_NEAR_PAIRS) != 0;
        Plan& pl = plans[k];
        int rc = plan_instance(s, (int)std::min<long long>(total_reps, 1 << 30), pl);
        if (rc) return rc;
        InstDesc& D = h_inst[k];
        D.state_off = (int)nsq; D.rep_off = (int)nrep; D.arch_off = (int)narch;
        nsq += (size_t)s.n * s.p.replicas;
        nrep += s.p.replicas;
        narch += s.n;
        size_t& lds = s.n >= kBigN ? lds_big : lds_small;
        lds = std::max(lds, pl.lds);
        lds_ladder = std::max(lds_ladder, ladder_lds_bytes(s.n, s.L, s.p.replicas));
        const long long iv = (s.p.sweeps + pl.spe - 1) / pl.spe;
        max_intervals = std::max(max_intervals, iv);
        for (int r0 = 0; r0 < s.p.replicas; r0 += pl.G)
            h_tiles.push_back(Tile{k, r0, std::min(pl.G, s.p.replicas - r0), pl.wide});
    }
    if (nsq > (size_t)INT32_MAX) return fail(SIS_EINVAL, "batch too large");
    n_big = order_tiles(h_tiles, [&](int k) { return specs[k].n; });
    alone = true;
    for (const Plan& pl : plans) alone = alone && pl.wide;
    n_sq = nsq; n_rep = nrep; n_arch = narch;

    // host initial state, slots, ladders
    std::vector<float4> h_state(nsq);
    std::vector<RepState> h_reps(nrep);
    std::vector<SlotState> h_slots(nrep);
    std::vector<int> h_ros(nrep);
    for (int k = 0; k < ni; ++k) {
        const InstSpec& s = specs[k];
        InstDesc& D = h_inst[k];
        std::vector<float4> st;
        initial_state(s, st);
        std::copy(st.begin(), st.end(), h_state.begin() + D.state_off);
        std::vector<RepState> reps;
        std::vector<SlotState> slots;
        std::vector<int> ros;
        instance_records(s, plans[k], D, reps, slots, ros);
        std::copy(reps.begin(), reps.end(), h_reps.begin() + D.rep_off);
        std::copy(slots.begin(), slots.end(), h_slots.begin() + D.rep_off);
        std::copy(ros.begin(), ros.end(), h_ros.begin() + D.rep_off);
    }

    HIPCHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIPCHECK(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking));
    // pinned host mirror of the descriptors (copied at the start and the end
    // of every run), and mapped, coherent sequence words: in the plain
    // schedule every exchange step writes its instance's interval count,
    // which the host polls for every interval
    HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&h_pin), sizeof(InstDesc) * ni, hipHostMallocDefault));
    HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&h_seq), sizeof(unsigned) * ni,
                           hipHostMallocMapped | hipHostMallocCoherent));
    HIPCHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&d_seq), h_seq, 0));
    std::copy(h_inst.begin(), h_inst.end(), h_pin);  // the status before any run
    HIPCHECK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    HIPCHECK(hipEventCreate(&ev0));
    HIPCHECK(hipEventCreate(&ev1));
    HIPCHECK(hipEventCreate(&evs0));
    HIPCHECK(hipEventCreate(&evs1));
    HIPCHECK(hipMalloc(&d_inst, sizeof(InstDesc) * ni));
    HIPCHECK(hipMalloc(&d_tiles, sizeof(Tile) * h_tiles.size()));
    HIPCHECK(hipMalloc(&d_reps, sizeof(RepState) * nrep));
    HIPCHECK(hipMalloc(&d_slots, sizeof(SlotState) * nrep));
    HIPCHECK(hipMalloc(&d_ros, sizeof(int) * nrep));
    HIPCHECK(hipMalloc(&d_state, sizeof(float4) * nsq));
    HIPCHECK(hipMalloc(&d_arch, sizeof(float4) * narch));
    HIPCHECK(hipMemcpyAsync(d_inst, h_inst.data(), sizeof(InstDesc) * ni, hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemcpyAsync(d_tiles, h_tiles.data(), sizeof(Tile) * h_tiles.size(), hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemcpyAsync(d_reps, h_reps.data(), sizeof(RepState) * nrep, hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemcpyAsync(d_slots, h_slots.data(), sizeof(SlotState) * nrep, hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemcpyAsync(d_ros, h_ros.data(), sizeof(int) * nrep, hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemcpyAsync(d_state, h_state.data(), sizeof(float4) * nsq, hipMemcpyHostToDevice, stream));
    std::vector<float4> h_arch(narch);
    for (int k = 0; k < ni; ++k)
        std::copy(h_state.begin() + h_inst[k].state_off,
                  h_state.begin() + h_inst[k].state_off + specs[k].n, h_arch.begin() + h_inst[k].arch_off);
    HIPCHECK(hipMemcpyAsync(d_arch, h_arch.data(), sizeof(float4) * narch, hipMemcpyHostToDevice, stream));
    if (ni >= kGraphChains) {
        // independent chains, instances assigned heaviest first to the lightest
        std::vector<int> order(ni);
        for (int k = 0; k < ni; ++k) order[k] = k;
        auto work = [&](int k) { return (double)specs[k].n * specs[k].p.replicas * (double)specs[k].p.sweeps; };
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work(a) > work(b); });
        std::vector<int> members[kGraphChains];
        double load[kGraphChains] = {};
        for (int k : order) {
            const int c = (int)(std::min_element(load, load + kGraphChains) - load);
            members[c].push_back(k);
            load[c] += work(k);
        }
        std::vector<Tile> ctiles;
        std::vector<int> clist;
        chains.resize(kGraphChains);
        for (int c = 0; c < kGraphChains; ++c) {
            Chain& ch = chains[c];
            std::vector<Tile> t;
            for (int k : members[c]) {
                const Plan& pl = plans[k];
                for (int r0 = 0; r0 < specs[k].p.replicas; r0 += pl.G)
                    t.push_back(Tile{k, r0, std::min(pl.G, specs[k].p.replicas - r0), pl.wide});
                size_t& lds = specs[k].n >= kBigN ? ch.lds_big : ch.lds_small;
                lds = std::max(lds, pl.lds);
                ch.lds_ladder = std::max(ch.lds_ladder, ladder_lds_bytes(specs[k].n, specs[k].L, specs[k].p.replicas));
            }
            ch.n_big = order_tiles(t, [&](int k) { return specs[k].n; });
            ch.tile_off = (int)ctiles.size();
            ch.ntiles = (int)t.size();
            ctiles.insert(ctiles.end(), t.begin(), t.end());
            ch.list_off = (int)clist.size();
            ch.ninst = (int)members[c].size();
            clist.insert(clist.end(), members[c].begin(), members[c].end());
            HIPCHECK(hipStreamCreateWithFlags(&ch.st, hipStreamNonBlocking));
            HIPCHECK(hipEventCreateWithFlags(&ch.done, hipEventDisableTiming));
        }
        HIPCHECK(hipEventCreateWithFlags(&ev_cfork, hipEventDisableTiming));
        HIPCHECK(hipMalloc(&d_ctiles, sizeof(Tile) * ctiles.size()));
        HIPCHECK(hipMalloc(&d_clist, sizeof(int) * clist.size()));
        HIPCHECK(hipMemcpyAsync(d_ctiles, ctiles.data(), sizeof(Tile) * ctiles.size(), hipMemcpyHostToDevice,
                                stream));
        HIPCHECK(hipMemcpyAsync(d_clist, clist.data(), sizeof(int) * clist.size(), hipMemcpyHostToDevice, stream));
        HIPCHECK(hipStreamSynchronize(stream));  // the host vectors go out of scope
    }
    HIPCHECK(hipStreamSynchronize(stream));
    return SIS_OK;
}

// One interval of one chain on the chain's stream: its sweep launches
// (four-wavefront tiles, then one-wavefront tiles), then the exchange step of
// its instances.  (Forking the two shapes of a chain as well measured slower:
// four streams, more than the device's hardware queues hold apart.)
int Batch::enqueue_chain(int ci) {
    Chain& c = chains[ci];
    const Tile* tiles = d_ctiles + c.tile_off;
    if (c.n_big > 0)
        HIPCHECK(launch_sweep(c.st, c.n_big, c.lds_big, tiles, d_inst, d_reps, d_slots, d_state, kSweepThreadsBig,
                              near_pairs, alone));
    if (c.ntiles > c.n_big)
        HIPCHECK(launch_sweep(c.st, c.ntiles - c.n_big, c.lds_small, tiles + c.n_big, d_inst, d_reps, d_slots,
                              d_state, kSweepThreads, near_pairs, alone));
    HIPCHECK(launch_ladder(c.st, c.ninst, c.lds_ladder, d_inst, d_reps, d_slots, d_ros, d_state, d_arch,
                           d_clist + c.list_off));
    return SIS_OK;
}

// The sweep launches of one interval: four-wavefront tiles on the batch
// stream and, when there are both kinds, one-wavefront tiles on a second
// stream forked from it and joined back, so each launch fills the compute
// units the other's last workgroups leave idle (the two touch disjoint
// replicas).  Captured into the batch's graph as a fork and a join.
int Batch::launch_tiles() {
    const int nt = (int)h_tiles.size();
    const bool both = n_big > 0 && nt > n_big;
    if (both) {
        HIPCHECK(hipEventRecord(ev_fork, stream));
        HIPCHECK(hipStreamWaitEvent(stream2, ev_fork, 0));
    }
    if (n_big > 0)
        HIPCHECK(launch_sweep(stream, n_big, lds_big, d_tiles, d_inst, d_reps, d_slots, d_state,
                              kSweepThreadsBig, near_pairs, alone));
    if (nt > n_big)
        HIPCHECK(launch_sweep(both ? stream2 : stream, nt - n_big, lds_small, d_tiles + n_big, d_inst, d_reps,
                              d_slots, d_state, kSweepThreads, near_pairs, alone));
    if (both) {
        HIPCHECK(hipEventRecord(ev_join, stream2));
        HIPCHECK(hipStreamWaitEvent(stream, ev_join, 0));
    }
    return SIS_OK;
}

int Batch::enqueue_interval() {
    if (int rc = launch_tiles()) return rc;
    HIPCHECK(launch_ladder(stream, (int)specs.size(), lds_ladder, d_inst, d_reps, d_slots, d_ros, d_state, d_arch));
    return SIS_OK;
}

// Plain schedule: one interval per launch pair, and the host waits for every
// interval's end (the reference's per-epoch blocking read,
// cuda/packing.cu:831).  The exchange step writes each instance's interval
// count into mapped host memory, which the host polls; the descriptors reach
// the host mirror by a copy at the start and the end of the run.  Every
// decision of the epoch loop is made on the device and the number of sweeps
// each interval runs is known in advance, so the host keeps one interval
// queued ahead of the one whose end it waits for: the device never idles on
// the host's round trip, and the host still sees every interval's end, one
// interval later.  The sweep kernel of every
// kSweepTimingEvery-th interval is timed with its own event pair on the
// launch stream.
int Batch::run_plain(long long intervals) {
    const int ni = (int)specs.size();
    HIPCHECK(hipMemcpyAsync(h_pin, d_inst, sizeof(InstDesc) * ni, hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    std::vector<unsigned> want(ni), iv_of(ni), pend(ni);
    std::vector<long long> done(ni), total(ni), spe(ni);
    for (int k = 0; k < ni; ++k) {
        iv_of[k] = h_pin[k].interval;
        done[k] = h_pin[k].sweeps_done;
        total[k] = h_pin[k].sweeps_total;
        spe[k] = h_pin[k].spe;
        __atomic_store_n(h_seq + k, iv_of[k], __ATOMIC_RELEASE);
    }
    bool have_pend = false, pend_timed = false;
    // the status of the interval launched before the last one, and its timing
    auto settle = [&]() -> int {
        if (int rc = wait_status(pend)) return rc;
        if (pend_timed) {
            HIPCHECK(hipEventSynchronize(evs1));
            float sms = 0;
            HIPCHECK(hipEventElapsedTime(&sms, evs0, evs1));
            sweep_ms += sms;
            sweep_launches += 1;
        }
        have_pend = false;
        return SIS_OK;
    };
    HIPCHECK(hipEventRecord(ev0, stream));
    for (long long iv = 0; iv < intervals; ++iv) {
        // stop once every instance has run its sweeps
        bool live = false;
        for (int k = 0; k < ni; ++k) {
            const bool l = done[k] < total[k];
            want[k] = l ? iv_of[k] + 1u : 0u;
            live = live || l;
        }
        if (!live) break;
        // one event pair: the previous timed interval has been read before
        // the pair is recorded again (kSweepTimingEvery >= 2)
        const bool timed = iv % kSweepTimingEvery == 0;
        if (timed) HIPCHECK(hipEventRecord(evs0, stream));
        if (int rc = launch_tiles()) return rc;
        if (timed) HIPCHECK(hipEventRecord(evs1, stream));
        HIPCHECK(launch_ladder(stream, ni, lds_ladder, d_inst, d_reps, d_slots, d_ros, d_state, d_arch, nullptr,
                               d_seq));
        for (int k = 0; k < ni; ++k)
            if (want[k]) {
                iv_of[k] = want[k];
                done[k] += std::min(spe[k], total[k] - done[k]);
            }
        intervals_done += 1;
        if (have_pend)
            if (int rc = settle()) return rc;
        pend.swap(want);
        pend_timed = timed;
        have_pend = true;
    }
    if (have_pend)
        if (int rc = settle()) return rc;
    HIPCHECK(hipEventRecord(ev1, stream));
    HIPCHECK(hipEventSynchronize(ev1));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, ev0, ev1));
    device_ms += ms;
    last_ms = ms;
    HIPCHECK(hipMemcpyAsync(h_pin, d_inst, sizeof(InstDesc) * ni, hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    return SIS_OK;
}

// Wait until every instance with want[k] != 0 has published interval count
// want[k] or a later one (the interval queued behind it may have ended
// too).  The counts alone are polled for the first kQueryAfterMs: a stream
// query enqueues a marker whose system-scope release sits between the
// kernels already queued and stalls the next launch by microseconds.  After
// that the stream is queried every few hundred polls, and an error ends the
// wait with SIS_EDEVICE.  The stream's completion and the count written by
// its last kernel reach the host by different paths, so after the stream has
// drained the count is polled for a while longer, and the device's own
// descriptors decide in the end.
int Batch::wait_status(const std::vector<unsigned>& want) {
    constexpr int kQueryAfterMs = 50;
    const int ni = (int)want.size();
    auto reached = [](unsigned have, unsigned w) { return (int)(have - w) >= 0; };
    auto published = [&]() {
        for (int k = 0; k < ni; ++k)
            if (want[k] != 0u && !reached(__atomic_load_n(h_seq + k, __ATOMIC_ACQUIRE), want[k])) return false;
        return true;
    };
    const auto w0 = std::chrono::steady_clock::now();
    bool query = false;
    for (unsigned polls = 1;; ++polls) {
        if (published()) return SIS_OK;
        if ((polls & 255u) != 0u) {
            __builtin_ia32_pause();
        } else if (!query) {
            query = std::chrono::steady_clock::now() - w0 > std::chrono::milliseconds(kQueryAfterMs);
        } else {
            const hipError_t q = hipStreamQuery(stream);
            if (q == hipSuccess) break;
            if (q != hipErrorNotReady)
                return fail(SIS_EDEVICE, std::string("device error: ") + hipGetErrorString(q));
        }
    }
    const auto t0 = std::chrono::steady_clock::now();
    while (std::chrono::steady_clock::now() - t0 < std::chrono::milliseconds(20))
        if (published()) return SIS_OK;
    std::vector<InstDesc> dev(ni);
    HIPCHECK(hipMemcpyAsync(dev.data(), d_inst, sizeof(InstDesc) * ni, hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    std::string msg;
    for (int k = 0; k < ni; ++k)
        if (want[k] != 0u && !reached(dev[k].interval, want[k]))
            msg += " instance " + std::to_string(k) + ": interval " + std::to_string(dev[k].interval) +
                   " after the step, " + std::to_string(want[k]) + " expected;";
    return msg.empty() ? SIS_OK : fail(SIS_EDEVICE, "exchange step did not advance:" + msg);
}

int Batch::run(long long intervals, int mode) {
    if (intervals <= 0) intervals = std::max(0ll, max_intervals - intervals_done);
    if (intervals == 0) {
        last_ms = 0;
        return SIS_OK;
    }
    if (mode == 0) return run_plain(intervals);
    const int per = (int)std::min<long long>(intervals, 50);
    int rc = run_graph(per, intervals / per);
    if (rc) return rc;
    const double graph_ms = last_ms;
    const long long rest = intervals % per;
    if (rest) {
        HIPCHECK(hipEventRecord(ev0, stream));
        for (long long k = 0; k < rest; ++k)
            if ((rc = enqueue_interval())) return rc;
        HIPCHECK(hipEventRecord(ev1, stream));
        HIPCHECK(hipEventSynchronize(ev1));
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, ev0, ev1));
        device_ms += ms;
        last_ms = graph_ms + ms;
        intervals_done += rest;
    }
    // the host mirror after a graph run (untimed): one copy per call
    HIPCHECK(hipMemcpyAsync(h_pin, d_inst, sizeof(InstDesc) * specs.size(), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    return SIS_OK;
}

// Batched schedule: `per_graph` intervals captured once and replayed
// `launches` times.  No host synchronisation in between.
int Batch::run_graph(int per_graph, long long launches) {
    per_graph = std::max(1, per_graph);
    if (!graph_exec || graph_intervals != per_graph) {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
        HIPCHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        int rc = SIS_OK;
        if (chains.empty()) {
            for (int k = 0; k < per_graph && rc == SIS_OK; ++k) rc = enqueue_interval();
        } else {
            // the chains fork from the batch stream and join it at the end
            if (hipEventRecord(ev_cfork, stream) != hipSuccess) rc = fail(SIS_EDEVICE, "graph fork failed");
            for (Chain& c : chains)
                if (rc == SIS_OK && hipStreamWaitEvent(c.st, ev_cfork, 0) != hipSuccess)
                    rc = fail(SIS_EDEVICE, "graph fork failed");
            for (int k = 0; k < per_graph && rc == SIS_OK; ++k)
                for (int ci = 0; ci < (int)chains.size() && rc == SIS_OK; ++ci) rc = enqueue_chain(ci);
            for (Chain& c : chains)
                if (rc == SIS_OK && (hipEventRecord(c.done, c.st) != hipSuccess ||
                                     hipStreamWaitEvent(stream, c.done, 0) != hipSuccess))
                    rc = fail(SIS_EDEVICE, "graph join failed");
        }
        if (rc) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(stream, &g);
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        HIPCHECK(hipStreamEndCapture(stream, &graph));
        HIPCHECK(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
        graph_intervals = per_graph;
    }
    HIPCHECK(hipEventRecord(ev0, stream));
    for (long long k = 0; k < launches; ++k) HIPCHECK(hipGraphLaunch(graph_exec, stream));
    HIPCHECK(hipEventRecord(ev1, stream));
    HIPCHECK(hipEventSynchronize(ev1));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, ev0, ev1));
    device_ms += ms;
    last_ms = ms;
    intervals_done += launches * per_graph;
    return SIS_OK;
}

// FP64 polish + certified side of each instance's best configuration.
int Batch::finalize(std::vector<FinalRecord>& rec) {
    const int ni = (int)specs.size();
    HIPCHECK(hipMemcpyAsync(h_inst.data(), d_inst, sizeof(InstDesc) * ni, hipMemcpyDeviceToHost, stream));
    std::vector<RepState> h_reps(n_rep);
    HIPCHECK(hipMemcpyAsync(h_reps.data(), d_reps, sizeof(RepState) * n_rep, hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    std::vector<FinalizeJob> jobs(ni);
    rec.assign(ni, FinalRecord{});
    long long qoff = 0;
    for (int k = 0; k < ni; ++k) {
        const InstDesc& D = h_inst[k];
        FinalizeJob& j = jobs[k];
        j.n = D.n;
        j.iters = specs[k].p.polish_iters;
        j.src64 = nullptr;
        j.out_off = qoff;
        qoff += 3ll * D.n;
        int best = 0;
        for (int r = 1; r < D.R; ++r)
            if (h_reps[D.rep_off + r].E < h_reps[D.rep_off + best].E) best = r;
        if (finalize_source(D, j.L, rec[k].target_L)) {
            j.src32 = d_arch + D.arch_off;
            rec[k].e32 = D.E_arch;
        } else {
            j.src32 = d_state + D.state_off + (size_t)best * D.n;
            rec[k].e32 = h_reps[D.rep_off + best].E;
        }
    }
    for (void* p : {(void*)d_jobs, (void*)d_fout, (void*)d_qout})
        if (p) (void)hipFree(p);
    d_jobs = nullptr;
    d_fout = nullptr;
    d_qout = nullptr;
    HIPCHECK(hipMalloc(&d_jobs, sizeof(FinalizeJob) * ni));
    HIPCHECK(hipMalloc(&d_fout, sizeof(FinalizeOut) * ni));
    HIPCHECK(hipMalloc(&d_qout, sizeof(double) * std::max<long long>(1, qoff)));
    HIPCHECK(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(FinalizeJob) * ni, hipMemcpyHostToDevice, stream));
    HIPCHECK(launch_finalize(stream, ni, d_jobs, d_fout, d_qout));
    std::vector<FinalizeOut> fo(ni);
    std::vector<double> q(std::max<long long>(1, qoff));
    HIPCHECK(hipMemcpyAsync(fo.data(), d_fout, sizeof(FinalizeOut) * ni, hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipMemcpyAsync(q.data(), d_qout, sizeof(double) * q.size(), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    for (int k = 0; k < ni; ++k) {
        const InstDesc& D = h_inst[k];
        FinalRecord& r = rec[k];
        r.q.assign(q.begin() + jobs[k].out_off, q.begin() + jobs[k].out_off + 3ll * D.n);
        r.e64_before = fo[k].e64_before;
        r.e64_after = fo[k].e64_after;
        r.lambda = fo[k].lambda;
        r.L_cert = fo[k].L_cert;
        r.desc = D;
    }
    return SIS_OK;
}

void fill_result(const FinalRecord& r, const Batch& b, int k, double wall_s, sis_result* out) {
    const InstDesc& D = r.desc;
    if (out->q_out) std::copy(r.q.begin(), r.q.end(), out->q_out);
    out->target_L = r.target_L;
    out->L_cert = r.L_cert;
    out->lambda = r.lambda;
    out->e32 = r.e32;
    out->e64_before = r.e64_before;
    out->e64_after = r.e64_after;
    // the caller's seed exported as it was (nothing improved on it) was never
    // evaluated by a sweep: no FP32 energy, no gap
    const bool seed_export = D.improve && D.seeded && !D.arch_valid;
    out->e_gap = seed_export ? 0.0 : std::fabs(r.e32 - r.e64_before);
    out->proposals = D.proposals;
    out->accepted = D.accepted;
    out->sweeps = (uint64_t)D.sweeps_done;
    out->exchanges_tried = D.ex_tried;
    out->exchanges_accepted = D.ex_acc;
    out->candidates = D.cand;
    out->near_pairs = D.near;
    out->shrinks = D.shrinks;
    out->feasible = D.arch_valid ? 1 : 0;
    out->replicas = D.R;
    out->status = SIS_OK;
    out->seconds = wall_s;
    out->device_seconds = b.device_ms * 1e-3;
    out->fp64_rejects = D.fp64_rejects;
    out->kicks = D.kicks;
    out->collective_tried = D.coll_tried;
    out->collective_accepted = D.coll_acc;
    out->feasible_checkpoints = D.improve ? D.certs : (uint64_t)D.feasible;
    (void)k;
}

InstSpec make_spec(const sis_instance* in, const sis_params* p) {
    InstSpec s;
    s.n = in->n;
    s.L = in->L;
    if (in->q) s.q.assign(in->q, in->q + 3 * in->n);
    s.p = *p;
    return s;
}

}  // namespace sis

// ===========================================================================
// C ABI
// ===========================================================================
using namespace sis;

extern "C" {

int sis_init(int device) { return ensure_device(device); }

const char* sis_last_error(void) { return t_err.c_str(); }
int sis_last_error_code(void) { return t_code; }

int sis_device_info(int* compute_units, char* name, int name_len) {
    int rc = ensure_device(-1);
    if (rc) return rc;
    const DeviceInfo& d = device_info();
    if (compute_units) *compute_units = d.cus;
    if (name && name_len > 0) {
        std::strncpy(name, d.name.c_str(), (size_t)name_len - 1);
        name[name_len - 1] = 0;
    }
    return SIS_OK;
}

void sis_default_params(sis_params* p) {
    if (p) default_params(p);
}

int sis_portfolio(const sis_instance* inst, const sis_params* params, int jitter, double* out) {
    if (!out) return fail(SIS_EINVAL, "null output");
    int rc = validate(inst, params);
    if (rc) return rc;
    std::vector<double> q;
    portfolio_state(make_spec(inst, params), jitter != 0, q);
    std::copy(q.begin(), q.end(), out);
    return SIS_OK;
}

int sis_occupancy(const sis_instance* inst, const sis_params* params, int* blocks_per_cu,
                  int* replicas_per_block) {
    int rc = validate(inst, params);
    if (rc) return rc;
    if ((rc = ensure_device(-1))) return rc;
    const InstSpec s = make_spec(inst, params);
    Plan pl;
    if ((rc = plan_instance(s, s.p.replicas, pl))) return rc;
    const int threads = sweep_threads(s.n);
    int nb = 0;
    HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, sweep_kernel_ptr(threads, false, pl.wide), threads,
                                                          pl.lds));
    if (blocks_per_cu) *blocks_per_cu = nb;
    if (replicas_per_block) *replicas_per_block = pl.G;
    return SIS_OK;
}

int sis_solve(const sis_instance* inst, const sis_params* params, sis_result* out) {
    if (!out) return fail(SIS_EINVAL, "null result");
    int rc = validate(inst, params);
    if (rc) return out->status = rc;
    if ((rc = ensure_device(-1))) return out->status = rc;
    auto t0 = std::chrono::steady_clock::now();
    Batch b;
    std::vector<InstSpec> v;
    v.push_back(make_spec(inst, params));
    if ((rc = b.build(std::move(v)))) return out->status = rc;
    if ((rc = b.run(0, 0))) return out->status = rc;
    std::vector<FinalRecord> rec;
    if ((rc = b.finalize(rec))) return out->status = rc;
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    fill_result(rec[0], b, 0, wall, out);
    return SIS_OK;
}

int sis_solve_batch(const sis_instance* insts, int count, const sis_params* params, sis_result* out) {
    if (!insts || !params || !out || count < 1) return fail(SIS_EINVAL, "bad batch arguments");
    int rc;
    for (int k = 0; k < count; ++k)
        if ((rc = validate(insts + k, params + k))) {
            for (int j = 0; j < count; ++j) out[j].status = rc;
            return rc;
        }
    if ((rc = ensure_device(-1))) return rc;
    auto t0 = std::chrono::steady_clock::now();
    Batch b;
    std::vector<InstSpec> v;
    for (int k = 0; k < count; ++k) v.push_back(make_spec(insts + k, params + k));
    if ((rc = b.build(std::move(v)))) return rc;
    if ((rc = b.run(0, 1))) return rc;
    std::vector<FinalRecord> rec;
    if ((rc = b.finalize(rec))) return rc;
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int k = 0; k < count; ++k) fill_result(rec[k], b, k, wall, out + k);
    return SIS_OK;
}

namespace {

// One autotuner candidate: replicas per workgroup, and the two settings that
// change the trajectory (0 / 0.0: each instance keeps its own).
struct TuneCand {
    int G, spe;
    double cell;
    double rate = 0;
};

// The caller's own value of a setting for instance k, or the candidate's.
inline int cand_spe(const TuneCand& c, const sis_params& p) { return c.spe ? c.spe : p.sweeps_per_exchange; }
inline double cand_cell(const TuneCand& c, const sis_params& p) { return c.cell > 0 ? c.cell : p.cell; }

// Speed probe: `probe_sweeps` sweeps (whole intervals) of every instance under
// the candidate, rate by HIP-event time of the batched graph.  1: candidate
// invalid for some instance.
int tune_speed(const sis_instance* insts, int count, const sis_params* params, int probe_sweeps, TuneCand& c) {
    std::vector<InstSpec> v;
    long long props = 0;
    for (int k = 0; k < count; ++k) {
        InstSpec s = make_spec(insts + k, params + k);
        s.p.replicas_per_group = c.G;
        s.p.sweeps_per_exchange = cand_spe(c, params[k]);
        s.p.cell = cand_cell(c, params[k]);
        // whole intervals only, so every candidate does the same work
        s.p.sweeps = (std::max(probe_sweeps, s.p.sweeps_per_exchange) / s.p.sweeps_per_exchange) *
                     s.p.sweeps_per_exchange;
        if (validate(insts + k, &s.p)) return 1;
        props += (long long)s.n * s.p.replicas * s.p.sweeps;
        v.push_back(std::move(s));
    }
    Batch b;
    int rc;
    if ((rc = b.build(std::move(v)))) return -rc;
    if ((rc = b.run(0, 1))) return -rc;
    c.rate = props / (b.device_ms * 1e-3);
    return 0;
}

long long gcd_ll(long long a, long long b) { return b ? gcd_ll(b, a % b) : a; }

// Largest loss of mean relative side reduction (upper two-standard-error
// bound) a trajectory-changing candidate may show in the quality probe.
constexpr double kQualityMargin = 1e-3;

// Quality probe: every instance run for its share of `sweeps` (a multiple of
// the caller's and the candidate's sweeps per exchange, so both runs end on
// an interval) with the polish off; per instance the relative side reduction
// 1 - min(L_cert, L0) / L0 of its certified export.
int tune_quality(const sis_instance* insts, int count, const sis_params* params, const TuneCand& c, int sweeps,
                 const TuneCand& ref, std::vector<double>& red) {
    std::vector<InstSpec> v;
    for (int k = 0; k < count; ++k) {
        InstSpec s = make_spec(insts + k, params + k);
        const long long a = cand_spe(c, params[k]), b = cand_spe(ref, params[k]);
        const long long l = a / gcd_ll(a, b) * b;
        s.p.replicas_per_group = 0;
        s.p.sweeps_per_exchange = (int)a;
        s.p.cell = cand_cell(c, params[k]);
        s.p.sweeps = std::max<long long>(l, (sweeps / l) * l);
        s.p.polish_iters = 0;
        v.push_back(std::move(s));
    }
    Batch b;
    int rc;
    if ((rc = b.build(std::move(v)))) return rc;
    if ((rc = b.run(0, 1))) return rc;
    std::vector<FinalRecord> rec;
    if ((rc = b.finalize(rec))) return rc;
    red.resize(count);
    for (int k = 0; k < count; ++k) {
        const double L0 = insts[k].L;
        red[k] = 1.0 - std::min(rec[k].L_cert, L0) / L0;
    }
    return SIS_OK;
}

}  // namespace

int sis_autotune(const sis_instance* insts, int count, const sis_params* params, int probe_sweeps,
                 sis_tuning* out) {
    if (!insts || !params || !out || count < 1 || probe_sweeps < 1) return fail(SIS_EINVAL, "bad autotune arguments");
    int rc;
    for (int k = 0; k < count; ++k)
        if ((rc = validate(insts + k, params + k))) return rc;
    if ((rc = ensure_device(-1))) return rc;
    // 0: the automatic tile size of every instance (by its n and the load)
    const int groups[5] = {0, 1, 2, 4, 8};
    const int spes[3] = {5, 10, 20};
    const double cells[2] = {0.7075, 0.8};
    *out = sis_tuning{};
    std::vector<TuneCand> cands;
    for (double cell : cells)
        for (int spe : spes)
            for (int G : groups) cands.push_back({G, spe, cell});
    // a candidate keeps the trajectory when its sweeps per exchange and cell
    // are every instance's own (tile size never changes results)
    auto keeps = [&](const TuneCand& c) {
        for (int k = 0; k < count; ++k)
            if (cand_spe(c, params[k]) != params[k].sweeps_per_exchange || cand_cell(c, params[k]) != params[k].cell)
                return false;
        return true;
    };
    bool own_on_grid = false;
    for (const TuneCand& c : cands) own_on_grid |= keeps(c);
    if (!own_on_grid)
        for (int G : groups) cands.push_back({G, 0, 0.0});
    std::vector<TuneCand> ran;
    for (TuneCand& c : cands) {
        rc = tune_speed(insts, count, params, probe_sweeps, c);
        if (rc < 0) return -rc;
        if (rc == 0) ran.push_back(c);
    }
    if (ran.empty()) return fail(SIS_EINVAL, "no candidate configuration was valid");
    std::stable_sort(ran.begin(), ran.end(), [](const TuneCand& a, const TuneCand& b) { return a.rate > b.rate; });
    // Faster settings that change the trajectory are taken only if they search
    // as well: the quality probe runs the batch for 50 x probe_sweeps sweeps
    // under the candidate and under the caller's settings, and the candidate
    // passes when the mean paired loss of relative side reduction, plus two
    // standard errors, is at most kQualityMargin (a non-inferiority test; for
    // one instance the loss itself).  At most three are checked, fastest
    // first; then the fastest candidate that keeps the trajectory wins.
    const TuneCand own{0, 0, 0.0};
    const int qsweeps = 50 * probe_sweeps;
    std::vector<double> base, red;
    const TuneCand* win = nullptr;
    const TuneCand* last = nullptr;
    double delta = 0, se = 0;
    int checked = 0;
    for (const TuneCand& c : ran) {
        if (keeps(c)) {
            win = &c;
            break;
        }
        if (checked == 3) continue;
        // both runs stop at the same sweep of every instance
        if ((rc = tune_quality(insts, count, params, own, qsweeps, c, base))) return rc;
        if ((rc = tune_quality(insts, count, params, c, qsweeps, own, red))) return rc;
        ++checked;
        double m = 0, v = 0;
        for (int k = 0; k < count; ++k) m += base[k] - red[k];
        m /= count;
        for (int k = 0; k < count; ++k) v += (base[k] - red[k] - m) * (base[k] - red[k] - m);
        const double e = count > 1 ? std::sqrt(v / (count - 1) / count) : 0.0;
        last = &c;
        delta = m;
        se = e;
        if (m + 2.0 * e <= kQualityMargin) {
            win = &c;
            break;
        }
    }
    if (!win) return fail(SIS_EINVAL, "no candidate that keeps the trajectory was valid");
    out->replicas_per_group = win->G;
    out->sweeps_per_exchange = cand_spe(*win, params[0]);
    out->cell = cand_cell(*win, params[0]);
    out->proposals_per_s = win->rate;
    out->candidates_tried = (int)ran.size();
    out->trajectory_kept = keeps(*win) ? 1 : 0;
    out->quality_checked = checked;
    out->quality_sweeps = checked ? qsweeps : 0;
    if (last) {
        out->checked_sweeps_per_exchange = cand_spe(*last, params[0]);
        out->checked_cell = cand_cell(*last, params[0]);
        out->quality_delta = delta;
        out->quality_se = se;
    }
    return SIS_OK;
}

struct sis_batch {
    sis::Batch b;
    std::chrono::steady_clock::time_point t0;
};

sis_batch* sis_batch_create(const sis_instance* insts, int count, const sis_params* params) {
    if (!insts || !params || count < 1) {
        fail(SIS_EINVAL, "bad batch arguments");
        return nullptr;
    }
    for (int k = 0; k < count; ++k)
        if (validate(insts + k, params + k)) return nullptr;
    if (ensure_device(-1)) return nullptr;
    auto* h = new sis_batch;
    h->t0 = std::chrono::steady_clock::now();
    std::vector<InstSpec> v;
    for (int k = 0; k < count; ++k) v.push_back(make_spec(insts + k, params + k));
    if (h->b.build(std::move(v))) {
        delete h;
        return nullptr;
    }
    return h;
}

int sis_batch_run(sis_batch* h, int64_t intervals, int mode, double* sweep_ms, int64_t* sweep_launches,
                  double* total_ms) {
    if (!h || (mode != 0 && mode != 1)) return fail(SIS_EINVAL, "bad batch handle or mode");
    const double s0 = h->b.sweep_ms;
    const long long l0 = h->b.sweep_launches;
    int rc = h->b.run(intervals, mode);
    if (rc) return rc;
    if (sweep_ms) *sweep_ms = h->b.sweep_ms - s0;
    if (sweep_launches) *sweep_launches = h->b.sweep_launches - l0;
    if (total_ms) *total_ms = h->b.last_ms;
    return SIS_OK;
}

int sis_batch_finalize(sis_batch* h, sis_result* out) {
    if (!h || !out) return fail(SIS_EINVAL, "bad batch handle");
    std::vector<FinalRecord> rec;
    int rc = h->b.finalize(rec);
    if (rc) return rc;
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - h->t0).count();
    for (size_t k = 0; k < rec.size(); ++k) fill_result(rec[k], h->b, (int)k, wall, out + k);
    return SIS_OK;
}

int sis_batch_destroy(sis_batch* h) {
    delete h;
    return SIS_OK;
}

static void ladder_desc_of(const InstDesc& D, sis_ladder_desc* d) {
    d->n = D.n; d->R = D.R; d->ladder = D.ladder; d->improve = D.improve; d->cool_period = D.cool_period;
    d->spe = D.spe; d->sweeps_total = D.sweeps_total; d->sweeps_done = D.sweeps_done;
    d->key0 = D.key0; d->key1 = D.key1; d->interval = D.interval;
    d->shrink = D.shrink; d->target_acc = D.target_acc; d->scale = D.scale; d->Lf = D.Lf;
    d->L = D.L; d->L_arch = D.L_arch; d->L_box_arch = D.L_box_arch;
    d->arch_valid = D.arch_valid; d->shrinks = D.shrinks; d->feasible = D.feasible;
    d->fp64_rejects = D.fp64_rejects; d->ex_tried = D.ex_tried; d->ex_acc = D.ex_acc; d->certs = D.certs;
}

int sis_batch_status(sis_batch* h, int k, sis_ladder_desc* d) {
    if (!h || !d || k < 0 || k >= (int)h->b.specs.size()) return fail(SIS_EINVAL, "bad batch handle or instance");
    ladder_desc_of(h->b.h_pin[k], d);
    return SIS_OK;
}

int sis_dev_batch_state(sis_batch* h, int k, sis_ladder_desc* d, sis_rep_state* reps, sis_slot_state* slots,
                        int* rep_of_slot, float* xycs, float* arch) {
    if (!h || k < 0 || k >= (int)h->b.specs.size()) return fail(SIS_EINVAL, "bad batch handle or instance");
    Batch& b = h->b;
    InstDesc D;
    SIS_HIPCHECK(hipMemcpyAsync(&D, b.d_inst + k, sizeof(InstDesc), hipMemcpyDeviceToHost, b.stream));
    SIS_HIPCHECK(hipStreamSynchronize(b.stream));
    const size_t R = (size_t)D.R, n = (size_t)D.n;
    if (reps) SIS_HIPCHECK(hipMemcpy(reps, b.d_reps + D.rep_off, sizeof(RepState) * R, hipMemcpyDeviceToHost));
    if (slots) SIS_HIPCHECK(hipMemcpy(slots, b.d_slots + D.rep_off, sizeof(SlotState) * R, hipMemcpyDeviceToHost));
    if (rep_of_slot) SIS_HIPCHECK(hipMemcpy(rep_of_slot, b.d_ros + D.rep_off, sizeof(int) * R, hipMemcpyDeviceToHost));
    if (xycs) SIS_HIPCHECK(hipMemcpy(xycs, b.d_state + D.state_off, sizeof(float4) * R * n, hipMemcpyDeviceToHost));
    if (arch) SIS_HIPCHECK(hipMemcpy(arch, b.d_arch + D.arch_off, sizeof(float4) * n, hipMemcpyDeviceToHost));
    if (d) ladder_desc_of(D, d);
    return SIS_OK;
}

}  // extern "C"
