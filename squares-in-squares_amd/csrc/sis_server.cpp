// Continuous-batching request loop.
//
// Replaces the reference's plain request loop (cuda/cpp_loop.py:82-106: one
// `packing` process per n, at most four in flight, admission by free VRAM and
// RAM, cuda/cpp_loop.py:41-45, then the C++ certifier gates,
// cuda/cpp_loop.py:17-28).  Here every admitted request keeps its replicas
// resident in one device arena.  Requests are split into size classes
// (n < 64, n < 256, the rest) served by lanes (one each, two for the last
// class); each lane has its own stream and runs rounds of
// `intervals_per_round` sweep + exchange steps over the tiles of all its
// resident requests, so the short launches of small instances are not held
// back by the long ones of large instances and the lanes' kernels overlap on
// the device.  Requests that arrive while others run are admitted at their
// lane's next round boundary as slots and memory free up.  Finished requests
// get the FP64 polish / certified side on a stream of their own and then the
// exact certificate on a host thread pool.  A request's arithmetic depends only
// on the request (its seed, n, parameters), never on what else is resident.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/sis_abi.h"
#include "sis_certify.h"
#include "sis_host.h"

namespace sis {

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator of [offset, offset + length) ranges.
class RangeAlloc {
public:
    explicit RangeAlloc(size_t cap = 0) { reset(cap); }
    void reset(size_t cap) {
        free_.clear();
        if (cap) free_[0] = cap;
    }
    bool alloc(size_t len, size_t& off) {
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= len) {
                off = it->first;
                const size_t rest = it->second - len;
                free_.erase(it);
                if (rest) free_[off + len] = rest;
                return true;
            }
        return false;
    }
    void release(size_t off, size_t len) {
        auto it = free_.emplace(off, len).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

private:
    std::map<size_t, size_t> free_;
};

struct Req {
    uint64_t ticket = 0;
    InstSpec spec;
    Plan plan;
    int lane = 0;
    std::vector<float4> init;   // initial replica state, built by the submitting thread
    double t_submit = 0, t_admit = -1, t_done = -1, t_cert = -1;
    int slot = -1;
    size_t sq_off = 0, rep_off = 0, arch_off = 0;
    size_t bytes = 0;
    bool done = false;          // result (and certificate, if asked) available
    sis_result res{};
    std::vector<double> q;
    std::string L50, cert_reason;
    double dilation = 0;
    bool certified = false;
    double deadline = 0;        // now_s() value after which the request stops; 0: none
    bool timed_out = false;     // stopped at a round boundary past its deadline
    int fail_code = 0;          // SIS_EDEVICE (its lane's round failed) or SIS_ETIMEOUT (never admitted)
    std::string fail_msg;
};

size_t request_bytes(int n, int R) {
    return (size_t)n * R * sizeof(float4) + (size_t)n * sizeof(float4) +
           (size_t)R * (sizeof(RepState) + sizeof(SlotState) + sizeof(int));
}

// Size classes n < 64, 64..255 and n >= 256.  The largest class (most of the
// work) is served by two lanes that take its requests in turn: each lane is a
// chain of launches that waits for its own last workgroups, and the two fill
// each other's tails (served stream at 20/s: 1.42e10 -> 1.52e10 proposals/s,
// p99 6.0 -> 4.7 s).  More lanes than that (a second mid-size lane, a third
// large one) lost p99 again (7.2-8.4 s): with the finalize stream they
// outnumber the device's hardware queues, which then serialise them.
constexpr int kMidLanes = 1, kBigLanes = 2;
constexpr int kLanes = 1 + kMidLanes + kBigLanes;
int size_class(int n) { return n < 64 ? 0 : (n < 256 ? 1 : 2); }
int lane_of(int cls, uint64_t seq) {
    return cls == 0 ? 0 : cls == 1 ? 1 + (int)(seq % kMidLanes) : 1 + kMidLanes + (int)(seq % kBigLanes);
}

// One size class of the loop (loop thread only).
struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t done_ev = nullptr;
    bool in_flight = false;
    bool dirty = true;
    bool planned_loaded = false;      // device load the tiles were planned for
    bool near_pairs = false;          // some active request counts near pairs (SIS_FLAG_NEAR_PAIRS)
    bool alone = false;               // every tile planned for an idle device (one-wavefront-per-SIMD kernels)
    std::vector<Req*> active;         // resident, in list order
    std::vector<Tile> tiles;          // four-wavefront tiles first (n_big of them)
    std::vector<int> list;            // descriptor slots of `active`
    int n_big = 0;
    size_t lds_big = 0, lds_small = 0, lds_ladder = 0;
    Tile* d_tiles = nullptr;
    int* d_list = nullptr;
    InstDesc* d_desc = nullptr;       // descriptors of `active` after the round
    InstDesc* h_desc = nullptr;       // pinned mirror
};

// A finalize launch: up to kFinJobs finished requests, device buffers and
// pinned mirrors allocated once.  Its requests keep their arena ranges and
// slots until it completes.
constexpr int kFinJobs = 32;
constexpr int kFinBlocks = 8;
struct FinBlock {
    FinalizeJob* d_jobs = nullptr;
    FinalizeOut* d_fo = nullptr;
    double* d_q = nullptr;
    FinalizeJob* h_jobs = nullptr;
    FinalizeOut* h_fo = nullptr;
    double* h_q = nullptr;
    hipEvent_t ev = nullptr;
    std::vector<Req*> reqs;
    std::vector<InstDesc> desc;       // descriptors at the end of each request
    bool busy = false;
};

}  // namespace

struct Server {
    sis_server_cfg cfg{};
    std::mutex mu;
    std::condition_variable cv, cert_cv;
    std::deque<Req*> queue[kLanes];   // submitted, not yet admitted (under mu)
    std::unordered_map<uint64_t, Req*> all;
    std::deque<Req*> cert_queue;
    uint64_t next_ticket = 1;
    uint64_t class_seq[3] = {0, 0, 0};  // requests submitted per size class (turns of its lanes)
    bool stop = false;
    bool loop_done = false;  // the device loop has exited (set under mu)
    std::thread loop;
    std::vector<std::thread> certifiers;
    std::string error;       // under mu once the loop runs
    int fault_lane = -1;     // test hook: the lane whose next round fails (under mu)
    size_t arena_bytes = 0;
    size_t peak_resident = 0;  // under mu
    int resident_reqs = 0;     // under mu

    // loop thread only
    std::string loop_err;
    size_t resident_bytes = 0;
    size_t cap_sq = 0, cap_rep = 0;
    int cap_inst = 0;
    RangeAlloc sq_alloc, rep_alloc, arch_alloc;
    std::vector<int> free_slots;
    std::vector<InstDesc> h_inst;
    Lane lanes[kLanes];
    FinBlock fins[kFinBlocks];
    hipStream_t fin_stream = nullptr;
    InstDesc* d_inst = nullptr;
    RepState* d_reps = nullptr;
    SlotState* d_slots = nullptr;
    int* d_ros = nullptr;
    float4* d_state = nullptr;
    float4* d_arch = nullptr;

    int init();
    void run();
    void certify_worker();
    int admit_one(Req* r, Lane& ln);
    int launch_round(Lane& ln);
    int round_done(Lane& ln);
    bool fail_lane(Lane& ln, const std::string& why);
    void release_req(Req* r);
    int finalize_start(std::vector<Req*>& done);
    int finalize_done(FinBlock& fb);
    bool work_left();
    void release_dev();
};

#define SRV_CHECK(x)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            loop_err = std::string(#x) + ": " + hipGetErrorString(e_);                   \
            return SIS_EDEVICE;                                                            \
        }                                                                                  \
    } while (0)

int Server::init() {
    if (ensure_device(-1)) return SIS_EDEVICE;
    cap_inst = std::max(1, cfg.max_instances);
    // arena: the admission budget (4 GiB when none is given), at most 90 % of
    // the device memory free now less the per-instance buffers; 90 % of it for
    // replica states, 10 % for the per-replica records (replica, slot, owner
    // and the lanes' tiles: 180 bytes a replica, 7 % of a request at n = 150,
    // half of one at n = 10, where the instance cap binds first)
    size_t free_b = 0, total_b = 0;
    SRV_CHECK(hipMemGetInfo(&free_b, &total_b));
    const double fixed = (double)cap_inst * (kMaxN * sizeof(float4) + kLanes * (sizeof(int) + 2 * sizeof(InstDesc)) +
                                             sizeof(InstDesc)) +
                         (double)kFinBlocks * kFinJobs * (sizeof(FinalizeJob) + sizeof(FinalizeOut) + 3.0 * kMaxN * 8);
    const double budget = cfg.hbm_budget_bytes > 0 ? cfg.hbm_budget_bytes : 4.0 * (1u << 30);
    const double usable = 0.9 * (double)free_b - fixed;
    if (!(usable > 0)) {
        loop_err = "not enough free device memory for the server";
        return SIS_EDEVICE;
    }
    const double arena = std::min(budget, usable);
    const double per_rep = sizeof(RepState) + sizeof(SlotState) + sizeof(int) + kLanes * sizeof(Tile);
    cap_sq = (size_t)(0.9 * arena / sizeof(float4));
    cap_rep = std::min<size_t>((size_t)(0.1 * arena / per_rep), (size_t)cap_inst * 4096);
    arena_bytes = (size_t)((double)cap_sq * sizeof(float4) + (double)cap_rep * per_rep);
    sq_alloc.reset(cap_sq);
    rep_alloc.reset(cap_rep);
    arch_alloc.reset((size_t)cap_inst * kMaxN);
    for (int s = cap_inst - 1; s >= 0; --s) free_slots.push_back(s);
    h_inst.assign(cap_inst, InstDesc{});
    SRV_CHECK(hipMalloc(&d_inst, sizeof(InstDesc) * cap_inst));
    SRV_CHECK(hipMalloc(&d_reps, sizeof(RepState) * cap_rep));
    SRV_CHECK(hipMalloc(&d_slots, sizeof(SlotState) * cap_rep));
    SRV_CHECK(hipMalloc(&d_ros, sizeof(int) * cap_rep));
    SRV_CHECK(hipMalloc(&d_state, sizeof(float4) * cap_sq));
    SRV_CHECK(hipMalloc(&d_arch, sizeof(float4) * (size_t)cap_inst * kMaxN));
    for (Lane& ln : lanes) {
        SRV_CHECK(hipStreamCreateWithFlags(&ln.stream, hipStreamNonBlocking));
        SRV_CHECK(hipEventCreateWithFlags(&ln.done_ev, hipEventDisableTiming));
        // a lane never has more tiles than replicas resident
        SRV_CHECK(hipMalloc(&ln.d_tiles, sizeof(Tile) * cap_rep));
        SRV_CHECK(hipMalloc(&ln.d_list, sizeof(int) * cap_inst));
        SRV_CHECK(hipMalloc(&ln.d_desc, sizeof(InstDesc) * cap_inst));
        SRV_CHECK(hipHostMalloc(&ln.h_desc, sizeof(InstDesc) * cap_inst));
    }
    SRV_CHECK(hipStreamCreateWithFlags(&fin_stream, hipStreamNonBlocking));
    for (FinBlock& fb : fins) {
        SRV_CHECK(hipMalloc(&fb.d_jobs, sizeof(FinalizeJob) * kFinJobs));
        SRV_CHECK(hipMalloc(&fb.d_fo, sizeof(FinalizeOut) * kFinJobs));
        SRV_CHECK(hipMalloc(&fb.d_q, sizeof(double) * 3 * kMaxN * kFinJobs));
        SRV_CHECK(hipHostMalloc(&fb.h_jobs, sizeof(FinalizeJob) * kFinJobs));
        SRV_CHECK(hipHostMalloc(&fb.h_fo, sizeof(FinalizeOut) * kFinJobs));
        SRV_CHECK(hipHostMalloc(&fb.h_q, sizeof(double) * 3 * kMaxN * kFinJobs));
        SRV_CHECK(hipEventCreateWithFlags(&fb.ev, hipEventDisableTiming));
    }
    return SIS_OK;
}

void Server::release_dev() {
    for (void* p : {(void*)d_inst, (void*)d_reps, (void*)d_slots, (void*)d_ros, (void*)d_state,
                    (void*)d_arch})
        if (p) (void)hipFree(p);
    for (Lane& ln : lanes) {
        if (ln.d_tiles) (void)hipFree(ln.d_tiles);
        if (ln.d_list) (void)hipFree(ln.d_list);
        if (ln.d_desc) (void)hipFree(ln.d_desc);
        if (ln.h_desc) (void)hipHostFree(ln.h_desc);
        if (ln.done_ev) (void)hipEventDestroy(ln.done_ev);
        if (ln.stream) (void)hipStreamDestroy(ln.stream);
    }
    for (FinBlock& fb : fins) {
        for (void* p : {(void*)fb.d_jobs, (void*)fb.d_fo, (void*)fb.d_q})
            if (p) (void)hipFree(p);
        for (void* p : {(void*)fb.h_jobs, (void*)fb.h_fo, (void*)fb.h_q})
            if (p) (void)hipHostFree(p);
        if (fb.ev) (void)hipEventDestroy(fb.ev);
    }
    if (fin_stream) (void)hipStreamDestroy(fin_stream);
}

// Device state of a newly admitted request, on its lane's stream (the lane is
// idle: admission happens at round boundaries).
int Server::admit_one(Req* r, Lane& ln) {
    const InstSpec& s = r->spec;
    const int n = s.n, R = s.p.replicas;
    InstDesc D{};
    std::vector<RepState> reps;
    std::vector<SlotState> slots;
    std::vector<int> ros;
    instance_records(s, r->plan, D, reps, slots, ros);
    D.state_off = (int)r->sq_off; D.rep_off = (int)r->rep_off; D.arch_off = (int)r->arch_off;
    h_inst[r->slot] = D;
    const std::vector<float4>& st = r->init;
    hipStream_t sm = ln.stream;
    SRV_CHECK(hipMemcpyAsync(d_state + r->sq_off, st.data(), sizeof(float4) * st.size(), hipMemcpyHostToDevice, sm));
    SRV_CHECK(hipMemcpyAsync(d_arch + r->arch_off, st.data(), sizeof(float4) * n, hipMemcpyHostToDevice, sm));
    SRV_CHECK(hipMemcpyAsync(d_reps + r->rep_off, reps.data(), sizeof(RepState) * R, hipMemcpyHostToDevice, sm));
    SRV_CHECK(hipMemcpyAsync(d_slots + r->rep_off, slots.data(), sizeof(SlotState) * R, hipMemcpyHostToDevice, sm));
    SRV_CHECK(hipMemcpyAsync(d_ros + r->rep_off, ros.data(), sizeof(int) * R, hipMemcpyHostToDevice, sm));
    SRV_CHECK(hipMemcpyAsync(d_inst + r->slot, &h_inst[r->slot], sizeof(InstDesc), hipMemcpyHostToDevice, sm));
    SRV_CHECK(hipStreamSynchronize(sm));  // the host sources are released next
    std::vector<float4>().swap(r->init);
    return SIS_OK;
}

// One round of a lane: intervals of sweep + exchange over its tiles, then the
// descriptors of its requests into pinned memory, then the lane's event.
int Server::launch_round(Lane& ln) {
    // the lanes share the device: tiles are sized by the replicas resident in
    // all of them, and re-planned when that load crosses the threshold the
    // tile rule uses (one replica per tile and staging for every tile size
    // below it), so the last requests of a drain run as they would alone
    long long total_reps = 0;
    for (const Lane& other : lanes)
        for (const Req* r : other.active) total_reps += r->spec.p.replicas;
    const bool loaded = device_loaded(total_reps);
    if (ln.dirty || loaded != ln.planned_loaded) {
        ln.tiles.clear();
        ln.list.clear();
        ln.lds_big = ln.lds_small = ln.lds_ladder = 0;
        ln.near_pairs = false;
        ln.alone = true;
        for (Req* r : ln.active) {
            ln.near_pairs = ln.near_pairs || (r->spec.p.flags & SIS_FLAG_NEAR_PAIRS) != 0;
            Plan pl;
            if (plan_instance(r->spec, (int)std::min<long long>(total_reps, 1 << 30), pl)) {
                loop_err = "tile planning failed";
                return SIS_EDEVICE;
            }
            r->plan.G = pl.G;  // grouping never changes a replica's arithmetic
            ln.alone = ln.alone && pl.wide;
            size_t& lds = r->spec.n >= kBigN ? ln.lds_big : ln.lds_small;
            lds = std::max(lds, pl.lds);
            ln.lds_ladder = std::max(ln.lds_ladder, ladder_lds_bytes(r->spec.n, r->spec.L, r->spec.p.replicas));
            for (int r0 = 0; r0 < r->spec.p.replicas; r0 += pl.G)
                ln.tiles.push_back(Tile{r->slot, r0, std::min(pl.G, r->spec.p.replicas - r0), pl.wide});
            ln.list.push_back(r->slot);
        }
        ln.n_big = order_tiles(ln.tiles, [&](int slot) { return h_inst[slot].n; });
        SRV_CHECK(hipMemcpyAsync(ln.d_tiles, ln.tiles.data(), sizeof(Tile) * ln.tiles.size(),
                                 hipMemcpyHostToDevice, ln.stream));
        SRV_CHECK(hipMemcpyAsync(ln.d_list, ln.list.data(), sizeof(int) * ln.list.size(),
                                 hipMemcpyHostToDevice, ln.stream));
        // the host vectors stay unchanged until the lane's next round boundary
        ln.dirty = false;
        ln.planned_loaded = loaded;
    }
    const int ipr = std::max(1, cfg.intervals_per_round);
    const int nt = (int)ln.tiles.size();
    {
        std::lock_guard<std::mutex> lk(mu);
        if (fault_lane == (int)(&ln - lanes)) {
            fault_lane = -1;
            loop_err = "injected launch failure";
            return SIS_EDEVICE;
        }
    }
    for (int k = 0; k < ipr; ++k) {
        if (ln.n_big > 0)
            SRV_CHECK(launch_sweep(ln.stream, ln.n_big, ln.lds_big, ln.d_tiles, d_inst, d_reps, d_slots, d_state,
                                   kSweepThreadsBig, ln.near_pairs, ln.alone));
        if (nt > ln.n_big)
            SRV_CHECK(launch_sweep(ln.stream, nt - ln.n_big, ln.lds_small, ln.d_tiles + ln.n_big, d_inst, d_reps,
                                   d_slots, d_state, kSweepThreads, ln.near_pairs, ln.alone));
        SRV_CHECK(launch_ladder(ln.stream, (int)ln.list.size(), ln.lds_ladder, d_inst, d_reps, d_slots, d_ros, d_state, d_arch,
                                ln.d_list));
    }
    SRV_CHECK(launch_gather_desc(ln.stream, (int)ln.list.size(), d_inst, ln.d_list, ln.d_desc));
    SRV_CHECK(hipMemcpyAsync(ln.h_desc, ln.d_desc, sizeof(InstDesc) * ln.list.size(), hipMemcpyDeviceToHost,
                             ln.stream));
    SRV_CHECK(hipEventRecord(ln.done_ev, ln.stream));
    ln.in_flight = true;
    return SIS_OK;
}

// A lane's round has completed: pick up its descriptors and hand finished
// requests to the finalize stream.
int Server::round_done(Lane& ln) {
    ln.in_flight = false;
    std::vector<Req*> done, keep;
    const double t = now_s();
    for (size_t k = 0; k < ln.active.size(); ++k) {
        Req* r = ln.active[k];
        h_inst[r->slot] = ln.h_desc[k];
        const InstDesc& D = h_inst[r->slot];
        const bool finished = D.sweeps_done >= D.sweeps_total;
        // past its deadline: finalized from where it is, at this round boundary
        if (!finished && r->deadline > 0 && t >= r->deadline) r->timed_out = true;
        (finished || r->timed_out ? done : keep).push_back(r);
    }
    if (done.empty()) return SIS_OK;
    ln.active.swap(keep);
    ln.dirty = true;
    return finalize_start(done);
}

int Server::finalize_start(std::vector<Req*>& done) {
    for (size_t first = 0; first < done.size(); first += kFinJobs) {
        FinBlock* fb = nullptr;
        for (;;) {
            for (FinBlock& f : fins)
                if (!f.busy) { fb = &f; break; }
            if (fb) break;
            // every block in flight: wait for the oldest to finish
            for (FinBlock& f : fins)
                if (f.busy) {
                    SRV_CHECK(hipEventSynchronize(f.ev));
                    if (int rc = finalize_done(f)) return rc;
                    break;
                }
        }
        const int k = (int)std::min<size_t>(kFinJobs, done.size() - first);
        fb->reqs.assign(done.begin() + first, done.begin() + first + k);
        fb->desc.resize(k);
        for (int j = 0; j < k; ++j) {
            Req* r = fb->reqs[j];
            const InstDesc& D = h_inst[r->slot];
            fb->desc[j] = D;
            FinalizeJob& job = fb->h_jobs[j];
            job = FinalizeJob{};
            job.n = D.n;
            job.iters = r->spec.p.polish_iters;
            job.out_off = 3ll * kMaxN * j;
            double target_L;
            if (finalize_source(D, job.L, target_L)) {
                job.src32 = d_arch + D.arch_off;
            } else {
                // the best replica, chosen on the device
                job.src32 = d_state + D.state_off;
                job.reps = d_reps + D.rep_off;
                job.R = D.R;
            }
        }
        SRV_CHECK(hipMemcpyAsync(fb->d_jobs, fb->h_jobs, sizeof(FinalizeJob) * k, hipMemcpyHostToDevice, fin_stream));
        SRV_CHECK(launch_finalize(fin_stream, k, fb->d_jobs, fb->d_fo, fb->d_q));
        SRV_CHECK(hipMemcpyAsync(fb->h_fo, fb->d_fo, sizeof(FinalizeOut) * k, hipMemcpyDeviceToHost, fin_stream));
        SRV_CHECK(hipMemcpyAsync(fb->h_q, fb->d_q, sizeof(double) * 3 * kMaxN * k, hipMemcpyDeviceToHost, fin_stream));
        SRV_CHECK(hipEventRecord(fb->ev, fin_stream));
        fb->busy = true;
    }
    return SIS_OK;
}

// A finalize launch has completed: results out, arena ranges and slots back.
int Server::finalize_done(FinBlock& fb) {
    const double t = now_s();
    std::lock_guard<std::mutex> lk(mu);
    for (size_t j = 0; j < fb.reqs.size(); ++j) {
        Req* r = fb.reqs[j];
        const InstDesc& D = fb.desc[j];
        const FinalizeOut& fo = fb.h_fo[j];
        FinalRecord rec;
        rec.q.assign(fb.h_q + 3 * kMaxN * j, fb.h_q + 3 * kMaxN * j + 3ll * D.n);
        double box;
        finalize_source(D, box, rec.target_L);
        rec.L_cert = fo.L_cert;
        rec.lambda = fo.lambda;
        rec.e32 = (D.arch_valid || (D.improve && D.seeded)) ? (double)D.E_arch : fo.e32;
        rec.e64_before = fo.e64_before;
        rec.e64_after = fo.e64_after;
        rec.desc = D;
        r->q = rec.q;
        sis_result res{};
        res.q_out = nullptr;
        Batch none;
        fill_result(rec, none, 0, t - r->t_admit, &res);
        res.device_seconds = 0;
        if (r->timed_out) res.status = SIS_ETIMEOUT;
        r->res = res;
        r->t_done = t;
        release_req(r);
        if (cfg.certify)
            cert_queue.push_back(r);
        else
            r->done = true;
    }
    fb.reqs.clear();
    fb.busy = false;
    cert_cv.notify_all();
    return SIS_OK;
}

// Arena ranges, slot and budget of a resident request back (under mu).
void Server::release_req(Req* r) {
    const int n = r->spec.n, R = r->spec.p.replicas;
    sq_alloc.release(r->sq_off, (size_t)n * R);
    rep_alloc.release(r->rep_off, R);
    arch_alloc.release(r->arch_off, n);
    free_slots.push_back(r->slot);
    resident_bytes -= r->bytes;
    resident_reqs -= 1;
    r->slot = -1;
}

// A round of this lane failed.  Its stream is drained first (launches the
// round enqueued before the failure still run, on the lane's own arena
// ranges); if that fails too the device itself is in error and the loop stops
// (returns false).  Otherwise every request of the lane fails with
// SIS_EDEVICE, its memory is released, and the other lanes go on.
bool Server::fail_lane(Lane& ln, const std::string& why) {
    const hipError_t e = hipStreamSynchronize(ln.stream);
    if (e != hipSuccess) {
        loop_err = why + "; then " + hipGetErrorString(e);
        return false;
    }
    (void)hipGetLastError();  // the failed launch's error is not sticky
    std::lock_guard<std::mutex> lk(mu);
    for (Req* r : ln.active) {
        release_req(r);
        r->fail_code = SIS_EDEVICE;
        r->fail_msg = "device round failed: " + why;
        r->t_done = now_s();
        r->done = true;
    }
    ln.active.clear();
    ln.in_flight = false;
    ln.dirty = true;
    return true;
}

bool Server::work_left() {
    for (const Lane& ln : lanes)
        if (ln.in_flight || !ln.active.empty()) return true;
    for (const FinBlock& fb : fins)
        if (fb.busy) return true;
    return false;
}

void Server::run() {
    auto die = [&](const std::string& why) {
        std::lock_guard<std::mutex> lk(mu);
        error = why.empty() ? std::string("device loop failed") : why;
        stop = true;
    };
    if (hipSetDevice(device_info().id) != hipSuccess) {
        die("hipSetDevice failed in the server loop");
        return;
    }
    for (;;) {
        bool progressed = false;
        // finished finalize launches
        for (FinBlock& fb : fins) {
            if (!fb.busy) continue;
            const hipError_t q = hipEventQuery(fb.ev);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) {
                die(std::string("finalize failed: ") + hipGetErrorString(q));
                return;
            }
            if (finalize_done(fb) != SIS_OK) {
                die(loop_err);
                return;
            }
            progressed = true;
        }
        for (int li = 0; li < kLanes; ++li) {
            Lane& ln = lanes[li];
            if (ln.in_flight) {
                const hipError_t q = hipEventQuery(ln.done_ev);
                if (q == hipErrorNotReady) continue;
                if (q != hipSuccess) {
                    if (!fail_lane(ln, hipGetErrorString(q))) {
                        die(std::string("device round failed: ") + loop_err);
                        return;
                    }
                    progressed = true;
                    continue;
                }
                if (round_done(ln) != SIS_OK) {
                    die(loop_err);
                    return;
                }
                progressed = true;
            }
            // admission into this lane: FIFO while a slot, the HBM budget and
            // arena ranges allow
            std::vector<Req*> admitted;
            {
                std::lock_guard<std::mutex> lk(mu);
                std::deque<Req*>& qu = queue[li];
                // requests whose deadline passed while queued are never admitted
                const double tq = now_s();
                for (auto it = qu.begin(); it != qu.end();) {
                    Req* r = *it;
                    if (r->deadline > 0 && tq >= r->deadline) {
                        r->fail_code = SIS_ETIMEOUT;
                        r->fail_msg = "deadline passed before admission";
                        r->t_done = tq;
                        r->done = true;
                        it = qu.erase(it);
                    } else {
                        ++it;
                    }
                }
                while (!qu.empty() && !free_slots.empty()) {
                    Req* r = qu.front();
                    const double budget = cfg.hbm_budget_bytes > 0 ? cfg.hbm_budget_bytes : 1e300;
                    if (resident_bytes > 0 && resident_bytes + r->bytes > budget) break;
                    size_t so, ro, ao;
                    const int n = r->spec.n, R = r->spec.p.replicas;
                    if (!sq_alloc.alloc((size_t)n * R, so)) break;
                    if (!rep_alloc.alloc(R, ro)) { sq_alloc.release(so, (size_t)n * R); break; }
                    if (!arch_alloc.alloc(n, ao)) { sq_alloc.release(so, (size_t)n * R); rep_alloc.release(ro, R); break; }
                    qu.pop_front();
                    r->sq_off = so; r->rep_off = ro; r->arch_off = ao;
                    r->slot = free_slots.back();
                    free_slots.pop_back();
                    r->t_admit = now_s();
                    resident_bytes += r->bytes;
                    resident_reqs += 1;
                    peak_resident = std::max(peak_resident, resident_bytes);
                    admitted.push_back(r);
                }
            }
            // every admitted request joins the lane before its state is
            // copied in, so that a failed copy fails all of them
            for (Req* r : admitted) ln.active.push_back(r);
            if (!admitted.empty()) ln.dirty = true;
            for (Req* r : admitted)
                if (admit_one(r, ln) != SIS_OK) {
                    if (!fail_lane(ln, loop_err)) {
                        die(loop_err);
                        return;
                    }
                    break;
                }
            if (!ln.active.empty()) {
                if (launch_round(ln) != SIS_OK) {
                    if (!fail_lane(ln, loop_err)) {
                        die(loop_err);
                        return;
                    }
                }
                progressed = true;
            }
        }
        if (progressed) continue;
        std::unique_lock<std::mutex> lk(mu);
        bool queued = false;
        for (const auto& qu : queue) queued = queued || !qu.empty();
        if (!work_left()) {
            if (queued) continue;  // admission is waiting for memory that nothing will free
            if (stop) break;
            cv.wait_for(lk, std::chrono::milliseconds(2));
        } else {
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(50));
        }
    }
}

void Server::certify_worker() {
    for (;;) {
        Req* r;
        {
            std::unique_lock<std::mutex> lk(mu);
            cert_cv.wait(lk, [&] { return !cert_queue.empty() || loop_done; });
            if (cert_queue.empty()) return;
            r = cert_queue.front();
            cert_queue.pop_front();
        }
        std::string L50, why;
        double dil = 0;
        const int rc = export_certify(r->spec.n, r->q.data(), &dil, L50, why);
        std::lock_guard<std::mutex> lk(mu);
        r->certified = rc == 0;
        r->L50 = L50;
        r->cert_reason = why;
        r->dilation = dil;
        r->t_cert = now_s();
        r->done = true;
    }
}

}  // namespace sis

using namespace sis;

struct sis_server {
    Server s;
};

extern "C" {

sis_server* sis_server_create(const sis_server_cfg* cfg) {
    if (!cfg || cfg->max_instances < 1 || cfg->max_instances > 100000 || cfg->intervals_per_round < 1 ||
        cfg->certifier_threads < 0 || cfg->certifier_threads > 256 || !(cfg->hbm_budget_bytes >= 0) ||
        !(cfg->request_timeout_s >= 0 && cfg->request_timeout_s < 1e9)) {
        fail(SIS_EINVAL, "bad server configuration");
        return nullptr;
    }
    auto* h = new sis_server;
    h->s.cfg = *cfg;
    if (h->s.init() != SIS_OK) {
        fail(SIS_EDEVICE, h->s.loop_err.empty() ? "server device set-up failed" : h->s.loop_err);
        h->s.release_dev();
        delete h;
        return nullptr;
    }
    h->s.loop = std::thread([h] { h->s.run(); });
    const int nc = cfg->certify ? std::max(1, cfg->certifier_threads) : 0;
    for (int k = 0; k < nc; ++k) h->s.certifiers.emplace_back([h] { h->s.certify_worker(); });
    return h;
}

int sis_submit(sis_server* h, const sis_instance* inst, const sis_params* params, uint64_t* ticket) {
    if (!h) return fail(SIS_EINVAL, "null server");
    return sis_submit_timeout(h, inst, params, h->s.cfg.request_timeout_s, ticket);
}

int sis_submit_timeout(sis_server* h, const sis_instance* inst, const sis_params* params, double timeout_s,
                       uint64_t* ticket) {
    if (!h || !ticket) return fail(SIS_EINVAL, "null server or ticket");
    if (!(timeout_s >= 0 && timeout_s < 1e9)) return fail(SIS_EINVAL, "timeout_s must be in [0, 1e9)");
    int rc = validate(inst, params);
    if (rc) return rc;
    auto* r = new Req;
    r->spec = make_spec(inst, params);
    if ((rc = plan_instance(r->spec, params->replicas, r->plan))) {
        delete r;
        return rc;
    }
    if ((size_t)inst->n * params->replicas > h->s.cap_sq || (size_t)params->replicas > h->s.cap_rep) {
        delete r;
        return fail(SIS_EINVAL, "request larger than the server arena");
    }
    r->bytes = request_bytes(inst->n, params->replicas);
    initial_state(r->spec, r->init);  // in the caller's thread, off the device loop
    r->t_submit = now_s();
    r->deadline = timeout_s > 0 ? r->t_submit + timeout_s : 0.0;
    {
        std::lock_guard<std::mutex> lk(h->s.mu);
        if (h->s.stop) {
            delete r;
            return fail(SIS_EDEVICE, h->s.error.empty() ? "server stopped" : h->s.error);
        }
        r->ticket = h->s.next_ticket++;
        const int cls = size_class(inst->n);
        r->lane = lane_of(cls, h->s.class_seq[cls]++);
        h->s.all[r->ticket] = r;
        h->s.queue[r->lane].push_back(r);
        *ticket = r->ticket;
    }
    h->s.cv.notify_all();
    return SIS_OK;
}

int sis_poll(sis_server* h, uint64_t ticket, sis_result* out) {
    if (!h || !out) return fail(SIS_EINVAL, "null server or result");
    std::lock_guard<std::mutex> lk(h->s.mu);
    auto it = h->s.all.find(ticket);
    if (it == h->s.all.end()) return fail(SIS_EINVAL, "unknown ticket");
    Req* r = it->second;
    if (!r->done) {
        if (h->s.stop && !h->s.error.empty()) return fail(SIS_EDEVICE, h->s.error);
        return SIS_PENDING;
    }
    if (r->fail_code) {
        out->status = r->fail_code;
        return fail(r->fail_code, r->fail_msg);
    }
    double* q = out->q_out;
    *out = r->res;
    out->q_out = q;
    if (q) std::copy(r->q.begin(), r->q.end(), q);
    if (h->s.cfg.certify && !r->certified) out->status = SIS_EINVAL;
    return SIS_OK;
}

int sis_server_stats(sis_server* h, double* arena_bytes, double* resident_bytes, double* peak_resident_bytes,
                     int* resident_requests) {
    if (!h) return fail(SIS_EINVAL, "null server");
    std::lock_guard<std::mutex> lk(h->s.mu);
    if (arena_bytes) *arena_bytes = (double)h->s.arena_bytes;
    if (resident_bytes) *resident_bytes = (double)h->s.resident_bytes;
    if (peak_resident_bytes) *peak_resident_bytes = (double)h->s.peak_resident;
    if (resident_requests) *resident_requests = h->s.resident_reqs;
    return SIS_OK;
}

int sis_server_inject_lane_fault(sis_server* h, int lane) {
    if (!h || lane < 0 || lane >= kLanes) return fail(SIS_EINVAL, "bad server or lane");
    std::lock_guard<std::mutex> lk(h->s.mu);
    h->s.fault_lane = lane;
    return SIS_OK;
}

int sis_ticket_times(sis_server* h, uint64_t ticket, double* queued_s, double* done_s, double* certified_s,
                     char* L50, int L50_len) {
    if (!h) return fail(SIS_EINVAL, "null server");
    std::lock_guard<std::mutex> lk(h->s.mu);
    auto it = h->s.all.find(ticket);
    if (it == h->s.all.end()) return fail(SIS_EINVAL, "unknown ticket");
    Req* r = it->second;
    if (queued_s) *queued_s = r->t_admit < 0 ? -1 : r->t_admit - r->t_submit;
    if (done_s) *done_s = r->t_done < 0 ? -1 : r->t_done - r->t_submit;
    if (certified_s) *certified_s = r->t_cert < 0 ? -1 : r->t_cert - r->t_submit;
    if (L50 && L50_len > 0) {
        const std::string& s = r->certified ? r->L50 : r->cert_reason;
        std::strncpy(L50, s.c_str(), (size_t)L50_len - 1);
        L50[L50_len - 1] = 0;
    }
    return r->done ? SIS_OK : SIS_PENDING;
}

int sis_ticket_release(sis_server* h, uint64_t ticket) {
    if (!h) return fail(SIS_EINVAL, "null server");
    std::lock_guard<std::mutex> lk(h->s.mu);
    auto it = h->s.all.find(ticket);
    if (it == h->s.all.end()) return fail(SIS_EINVAL, "unknown ticket");
    // done is set once the device loop and the certifier are through with it
    if (!it->second->done) return SIS_PENDING;
    delete it->second;
    h->s.all.erase(it);
    return SIS_OK;
}

int sis_server_destroy(sis_server* h) {
    if (!h) return SIS_OK;
    {
        std::lock_guard<std::mutex> lk(h->s.mu);
        h->s.stop = true;
    }
    h->s.cv.notify_all();
    if (h->s.loop.joinable()) h->s.loop.join();
    {
        std::lock_guard<std::mutex> lk(h->s.mu);
        h->s.loop_done = true;
    }
    h->s.cert_cv.notify_all();
    for (auto& t : h->s.certifiers) t.join();
    h->s.release_dev();
    for (auto& kv : h->s.all) delete kv.second;
    delete h;
    return SIS_OK;
}

}  // extern "C"
