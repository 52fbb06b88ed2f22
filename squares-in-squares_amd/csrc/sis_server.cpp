// Continuous-batching request loop.
//
// Replaces the reference's plain request loop (cuda/cpp_loop.py:82-106: one
// `packing` process per n, at most four in flight, admission by free VRAM and
// RAM, cuda/cpp_loop.py:41-45, then the C++ certifier gates,
// cuda/cpp_loop.py:17-28).  Here every admitted request keeps its replicas
// resident in one device arena; each round runs `intervals_per_round`
// sweep + exchange steps over the tiles of all resident requests at once, so
// requests that arrive while others run are admitted as soon as memory frees
// up and their work overlaps on the device.  Finished requests get the FP64
// polish / certified side on the device and then the exact certificate on a
// host thread pool.  A request's arithmetic depends only on the request
// (its seed, n, parameters), never on what else is resident.
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
    double t_submit = 0, t_admit = -1, t_done = -1, t_cert = -1;
    int slot = -1;
    size_t sq_off = 0, rep_off = 0, arch_off = 0;
    size_t bytes = 0;
    bool done = false;       // result (and certificate, if asked) available
    sis_result res{};
    std::vector<double> q;
    std::string L50, cert_reason;
    double dilation = 0;
    bool certified = false;
};

size_t request_bytes(int n, int R) {
    return (size_t)n * R * sizeof(float4) + (size_t)n * sizeof(float4) +
           (size_t)R * (sizeof(RepState) + sizeof(SlotState) + sizeof(int));
}

}  // namespace

struct Server {
    sis_server_cfg cfg{};
    std::mutex mu;
    std::condition_variable cv, cert_cv;
    std::deque<Req*> queue;
    std::vector<Req*> active;
    std::unordered_map<uint64_t, Req*> all;
    std::deque<Req*> cert_queue;
    uint64_t next_ticket = 1;
    bool stop = false;
    bool loop_done = false;  // the device loop has exited (set under mu)
    std::thread loop;
    std::vector<std::thread> certifiers;
    std::string error;
    size_t resident_bytes = 0;

    // device arena
    size_t cap_sq = 0, cap_rep = 0;
    int cap_inst = 0;
    RangeAlloc sq_alloc, rep_alloc, arch_alloc;
    std::vector<int> free_slots;
    hipStream_t stream = nullptr;
    InstDesc* d_inst = nullptr;
    RepState* d_reps = nullptr;
    SlotState* d_slots = nullptr;
    int* d_ros = nullptr;
    float4* d_state = nullptr;
    float4* d_arch = nullptr;
    Tile* d_tiles = nullptr;
    size_t tiles_cap = 0;
    int* d_list = nullptr;
    std::vector<InstDesc> h_inst;

    int init();
    void run();
    void certify_worker();
    int admit_one(Req* r);
    int finish(std::vector<Req*>& done);
    void release_dev();
};

#define SRV_CHECK(x)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            error = std::string(#x) + ": " + hipGetErrorString(e_);                      \
            return SIS_EDEVICE;                                                            \
        }                                                                                  \
    } while (0)

int Server::init() {
    if (ensure_device(-1)) return SIS_EDEVICE;
    cap_inst = std::max(1, cfg.max_instances);
    // arena: enough squares for the admission cap, bounded to 4 GiB of state
    const double budget = cfg.hbm_budget_bytes > 0 ? cfg.hbm_budget_bytes : 4.0 * (1u << 30);
    cap_sq = (size_t)std::min(budget, 4.0 * (1u << 30)) / sizeof(float4);
    cap_rep = std::min<size_t>(cap_sq / 10, (size_t)cap_inst * 4096);
    sq_alloc.reset(cap_sq);
    rep_alloc.reset(cap_rep);
    arch_alloc.reset((size_t)cap_inst * kMaxN);
    for (int s = cap_inst - 1; s >= 0; --s) free_slots.push_back(s);
    h_inst.assign(cap_inst, InstDesc{});
    SRV_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    SRV_CHECK(hipMalloc(&d_inst, sizeof(InstDesc) * cap_inst));
    SRV_CHECK(hipMalloc(&d_reps, sizeof(RepState) * cap_rep));
    SRV_CHECK(hipMalloc(&d_slots, sizeof(SlotState) * cap_rep));
    SRV_CHECK(hipMalloc(&d_ros, sizeof(int) * cap_rep));
    SRV_CHECK(hipMalloc(&d_state, sizeof(float4) * cap_sq));
    SRV_CHECK(hipMalloc(&d_arch, sizeof(float4) * (size_t)cap_inst * kMaxN));
    SRV_CHECK(hipMalloc(&d_list, sizeof(int) * cap_inst));
    return SIS_OK;
}

void Server::release_dev() {
    for (void* p : {(void*)d_inst, (void*)d_reps, (void*)d_slots, (void*)d_ros, (void*)d_state,
                    (void*)d_arch, (void*)d_tiles, (void*)d_list})
        if (p) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
}

// Device state of a newly admitted request (called with the lock released,
// only from the loop thread).
int Server::admit_one(Req* r) {
    const InstSpec& s = r->spec;
    const int n = s.n, R = s.p.replicas;
    InstDesc D{};
    D.n = n; D.R = R; D.ladder = r->plan.ladder; D.G = r->plan.G;
    D.state_off = (int)r->sq_off; D.rep_off = (int)r->rep_off; D.arch_off = (int)r->arch_off;
    D.M = r->plan.M; D.spe = r->plan.spe;
    D.sweeps_total = s.p.sweeps; D.sweeps_done = 0;
    D.L = s.L; D.L_arch = s.L; D.Lf = (float)s.L; D.scale = 1.0f;
    D.w = (float)s.p.cell; D.shrink = (float)s.p.shrink; D.target_acc = (float)s.p.target_accept;
    D.key0 = (uint32_t)s.p.rng_seed; D.key1 = (uint32_t)(s.p.rng_seed >> 32);
    std::vector<float4> st;
    initial_state(s, st);
    std::vector<RepState> reps(R);
    std::vector<SlotState> slots(R);
    std::vector<int> ros(R);
    const int lad = D.ladder;
    for (int k = 0; k < R; ++k) {
        const int pos = k % lad;
        const double T = lad == 1 ? s.p.tmin : s.p.tmin * std::pow(s.p.tmax / s.p.tmin, (double)pos / (lad - 1));
        slots[k] = SlotState{(float)(1.4426950408889634 / T), (float)(std::sqrt(3.0) * s.p.move_sigma),
                             (float)std::min(0.7794229, std::sqrt(3.0) * s.p.rot_sigma), (float)T};
        reps[k] = RepState{0.0f, k, 0u, 0u};
        ros[k] = k;
    }
    h_inst[r->slot] = D;
    SRV_CHECK(hipMemcpyAsync(d_state + r->sq_off, st.data(), sizeof(float4) * st.size(), hipMemcpyHostToDevice, stream));
    SRV_CHECK(hipMemcpyAsync(d_arch + r->arch_off, st.data(), sizeof(float4) * n, hipMemcpyHostToDevice, stream));
    SRV_CHECK(hipMemcpyAsync(d_reps + r->rep_off, reps.data(), sizeof(RepState) * R, hipMemcpyHostToDevice, stream));
    SRV_CHECK(hipMemcpyAsync(d_slots + r->rep_off, slots.data(), sizeof(SlotState) * R, hipMemcpyHostToDevice, stream));
    SRV_CHECK(hipMemcpyAsync(d_ros + r->rep_off, ros.data(), sizeof(int) * R, hipMemcpyHostToDevice, stream));
    SRV_CHECK(hipMemcpyAsync(d_inst + r->slot, &h_inst[r->slot], sizeof(InstDesc), hipMemcpyHostToDevice, stream));
    SRV_CHECK(hipStreamSynchronize(stream));
    return SIS_OK;
}

// FP64 finalisation of finished requests; frees their arena ranges.
int Server::finish(std::vector<Req*>& done) {
    const int k = (int)done.size();
    std::vector<FinalizeJob> jobs(k);
    long long qoff = 0;
    std::vector<double> e32(k, 0.0);
    for (int j = 0; j < k; ++j) {
        Req* r = done[j];
        const InstDesc& D = h_inst[r->slot];
        FinalizeJob& job = jobs[j];
        job.n = D.n;
        job.iters = r->spec.p.polish_iters;
        job.src64 = nullptr;
        job.out_off = qoff;
        qoff += 3ll * D.n;
        if (D.arch_valid) {
            job.src32 = d_arch + D.arch_off;
            job.L = D.L_arch;
        } else {
            std::vector<RepState> reps(D.R);
            SRV_CHECK(hipMemcpyAsync(reps.data(), d_reps + D.rep_off, sizeof(RepState) * D.R, hipMemcpyDeviceToHost, stream));
            SRV_CHECK(hipStreamSynchronize(stream));
            int best = 0;
            for (int q = 1; q < D.R; ++q)
                if (reps[q].E < reps[best].E) best = q;
            job.src32 = d_state + D.state_off + (size_t)best * D.n;
            job.L = D.L;
            e32[j] = reps[best].E;
        }
    }
    DevBuf<FinalizeJob> d_jobs;
    DevBuf<FinalizeOut> d_fo;
    DevBuf<double> d_q;
    SRV_CHECK(d_jobs.alloc(k));
    SRV_CHECK(d_fo.alloc(k));
    SRV_CHECK(d_q.alloc(std::max<long long>(1, qoff)));
    SRV_CHECK(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(FinalizeJob) * k, hipMemcpyHostToDevice, stream));
    SRV_CHECK(launch_finalize(stream, k, d_jobs, d_fo, d_q));
    std::vector<FinalizeOut> fo(k);
    std::vector<double> q(std::max<long long>(1, qoff));
    SRV_CHECK(hipMemcpyAsync(fo.data(), d_fo, sizeof(FinalizeOut) * k, hipMemcpyDeviceToHost, stream));
    SRV_CHECK(hipMemcpyAsync(q.data(), d_q, sizeof(double) * q.size(), hipMemcpyDeviceToHost, stream));
    SRV_CHECK(hipStreamSynchronize(stream));
    const double t = now_s();
    for (int j = 0; j < k; ++j) {
        Req* r = done[j];
        const InstDesc& D = h_inst[r->slot];
        FinalRecord rec;
        rec.q.assign(q.begin() + jobs[j].out_off, q.begin() + jobs[j].out_off + 3ll * D.n);
        rec.target_L = jobs[j].L;
        rec.L_cert = fo[j].L_cert;
        rec.lambda = fo[j].lambda;
        rec.e32 = e32[j];
        rec.e64_before = fo[j].e64_before;
        rec.e64_after = fo[j].e64_after;
        rec.desc = D;
        r->q = rec.q;
        sis_result res{};
        res.q_out = nullptr;
        Batch dummy;
        fill_result(rec, dummy, 0, t - r->t_admit, &res);
        res.device_seconds = 0;
        r->res = res;
        r->t_done = t;
    }
    return SIS_OK;
}

void Server::run() {
    if (hipSetDevice(device_info().id) != hipSuccess) {
        std::lock_guard<std::mutex> lk(mu);
        error = "hipSetDevice failed in the server loop";
        stop = true;
        return;
    }
    std::vector<Tile> tiles;
    std::vector<int> list;
    bool dirty = true;
    size_t max_lds = 0;
    for (;;) {
        std::vector<Req*> admitted;
        {
            std::unique_lock<std::mutex> lk(mu);
            if (active.empty() && queue.empty()) {
                if (stop) break;
                cv.wait_for(lk, std::chrono::milliseconds(2));
                continue;
            }
            // admission: FIFO while a slot, the HBM budget and arena ranges allow
            while (!queue.empty() && !free_slots.empty()) {
                Req* r = queue.front();
                const double budget = cfg.hbm_budget_bytes > 0 ? cfg.hbm_budget_bytes : 1e300;
                if (resident_bytes > 0 && resident_bytes + r->bytes > budget) break;
                size_t so, ro, ao;
                const int n = r->spec.n, R = r->spec.p.replicas;
                if (!sq_alloc.alloc((size_t)n * R, so)) break;
                if (!rep_alloc.alloc(R, ro)) { sq_alloc.release(so, (size_t)n * R); break; }
                if (!arch_alloc.alloc(n, ao)) { sq_alloc.release(so, (size_t)n * R); rep_alloc.release(ro, R); break; }
                queue.pop_front();
                r->sq_off = so; r->rep_off = ro; r->arch_off = ao;
                r->slot = free_slots.back();
                free_slots.pop_back();
                r->t_admit = now_s();
                resident_bytes += r->bytes;
                admitted.push_back(r);
            }
        }
        for (Req* r : admitted) {
            if (admit_one(r) != SIS_OK) {
                std::lock_guard<std::mutex> lk(mu);
                stop = true;
                return;
            }
            active.push_back(r);
            dirty = true;
        }
        if (active.empty()) continue;
        if (dirty) {
            tiles.clear();
            list.clear();
            max_lds = 0;
            long long total_reps = 0;
            for (Req* r : active) total_reps += r->spec.p.replicas;
            for (Req* r : active) {
                Plan pl;
                plan_instance(r->spec, (int)std::min<long long>(total_reps, 1 << 30), pl);
                r->plan.G = pl.G;  // grouping never changes a replica's arithmetic
                h_inst[r->slot].G = pl.G;
                max_lds = std::max(max_lds, pl.lds);
                for (int r0 = 0; r0 < r->spec.p.replicas; r0 += pl.G)
                    tiles.push_back(Tile{r->slot, r0, std::min(pl.G, r->spec.p.replicas - r0), 0});
                list.push_back(r->slot);
            }
            std::stable_sort(tiles.begin(), tiles.end(), [&](const Tile& a, const Tile& b) {
                return (long long)a.g * h_inst[a.inst].n > (long long)b.g * h_inst[b.inst].n;
            });
            if (tiles.size() > tiles_cap) {
                if (d_tiles) (void)hipFree(d_tiles);
                tiles_cap = tiles.size() * 2;
                if (hipMalloc(&d_tiles, sizeof(Tile) * tiles_cap) != hipSuccess) {
                    std::lock_guard<std::mutex> lk(mu);
                    error = "tile allocation failed";
                    stop = true;
                    return;
                }
            }
            (void)hipMemcpyAsync(d_tiles, tiles.data(), sizeof(Tile) * tiles.size(), hipMemcpyHostToDevice, stream);
            (void)hipMemcpyAsync(d_list, list.data(), sizeof(int) * list.size(), hipMemcpyHostToDevice, stream);
            dirty = false;
        }
        const int ipr = std::max(1, cfg.intervals_per_round);
        bool ok = true;
        for (int k = 0; k < ipr && ok; ++k) {
            ok = launch_sweep(stream, (int)tiles.size(), max_lds, d_tiles, d_inst, d_reps, d_slots, d_state) == hipSuccess;
            ok = ok && launch_ladder(stream, (int)list.size(), d_inst, d_reps, d_slots, d_ros, d_state, d_arch,
                                     d_list) == hipSuccess;
        }
        ok = ok && hipMemcpyAsync(h_inst.data(), d_inst, sizeof(InstDesc) * cap_inst, hipMemcpyDeviceToHost,
                                  stream) == hipSuccess;
        ok = ok && hipStreamSynchronize(stream) == hipSuccess;
        if (!ok) {
            std::lock_guard<std::mutex> lk(mu);
            error = std::string("device round failed: ") + hipGetErrorString(hipGetLastError());
            stop = true;
            return;
        }
        std::vector<Req*> done, keep;
        for (Req* r : active) {
            const InstDesc& D = h_inst[r->slot];
            (D.sweeps_done >= D.sweeps_total ? done : keep).push_back(r);
        }
        if (!done.empty()) {
            if (finish(done) != SIS_OK) {
                std::lock_guard<std::mutex> lk(mu);
                stop = true;
                return;
            }
            std::lock_guard<std::mutex> lk(mu);
            for (Req* r : done) {
                const int n = r->spec.n, R = r->spec.p.replicas;
                sq_alloc.release(r->sq_off, (size_t)n * R);
                rep_alloc.release(r->rep_off, R);
                arch_alloc.release(r->arch_off, n);
                free_slots.push_back(r->slot);
                resident_bytes -= r->bytes;
                if (cfg.certify) {
                    cert_queue.push_back(r);
                } else {
                    r->done = true;
                }
            }
            cert_cv.notify_all();
            active.swap(keep);
            dirty = true;
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
        cfg->certifier_threads < 0 || cfg->certifier_threads > 256) {
        fail(SIS_EINVAL, "bad server configuration");
        return nullptr;
    }
    auto* h = new sis_server;
    h->s.cfg = *cfg;
    if (h->s.init() != SIS_OK) {
        fail(SIS_EDEVICE, h->s.error.empty() ? "server device set-up failed" : h->s.error);
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
    if (!h || !ticket) return fail(SIS_EINVAL, "null server or ticket");
    int rc = validate(inst, params);
    if (rc) return rc;
    auto* r = new Req;
    r->spec = make_spec(inst, params);
    if ((rc = plan_instance(r->spec, params->replicas, r->plan))) {
        delete r;
        return rc;
    }
    r->bytes = request_bytes(inst->n, params->replicas);
    r->t_submit = now_s();
    {
        std::lock_guard<std::mutex> lk(h->s.mu);
        if (h->s.stop) {
            delete r;
            return fail(SIS_EDEVICE, h->s.error.empty() ? "server stopped" : h->s.error);
        }
        if ((size_t)inst->n * params->replicas > h->s.cap_sq || (size_t)params->replicas > h->s.cap_rep) {
            delete r;
            return fail(SIS_EINVAL, "request larger than the server arena");
        }
        r->ticket = h->s.next_ticket++;
        h->s.all[r->ticket] = r;
        h->s.queue.push_back(r);
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
    double* q = out->q_out;
    *out = r->res;
    out->q_out = q;
    if (q) std::copy(r->q.begin(), r->q.end(), q);
    if (h->s.cfg.certify && !r->certified) out->status = SIS_EINVAL;
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
