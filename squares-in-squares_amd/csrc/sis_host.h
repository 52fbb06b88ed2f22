// Host-side types of the runtime (resident batches, plans, results).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/sis_abi.h"
#include "sis_internal.h"
#include "sis_kernels.h"

namespace sis {

struct DeviceInfo {
    int id = -1;
    int cus = 0;
    size_t lds_per_block = 0;
    std::string name, arch;
};

struct InstSpec {
    int n = 0;
    double L = 0;
    std::vector<double> q;      // n*3 or empty
    sis_params p{};
};

struct Plan {
    int G = 1, M = 0, spe = 1, ladder = 1, ntiles = 0;
    int wide = 0;               // tiles stage proposals up to kMaxN squares
    size_t lds = 0;
};

struct FinalRecord {
    std::vector<double> q;
    double target_L = 0, L_cert = 0, lambda = 0, e32 = 0, e64_before = 0, e64_after = 0;
    InstDesc desc{};
};

class Batch {
public:
    ~Batch();
    int build(std::vector<InstSpec>&& in);
    int launch_tiles();
    int enqueue_interval();
    int enqueue_chain(int c);
    int run(long long intervals, int mode);
    // The plain schedule times one sweep launch in this many with a HIP
    // event pair: the two event markers around every launch cost about 5 us
    // of a 122 us interval at n = 50 (1.04e9 against 1.08e9 proposals/s with
    // every eighth launch timed, 1.09e9 with none).
    static constexpr int kSweepTimingEvery = 8;
    // the plain loop reads a timed interval's event pair one interval later,
    // before the pair is recorded again
    static_assert(kSweepTimingEvery >= 2, "one event pair per timed interval");
    int run_plain(long long intervals);
    int wait_status(const std::vector<unsigned>& want);
    int run_graph(int per_graph, long long launches);
    int finalize(std::vector<FinalRecord>& rec);
    void release();

    std::vector<InstSpec> specs;
    std::vector<Plan> plans;
    std::vector<InstDesc> h_inst;
    std::vector<Tile> h_tiles;  // tiles of large instances (four wavefronts) first
    int n_big = 0;              // number of those
    size_t lds_big = 0, lds_small = 0, lds_ladder = 0;
    size_t n_sq = 0, n_rep = 0, n_arch = 0;
    long long max_intervals = 0, intervals_done = 0;
    double device_ms = 0, last_ms = 0, sweep_ms = 0;
    long long sweep_launches = 0;
    bool near_pairs = false;  // some instance asks for near-pair counts (SIS_FLAG_NEAR_PAIRS)
    bool alone = false;       // every tile planned for an idle device: the one-wavefront-per-SIMD kernels

    hipStream_t stream = nullptr, stream2 = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evs0 = nullptr, evs1 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_intervals = 0;

    // The batched schedule runs the instances as kGraphChains independent
    // chains of launches (sweeps + exchange per interval), each on its own
    // stream, so that one chain's last workgroups and exchange step overlap
    // the other's sweeps.  Instances are split by work; results do not depend
    // on it.  Two chains: 1.59e10 -> 1.65e10 proposals/s on 512 mixed-n
    // instances; three measured 1.61-1.64e10.
    static constexpr int kGraphChains = 2;
    struct Chain {
        hipStream_t st = nullptr;
        hipEvent_t done = nullptr;
        int tile_off = 0, ntiles = 0, n_big = 0;  // into d_ctiles
        int list_off = 0, ninst = 0;              // into d_clist
        size_t lds_big = 0, lds_small = 0, lds_ladder = 0;
    };
    std::vector<Chain> chains;
    Tile* d_ctiles = nullptr;
    int* d_clist = nullptr;
    hipEvent_t ev_cfork = nullptr;

    InstDesc* d_inst = nullptr;
    InstDesc* h_pin = nullptr;  // pinned host mirror of the descriptors
    unsigned* h_seq = nullptr;  // mapped host words: interval count of each instance
    unsigned* d_seq = nullptr;  // their device address
    Tile* d_tiles = nullptr;
    RepState* d_reps = nullptr;
    SlotState* d_slots = nullptr;
    int* d_ros = nullptr;
    float4* d_state = nullptr;
    float4* d_arch = nullptr;
    FinalizeJob* d_jobs = nullptr;
    FinalizeOut* d_fout = nullptr;
    double* d_qout = nullptr;
};

int fail(int code, const std::string& msg);

#define SIS_HIPCHECK(x)                                                                    \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            return ::sis::fail(SIS_EDEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Device buffer released on scope exit (error paths included).
template <class T>
struct DevBuf {
    T* p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t count) { return hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * (count ? count : 1)); }
    operator T*() const { return p; }
};
int ensure_device(int device);
// Address of the allocation-free solve's device buffer on `device` (current).
int noalloc_setup(int device);
const DeviceInfo& device_info();
void default_params(sis_params* p);
int validate(const sis_instance* in, const sis_params* p);
void initial_state(const InstSpec& s, std::vector<float4>& out);
void portfolio_state(const InstSpec& s, bool jitter, std::vector<double>& q);
// sis_portfolio.cpp
void portfolio_layouts(int n, int R, int ladder, double seed_L, const double* seed, std::vector<double>& out);
double portfolio_jitter(int cls, double u);
int plan_instance(const InstSpec& s, int batch_replicas, Plan& pl);
// Initial device records of one instance (the caller sets the descriptor's
// state, replica and archive offsets).
void instance_records(const InstSpec& s, const Plan& pl, InstDesc& D, std::vector<RepState>& reps,
                      std::vector<SlotState>& slots, std::vector<int>& ros);
double slot_box(const InstSpec& s, int ladder, int slot, double Lb);
bool finalize_source(const InstDesc& D, double& box, double& target_L);
// A launch set of this many replicas keeps every compute unit busy (eight per
// compute unit): below it tiles hold one replica each and stage every size.
bool device_loaded(long long replicas);
InstSpec make_spec(const sis_instance* in, const sis_params* p);
void fill_result(const FinalRecord& r, const Batch& b, int k, double wall_s, sis_result* out);
// the sweep kernel of a workgroup size; near_pairs: the form that counts near pairs
const void* sweep_kernel_ptr(int threads, bool near_pairs = false, bool alone = false);
// Launch order of a tile set: tiles of instances with kBigN squares or more
// (four-wavefront workgroups) first, then the rest, each part heaviest first
// so the dispatcher hands the long tiles out while compute units free up.
// Returns the number of large-instance tiles.
template <class NOf>
int order_tiles(std::vector<Tile>& tiles, NOf n_of) {
    std::stable_sort(tiles.begin(), tiles.end(), [&](const Tile& a, const Tile& b) {
        const int na = n_of(a.inst), nb = n_of(b.inst);
        const bool ba = na >= kBigN, bb = nb >= kBigN;
        if (ba != bb) return ba;
        return (long long)a.g * na > (long long)b.g * nb;
    });
    int big = 0;
    while (big < (int)tiles.size() && n_of(tiles[big].inst) >= kBigN) ++big;
    return big;
}

}  // namespace sis
