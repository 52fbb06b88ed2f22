// Allocation-free plain solve (sis_solve_noalloc): the plain form's kernels
// and records, with every device buffer carved from one file-scope device
// array instead of hipMalloc.  For callers that cannot allocate device memory.
//
// The array `scratch` exists once per device the library runs on and lives
// until the process ends.  Host code never passes its name to a kernel or a
// copy: the device address is taken with hipGetSymbolAddress when the
// library sets a device up (ensure_device) and kept per device.  Calls on one
// device share the array and would overwrite each other's state; the library
// serialises them with a per-device lock.  Every byte a call reads is written
// by that call first (records, states and descriptor are copied in whole), so
// nothing left by a previous call needs clearing.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sis_abi.h"
#include "sis_host.h"
#include "sis_internal.h"
#include "sis_kernels.h"

namespace sis {

constexpr size_t kScratchBytes = (size_t)32 << 20;
constexpr int kScratchDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu[kScratchDevices];
static unsigned char* g_scratch_ptr[kScratchDevices] = {};

// Called by ensure_device once per device, with that device current.
int noalloc_setup(int device) {
    if (device < 0 || device >= kScratchDevices) return fail(SIS_EINVAL, "device index beyond the scratch table");
    if (g_scratch_ptr[device]) return SIS_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(SIS_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<unsigned char*>(p);
    return SIS_OK;
}

namespace {

// Byte layout of one solve inside the scratch array (256-byte aligned parts).
struct ScratchLayout {
    size_t desc, tiles, reps, slots, ros, state, arch, job, fout, qout, total;
};

ScratchLayout scratch_layout(int n, int R, int ntiles) {
    ScratchLayout l{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off += (bytes + 255) & ~(size_t)255;
        return at;
    };
    l.desc = take(sizeof(InstDesc));
    l.tiles = take(sizeof(Tile) * (size_t)ntiles);
    l.reps = take(sizeof(RepState) * (size_t)R);
    l.slots = take(sizeof(SlotState) * (size_t)R);
    l.ros = take(sizeof(int) * (size_t)R);
    l.state = take(sizeof(float4) * (size_t)n * R);
    l.arch = take(sizeof(float4) * (size_t)n);
    l.job = take(sizeof(FinalizeJob));
    l.fout = take(sizeof(FinalizeOut));
    l.qout = take(sizeof(double) * 3 * (size_t)n);
    l.total = off;
    return l;
}

#define NA_CHECK(x)                                                                        \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            return fail(SIS_EDEVICE, std::string(#x) + ": " + hipGetErrorString(e_));     \
    } while (0)

int solve_noalloc(const sis_instance* inst, const sis_params* params, sis_result* out) {
    const auto t0 = std::chrono::steady_clock::now();
    const int dev = device_info().id;
    if (dev < 0 || dev >= kScratchDevices || !g_scratch_ptr[dev])
        return fail(SIS_EDEVICE, "the scratch array of this device is not set up");
    const InstSpec spec = make_spec(inst, params);
    Plan pl;
    if (int rc = plan_instance(spec, spec.p.replicas, pl)) return rc;
    const int n = spec.n, R = spec.p.replicas;
    const ScratchLayout lay = scratch_layout(n, R, pl.ntiles);
    if (lay.total > kScratchBytes)
        return fail(SIS_EINVAL, "n x replicas too large for the allocation-free solve: needs " +
                                    std::to_string(lay.total) + " bytes of its " + std::to_string(kScratchBytes) +
                                    "-byte device buffer (about 2 million squares at most)");
    std::lock_guard<std::mutex> lk(g_scratch_mu[dev]);
    unsigned char* base = g_scratch_ptr[dev];
    InstDesc* d_inst = reinterpret_cast<InstDesc*>(base + lay.desc);
    Tile* d_tiles = reinterpret_cast<Tile*>(base + lay.tiles);
    RepState* d_reps = reinterpret_cast<RepState*>(base + lay.reps);
    SlotState* d_slots = reinterpret_cast<SlotState*>(base + lay.slots);
    int* d_ros = reinterpret_cast<int*>(base + lay.ros);
    float4* d_state = reinterpret_cast<float4*>(base + lay.state);
    float4* d_arch = reinterpret_cast<float4*>(base + lay.arch);
    FinalizeJob* d_job = reinterpret_cast<FinalizeJob*>(base + lay.job);
    FinalizeOut* d_fout = reinterpret_cast<FinalizeOut*>(base + lay.fout);
    double* d_qout = reinterpret_cast<double*>(base + lay.qout);

    // host records: the same builders as the plain form
    InstDesc D{};
    std::vector<RepState> reps;
    std::vector<SlotState> slots;
    std::vector<int> ros;
    instance_records(spec, pl, D, reps, slots, ros);
    D.state_off = 0;
    D.rep_off = 0;
    D.arch_off = 0;
    std::vector<float4> st;
    initial_state(spec, st);
    std::vector<Tile> tiles;
    for (int r0 = 0; r0 < R; r0 += pl.G) tiles.push_back(Tile{0, r0, std::min(pl.G, R - r0), pl.wide});
    NA_CHECK(hipMemcpy(d_inst, &D, sizeof(InstDesc), hipMemcpyHostToDevice));
    NA_CHECK(hipMemcpy(d_tiles, tiles.data(), sizeof(Tile) * tiles.size(), hipMemcpyHostToDevice));
    NA_CHECK(hipMemcpy(d_reps, reps.data(), sizeof(RepState) * R, hipMemcpyHostToDevice));
    NA_CHECK(hipMemcpy(d_slots, slots.data(), sizeof(SlotState) * R, hipMemcpyHostToDevice));
    NA_CHECK(hipMemcpy(d_ros, ros.data(), sizeof(int) * R, hipMemcpyHostToDevice));
    NA_CHECK(hipMemcpy(d_state, st.data(), sizeof(float4) * st.size(), hipMemcpyHostToDevice));
    NA_CHECK(hipMemcpy(d_arch, st.data(), sizeof(float4) * n, hipMemcpyHostToDevice));

    // the exchange intervals: sweeps, then the exchange step, on the null stream
    const long long intervals = (spec.p.sweeps + pl.spe - 1) / pl.spe;
    const size_t lds_ladder = ladder_lds_bytes(n, spec.L, R);
    const int threads = sweep_threads(n);
    const auto t1 = std::chrono::steady_clock::now();
    for (long long k = 0; k < intervals; ++k) {
        NA_CHECK(launch_sweep(nullptr, (int)tiles.size(), pl.lds, d_tiles, d_inst, d_reps, d_slots, d_state, threads,
                              (spec.p.flags & SIS_FLAG_NEAR_PAIRS) != 0, pl.wide));
        NA_CHECK(launch_ladder(nullptr, 1, lds_ladder, d_inst, d_reps, d_slots, d_ros, d_state, d_arch));
    }
    NA_CHECK(hipStreamSynchronize(nullptr));
    const double device_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();

    // FP64 polish + certified side of the archive or of the best replica
    NA_CHECK(hipMemcpy(&D, d_inst, sizeof(InstDesc), hipMemcpyDeviceToHost));
    NA_CHECK(hipMemcpy(reps.data(), d_reps, sizeof(RepState) * R, hipMemcpyDeviceToHost));
    FinalizeJob job{};
    FinalRecord rec;
    job.n = n;
    job.iters = spec.p.polish_iters;
    job.out_off = 0;
    if (finalize_source(D, job.L, rec.target_L)) {
        job.src32 = d_arch;
        rec.e32 = D.E_arch;
    } else {
        int best = 0;
        for (int r = 1; r < R; ++r)
            if (reps[r].E < reps[best].E) best = r;
        job.src32 = d_state + (size_t)best * n;
        rec.e32 = reps[best].E;
    }
    NA_CHECK(hipMemcpy(d_job, &job, sizeof(job), hipMemcpyHostToDevice));
    NA_CHECK(launch_finalize(nullptr, 1, d_job, d_fout, d_qout));
    FinalizeOut fo;
    rec.q.resize(3 * (size_t)n);
    NA_CHECK(hipMemcpy(&fo, d_fout, sizeof(fo), hipMemcpyDeviceToHost));
    NA_CHECK(hipMemcpy(rec.q.data(), d_qout, sizeof(double) * rec.q.size(), hipMemcpyDeviceToHost));
    rec.e64_before = fo.e64_before;
    rec.e64_after = fo.e64_after;
    rec.lambda = fo.lambda;
    rec.L_cert = fo.L_cert;
    rec.desc = D;
    Batch timing;  // holds no device memory: only the time the result reports
    timing.device_ms = device_s * 1e3;
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    fill_result(rec, timing, 0, wall, out);
    return SIS_OK;
}

}  // namespace
}  // namespace sis

using namespace sis;

extern "C" int sis_solve_noalloc(const sis_instance* inst, const sis_params* params, sis_result* out) {
    if (!out) return fail(SIS_EINVAL, "null result");
    int rc = validate(inst, params);
    if (rc) return out->status = rc;
    if ((rc = ensure_device(-1))) return out->status = rc;
    rc = solve_noalloc(inst, params, out);
    out->status = rc;
    return rc;
}
