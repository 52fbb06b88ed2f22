// C ABI parity entry points: run the hot-path kernels on caller data so the
// tests can compare them with the CPU oracle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "../../include/sis_abi.h"
#include "sis_host.h"

using namespace sis;

extern "C" {

int sis_dev_depth32(int npairs, const float* in, float* depth, float* depth2) {
    if (npairs < 0 || (npairs > 0 && (!in || !depth || !depth2))) return fail(SIS_EINVAL, "bad arguments");
    if (npairs == 0) return SIS_OK;
    int rc = ensure_device(-1);
    if (rc) return rc;
    DevBuf<float> d_in, d_o, d_o2;
    SIS_HIPCHECK(d_in.alloc(8 * (size_t)npairs));
    SIS_HIPCHECK(d_o.alloc(npairs));
    SIS_HIPCHECK(d_o2.alloc(npairs));
    SIS_HIPCHECK(hipMemcpy(d_in, in, sizeof(float) * 8 * (size_t)npairs, hipMemcpyHostToDevice));
    SIS_HIPCHECK(launch_depth32(nullptr, npairs, d_in, d_o, d_o2));
    SIS_HIPCHECK(hipMemcpy(depth, d_o, sizeof(float) * npairs, hipMemcpyDeviceToHost));
    SIS_HIPCHECK(hipMemcpy(depth2, d_o2, sizeof(float) * npairs, hipMemcpyDeviceToHost));
    return SIS_OK;
}

int sis_dev_cellcoord(int count, const float* x, float shift, float w, int lo, int hi, int* out) {
    if (count < 0 || (count > 0 && (!x || !out)) || !(w > 0) || lo > hi)
        return fail(SIS_EINVAL, "bad arguments");
    if (count == 0) return SIS_OK;
    int rc = ensure_device(-1);
    if (rc) return rc;
    DevBuf<float> d_x;
    DevBuf<int> d_o;
    SIS_HIPCHECK(d_x.alloc(count));
    SIS_HIPCHECK(d_o.alloc(count));
    SIS_HIPCHECK(hipMemcpy(d_x, x, sizeof(float) * count, hipMemcpyHostToDevice));
    SIS_HIPCHECK(launch_cellcoord(nullptr, count, d_x, shift, w, lo, hi, d_o));
    SIS_HIPCHECK(hipMemcpy(out, d_o, sizeof(int) * count, hipMemcpyDeviceToHost));
    return SIS_OK;
}

int sis_dev_accept_exp2(int count, const float* x, float* out) {
    if (count < 0 || (count > 0 && (!x || !out))) return fail(SIS_EINVAL, "bad arguments");
    if (count == 0) return SIS_OK;
    int rc = ensure_device(-1);
    if (rc) return rc;
    DevBuf<float> d_x, d_o;
    SIS_HIPCHECK(d_x.alloc(count));
    SIS_HIPCHECK(d_o.alloc(count));
    SIS_HIPCHECK(hipMemcpy(d_x, x, sizeof(float) * count, hipMemcpyHostToDevice));
    SIS_HIPCHECK(launch_accept_exp2(nullptr, count, d_x, d_o));
    SIS_HIPCHECK(hipMemcpy(out, d_o, sizeof(float) * count, hipMemcpyDeviceToHost));
    return SIS_OK;
}

int sis_dev_sweeps(int n, int R, float L, float cell, int group, const float* inv_t_log2e,
                   const float* step_xy, const float* step_th, uint32_t key0, uint32_t key1,
                   uint32_t sweep0, int sweeps, int sweeps_per_shift, float* xycs, int* accepted,
                   float* energy, uint64_t counts[3]) {
    uint64_t c6[6] = {0, 0, 0, 0, 0, 0};
    const int rc = sis_dev_epoch(n, R, L, cell, group, inv_t_log2e, step_xy, step_th, key0, key1, sweep0,
                                 sweeps, sweeps_per_shift, 0, 0, 0u, xycs, accepted, energy, c6);
    if (counts)
        for (int k = 0; k < 3; ++k) counts[k] = c6[k];
    return rc;
}

int sis_dev_epoch(int n, int R, float L, float cell, int group, const float* inv_t_log2e,
                  const float* step_xy, const float* step_th, uint32_t key0, uint32_t key1,
                  uint32_t sweep0, int sweeps, int sweeps_per_shift, int collective, int kick_every,
                  uint32_t interval, float* xycs, int* accepted, float* energy, uint64_t counts[6]) {
    if (R < 1 || R > 4096) return fail(SIS_EINVAL, "bad arguments");
    std::vector<float> box(R, L);
    return sis_dev_launch(n, R, L, box.data(), cell, group, inv_t_log2e, step_xy, step_th, key0, key1, sweep0,
                          sweeps, sweeps_per_shift, collective, kick_every, interval, 0, xycs, accepted, energy,
                          counts, nullptr);
}

int sis_dev_launch(int n, int R, float Lgrid, const float* box, float cell, int group, const float* inv_t_log2e,
                   const float* step_xy, const float* step_th, uint32_t key0, uint32_t key1, uint32_t sweep0,
                   int sweeps, int sweeps_per_shift, int collective, int kick_every, uint32_t interval,
                   int certify, float* xycs, int* accepted, float* energy, uint64_t counts[6],
                   sis_rep_state* reps_out) {
    if (n < 1 || n > kMaxN || R < 1 || R > 4096 || !(Lgrid > 0 && Lgrid <= 64) || !(cell >= kMinCell) ||
        sweeps < 1 || sweeps_per_shift < 1 || sweeps_per_shift > 100000 || group < 0 || collective < 0 ||
        collective > 64 || kick_every < 0 || !inv_t_log2e || !step_xy || !step_th || !xycs || !box ||
        (certify != 0 && certify != 1))
        return fail(SIS_EINVAL, "bad arguments");
    for (int r = 0; r < R; ++r)
        if (!(box[r] > 0 && box[r] <= Lgrid)) return fail(SIS_EINVAL, "every box must lie in (0, Lgrid]");
    int rc = ensure_device(-1);
    if (rc) return rc;
    const int M = grid_cells(Lgrid, cell) + 4;
    if (M > 255) return fail(SIS_EINVAL, "cell grid too wide");
    int G = group > 0 ? std::min(group, R) : 1;
    const size_t hard = device_info().lds_per_block;
    // one replica per tile stages proposals up to kMaxN squares, as a launch
    // with compute units to spare does
    const bool wide = G == 1;
    while (G > 1 && sweep_lds_bytes(n, M, G, false) > hard) --G;
    const size_t lds = sweep_lds_bytes(n, M, G, wide);
    if (lds > hard) return fail(SIS_EINVAL, "replica does not fit LDS");
    for (int r = 0; r < R; ++r)
        for (int i = 0; i < n; ++i) {
            const float* v = xycs + 4 * ((size_t)r * n + i);
            if (!(v[0] >= 0.0f && v[0] <= Lgrid && v[1] >= 0.0f && v[1] <= Lgrid))
                return fail(SIS_EINVAL, "centres must lie in [0, Lgrid]");
        }
    InstDesc D{};
    D.n = n; D.R = R; D.ladder = R; D.G = G; D.state_off = 0; D.rep_off = 0; D.M = M;
    D.spe = sweeps; D.spsh = sweeps_per_shift;
    D.sweeps_total = (long long)sweep0 + sweeps; D.sweeps_done = sweep0;
    D.L = Lgrid; D.L_arch = Lgrid; D.Lf = Lgrid; D.scale = 1.0f; D.w = cell; D.key0 = key0; D.key1 = key1;
    D.collective = collective; D.kick_every = kick_every; D.interval = interval; D.improve = certify;
    D.certify_every = 1;
    D.Lmin = std::sqrt((double)n);
    std::vector<Tile> tiles;
    for (int r0 = 0; r0 < R; r0 += G) tiles.push_back(Tile{0, r0, std::min(G, R - r0), wide ? 1 : 0});
    std::vector<RepState> reps(R);
    std::vector<SlotState> slots(R);
    for (int r = 0; r < R; ++r) {
        reps[r] = RepState{0.0f, r, 0u, 0u, kCertNone, 1.0, 0.0, 0.0};
        slots[r] = SlotState{};
        slots[r].inv_t_log2e = inv_t_log2e[r];
        slots[r].step_xy = step_xy[r];
        slots[r].step_th = step_th[r];
        slots[r].L = box[r];
        slots[r].T = 1.0f;
    }
    DevBuf<InstDesc> d_inst;
    DevBuf<Tile> d_tiles;
    DevBuf<RepState> d_reps;
    DevBuf<SlotState> d_slots;
    DevBuf<float4> d_state;
    SIS_HIPCHECK(d_inst.alloc(1));
    SIS_HIPCHECK(d_tiles.alloc(tiles.size()));
    SIS_HIPCHECK(d_reps.alloc(R));
    SIS_HIPCHECK(d_slots.alloc(R));
    SIS_HIPCHECK(d_state.alloc((size_t)n * R));
    SIS_HIPCHECK(hipMemcpy(d_inst, &D, sizeof(InstDesc), hipMemcpyHostToDevice));
    SIS_HIPCHECK(hipMemcpy(d_tiles, tiles.data(), sizeof(Tile) * tiles.size(), hipMemcpyHostToDevice));
    SIS_HIPCHECK(hipMemcpy(d_reps, reps.data(), sizeof(RepState) * R, hipMemcpyHostToDevice));
    SIS_HIPCHECK(hipMemcpy(d_slots, slots.data(), sizeof(SlotState) * R, hipMemcpyHostToDevice));
    SIS_HIPCHECK(hipMemcpy(d_state, xycs, sizeof(float4) * (size_t)n * R, hipMemcpyHostToDevice));
    SIS_HIPCHECK(launch_sweep(nullptr, (int)tiles.size(), lds, d_tiles, d_inst, d_reps, d_slots, d_state,
                              sweep_threads(n), true));
    SIS_HIPCHECK(hipDeviceSynchronize());
    SIS_HIPCHECK(hipMemcpy(xycs, d_state, sizeof(float4) * (size_t)n * R, hipMemcpyDeviceToHost));
    SIS_HIPCHECK(hipMemcpy(reps.data(), d_reps, sizeof(RepState) * R, hipMemcpyDeviceToHost));
    SIS_HIPCHECK(hipMemcpy(&D, d_inst, sizeof(InstDesc), hipMemcpyDeviceToHost));
    for (int r = 0; r < R; ++r) {
        if (accepted) accepted[r] = (int)reps[r].acc;
        if (energy) energy[r] = reps[r].E;
        if (reps_out) reps_out[r] = reps[r];
    }
    if (counts) {
        // a launch without the exchange step leaves its counters in the
        // replica records (each tile's in its first replica)
        counts[0] = (uint64_t)n * (uint64_t)R * (uint64_t)sweeps;
        for (int k = 1; k < 6; ++k) counts[k] = 0;
        for (int r = 0; r < R; ++r) {
            counts[1] += reps[r].cand;
            counts[2] += reps[r].near_pairs;
            counts[3] += reps[r].coll_tried;
            counts[4] += reps[r].coll_acc;
            counts[5] += reps[r].kicks;
        }
    }
    return SIS_OK;
}

// Descriptor <-> the probe's view of it.
static void desc_to_ladder(const InstDesc& D, sis_ladder_desc* d) {
    d->n = D.n; d->R = D.R; d->ladder = D.ladder; d->improve = D.improve; d->cool_period = D.cool_period;
    d->spe = D.spe; d->sweeps_total = D.sweeps_total; d->sweeps_done = D.sweeps_done;
    d->key0 = D.key0; d->key1 = D.key1; d->interval = D.interval;
    d->shrink = D.shrink; d->target_acc = D.target_acc; d->scale = D.scale; d->Lf = D.Lf;
    d->L = D.L; d->L_arch = D.L_arch; d->L_box_arch = D.L_box_arch;
    d->arch_valid = D.arch_valid; d->shrinks = D.shrinks; d->feasible = D.feasible; d->fp64_rejects = D.fp64_rejects;
    d->ex_tried = D.ex_tried; d->ex_acc = D.ex_acc; d->certs = D.certs;
}

int sis_dev_ladder(sis_ladder_desc* d, sis_rep_state* reps, sis_slot_state* slots, int* rep_of_slot,
                   float* xycs, float* arch) {
    if (!d || !reps || !slots || !rep_of_slot || !xycs || !arch) return fail(SIS_EINVAL, "null argument");
    const int n = d->n, R = d->R, lad = d->ladder;
    if (n < 1 || n > kMaxN || R < 1 || R > 4096 || lad < 1 || R % lad != 0 || d->spe < 1 ||
        (d->improve != 0 && d->improve != 1) || d->cool_period < 0 || d->cool_period == 1 ||
        !(d->L > 0 && d->L <= 64) || !(d->L_arch > 0 && d->L_arch <= 64))
        return fail(SIS_EINVAL, "bad ladder descriptor");
    for (int r = 0; r < R; ++r)
        if (reps[r].slot < 0 || reps[r].slot >= R || rep_of_slot[reps[r].slot] != r)
            return fail(SIS_EINVAL, "slot owners must invert the replicas' slots");
    int rc = ensure_device(-1);
    if (rc) return rc;
    InstDesc D{};
    D.n = n; D.R = R; D.ladder = lad; D.G = 1; D.state_off = 0; D.rep_off = 0; D.arch_off = 0;
    D.spe = d->spe; D.sweeps_total = d->sweeps_total; D.sweeps_done = d->sweeps_done;
    D.key0 = d->key0; D.key1 = d->key1; D.interval = d->interval;
    D.shrink = d->shrink; D.target_acc = d->target_acc; D.scale = d->scale; D.Lf = d->Lf;
    D.L = d->L; D.L_arch = d->L_arch; D.L_box_arch = d->L_box_arch; D.Lmin = std::sqrt((double)n);
    D.arch_valid = d->arch_valid; D.shrinks = d->shrinks; D.feasible = d->feasible; D.fp64_rejects = d->fp64_rejects;
    D.ex_tried = d->ex_tried; D.ex_acc = d->ex_acc; D.certs = d->certs;
    D.improve = d->improve; D.cool_period = d->cool_period;
    DevBuf<InstDesc> d_inst;
    DevBuf<RepState> d_reps;
    DevBuf<SlotState> d_slots;
    DevBuf<int> d_ros;
    DevBuf<float4> d_state, d_arch;
    SIS_HIPCHECK(d_inst.alloc(1));
    SIS_HIPCHECK(d_reps.alloc(R));
    SIS_HIPCHECK(d_slots.alloc(R));
    SIS_HIPCHECK(d_ros.alloc(R));
    SIS_HIPCHECK(d_state.alloc((size_t)n * R));
    SIS_HIPCHECK(d_arch.alloc(n));
    SIS_HIPCHECK(hipMemcpy(d_inst, &D, sizeof(InstDesc), hipMemcpyHostToDevice));
    SIS_HIPCHECK(hipMemcpy(d_reps, reps, sizeof(RepState) * R, hipMemcpyHostToDevice));
    SIS_HIPCHECK(hipMemcpy(d_slots, slots, sizeof(SlotState) * R, hipMemcpyHostToDevice));
    SIS_HIPCHECK(hipMemcpy(d_ros, rep_of_slot, sizeof(int) * R, hipMemcpyHostToDevice));
    SIS_HIPCHECK(hipMemcpy(d_state, xycs, sizeof(float4) * (size_t)n * R, hipMemcpyHostToDevice));
    SIS_HIPCHECK(hipMemcpy(d_arch, arch, sizeof(float4) * n, hipMemcpyHostToDevice));
    SIS_HIPCHECK(launch_ladder(nullptr, 1, ladder_lds_bytes(n, D.L, R), d_inst, d_reps, d_slots, d_ros, d_state,
                               d_arch));
    SIS_HIPCHECK(hipDeviceSynchronize());
    SIS_HIPCHECK(hipMemcpy(&D, d_inst, sizeof(InstDesc), hipMemcpyDeviceToHost));
    SIS_HIPCHECK(hipMemcpy(reps, d_reps, sizeof(RepState) * R, hipMemcpyDeviceToHost));
    SIS_HIPCHECK(hipMemcpy(slots, d_slots, sizeof(SlotState) * R, hipMemcpyDeviceToHost));
    SIS_HIPCHECK(hipMemcpy(rep_of_slot, d_ros, sizeof(int) * R, hipMemcpyDeviceToHost));
    SIS_HIPCHECK(hipMemcpy(xycs, d_state, sizeof(float4) * (size_t)n * R, hipMemcpyDeviceToHost));
    SIS_HIPCHECK(hipMemcpy(arch, d_arch, sizeof(float4) * n, hipMemcpyDeviceToHost));
    desc_to_ladder(D, d);
    return SIS_OK;
}

int sis_dev_zero64(int n, double L, const float* xycs, int* zero, double* e64) {
    if (n < 1 || n > kMaxN || !(L > 0 && L <= 64) || !xycs || !zero || !e64) return fail(SIS_EINVAL, "bad arguments");
    int rc = ensure_device(-1);
    if (rc) return rc;
    DevBuf<float4> d_q;
    DevBuf<int> d_z;
    DevBuf<double> d_e;
    SIS_HIPCHECK(d_q.alloc(n));
    SIS_HIPCHECK(d_z.alloc(1));
    SIS_HIPCHECK(d_e.alloc(1));
    SIS_HIPCHECK(hipMemcpy(d_q, xycs, sizeof(float4) * n, hipMemcpyHostToDevice));
    SIS_HIPCHECK(launch_zero64(n, L, d_q, d_z, d_e));
    SIS_HIPCHECK(hipDeviceSynchronize());
    SIS_HIPCHECK(hipMemcpy(zero, d_z, sizeof(int), hipMemcpyDeviceToHost));
    SIS_HIPCHECK(hipMemcpy(e64, d_e, sizeof(double), hipMemcpyDeviceToHost));
    return SIS_OK;
}

int sis_dev_polish(int n, double L, int iters, double* q, double* energy_before,
                   double* energy_after, double* lam, double* L_cert, double* q_export) {
    if (n < 1 || n > kMaxN || !(L > 0 && L <= 64) || iters < 0 || !q) return fail(SIS_EINVAL, "bad arguments");
    int rc = ensure_device(-1);
    if (rc) return rc;
    DevBuf<double> d_q, d_out;
    DevBuf<FinalizeJob> d_job;
    DevBuf<FinalizeOut> d_fo;
    SIS_HIPCHECK(d_q.alloc(3 * (size_t)n));
    SIS_HIPCHECK(d_out.alloc(3 * (size_t)n));
    SIS_HIPCHECK(d_job.alloc(1));
    SIS_HIPCHECK(d_fo.alloc(1));
    SIS_HIPCHECK(hipMemcpy(d_q, q, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    FinalizeJob j{};
    j.n = n; j.iters = iters; j.L = L; j.src32 = nullptr; j.src64 = d_q; j.out_off = 0;
    SIS_HIPCHECK(hipMemcpy(d_job, &j, sizeof(j), hipMemcpyHostToDevice));
    SIS_HIPCHECK(launch_finalize(nullptr, 1, d_job, d_fo, d_out));
    SIS_HIPCHECK(hipDeviceSynchronize());
    FinalizeOut fo;
    SIS_HIPCHECK(hipMemcpy(&fo, d_fo, sizeof(fo), hipMemcpyDeviceToHost));
    if (q_export) SIS_HIPCHECK(hipMemcpy(q_export, d_out, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    if (energy_before) *energy_before = fo.e64_before;
    if (energy_after) *energy_after = fo.e64_after;
    if (lam) *lam = fo.lambda;
    if (L_cert) *L_cert = fo.L_cert;
    return SIS_OK;
}

}  // extern "C"
