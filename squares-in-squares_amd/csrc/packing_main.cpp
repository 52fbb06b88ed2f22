// `packing seed.txt epochs steps chains rng_seed out.json` on MI355X.
//
// Keeps the reference's process contract (cuda/packing.cu:768-868): the argv
// ranges and exit code 2, the seed text `n L` + n lines `x y angle`, the
// PACK_* environment knobs with their ranges and "invalid NAME" message, and
// the raw JSON fields at precision 17.  The search itself runs through the C
// ABI of libsisamd (colour-checkerboard FP32 sweeps, replica exchange, FP64
// polish and certified side on the GPU).
//
// Mapping of the reference arguments onto this engine:
//   epochs  exchange rounds of the search (0: 50/40/30 by n; >120 capped to 80)
//   steps   sweeps per epoch, one sweep = n proposals (0: 64/48/32 by n)
//   chains  replicas, multiple of 8, in ladders of 8 temperatures
//           (0: enough to give every compute unit two; PACK_AUTOTUNE raises it)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/sis_abi.h"

static double envf(const char* name, double fallback, double low, double high) {
    const char* s = std::getenv(name);
    if (!s) return fallback;
    char* end = nullptr;
    const double v = std::strtod(s, &end);
    if (*end || !std::isfinite(v) || v < low || v > high) {
        std::cerr << "invalid " << name << "\n";
        std::exit(2);
    }
    return v;
}

static void json_escape(std::ostream& o, const char* s) {
    for (; *s; ++s) {
        if (*s == '"' || *s == '\\') o << '\\';
        o << *s;
    }
}

int main(int argc, char** argv) {
    if (argc != 7) {
        std::cerr << "packing seed.txt epochs steps chains random_seed output.json\n";
        return 2;
    }
    int epochs = std::atoi(argv[2]), steps = std::atoi(argv[3]), chains = std::atoi(argv[4]);
    if (epochs < 0 || epochs > 100000 || steps < 0 || steps > 4096 || chains < 0 || chains > 4096 ||
        (chains > 0 && chains % 8))
        return 2;
    std::ifstream in(argv[1]);
    int n = 0;
    double L = 0;
    if (!(in >> n >> L)) return 2;
    if (n < 1 || n > 1000 || !std::isfinite(L) || L <= 0 || L > 64) return 2;
    std::vector<double> seed(3 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        in >> seed[3 * i] >> seed[3 * i + 1] >> seed[3 * i + 2];
        if (!in || !std::isfinite(seed[3 * i]) || !std::isfinite(seed[3 * i + 1]) ||
            !std::isfinite(seed[3 * i + 2]))
            return 2;
    }
    const double tmin = envf("PACK_TMIN", 1e-11, 1e-12, 1.0);
    const double tmax = envf("PACK_TMAX", 1e-4, 1e-12, 1.0);
    const double move = envf("PACK_MOVE", .02, .000001, .3);
    const double rot = envf("PACK_ROTATION", .12, .000001, 1.57);
    if (tmax < tmin) return 2;
    const double shrink = envf("PACK_SHRINK", .001, .000001, .1);
    const int autotune = (int)envf("PACK_AUTOTUNE", 1, 0, 1);
    const int max_chains = (int)envf("PACK_MAX_CHAINS", 768, 8, 4096);
    const unsigned long long rng_seed = std::strtoull(argv[5], nullptr, 10);

    if (sis_init(0) != SIS_OK) {
        std::cerr << sis_last_error() << "\n";
        return 2;
    }
    int cus = 0;
    char name[256] = {0};
    sis_device_info(&cus, name, sizeof(name));
    const int wave_chains = std::max(8, ((2 * cus) / 8) * 8);
    if (chains == 0 || (autotune && chains >= 64)) {
        const int target = chains == 0 ? wave_chains : std::max(chains, std::min(wave_chains, max_chains));
        chains = ((std::min(target, max_chains) + 7) / 8) * 8;
    }
    if (epochs == 0) epochs = n <= 64 ? 50 : n <= 256 ? 40 : 30;
    if (steps == 0) steps = n <= 64 ? 64 : n <= 256 ? 48 : 32;
    const int eff_epochs = epochs > 120 ? std::min(epochs, 80) : epochs;

    sis_params p;
    sis_default_params(&p);
    p.tmin = tmin;
    p.tmax = tmax;
    p.move_sigma = move;
    p.rot_sigma = std::min(rot, 0.45);
    p.target_accept = 0.25;
    p.shrink = shrink;
    p.replicas = chains;
    p.ladder = 8;
    p.sweeps = (int64_t)eff_epochs * steps;
    p.sweeps_per_exchange = std::max(1, std::min(steps, 10));
    // the reference's epoch mixes collective moves into its steps
    // (cuda/packing.cu:564-616) and kicks once per epoch (:696-745)
    p.collective_moves = 4;
    p.kick_every = std::max(1, steps / p.sweeps_per_exchange);
    p.polish_iters = 200;
    p.rng_seed = rng_seed;
    p.init_mode = 2;  // portfolio starts + structured jitter, as cuda/packing.cu:806,821
    sis_instance inst{n, L, seed.data()};
    std::vector<double> q(3 * (size_t)n);
    sis_result r{};
    r.q_out = q.data();
    if (sis_solve(&inst, &p, &r) != SIS_OK) {
        std::cerr << sis_last_error() << "\n";
        return 2;
    }
    // The reference keeps the seed until a certified side beats the seed's L.
    const bool improved = r.feasible && r.L_cert < L;
    const double bestL = improved ? r.L_cert : L;
    const std::vector<double>& best = improved ? q : seed;

    std::ofstream out(argv[6]);
    if (!out) return 2;
    out.precision(17);
    out << "{\"n\":" << n << ",\"target_L\":" << bestL << ",\"seconds\":" << r.seconds
        << ",\"chain_steps\":" << (long long)chains * eff_epochs * steps << ",\"proposals\":" << r.proposals
        << ",\"feasible_checkpoints\":" << std::max(1, r.feasible ? r.shrinks + 1 : 0)
        << ",\"shrinks\":" << (improved ? r.shrinks : 0) << ",\"chains\":" << chains
        << ",\"sm_count\":" << cus << ",\"blocks_per_sm\":" << 2 << ",\"wave_chains\":" << wave_chains
        << ",\"device\":\"";
    json_escape(out, name);
    out << "\",\"q\":[";
    for (int i = 0; i < n; ++i) {
        if (i) out << ",";
        out << "[" << best[3 * i] << "," << best[3 * i + 1] << "," << best[3 * i + 2] << "]";
    }
    out << "]}\n";
    return out ? 0 : 2;
}
