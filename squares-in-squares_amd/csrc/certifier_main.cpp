// `certifier submission.csv [--max-n N] [--threads T] [--roundtrip] [--details out.json]`
//
// The process contract of the reference's certifier (cuda/certifier.cpp:
// 933-1077): CSV `id,n,square_id,x,y,angle` with consecutive ids, exit 2 on a
// malformed file, every n in 1..max-n certified on a thread pool (largest n
// first), score L/sqrt(n) - 1 for an accepted packing and ceil(sqrt n)/sqrt n
// - 1 for a missing or rejected one, the JSON summary on stdout and the
// per-n details on request.  The decisions and the 50-decimal L come from
// this library's exact certifier (sis_certify_strings).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sis_abi.h"

namespace {

constexpr int kMaxN = 1000;

struct Row {
    int sid;
    std::string x, y, a;
};

struct Result {
    double score = 0;
    std::string L;
    std::string reason = "missing from submission";
    bool ok = false;
};

double infeasible_score(int n) { return std::ceil(std::sqrt((double)n)) / std::sqrt((double)n) - 1.0; }

// The float round trip of a CSV writer (cuda/certifier.cpp:914-929): the
// value read as a double, printed with 15, 16 or 17 significant digits
// (the first that reads back the same), ".0" appended when the text has
// neither a point nor an exponent.
std::string roundtrip(const std::string& s) {
    const double v = std::strtod(s.c_str(), nullptr);
    char buf[64];
    for (int p = 15; p <= 17; ++p) {
        std::snprintf(buf, sizeof(buf), "%.*g", p, v);
        if (std::strtod(buf, nullptr) == v) break;
    }
    std::string out(buf);
    if (out.find_first_of(".eE") == std::string::npos) out += ".0";
    return out;
}

Result evaluate(int n, std::vector<Row>& rows) {
    Result res;
    res.score = infeasible_score(n);
    if ((int)rows.size() != n) {
        res.reason = "n=" + std::to_string(n) + ": expected " + std::to_string(n) + " rows, got " +
                     std::to_string(rows.size());
        return res;
    }
    std::sort(rows.begin(), rows.end(), [](const Row& a, const Row& b) { return a.sid < b.sid; });
    for (int i = 0; i < n; ++i)
        if (rows[i].sid != i) {
            res.reason = "n=" + std::to_string(n) + ": square_ids should be 0.." + std::to_string(n - 1);
            return res;
        }
    std::vector<const char*> xs(n), ys(n), as(n);
    for (int i = 0; i < n; ++i) {
        xs[i] = rows[i].x.c_str();
        ys[i] = rows[i].y.c_str();
        as[i] = rows[i].a.c_str();
    }
    char L50[128];
    if (sis_certify_strings(n, xs.data(), ys.data(), as.data(), L50, sizeof(L50)) != SIS_OK) {
        res.reason = sis_last_error();
        return res;
    }
    res.ok = true;
    res.reason = "ok";
    res.L = L50;
    res.score = std::max(0.0, std::strtod(L50, nullptr) / std::sqrt((double)n) - 1.0);
    return res;
}

std::string json_str(const std::string& s) {
    std::string o;
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        o += c;
    }
    return o;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "Usage: certifier <submission.csv> [--max-n N] [--threads T] [--roundtrip] [--details out.json]\n";
        return 2;
    }
    const std::string csv = argv[1];
    int max_n = kMaxN;
    int threads = (int)std::max(1u, std::thread::hardware_concurrency());
    bool rt = false;
    std::string details;
    for (int i = 2; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--max-n" && i + 1 < argc) max_n = std::atoi(argv[++i]);
        else if (a == "--threads" && i + 1 < argc) threads = std::max(1, std::atoi(argv[++i]));
        else if (a == "--roundtrip") rt = true;
        else if (a == "--details" && i + 1 < argc) details = argv[++i];
    }
    max_n = std::min(std::max(max_n, 0), kMaxN);
    const auto t0 = std::chrono::steady_clock::now();
    std::ifstream in(csv);
    if (!in) {
        std::cerr << "Cannot open " << csv << "\n";
        return 2;
    }
    std::string line;
    if (!std::getline(in, line)) {
        std::cerr << "Empty CSV\n";
        return 2;
    }
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line != "id,n,square_id,x,y,angle") {
        std::cerr << "CSV header must be id,n,square_id,x,y,angle\n";
        return 2;
    }
    std::vector<std::vector<Row>> puzzles(max_n + 1);
    long long expect = 0;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty()) continue;
        size_t c[5];
        size_t from = 0;
        for (int k = 0; k < 5; ++k) {
            c[k] = line.find(',', from);
            if (c[k] == std::string::npos) return 2;
            from = c[k] + 1;
        }
        if (line.find(',', from) != std::string::npos) return 2;
        if (std::strtoll(line.c_str(), nullptr, 10) != expect) {
            std::cerr << "Row id mismatch at " << expect << "\n";
            return 2;
        }
        ++expect;
        const int n = std::atoi(line.c_str() + c[0] + 1);
        const int sid = std::atoi(line.c_str() + c[1] + 1);
        if (n < 1 || n > kMaxN) return 2;
        if (n > max_n) continue;
        Row r{sid, line.substr(c[2] + 1, c[3] - c[2] - 1), line.substr(c[3] + 1, c[4] - c[3] - 1),
              line.substr(c[4] + 1)};
        if (rt) {
            r.x = roundtrip(r.x);
            r.y = roundtrip(r.y);
            r.a = roundtrip(r.a);
        }
        puzzles[n].push_back(std::move(r));
    }
    std::vector<Result> res(max_n + 1);
    std::atomic<int> next{max_n};
    std::vector<std::thread> pool;
    for (int k = 0; k < threads; ++k)
        pool.emplace_back([&] {
            for (;;) {
                const int n = next.fetch_sub(1);
                if (n < 1) return;
                if (puzzles[n].empty()) {
                    res[n].score = infeasible_score(n);
                } else {
                    res[n] = evaluate(n, puzzles[n]);
                }
            }
        });
    for (auto& t : pool) t.join();
    double total = 0;
    int accepted = 0, missing = 0, rejected = 0;
    for (int n = 1; n <= max_n; ++n) {
        total += res[n].score;
        if (res[n].ok) ++accepted;
        else if (res[n].reason == "missing from submission") ++missing;
        else ++rejected;
    }
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    auto summary = [&](std::ostream& o) {
        o << std::setprecision(17) << "{\"score\":" << total << ",\"max_n\":" << max_n << ",\"accepted\":" << accepted
          << ",\"missing\":" << missing << ",\"rejected\":" << rejected << ",\"seconds\":" << sec
          << ",\"threads\":" << threads << ",\"roundtrip\":" << (rt ? "true" : "false");
    };
    if (!details.empty()) {
        std::ofstream out(details);
        summary(out);
        out << ",\"results\":[";
        for (int n = 1; n <= max_n; ++n) {
            if (n > 1) out << ",";
            out << "{\"n\":" << n << ",\"score\":" << res[n].score
                << ",\"L\":" << (res[n].ok ? "\"" + res[n].L + "\"" : std::string("null")) << ",\"reason\":\""
                << json_str(res[n].reason) << "\"}";
        }
        out << "]}\n";
    }
    summary(std::cout);
    std::cout << "}\n";
    return 0;
}
