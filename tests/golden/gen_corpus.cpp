// Regenerates the input corpora of the reference's device tests with libstdc++
// (the same generators and call order as the reference):
//   selftest.f32  cuda/tests/selftest.cu:33-38   std::mt19937(923), 100 blocks x 50 squares,
//                 x, y ~ U(-0.2, 12.2), angle ~ U(-3.14, 3.14) (float), then the four
//                 edge fixtures (2,2,0) (3,2,0) (3.000001,2,0) (2.999999,2,0).
//   seeded.f64    cuda/tests/seeded_test.cu:15-18 std::mt19937(832), 10000 pairs of
//                 (x, y, a) ~ U(-1, 1) (double), p then z.
// Usage: gen_corpus <out_dir>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

int main(int argc, char** argv) {
    if (argc != 2) return 2;
    const std::string dir = argv[1];
    {
        std::mt19937 rng(923);
        std::uniform_real_distribution<float> coord(-.2f, 12.2f), angle(-3.14f, 3.14f);
        std::vector<float> q(100 * 50 * 3);
        for (size_t k = 0; k < q.size(); k += 3) {
            // braced initialisation in the reference evaluates left to right
            float x = coord(rng), y = coord(rng), a = angle(rng);
            q[k] = x; q[k + 1] = y; q[k + 2] = a;
        }
        const float edge[4][3] = {{2, 2, 0}, {3, 2, 0}, {3.000001f, 2, 0}, {2.999999f, 2, 0}};
        for (int i = 0; i < 4; ++i)
            for (int d = 0; d < 3; ++d) q[3 * i + d] = edge[i][d];
        FILE* f = std::fopen((dir + "/selftest.f32").c_str(), "wb");
        if (!f) return 1;
        std::fwrite(q.data(), sizeof(float), q.size(), f);
        std::fclose(f);
    }
    {
        std::mt19937 r(832);
        std::uniform_real_distribution<double> u(-1, 1);
        std::vector<double> q(2 * 10000 * 3);
        for (int i = 0; i < 10000; ++i) {
            for (int d = 0; d < 3; ++d) q[6 * i + d] = u(r);
            for (int d = 0; d < 3; ++d) q[6 * i + 3 + d] = u(r);
        }
        FILE* f = std::fopen((dir + "/seeded.f64").c_str(), "wb");
        if (!f) return 1;
        std::fwrite(q.data(), sizeof(double), q.size(), f);
        std::fclose(f);
    }
    return 0;
}
