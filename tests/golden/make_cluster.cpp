// Writes the N=128 random cluster of the reference's Barnes-Hut audit test
// (engine/tests/test_phase18_19.cpp:38-66: std::mt19937_64(12345), positions
// U(-5, 5), velocities U(-0.01, 0.01), masses U(0.5, 5), radius 0.001 cbrt(m))
// as raw little-endian doubles: px, py, pz, vx, vy, vz, mass, radius.
// Build and run: g++ -O2 -std=c++17 make_cluster.cpp -o /tmp/mc && /tmp/mc cluster128.bin
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

int main(int argc, char** argv) {
    const int n = 128;
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> pos(-10.0 / 2.0, 10.0 / 2.0);
    std::uniform_real_distribution<double> vel(-0.01, 0.01);
    std::uniform_real_distribution<double> mass(0.5, 5.0);
    std::vector<double> f[8];
    for (auto& v : f) v.resize(n);
    for (int i = 0; i < n; i++) {
        f[0][i] = pos(rng); f[1][i] = pos(rng); f[2][i] = pos(rng);
        f[3][i] = vel(rng); f[4][i] = vel(rng); f[5][i] = vel(rng);
        f[6][i] = mass(rng);
        f[7][i] = 0.001 * std::cbrt(f[6][i]);
    }
    FILE* fp = std::fopen(argc > 1 ? argv[1] : "cluster128.bin", "wb");
    if (!fp) return 1;
    for (auto& v : f) std::fwrite(v.data(), sizeof(double), n, fp);
    std::fclose(fp);
    return 0;
}
