// The uneven-cavity workload: n square cavities, per cavity N drawn uniformly from
// {64, 96, 128, 192, 256}, then Re log-uniform in [100, 3200], both from one std::mt19937_64(seed)
// in that order. Writes "N Re" lines (Re to 17 significant digits).
//   g++ -O2 -std=c++17 cavities.cpp -o cavities && ./cavities 64 0 > cavities_seed0.txt
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

int main(int argc, char** argv) {
    const int n = argc > 1 ? std::atoi(argv[1]) : 64;
    const unsigned long long seed = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 0;
    static const int sizes[5] = {64, 96, 128, 192, 256};
    std::mt19937_64 g(seed);
    std::uniform_int_distribution<int> pick(0, 4);
    std::uniform_real_distribution<double> lre(std::log10(100.0), std::log10(3200.0));
    for (int k = 0; k < n; ++k) {
        const int N = sizes[pick(g)];
        const double re = std::pow(10.0, lre(g));
        std::printf("%d %.17g\n", N, re);
    }
    return 0;
}
