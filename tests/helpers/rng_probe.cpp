// Prints libstdc++ draws used by the reference's initOmg / permute_matrix
// (std::default_random_engine, normal / uniform_real distributions,
// std::shuffle), so the oracle's Python restatement can be checked against the
// real library.  Usage: rng_probe <mode> <seed> <n>
//   mode normal|uniform: n draws, one per line (%.17g)
//   mode shuffle: permutation of 0..n-1 with a default-constructed engine
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

int main(int argc, char** argv) {
  if (argc != 4) return 2;
  const char* mode = argv[1];
  const int seed = std::atoi(argv[2]);
  const long n = std::atol(argv[3]);
  if (!std::strcmp(mode, "shuffle")) {
    std::vector<int> v(n);
    std::iota(v.begin(), v.end(), 0);
    auto rng = std::default_random_engine{};
    std::shuffle(v.data(), v.data() + v.size(), rng);
    for (int x : v) std::printf("%d\n", x);
    return 0;
  }
  auto rng = std::default_random_engine{};
  rng.seed(seed);
  if (!std::strcmp(mode, "normal")) {
    std::normal_distribution<double> d{0, 1};
    for (long i = 0; i < n; ++i) std::printf("%.17g\n", d(rng));
  } else {
    std::uniform_real_distribution<double> d{-1, 1};
    for (long i = 0; i < n; ++i) std::printf("%.17g\n", d(rng));
  }
  return 0;
}
