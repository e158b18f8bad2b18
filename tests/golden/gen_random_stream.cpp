// Golden vectors for the Omega stream of Matrix::Random (include/matrix.h:50-57):
// std::mt19937 seeded with a fixed value, uniform_real_distribution<double>(0,1),
// values in draw (= column-major storage) order.  Build and run:
//   g++ -O2 -std=c++17 gen_random_stream.cpp -o /tmp/gen && /tmp/gen > random_stream.json
#include <cstdio>
#include <random>

int main() {
  const unsigned seeds[] = {5489u, 7u, 12345u};
  std::printf("{\n");
  for (int s = 0; s < 3; ++s) {
    std::mt19937 g(seeds[s]);
    std::uniform_real_distribution<double> d(0.0, 1.0);
    std::printf("  \"%u\": [", seeds[s]);
    for (int i = 0; i < 2000; ++i) std::printf("%s%.17g", i ? ", " : "", d(g));
    std::printf("]%s\n", s < 2 ? "," : "");
  }
  std::printf("}\n");
  return 0;
}
