// wloop_temporal_bench -n N (benchmarks/Gauge/wloop_temporal_bench.cpp):
// temporal Wilson loops (L,T) = (N/4,N/4), (N/2,N/4), (N/2,N/2) on cold
// fields, fastest of 20, reference bandwidth 16 (V + 2 V Nc^2 + Nd V Nc^2) / t.
#include <vector>

#include "bench_common.hpp"

template <size_t Nd, size_t Nc>
void one(size_t n) {
  const char *grp[] = {"", "U(1)", "SU(2)", "SU(3)"};
  printf("%s Gauge %zuD: \n", grp[Nc], Nd);
  double V = 1.0;
  for (size_t d = 0; d < Nd; ++d) V *= (double)n;
  const int a = std::max<int>(1, n / 4), b = std::max<int>(1, n / 2);
  std::vector<klft::Array<klft::index_t, 2>> pairs = {{a, a}, {b, a}, {b, b}};
  std::vector<klft::Array<klft::real_t, 3>> out;
  double t;
  if constexpr (Nd == 4) {
    klft::deviceGaugeField<4, Nc> g(n, n, n, n, klft::identitySUN<Nc>());
    t = min_time([&] { out.clear(); klft::WilsonLoop_temporal<4, Nc>(g, pairs, out); });
  } else if constexpr (Nd == 3) {
    klft::deviceGaugeField<3, Nc> g(n, n, n, klft::identitySUN<Nc>());
    t = min_time([&] { out.clear(); klft::WilsonLoop_temporal<3, Nc>(g, pairs, out); });
  } else {
    klft::deviceGaugeField<2, Nc> g(n, n, klft::identitySUN<Nc>());
    t = min_time([&] { out.clear(); klft::WilsonLoop_temporal<2, Nc>(g, pairs, out); });
  }
  for (const auto &w : out)
    printf("L = %d, T = %d: Calculated W_temp: %11.4e      Expected: %11.4e\n", (int)w[0], (int)w[1], w[2], 1.0);
  printf("Wilson Loop Kernel Time:   %11.4e s\n", t);
  printf("Wilson Loop BW:            %11.4f GB/s\n", 1.0e-9 * 16.0 * (V + 2 * V * Nc * Nc + Nd * V * Nc * Nc) / t);
  printf(HLINE);
}

int main(int argc, char *argv[]) {
  printf(HLINE);
  printf("SU(N) GaugeField 2D, 3D and 4D temporal Wilson loop test and benchmark\n");
  printf(HLINE);
  size_t n;
  int rc = parse_n(argc, argv, n);
  if (rc == -2) return 0;
  if (rc) return rc;
  printf("Reports fastest timing per kernel\n");
  printf("Lattice extent: %ld\n", (long)n);
  printf(HLINE);
  try {
    one<4, 1>(n); one<4, 2>(n); one<4, 3>(n);
    one<3, 1>(n); one<3, 2>(n); one<3, 3>(n);
    one<2, 1>(n); one<2, 2>(n); one<2, 3>(n);
  } catch (const klft::Error &e) {
    printf("Error: %s\n", e.what());
    return -1;
  }
  return 0;
}
