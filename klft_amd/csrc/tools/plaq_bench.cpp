// plaq_bench -n N (benchmarks/Gauge/plaq_bench.cpp): plaquette of cold 2D, 3D
// and 4D U(1)/SU(2)/SU(3) fields, fastest of 20 calls, with the reference's
// bandwidth formula 16 (V + Nd V Nc^2) / t and the algorithmic 16 Nd V Nc^2 / t.
#include "bench_common.hpp"

template <size_t Nd, size_t Nc>
void one(size_t n) {
  const char *grp[] = {"", "U(1)", "SU(2)", "SU(3)"};
  printf("%s Gauge %zuD: \n", grp[Nc], Nd);
  double V = 1.0;
  for (size_t d = 0; d < Nd; ++d) V *= (double)n;
  printf("- Size of %s GaugeField:                  %12.2f MB\n", grp[Nc], 1.0e-6 * Nd * V * Nc * Nc * 16.0);
  printf(HLINE);
  klft::real_t plaq = 0.0;
  double t;
  if constexpr (Nd == 4) {
    klft::deviceGaugeField<4, Nc> g(n, n, n, n, klft::identitySUN<Nc>());
    t = min_time([&] { plaq = klft::GaugePlaquette<4, Nc>(g); });
  } else if constexpr (Nd == 3) {
    klft::deviceGaugeField<3, Nc> g(n, n, n, klft::identitySUN<Nc>());
    t = min_time([&] { plaq = klft::GaugePlaquette<3, Nc>(g); });
  } else {
    klft::deviceGaugeField<2, Nc> g(n, n, klft::identitySUN<Nc>());
    t = min_time([&] { plaq = klft::GaugePlaquette<2, Nc>(g); });
  }
  printf("Calculated Plaquette value: %11.4e      Expected value: %11.4e\n", plaq, 1.0);
  printf("Plaquette Kernel Time:     %11.4e s\n", t);
  printf("Plaquette BW:              %11.4f GB/s\n", 1.0e-9 * 16.0 * (V + Nd * V * Nc * Nc) / t);
  printf("Plaquette algorithmic BW:  %11.4f GB/s\n", 1.0e-9 * 16.0 * Nd * V * Nc * Nc / t);
  printf(HLINE);
}

int main(int argc, char *argv[]) {
  printf(HLINE);
  printf("SU(N) GaugeField 2D, 3D and 4D plaquette kernel test and benchmark\n");
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
