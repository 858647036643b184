// metropolis_bench -n N (benchmarks/Gauge/metropolis_bench.cpp): one sweep of
// a cold N^d field with the MetropolisParams struct defaults, fastest of 20,
// printing the acceptance rate, the time and link-updates/s.
#include "bench_common.hpp"

template <size_t Nd, size_t Nc>
void one(size_t n) {
  const char *grp[] = {"", "U(1)", "SU(2)", "SU(3)"};
  printf("%s Gauge %zuD: \n", grp[Nc], Nd);
  klft::MetropolisParams p;
  p.Ndims = (int)Nd;
  p.Nd = Nd;
  p.Nc = Nc;
  p.L0 = p.L1 = p.L2 = p.L3 = (int)n;
  klft::RNGType rng(p.seed);
  double V = 1.0;
  for (size_t d = 0; d < Nd; ++d) V *= (double)n;
  klft::real_t acc = 0.0;
  double t;
  if constexpr (Nd == 4) {
    klft::deviceGaugeField<4, Nc> g(n, n, n, n, klft::identitySUN<Nc>());
    t = min_time([&] { acc = klft::sweep_Metropolis<4, Nc>(g, p, rng); });
  } else if constexpr (Nd == 3) {
    klft::deviceGaugeField<3, Nc> g(n, n, n, klft::identitySUN<Nc>());
    t = min_time([&] { acc = klft::sweep_Metropolis<3, Nc>(g, p, rng); });
  } else {
    klft::deviceGaugeField<2, Nc> g(n, n, klft::identitySUN<Nc>());
    t = min_time([&] { acc = klft::sweep_Metropolis<2, Nc>(g, p, rng); });
  }
  printf("Acceptance rate: %f\n", acc);
  printf("Metropolis Sweep Time:     %11.4e s\n", t);
  printf("Link updates per second:   %11.4e\n", Nd * V / t);
  printf(HLINE);
}

int main(int argc, char *argv[]) {
  printf(HLINE);
  printf("SU(N) GaugeField 2D, 3D and 4D Metropolis sweep benchmark\n");
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
