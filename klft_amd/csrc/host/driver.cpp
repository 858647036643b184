// klft::Metropolis as a C entry point (lib/Metropolis.cpp:36-166): environment
// (KLFT_VERBOSITY, KLFT_TUNING, KLFT_CACHE_FILE), YAML input, cold start,
// dispatch on (Ndims, Nc), run_metropolis, tuning-cache write.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "klft_amd.hpp"

namespace {

template <size_t Nd, size_t Nc>
int run(const klft::MetropolisParams &mp, klft::GaugeObservableParams &gp, const klft::RNGType &rng) {
  if constexpr (Nd == 4) {
    klft::deviceGaugeField<4, Nc> g(mp.L0, mp.L1, mp.L2, mp.L3, klft::identitySUN<Nc>());
    klft::check(klft_field_set_proposal(g.h, mp.proposal));
    return klft::run_metropolis<4, Nc>(g, mp, gp, rng);
  } else if constexpr (Nd == 3) {
    klft::deviceGaugeField<3, Nc> g(mp.L0, mp.L1, mp.L2, klft::identitySUN<Nc>());
    klft::check(klft_field_set_proposal(g.h, mp.proposal));
    return klft::run_metropolis<3, Nc>(g, mp, gp, rng);
  } else {
    klft::deviceGaugeField<2, Nc> g(mp.L0, mp.L1, klft::identitySUN<Nc>());
    klft::check(klft_field_set_proposal(g.h, mp.proposal));
    return klft::run_metropolis<2, Nc>(g, mp, gp, rng);
  }
}

template <size_t Nd>
int by_nc(const klft::MetropolisParams &mp, klft::GaugeObservableParams &gp, const klft::RNGType &rng) {
  switch (mp.Nc) {
    case 1: return run<Nd, 1>(mp, gp, rng);
    case 2: return run<Nd, 2>(mp, gp, rng);
    case 3: return run<Nd, 3>(mp, gp, rng);
  }
  printf("Error: Unsupported gauge group\n");
  return -1;
}

}  // namespace

extern "C" int klft_metropolis_main(const char *input_file) {
  using namespace klft;
  const char *v = std::getenv("KLFT_VERBOSITY");
  setVerbosity(v ? std::atoi(v) : 0);
  const char *t = std::getenv("KLFT_TUNING");
  setTuning(t ? std::atoi(t) : 0);
  if (KLFT_TUNING) {
    const char *cache = std::getenv("KLFT_CACHE_FILE");
    if (cache) {
      if (KLFT_VERBOSITY > 0) printf("Reading cache file: %s\n", cache);
      klft_tune_cache_read(cache);
    }
  }
  MetropolisParams mp;
  GaugeObservableParams gp;
  const std::string file = input_file ? input_file : "";
  if (!parseInputFile(file, mp)) {
    printf("Error parsing input file\n");
    return -1;
  }
  if (!parseInputFile(file, gp)) {
    printf("Error parsing input file\n");
    return -1;
  }
  mp.print();
  const RNGType rng((uint64_t)(int64_t)mp.seed);
  int rc = 0;
  try {
    if (mp.Ndims == 4)
      rc = by_nc<4>(mp, gp, rng);
    else if (mp.Ndims == 3)
      rc = by_nc<3>(mp, gp, rng);
    else if (mp.Ndims == 2)
      rc = by_nc<2>(mp, gp, rng);
  } catch (const klft::Error &e) {
    printf("Error: %s\n", e.what());
    return -1;
  }
  if (rc != 0) return -1;
  if (KLFT_TUNING) {
    const char *cache = std::getenv("KLFT_CACHE_FILE");
    if (cache)
      klft_tune_cache_write(cache);
    else
      printf("KLFT_CACHE_FILE not set\n");
  }
  return 0;
}
