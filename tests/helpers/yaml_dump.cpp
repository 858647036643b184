// Test helper: parses an input file with the drop-in parser (klft_amd.hpp)
// and prints the resulting parameters as JSON, or {"ok": false}.
#include <cstdio>

#include "klft_amd.hpp"

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  klft::MetropolisParams mp;
  klft::GaugeObservableParams gp;
  const bool a = klft::parseInputFile(argv[1], mp);
  const bool b = a && klft::parseInputFile(argv[1], gp);
  if (!a || !b) {
    printf("\n{\"ok\": false}\n");
    return 0;
  }
  auto pairs = [](const std::vector<klft::Array<int, 2>> &v) {
    std::string s = "[";
    for (size_t i = 0; i < v.size(); ++i)
      s += (i ? ", [" : "[") + std::to_string(v[i][0]) + ", " + std::to_string(v[i][1]) + "]";
    return s + "]";
  };
  printf("\n{\"ok\": true, \"Ndims\": %d, \"L\": [%d, %d, %d, %d], \"nHits\": %d, \"nSweep\": %d, \"seed\": %d, "
         "\"Nd\": %zu, \"Nc\": %zu, \"beta\": %.17g, \"delta\": %.17g, \"measurement_interval\": %zu, "
         "\"measure_plaquette\": %s, \"measure_wilson_loop_temporal\": %s, \"measure_wilson_loop_mu_nu\": %s, "
         "\"W_temp_L_T_pairs\": %s, \"W_mu_nu_pairs\": %s, \"W_Lmu_Lnu_pairs\": %s, \"plaquette_filename\": \"%s\", "
         "\"W_temp_filename\": \"%s\", \"W_mu_nu_filename\": \"%s\", \"write_to_file\": %s}\n",
         mp.Ndims, mp.L0, mp.L1, mp.L2, mp.L3, mp.nHits, mp.nSweep, mp.seed, mp.Nd, mp.Nc, mp.beta, mp.delta,
         gp.measurement_interval, gp.measure_plaquette ? "true" : "false",
         gp.measure_wilson_loop_temporal ? "true" : "false", gp.measure_wilson_loop_mu_nu ? "true" : "false",
         pairs(gp.W_temp_L_T_pairs).c_str(), pairs(gp.W_mu_nu_pairs).c_str(), pairs(gp.W_Lmu_Lnu_pairs).c_str(),
         gp.plaquette_filename.c_str(), gp.W_temp_filename.c_str(), gp.W_mu_nu_filename.c_str(),
         gp.write_to_file ? "true" : "false");
  return 0;
}
