// `hmc` driver (hmc.cpp:19-73, lib/HMC_SU2.cpp:8-75): same flags and defaults;
// --gauge-group accepts SU2, U1 and (added) SU3, --precision f64|f32|f32_rows
// selects the link storage.  Hot start: the reference's
// gauge_field.set_random(0.5, RNG(seed*2)) (lib/HMC_SU2.cpp:41-42), read as
// every link a randSUN(0.5) draw (the deviceGaugeField(L.., rng, delta)
// start, GaugeField.hpp:96-112) from the stream keyed by seed*2.  Output:
// "traj, accept, plaquette, time, acceptance rate" with one line per
// trajectory.
#include <chrono>
#include <fstream>
#include <iostream>
#include <string>

#include "klft_amd.hpp"

using real_t = double;

int main(int argc, char **argv) {
  std::string gauge_group = "SU2", precision = "f64", outfilename = "";
  int ndim = 4;
  size_t LX = 8, LY = 8, LZ = 8, LT = 16, n_traj = 1000, n_steps = 50, seed = 1234;
  real_t tau = 1.0, beta = 2.0;
  for (int i = 1; i < argc; i++) {
    const std::string a(argv[i]);
    const bool has = i + 1 < argc;
    if (a == "--gauge-group" && has) gauge_group = argv[i + 1];
    if (a == "--ndim" && has) ndim = std::stoi(argv[i + 1]);
    if (a == "--LX" && has) LX = std::stoi(argv[i + 1]);
    if (a == "--LY" && has) LY = std::stoi(argv[i + 1]);
    if (a == "--LZ" && has) LZ = std::stoi(argv[i + 1]);
    if (a == "--LT" && has) LT = std::stoi(argv[i + 1]);
    if (a == "--n-traj" && has) n_traj = std::stoi(argv[i + 1]);
    if (a == "--n-steps" && has) n_steps = std::stoi(argv[i + 1]);
    if (a == "--tau" && has) tau = std::stod(argv[i + 1]);
    if (a == "--beta" && has) beta = std::stod(argv[i + 1]);
    if (a == "--seed" && has) seed = std::stoi(argv[i + 1]);
    if (a == "--outfilename" && has) outfilename = argv[i + 1];
    if (a == "--precision" && has) precision = argv[i + 1];
    if (a == "--help") {
      std::cout << "Usage: " << argv[0] << " [options]" << std::endl;
      std::cout << "Options:" << std::endl;
      std::cout << "--gauge-group SU2, U1 or SU3" << std::endl;
      std::cout << "--ndim 2, 3, or 4" << std::endl;
      std::cout << "--LX lattice size in x direction" << std::endl;
      std::cout << "--LY lattice size in y direction" << std::endl;
      std::cout << "--LZ lattice size in z direction" << std::endl;
      std::cout << "--LT lattice size in t direction" << std::endl;
      std::cout << "--n-traj number of trajectories" << std::endl;
      std::cout << "--n-steps number of steps in leapfrog" << std::endl;
      std::cout << "--tau trajectory length" << std::endl;
      std::cout << "--beta inverse coupling constant" << std::endl;
      std::cout << "--seed random number generator seed" << std::endl;
      std::cout << "--outfilename output filename" << std::endl;
      std::cout << "--precision f64, f32 or f32_rows" << std::endl;
      return 0;
    }
  }
  const int nc = gauge_group == "U1" ? 1 : (gauge_group == "SU2" ? 2 : (gauge_group == "SU3" ? 3 : 0));
  if (nc == 0 || ndim < 2 || ndim > 4) return 0;  // the reference runs nothing for other choices
  const int storage = precision == "f32" ? KLFT_STORAGE_F32
                                         : (precision == "f32_rows" ? KLFT_STORAGE_F32_ROWS : KLFT_STORAGE_F64);
  int dims[4] = {(int)LX, (int)LY, (int)LZ, (int)LT};
  if (ndim == 3) dims[2] = (int)LT;
  if (ndim == 2) dims[1] = (int)LT;
  std::ofstream outfile;
  if (outfilename != "") {
    outfile.open(outfilename);
    outfile << "traj, accept, plaquette, time, acceptance rate" << std::endl;
  }
  klft_field_t f = nullptr;
  klft_hmc_t h = nullptr;
  try {
    klft::check(klft_field_create(ndim, nc, dims, storage, &f));
    klft::check(klft_field_set_random_sun(f, 0.5, (uint64_t)seed * 2));
    klft_hmc_params p{beta, tau, (int)n_steps, (uint64_t)seed};
    klft::check(klft_hmc_create(f, &p, &h));
    real_t plaq = 0.0;
    klft::check(klft_plaquette(f, 1, &plaq, nullptr));
    std::cout << "Starting Plaquette: " << plaq << std::endl;
    std::cout << "Starting HMC: " << std::endl;
    size_t n_accept = 0;
    const auto hmc_start = std::chrono::high_resolution_clock::now();
    for (size_t i = 0; i < n_traj; i++) {
      const auto t0 = std::chrono::high_resolution_clock::now();
      double dH = 0.0;
      int acc = 0;
      klft::check(klft_hmc_trajectory(h, i, &dH, &acc));
      const std::chrono::duration<double> dt = std::chrono::high_resolution_clock::now() - t0;
      const bool accept = acc != 0;
      if (accept) n_accept++;
      klft::check(klft_plaquette(f, 1, &plaq, nullptr));
      std::cout << "Traj: " << i << " Accept: " << accept << " Plaquette: " << plaq << " Time: " << dt.count()
                << " Acceptance Rate: " << real_t(n_accept) / real_t(i + 1) << std::endl;
      if (outfilename != "")
        outfile << i << ", " << accept << ", " << plaq << ", " << dt.count() << ", "
                << real_t(n_accept) / real_t(i + 1) << std::endl;
    }
    const std::chrono::duration<double> total = std::chrono::high_resolution_clock::now() - hmc_start;
    std::cout << "HMC Time: " << total.count() << std::endl;
  } catch (const klft::Error &e) {
    std::cout << "Error: " << e.what() << std::endl;
    if (h) klft_hmc_destroy(h);
    if (f) klft_field_destroy(f);
    return -1;
  }
  klft_hmc_destroy(h);
  klft_field_destroy(f);
  outfile.close();
  return 0;
}
