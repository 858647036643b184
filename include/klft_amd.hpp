// klft_amd.hpp -- klft's C++ API over the MI355X C ABI (klft_hip.h).
//
// Code written against the reference's headers (Metropolis.hpp,
// GaugePlaquette.hpp, WilsonLoop.hpp, GaugeObservable.hpp, InputParser.hpp,
// Metropolis_Params.hpp, HMC_Params.hpp, klft.hpp) compiles against this
// header with the same names, argument meaning, return values and output
// files.  Differences a caller can see:
//   * the random generator argument is an RNGType carrying the seed and a
//     sweep counter (the reference's Kokkos pool is replaced by the keyed
//     Philox streams of the library);
//   * odd lattice extents are rejected (the reference silently skips a slice,
//     Metropolis.hpp:127-130) and parameter mismatches return -1 instead of
//     failing an assert;
//   * deviceGaugeField<Nd,Nc> covers the reference's deviceGaugeField,
//     deviceGaugeField3D and deviceGaugeField2D (one template on Nd).
#pragma once
#include <array>
#include <chrono>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "klft_hip.h"
#include "klft_yaml.hpp"

namespace klft {

using real_t = double;
using index_t = int;
using complex_t = std::complex<real_t>;
template <typename T, size_t N>
using Array = std::array<T, N>;
template <size_t N>
using IndexArray = std::array<index_t, N>;
template <size_t Nc>
using SUN = Array<Array<complex_t, Nc>, Nc>;

// GLOBAL.hpp:236-256
inline int KLFT_VERBOSITY = 0;
inline int KLFT_TUNING = 0;
inline void setVerbosity(int v) { KLFT_VERBOSITY = v; }
inline void setTuning(int t) {
  KLFT_TUNING = t;
  klft_tuning(t);
}

template <size_t Nc>
inline SUN<Nc> identitySUN() {
  SUN<Nc> a{};
  for (size_t i = 0; i < Nc; ++i) a[i][i] = complex_t(1.0, 0.0);
  return a;
}

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};
inline void check(int rc) {
  if (rc != KLFT_OK) throw Error(klft_last_error());
}

// Metropolis_Params.hpp:28-99
struct MetropolisParams {
  index_t Ndims = 4, L0 = 4, L1 = 4, L2 = 4, L3 = 4;
  index_t nHits = 10, nSweep = 1000, seed = 1234;
  size_t Nd = 4, Nc = 2;
  real_t beta = 1.0, delta = 0.1;
  // SU(3) proposal law (no counterpart in the reference): KLFT_PROPOSAL_REFERENCE
  // or, from the optional input key "proposal: symmetric", KLFT_PROPOSAL_SYMMETRIC
  int proposal = KLFT_PROPOSAL_REFERENCE;
  void print() const {
    if (KLFT_VERBOSITY > 0) {
      printf("Metropolis Parameters:\n");
      printf("General Parameters:\n");
      printf("Ndims: %d\n", Ndims);
      printf("L0: %d\n", L0);
      printf("L1: %d\n", L1);
      printf("L2: %d\n", L2);
      printf("L3: %d\n", L3);
      printf("nHits: %d\n", nHits);
      printf("nSweep: %d\n", nSweep);
      printf("seed: %d\n", seed);
      printf("GaugeField Parameters:\n");
      printf("Nd: %zu\n", Nd);
      printf("Nc: %zu\n", Nc);
      printf("Wilson Action Parameters:\n");
      printf("beta: %.4f\n", beta);
      printf("delta: %.4f\n", delta);
      if (proposal == KLFT_PROPOSAL_SYMMETRIC) printf("proposal: symmetric\n");
    }
  }
};

// HMC_Params.hpp:5-15
class HMC_Params {
 public:
  size_t n_steps;
  double tau;
  HMC_Params(size_t n, double t) : n_steps(n), tau(t) {}
  double get_tau() const { return tau; }
  void set_tau(const double &t) { tau = t; }
  size_t get_n_steps() const { return n_steps; }
  void set_n_steps(const size_t &n) { n_steps = n; }
};

// GaugeObservable.hpp:31-61
struct GaugeObservableParams {
  size_t measurement_interval = 0;
  bool measure_plaquette = false;
  bool measure_wilson_loop_temporal = false;
  bool measure_wilson_loop_mu_nu = false;
  std::vector<Array<index_t, 2>> W_temp_L_T_pairs;
  std::vector<Array<index_t, 2>> W_mu_nu_pairs;
  std::vector<Array<index_t, 2>> W_Lmu_Lnu_pairs;
  std::vector<size_t> measurement_steps;
  std::vector<real_t> plaquette_measurements;
  std::vector<std::vector<Array<real_t, 3>>> W_temp_measurements;
  std::vector<std::vector<Array<real_t, 5>>> W_mu_nu_measurements;
  std::string plaquette_filename;
  std::string W_temp_filename;
  std::string W_mu_nu_filename;
  bool write_to_file = false;
};

// The generator handle of sweep_Metropolis: seed plus the index of the next
// sweep, which keys the Philox streams of that sweep.
struct RNGType {
  uint64_t seed;
  mutable uint64_t sweep = 0;
  explicit RNGType(uint64_t s) : seed(s) {}
};

inline int storage_from_env() {
  const char *e = std::getenv("KLFT_STORAGE");
  if (!e) return KLFT_STORAGE_F64;
  const std::string s(e);
  if (s == "f32") return KLFT_STORAGE_F32;
  if (s == "f32_rows") return KLFT_STORAGE_F32_ROWS;
  return KLFT_STORAGE_F64;
}

// deviceGaugeField<Nd,Nc> (GaugeField.hpp:30-620): a handle to links held on
// the GPU in the library's layout.
template <size_t Nd, size_t Nc>
class deviceGaugeField {
 public:
  IndexArray<4> dimensions{1, 1, 1, 1};
  klft_field_t h = nullptr;

  template <typename... L>
  explicit deviceGaugeField(const SUN<Nc> &init, int storage, L... extents) {
    static_assert(sizeof...(L) == Nd, "one extent per dimension");
    const index_t e[] = {(index_t)extents...};
    create(e, storage);
    set(init);
  }
  // deviceGaugeField(L0, .., L_{Nd-1}, identitySUN<Nc>()) and friends
  deviceGaugeField(index_t L0, index_t L1, index_t L2, index_t L3, const SUN<Nc> &init) {
    static_assert(Nd == 4, "4D constructor");
    const index_t e[] = {L0, L1, L2, L3};
    create(e, storage_from_env());
    set(init);
  }
  deviceGaugeField(index_t L0, index_t L1, index_t L2, const SUN<Nc> &init) {
    static_assert(Nd == 3, "3D constructor");
    const index_t e[] = {L0, L1, L2};
    create(e, storage_from_env());
    set(init);
  }
  deviceGaugeField(index_t L0, index_t L1, const SUN<Nc> &init) {
    static_assert(Nd == 2, "2D constructor");
    const index_t e[] = {L0, L1};
    create(e, storage_from_env());
    set(init);
  }
  deviceGaugeField(const deviceGaugeField &) = delete;
  deviceGaugeField &operator=(const deviceGaugeField &) = delete;
  ~deviceGaugeField() {
    if (h) klft_field_destroy(h);
  }

  std::vector<complex_t> download() const {
    std::vector<complex_t> v(volume() * Nd * Nc * Nc);
    check(klft_field_download(h, reinterpret_cast<double *>(v.data())));
    return v;
  }
  void upload(const std::vector<complex_t> &v) {
    check(klft_field_upload(h, reinterpret_cast<const double *>(v.data())));
  }
  size_t volume() const {
    size_t v = 1;
    for (size_t d = 0; d < Nd; ++d) v *= dimensions[d];
    return v;
  }

 private:
  void create(const index_t *e, int storage) {
    int dims[4] = {1, 1, 1, 1};
    for (size_t d = 0; d < Nd; ++d) dims[d] = dimensions[d] = e[d];
    check(klft_field_create((int)Nd, (int)Nc, dims, storage, &h));
  }
  void set(const SUN<Nc> &init) {
    if (init == identitySUN<Nc>()) {
      check(klft_field_set_identity(h));
      return;
    }
    std::vector<complex_t> v(volume() * Nd * Nc * Nc);
    for (size_t i = 0; i < volume() * Nd; ++i)
      for (size_t a = 0; a < Nc; ++a)
        for (size_t b = 0; b < Nc; ++b) v[(i * Nc + a) * Nc + b] = init[a][b];
    upload(v);
  }
};

template <size_t Nd, size_t Nc>
struct DeviceGaugeFieldType {
  using type = deviceGaugeField<Nd, Nc>;
};

// sweep_Metropolis (Metropolis.hpp:118-178): acceptance rate of one sweep
template <size_t rank, size_t Nc>
real_t sweep_Metropolis(deviceGaugeField<rank, Nc> &g, const MetropolisParams &params,
                        const RNGType &rng) {
  klft_metropolis_params p{params.nHits, params.beta, params.delta, rng.seed};
  double acc = 0.0;
  check(klft_metropolis_sweep(g.h, &p, rng.sweep++, &acc, nullptr));
  return acc;
}

// GaugePlaquette (GaugePlaquette.hpp:89-135)
template <size_t rank, size_t Nc>
real_t GaugePlaquette(const deviceGaugeField<rank, Nc> &g, const bool normalize = true) {
  double p = 0.0;
  check(klft_plaquette(g.h, normalize ? 1 : 0, &p, nullptr));
  return p;
}

// WilsonLoop_mu_nu (WilsonLoop.hpp:170-245): appends {mu, nu, Lmu, Lnu, W}
template <size_t rank, size_t Nc>
void WilsonLoop_mu_nu(const deviceGaugeField<rank, Nc> &g, const index_t mu, const index_t nu,
                      const std::vector<Array<index_t, 2>> &Lmu_nu_pairs,
                      std::vector<Array<real_t, 5>> &Wmunu_vals, const bool normalize = true) {
  std::vector<int> flat;
  for (const auto &p : Lmu_nu_pairs) {
    flat.push_back(p[0]);
    flat.push_back(p[1]);
  }
  std::vector<double> out(Lmu_nu_pairs.size() + 1);
  check(klft_wilson_loop_mu_nu(g.h, mu, nu, flat.data(), (int)Lmu_nu_pairs.size(), normalize ? 1 : 0,
                               out.data(), nullptr));
  for (size_t i = 0; i < Lmu_nu_pairs.size(); ++i)
    Wmunu_vals.push_back(Array<real_t, 5>{(real_t)mu, (real_t)nu, (real_t)Lmu_nu_pairs[i][0],
                                          (real_t)Lmu_nu_pairs[i][1], out[i]});
}

// WilsonLoop_temporal (WilsonLoop.hpp:250-304): appends {L, T, W}
template <size_t rank, size_t Nc>
void WilsonLoop_temporal(const deviceGaugeField<rank, Nc> &g,
                         const std::vector<Array<index_t, 2>> &L_T_pairs,
                         std::vector<Array<real_t, 3>> &Wtemporal_vals, const bool normalize = true) {
  std::vector<int> flat;
  for (const auto &p : L_T_pairs) {
    flat.push_back(p[0]);
    flat.push_back(p[1]);
  }
  std::vector<double> out(L_T_pairs.size() + 1);
  check(klft_wilson_loop_temporal(g.h, flat.data(), (int)L_T_pairs.size(), normalize ? 1 : 0, out.data()));
  for (size_t i = 0; i < L_T_pairs.size(); ++i)
    Wtemporal_vals.push_back(Array<real_t, 3>{(real_t)L_T_pairs[i][0], (real_t)L_T_pairs[i][1], out[i]});
}

// stapleField (Gauge_Util.hpp:27-97): the staple of every link into host memory
template <size_t Nd, size_t Nc>
std::vector<complex_t> stapleField(const deviceGaugeField<Nd, Nc> &g) {
  int nd, nc, dims[4], storage;
  check(klft_field_info(g.h, &nd, &nc, dims, &storage));
  klft_field_t out = nullptr;
  check(klft_field_create(nd, nc, dims, storage == KLFT_STORAGE_F64 ? KLFT_STORAGE_F64 : KLFT_STORAGE_F32,
                          &out));
  std::vector<complex_t> v(g.volume() * Nd * Nc * Nc);
  int rc = klft_staple_field(g.h, out);
  if (rc == KLFT_OK) rc = klft_field_download(out, reinterpret_cast<double *>(v.data()));
  klft_field_destroy(out);
  check(rc);
  return v;
}

// measureGaugeObservables (GaugeObservable.hpp:64-127)
template <size_t rank, size_t Nc>
void measureGaugeObservables(const deviceGaugeField<rank, Nc> &g, GaugeObservableParams &params,
                             const size_t step) {
  if ((params.measurement_interval == 0) || (step % params.measurement_interval != 0) || (step == 0)) return;
  if (KLFT_VERBOSITY > 0) {
    printf("Measurement of Gauge Observables\n");
    printf("step: %zu\n", step);
  }
  if (params.measure_plaquette) {
    const real_t P = GaugePlaquette<rank, Nc>(g);
    params.plaquette_measurements.push_back(P);
    if (KLFT_VERBOSITY > 0) printf("plaquette: %11.6f\n", P);
  }
  if (params.measure_wilson_loop_temporal) {
    if (KLFT_VERBOSITY > 0) {
      printf("temporal Wilson loop:\n");
      printf("L, T, W_temp\n");
    }
    std::vector<Array<real_t, 3>> temp;
    WilsonLoop_temporal<rank, Nc>(g, params.W_temp_L_T_pairs, temp);
    if (KLFT_VERBOSITY > 0)
      for (const auto &m : temp) printf("%d, %d, %11.6f\n", (index_t)m[0], (index_t)m[1], m[2]);
    params.W_temp_measurements.push_back(temp);
  }
  if (params.measure_wilson_loop_mu_nu) {
    if (KLFT_VERBOSITY > 0) {
      printf("Wilson loop in the mu-nu plane:\n");
      printf("mu, nu, Lmu, Lnu, W_mu_nu\n");
    }
    std::vector<Array<real_t, 5>> temp;
    for (const auto &pmn : params.W_mu_nu_pairs) {
      WilsonLoop_mu_nu<rank, Nc>(g, pmn[0], pmn[1], params.W_Lmu_Lnu_pairs, temp);
      if (KLFT_VERBOSITY > 0)  // re-prints the accumulated list per plane, as the reference
        for (const auto &m : temp)
          printf("%d, %d, %d, %d, %11.6f\n", (index_t)m[0], (index_t)m[1], (index_t)m[2], (index_t)m[3], m[4]);
    }
    params.W_mu_nu_measurements.push_back(temp);
  }
  params.measurement_steps.push_back(step);
}

// flush functions (GaugeObservable.hpp:130-230): default ostream formatting
inline void flushPlaquette(std::ofstream &file, const GaugeObservableParams &params, const bool HEADER = true) {
  if (!file.is_open()) {
    printf("Error: file is not open\n");
    return;
  }
  if (!params.measure_plaquette) {
    printf("Error: no plaquette measurements available\n");
    return;
  }
  if (HEADER) file << "# step, plaquette\n";
  for (size_t i = 0; i < params.measurement_steps.size(); ++i)
    file << params.measurement_steps[i] << ", " << params.plaquette_measurements[i] << "\n";
}

inline void flushWilsonLoopTemporal(std::ofstream &file, const GaugeObservableParams &params,
                                    const bool HEADER = true) {
  if (!file.is_open()) {
    printf("Error: file is not open\n");
    return;
  }
  if (!params.measure_wilson_loop_temporal) {
    printf("Error: no temporal Wilson loop measurements available\n");
    return;
  }
  if (HEADER) file << "# step, L, T, W_temp\n";
  for (size_t i = 0; i < params.measurement_steps.size(); ++i)
    for (const auto &m : params.W_temp_measurements[i])
      file << params.measurement_steps[i] << ", " << m[0] << ", " << m[1] << ", " << m[2] << "\n";
}

inline void flushWilsonLoopMuNu(std::ofstream &file, const GaugeObservableParams &params,
                                const bool HEADER = true) {
  if (!file.is_open()) {
    printf("Error: file is not open\n");
    return;
  }
  if (!params.measure_wilson_loop_mu_nu) {
    printf("Error: no mu-nu Wilson loop measurements available\n");
    return;
  }
  if (HEADER) file << "# step, mu, nu, Lmu, Lnu, W_mu_nu\n";
  for (size_t i = 0; i < params.measurement_steps.size(); ++i)
    for (const auto &m : params.W_mu_nu_measurements[i])
      file << params.measurement_steps[i] << ", " << m[0] << ", " << m[1] << ", " << m[2] << ", " << m[3]
           << ", " << m[4] << "\n";
}

inline void flushAllGaugeObservables(const GaugeObservableParams &params, const bool HEADER = true) {
  if (!params.write_to_file) {
    printf("write_to_file is not enabled\n");
    return;
  }
  if (params.plaquette_filename != "") {
    std::ofstream file(params.plaquette_filename, std::ios::app);
    flushPlaquette(file, params, HEADER);
    file.close();
  }
  if (params.W_temp_filename != "") {
    std::ofstream file(params.W_temp_filename, std::ios::app);
    flushWilsonLoopTemporal(file, params, HEADER);
    file.close();
  }
  if (params.W_mu_nu_filename != "") {
    std::ofstream file(params.W_mu_nu_filename, std::ios::app);
    flushWilsonLoopMuNu(file, params, HEADER);
    file.close();
  }
}

inline void clearAllGaugeObservables(GaugeObservableParams &params) {
  params.measurement_steps.clear();
  params.plaquette_measurements.clear();
  params.W_temp_measurements.clear();
  params.W_mu_nu_measurements.clear();
}

// run_metropolis (Metropolis.hpp:180-223).  With KLFT_VERBOSITY == 0 the
// sweeps between two measurement steps are enqueued without reading the
// acceptance back (nothing prints it), which saves a host round trip per
// sweep; the field follows the same sequence of keyed sweeps either way.
template <size_t rank, size_t Nc>
int run_metropolis(deviceGaugeField<rank, Nc> &g, const MetropolisParams &mp, GaugeObservableParams &gp,
                   const RNGType &rng) {
  const auto &dims = g.dimensions;
  bool ok = mp.Ndims == (index_t)rank && mp.Nd == rank && mp.Nc == Nc && mp.L0 == dims[0] && mp.L1 == dims[1];
  if (rank > 2) ok = ok && mp.L2 == dims[2];
  if (rank > 3) ok = ok && mp.L3 == dims[3];
  if (!ok) {
    printf("Error: MetropolisParams do not match the gauge field\n");
    return -1;
  }
  const size_t nsweep = mp.nSweep > 0 ? (size_t)mp.nSweep : 0;
  size_t step = 0;
  while (step < nsweep) {
    if (KLFT_VERBOSITY > 0) {
      const auto t0 = std::chrono::steady_clock::now();
      const real_t acc = sweep_Metropolis<rank, Nc>(g, mp, rng);
      const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      printf("Step: %ld, Acceptance rate: %f, Time: %f\n", (long)step, acc, t);
      measureGaugeObservables<rank, Nc>(g, gp, step);
      ++step;
      continue;
    }
    // next step that measures (or the end)
    size_t stop = nsweep;
    if (gp.measurement_interval > 0) {
      size_t k = step + 1;
      const size_t m = ((k + gp.measurement_interval - 1) / gp.measurement_interval) * gp.measurement_interval;
      stop = std::min(nsweep, m + 1);
      if (step == 0 && stop > nsweep) stop = nsweep;
    }
    klft_metropolis_params p{mp.nHits, mp.beta, mp.delta, rng.seed};
    check(klft_metropolis_run(g.h, &p, rng.sweep, (int)(stop - step), nullptr));
    rng.sweep += stop - step;
    measureGaugeObservables<rank, Nc>(g, gp, stop - 1);
    step = stop;
  }
  flushAllGaugeObservables(gp);
  return 0;
}

// parseInputFile (InputParser.hpp:33-125)
inline bool parseInputFile(const std::string &filename, MetropolisParams &mp) {
  try {
    const klft_yaml::Node config = klft_yaml::LoadFile(filename);
    if (config["MetropolisParams"]) {
      const auto &m = config["MetropolisParams"];
      mp.Ndims = m["Ndims"].as<index_t>(4);
      mp.L0 = m["L0"].as<index_t>(32);
      mp.L1 = m["L1"].as<index_t>(32);
      mp.L2 = m["L2"].as<index_t>(32);
      mp.L3 = m["L3"].as<index_t>(32);
      mp.nHits = m["nHits"].as<index_t>(10);
      mp.nSweep = m["nSweep"].as<index_t>(1000);
      mp.seed = m["seed"].as<index_t>(1234);
      mp.Nd = m["Nd"].as<size_t>(4);
      mp.Nc = m["Nc"].as<size_t>(2);
      mp.beta = m["beta"].as<double>(1.0);
      mp.delta = m["delta"].as<double>(0.1);
      const std::string law = m["proposal"].as<std::string>("reference");
      if (law == "symmetric") {
        mp.proposal = KLFT_PROPOSAL_SYMMETRIC;
      } else if (law == "reference") {
        mp.proposal = KLFT_PROPOSAL_REFERENCE;
      } else {
        printf("Error parsing input file: unknown proposal %s\n", law.c_str());
        return false;
      }
    } else {
      printf("Error: MetropolisParams not found in input file\n");
      return false;
    }
    return true;
  } catch (const klft_yaml::YamlError &e) {
    printf("Error parsing input file: %s\n", e.what());
    return false;
  }
}

inline bool parseInputFile(const std::string &filename, GaugeObservableParams &gp) {
  try {
    const klft_yaml::Node config = klft_yaml::LoadFile(filename);
    if (config["GaugeObservableParams"]) {
      const auto &g = config["GaugeObservableParams"];
      gp.measurement_interval = g["measurement_interval"].as<size_t>(0);
      gp.measure_plaquette = g["measure_plaquette"].as<bool>(false);
      gp.measure_wilson_loop_temporal = g["measure_wilson_loop_temporal"].as<bool>(false);
      gp.measure_wilson_loop_mu_nu = g["measure_wilson_loop_mu_nu"].as<bool>(false);
      auto pairs = [&](const char *key, std::vector<Array<index_t, 2>> &dst) {
        if (g[key])
          for (const auto &pr : g[key]) dst.push_back(IndexArray<2>{pr[0].as<index_t>(), pr[1].as<index_t>()});
      };
      pairs("W_temp_L_T_pairs", gp.W_temp_L_T_pairs);
      pairs("W_mu_nu_pairs", gp.W_mu_nu_pairs);
      pairs("W_Lmu_Lnu_pairs", gp.W_Lmu_Lnu_pairs);
      gp.plaquette_filename = g["plaquette_filename"].as<std::string>("");
      gp.W_temp_filename = g["W_temp_filename"].as<std::string>("");
      gp.W_mu_nu_filename = g["W_mu_nu_filename"].as<std::string>("");
      gp.write_to_file = g["write_to_file"].as<bool>(false);
    } else {
      printf("Error: GaugeObservableParams not found in input file\n");
      return false;
    }
    return true;
  } catch (const klft_yaml::YamlError &e) {
    printf("Error parsing input file: %s\n", e.what());
    return false;
  }
}

// klft::Metropolis (klft.hpp:5, lib/Metropolis.cpp:36-166)
inline int Metropolis(const std::string &input_file) { return klft_metropolis_main(input_file.c_str()); }

// HMC over a gauge field (HMC.hpp:48-71 with the leapfrog of
// Integrator.hpp:26-53); hmc_step returns whether the trajectory was accepted
template <size_t Nd, size_t Nc>
class HMC {
 public:
  HMC(deviceGaugeField<Nd, Nc> &g, const HMC_Params &params, real_t beta, uint64_t seed) {
    klft_hmc_params p{beta, params.tau, (int)params.n_steps, seed};
    check(klft_hmc_create(g.h, &p, &h_));
  }
  ~HMC() {
    if (h_) klft_hmc_destroy(h_);
  }
  HMC(const HMC &) = delete;
  HMC &operator=(const HMC &) = delete;
  bool hmc_step() {
    int acc = 0;
    check(klft_hmc_trajectory(h_, traj_++, &delta_H_, &acc));
    return acc != 0;
  }
  real_t get_delta_H() const { return delta_H_; }

 private:
  klft_hmc_t h_ = nullptr;
  uint64_t traj_ = 0;
  real_t delta_H_ = 0.0;
};

}  // namespace klft
