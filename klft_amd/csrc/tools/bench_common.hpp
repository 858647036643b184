// Shared pieces of the benchmark drivers (benchmarks/Gauge/*.cpp): option
// parsing (-n N), the min-over-20 timer (Tuner.hpp:31-33) and the banner.
#pragma once
#include <getopt.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <string>

#include "klft_amd.hpp"

#define HLINE "====================================================================\n"
constexpr int STREAM_NTIMES = 20;

inline int parse_n(int argc, char **argv, size_t &n) {
  n = 32;
  const std::string help =
      "  -n <N>, --nelements <N>\n"
      "     Length of the gauge field dimensions.\n"
      "     Default: 32\n"
      "  -h, --help\n"
      "     Prints this message.\n";
  static struct option opts[] = {
      {"nelements", required_argument, NULL, 'n'}, {"help", no_argument, NULL, 'h'}, {NULL, 0, NULL, 0}};
  int c, idx = 0;
  while ((c = getopt_long(argc, argv, "n:h", opts, &idx)) != -1) switch (c) {
      case 'n': n = std::atoi(optarg); break;
      case 'h': printf("%s", help.c_str()); return -2;
      case 0: break;
      default: printf("%s", help.c_str()); return -1;
    }
  return 0;
}

template <class F>
double min_time(F &&fn) {
  double best = std::numeric_limits<double>::max();
  for (int k = 0; k < STREAM_NTIMES; ++k) {
    const auto t0 = std::chrono::steady_clock::now();
    fn();
    best = std::min(best, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  }
  return best;
}
