// Test helper: prints words and doubles of the library's Philox streams
// (klft_amd/csrc/philox.hpp, the same header the kernels use) for comparison
// with the published / rocRAND known answers and with the oracle's streams.
#include <cstdio>
#include <cstdlib>

#include "philox.hpp"

int main(int argc, char **argv) {
  if (argc == 8) {  // words: c0 c1 c2 c3 k0 k1 -> one block
    uint32_t o[4];
    klft_rng::philox4x32_10(strtoul(argv[1], 0, 16), strtoul(argv[2], 0, 16), strtoul(argv[3], 0, 16),
                            strtoul(argv[4], 0, 16), strtoul(argv[5], 0, 16), strtoul(argv[6], 0, 16), o);
    printf("%08x %08x %08x %08x\n", o[0], o[1], o[2], o[3]);
    return 0;
  }
  if (argc == 9) {  // stream doubles: seed tag c0 c1 c2 chain mu n
    klft_rng::Stream s(strtoull(argv[1], 0, 10), atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), atoi(argv[5]),
                       atoi(argv[6]), atoi(argv[7]));
    const int n = atoi(argv[8]);
    for (int i = 0; i < n; ++i) printf("%.17g\n", s.u01());
    return 0;
  }
  return 1;
}
