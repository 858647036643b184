#!/bin/sh
# Writes philox_kat.txt: Philox4x32-10 outputs computed by rocRAND's header
# implementation (readable source under /opt/rocm/include), compiled here for
# the host.  Used to pin both the oracle's and the library's Philox.
set -e
here=$(cd "$(dirname "$0")" && pwd)
mkdir -p "$here/_bin"
cat > "$here/_bin/kat.cpp" <<'CPP'
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>
// ten_rounds is a protected member of the engine; expose it
struct Engine : rocrand_device::philox4x32_10_engine {
  uint4 rounds(uint4 c, uint2 k) { return ten_rounds(c, k); }
};
int main() {
  const unsigned ctr[][4] = {{0, 0, 0, 0}, {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
                             {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u},
                             {1, 2, 3, 4}, {7, 0, 0x10000000u, 0xdeadbeefu}};
  const unsigned key[][2] = {{0, 0}, {0xffffffffu, 0xffffffffu}, {0xa4093822u, 0x299f31d0u},
                             {5, 6}, {32091, 0}};
  for (int i = 0; i < 5; ++i) {
    Engine e;
    uint4 c = {ctr[i][0], ctr[i][1], ctr[i][2], ctr[i][3]};
    uint2 k = {key[i][0], key[i][1]};
    uint4 r = e.rounds(c, k);
    printf("%08x %08x %08x %08x  %08x %08x  %08x %08x %08x %08x\n", c.x, c.y, c.z, c.w, k.x, k.y, r.x, r.y,
           r.z, r.w);
  }
  return 0;
}
CPP
/opt/rocm/bin/hipcc -x c++ -O1 -I/opt/rocm/include -D__HIP_PLATFORM_AMD__ -o "$here/_bin/kat" "$here/_bin/kat.cpp" 2>/dev/null
"$here/_bin/kat" > "$here/philox_kat.txt"
