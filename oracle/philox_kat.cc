// Known-answer generator for Philox4x32-10.  Built from Random123's
// philox.h as vendored in the reference (docking/ising/src/Random123/
// philox.h:206-236) by oracle/Makefile; prints "c0 c1 c2 c3 k0 k1 -> r0 r1
// r2 r3" lines in hex for a fixed list of counters and keys.
#include <cstdio>
#include <cstdint>
#include <Random123/philox.h>

int main()
{
  const uint32_t cases[][6] = {
      {0u, 0u, 0u, 0u, 0u, 0u},
      {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
      {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u},
      {0u, 0u, 0u, 0u, 1u, 0u},
      {5u, 17u, 0u, 0u, 1u, 0u},
      {5u, 17u, 1u, 0u, 1u, 0u},
      {123u, 4999u, 2u, 3u, 7u, 42u},
      {16383u, 4999u, 1u, 0u, 1u, 0u},
      {0x80000000u, 1u, 2u, 3u, 0x9e3779b9u, 0xbb67ae85u},
  };
  r123::Philox4x32 g;
  for (const auto &c : cases) {
    r123::Philox4x32::ctr_type ctr = {{c[0], c[1], c[2], c[3]}};
    r123::Philox4x32::key_type key = {{c[4], c[5]}};
    r123::Philox4x32::ctr_type r = g(ctr, key);
    std::printf("%08x %08x %08x %08x %08x %08x %08x %08x %08x %08x\n", c[0], c[1], c[2],
                c[3], c[4], c[5], r[0], r[1], r[2], r[3]);
  }
  return 0;
}
