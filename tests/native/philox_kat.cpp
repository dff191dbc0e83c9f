// CPU check of the renderer's counter-based generator (philox and u01 in
// renderx_amd/csrc/rx_dev.h): the published Philox4x32-10 known-answer
// vectors (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
// 1, 2, 3", SC'11; the Random123 kat_vectors file), then independence of the
// streams the renderer derives from it: key = (pixel, seed key), counter =
// (sample, bounce, stream, 0).  Prints one line per check; exits 1 on a failure.
#include <cmath>
#include <cstdio>

#include "../../renderx_amd/csrc/rx_dev.h"

using namespace rx;

static int fails = 0;

static void kat(U4 c, uint32_t k0, uint32_t k1, U4 want) {
  const U4 r = philox(c, k0, k1);
  const bool ok = r.x == want.x && r.y == want.y && r.z == want.z && r.w == want.w;
  printf("kat %08x %08x %08x %08x / %08x %08x -> %08x %08x %08x %08x %s\n", c.x, c.y, c.z, c.w, k0, k1, r.x, r.y,
         r.z, r.w, ok ? "ok" : "FAIL");
  fails += !ok;
}

int main() {
  kat(U4{0u, 0u, 0u, 0u}, 0u, 0u, U4{0x6627e8d5u, 0xe169c58du, 0xbc57ac4cu, 0x9b00dbd8u});
  kat(U4{0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu}, 0xffffffffu, 0xffffffffu,
      U4{0x408f276du, 0x41c83b0eu, 0xa20bc7c6u, 0x6d5451fdu});
  kat(U4{0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u}, 0xa4093822u, 0x299f31d0u,
      U4{0xd16cfe09u, 0x94fdccebu, 0x5001e420u, 0x24126ea1u});

  // u01: the reference's 24-bit float form, in [0, 1 - 2^-24]
  const bool u_ok = u01(0u) == 0.0f && u01(0xffffffffu) == 1.0f - 0x1p-24f && u01(0x80000000u) == 0.5f;
  printf("u01 range %s\n", u_ok ? "ok" : "FAIL");
  fails += !u_ok;

  // Streams of neighbouring pixels, samples and bounces: means, variances and
  // the correlation of each stream's first draw with its neighbour's.
  const uint32_t key1 = 0x9E3779B9u * 7u;
  const int NP = 4096, NS = 64;
  double s = 0, s2 = 0, cpix = 0, csmp = 0, cb = 0;
  long n = 0;
  int hist[16] = {0};
  for (int p = 0; p < NP; p++)
    for (int k = 0; k < NS; k++) {
      const float a = u01(philox(U4{(uint32_t)k, 0u, 0u, 0u}, (uint32_t)p, key1).x) - 0.5f;
      const float b = u01(philox(U4{(uint32_t)k, 0u, 0u, 0u}, (uint32_t)p + 1u, key1).x) - 0.5f;
      const float c = u01(philox(U4{(uint32_t)k + 1u, 0u, 0u, 0u}, (uint32_t)p, key1).x) - 0.5f;
      const float d = u01(philox(U4{(uint32_t)k, 1u, 0u, 0u}, (uint32_t)p, key1).x) - 0.5f;
      s += a;
      s2 += a * a;
      cpix += a * b;
      csmp += a * c;
      cb += a * d;
      hist[(int)((a + 0.5f) * 16)]++;
      n++;
    }
  const double mean = s / n, var = s2 / n, se = std::sqrt(1.0 / 12.0 / n);
  const double r_pix = cpix / n * 12.0, r_smp = csmp / n * 12.0, r_b = cb / n * 12.0;
  double chi2 = 0;
  for (int i = 0; i < 16; i++) chi2 += (hist[i] - n / 16.0) * (hist[i] - n / 16.0) / (n / 16.0);
  // 5-sigma bounds; chi2 with 15 degrees of freedom below 45 (p ~ 1e-4)
  const bool ind = std::fabs(mean) < 5 * se && std::fabs(var - 1.0 / 12.0) < 5 * 0.075 / std::sqrt((double)n) &&
                   std::fabs(r_pix) < 5 / std::sqrt((double)n) && std::fabs(r_smp) < 5 / std::sqrt((double)n) &&
                   std::fabs(r_b) < 5 / std::sqrt((double)n) && chi2 < 45;
  printf("streams n=%ld mean %+.2e var %.6f corr pixel %+.2e sample %+.2e bounce %+.2e chi2 %.1f %s\n", n, mean,
         var, r_pix, r_smp, r_b, chi2, ind ? "ok" : "FAIL");
  fails += !ind;
  return fails ? 1 : 0;
}
