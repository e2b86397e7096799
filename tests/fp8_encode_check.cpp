// Host-only check of spk::f32_to_e4m3 (specklesim_amd/csrc/spk_fp8.h): every
// input's code must decode to the representable E4M3FN value nearest to it
// (ties to the even code), found by brute force over the 127 finite
// magnitudes, decoded from the format's definition: normal (1 + m/8) 2^(e-7),
// subnormal (m/8) 2^-6, largest finite 448.  Inputs: exhaustive walks over
// float bit patterns below 448 (every 4096th) and around every code's
// midpoints, plus saturation and signed zero.  Prints "mismatches <n>".
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "spk_fp8.h"

static double decode(uint32_t c) {
  const int e = (c >> 3) & 15, m = c & 7;
  const double v = e ? std::ldexp(1.0 + m / 8.0, e - 7) : std::ldexp(m / 8.0, -6);
  return (c & 0x80u) ? -v : v;
}

static long bad = 0, checked = 0;

static void check(float x) {
  const uint32_t c = spk::f32_to_e4m3(x);
  ++checked;
  if ((c & 0x7Fu) == 0x7Fu) {
    if (bad++ < 10) std::printf("x=%.9g gives NaN code 0x%02x\n", x, c);
    return;
  }
  const double ax = std::fabs((double)x);
  int best = 0;
  double bd = 1e300;
  for (int k = 0; k <= 0x7E; ++k) {
    const double d = std::fabs(decode(k) - ax);
    if (d < bd || (d == bd && (k & 1) == 0)) {
      bd = d;
      best = k;
    }
  }
  const uint32_t want = (x < 0.f ? 0x80u : 0u) | (uint32_t)best;
  if (c != want && !(x == 0.f && (c & 0x7Fu) == 0u)) {
    if (bad++ < 10) std::printf("x=%.9g got 0x%02x want 0x%02x\n", x, c, want);
  }
}

int main() {
  for (uint32_t u = 0; u < 0x43E00000u; u += 4096) {  // 0x43E00000 = 448.0f
    float y;
    std::memcpy(&y, &u, 4);
    check(y);
    check(-y);
  }
  for (int k = 0; k < 0x7E; ++k) {  // each pair of neighbouring codes: midpoint and its neighbours
    const float mid = (float)((decode(k) + decode(k + 1)) / 2);
    check(mid);
    check(std::nextafterf(mid, 0.f));
    check(std::nextafterf(mid, 1e9f));
    check((float)decode(k));
  }
  for (float x : {448.f, 449.f, 464.f, 480.f, 1e6f, 3.4e38f, -448.f, -1e30f, 0.f, -0.f})
    check(x);
  const uint32_t sat = spk::f32_to_e4m3(1e30f), nsat = spk::f32_to_e4m3(-1e30f);
  if (sat != 0x7Eu || nsat != 0xFEu) ++bad;
  std::printf("checked %ld mismatches %ld\n", checked, bad);
  return bad != 0;
}
