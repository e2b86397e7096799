// fp8 OCP E4M3FN (bias 7, no infinities, 0x7F/0xFF NaN, largest finite 448)
// encoding for the near-field preconditioner's fp8 copy: round to nearest,
// ties to even, subnormals kept (m 2^-9), magnitudes from 448 up (and NaN)
// saturated to 448.  The same code runs on the host (the encoder's test) and
// in the encoding kernel; decoding is the hardware's v_cvt_pk_f32_fp8.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace spk {

__host__ __device__ inline uint32_t f32_to_e4m3(float x) {
  const uint32_t sg = x < 0.f ? 0x80u : 0u;
  const float a = fabsf(x);
  if (!(a < 448.f)) return sg | 0x7Eu;
  if (a < 0.015625f) return sg | (uint32_t)rintf(a * 512.f);  // 8 -> 0x08, the least normal
  int E;
  const float f = frexpf(a, &E);  // a = f 2^E, f in [0.5, 1)
  int e = E - 1;
  uint32_t mnt = (uint32_t)rintf((f * 2.f - 1.f) * 8.f);
  if (mnt == 8u) {
    mnt = 0u;
    ++e;
  }
  if (e + 7 > 15 || (e + 7 == 15 && mnt == 7u)) return sg | 0x7Eu;
  return sg | ((uint32_t)(e + 7) << 3) | mnt;
}

}  // namespace spk
