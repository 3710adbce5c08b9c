// transpose32.h -- 32x32 bit transpose across the 32 lanes of a half wave:
// on entry lane l holds word x_l; on exit lane l holds the word whose bit j is
// bit l of x_j.  Five butterfly steps of lane exchanges and masked merges.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

__device__ __forceinline__ uint32_t cuot_transpose32(uint32_t x, int lane) {
  const uint32_t masks[5] = {0x0000FFFFu, 0x00FF00FFu, 0x0F0F0F0Fu, 0x33333333u, 0x55555555u};
#pragma unroll
  for (int s = 0; s < 5; s++) {
    const int j = 16 >> s;
    const uint32_t m = masks[s];
    const uint32_t xp = (uint32_t)__shfl_xor((int)x, j, 64);
    if ((lane & j) == 0)
      x = (x & m) | ((xp & m) << j);
    else
      x = ((xp >> j) & m) | (x & ~m);
  }
  return x;
}
