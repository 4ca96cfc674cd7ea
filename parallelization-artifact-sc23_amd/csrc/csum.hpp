// Neumaier-compensated FP32 summation (the improved Kahan-Babuska sum: a
// running error term added once at the end), per lane and across a wave.
#pragma once

#include "device_common.hpp"

namespace indigo {

// Neumaier-compensated accumulator: value = s + c.
struct CSum {
  float s = 0.f, c = 0.f;
  IG_DEV void add(float x)
  {
    const float t = s + x;
    c += (fabsf(s) >= fabsf(x)) ? ((s - t) + x) : ((x - t) + s);
    s = t;
  }
  IG_DEV void merge(float s2, float c2)
  {
    add(s2);
    c += c2;
  }
  IG_DEV float value() const { return s + c; }
};

IG_DEV CSum wave_merge(CSum a)
{
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) {
    const float s2 = __shfl_xor(a.s, d, kWave);
    const float c2 = __shfl_xor(a.c, d, kWave);
    a.merge(s2, c2);
  }
  return a;
}

}  // namespace indigo
