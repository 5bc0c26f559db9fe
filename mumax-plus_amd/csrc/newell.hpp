// FP64 Newell demag tensor, device side.
//
// Indefinite integrals of Newell, Williams & Dunlop, JGR 98 (1993) 9551,
// eqs. 27 (Nxx) and 32 (Nxy); the 27-point weighted sum with weights
// 8/-4/2/-1 and the component permutations follow the reference
// (src/physics/newell.cu:7-139).  Lattice coordinates are formed in double.
#pragma once

#include <hip/hip_runtime.h>

namespace mpa {

__device__ __forceinline__ double newell_fxx(int ix, int iy, int iz, double cx, double cy, double cz) {
  const double x = ix * cx, y = iy * cy, z = iz * cz;
  const double R = sqrt(x * x + y * y + z * z);
  double r = 0.0;
  if (ix != 0 || iz != 0) r += (y / 2) * (z * z - x * x) * asinh(y / sqrt(x * x + z * z));
  if (ix != 0 || iy != 0) r += (z / 2) * (y * y - x * x) * asinh(z / sqrt(x * x + y * y));
  if (ix != 0) r -= x * y * z * atan((y * z) / (x * R));
  r += (1.0 / 6.0) * (2 * x * x - y * y - z * z) * R;
  return r;
}

__device__ __forceinline__ double newell_fxy(int ix, int iy, int iz, double cx, double cy, double cz) {
  if (iy == 0 || ix == 0) return 0.0;
  const double x = ix * cx, y = iy * cy, z = iz * cz;
  const double R = sqrt(x * x + y * y + z * z);
  double r = 0.0;
  r += (x * y * z) * asinh(z / sqrt(x * x + y * y));
  r += (y / 6) * (3 * z * z - y * y) * asinh(x / sqrt(y * y + z * z));
  r += (x / 6) * (3 * z * z - x * x) * asinh(y / sqrt(x * x + z * z));
  r -= (z * y * y / 2) * atan((x * z) / (y * R));
  r -= (z * x * x / 2) * atan((y * z) / (x * R));
  r -= x * y * R / 3;
  if (iz != 0) r -= (z * z * z / 6) * atan((x * y) / (z * R));
  return r;
}

__device__ __forceinline__ int newell_weight(int dx, int dy, int dz) {
  switch (dx * dx + dy * dy + dz * dz) {
    case 0: return 8;
    case 1: return -4;
    case 2: return 2;
    case 3: return -1;
    default: return 0;
  }
}

template <bool XY>
__device__ double newell_sum(int ix, int iy, int iz, double cx, double cy, double cz) {
  if (XY && (ix == 0 || iy == 0)) return 0.0;
  double r = 0.0;
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -1; dz <= 1; ++dz) {
        const int w = newell_weight(dx, dy, dz);
        const double f = XY ? newell_fxy(ix + dx, iy + dy, iz + dz, cx, cy, cz)
                            : newell_fxx(ix + dx, iy + dy, iz + dz, cx, cy, cz);
        r += w * f;
      }
  return r / (4 * M_PI * cx * cy * cz);
}

// All six components at lattice offset (ix, iy, iz): xx, yy, zz, xy, xz, yz.
__device__ __forceinline__ void newell6(int ix, int iy, int iz, double cx, double cy, double cz, double out[6]) {
  out[0] = newell_sum<false>(ix, iy, iz, cx, cy, cz);
  out[1] = newell_sum<false>(iy, iz, ix, cy, cz, cx);
  out[2] = newell_sum<false>(iz, ix, iy, cz, cx, cy);
  out[3] = newell_sum<true>(ix, iy, iz, cx, cy, cz);
  out[4] = newell_sum<true>(ix, iz, iy, cx, cz, cy);
  out[5] = newell_sum<true>(iy, iz, ix, cy, cz, cx);
}

}  // namespace mpa
