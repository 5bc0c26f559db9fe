// Stray field of one or more magnets on another grid.
//
// Replaces StrayField / StrayFieldFFTExecutor between different grids
// (src/physics/strayfield.cpp, strayfieldfft.cu:149-246 with the kernel of
// strayfieldkernel.cu:80-99 for a source and a destination grid) and
// StrayFieldBruteExecutor (strayfieldbrute.cu:15-88).
//
// The FFT form uses the identity the reference's own test states
// (test/test_strayfield.py:86-136): the field of sources inside a box,
// evaluated on a destination inside the same box, is the box's demag field of
// the magnetisation that is the sources' inside them and zero elsewhere.  The
// box is the bounding grid of the sources and the destination; its
// convolution is a DemagPlan (same kernel, same periodic images).
#pragma once

#include "common.hpp"
#include "demag.hpp"
#include "profiler.hpp"

namespace mpa {

constexpr int MAX_STRAY_SOURCES = 16;

// One source magnet: M = Ms m (+ Ms2 m2 for the two sublattices of an
// antiferromagnet), zero outside its geometry.  origin and n in cells of the
// world grid.
struct StraySource {
  const float* m;
  PS msat;
  const float* m2;
  PS msat2;
  const uint8_t* geom;
  int origin[3];
  int n[3];
};

// A destination window: cells origin .. origin + n - 1 of the world grid;
// out holds 3 * n0 n1 n2 floats (SoA); cells outside geom (when given) are
// zero.
struct StrayTarget {
  int origin[3];
  int n[3];
  const uint8_t* geom;
  float* out;
};

class StrayPlan {
 public:
  StrayPlan(const int box_origin[3], const int box_size[3], const double cs[3], const int mastergrid[3],
            const int pbc[3]);
  bool matches(const int box_origin[3], const int box_size[3]) const;
  // out = (or, with accumulate, out +=) the field of the sources on the
  // target, in tesla.
  void exec(const StraySource* src, int nsrc, const StrayTarget& dst, bool accumulate, hipStream_t s,
            Profiler* prof);

 private:
  int o_[3], n_[3];
  DemagPlan plan_;
  DeviceArray<float> M_, H_;
};

// Bounding grid of the sources and the destination window.
void stray_box(const StraySource* src, int nsrc, const StrayTarget& dst, int origin[3], int size[3]);

// Direct sum over the source cells with the kernel evaluated at every
// destination-source offset (FP64, stored as float), for tests and small
// grids: the reference's "brute" method.
void stray_brute(const StraySource& src, const StrayTarget& dst, const double cs[3], const int mastergrid[3],
                 const int pbc[3], hipStream_t s);

}  // namespace mpa
