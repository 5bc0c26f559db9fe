// Stray fields between grids: box embedding around the demag convolution,
// and the direct sum (see stray.hpp).
#include "stray.hpp"

#include <algorithm>

namespace mpa {

namespace {

constexpr int BLOCK = 256;
constexpr double MU0 = 1.25663706212e-6;  // src/physics/constants.hpp:22, as in the demag plan

struct SourceList {
  StraySource s[MAX_STRAY_SOURCES];
  int n;
};

unsigned blocks_for(long long n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

// M (3 x box cells) = the sum of the sources' Ms m over the box, zero where no
// source (or no source geometry) covers a cell.
__global__ void __launch_bounds__(BLOCK) k_embed(SourceList L, int bx, int by, int bz, int ox, int oy, int oz,
                                                 float* __restrict__ M) {
  const long long nb = (long long)bx * by * bz;
  const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
  if (i >= nb) return;
  const int x = (int)(i % bx), y = (int)((i / bx) % by), z = (int)(i / ((long long)bx * by));
  const int w[3] = {ox + x, oy + y, oz + z};
  float v[3] = {0.f, 0.f, 0.f};
  for (int k = 0; k < L.n; ++k) {
    const StraySource& S = L.s[k];
    int l[3];
    bool in = true;
    for (int a = 0; a < 3; ++a) {
      l[a] = w[a] - S.origin[a];
      in = in && l[a] >= 0 && l[a] < S.n[a];
    }
    if (!in) continue;
    const int ns = S.n[0] * S.n[1] * S.n[2];
    const int cell = (l[2] * S.n[1] + l[1]) * S.n[0] + l[0];
    if (S.geom && !S.geom[cell]) continue;
    const float ms = S.msat.at(cell);
    for (int c = 0; c < 3; ++c) v[c] += ms * S.m[(size_t)c * ns + cell];
    if (S.m2) {
      const float ms2 = S.msat2.at(cell);
      for (int c = 0; c < 3; ++c) v[c] += ms2 * S.m2[(size_t)c * ns + cell];
    }
  }
  for (int c = 0; c < 3; ++c) M[(size_t)c * nb + i] = v[c];
}

// out (3 x target cells) = H at the target's cells of the box.
__global__ void __launch_bounds__(BLOCK) k_extract(const float* __restrict__ H, int bx, int by, int bz, int ox,
                                                   int oy, int oz, StrayTarget t, int accumulate) {
  const int nd = t.n[0] * t.n[1] * t.n[2];
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= nd) return;
  const int x = i % t.n[0], y = (i / t.n[0]) % t.n[1], z = i / (t.n[0] * t.n[1]);
  const long long nb = (long long)bx * by * bz;
  const long long b = ((long long)(t.origin[2] + z - oz) * by + (t.origin[1] + y - oy)) * bx + (t.origin[0] + x - ox);
  const bool in = !t.geom || t.geom[i];
  for (int c = 0; c < 3; ++c) {
    const size_t o = (size_t)c * nd + i;
    const float v = in ? H[(size_t)c * nb + b] : 0.f;
    t.out[o] = accumulate ? t.out[o] + v : v;
  }
}

// h(d) = -mu0 sum_s N(d - s) . M(s) over the source cells, in FP64; the kernel
// grid K (6, kz, ky, kx) holds N at offsets ko + j.
__global__ void __launch_bounds__(BLOCK) k_brute(const float* __restrict__ K, int kx, int ky, int kz, int kox,
                                                 int koy, int koz, StraySource S, StrayTarget t) {
  const int nd = t.n[0] * t.n[1] * t.n[2];
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= nd) return;
  const int dx = t.origin[0] + i % t.n[0], dy = t.origin[1] + (i / t.n[0]) % t.n[1],
            dz = t.origin[2] + i / (t.n[0] * t.n[1]);
  const int ns = S.n[0] * S.n[1] * S.n[2];
  const long long nk = (long long)kx * ky * kz;
  double h[3] = {0.0, 0.0, 0.0};
  for (int sz = 0; sz < S.n[2]; ++sz)
    for (int sy = 0; sy < S.n[1]; ++sy)
      for (int sx = 0; sx < S.n[0]; ++sx) {
        const int cell = (sz * S.n[1] + sy) * S.n[0] + sx;
        if (S.geom && !S.geom[cell]) continue;
        double mv[3];
        const double ms = S.msat.at(cell);
        for (int c = 0; c < 3; ++c) mv[c] = ms * S.m[(size_t)c * ns + cell];
        if (S.m2) {
          const double ms2 = S.msat2.at(cell);
          for (int c = 0; c < 3; ++c) mv[c] += ms2 * S.m2[(size_t)c * ns + cell];
        }
        const long long r = ((long long)(dz - S.origin[2] - sz - koz) * ky + (dy - S.origin[1] - sy - koy)) * kx +
                            (dx - S.origin[0] - sx - kox);
        const double nxx = K[r], nyy = K[nk + r], nzz = K[2 * nk + r];
        const double nxy = K[3 * nk + r], nxz = K[4 * nk + r], nyz = K[5 * nk + r];
        h[0] += nxx * mv[0] + nxy * mv[1] + nxz * mv[2];
        h[1] += nxy * mv[0] + nyy * mv[1] + nyz * mv[2];
        h[2] += nxz * mv[0] + nyz * mv[1] + nzz * mv[2];
      }
  const bool in = !t.geom || t.geom[i];
  for (int c = 0; c < 3; ++c) t.out[(size_t)c * nd + i] = in ? (float)(-MU0 * h[c]) : 0.f;
}

DemagGeometry box_geometry(const int size[3], const double cs[3], const int mastergrid[3], const int pbc[3]) {
  DemagGeometry g;
  for (int a = 0; a < 3; ++a) {
    g.n[a] = size[a];
    g.cs[a] = cs[a];
    g.mastergrid[a] = mastergrid[a];
    g.pbc[a] = pbc[a];
  }
  return g;
}

}  // namespace

void stray_box(const StraySource* src, int nsrc, const StrayTarget& dst, int origin[3], int size[3]) {
  for (int a = 0; a < 3; ++a) {
    int lo = dst.origin[a], hi = dst.origin[a] + dst.n[a];
    for (int k = 0; k < nsrc; ++k) {
      lo = std::min(lo, src[k].origin[a]);
      hi = std::max(hi, src[k].origin[a] + src[k].n[a]);
    }
    origin[a] = lo;
    size[a] = hi - lo;
  }
}

StrayPlan::StrayPlan(const int box_origin[3], const int box_size[3], const double cs[3], const int mastergrid[3],
                     const int pbc[3])
    : plan_(box_geometry(box_size, cs, mastergrid, pbc)) {
  for (int a = 0; a < 3; ++a) {
    o_[a] = box_origin[a];
    n_[a] = box_size[a];
  }
  const size_t nb = (size_t)n_[0] * n_[1] * n_[2];
  M_.allocate(3 * nb);
  H_.allocate(3 * nb);
}

bool StrayPlan::matches(const int box_origin[3], const int box_size[3]) const {
  for (int a = 0; a < 3; ++a)
    if (o_[a] != box_origin[a] || n_[a] != box_size[a]) return false;
  return true;
}

void StrayPlan::exec(const StraySource* src, int nsrc, const StrayTarget& dst, bool accumulate, hipStream_t s,
                     Profiler* prof) {
  if (nsrc < 1 || nsrc > MAX_STRAY_SOURCES) throw Error(MPA_ERR_INVALID, "stray field: 1 to 16 source magnets");
  for (int a = 0; a < 3; ++a) {
    for (int k = 0; k < nsrc; ++k)
      if (src[k].origin[a] < o_[a] || src[k].origin[a] + src[k].n[a] > o_[a] + n_[a])
        throw Error(MPA_ERR_RUNTIME, "stray field: source outside the plan's box");
    if (dst.origin[a] < o_[a] || dst.origin[a] + dst.n[a] > o_[a] + n_[a])
      throw Error(MPA_ERR_RUNTIME, "stray field: destination outside the plan's box");
  }
  SourceList L{};
  L.n = nsrc;
  for (int k = 0; k < nsrc; ++k) L.s[k] = src[k];
  const long long nb = (long long)n_[0] * n_[1] * n_[2];
  {
    Timed t(prof, "k_stray_embed", 12.0 * (double)nb * 2, s);
    k_embed<<<blocks_for(nb), BLOCK, 0, s>>>(L, n_[0], n_[1], n_[2], o_[0], o_[1], o_[2], M_.get());
    checkLaunch("k_embed");
  }
  plan_.exec(M_.get(), PS{nullptr, 1.f}, nullptr, H_.get(), s, prof);
  const long long nd = (long long)dst.n[0] * dst.n[1] * dst.n[2];
  {
    Timed t(prof, "k_stray_extract", (accumulate ? 36.0 : 24.0) * (double)nd, s);
    k_extract<<<blocks_for(nd), BLOCK, 0, s>>>(H_.get(), n_[0], n_[1], n_[2], o_[0], o_[1], o_[2], dst,
                                               accumulate ? 1 : 0);
    checkLaunch("k_extract");
  }
}

void stray_brute(const StraySource& src, const StrayTarget& dst, const double cs[3], const int mastergrid[3],
                 const int pbc[3], hipStream_t s) {
  // offsets d - s over the two grids
  int ksize[3], korigin[3];
  for (int a = 0; a < 3; ++a) {
    ksize[a] = dst.n[a] + src.n[a] - 1;
    korigin[a] = dst.origin[a] - (src.origin[a] + src.n[a] - 1);
  }
  const long long nk = (long long)ksize[0] * ksize[1] * ksize[2];
  DeviceArray<float> K((size_t)6 * nk);
  const int dummy[3] = {1, 1, 1};
  newell_grid_device(K.get(), ksize, korigin, box_geometry(dummy, cs, mastergrid, pbc), s);
  const long long nd = (long long)dst.n[0] * dst.n[1] * dst.n[2];
  k_brute<<<blocks_for(nd), BLOCK, 0, s>>>(K.get(), ksize[0], ksize[1], ksize[2], korigin[0], korigin[1],
                                           korigin[2], src, dst);
  checkLaunch("k_brute");
  HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace mpa
