// Local effective-field terms, LLG torque and energy densities of a
// ferromagnet, fused into one stencil kernel.
#pragma once

#include <vector>

#include "common.hpp"

namespace mpa {

enum TermBits : unsigned {
  TERM_EXCH = 1u << 0,
  TERM_UNI = 1u << 1,
  TERM_CUBIC = 1u << 2,
  TERM_EXT = 1u << 3,
  TERM_DMI = 1u << 4,
  TERM_DEMAG = 1u << 5,
  TERM_ME = 1u << 6,         // magnetoelastic field
  TERM_AFM_SITE = 1u << 7,   // homogeneous (same-cell) antiferromagnetic exchange
  TERM_AFM_NN = 1u << 8,     // inhomogeneous (neighbour) antiferromagnetic exchange
  TERM_ZHANGLI = 1u << 9,    // Zhang-Li spin-transfer torque
  TERM_SLONC = 1u << 10,     // Slonczewski spin-transfer torque
  TERM_THERMAL = 1u << 11,   // thermal field (temperature not zero)
};
constexpr unsigned TERM_STT = TERM_ZHANGLI | TERM_SLONC;

// DMI tensor entries in the member order of DmiTensor (src/physics/dmitensor.hpp:12-22).
enum DmiIdx { D_XXY = 0, D_XYZ, D_XXZ, D_YXY, D_YYZ, D_YXZ, D_ZXY, D_ZYZ, D_ZXZ, D_COUNT };

// Everything the stencil needs, passed to kernels by value.
struct FMDesc {
  int nx, ny, nz, ncell;
  int ox, oy, oz;     // magnet grid origin (cells)
  int mgx, mgy, mgz;  // mastergrid size (0 = not periodic)
  int mox, moy, moz;  // mastergrid origin
  float cx, cy, cz;
  const uint8_t* geom;
  PS msat, aex, alpha, ku1, ku2, kc1, kc2, kc3, frozen;
  PV anisU, anisC1, anisC2, bias;
  PS dmi[D_COUNT];
  float wb[3];  // world bias field (T)
  int openbc;
  unsigned terms;
  // magnetoelastic coupling (magnetoelasticfield.cu:44-103): dynamic strain
  // (6N: xx yy zz xy xz yz) of the magnet owning the displacement, or a rigid
  // applied strain (normal xx yy zz, shear xy xz yz) when me_rigid is set
  PS B1, B2;
  const float* strain;
  PV rnorm, rshear;
  int me_rigid;
  // antiferromagnet sublattice: sister magnetisation and msat, host
  // parameters and the inter-sublattice DMI tensor
  int sublattice;
  const float* m2;
  PS msat2, afmex_cell, afmex_nn, latcon;
  PS hdmi[D_COUNT];
  // stray field of the other magnets (3N, tesla), part of the external field
  // (zeeman.cu:70-88); nullptr when none
  const float* hstray;
  // spin-transfer torques (stt.cu:27-195)
  PS pol, xi, lambda, epsp, fthick;
  PV jcur, fixed;
  int fixed_on_top;
  // regions (system.hpp:79-99): per-cell region index, or nullptr when the
  // regions change no term (no regions, or default inter-region values);
  // exchange between cells of different regions (exchange.cu:79-99,
  // afmexchange.cu:116-143)
  const uint32_t* regions;
  IP inter_ex, scale_ex;    // this ferromagnet's inter_exchange / scale_exchange
  IP inter_afm, scale_afm;  // the host's inter_afmex_nn / scale_afmex_nn
  // thermal field (thermalnoise.cu:21-70): per cell sqrt(pref alpha T /
  // ((1 + alpha^2) Ms)) times three normal deviates of (seed, cell, purpose,
  // draw); added to the torque times noise_scale = 1/sqrt(dt) when `noise`
  // (rungekutta.cpp:95-106)
  PS temperature;
  int noise;
  float noise_pref, noise_scale;
  unsigned noise_seed, noise_purpose;
  unsigned long long noise_draw;
  // uniform path: z planes per block of the LDS halo-tile kernel, 0 for the
  // one-cell-per-thread kernel
  int tile_zc;
};

enum FieldMode {
  MODE_FIELD = 0,    // out = sum of the selected terms (T)
  MODE_LLG = 1,      // out = LLG torque of the selected terms (1/s)
  MODE_DAMPING = 2,  // out = damping-only torque (relax)
  MODE_EDENS = 3,    // out (1 comp) = energy density of the selected terms (J/m3)
  MODE_STT = 4,      // out = spin-transfer torque (1/s)
  MODE_NOISE = 5,    // out = thermal field quantity (thermalnoise.cu:46-69)
};

// out may alias nothing.  hdemag (3N) is added when TERM_DEMAG is set.
void launch_field(const FMDesc& d, const float* m, const float* hdemag, float* out, FieldMode mode, hipStream_t s);

// x_next = normalise(x0 + sum_j c_j k_j + cself * k) at the cell, zero
// outside the geometry, with k the derivative just computed: the stage
// combination of the Runge-Kutta stepper fused into the field kernel that
// produces the last derivative it needs.  Same arithmetic, term order and
// normalisation as k_combine_batch.  x0 == nullptr takes the cell's own
// state as x0 (the first stage of the next step, whose x0 is the state the
// field was evaluated at).
struct NextStage {
  const float* x0;
  const float* k[7];
  float c[7];
  float cself;
  int nk;
  float* out;
  int N;
  __device__ __forceinline__ void operator()(int idx, float3 mc, float3 r, bool in) const {
    apply<-1>(idx, mc, r, in);
  }
  // NK >= 0: nk == NK known at compile time, so every load is issued before
  // the sums (NK < 0: nk read at run time).
  template <int NK>
  __device__ __forceinline__ void apply(int idx, float3 mc, float3 r, bool in) const {
    float v[3] = {mc.x, mc.y, mc.z};
    const float rr[3] = {r.x, r.y, r.z};
#pragma unroll
    for (int c3 = 0; c3 < 3; ++c3) {
      float a = x0 ? x0[(size_t)c3 * N + idx] : v[c3];
#pragma unroll
      for (int j = 0; j < (NK >= 0 ? NK : 7); ++j) {
        if (NK < 0 && j >= nk) break;
        a += c[j] * k[j][(size_t)c3 * N + idx];
      }
      a += cself * rr[c3];
      v[c3] = a;
    }
    const float inv = rsqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
#pragma unroll
    for (int c3 = 0; c3 < 3; ++c3) out[(size_t)c3 * N + idx] = in ? v[c3] * inv : 0.f;
  }
};

// launch_field on the uniform-parameter LLG path that also writes the next
// Runge-Kutta stage state (NextStage above); field_next_fusable
// tells whether a magnet is on that path.
bool field_next_fusable(const FMDesc& d, FieldMode mode);
void launch_field_next(const FMDesc& d, const float* m, const float* hdemag, float* out, FieldMode mode,
                       const NextStage& next, hipStream_t s);

// The tile-schedule autotuner's record of the uniform field kernel (z planes
// per workgroup chosen per grid size and kernel variant when FMDesc.tile_zc
// is -1): one line per grid, the choice and the microseconds per launch of
// each candidate; and its reset.
std::string field_tile_report();
void field_tile_clear();

// launch_field for the two sublattices of an antiferromagnet in one launch
// (two separate launches when their parameter layouts differ).
void launch_field_pair(const FMDesc& d0, const float* m0, const float* h0, float* o0, const FMDesc& d1,
                       const float* m1, const float* h1, float* o1, FieldMode mode, hipStream_t s);

// The torques launch_field_pair gives in `mode` (MODE_LLG or MODE_DAMPING)
// with demag, from the sublattices' fields without demag (hl0, hl1:
// launch_field_pair in MODE_FIELD with TERM_DEMAG taken out of the
// descriptors' terms) and the demag fields hd0, hd1, for descriptors that
// torque_pair_splittable accepts (uniform parameters on the general
// sublattice path, demag on, no spin-transfer torque, no thermal noise, not
// frozen, no stray field of other magnets).  Lets the field without demag
// run beside the demag passes.
// The two sublattices' fields and torques in one launch (launch_field_pair's
// pair kernel) that also writes each sublattice's next Runge-Kutta stage
// state (NextStage, 0 to 5 earlier derivatives, x0 == nullptr only with
// none: the next step's first stage); fusable: the pair
// takes the one-launch pair kernel in MODE_LLG without thermal noise.
bool field_pair_next_fusable(const FMDesc& d0, const FMDesc& d1, FieldMode mode);
void launch_field_pair_next(const FMDesc& d0, const float* m0, const float* h0, float* o0, const NextStage& n0,
                            const FMDesc& d1, const float* m1, const float* h1, float* o1, const NextStage& n1,
                            FieldMode mode, hipStream_t s);

bool torque_pair_splittable(const FMDesc& d0, const FMDesc& d1, FieldMode mode);
void launch_torque_pair(const FMDesc& d0, const float* m0, const float* hl0, const float* hd0, float* o0,
                        const FMDesc& d1, const float* m1, const float* hl1, const float* hd1, float* o1,
                        FieldMode mode, hipStream_t s);

// Per cell the largest angle (rad) between m and an in-geometry neighbour with
// msat != 0 and a non-zero harmonic-mean exchange constant; 0 outside the
// geometry and where msat = 0 (k_maxangle, exchange.cu:271-313).
void launch_max_angle(const FMDesc& d, const float* m, float* out, hipStream_t s);

// Per-cell values of a parameter with time-dependent terms:
//   out[c N + i] = (base ? base[c N + i] : 0) + add[c] + sum_k coef[k][c] mask[k][c N + i]
// for c < ncomp (dynamic_parameter.cpp:8-29 plus the static field).
struct ParamFill {
  static constexpr int MAX = 8;
  const float* mask[MAX];
  float coef[MAX][3];
  float add[3];
  int nmask;
};
void launch_param_fill(float* out, const float* base, const ParamFill& f, int ncomp, int N, hipStream_t s);

// The field kernels of many ferromagnets with uniform parameters and no
// stray field in one launch.  Their cells are packed: component c of cell i
// of a member is at c * cs + off + i, and every member starts on a multiple
// of 256 cells (one block never spans two members).
bool field_batchable(const FMDesc& d);
struct FieldBatchMember {
  FMDesc d;  // d.geom: the member's own mask or nullptr
  int off;
  int mode;  // MODE_FIELD, MODE_LLG or MODE_DAMPING
  int slot;  // the member's record in BatchNext::st
};

struct BatchMember;
// The next Runge-Kutta stage state of every active batch member, formed in
// the batched field launch from the derivative k it has just computed:
//   x_next = normalise(x0 + sum_{j < nk} (dt a_j) k_j + (dt a_nk) k),
// zero outside the member's geometry, dt the member's own step (each
// coefficient rounded once from double); inactive members are left alone.
// The arithmetic, term order and normalisation of the batch's stage
// combination kernel.
struct BatchNext {
  const float* x0;
  const float* k[7];
  double a[7];  // a[0..nk]: a[nk] multiplies the derivative just computed
  int nk;
  float* xout;
  const BatchMember* st;
};
// The step's error norms, formed in the batched field launch of the last
// stage from its derivative k in registers (the arithmetic of the batch's
// error-norm kernel): per cell of an active member |sum_{j < nk} (dt e_j)
// k_j + (dt e_nk) k| when the member steps adaptively, and |k| (the torque at
// the new state) when it relaxes; each block's maxima merged into the
// member's errbits / tqbits (float bits, atomic max).
struct BatchErr {
  const float* k[7];
  double e[7];  // e[0..nk]: e[nk] multiplies the derivative just computed
  int nk;
  unsigned* errbits;
  unsigned* tqbits;
  const BatchMember* st;
};

class FieldBatch {
 public:
  FieldBatch(const std::vector<FieldBatchMember>& members, int cs);
  // OUT = field or torque (each member's mode) of X, with HD where the
  // member has TERM_DEMAG, for every member; with `next`, also the next
  // stage state of every active member (BatchNext).
  void launch(const float* X, const float* HD, float* OUT, hipStream_t s, const BatchNext* next = nullptr) const;
  // launch without a next stage, forming the error norms (BatchErr).
  void launch_err(const float* X, const float* HD, float* OUT, hipStream_t s, const BatchErr& err) const;
  unsigned nblocks() const { return nblocks_; }
  // The device member table (FBInst entries) and block -> member map.
  const void* table() const { return tab_.get(); }
  const int* block_members() const { return b2i_.get(); }

 private:
  int cs_;
  unsigned nblocks_ = 0;
  DeviceArray<char> tab_;
  DeviceArray<int> b2i_;
};

}  // namespace mpa
