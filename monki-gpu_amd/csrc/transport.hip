// MONKI forward photon transport on gfx950: one lane per photon history.
//
// Physics follows Victor Trees's Fortran transport (fortran/monki/upstream/PROGRAM/mc.f90:314-1474
// and its helper modules), with the counter-based random streams and float32 arithmetic of the
// reference GPU kernel (kernels/metal/transport.metal:339-458). Every device routine is a
// template over the arithmetic type so the same code runs in float32 (product) and float64
// (precision reference on identical Philox streams).
//
// Layout on the device: packed float32 tables (layout v1, see monki/execution/hip_tables.py),
// float32 tally banks [round][shard][view][pixel][layer paths | Stokes], per-lane layer paths in LDS.

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "monki_hip.h"
#include "proposal.h"

namespace {

constexpr int kMaxLayers = 64;
constexpr int kMaxSpecies = 4;
constexpr int kMaxViews = 3;
constexpr int kHeaderWords = 72;
constexpr int kLayoutVersion = 2;
constexpr int kBlock32 = 256;
constexpr int kBlock64 = 128;
constexpr int kCounters = 8;
constexpr int kTraceWords = 8;
constexpr uint32_t kSamplingAttempts = 4096;
constexpr int kCdfBuckets = 1024;

enum FailureFlag : unsigned {
  kNonfiniteProbability = 1u,
  kDetectorLimit = 2u,
  kSamplingLimit = 4u,
  kInvalidStokes = 8u,
  kHistoryLimit = 16u,
  kNonfiniteTally = 32u,  // a round's path or weight tally accumulated on the device
};

enum Status : int { kRunning = 0, kEscaped = 1, kAbsorbed = 2, kFailed = 3 };

// Transport modes: plain (the reference algorithm), logging of scattering events for the
// proposal's training, and guided sampling with the learned proposal.
// Detector-directed sampling draws from vMF lobes around the view directions instead of a
// network's mixture.
// kStokes4: the plain tracer with the full Stokes vector (nmat 4: I, Q, U, V). kBackward,
// kBackward4: backward transport (nmat 1/3 and nmat 4). kScats, kScats4: the single/multiple-
// scattering decomposition (calc_radiances_and_scats) in place of the layer paths; kBackScats,
// kBackScats4 the same with backward transport. kSurface, kSurface4: the plain tracer toward
// a detector at the ground. kWide: the plain tracer (float32, forward, TOA, nmat 1 or 3) for
// scenes whose layer rows would leave fewer than three waves per SIMD: LDS holds the lower
// layers' paths of each lane, the upper (uniform) layers' paths are the first flight's whole
// crossings from TOA (a layer index per lane) until the photon has an event up there or flies
// back down through them, then a row of the lane's own in global memory (same values).
enum Mode : int {
  kPlain = 0, kLog = 1, kGuided = 2, kDirected = 3, kStokes4 = 4, kBackward = 5, kBackward4 = 6,
  kScats = 7, kScats4 = 8, kSurface = 9, kSurface4 = 10, kBackScats = 11, kBackScats4 = 12, kWide = 13
};
constexpr bool full_stokes(int mode) {
  return mode == kStokes4 || mode == kBackward4 || mode == kScats4 || mode == kSurface4 || mode == kBackScats4;
}
// kSurface, kSurface4: detector and result grid at the ground (views point down).
constexpr bool downward(int mode) { return mode == kSurface || mode == kSurface4; }
constexpr bool backward(int mode) {
  return mode == kBackward || mode == kBackward4 || mode == kBackScats || mode == kBackScats4;
}
constexpr bool decomposition(int mode) {
  return mode == kScats || mode == kScats4 || mode == kBackScats || mode == kBackScats4;
}

// How the steered modes (guided, detector-directed) sample scattering directions. A steered
// event draws from the physical sampler (phase function, or Lambert at the surface) with
// probability alpha and from the proposal q otherwise, and multiplies the weight by
// p / (alpha p + (1 - alpha) q). Atmospheric events of the scatterer kinds in species_mask
// (bit 0 gas, bit s + 1 tabulated species s) up to scattering order max_order are steered, and
// surface reflections when `surface` is set. The proposal holds lobes around the view
// directions shaped by the bound CDF of tabulated species lobe_species (vMF lobes of
// concentration kappa when it is negative); in the guided mode it is the mixture
// lobe_fraction x lobes + (1 - lobe_fraction) x the network's vMF mixture.
// Weight window (split_above > 0), on the product g of the photon's steering factors (its
// weight over the plain tracer's): a photon whose g exceeds split_above after a scattering
// continues as up to kMaxCopies copies of equal weight (the others resume later on random
// streams of their own); one whose g is below roulette_below before a scattering survives with
// probability g / roulette_weight, its g raised to roulette_weight. Both keep the estimator
// unbiased and bound the products of the steering factors; physical factors (albedos) are
// left alone, as in the plain tracer.
struct Guidance {
  float alpha;
  uint32_t species_mask;
  uint32_t max_order;
  float kappa;
  int lobe_species;
  float lobe_fraction;
  int surface;
  float split_above, roulette_below, roulette_weight;
  // Guided mode: scattering orders up to which the network's mixture takes part in the proposal
  // (beyond them the view lobes alone, and the network does not run).
  uint32_t network_orders;
};

// Unsteered runs.
__host__ __device__ constexpr Guidance unsteered() {
  return Guidance{1.0f, 0u, 0u, 0.0f, -1, 0.0f, 0, 0.0f, 0.0f, 0.0f, 0xFFFFFFFFu};
}

struct SpeciesTables {
  int lut_size;
  float ssalb;
  // Host-derived: trapezoid integral of F11 over cos / 2 (phase-function normalization) and
  // the largest ratio, over node intervals, of the bound F11 (+|F12|) at a node to its
  // interval mean (the constant that makes tabulated sampling exact).
  float norm, rho;
  // Cosines on the equal-angle grid cos(pi (1 - i / (n - 1))): brackets by acos, not search.
  int equal_angle;
  // cdf_start[b], b = 0..kCdfBuckets: the CDF interval holding b / kCdfBuckets (device table
  // built at session creation; null: binary search).
  const uint16_t *cdf_start;
  // cosine[0, n) ascending; cosine[n, 2n) = 1 - cos; cosine[2n, 3n) = 1 + cos.
  const float *cosine, *cdf, *f11, *f22, *f33, *f12;
  const float *f34, *f44;  // nmat 4 only (the two arrays after F12), else null
};

// Scene constants and table pointers, passed by value as a kernel argument.
struct Scene {
  int nx, ny, nz, nmat, views, species, paths, uniform_top, shards;
  int backward;  // backward transport: weight-matrix rows, injection-column pixels, no mu0/mu
  // Scattering decomposition: banks hold 8-value records [single (nmat) | multiple (nmat)],
  // one per view and pixel, then one per view and event cell (level 0 the surface).
  int scats;
  int tally_count;  // output values per round: Stokes | paths | weights
  int stokes_count, path_count;
  // Device banks per round and shard: one record per (view, pixel) holding the layer paths
  // then the Stokes components. Records start every record_stride values (the record length
  // rounded up to 8: 32-byte aligned in float32), which keeps the atomics of one record on the
  // fewest 32-byte sectors; the padding is never written.
  int record_stride;
  int bank_count;
  size_t bank_base;  // first bank element of this scene in a multi-scene session
  float xunit, yunit, depol, theta0, phi0, phi;
  float view[kMaxViews];
  // Geometry derived once on the host in float64 from the float32 angles: launch direction,
  // |cos theta0|, and per view cos, sin, the horizontal direction and mu0/mu.
  double launch_a, launch_b, launch_c, mu0;
  double view_mu[kMaxViews], view_sine[kMaxViews], view_a[kMaxViews], view_b[kMaxViews];
  double view_geometry[kMaxViews], view_inv_mu[kMaxViews];
  const float *zunit, *albedo, *beta, *cumfrac, *ssalb_gas, *height_above, *tau_above;
  // Factored extinction (one tabulated species), else null: beta = species_beta[cell] +
  // gas_beta[layer] and the species fraction = species_beta[cell] / beta, exactly the float32
  // tables. Scenes of a session with equal species fields share one device copy of it.
  const float *species_beta, *gas_beta;
  const float *uniform;  // per layer: 1 if every column has the same extinction and mix
  SpeciesTables sp[kMaxSpecies];
  // Detector and result grid at the surface: views point down (view_mu < 0), the detector ray
  // ends at the ground and registers at its exit column (main.f90:711-745). Host-side choice
  // of the kernel instantiation; last, so that the hot fields keep their offsets.
  int detector_surface;
  // Result grid on the other level than the detector (main.f90:747-757): contributions register
  // where the line of sight through the event meets that level.
  int register_opposite;
};

struct Controls {
  uint32_t count, first, round, seed, max_steps, profile, per_round, rounds;
};

// sin and cos with one argument reduction.
__device__ inline void sin_cos(float x, float *s, float *c) { sincosf(x, s, c); }
__device__ inline void sin_cos(double x, double *s, double *c) { sincos(x, s, c); }


// ---------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al. 2011, Random123 constants). Counter (photon, round, block,
// stream), key (seed, 0): identical to transport.metal:64-92 for stream 0, so a photon's
// draws do not depend on dispatch size, sharding or grouping.
// ---------------------------------------------------------------------------------------
struct Words {
  uint32_t x, y, z, w;
};

__host__ __device__ inline Words philox(Words c, uint32_t k0, uint32_t k1) {
  for (int i = 0; i < 10; ++i) {
    const uint64_t p0 = uint64_t(0xD2511F53u) * c.x;
    const uint64_t p1 = uint64_t(0xCD9E8D57u) * c.z;
    c = Words{uint32_t(p1 >> 32) ^ c.y ^ k0, uint32_t(p1), uint32_t(p0 >> 32) ^ c.w ^ k1,
              uint32_t(p0)};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// Blocks are consumed strictly in order. `next` holds the block after `words` once prefetch
// has run: lanes of a wave use their draws out of phase, so a block generated where a lane
// runs out would run the generator for the whole wave at almost every draw; prefetch at a
// few wave-uniform points instead generates the look-ahead blocks of all lanes together.
struct Stream {
  uint32_t photon, round, seed, stream, block;  // block: the next block to generate
  Words words, next;
  int lane;    // next word of `words` (4: used up)
  bool ready;  // `next` holds block `block - 1`
};

__device__ inline void prefetch(Stream &rng) {
  if (!rng.ready) {
    rng.next = philox(Words{rng.photon, rng.round, rng.block++, rng.stream}, rng.seed, 0u);
    rng.ready = true;
  }
}

// Open (0,1) float32 grid of half steps: ((w >> 9) + 0.5) * 2^-23 (transport.metal:89-91).
__device__ inline float uniform(Stream &rng) {
  if (rng.lane == 4) {
    if (rng.ready) {
      rng.words = rng.next;
      rng.ready = false;
    } else {
      rng.words = philox(Words{rng.photon, rng.round, rng.block++, rng.stream}, rng.seed, 0u);
    }
    rng.lane = 0;
  }
  uint32_t w = rng.lane == 0 ? rng.words.x
               : rng.lane == 1 ? rng.words.y
               : rng.lane == 2 ? rng.words.z
                               : rng.words.w;
  ++rng.lane;
  return (float(w >> 9) + 0.5f) * 0x1p-23f;
}

// Per-lane event counters, reduced per wave before one global atomic each.
struct Counts {
  uint32_t photons, events, surface, walk, detector, lut, attempts, atomics;
};

template <typename R> struct Elements {
  R f11, f22, f33, f12;
  R f34, f44;  // filled for the full Stokes vector only
};

// V is carried by the full-Stokes (nmat 4) instantiations only; elsewhere it stays zero and
// the compiler drops it.
template <typename R> struct Stokes {
  R i, q, u, v;
};

template <typename R> struct Photon {
  int i, j, k;
  R rx, ry, rz;  // distances to the faces being approached (mc.f90:280-300)
  R a, b, c;     // direction cosines, c positive upwards
  R phi;         // azimuth of the direction (rad)
};

template <typename R> __device__ inline R clamp1(R x) {
  return x < R(-1) ? R(-1) : (x > R(1) ? R(1) : x);
}

// Largest index low with axis[low] <= value, in [0, n-2] (transport.metal:98-108).
__device__ inline int bracket(const float *__restrict__ axis, int n, float value, uint32_t &reads) {
  int low = 0, high = n - 1;
  while (high - low > 1) {
    const int middle = (low + high) >> 1;
    ++reads;
    if (axis[middle] <= value)
      low = middle;
    else
      high = middle;
  }
  return low;
}

// Interval of a tabulated cosine grid holding value: the equal-angle estimate from acos, moved
// to the exact interval (the one the binary search finds) by comparisons with the nodes.
template <typename R>
__device__ inline int bracket_angle(const float *__restrict__ axis, int n, R value, uint32_t &reads) {
  const R angle = acos(fmin(R(1), fmax(R(-1), value)));
  int low = int(R(n - 1) * (R(1) - angle * R(0.31830988618379067154)));
  low = low < 0 ? 0 : (low > n - 2 ? n - 2 : low);
  while (low > 0 && R(axis[low]) > value) --low;
  while (low < n - 2 && R(axis[low + 1]) <= value) ++low;
  reads += 2;
  return low;
}

// CDF interval of a draw: the bucket table gives the interval of the bucket's lower edge, then
// forward steps to the largest interval start <= draw (what the binary search finds).
__device__ inline int cdf_interval(const float *__restrict__ cdf, const uint16_t *__restrict__ start, int n,
                                   float draw, uint32_t &reads) {
  int low = start[int(draw * float(kCdfBuckets))];
  while (low < n - 2 && cdf[low + 1] <= draw) ++low;
  reads += 2;
  return low;
}

template <typename R>
__device__ inline int bracket_r(const float *__restrict__ axis, int n, R value, uint32_t &reads) {
  int low = 0, high = n - 1;
  while (high - low > 1) {
    const int middle = (low + high) >> 1;
    ++reads;
    if (R(axis[middle]) <= value)
      low = middle;
    else
      high = middle;
  }
  return low;
}

// ray_scat_matrix / mie_scat_matrix (mc_routines_general_mod.f90:74-136, 257-388):
// F11, F22, F33, F12 divided by 4 (radiance normalized for incident flux pi).
// The scattering angle is given as cos together with 1 - cos and 1 + cos. Tabulated elements
// interpolate linearly in cos as the Fortran does; near the peaks the interpolation weight is
// taken from 1 - cos (or 1 + cos) against the table's exact 1 - cos (1 + cos) nodes, because a
// float32 cos cannot resolve the forward peak (F11 changes ~1e8 per unit cos there).
// V4: also F34 and F44 (nmat 4): Rayleigh F34 = 0, F44 = Delta Delta' 1.5 cos with
// Delta' = (1 - 2 depol) / (1 - depol); tabulated F34, F44 interpolated like the others.
template <typename R, bool V4 = false>
// known: the table interval of cosine when the caller has it (the sampler), else -1.
__device__ inline Elements<R> elements(const Scene &s, int species, R cosine, R one_minus,
                                       R one_plus, uint32_t &reads, int known = -1) {
  cosine = clamp1(cosine);
  Elements<R> f;
  f.f34 = R(0);
  f.f44 = R(0);
  if (species < 0) {
    const R depol = R(s.depol);
    const R delta = (R(1) - depol) / (R(1) + depol / R(2));
    const R sin2 = one_minus * one_plus;
    const R diagonal = delta * R(0.75) * (R(2) - sin2);
    f.f11 = diagonal + R(1) - delta;
    f.f22 = diagonal;
    f.f33 = delta * R(1.5) * cosine;
    f.f12 = -delta * R(0.75) * sin2;
    if (V4) f.f44 = delta * ((R(1) - R(2) * depol) / (R(1) - depol) * R(1.5) * cosine);
  } else {
    const SpeciesTables &t = s.sp[species];
    const int n = t.lut_size;
    const int low = known >= 0       ? known
                    : t.equal_angle ? bracket_angle<R>(t.cosine, n, cosine, reads)
                                    : bracket_r<R>(t.cosine, n, cosine, reads);
    R w;
    if (cosine > R(0.5)) {
      const R o0 = R(t.cosine[n + low]), o1 = R(t.cosine[n + low + 1]);
      w = (o0 - one_minus) / (o0 - o1);
    } else if (cosine < R(-0.5)) {
      const R o0 = R(t.cosine[2 * n + low]), o1 = R(t.cosine[2 * n + low + 1]);
      w = (one_plus - o0) / (o1 - o0);
    } else {
      const R c0 = R(t.cosine[low]), c1 = R(t.cosine[low + 1]);
      w = (cosine - c0) / (c1 - c0);
    }
    reads += 10;
    f.f11 = R(t.f11[low]) + w * (R(t.f11[low + 1]) - R(t.f11[low]));
    f.f22 = R(t.f22[low]) + w * (R(t.f22[low + 1]) - R(t.f22[low]));
    f.f33 = R(t.f33[low]) + w * (R(t.f33[low + 1]) - R(t.f33[low]));
    f.f12 = R(t.f12[low]) + w * (R(t.f12[low + 1]) - R(t.f12[low]));
    if (V4) {
      f.f34 = R(t.f34[low]) + w * (R(t.f34[low + 1]) - R(t.f34[low]));
      f.f44 = R(t.f44[low]) + w * (R(t.f44[low + 1]) - R(t.f44[low]));
      reads += 4;
    }
  }
  if (s.nmat == 1) f.f12 = R(0);
  f.f11 *= R(0.25);
  f.f22 *= R(0.25);
  f.f33 *= R(0.25);
  f.f12 *= R(0.25);
  if (V4) {
    f.f34 *= R(0.25);
    f.f44 *= R(0.25);
  }
  return f;
}

template <typename R> struct Rotation {
  R c1, s1, c2, s2;  // cos/sin of twice the rotation angles i1, i2
};

// stokes_rot_mat (mc_routines_general_mod.f90:942-1174): coplanar, both-vertical,
// scattered-vertical, incident-vertical and general cases. The coplanarity threshold is
// 1e-16 in float64 (as mc.f90) and 1e-6 in float32, above the float32 spacing of azimuths up
// to 2 pi, so exact forward/backward coplanar geometry is recognised after rounding.
// The azimuths enter only through their difference: cos and sin of phio - phi1 (cd, sd)
// replace the four products of cos/sin phio and phi1 of the Fortran.
template <typename R>
__device__ inline Rotation<R> rotation_delta(R cosine, R uo, R u1, R delta, R cd, R sd) {
  const R eps = sizeof(R) == 4 ? R(1e-6) : R(1e-16);
  const R pi = R(3.14159265358979323846);
  const R dphi = fabs(delta);
  const R rem = fmod(dphi, pi);
  const bool vertical_o = sizeof(R) == 4 ? fabs(uo) == R(1) : fabs(uo) > R(1) - R(1e-16);
  const bool vertical_1 = sizeof(R) == 4 ? fabs(u1) == R(1) : fabs(u1) > R(1) - R(1e-16);
  if (fmin(rem, pi - rem) < eps || (vertical_o && vertical_1)) return Rotation<R>{R(1), R(0), R(1), R(0)};

  const R tiny = R(1e-16);
  const R fo = R(1) - uo * uo, f1 = R(1) - u1 * u1;
  const R sino = fo < tiny ? R(0) : sqrt(fo);
  const R sin1 = f1 < tiny ? R(0) : sqrt(f1);
  R cosi1, cosi2, q;
  if (vertical_1) {
    cosi1 = copysign(R(1), u1);
    cosi2 = -cosi1 * cd;
    q = sino * sd;
  } else if (vertical_o) {
    cosi2 = copysign(R(1), uo);
    cosi1 = -cosi2 * cd;
    q = sin1 * sd;
  } else {
    const R e = R(1e-16);
    const R sc = R(1) - cosine * cosine;
    cosi1 = (u1 - uo * cosine) / sqrt(fmax(e, (R(1) - uo * uo) * sc));
    cosi2 = (uo - u1 * cosine) / sqrt(fmax(e, (R(1) - u1 * u1) * sc));
    q = sino * sin1 * sd;
  }
  // Cancellation, or the 1e-16 floor under a near-zero scattering angle, can push |cos i|
  // past one; clamp as transport.metal:158-159 does (the Fortran only warns, :1142-1167).
  cosi1 = clamp1(cosi1);
  cosi2 = clamp1(cosi2);
  const R sign = q >= R(0) ? R(1) : R(-1);
  const R g1 = R(1) - cosi1 * cosi1, g2 = R(1) - cosi2 * cosi2;
  const R sini1 = g1 < tiny ? R(0) : sqrt(g1) * sign;
  const R sini2 = g2 < tiny ? R(0) : sqrt(g2) * sign;
  return Rotation<R>{cosi1 * cosi1 - sini1 * sini1, R(2) * sini1 * cosi1,
                     cosi2 * cosi2 - sini2 * sini2, R(2) * sini2 * cosi2};
}

template <typename R>
__device__ inline Rotation<R> rotation(R cosine, R uo, R u1, R phio, R phi1) {
  const R delta = phio - phi1;
  R sd, cd;
  sin_cos(delta, &sd, &cd);
  return rotation_delta<R>(cosine, uo, u1, delta, cd, sd);
}

// matmul_stokes_rot then P * S (mc_routines_general_mod.f90:1184-1242): rotate Q, U into the
// scattering plane, apply F, rotate out. V4 adds the circular component: U gains F34 V and
// V' = -F34 U_rotated + F44 V (F43 = -F34; V is unchanged by rotations).
template <typename R, bool V4 = false>
__device__ inline Stokes<R> scatter(const Elements<R> &f, const Rotation<R> &r, const Stokes<R> &s) {
  const R rq = r.c1 * s.q - r.s1 * s.u;
  const R ru = r.s1 * s.q + r.c1 * s.u;
  const R q = f.f12 * s.i + f.f22 * rq;
  const R u = V4 ? f.f33 * ru + f.f34 * s.v : f.f33 * ru;
  return Stokes<R>{f.f11 * s.i + f.f12 * rq, r.c2 * q - r.s2 * u, r.s2 * q + r.c2 * u,
                   V4 ? f.f44 * s.v - f.f34 * ru : R(0)};
}

// rayl_scatter_pol / mie_scatter_pol_bound, forward sign -1 (mc_routines_general_mod.f90:
// 492-932): rejection sampling of (cos theta, sigma) from F11 + F12 (Q cos 2s - U sin 2s),
// uniform cosine under the Rayleigh bound F11(1), inverse bound-CDF for tabulated species.
//
// Exact = true draws from the phase function itself: the reference accepts a tabulated draw
// with probability pdf / bound(cos), under a proposal that is uniform in cos within each node
// interval, so its density carries the factor (interval mean of the bound) / bound(cos). The
// exact variant accepts with pdf / (rho * interval mean) instead, which removes that factor.
template <typename R, bool Exact = false>
__device__ inline bool sample_direction(const Scene &s, int species, const Stokes<R> &st, R uo,
                                        Stream &rng, R &u1, R &dphi, R &sampled,
                                        R &sampled_minus, R &sampled_plus, Counts &n,
                                        int &sampled_interval) {
  const R pi = R(3.14159265358979323846);
  const R eps = R(1e-16);
  for (uint32_t attempt = 0; attempt < kSamplingAttempts; ++attempt) {
    ++n.attempts;
    // sin^2 of the scattering angle as (1 - cos)(1 + cos) from factors that keep their
    // precision near both peaks: exact for Rayleigh (cos = 2u - 1), interpolated from the
    // 1 -/+ cos tables for tabulated species.
    R cosine, sin2, one_minus, one_plus;
    R exact_bound = R(0);
    int interval = -1;
    if (species < 0) {
      const float u = uniform(rng);
      cosine = R(2) * R(u) - R(1);
      one_minus = R(2) - R(2) * R(u);
      one_plus = R(2) * R(u);
      sin2 = one_minus * one_plus;
    } else {
      const SpeciesTables &t = s.sp[species];
      const float draw = uniform(rng);
      const int low = t.cdf_start ? cdf_interval(t.cdf, t.cdf_start, t.lut_size, draw, n.lut)
                                  : bracket(t.cdf, t.lut_size, draw, n.lut);
      interval = low;
      const R d0 = R(t.cdf[low]), d1 = R(t.cdf[low + 1]);
      const R w = (R(draw) - d0) / (d1 - d0);
      const float *omc = t.cosine + t.lut_size, *opc = t.cosine + 2 * t.lut_size;
      one_minus = R(omc[low]) + w * (R(omc[low + 1]) - R(omc[low]));
      one_plus = R(opc[low]) + w * (R(opc[low + 1]) - R(opc[low]));
      cosine = R(t.cosine[low]) + w * (R(t.cosine[low + 1]) - R(t.cosine[low]));
      sin2 = one_minus * one_plus;
      n.lut += 8;
      if (Exact) {
        const R b0 = R(t.f11[low]) + (s.nmat > 1 ? fabs(R(t.f12[low])) : R(0));
        const R b1 = R(t.f11[low + 1]) + (s.nmat > 1 ? fabs(R(t.f12[low + 1])) : R(0));
        exact_bound = R(t.rho) * R(0.125) * (b0 + b1);  // rho x interval mean, in F/4 units
      }
    }
    const R sigma = R(2) * pi * R(uniform(rng)) - pi;
    const Elements<R> f = elements<R>(s, species, cosine, one_minus, one_plus, n.lut, interval);
    R ss, cs;
    sin_cos(sigma, &ss, &cs);
    // cos 2 sigma and sin 2 sigma from the half angle.
    const R pdf = f.f11 + f.f12 * (st.q * (cs * cs - ss * ss) - st.u * (R(2) * ss * cs));
    const R bound = species < 0 ? elements<R>(s, -1, R(1), R(0), R(2), n.lut).f11
                    : Exact     ? exact_bound
                                : f.f11 + fabs(f.f12);
    if (R(uniform(rng)) * bound > pdf) continue;

    sampled = cosine;
    sampled_minus = one_minus;
    sampled_plus = one_plus;
    sampled_interval = interval;
    const R sin_in = sqrt(fmax(eps, R(1) - uo * uo));
    const R sin_sca = sqrt(fmax(eps, sin2));
    u1 = clamp1(uo * cosine + sin_in * sin_sca * cs);
    // Azimuth change of the spherical triangle: the Fortran's acos((cos - uo u1) /
    // (sin_in sin_out)) written as an atan2 of its sine and cosine, which stays accurate for
    // the small azimuth changes of forward-peaked scattering (acos near 1 does not).
    const R magnitude = atan2(sin_sca * fabs(ss), cosine * sin_in - uo * sin_sca * cs);
    dphi = -copysign(magnitude, sigma);
    return true;
  }
  return false;
}

// Extinction of a cell (1/km) in layer k: the table, or the factored parts with the float32
// arithmetic that produced the table (the same value).
__device__ inline float cell_extinction(const Scene &s, int cell, int k) {
  return s.species_beta ? s.species_beta[cell] + s.gas_beta[k] : s.beta[cell];
}

template <typename R>
__device__ inline int cell_index(const Scene &s, int i, int j, int k) {
  return i + s.nx * (j + s.ny * k);
}

// New direction: reverse the face distances whose axis changed sign (mc.f90:1418-1421).
template <typename R>
__device__ inline void change_direction(const Scene &s, Photon<R> &p, R a, R b, R c, R phi) {
  if (p.a * a < R(0)) p.rx = fmax(R(0), R(s.xunit) - p.rx);
  if (p.b * b < R(0)) p.ry = fmax(R(0), R(s.yunit) - p.ry);
  if (p.c * c < R(0)) p.rz = fmax(R(0), R(s.zunit[p.k]) - p.rz);
  p.a = a;
  p.b = b;
  p.c = c;
  p.phi = phi;
}

template <typename R> __device__ inline R boundary_distance(const Photon<R> &p) {
  const R eps = R(1e-16);
  const R dx = p.rx / fmax(fabs(p.a), eps);
  const R dy = p.ry / fmax(fabs(p.b), eps);
  const R dz = p.rz / fmax(fabs(p.c), eps);
  return fmin(dx, fmin(dy, dz));
}

// ADVANCE_internal (mc.f90:1531-1580): move to the nearest face, x before y before z on ties,
// periodic in x and y. Returns the distance travelled.
template <typename R, bool LAYERS = false>
__device__ inline R advance(const Scene &s, Photon<R> &p, const float *layers = nullptr, int row = 0) {
  const R eps = R(1e-16);
  const R aa = fabs(p.a), ab = fabs(p.b), ac = fabs(p.c);
  const R dx = p.rx / fmax(aa, eps);
  const R dy = p.ry / fmax(ab, eps);
  const R dz = p.rz / fmax(ac, eps);
  const R d = fmin(dx, fmin(dy, dz));
  if (dx == d) {
    p.i += p.a >= R(0) ? 1 : -1;
    p.i = p.i < 0 ? s.nx - 1 : (p.i >= s.nx ? 0 : p.i);
    p.rx = R(s.xunit);
    p.ry = fmax(R(0), p.ry - ab * d);
    p.rz = fmax(R(0), p.rz - ac * d);
  } else if (dy == d) {
    p.j += p.b >= R(0) ? 1 : -1;
    p.j = p.j < 0 ? s.ny - 1 : (p.j >= s.ny ? 0 : p.j);
    p.rx = fmax(R(0), p.rx - aa * d);
    p.ry = R(s.yunit);
    p.rz = fmax(R(0), p.rz - ac * d);
  } else {
    p.k += p.c >= R(0) ? 1 : -1;
    p.rx = fmax(R(0), p.rx - aa * d);
    p.ry = fmax(R(0), p.ry - ab * d);
    // Layer thickness from the wave's LDS layer table where the caller has one.
    if (p.k >= 0 && p.k < s.nz) p.rz = LAYERS ? R(layers[row + p.k]) : R(s.zunit[p.k]);
  }
  return d;
}

// advance with the reciprocal direction magnitudes given (a ray whose direction does not change).
template <typename R, bool LAYERS = false>
__device__ inline R advance_by(const Scene &s, Photon<R> &p, R ia, R ib, R ic, const float *layers = nullptr,
                               int row = 0) {
  const R aa = fabs(p.a), ab = fabs(p.b), ac = fabs(p.c);
  const R dx = p.rx * ia;
  const R dy = p.ry * ib;
  const R dz = p.rz * ic;
  const R d = fmin(dx, fmin(dy, dz));
  if (dx == d) {
    p.i += p.a >= R(0) ? 1 : -1;
    p.i = p.i < 0 ? s.nx - 1 : (p.i >= s.nx ? 0 : p.i);
    p.rx = R(s.xunit);
    p.ry = fmax(R(0), p.ry - ab * d);
    p.rz = fmax(R(0), p.rz - ac * d);
  } else if (dy == d) {
    p.j += p.b >= R(0) ? 1 : -1;
    p.j = p.j < 0 ? s.ny - 1 : (p.j >= s.ny ? 0 : p.j);
    p.rx = fmax(R(0), p.rx - aa * d);
    p.ry = R(s.yunit);
    p.rz = fmax(R(0), p.rz - ac * d);
  } else {
    p.k += p.c >= R(0) ? 1 : -1;
    p.rx = fmax(R(0), p.rx - aa * d);
    p.ry = fmax(R(0), p.ry - ab * d);
    if (p.k >= 0 && p.k < s.nz) p.rz = LAYERS ? R(layers[row + p.k]) : R(s.zunit[p.k]);
  }
  return d;
}

// x reduced to [0, span): floor-based (fmod's exact remainder loop is not needed at the
// quotients of a periodic grid, and costs far more).
template <typename T> __device__ inline T periodic(T x, T span) {
  T r = x - span * floor(x / span);
  if (r < T(0)) r += span;
  if (r >= span) r -= span;
  return r;
}

// Move a distance along one horizontal axis inside a horizontally uniform layer: the photon
// keeps its direction, so the cell index and the distance to the approached face follow from
// the periodic position directly (positions in float64; a near-horizontal photon may travel
// thousands of cells in one flight). Returns the number of faces crossed.
template <typename R>
__device__ inline uint32_t slide_axis(int &index, R &remaining, R direction, double distance, float unit,
                                      int count) {
  if (direction == R(0)) return 0;
  const double u = double(unit);
  const double offset = direction > R(0) ? u - double(remaining) : double(remaining);
  const double travel = double(fabs(direction)) * distance;
  const double moved = direction > R(0) ? offset + travel : u - offset + travel;
  const uint32_t crossed = uint32_t(fmin(moved / u, 4.0e9));
  const double span = u * double(count);
  const double x = periodic<double>(double(index) * u + offset + double(direction) * distance, span);
  int cell = int(x / u);
  cell = cell < 0 ? 0 : (cell >= count ? count - 1 : cell);
  const double local = fmin(fmax(x - double(cell) * u, 0.0), u);
  index = cell;
  remaining = R(direction > R(0) ? u - local : local);
  return crossed;
}

// Column reached by a straight ray through uniform layers (transport.metal:270-279 and the
// analytical register mapping of mc_routines_cartesian_mod.f90:1076-1105).
template <typename R>
__device__ inline int projected_column(int index, R remaining, R direction, R distance, R unit,
                                       int count) {
  const R local = direction >= R(0) ? unit - remaining : remaining;
  const R position = periodic<R>(R(index) * unit + local + direction * distance, unit * R(count));
  const int column = int(position / unit);
  return column < 0 ? 0 : (column < count - 1 ? column : count - 1);
}

struct DetectorHit {
  int i, j;
  bool ok;
};

// calc_radiances_and_paths_pol (mc_routines_cartesian_mod.f90:849-1222) for one view:
// optical depth from the event to the top of the last physical layer and the TOA pixel.
// Cells are walked one by one below uniform_top; above it the vertical prefix gives the exact
// slant integral (transport.metal:299-313).
template <typename R, bool LAYERS = false, bool DOWN = false>
__device__ inline DetectorHit detector_walk(const Scene &s, const Photon<R> &event, R va, R vb,
                                            R mu, R &tau, uint32_t max_steps, Counts &n,
                                            const float *layers = nullptr, int row = 0) {
  Photon<R> ray = event;
  change_direction<R>(s, ray, va, vb, mu, R(s.phi));
  tau = R(0);
  uint32_t steps = 0;
  // The cell walk is software-pipelined: a cell's extinction is loaded when the walk enters the
  // cell and added to tau one cell later, so the load's latency overlaps the next face
  // crossing instead of stalling every step (the sum and its order are unchanged). The factored
  // parts stay separate until the add for the same reason.
  R held_d = R(0);
  float held_part = 0.0f, held_gas = 0.0f;
  const R ia = R(1) / fmax(fabs(ray.a), R(1e-16)), ib = R(1) / fmax(fabs(ray.b), R(1e-16)),
          ic = R(1) / fmax(fabs(ray.c), R(1e-16));
  while (ray.k >= 0 && ray.k < s.nz - 1) {
    if (++steps > max_steps) return DetectorHit{0, 0, false};
    if (!DOWN && ray.k >= s.uniform_top) {
      tau += held_d * R(held_part + held_gas);
      // Layers at and above uniform_top are uniform: the wave's LDS layer table holds their
      // extinction, height above and optical depth above.
      const R height_above = LAYERS ? R(layers[2 * row + ray.k]) : R(s.height_above[ray.k]);
      const R tau_above = LAYERS ? R(layers[3 * row + ray.k]) : R(s.tau_above[ray.k]);
      const R beta = LAYERS ? R(layers[ray.k]) : R(s.beta[cell_index<R>(s, ray.i, ray.j, ray.k)]);
      const R distance = (ray.rz + height_above) / mu;
      tau += (ray.rz * beta + tau_above) / mu;
      // Explicit-equivalent cell count: layers left plus horizontal faces crossed.
      const R hx = fabs(va) * distance, hy = fabs(vb) * distance;
      n.detector += uint32_t(s.nz - 1 - ray.k) +
                    uint32_t((R(s.xunit) - ray.rx + hx) / R(s.xunit)) +
                    uint32_t((R(s.yunit) - ray.ry + hy) / R(s.yunit));
      const int i = projected_column<R>(ray.i, ray.rx, va, distance, R(s.xunit), s.nx);
      const int j = projected_column<R>(ray.j, ray.ry, vb, distance, R(s.yunit), s.ny);
      return DetectorHit{i, j, true};
    }
    const int layer = ray.k;
    const int cell = cell_index<R>(s, ray.i, ray.j, layer);
    const float part = s.species_beta ? s.species_beta[cell] : s.beta[cell];
    const float gas = s.species_beta ? s.gas_beta[layer] : 0.0f;
    const R d = advance_by<R, LAYERS>(s, ray, ia, ib, ic, layers, row);
    tau += held_d * R(held_part + held_gas);
    held_d = d;
    held_part = part;
    held_gas = gas;
    ++n.detector;
  }
  tau += held_d * R(held_part + held_gas);
  return DetectorHit{ray.i, ray.j, true};
}

template <typename T> __device__ inline void add_tally(T *target, T value) {
  __hip_atomic_fetch_add(target, value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename R> struct Trace {
  R events, surface, status, weight, stokes_i, weights, paths, layer;
};

// State of one photon history between events. The layer paths of the photon (path_layers
// of mc.f90) live in LDS, one row of layers per lane.
template <typename R> struct History {
  Photon<R> p;
  Stokes<R> st;
  R weight, remains;
  Stream rng;
  uint32_t round_offset;
  uint32_t steps;
  int species;  // -1 gas, else tabulated species of the pending event
  bool surface;
  R path_total;  // sum of the layer paths, for the per-photon records
  float contribution;  // sum of the detector weights C1 of the history so far (all views)
  uint32_t order;      // atmospheric events so far
  // Backward transport: rows 2..4 of the weight matrix (mc.f90:411-418, 948-955); row 1 is
  // `st`. Each row is carried like a Stokes vector (W <- W P^T transforms every row r into
  // P r), all divided by the same norm. Unused, and dropped by the compiler, elsewhere.
  Stokes<R> w2, w3, w4;
  // Backward decomposition: the cell of the first scattering event, where every later event's
  // 3-D contribution is registered (calc_radiances_and_scats :509-521).
  int first_cell;
  // Steered modes: the product of the steering factors in `weight` (weight over the weight the
  // plain tracer would carry), which the weight window keeps near one.
  R steer;
};

// Local estimates of one event, parked in LDS until the wave's cooperative tally reads them.
// Struct of arrays over the block's lanes: lane writes are contiguous, and the tally's reads
// of one source lane are broadcasts.
constexpr int kPendingReals = 1 + 3 * kMaxViews;
constexpr int kPendingReals4 = 1 + 4 * kMaxViews;  // with V (nmat 4)
constexpr int kPendingWords = 2 + kMaxViews;
constexpr int pending_reals(bool v4) { return v4 ? kPendingReals4 : kPendingReals; }

template <typename R> struct Pending {
  R *real;     // rz | weight x views | q x views | u x views [| v x views with nmat 4]
  int *word;   // layer | round offset | pixel x views
  int stride;  // lanes per block
  // Vertical extent of the detector ray inside the event layer.
  __device__ R &rz() const { return real[0]; }
  // C1 = I exp(-tau), the detector weight of the view; Q and U contributions times mu0/mu.
  __device__ R &weight(int v) const { return real[(1 + v) * stride]; }
  __device__ R &q(int v) const { return real[(1 + kMaxViews + v) * stride]; }
  __device__ R &u(int v) const { return real[(1 + 2 * kMaxViews + v) * stride]; }
  __device__ R &v(int v) const { return real[(1 + 3 * kMaxViews + v) * stride]; }
  __device__ int &layer() const { return word[0]; }
  __device__ int &round() const { return word[stride]; }
  // Output pixel i + nx j, or -1 when the view gets nothing from this event.
  __device__ int &pixel(int v) const { return word[(2 + v) * stride]; }
  __device__ Pending at(int lane) const { return Pending{real + lane, word + lane, stride}; }
};

// Open tally records of a lane (modes with layer paths), in the same LDS area as Pending.
// Consecutive events of one photon in one layer k whose detector rays toward a view reach the
// same pixel share that view's record:
//   S = sum of their detector weights C1, E = sum of C1 (path in layer k so far + detector-ray
//   length inside layer k), and the sums of their Q, U (V) contributions.
// The record is written out before the photon's layer paths change outside layer k, so each
// layer j != k holds the same path for all its events, and the events' tally over layer j is
// S (path_j + detector-ray length in j) — one atomic instruction for the whole run of events.
constexpr int kRecordReals = 4 * kMaxViews;
constexpr int kRecordReals4 = 5 * kMaxViews;  // with V
constexpr int kRecordWords = kMaxViews;
constexpr int record_reals(bool v4) { return v4 ? kRecordReals4 : kRecordReals; }

template <typename R> struct Record {
  R *real;     // E x views | S x views | q x views | u x views [| v x views with nmat 4]
  int *word;   // pixel x views
  int stride;  // lanes per block
  __device__ R &e(int v) const { return real[v * stride]; }
  __device__ R &s(int v) const { return real[(kMaxViews + v) * stride]; }
  __device__ R &q(int v) const { return real[(2 * kMaxViews + v) * stride]; }
  __device__ R &u(int v) const { return real[(3 * kMaxViews + v) * stride]; }
  __device__ R &v(int v) const { return real[(4 * kMaxViews + v) * stride]; }
  // Output pixel of the open record, -1 when the view has none.
  __device__ int &pixel(int v) const { return word[v * stride]; }
  __device__ Record at(int lane) const { return Record{real + lane, word + lane, stride}; }
};

// Bytes per lane of the area that holds either layout.
template <typename R> constexpr size_t pending_bytes(bool v4) {
  return pending_reals(v4) * sizeof(R) + kPendingWords * sizeof(int) >
                 record_reals(v4) * sizeof(R) + kRecordWords * sizeof(int)
             ? pending_reals(v4) * sizeof(R) + kPendingWords * sizeof(int)
             : record_reals(v4) * sizeof(R) + kRecordWords * sizeof(int);
}

template <typename R> __device__ inline bool finite3(const Stokes<R> &v) {
  return isfinite(v.i) && isfinite(v.q) && isfinite(v.u);
}

template <typename R, bool V4> __device__ inline bool finite_stokes(const Stokes<R> &v) {
  return finite3(v) && (!V4 || isfinite(v.v));
}

// Launch at TOA in the vacuum layer of column photon mod (nx*ny) (mc.f90:351-464).
template <typename R>
__device__ void launch_photon(const Scene &s, const Controls &c, uint32_t id, R *history,
                              History<R> &h, int held = 1 << 30) {
  const uint32_t identity = c.first + id;
  h.round_offset = identity / c.per_round;
  const uint32_t photon = identity % c.per_round;
  h.rng = Stream{photon, c.round + h.round_offset, c.seed, 0u, 0u, Words{0, 0, 0, 0}, Words{0, 0, 0, 0}, 4, false};
  const int column = int(photon % uint32_t(s.nx * s.ny));
  Photon<R> &p = h.p;
  p.i = column % s.nx;
  p.j = column / s.nx;
  p.k = s.nz - 1;
  p.a = R(s.launch_a);
  p.b = R(s.launch_b);
  p.c = R(s.launch_c);
  p.phi = R(s.phi0);
  p.rx = R(s.xunit) * R(uniform(h.rng));
  p.ry = R(s.yunit) * R(uniform(h.rng));
  p.rz = R(s.zunit[p.k]);
  for (int k = 0; k < s.nz && k < held; ++k) history[k] = R(0);
  h.st = Stokes<R>{R(1), R(0), R(0)};
  h.w2 = Stokes<R>{R(0), R(1), R(0), R(0)};
  h.w3 = Stokes<R>{R(0), R(0), R(1), R(0)};
  h.w4 = Stokes<R>{R(0), R(0), R(0), R(1)};
  h.weight = R(1);
  h.remains = -log(R(uniform(h.rng)));
  h.steps = 0;
  h.path_total = R(0);
  h.contribution = 0.0f;
  h.order = 0;
  h.first_cell = 0;
  h.steer = R(1);
}

// Result grid on the other level than the detector (mc_routines_cartesian_mod.f90:1041-1105,
// direction -1): follow the line of sight back from the event, vertically to the ground for a
// TOA detector or to TOA for a ground detector, and return the column reached there. rz is
// the event's distance to the layer face toward the detector.
template <typename R, bool DOWN>
__device__ inline int opposite_grid_pixel(const Scene &s, const Photon<R> &p, R rz, int view) {
  const R thickness = R(s.zunit[p.k]);
  const R above = R(s.height_above[p.k]);
  const R vertical = DOWN ? above + thickness - rz : R(s.height_above[0] + s.zunit[0]) - above - rz;
  const R back = -vertical * R(s.view_inv_mu[view]);
  const R x = R(p.i) * R(s.xunit) + (p.a >= R(0) ? R(s.xunit) - p.rx : p.rx) + R(s.view_a[view]) * back;
  const R y = R(p.j) * R(s.yunit) + (p.b >= R(0) ? R(s.yunit) - p.ry : p.ry) + R(s.view_b[view]) * back;
  const int i = min(int(periodic<R>(x, R(s.xunit) * R(s.nx)) / R(s.xunit)), s.nx - 1);
  const int j = min(int(periodic<R>(y, R(s.yunit) * R(s.ny)) / R(s.yunit)), s.ny - 1);
  return max(i, 0) + s.nx * max(j, 0);
}

// Paths of a kWide lane in the upper layers low .. nz - 1 (all horizontally uniform). While
// `residual` is clear, the photon's first flight from TOA crossed layers descent .. nz - 1
// whole (descent nz: none yet) and the other upper layers hold nothing; once set, the lane's
// row in global memory holds every upper layer's path (layer k at k - low). `initial`: no
// event yet, the current flight is the first one.
// One word: descent (bits 0-6) | residual << 7 | initial << 8.
struct Upper {
  uint32_t word;
  __device__ int descent() const { return int(word & 127u); }
  __device__ bool residual() const { return (word >> 7) & 1u; }
  __device__ bool initial() const { return (word >> 8) & 1u; }
  __device__ void set_descent(int k) { word = (word & ~127u) | uint32_t(k); }
};

// A whole crossing of layer k at direction cosine c, as the free-path walk measures it.
template <typename R>
__device__ inline R whole_crossing(const float *layers, int row, int k, R c) {
  return R(layers[row + k]) / fmax(fabs(c), R(1e-16));
}

// From here on the lane's global row holds the upper paths: the first flight's crossings.
template <typename R>
__device__ inline void upper_materialize(const Scene &s, Upper &u, float *upper, int low, const float *layers,
                                         int row) {
  if (u.residual()) return;
  const R c0 = R(s.launch_c);
  const int descent = u.descent();
  for (int k = low; k < s.nz; ++k)
    upper[k - low] = k >= descent ? float(whole_crossing<R>(layers, row, k, c0)) : 0.0f;
  u.word = (u.word & 256u) | 128u | uint32_t(s.nz);
}

// The lane's path in upper layer k.
template <typename R>
__device__ inline R upper_path(const Scene &s, const Upper &u, const float *upper, int low, const float *layers,
                               int row, int k) {
  if (u.residual()) return R(upper[k - low]);
  return k >= u.descent() ? whole_crossing<R>(layers, row, k, R(s.launch_c)) : R(0);
}

// Advance one photon to its next collision or surface hit (mc.f90:476-677). Returns kRunning
// at the event (`surface` set for a surface hit), kEscaped or kFailed. layer_path: the path
// the flight added to the layer it ends in (with layer paths), 0 on entry.
// layers (LAYERS): the wave's LDS copy of four per-layer rows (row values apart): extinction of
// horizontally uniform layers (-1 for varying ones), thickness, height above and optical depth
// above, read instead of the global tables on every layer crossing of the free path and by
// the detector walk.
// WIDE (kWide): the paths of layers low .. nz - 1 live outside LDS (Upper); a flight's
// crossings there are added when it ends in an event up there or flies back down below `low`,
// never for a flight that escapes.
template <typename R, bool LAYERS, bool WIDE = false>
__device__ inline int walk(const Scene &s, const Controls &c, unsigned *failure, R *history, History<R> &h,
                           Counts &n, const float *layers, int row, bool &surface, R &layer_path,
                           Upper *u = nullptr, float *upper = nullptr, int low = 1 << 30) {
  const R eps = R(1e-16);
  Photon<R> &p = h.p;
  double slide = 0.0;  // flight through uniform layers not yet applied to the x/y cells
  // WIDE: upper layers this flight crossed (pending_lo .. pending_hi, none while lo > hi), the
  // first of them with a length of its own (the flight may start inside it).
  int pending_lo = s.nz, pending_hi = -1, pending_start = -1;
  R pending_first = R(0);
  auto add_pending = [&]() {
    if (WIDE && pending_lo <= pending_hi) {
      upper_materialize<R>(s, *u, upper, low, layers, row);
      for (int k = pending_lo; k <= pending_hi; ++k)
        upper[k - low] += float(k == pending_start ? pending_first : whole_crossing<R>(layers, row, k, p.c));
      pending_lo = s.nz;
      pending_hi = -1;
    }
  };
  auto flush_slide = [&]() {
    if (slide != 0.0) {
      n.walk += slide_axis<R>(p.i, p.rx, p.a, slide, s.xunit, s.nx) +
                slide_axis<R>(p.j, p.ry, p.b, slide, s.yunit, s.ny);
      slide = 0.0;
    }
  };

  for (;;) {
    if (++h.steps > c.max_steps) {
      atomicOr(failure, kHistoryLimit);
      return kFailed;
    }
    // Entering a varying layer: apply the deferred horizontal move before the column matters.
    const float layer_extinction = LAYERS ? layers[p.k] : 0.0f;
    const bool uniform_layer = LAYERS ? layer_extinction >= 0.0f : s.uniform[p.k] != 0.0f;
    if (!uniform_layer) flush_slide();
    const R extinction = LAYERS && uniform_layer ? R(layer_extinction)
                         : (p.k < s.nz - 1 && p.k >= 0) ? R(cell_extinction(s, cell_index<R>(s, p.i, p.j, p.k), p.k))
                                                        : R(0);
    const R event_distance = extinction > eps ? h.remains / extinction : R(INFINITY);
    if (uniform_layer) {
      // Horizontally uniform layer: straight to the event or to the next z face. The path is
      // the one the cell-by-cell walk takes (x/y faces change nothing here).
      const R to_face = p.rz / fmax(fabs(p.c), eps);
      const bool event = event_distance < to_face;
      const R d = event ? event_distance : to_face;
      // The horizontal move is deferred: one slide for the whole flight through uniform
      // layers (the direction does not change within a flight).
      slide += double(d);
      if (s.paths) {
        if (!WIDE || p.k < low) {
          history[p.k] += d;
        } else if (event) {
          add_pending();
          upper_materialize<R>(s, *u, upper, low, layers, row);
          upper[p.k - low] += float(d);
        } else if (u->initial()) {
          u->set_descent(p.k);  // a whole crossing of the first flight from TOA
        } else {
          if (pending_lo > pending_hi) {
            pending_start = p.k;
            pending_first = d;
          }
          pending_lo = min(pending_lo, p.k);
          pending_hi = max(pending_hi, p.k);
        }
        h.path_total += d;
        layer_path += d;
      }
      if (event) {
        flush_slide();
        p.rz = fmax(R(0), p.rz - fabs(p.c) * d);
        break;
      }
      ++n.walk;
      h.remains -= d * extinction;
      p.k += p.c >= R(0) ? 1 : -1;
      if (p.k >= s.nz) return kEscaped;
      // Back down below the upper layers: this flight's crossings there are complete.
      if (WIDE && p.k == low - 1 && p.c < R(0)) add_pending();
      if (p.k >= 0) {
        p.rz = LAYERS ? R(layers[row + p.k]) : R(s.zunit[p.k]);
        layer_path = R(0);
        continue;
      }
      flush_slide();
      surface = true;
      p.k = 0;
      p.rz = R(0);
      break;
    }
    if (event_distance < boundary_distance<R>(p)) {
      p.rx = fmax(R(0), p.rx - fabs(p.a) * event_distance);
      p.ry = fmax(R(0), p.ry - fabs(p.b) * event_distance);
      p.rz = fmax(R(0), p.rz - fabs(p.c) * event_distance);
      if (s.paths) {
        history[p.k] += event_distance;
        h.path_total += event_distance;
        layer_path += event_distance;
      }
      break;
    }
    const int old_layer = p.k;
    const R d = advance<R, LAYERS>(s, p, layers, row);
    ++n.walk;
    if (s.paths) {
      history[old_layer] += d;
      h.path_total += d;
      layer_path += d;
    }
    h.remains -= d * extinction;
    if (p.k >= s.nz) return kEscaped;
    if (p.k >= 0) {
      if (p.k != old_layer) layer_path = R(0);
      continue;
    }
    surface = true;
    p.k = 0;
    p.rz = R(0);
    break;
  }
  return kRunning;
}

// Interaction at the event (mc.f90:681-727, 1028-1074): the surface albedo, or the species
// drawn from the cell's fractions and its single-scattering albedo. Returns kAbsorbed when the
// weight falls below 1e-16, else kRunning.
template <typename R, bool TRACE, bool SCATS>
__device__ inline int interact(const Scene &s, History<R> &h, Counts &n, Trace<R> &trace, bool surface) {
  const R eps = R(1e-16);
  const Photon<R> &p = h.p;
  int species = -1;
  if (surface) {
    h.weight *= R(s.albedo[p.i + s.nx * p.j]);
    ++n.surface;
    if (SCATS) ++h.order;  // the scattering order counts surface reflections (mc.f90:689)
    if (TRACE) trace.surface += R(1);
  } else {
    const float draw = uniform(h.rng);
    const int cell = cell_index<R>(s, p.i, p.j, p.k);
    const int cells = s.nx * s.ny * s.nz;
    if (s.species_beta) {
      // One tabulated species: its fraction from the factored parts (the table's value).
      const float part = s.species_beta[cell];
      if (draw < part / (part + s.gas_beta[p.k])) species = 0;
    } else {
      for (int t = 0; t < s.species; ++t) {
        if (draw < s.cumfrac[t * cells + cell]) {
          species = t;
          break;
        }
      }
    }
    h.weight *= species < 0 ? R(s.ssalb_gas[p.k]) : R(s.sp[species].ssalb);
    ++n.events;
    ++h.order;
    if (TRACE) trace.events += R(1);
  }
  h.species = species;
  h.surface = surface;
  return h.weight < eps ? kAbsorbed : kRunning;
}

// Detector ray inside the event layer: the distance to its top face (views point upward), or
// to its bottom face for the downward views of a surface detector.
template <typename R, bool LAYERS, bool DOWN>
__device__ inline R event_rz(const Scene &s, const Photon<R> &p, const float *layers, int row) {
  return (p.c < R(0)) != DOWN ? fmax(R(0), (LAYERS ? R(layers[row + p.k]) : R(s.zunit[p.k])) - p.rz) : p.rz;
}

// Local estimate of one event toward one view: C1 = I exp(-tau) and the Q, U (V) contributions
// times mu0/mu, and the output pixel (-1 when the view gets nothing from the event).
template <typename R> struct Estimate {
  R weight, q, u, v;
  int pixel;
};

// The local estimate toward `view` (mc.f90:738-822, 1109-1207; detector walk and pixel,
// mc_routines_cartesian_mod.f90:849-1222) of the event the history is at. delta: the
// photon-view azimuth difference, with its cosine and sine; rz: event_rz. Sets `status` to
// kFailed (and a failure flag) on a non-finite estimate or a detector walk over its limit.
template <typename R, bool TRACE, bool V4, bool BACK, bool LAYERS, bool DOWN>
__device__ inline Estimate<R> estimate(const Scene &s, const Controls &c, unsigned *failure, History<R> &h,
                                       int view, R delta, R cos_delta, R sin_delta, R rz, Counts &n,
                                       Trace<R> &trace, const float *layers, int row, int &status) {
  const R eps = R(1e-16);
  const Photon<R> &p = h.p;
  const int species = h.species;
  Estimate<R> out{R(0), R(0), R(0), R(0), -1};
  {
    const R mu = R(s.view_mu[view]);
    Stokes<R> prob;
    if (h.surface) {
      if (BACK) {
        // W As S0 = the first column of W (mc.f90:754-755).
        const R scale = (DOWN ? R(0) : mu) * h.weight;
        prob = Stokes<R>{h.st.i * scale, h.w2.i * scale, h.w3.i * scale, V4 ? h.w4.i * scale : R(0)};
      } else {
        // Surface reflections only travel upward (mc.f90:804-808): nothing toward a view that
        // points down.
        prob = Stokes<R>{DOWN ? R(0) : h.st.i * mu * h.weight, R(0), R(0)};
      }
    } else {
      // Sines floored at sqrt(1e-16) as mc.f90:1112-1113 does: the forward peak of a Mie
      // table is steep enough for this to matter in float64.
      const R cosine = clamp1(p.c * mu + sqrt(fmax(eps, R(1) - p.c * p.c)) * R(s.view_sine[view]) *
                                             cos_delta);
      // 1 -/+ cos from the chord between the photon and view directions: accurate to the
      // direction's own rounding even within a fraction of a degree of the peaks.
      const R vx = R(s.view_a[view]), vy = R(s.view_b[view]);
      const R dx = p.a - vx, dy = p.b - vy, dz = p.c - mu;
      const R ex = p.a + vx, ey = p.b + vy, ez = p.c + mu;
      const R one_minus = R(0.5) * (dx * dx + dy * dy + dz * dz);
      const R one_plus = R(0.5) * (ex * ex + ey * ey + ez * ez);
      const Elements<R> f = elements<R, V4>(s, species, cosine, one_minus, one_plus, n.lut);
      const Rotation<R> r = s.nmat > 1 ? rotation_delta<R>(cosine, p.c, mu, delta, cos_delta, sin_delta)
                                       : Rotation<R>{R(1), R(0), R(1), R(0)};
      if (BACK) {
        // W P^T S0 (mc.f90:1150-1151): each weight-matrix row dotted with the first row of the
        // rotated phase matrix, (F11, F12 C1, -F12 S1, 0).
        const R a0 = f.f11, a1 = f.f12 * r.c1, a2 = -f.f12 * r.s1;
        prob = Stokes<R>{a0 * h.st.i + a1 * h.st.q + a2 * h.st.u, a0 * h.w2.i + a1 * h.w2.q + a2 * h.w2.u,
                         a0 * h.w3.i + a1 * h.w3.q + a2 * h.w3.u,
                         V4 ? a0 * h.w4.i + a1 * h.w4.q + a2 * h.w4.u : R(0)};
      } else {
        prob = scatter<R, V4>(f, r, h.st);
      }
      prob.i *= h.weight;
      prob.q *= h.weight;
      prob.u *= h.weight;
      if (V4) prob.v *= h.weight;
    }
    if (!finite_stokes<R, V4>(prob)) {
      atomicOr(failure, kNonfiniteProbability);
      status = kFailed;
      return out;
    }
    if (prob.i <= eps) return out;
    R tau;
    const DetectorHit hit =
        detector_walk<R, LAYERS, DOWN>(s, p, R(s.view_a[view]), R(s.view_b[view]), mu, tau, c.max_steps, n, layers,
                                       row);
    if (!hit.ok) {
      atomicOr(failure, kDetectorLimit);
      status = kFailed;
      return out;
    }
    const R attenuation = exp(-tau);
    const R geometry = R(s.view_geometry[view]);  // mu0 / mu (mc.f90:1456-1474)
    const R ci = prob.i * attenuation;
    // Backward transport registers at the injection column (mc_routines_cartesian_mod.f90:
    // 1035-1039), the launch column of the photon.
    if (BACK) {
      out.pixel = int(h.rng.photon % uint32_t(s.nx * s.ny));
    } else if (__builtin_expect(s.register_opposite, 0)) {
      out.pixel = opposite_grid_pixel<R, DOWN>(s, p, rz, view);
    } else {
      out.pixel = hit.i + s.nx * hit.j;
    }
    out.weight = ci;
    h.contribution += float(ci);
    out.q = prob.q * attenuation * geometry;
    out.u = prob.u * attenuation * geometry;
    if (V4) out.v = prob.v * attenuation * geometry;
    if (TRACE) {
      trace.stokes_i += ci * geometry;
      if (s.paths) {
        trace.weights += ci;
        // Layers the detector ray crosses after the event layer: above it, or below it for a
        // surface detector.
        const R beyond = DOWN
                             ? R(s.height_above[0] + s.zunit[0]) - R(s.height_above[p.k]) - R(s.zunit[p.k])
                             : R(s.height_above[p.k]);
        trace.paths += ci * (h.path_total + (rz + beyond) / fabs(mu));
      }
    }
  }
  return out;
}

// One event of a history whose local estimates are tallied one event at a time (the
// scattering decomposition): the flight, the interaction, and every view's local estimate
// parked in `out`. Returns kRunning while the history continues, else its end status.
template <typename R, bool TRACE, bool V4 = false, bool BACK = false, bool SCATS = false, bool LAYERS = false,
          bool DOWN = false>
__device__ int collide(const Scene &s, const Controls &c, unsigned *failure, R *history,
                       History<R> &h, const Pending<R> &out, Counts &n, Trace<R> &trace,
                       const float *layers = nullptr, int row = 0) {
  const Photon<R> &p = h.p;
  bool surface = false;
  R layer_path = R(0);
  int status = walk<R, LAYERS>(s, c, failure, history, h, n, layers, row, surface, layer_path);
  if (status != kRunning) return status;
  status = interact<R, TRACE, SCATS>(s, h, n, trace, surface);
  if (status != kRunning) return status;

  if (SCATS) {
    // Event cell (calc_radiances_and_scats :503-507): level 0 for the surface, else the layer;
    // complemented when the scattering order exceeds one.
    // (Events never occur in the vacuum layer, so level <= nz - 1; the bound keeps it so.)
    const int level = (p.k == 0 && p.rz < R(1e-16)) ? 0 : min(p.k + 1, s.nz - 1);
    int cell = p.i + s.nx * (p.j + s.ny * level);
    if (BACK) {
      if (h.order == 1)
        h.first_cell = cell;
      else
        cell = h.first_cell;
    }
    out.layer() = h.order == 1 ? cell : ~cell;
  } else {
    out.layer() = p.k;
  }
  out.round() = int(h.round_offset);
  const R rz = event_rz<R, LAYERS, DOWN>(s, p, layers, row);
  out.rz() = rz;
  // All views share one azimuth: cos and sin of the photon-view azimuth difference, once.
  const R delta = p.phi - R(s.phi);
  R sin_delta, cos_delta;
  sin_cos(delta, &sin_delta, &cos_delta);
#pragma unroll 1
  for (int view = 0; view < s.views; ++view) {
    const Estimate<R> e = estimate<R, TRACE, V4, BACK, LAYERS, DOWN>(s, c, failure, h, view, delta, cos_delta,
                                                                     sin_delta, rz, n, trace, layers, row, status);
    if (status != kRunning) return status;
    out.pixel(view) = e.pixel;
    out.weight(view) = e.weight;
    out.q(view) = e.q;
    out.u(view) = e.u;
    if (V4) out.v(view) = e.v;
  }
  return kRunning;
}

// Proposals of the steered modes.
//
// ViewLobes: lobes around the view directions shaped by a phase function. A lobe picks a view
// uniformly, draws a scattering cosine from the bound CDF of tabulated species `species` (the
// plain tracer's sampler, 1 - cos interpolated from the 1 - cos nodes so that the forward peak
// keeps its resolution) and measures it from the view direction, with a uniform azimuth about
// it. The CDF interpolates linearly between nodes, so the cosine density is constant on each
// node interval, and the direction density is
//   q(w) = (1 / V) sum_v g(w . view_v) / (2 pi),  g = (cdf[i+1] - cdf[i]) / (cos[i+1] - cos[i])
// on the interval i holding w . view_v. At the next event the local estimate toward view v
// evaluates the same phase function at the same cosine, which exceeds g by at most the table's
// node-to-interval-mean ratio; a photon steered into the forward peak toward a view therefore
// carries that peak's inverse in its weight factor p / (alpha p + (1 - alpha) q), and the
// forward-peak spikes of the local estimate (rare, huge contributions) become frequent, small
// ones. Without a tabulated species: vMF lobes of concentration kappa.
struct ViewLobes {
  const Scene *s;
  int species;
  float kappa;
};

__device__ inline void steer_sample(const ViewLobes &l, float u0, float u1, float u2, float &wx, float &wy,
                                    float &wz) {
  const Scene &s = *l.s;
  const int v = min(int(u0 * float(s.views)), s.views - 1);
  const float vx = float(s.view_a[v]), vy = float(s.view_b[v]), vz = float(s.view_mu[v]);
  if (l.species < 0) {
    monki_net::vmf_sample(vx, vy, vz, l.kappa, u1, u2, wx, wy, wz);
    return;
  }
  const SpeciesTables &t = s.sp[l.species];
  uint32_t reads = 0;
  const int i = t.cdf_start ? cdf_interval(t.cdf, t.cdf_start, t.lut_size, u1, reads)
                            : bracket(t.cdf, t.lut_size, u1, reads);
  const float d0 = t.cdf[i], d1 = t.cdf[i + 1];
  const float w = d1 > d0 ? (u1 - d0) / (d1 - d0) : 0.5f;
  const float *omc = t.cosine + t.lut_size;
  const float one_minus = omc[i] + w * (omc[i + 1] - omc[i]);
  const float cosine = t.cosine[i] + w * (t.cosine[i + 1] - t.cosine[i]);
  const float sine = sqrtf(fmaxf(0.0f, one_minus * (2.0f - one_minus)));
  monki_net::frame_direction(vx, vy, vz, cosine, sine, 6.28318530717958647692f * u2, wx, wy, wz);
}

__device__ inline float steer_density(const ViewLobes &l, float wx, float wy, float wz) {
  const Scene &s = *l.s;
  float q = 0.0f;
  for (int v = 0; v < s.views; ++v) {
    const float dot = fminf(1.0f, fmaxf(-1.0f, float(s.view_a[v]) * wx + float(s.view_b[v]) * wy +
                                                   float(s.view_mu[v]) * wz));
    if (l.species < 0) {
      q += expf(monki_net::log_vmf(l.kappa, dot));
    } else {
      const SpeciesTables &t = s.sp[l.species];
      uint32_t reads = 0;
      const int i = t.equal_angle ? bracket_angle<float>(t.cosine, t.lut_size, dot, reads)
                                  : bracket_r<float>(t.cosine, t.lut_size, dot, reads);
      q += (t.cdf[i + 1] - t.cdf[i]) / (t.cosine[i + 1] - t.cosine[i]) * 0.15915494309189533577f;
    }
  }
  return q / float(s.views);
}

// The guided proposal: lobe_fraction x ViewLobes + (1 - lobe_fraction) x the network's mixture
// for this event.
struct GuidedProposal {
  const monki_net::Mixture *net;
  ViewLobes lobes;
  float lobe_fraction;
};

__device__ inline void steer_sample(const GuidedProposal &g, float u0, float u1, float u2, float &wx, float &wy,
                                    float &wz) {
  // u0 picks the part and is then reused, rescaled, inside it.
  if (u0 < g.lobe_fraction) {
    steer_sample(g.lobes, u0 / g.lobe_fraction, u1, u2, wx, wy, wz);
  } else {
    const float rest = (u0 - g.lobe_fraction) / (1.0f - g.lobe_fraction);
    monki_net::mixture_sample(*g.net, fminf(rest, 0.99999994f), u1, u2, wx, wy, wz);
  }
}

__device__ inline float steer_density(const GuidedProposal &g, float wx, float wy, float wz) {
  const float lobes = g.lobe_fraction > 0.0f ? steer_density(g.lobes, wx, wy, wz) : 0.0f;
  const float net = g.lobe_fraction < 1.0f ? monki_net::mixture_density(*g.net, wx, wy, wz) : 0.0f;
  return g.lobe_fraction * lobes + (1.0f - g.lobe_fraction) * net;
}

// Continuation direction after the event (mc.f90:1266-1431; GROUND_LAMB2 at the surface).
// Steered events (guided and detector-directed modes; see Guidance) draw from the physical
// sampler with probability alpha, otherwise from the proposal, and multiply the weight by
// p / (alpha p + (1 - alpha) q): p the density of the physical sampler for the photon's Stokes
// vector (phase function; Lambert mu / pi at the surface), q the proposal's. The factor is at
// most 1 / alpha. Atmospheric events use `proposal`; surface reflections the view lobes.
// Events outside the steering (kind, order, surface) are sampled exactly as in the plain mode.
template <typename R, bool TRACE, int MODE = kPlain, typename Proposal = ViewLobes>
__device__ int scatter_photon(const Scene &s, unsigned *failure, History<R> &h, Counts &n,
                              Trace<R> &trace, const Proposal *proposal = nullptr,
                              const ViewLobes *lobes = nullptr, Guidance guidance = unsteered()) {
  constexpr bool V4 = full_stokes(MODE);
  constexpr bool BACK = backward(MODE);
  constexpr bool STEER = MODE == kGuided || MODE == kDirected;
  const float alpha = guidance.alpha;
  const R pi = R(3.14159265358979323846);
  Photon<R> &p = h.p;
  R mu, phi;
  R da = R(0), db = R(0), dc = R(0);
  bool direct = false;  // direction cosines given (steered proposal draws)
  if (h.surface) {
    const bool steer = STEER && guidance.surface;
    const bool lambert = !steer || uniform(h.rng) < alpha;
    if (lambert) {
      mu = sqrt(R(uniform(h.rng)));
      phi = R(2) * pi * R(uniform(h.rng));
    } else {
      float wx, wy, wz;
      const float u0 = uniform(h.rng), u1 = uniform(h.rng), u2 = uniform(h.rng);
      steer_sample(*lobes, u0, u1, u2, wx, wy, wz);
      da = R(wx), db = R(wy), dc = R(wz);
      direct = true;
      mu = dc;
      phi = atan2(db, da);
    }
    if (steer) {
      if (lambert) {
        const R sine = sqrt(fmax(R(0), R(1) - mu * mu));
        da = cos(phi) * sine, db = sin(phi) * sine, dc = mu;
      }
      // Lambertian density over the upper hemisphere; a proposal draw into the ground ends the
      // history (p = 0 there).
      const R lambert_density = mu > R(0) ? mu / pi : R(0);
      const R q = R(steer_density(*lobes, float(da), float(db), float(dc)));
      const R factor = lambert_density / (R(alpha) * lambert_density + R(1.0f - alpha) * q);
      h.weight *= factor;
      h.steer *= factor;
      if (!(h.weight > R(0))) return kAbsorbed;
    }
    h.st = Stokes<R>{R(1), R(0), R(0)};
    if (BACK) {
      // W As keeps the first column of W (mc.f90:894-903; W11 is one here).
      h.w2 = Stokes<R>{h.w2.i, R(0), R(0), R(0)};
      h.w3 = Stokes<R>{h.w3.i, R(0), R(0), R(0)};
      h.w4 = Stokes<R>{h.w4.i, R(0), R(0), R(0)};
    }
  } else {
    R cosine, one_minus, one_plus;
    int interval = -1;  // table interval of the sampled cosine (phase-function draws)
    bool from_phase = true;
    // Guided mode beyond the network's orders: the lobes alone, or no steering when the proposal
    // has no lobes (lobe_fraction 0).
    const bool guide = STEER && ((guidance.species_mask >> (h.species + 1)) & 1u) && h.order <= guidance.max_order &&
                       (MODE != kGuided || h.order <= guidance.network_orders || guidance.lobe_fraction > 0.0f);
    if (guide) from_phase = uniform(h.rng) < alpha;
    if (from_phase) {
      R dphi;
      const bool sampled =
          guide
              ? sample_direction<R, true>(s, h.species, h.st, p.c, h.rng, mu, dphi, cosine, one_minus, one_plus, n,
                                          interval)
              : sample_direction<R, false>(s, h.species, h.st, p.c, h.rng, mu, dphi, cosine, one_minus, one_plus, n,
                                           interval);
      if (!sampled) {
        atomicOr(failure, kSamplingLimit);
        return kFailed;
      }
      phi = p.phi + dphi;
      // float32 keeps the azimuth in [-pi, pi]: every use is periodic, and an azimuth that
      // random-walks away from zero loses the resolution the coplanarity test needs.
      if (sizeof(R) == 4) phi = remainder(phi, R(2) * pi);
    } else {
      float wx, wy, wz;
      const float u0 = uniform(h.rng), u1 = uniform(h.rng), u2 = uniform(h.rng);
      steer_sample(*proposal, u0, u1, u2, wx, wy, wz);
      da = R(wx), db = R(wy), dc = R(wz);
      direct = true;
      mu = dc;
      phi = atan2(db, da);
      const R ex = p.a - da, ey = p.b - db, ez = p.c - dc;
      const R fx = p.a + da, fy = p.b + db, fz = p.c + dc;
      one_minus = R(0.5) * (ex * ex + ey * ey + ez * ez);
      one_plus = R(0.5) * (fx * fx + fy * fy + fz * fz);
      cosine = clamp1(p.a * da + p.b * db + p.c * dc);
    }
    // The sampled scattering cosine is used directly, as in mc.f90.
    if (s.nmat > 1 || guide) {
      const Elements<R> f = elements<R, V4>(s, h.species, cosine, one_minus, one_plus, n.lut, interval);
      const Rotation<R> r = s.nmat > 1 ? rotation<R>(cosine, p.c, mu, p.phi, phi)
                                       : Rotation<R>{R(1), R(0), R(1), R(0)};
      Stokes<R> st = scatter<R, V4>(f, r, h.st);
      const R norm = st.i;
      if (guide) {
        if (!direct) {
          const R sine = sqrt(fmax(R(0), R(1) - mu * mu));
          da = cos(phi) * sine, db = sin(phi) * sine, dc = mu;
        }
        const float table_norm = h.species < 0 ? 1.0f : s.sp[h.species].norm;
        const R density = norm / (pi * R(table_norm));
        const R q = R(steer_density(*proposal, float(da), float(db), float(dc)));
        const R factor = density / (R(alpha) * density + R(1.0f - alpha) * q);
        h.weight *= factor;
        h.steer *= factor;
        // A direction neither sampler can draw (both densities zero) ends the history.
        if (!(h.weight > R(0))) return kAbsorbed;
      }
      if (s.nmat > 1) {
        st.i = R(1);
        st.q /= norm;
        st.u /= norm;
        if (V4) st.v /= norm;
        const R polarized = V4 ? st.q * st.q + st.u * st.u + st.v * st.v : st.q * st.q + st.u * st.u;
        if (!finite_stokes<R, V4>(st) || sqrt(polarized) > R(1.01)) {
          atomicOr(failure, kInvalidStokes);
          return kFailed;
        }
        h.st = st;
        if (BACK) {
          // The other rows of W P^T, normalized by the same W11 (mc.f90:1393-1395).
          auto carry = [&](Stokes<R> &row) {
            Stokes<R> w = scatter<R, V4>(f, r, row);
            w.i /= norm;
            w.q /= norm;
            w.u /= norm;
            if (V4) w.v /= norm;
            row = w;
          };
          carry(h.w2);
          carry(h.w3);
          if (V4) carry(h.w4);
        }
      }
    }
    h.remains = -log(R(uniform(h.rng)));
  }
  if (direct) {
    change_direction<R>(s, p, da, db, dc, phi);
  } else {
    const R sine = sqrt(fmax(R(0), R(1) - mu * mu));
    change_direction<R>(s, p, cos(phi) * sine, sin(phi) * sine, mu, phi);
  }
  if (TRACE) {
    trace.weight = h.weight;
    trace.layer = R(p.k);
  }
  return kRunning;
}

// Proposal inputs of an atmospheric event: periodic horizontal position (sin, cos of the
// phase in the domain), height over the top of the grid, direction of travel, log extinction
// of the cell, single-scattering albedo of the scatterer, scatterer kind (0 gas, else
// (species + 1) / species count) and scattering order.
template <typename R>
__device__ inline void event_features(const Scene &s, const History<R> &h, float (&f)[monki_net::kInputs]) {
  const Photon<R> &p = h.p;
  const float two_pi = 6.28318530717958647692f;
  const float top = s.height_above[0] + s.zunit[0];
  const float x = float(p.i) * s.xunit + float(p.a >= R(0) ? R(s.xunit) - p.rx : p.rx);
  const float y = float(p.j) * s.yunit + float(p.b >= R(0) ? R(s.yunit) - p.ry : p.ry);
  const float bottom = top - s.height_above[p.k] - s.zunit[p.k];
  const float z = bottom + float(p.c >= R(0) ? R(s.zunit[p.k]) - p.rz : p.rz);
  const float px = two_pi * x / (float(s.nx) * s.xunit), py = two_pi * y / (float(s.ny) * s.yunit);
  f[0] = sinf(px);
  f[1] = cosf(px);
  f[2] = sinf(py);
  f[3] = cosf(py);
  f[4] = 2.0f * z / top - 1.0f;
  f[5] = float(p.a);
  f[6] = float(p.b);
  f[7] = float(p.c);
  f[8] = log1pf(s.beta[cell_index<R>(s, p.i, p.j, p.k)]) / 3.0f;
  f[9] = h.species < 0 ? s.ssalb_gas[p.k] : s.sp[h.species].ssalb;
  f[10] = h.species < 0 ? 0.0f : float(h.species + 1) / float(s.species);
  f[11] = fminf(float(h.order), 64.0f) / 32.0f - 1.0f;
}

__device__ inline float read_lane(float v, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
__device__ inline double read_lane(double v, int src) {
  const unsigned long long bits = static_cast<unsigned long long>(__double_as_longlong(v));
  const unsigned lo = unsigned(__builtin_amdgcn_readlane(int(unsigned(bits)), src));
  const unsigned hi = unsigned(__builtin_amdgcn_readlane(int(unsigned(bits >> 32)), src));
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// The flight of a lane since its open records' last event, when it left their layer (`stale`):
// the layer it ended in (nz after an escape), the path it added there, and 1 / max(|cos|,
// 1e-16) of its direction. The layers between were crossed whole: thickness times that each.
template <typename R> struct Flight {
  bool stale;
  int end_layer;
  R end_path, inverse;
};

// Wave-cooperative tally (mc_routines_cartesian_mod.f90:1173-1186) of open records (Record):
// for each lane in `lanes`, the records of the views in its `views` bit mask. A record is one
// per-(view, pixel) bank record [layer paths 0..nz-1 | Stokes 0..nmat-1] added by a single
// atomic instruction, lane t adding slot t: contiguous lanes make it a few cache-line requests
// instead of one request per value. Path slots (C1 (path_layers + path_layers_angle) summed
// over the record's events): the record's E in its layer k, S (path + full-layer detector
// length) in the layers the detector ray crosses after k, S path in the others, the path read
// from the source lane's LDS row, where a stale lane's paths are its current ones minus its
// flight's; Stokes slots: S mu0/mu, Q, U (V). views, open_layer, round_offset and flight are
// the lanes' own values (read from the source lane). The source lanes run one after another:
// the per-view constants are set up once, outside that loop. Record offsets within a bank are
// 32-bit (session creation keeps them below 2^31 / 8, pixel indexes below 2^24): 24-bit
// multiplies instead of 64-bit index arithmetic. WIDE (kWide): the LDS rows hold layers
// 0 .. low - 1; a source lane's upper paths are its first flight's whole crossings from TOA
// (layers from its descent index up) or its row in global memory (wave_upper, `stride` floats
// per lane) once it has one (upper_word: descent | residual << 7).
template <typename R, bool V4, bool DOWN, bool WIDE = false>
__device__ inline void flush_records(const Scene &s, R *__restrict__ banks, int shard, uint64_t lanes, int views,
                                     const Record<R> &wave_records, const R *wave_history, int row, int open_layer,
                                     uint32_t round_offset, const Flight<R> &flight, Counts &n, int low = 1 << 30,
                                     const float *wave_upper = nullptr, int stride = 0, int upper_word = 0,
                                     const float *layers = nullptr, int table_row = 0, int open_high = -1) {
  const int lane = __lane_id();
  const int record = s.paths ? s.nz + s.nmat : s.nmat;
  const int stokes_at = s.paths ? s.nz : 0;
  const bool path_lane = s.paths && lane < s.nz;
  const bool tally_lane = lane < record;
  const int component = lane - stokes_at;  // Stokes component of a non-path lane
  const uint32_t view_values = uint32_t(s.nx * s.ny) * uint32_t(s.record_stride);
  const R lane_layer = lane < s.nz - 1 ? R(s.zunit[lane]) : R(0);
  const R lane_thickness = lane < s.nz ? R(s.zunit[lane]) : R(0);
  // WIDE: this lane's layer crossed whole by a first flight from TOA.
  const bool upper_lane = WIDE && path_lane && lane >= low;
  const R lane_descent = upper_lane ? whole_crossing<R>(layers, table_row, lane, R(s.launch_c)) : R(0);
  R geometry[kMaxViews], layer_length[kMaxViews];
  uint32_t lane_index[kMaxViews];
#pragma unroll
  for (int view = 0; view < kMaxViews; ++view) {
    geometry[view] = component == 0 ? R(s.view_geometry[view]) : R(0);
    layer_length[view] = lane_layer * R(s.view_inv_mu[view]);
    lane_index[view] = uint32_t(view) * view_values + uint32_t(lane);
  }
  // The source lane's small words in one register: views (3 bits), layer + 1 (7), stale (1),
  // end layer + 1 (7); layers stay below 64.
  const int word = (views & 7) | ((open_layer + 1) << 3) | (int(flight.stale) << 10) | ((flight.end_layer + 1) << 11) |
                   (WIDE ? (upper_word << 18) | (int(open_high > open_layer) << 26) : 0);
  for (uint64_t mask = lanes; mask; mask &= mask - 1) {
    const int src = __ffsll(static_cast<long long>(mask)) - 1;
    const int packed = __builtin_amdgcn_readlane(word, src);
    const int chosen = packed & 7;
    const int layer = ((packed >> 3) & 127) - 1;
    const uint32_t round = uint32_t(__builtin_amdgcn_readlane(int(round_offset), src));
    const Record<R> source = wave_records.at(src);
    R *bank = banks + s.bank_base + (size_t(round) * s.shards + shard) * size_t(s.bank_count);
    R path = path_lane && !upper_lane ? wave_history[src * row + lane] : R(0);
    if (upper_lane) {
      const int descent = (packed >> 18) & 127;
      path = (packed >> 25) & 1 ? R(wave_upper[size_t(src) * stride + (lane - low)])
                                : (lane >= descent ? lane_descent : R(0));
    }
    // WIDE: records of two adjacent layers, layer and high (E of high in the fifth field).
    const int high = WIDE ? layer + ((packed >> 26) & 1) : layer;
    if ((packed >> 10) & 1) {
      const int end = ((packed >> 11) & 127) - 1;
      const bool between = end > high ? (lane > high && lane < end) : (lane < layer && lane > end);
      const R crossed = between ? lane_thickness * read_lane(flight.inverse, src) : R(0);
      const R added = lane == end ? read_lane(flight.end_path, src) : crossed;
      // (A layer first entered on this flight comes back as rounding around zero.) WIDE: an
      // escaping flight's upper crossings were never added.
      if (path_lane && !(upper_lane && end >= s.nz)) path = fmax(R(0), path - added);
    }
    const bool at_layer = lane == layer, at_high = WIDE && high != layer && lane == high;
    const bool beyond = DOWN ? lane < layer : lane > high;
#pragma unroll
    for (int view = 0; view < kMaxViews; ++view) {
      if (!((chosen >> view) & 1)) continue;
      const int pixel = source.pixel(view);
      const R weight = source.s(view);
      R stokes = weight * geometry[view];
      stokes = component == 1 ? source.q(view) : stokes;
      stokes = component == 2 ? source.u(view) : stokes;
      if (V4) stokes = component == 3 ? source.v(view) : stokes;
      const R value = path_lane ? (at_layer  ? source.e(view)
                                   : at_high ? source.v(view)
                                             : weight * (path + (beyond ? layer_length[view] : R(0))))
                                : stokes;
      if (pixel >= 0 && tally_lane && value != R(0)) {
        add_tally<R>(bank + (__umul24(uint32_t(pixel), uint32_t(s.record_stride)) + lane_index[view]), value);
        ++n.atomics;
      }
    }
  }
}

// Wave-cooperative tally of the scattering decomposition (calc_radiances_and_scats,
// mc_routines_cartesian_mod.f90:405-836). Per view, one atomic instruction adds the event's
// Stokes contribution to two 8-value records [single (nmat) | multiple (nmat)]: lanes
// 0 .. 2 nmat - 1 to the record of the detector pixel (Stokes_sing / Stokes_mult), lanes
// 2 nmat .. 4 nmat - 1 to the record of the event cell (the 3-D contributions); the half
// matching the event's scattering order (1: single) carries the value. The pending layer
// word holds the cell (i + nx (j + ny level)), complemented for orders above one.
template <typename R, bool V4>
__device__ inline void tally_wave_scats(const Scene &s, R *__restrict__ banks, int shard, bool has,
                                        const Pending<R> &wave_pending, Counts &n) {
  const int lane = __lane_id();
  const int nmat = s.nmat;
  const bool cell_lane = lane >= 2 * nmat;
  const int slot = cell_lane ? lane - 2 * nmat : lane;
  const bool tally_lane = lane < 4 * nmat;
  const int component = slot % nmat, group = slot / nmat;  // group 0: single scattering
  const uint32_t plane = uint32_t(s.nx * s.ny);
  const uint32_t cells = plane * uint32_t(s.nz);
  const uint32_t pixel_values = plane * uint32_t(s.views) * 8u;
  R geometry[kMaxViews];
#pragma unroll
  for (int view = 0; view < kMaxViews; ++view) geometry[view] = R(s.view_geometry[view]);
  const R stokes_q = component == 1 ? R(1) : R(0), stokes_u = component == 2 ? R(1) : R(0);
  const R stokes_v = component == 3 ? R(1) : R(0), stokes_i = component == 0 ? R(1) : R(0);
  for (uint64_t mask = __ballot(has); mask; mask &= mask - 1) {
    const int src = __ffsll(static_cast<long long>(mask)) - 1;
    const Pending<R> source = wave_pending.at(src);
    const int code = source.layer();
    const bool single = code >= 0;
    const uint32_t cell = uint32_t(single ? code : ~code);
    R *bank = banks + s.bank_base + (size_t(source.round()) * s.shards + shard) * size_t(s.bank_count);
    const bool carries = tally_lane && group == (single ? 0 : 1);
#pragma unroll
    for (int view = 0; view < kMaxViews; ++view) {
      if (view >= s.views) break;
      const int pixel = source.pixel(view);
      const R q = source.q(view), u = source.u(view);
      R value = source.weight(view) * geometry[view] * stokes_i + q * stokes_q + u * stokes_u;
      if (V4) value += source.v(view) * stokes_v;
      const uint32_t offset = cell_lane ? pixel_values + (uint32_t(view) * cells + cell) * 8u + uint32_t(slot)
                                        : (uint32_t(view) * plane + uint32_t(pixel)) * 8u + uint32_t(slot);
      if (pixel >= 0 && carries && value != R(0)) {
        add_tally<R>(bank + offset, value);
        ++n.atomics;
      }
    }
  }
}

__device__ inline unsigned long long wave_sum(unsigned long long v) {
  for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_xor(v, offset, 64);
  return v;
}

__device__ inline void flush_counts(Counts &n, unsigned long long *counters) {
  const uint32_t values[kCounters] = {n.photons, n.events, n.surface, n.walk,
                                      n.detector, n.lut, n.attempts, n.atomics};
  for (int i = 0; i < kCounters; ++i)
    if (values[i]) atomicAdd(counters + i, static_cast<unsigned long long>(values[i]));
  n = Counts{0, 0, 0, 0, 0, 0, 0, 0};
}

// The proposal's mixture for one 32-event block of the wave's guided events: the network on
// the matrix cores, all lanes together. Guided events are numbered by lane (rank r = guided
// lanes below the lane); block nb holds ranks 32 nb .. 32 nb + 31, sample column r % 32 taking
// the features of its lane, so a wave with at most 32 guided events runs one pass. Not
// inlined: the network needs most of the register file, and a call saves the photon state
// around it once instead of the compiler spilling it piecemeal through the network.
__device__ __attribute__((noinline)) void guided_mixture(const monki_net::Net &net,
                                                          const float (&features)[monki_net::kInputs],
                                                          unsigned long long guided_lanes, int nb, bool mine,
                                                          monki_net::Mixture &mixture) {
  const int lane = __lane_id();
  const int sample = 32 * nb + (lane & 31);
  const int source = sample < __popcll(guided_lanes) ? monki_net::nth_set_lane(guided_lanes, sample) : -1;
  monki_net::bf16x8 x[8];
  monki_net::scatter_inputs(features, source, x);
  monki_net::floatx16 raw[2];
  monki_net::forward_block(net, x, raw, nullptr);
  float o[monki_net::kOutputs];
  monki_net::gather_outputs(raw, __popcll(guided_lanes & __lanemask_lt()) & 31, mine, o);
  if (mine) mixture = monki_net::mixture_from(o);
}

// Weight-window copies (steered modes). Each lane keeps a stack of kSplitDepth entries in global
// memory; an entry holds the photon state right after a scattering (position, direction,
// Stokes vector, weight of one copy, order, counters) and the photon's layer paths, plus the
// number of copies still to run. A lane whose history ends resumes the top entry's next copy
// on a random stream of its own (stream = number of copies started in this photon's tree, so
// histories stay independent of lanes and timing) before it takes a new photon. Copies of one
// photon tally like the photon itself. At most kTreeCopies copies start per photon, and the
// history step limit counts from the photon's launch, so every tree ends.
constexpr int kSplitDepth = 8;
constexpr int kSplitWords = 21;  // state words before the layer paths
constexpr uint32_t kMaxCopies = 4;
constexpr uint32_t kTreeCopies = 4096;

template <typename R>
__device__ inline void push_copies(uint32_t *entry, const History<R> &h, const R *history, int nz, uint32_t copies) {
  const Photon<R> &p = h.p;
  entry[0] = uint32_t(p.i);
  entry[1] = uint32_t(p.j);
  entry[2] = uint32_t(p.k);
  entry[3] = __float_as_uint(float(p.rx));
  entry[4] = __float_as_uint(float(p.ry));
  entry[5] = __float_as_uint(float(p.rz));
  entry[6] = __float_as_uint(float(p.a));
  entry[7] = __float_as_uint(float(p.b));
  entry[8] = __float_as_uint(float(p.c));
  entry[9] = __float_as_uint(float(p.phi));
  entry[10] = __float_as_uint(float(h.st.i));
  entry[11] = __float_as_uint(float(h.st.q));
  entry[12] = __float_as_uint(float(h.st.u));
  entry[13] = __float_as_uint(float(h.weight));
  entry[14] = h.order;
  entry[15] = h.round_offset;
  entry[16] = h.steps;
  entry[17] = __float_as_uint(float(h.path_total));
  entry[18] = __float_as_uint(h.contribution);
  entry[19] = copies;
  entry[20] = __float_as_uint(float(h.steer));
  for (int k = 0; k < nz; ++k) entry[kSplitWords + k] = __float_as_uint(float(history[k]));
}

// Restore the next copy of `entry` into h and the lane's layer paths; returns the copies left.
template <typename R>
__device__ inline uint32_t pop_copy(uint32_t *entry, History<R> &h, R *history, int nz) {
  Photon<R> &p = h.p;
  p.i = int(entry[0]);
  p.j = int(entry[1]);
  p.k = int(entry[2]);
  p.rx = R(__uint_as_float(entry[3]));
  p.ry = R(__uint_as_float(entry[4]));
  p.rz = R(__uint_as_float(entry[5]));
  p.a = R(__uint_as_float(entry[6]));
  p.b = R(__uint_as_float(entry[7]));
  p.c = R(__uint_as_float(entry[8]));
  p.phi = R(__uint_as_float(entry[9]));
  h.st = Stokes<R>{R(__uint_as_float(entry[10])), R(__uint_as_float(entry[11])), R(__uint_as_float(entry[12])), R(0)};
  h.weight = R(__uint_as_float(entry[13]));
  h.order = entry[14];
  h.round_offset = entry[15];
  h.steps = entry[16];
  h.path_total = R(__uint_as_float(entry[17]));
  h.contribution = __uint_as_float(entry[18]);
  h.steer = R(__uint_as_float(entry[20]));
  for (int k = 0; k < nz; ++k) history[k] = R(__uint_as_float(entry[kSplitWords + k]));
  const uint32_t left = entry[19] - 1u;
  entry[19] = left;
  return left;
}

// Persistent transport. Each iteration, every lane advances its photon to the next event and
// evaluates the local estimates (divergent work), the wave then tallies all pending events
// cooperatively, and the lanes sample their continuation directions. A lane whose history
// ended takes the next photon id from a wave-aggregated work counter, so a wave never idles
// behind its longest history. Photon ids (and with them the random streams and the order of
// draws) do not depend on which lane runs them. Tally banks are chosen per wave.
// Three waves per SIMD is what the LDS rows of a 256-lane block allow; the register budget is
// capped to match (float32), which measured 1.35x faster than the compiler's own choice of
// two waves per SIMD on the layered-cloud round. The guided mode evaluates the proposal network
// for the wave's scattering events between the tally and the scatter phases (matrix cores,
// all lanes together, 32 events per pass); the network's call fits the same budget without
// spilling (box-cloud round 223 -> 193 ms against two waves per SIMD).
template <typename R, bool TRACE, int MODE>
__global__ void __launch_bounds__(kBlock32)
    __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 3 : 1)))
    transport_kernel(const Scene *__restrict__ scenes, const Controls *__restrict__ controls,
                     uint32_t scene_count, int row, R *banks, unsigned *failure, unsigned *work,
                     unsigned long long *counters, float *records, uint32_t record_count,
                     monki_net::Net net, monki_net::EventLog log, Guidance guidance, uint32_t *splits,
                     int split_stride, int wide) {
  constexpr bool V4 = full_stokes(MODE);
  // kWide: LDS layer rows of `wide` values (layers 0 .. wide - 1); splits / split_stride are the
  // lanes' rows of upper-layer paths in global memory.
  constexpr bool WIDE = MODE == kWide;
  const int hrow = WIDE ? wide : row;
  constexpr bool BACK = backward(MODE);
  constexpr bool SCATS = decomposition(MODE);
  constexpr bool DOWN = downward(MODE);
  constexpr bool STEER = MODE == kGuided || MODE == kDirected;
  // This lane's weight-window stack (steered modes), its entries in use, and the copies started
  // in the current photon's tree.
  uint32_t *stack = STEER && splits ? splits + size_t(blockIdx.x * blockDim.x + threadIdx.x) * kSplitDepth * split_stride
                                    : nullptr;
  int top = 0;
  uint32_t tree = 0;
  uint32_t logged = 0;  // events of the photon logged so far (logging mode)
  // Logging with a path target: sum over the photon's estimates so far of their detector weight
  // times the weighted path at their event, and the weighted path at the last event.
  float path_contribution = 0.0f, logged_path = 0.0f;
  const bool window = STEER && stack && guidance.split_above > 0.0f;
  extern __shared__ unsigned char lds_raw[];
  // Layer rows of odd length (the largest nz of the scenes, | 1): lanes reading one row, or
  // each its own row at one layer, touch distinct LDS banks.
  R *history = reinterpret_cast<R *>(lds_raw) + threadIdx.x * hrow;
  const R *wave_history = reinterpret_cast<R *>(lds_raw) + (threadIdx.x & ~63u) * hrow;
  // (The wave's first lane index is wave-uniform: the row base stays in scalar registers.)
  const uint32_t wave_lane0 = __builtin_amdgcn_readfirstlane(blockIdx.x * blockDim.x + (threadIdx.x & ~63u));
  const float *const wave_upper =
      WIDE ? reinterpret_cast<const float *>(splits) + size_t(wave_lane0) * split_stride : nullptr;
  float *const upper =
      WIDE ? reinterpret_cast<float *>(splits) + (size_t(wave_lane0) + __lane_id()) * split_stride : nullptr;
  Upper u{0u};
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
  // Scenes of a batch: waves start spread over them, each wave works on one scene at a time
  // (scene constants stay wave-uniform) and moves to the next one when its scene has no
  // photons left and its lanes are idle; a wave visits every scene at most once.
  uint32_t current = __builtin_amdgcn_readfirstlane(uint32_t((uint64_t(wave) * scene_count) / waves));
  uint32_t visited = 0;
  const int block = int(blockDim.x);
  // After the layer rows: the block's pending local estimates (scattering decomposition) or
  // open records (every other mode), struct of arrays over the block's lanes.
  R *const area = reinterpret_cast<R *>(lds_raw) + size_t(block) * hrow;
  // kWide records span up to two adjacent layers: the record area of the full Stokes vector,
  // whose fifth field holds the E of the pair's upper layer.
  constexpr bool PAIRS = WIDE;
  const Pending<R> wave_pending{area + (threadIdx.x & ~63u),
                                reinterpret_cast<int *>(area + size_t(block) * pending_reals(V4 || PAIRS)) +
                                    (threadIdx.x & ~63u),
                                block};
  const Pending<R> pending = wave_pending.at(__lane_id());
  const Record<R> wave_records{area + (threadIdx.x & ~63u),
                               reinterpret_cast<int *>(area + size_t(block) * record_reals(V4 || PAIRS)) +
                                   (threadIdx.x & ~63u),
                               block};
  const Record<R> record = wave_records.at(__lane_id());
  // Per-wave layer table after that area (plain-width Stokes modes only: the full Stokes
  // vector's rows leave no room at three blocks per CU).
  constexpr bool LAYERS = !V4;
  float *layers = reinterpret_cast<float *>(reinterpret_cast<unsigned char *>(area) +
                                            size_t(block) * pending_bytes<R>(V4 || PAIRS)) +
                  (threadIdx.x >> 6) * 4 * row;
  uint32_t layers_scene = 0xffffffffu;
  Counts n{0, 0, 0, 0, 0, 0, 0, 0};
  History<R> h;
  Trace<R> trace;
  bool active = false;
  uint32_t id = 0;
  // Layer of the lane's open records, -1 while it has none (every mode but the decomposition).
  constexpr bool MERGE = !SCATS;
  int open_layer = -1;
  // PAIRS: the upper layer of the records' pair (open_layer itself for one-layer records).
  int open_high = -1;

  for (;;) {
    const Scene &s = scenes[current];
    const Controls &c = controls[current];
    const int shard = int(wave % uint32_t(s.shards));
    if (LAYERS && layers_scene != current) {
      // The wave's scene changed (or first iteration): load its layer table (wave-uniform).
      for (int k = __lane_id(); k < s.nz; k += 64) {
        layers[k] = s.uniform[k] != 0.0f ? s.beta[cell_index<R>(s, 0, 0, k)] : -1.0f;
        layers[row + k] = s.zunit[k];
        layers[2 * row + k] = s.height_above[k];
        layers[3 * row + k] = s.tau_above[k];
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      layers_scene = current;
    }
    const unsigned long long idle = __ballot(!active);
    if (idle) {
      const int leader = __ffsll(static_cast<long long>(idle)) - 1;
      uint32_t base = 0;
      if (__lane_id() == leader) base = atomicAdd(work + current, uint32_t(__popcll(idle)));
      base = __shfl(base, leader, 64);
      if (!active) {
        id = base + uint32_t(__popcll(idle & __lanemask_lt()));
        if (id < c.count) {
          launch_photon<R>(s, c, id, history, h, hrow);
          if (WIDE) u = Upper{uint32_t(s.nz) | 256u};
          trace = Trace<R>{R(0), R(0), R(0), R(1), R(0), R(0), R(0), R(0)};
          ++n.photons;
          active = true;
          tree = 0;
          logged = 0;
          path_contribution = 0.0f;
        }
      }
    }
    if (__ballot(active) == 0) {
      if (++visited >= scene_count) break;
      current = __builtin_amdgcn_readfirstlane((current + 1) % scene_count);
      continue;
    }

    int status = kRunning;
    prefetch(h.rng);
    Flight<R> flight{false, 0, R(0), R(1)};
    if (!MERGE) {
      if (active)
        status = collide<R, TRACE, V4, BACK, SCATS, LAYERS, DOWN>(s, c, failure, history, h, pending, n, trace,
                                                                  layers, row);
      // Lanes read each other's LDS rows below: keep every access of this wave in order.
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      tally_wave_scats<R, V4>(s, banks, shard, active && status == kRunning, wave_pending, n);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    } else {
      bool surface = false;
      R layer_path = R(0);
      if (active) {
        status = walk<R, LAYERS, WIDE>(s, c, failure, history, h, n, layers, row, surface, layer_path, &u, upper, hrow);
        if (WIDE) u.word &= ~256u;
      }
      // A flight that left the layer of the open records changed the layer paths the records
      // hold outside that layer: their write-out takes this flight's part off again.
      flight = Flight<R>{active && open_layer >= 0 && (PAIRS ? h.p.k < open_layer || h.p.k > open_high
                                                             : h.p.k != open_layer),
                         h.p.k, layer_path, R(1) / fmax(fabs(h.p.c), R(1e-16))};
      if (active && status == kRunning) status = interact<R, TRACE, false>(s, h, n, trace, surface);
      const bool event = active && status == kRunning;
      R rz = R(0), delta = R(0), sin_delta = R(0), cos_delta = R(0), path = R(0);
      if (event) {
        rz = event_rz<R, LAYERS, DOWN>(s, h.p, layers, row);
        // All views share one azimuth: cos and sin of the photon-view azimuth difference, once.
        delta = h.p.phi - R(s.phi);
        sin_cos(delta, &sin_delta, &cos_delta);
        path = WIDE && h.p.k >= hrow ? upper_path<R>(s, u, upper, hrow, layers, row, h.p.k) : history[h.p.k];
      }
      // The local estimates toward every view, lane by lane (the detector walks of one lane run
      // back to back, so the wave waits once for its longest lane).
      Estimate<R> e0{R(0), R(0), R(0), R(0), -1}, e1 = e0, e2 = e0;
      // Logging with a path target: the weighted path of the history up to this event (its
      // layer paths times the layer weights) and the detector weight before its estimates.
      float event_path = 0.0f, weight_before = 0.0f;
      if (MODE == kLog && log.layer_weight && event) {
        for (int k = 0; k < s.nz; ++k) event_path += float(history[k]) * log.layer_weight[k];
        weight_before = h.contribution;
      }
      if (event) {
#pragma unroll 1
        for (int view = 0; view < s.views; ++view) {
          const Estimate<R> e = estimate<R, TRACE, V4, BACK, LAYERS, DOWN>(s, c, failure, h, view, delta, cos_delta,
                                                                           sin_delta, rz, n, trace, layers, row, status);
          if (status != kRunning) break;
          if (view == 0)
            e0 = e;
          else if (view == 1)
            e1 = e;
          else
            e2 = e;
        }
      }
      if (MODE == kLog && log.layer_weight && event) {
        path_contribution += (h.contribution - weight_before) * event_path;
        logged_path = event_path;
      }
      if constexpr (PAIRS) {
        // Records of up to two adjacent layers. A record whose events all lie in one layer takes
        // in the next layer up or down when the photon's flight went straight there and its
        // pixel stays (or the view gets nothing from the event): its past events' E for the new
        // layer is S (that layer's path before this flight [+ the full detector-ray length when
        // the layer lies above theirs]); every event then adds C1 (path + detector-ray length)
        // to the E of both layers. The layer paths outside the pair stay those of the events.
        const int k = h.p.k;
        const bool left = flight.stale;
        const bool adjacent = event && left && open_layer == open_high && (k == open_layer + 1 || k == open_layer - 1);
        int survive = 0;
#pragma unroll 1
        for (int view = 0; view < s.views; ++view) {
          const int pixel = view == 0 ? e0.pixel : (view == 1 ? e1.pixel : e2.pixel);
          const int held = event && open_layer >= 0 ? record.pixel(view) : -1;
          if (held >= 0 && (!left || adjacent) && (pixel < 0 || pixel == held)) survive |= 1 << view;
        }
        const bool extend = adjacent && survive != 0;
        int flushed = 0;
#pragma unroll 1
        for (int view = 0; view < s.views; ++view) {
          const int held = event && open_layer >= 0 ? record.pixel(view) : -1;
          if (held >= 0 && !((survive >> view) & 1)) flushed |= 1 << view;
        }
        const unsigned long long flushing = __ballot(flushed != 0);
        if (flushing) {
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          __builtin_amdgcn_wave_barrier();
          flush_records<R, V4, DOWN, WIDE>(s, banks, shard, flushing, flushed, wave_records, wave_history, hrow,
                                           open_layer, h.round_offset, flight, n, hrow, wave_upper, split_stride,
                                           int(u.word & 255u), layers, row, open_high);
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          __builtin_amdgcn_wave_barrier();
        }
        const bool fresh = left || open_layer < 0;
        const int low = extend ? min(open_layer, k) : (fresh ? k : open_layer);
        const int high = extend ? max(open_layer, k) : (fresh ? k : open_high);
        auto path_at = [&](int layer) {
          return layer >= hrow ? upper_path<R>(s, u, upper, hrow, layers, row, layer) : history[layer];
        };
        R path_low = R(0), path_high = R(0);
        if (event) {
          path_low = path_at(low);
          path_high = high != low ? path_at(high) : path_low;
        }
        const R path_new = k == low ? path_low : path_high;
        bool open = false;
#pragma unroll 1
        for (int view = 0; view < s.views; ++view) {
          const Estimate<R> e = view == 0 ? e0 : (view == 1 ? e1 : e2);
          const bool keep = (survive >> view) & 1;
          const R inv_mu = R(s.view_inv_mu[view]);
          if (keep && extend) {
            // The past events' E for the layer the record takes in.
            const R before = path_new - flight.end_path;
            const R past = record.s(view) * (k > open_layer ? before + R(layers[row + k]) * inv_mu : before);
            if (k > open_layer) {
              record.v(view) = past;
            } else {
              record.v(view) = record.e(view);
              record.e(view) = past;
            }
          }
          if (e.pixel >= 0) {
            // Detector-ray length in the event layer (its part above the event), the full
            // layer above it, none below.
            const R part_low = e.weight * (path_low + (low == k ? rz * inv_mu : R(0)));
            const R part_high =
                e.weight * (path_high + (high == k ? rz * inv_mu : R(layers[row + high]) * inv_mu));
            if (keep) {
              record.e(view) += part_low;
              if (high != low) record.v(view) += part_high;
              record.s(view) += e.weight;
              record.q(view) += e.q;
              record.u(view) += e.u;
            } else {
              record.e(view) = part_low;
              record.v(view) = high != low ? part_high : R(0);
              record.s(view) = e.weight;
              record.q(view) = e.q;
              record.u(view) = e.u;
              record.pixel(view) = e.pixel;
            }
            open = true;
          } else if (event) {
            if (!keep) record.pixel(view) = -1;
            open = open || keep;
          }
        }
        if (event) {
          open_layer = open ? low : -1;
          open_high = open ? high : -1;
          flight.stale = false;
        }
      } else {
        // Per view: the event's local estimate joins the open record of the same pixel and layer,
        // or the open record is written out first (another pixel, or the photon left its layer)
        // and the event opens a new one. One write-out for all views of all lanes. (Lanes read
        // each other's LDS rows and records: every access of the wave in order around it.)
        int flushed = 0;
#pragma unroll 1
        for (int view = 0; view < s.views; ++view) {
          const int pixel = view == 0 ? e0.pixel : (view == 1 ? e1.pixel : e2.pixel);
          const int held = event && open_layer >= 0 ? record.pixel(view) : -1;
          if (held >= 0 && (flight.stale || (pixel >= 0 && pixel != held))) flushed |= 1 << view;
        }
        const unsigned long long flushing = __ballot(flushed != 0);
        if (flushing) {
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          __builtin_amdgcn_wave_barrier();
          flush_records<R, V4, DOWN, WIDE>(s, banks, shard, flushing, flushed, wave_records, wave_history, hrow,
                                           open_layer, h.round_offset, flight, n, hrow, wave_upper, split_stride,
                                           int(u.word & 255u), layers, row);
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          __builtin_amdgcn_wave_barrier();
        }
        bool open = false;
#pragma unroll 1
        for (int view = 0; view < s.views; ++view) {
          const Estimate<R> e = view == 0 ? e0 : (view == 1 ? e1 : e2);
          const int held = event && open_layer >= 0 ? record.pixel(view) : -1;
          const bool same = !flight.stale && held == e.pixel;
          if (e.pixel >= 0) {
            const R part = e.weight * (path + rz * R(s.view_inv_mu[view]));
            if (same) {
              record.e(view) += part;
              record.s(view) += e.weight;
              record.q(view) += e.q;
              record.u(view) += e.u;
              if (V4) record.v(view) += e.v;
            } else {
              record.e(view) = part;
              record.s(view) = e.weight;
              record.q(view) = e.q;
              record.u(view) = e.u;
              if (V4) record.v(view) = e.v;
              record.pixel(view) = e.pixel;
            }
            open = true;
          } else if (event) {
            // Nothing toward this view: a record of the same layer stays open.
            if (held < 0 || flight.stale) record.pixel(view) = -1;
            open = open || (held >= 0 && !flight.stale);
          }
        }
        if (event) {
          // The records now hold the layer paths of this event.
          open_layer = open ? h.p.k : -1;
          flight.stale = false;
        }
      }
      }
    // Russian roulette of the weight window, after the event's local estimates (a photon steered
    // toward a view has used its small weight there) and before its scattering.
    if (window && active && status == kRunning && h.steer < R(guidance.roulette_below)) {
      if (R(uniform(h.rng)) * R(guidance.roulette_weight) < h.steer) {
        h.weight *= R(guidance.roulette_weight) / h.steer;
        h.steer = R(guidance.roulette_weight);
      } else {
        status = kAbsorbed;
      }
    }
    const bool scatters = active && status == kRunning && !h.surface;
    float features[monki_net::kInputs];
    monki_net::Mixture mixture;
    // The network runs when some lane's event is guided (kind and order within the guidance,
    // order within the network's orders).
    const bool guided_event = MODE == kGuided && scatters &&
                              ((guidance.species_mask >> (h.species + 1)) & 1u) && h.order <= guidance.max_order &&
                              h.order <= guidance.network_orders;
    if (MODE == kLog || MODE == kGuided) {
#pragma unroll
      for (int k = 0; k < monki_net::kInputs; ++k) features[k] = 0.0f;
      if (MODE == kLog ? scatters : guided_event) event_features<R>(s, h, features);
    }
    const unsigned long long guided_lanes =
        MODE == kGuided && guidance.lobe_fraction < 1.0f ? __ballot(guided_event) : 0ull;
    if (MODE == kGuided && guided_lanes) {
      const int rank = __popcll(guided_lanes & __lanemask_lt());
      for (int nb = 0; 32 * nb < __popcll(guided_lanes); ++nb)
        guided_mixture(net, features, guided_lanes, nb, guided_event && (rank >> 5) == nb, mixture);
    }
    prefetch(h.rng);
    const ViewLobes lobes{&s, guidance.lobe_species, guidance.kappa};
    if (MODE == kDirected) {
      if (active && status == kRunning)
        status = scatter_photon<R, TRACE, MODE>(s, failure, h, n, trace, &lobes, &lobes, guidance);
    } else if (MODE == kGuided) {
      const GuidedProposal guided{&mixture, lobes,
                                  h.order <= guidance.network_orders ? guidance.lobe_fraction : 1.0f};
      if (active && status == kRunning)
        status = scatter_photon<R, TRACE, MODE>(s, failure, h, n, trace, &guided, &lobes, guidance);
    } else if (active && status == kRunning) {
      status = scatter_photon<R, TRACE, MODE>(s, failure, h, n, trace, &lobes, &lobes, guidance);
    }
    // Splitting of the weight window: the photon continues as `copies` copies of equal weight.
    if (window && active && status == kRunning && h.steer > R(guidance.split_above) && top < kSplitDepth &&
        tree < kTreeCopies) {
      const uint32_t copies = min(uint32_t(ceil(float(h.steer))), kMaxCopies);
      h.weight /= R(copies);
      h.steer /= R(copies);
      push_copies<R>(stack + size_t(top) * split_stride, h, history, s.nz, copies - 1u);
      ++top;
    }
    if (MODE == kLog && scatters && status == kRunning) {
      // The counting pass only counts the photon's events; the writing pass puts its k-th event
      // at the photon's start (prefix sum of the counts) + k: samples in photon, then event
      // order, whatever the timing of the waves.
      const uint32_t slot = log.counting ? 0xffffffffu : log.starts[id] + logged;
      ++logged;
      if (slot < log.capacity) {
#pragma unroll
        for (int k = 0; k < monki_net::kInputs; ++k) log.features[size_t(slot) * monki_net::kInputs + k] = features[k];
        log.directions[size_t(slot) * 3] = float(h.p.a);
        log.directions[size_t(slot) * 3 + 1] = float(h.p.b);
        log.directions[size_t(slot) * 3 + 2] = float(h.p.c);
        log.cumulative[slot] = h.contribution;
        log.photon[slot] = id;
        if (log.layer_weight) {
          log.path_cumulative[slot] = path_contribution;
          log.event_path[slot] = logged_path;
        }
      }
    }
    if (MERGE) {
      // Histories (and copies of split ones) that ended write out their open records before a
      // new photon or copy takes the lane's layer row.
      const bool ends = active && status != kRunning;
      const unsigned long long ending = __ballot(ends && open_layer >= 0);
      if (ending) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        flush_records<R, V4, DOWN, WIDE>(s, banks, shard, ending, (1 << s.views) - 1, wave_records, wave_history,
                                         hrow, open_layer, h.round_offset, flight, n, hrow, wave_upper, split_stride,
                                         int(u.word & 255u), layers, row, open_high);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
      if (ends) open_layer = open_high = -1;
    }
    if (!active || status == kRunning) continue;
    if (STEER && top > 0) {
      // The next copy of the photon resumes right after the scattering it was split at.
      if (pop_copy<R>(stack + size_t(top - 1) * split_stride, h, history, s.nz) == 0u) --top;
      ++tree;
      h.rng = Stream{h.rng.photon, h.rng.round, h.rng.seed, tree, 0u, Words{0, 0, 0, 0}, Words{0, 0, 0, 0}, 4, false};
      h.remains = -::log(R(uniform(h.rng)));
      continue;
    }
    active = false;
    if (MODE == kLog && id < log.photons) {
      log.totals[id] = h.contribution;
      if (log.layer_weight) log.path_totals[id] = path_contribution;
      if (log.counting) log.starts[id] = logged;
    }
    if (TRACE && id < record_count) {
      float *out = records + size_t(id) * kTraceWords;
      out[0] = float(trace.events);
      out[1] = float(trace.surface);
      out[2] = float(status);
      out[3] = float(trace.weight);
      out[4] = float(trace.stokes_i);
      out[5] = float(trace.weights);
      out[6] = float(trace.paths);
      out[7] = float(trace.layer);
    }
    // Keep the 32-bit lane counters far from overflow on very long dispatches.
    if ((n.walk | n.detector | n.lut | n.attempts | n.atomics) >= (1u << 30)) flush_counts(n, counters);
  }

  const uint32_t values[kCounters] = {n.photons, n.events, n.surface, n.walk,
                                      n.detector, n.lut, n.attempts, n.atomics};
#pragma unroll
  for (int i = 0; i < kCounters; ++i) {
    const unsigned long long total = wave_sum(static_cast<unsigned long long>(values[i]));
    if (__lane_id() == 0 && total) atomicAdd(counters + i, total);
  }
}

// Sum the shard banks of each round in float64: out[r][t] = sum_s banks[r][s][t].
template <typename B>
__global__ void reduce_shards(const B *__restrict__ banks, double *__restrict__ out, int bank_count,
                              int shards, int rounds) {
  const size_t index = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (index >= size_t(bank_count) * rounds) return;
  const size_t r = index / bank_count, t = index % bank_count;
  const B *base = banks + r * size_t(shards) * bank_count + t;
  double sum = 0.0;
  for (int sh = 0; sh < shards; ++sh) sum += double(base[size_t(sh) * bank_count]);
  out[index] = sum;
}

// Output tallies of each round, in Fortran order: Stokes (nmat, nx, ny, views), layer paths
// (nx, ny, nz, views) and detector weights (nx, ny, views), from the banks' per-(view, pixel)
// records. The weight C1 of every estimate is its Stokes-I contribution over mu0/mu.
__global__ void assemble_tallies(Scene s, const double *__restrict__ reduced,
                                 double *__restrict__ out, int rounds) {
  const int plane = s.nx * s.ny;
  if (s.scats) {
    // [Stokes | Stokes_sing | Stokes_mult | contrib_sing | contrib_mult], Fortran order; the
    // Stokes tally is the sum of the single- and multiple-scattering records.
    const size_t bank = size_t(s.bank_count);
    const size_t index = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (index >= bank * rounds) return;
    const size_t r = index / bank, t = index % bank;
    const size_t pixel_values = size_t(plane) * s.views * 8;
    const bool cell_region = t >= pixel_values;
    const size_t u = cell_region ? t - pixel_values : t;
    const int slot = int(u % 8);
    if (slot >= 2 * s.nmat) return;  // record padding
    const int m = slot % s.nmat, group = slot / s.nmat;
    const size_t record = u / 8;
    const double value = reduced[r * bank + t];
    double *dst = out + r * size_t(s.tally_count);
    if (!cell_region) {
      const size_t at = m + size_t(s.nmat) * record;  // record = pixel + plane view
      dst[size_t(1 + group) * s.stokes_count + at] = value;
      if (group == 0) dst[at] = value + reduced[r * bank + t + s.nmat];
    } else {
      const size_t contrib = size_t(s.nmat) * s.path_count;  // record = cell + cells view
      dst[3 * size_t(s.stokes_count) + group * contrib + m + size_t(s.nmat) * record] = value;
    }
    return;
  }
  const int record = s.paths ? s.nz + s.nmat : s.nmat;
  const int stride = s.record_stride;
  const size_t per_round = size_t(plane) * s.views * stride;
  const size_t index = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (index >= per_round * rounds) return;
  const size_t r = index / per_round, t = index % per_round;
  const int slot = int(t % stride);
  if (slot >= record) return;  // record padding
  const int pixel = int((t / stride) % plane);
  const int view = int(t / (size_t(stride) * plane));
  const double value = reduced[r * size_t(s.bank_count) + t];
  double *dst = out + r * size_t(s.tally_count);
  const int stokes_at = s.paths ? s.nz : 0;
  if (slot >= stokes_at) {
    const int m = slot - stokes_at;
    dst[m + size_t(s.nmat) * (pixel + size_t(plane) * view)] = value;
    if (s.paths && m == 0) {
      const double geometry = s.view_geometry[view];
      dst[size_t(s.stokes_count) + s.path_count + pixel + size_t(plane) * view] = value / geometry;
    }
  } else {
    dst[size_t(s.stokes_count) + pixel + size_t(plane) * (slot + size_t(s.nz) * view)] = value;
  }
}

// Rounds accumulated on the device (batch sessions). Per scene, in doubles:
//   path_sum (nx, ny, nz, views) | weight_sum (nx, ny, views) |
//   path_mean, path_m2, covariance (nx, ny, nz - 1, views) | weight_mean, weight_m2 (nx, ny, views)
// The sums are those of the rounds' outputs; the moments are the online (Welford) moments of
// the AMF estimator over rounds, physical layers only, each round's detector weight paired
// with every layer of its pixel and view. Both kernels follow the host update step by step
// without contractions, so the device moments equal the host's to the bit. Path elements run
// first: they read the weight mean of the previous round.
struct MomentLayout {
  size_t path_sum, weight_sum, path_mean, path_m2, covariance, weight_mean, weight_m2, total;
};

__host__ __device__ inline MomentLayout moment_layout(const Scene &s) {
  const size_t plane = size_t(s.nx) * s.ny;
  const size_t paths = plane * s.nz * s.views, physical = plane * (s.nz - 1) * s.views;
  const size_t weights = plane * s.views;
  MomentLayout m;
  m.path_sum = 0;
  m.weight_sum = m.path_sum + paths;
  m.path_mean = m.weight_sum + weights;
  m.path_m2 = m.path_mean + physical;
  m.covariance = m.path_m2 + physical;
  m.weight_mean = m.covariance + physical;
  m.weight_m2 = m.weight_mean + weights;
  m.total = m.weight_m2 + weights;
  return m;
}

__global__ void accumulate_paths(Scene s, const Controls *__restrict__ controls, const unsigned *__restrict__ counts,
                                 const double *__restrict__ round, double *__restrict__ acc,
                                 unsigned *__restrict__ failure) {
#pragma clang fp contract(off)
  if (controls->count == 0) return;  // scene not traced this round
  const size_t plane = size_t(s.nx) * s.ny;
  const size_t index = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (index >= plane * s.nz * s.views) return;
  const MomentLayout m = moment_layout(s);
  const double value = round[s.stokes_count + index];
  const size_t pixel = index % plane;
  const int k = int((index / plane) % s.nz), view = int(index / (plane * s.nz));
  const double weight = round[s.stokes_count + s.path_count + pixel + plane * view];
  if (!isfinite(value) || !isfinite(weight)) atomicOr(failure, kNonfiniteTally);
  acc[m.path_sum + index] += value;
  if (k == s.nz - 1) return;  // the vacuum layer enters the sums only
  const double n = double(*counts + 1u);
  const size_t e = pixel + plane * (size_t(k) + size_t(s.nz - 1) * view);
  const double weight_mean = acc[m.weight_mean + pixel + plane * view];
  const double path_delta = value - acc[m.path_mean + e];
  const double weight_delta = weight - weight_mean;
  const double path_mean = acc[m.path_mean + e] + path_delta / n;
  const double new_weight_mean = weight_mean + weight_delta / n;
  acc[m.path_mean + e] = path_mean;
  acc[m.path_m2 + e] = acc[m.path_m2 + e] + path_delta * (value - path_mean);
  acc[m.covariance + e] = acc[m.covariance + e] + path_delta * (weight - new_weight_mean);
}

__global__ void accumulate_weights(Scene s, const Controls *__restrict__ controls, const unsigned *__restrict__ counts,
                                   const double *__restrict__ round, double *__restrict__ acc) {
#pragma clang fp contract(off)
  if (controls->count == 0) return;
  const size_t weights = size_t(s.nx) * s.ny * s.views;
  const size_t index = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (index >= weights) return;
  const MomentLayout m = moment_layout(s);
  const double weight = round[s.stokes_count + s.path_count + index];
  const double n = double(*counts + 1u);
  acc[m.weight_sum + index] += weight;
  const double delta = weight - acc[m.weight_mean + index];
  const double mean = acc[m.weight_mean + index] + delta / n;
  acc[m.weight_mean + index] = mean;
  acc[m.weight_m2 + index] = acc[m.weight_m2 + index] + delta * (weight - mean);
}

// Rounds accumulated per scene: one more for every scene traced in this round.
__global__ void count_rounds(const Controls *__restrict__ controls, unsigned *__restrict__ counts, uint32_t scenes) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < scenes && controls[i].count) counts[i] += 1u;
}

// ---------------------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------------------
// One transport session: the scenes of a batch (one for the classic entry points), their
// tables, tally banks, reduced and assembled tallies, and the captured graph of a batch round.
struct Session {
  int device = -1;
  char name[256] = {0};
  std::vector<Scene> scenes;  // device-pointer views of each scene's tables
  std::vector<size_t> reduced_base, assembled_base;  // per scene, in doubles
  Scene *d_scenes = nullptr;
  Controls *d_controls = nullptr;
  unsigned *work = nullptr;  // per-scene photon counters
  float *tables = nullptr;
  uint16_t *buckets = nullptr;  // CDF bucket tables of every tabulated species
  void *banks = nullptr;  // float or double banks, sized for double
  size_t bank_elements = 0;  // all scenes, all round slots
  double *reduced = nullptr;
  double *assembled = nullptr;
  size_t assembled_total = 0;  // doubles of all scenes for the round capacity
  unsigned *failure = nullptr;  // failure flags
  unsigned long long *counters = nullptr;
  int compute_units = 0;
  int row = 1;  // LDS row length: largest nz | 1
  int block = kBlock32;
  float *records = nullptr;
  size_t record_capacity = 0;
  uint32_t round_capacity = 0;
  hipStream_t stream = nullptr;
  hipEvent_t start = nullptr, stop = nullptr;
  double gpu_seconds = 0.0;
  double kernel_seconds = 0.0;
  void *proposal = nullptr;  // attached learned proposal (guided runs)
  Guidance guidance = unsteered();  // steered runs
  int steer = kPlain;  // batch rounds: kPlain, kDirected or kGuided
  uint32_t *splits = nullptr;  // weight-window copies of the steered modes, per lane
  size_t split_words = 0;
  // kWide: LDS layer rows of `wide` values (0: every plain launch keeps whole rows in LDS) and
  // each lane's row of upper-layer paths in global memory (`upper_stride` floats).
  int wide = 0;
  int upper_stride = 0;
  float *upper = nullptr;
  uint32_t upper_blocks = 0;  // blocks whose lanes have a global row
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  // Rounds accumulated on the device (batch sessions with layer paths): per-scene sums and
  // moments (allocated by the first accumulating round), rounds per scene, and that round's
  // captured graph.
  double *moments = nullptr;
  std::vector<size_t> moments_base;
  unsigned *moment_counts = nullptr;
  hipGraph_t accumulate_graph = nullptr;
  hipGraphExec_t accumulate_exec = nullptr;
};

void set_error(char *error, size_t size, const char *fmt, const char *detail = "") {
  if (error && size) snprintf(error, size, fmt, detail);
}

bool check(hipError_t status, char *error, size_t size, const char *what) {
  if (status == hipSuccess) return true;
  if (error && size) snprintf(error, size, "%s: %s", what, hipGetErrorString(status));
  return false;
}

void release(Session *s) {
  if (!s) return;
  if (s->device >= 0) (void)hipSetDevice(s->device);
  if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
  if (s->upper) (void)hipFree(s->upper);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  if (s->accumulate_exec) (void)hipGraphExecDestroy(s->accumulate_exec);
  if (s->accumulate_graph) (void)hipGraphDestroy(s->accumulate_graph);
  if (s->start) (void)hipEventDestroy(s->start);
  if (s->stop) (void)hipEventDestroy(s->stop);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  // s->work points into the failure allocation.
  void *buffers[] = {s->tables, s->buckets, s->banks, s->reduced, s->assembled, s->failure, s->counters, s->records,
                     s->d_scenes, s->d_controls, s->moments, s->moment_counts, s->splits};
  for (void *b : buffers) (void)hipFree(b);
  delete s;
}

// Zero `words` 32-bit words (a kernel, not a memset node, so that it can be captured).
__global__ void fill_zero(uint32_t *__restrict__ p, size_t words) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < words; i += stride) p[i] = 0u;
}

void launch_fill(uint32_t *p, size_t words, hipStream_t stream) {
  const size_t blocks = (words + 255) / 256;
  hipLaunchKernelGGL(fill_zero, dim3(unsigned(blocks < 4096 ? (blocks ? blocks : 1) : 4096)), dim3(256), 0, stream,
                     p, words);
}

bool word_int(const float *t, int index, int lo, int hi, int &out) {
  const float v = t[index];
  if (!(v == floorf(v)) || v < float(lo) || v > float(hi)) return false;
  out = int(v);
  return true;
}

// Parse and validate the header; pointers are offsets into `base` (host or device).
bool parse_tables(const float *t, size_t count, const float *base, Scene &s, char *error, size_t size) {
  if (count < size_t(kHeaderWords) || t[0] != float(kLayoutVersion)) {
    set_error(error, size, "Unsupported packed table layout");
    return false;
  }
  for (size_t i = 0; i < count; ++i) {
    if (!isfinite(t[i])) {
      set_error(error, size, "Nonfinite packed transport tables");
      return false;
    }
  }
  int length = 0;
  if (!word_int(t, 1, 1, 1 << 16, s.nx) || !word_int(t, 2, 1, 1 << 16, s.ny) ||
      !word_int(t, 3, 2, kMaxLayers, s.nz) || !word_int(t, 4, 1, 4, s.nmat) ||
      !word_int(t, 5, 1, kMaxViews, s.views) || !word_int(t, 6, 0, kMaxSpecies, s.species) ||
      !word_int(t, 7, 0, 1, s.paths) || !word_int(t, 8, 0, kMaxLayers, s.uniform_top) ||
      !word_int(t, 9, 1, 64, s.shards) || !word_int(t, 26, kHeaderWords, 1 << 24, length) ||
      !word_int(t, 28, 0, 1, s.backward) || !word_int(t, 29, 0, 1, s.scats) ||
      !word_int(t, 30, 0, 1, s.detector_surface) || !word_int(t, 31, 0, 1, s.register_opposite)) {
    set_error(error, size, "Invalid packed table dimensions");
    return false;
  }
  if (s.nmat == 2 || size_t(length) != count) {
    set_error(error, size, "Packed table size or Stokes size mismatch");
    return false;
  }
  // One tally record (layer paths, then Stokes) spans at most one wave of 64 lanes.
  if ((s.paths ? s.nz + s.nmat : s.nmat) > 64) {
    set_error(error, size, "Layers plus Stokes components exceed 64 per tally record");
    return false;
  }
  const long cells = long(s.nx) * s.ny * s.nz;
  s.xunit = t[10];
  s.yunit = t[11];
  s.depol = t[12];
  s.theta0 = t[13];
  s.phi0 = t[14];
  s.phi = t[15];
  for (int v = 0; v < kMaxViews; ++v) s.view[v] = v < s.views ? t[16 + v] : 0.0f;
  const double t0 = s.theta0, p0 = s.phi0, ph = s.phi;
  s.launch_a = sin(t0) * cos(p0);
  s.launch_b = sin(t0) * sin(p0);
  s.launch_c = cos(t0);
  s.mu0 = fabs(cos(t0));
  for (int v = 0; v < kMaxViews; ++v) {
    const double theta = s.view[v], mu = cos(theta);
    s.view_mu[v] = mu;
    s.view_sine[v] = sqrt(fmax(1e-16, 1.0 - mu * mu));
    s.view_a[v] = sin(theta) * cos(ph);
    s.view_b[v] = sin(theta) * sin(ph);
    // mu0 / mu converts source to detector area in forward transport only (mc.f90:1456).
    // |mu|: the views of a surface detector point down (mc.f90:1456-1474 takes DABS).
    s.view_geometry[v] = v < s.views ? (s.backward ? 1.0 : s.mu0 / fabs(mu)) : 0.0;
    s.view_inv_mu[v] = v < s.views ? 1.0 / fabs(mu) : 0.0;
  }
  struct Span {
    int word;
    long length;
    const float **target;
  } spans[] = {{19, s.nz, &s.zunit},        {20, long(s.nx) * s.ny, &s.albedo},
               {21, cells, &s.beta},         {22, cells * s.species, &s.cumfrac},
               {23, s.nz, &s.ssalb_gas},     {24, s.nz, &s.height_above},
               {25, s.nz, &s.tau_above},     {27, s.nz, &s.uniform}};
  for (const Span &span : spans) {
    int offset = 0;
    if (!word_int(t, span.word, kHeaderWords, length, offset) || offset + span.length > length) {
      set_error(error, size, "Packed table offset out of range");
      return false;
    }
    *span.target = base + offset;
  }
  // Every layer from uniform_top up is horizontally uniform (the detector walk's prefix and the
  // split layer rows of tall scenes rely on it).
  {
    const float *flags = t + (s.uniform - base);
    for (int k = s.uniform_top; k < s.nz; ++k) {
      if (flags[k] == 0.0f) {
        set_error(error, size, "A layer above the uniform top is not uniform");
        return false;
      }
    }
  }
  // Factored extinction (words 64-65): one tabulated species, both parts in range, and on the
  // host pass every cell reproduces the extinction and fraction tables bit for bit.
  s.species_beta = s.gas_beta = nullptr;
  if (t[64] != 0.0f || t[65] != 0.0f) {
    int species_at = 0, gas_at = 0;
    if (s.species != 1 || !word_int(t, 64, kHeaderWords, length, species_at) || species_at + cells > length ||
        !word_int(t, 65, kHeaderWords, length, gas_at) || gas_at + s.nz > length) {
      set_error(error, size, "Invalid factored extinction");
      return false;
    }
    if (!base) {
      const float *beta = t + (s.beta - base), *fraction = t + (s.cumfrac - base);
      const float *part = t + species_at, *gas = t + gas_at;
      for (long cell = 0; cell < cells; ++cell) {
        const float total = part[cell] + gas[cell / (long(s.nx) * s.ny)];
        if (total != beta[cell] || (total > 0.0f && part[cell] / total != fraction[cell])) {
          set_error(error, size, "Factored extinction does not reproduce the tables");
          return false;
        }
      }
    }
    s.species_beta = base + species_at;
    s.gas_beta = base + gas_at;
  }
  for (int sp = 0; sp < kMaxSpecies; ++sp) {
    SpeciesTables &tab = s.sp[sp];
    tab = SpeciesTables{0,       0.0f,    1.0f,    1.0f,    0,       nullptr, nullptr,
                        nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    if (sp >= s.species) continue;
    const int h = 32 + 8 * sp;
    if (!word_int(t, h, 2, 1 << 20, tab.lut_size)) {
      set_error(error, size, "Invalid scattering LUT size");
      return false;
    }
    tab.ssalb = t[h + 1];
    const float **targets[6] = {&tab.cosine, &tab.cdf, &tab.f11, &tab.f22, &tab.f33, &tab.f12};
    for (int a = 0; a < 6; ++a) {
      int offset = 0;
      const int extent = a == 0 ? 3 * tab.lut_size : tab.lut_size;
      if (!word_int(t, h + 2 + a, kHeaderWords, length, offset) || offset + extent > length) {
        set_error(error, size, "Scattering table offset out of range");
        return false;
      }
      *targets[a] = base + offset;
    }
    if (s.nmat == 4) {
      // F34 and F44 follow F12 directly.
      if (size_t(tab.f12 - base) + 3 * size_t(tab.lut_size) > size_t(length)) {
        set_error(error, size, "Scattering table offset out of range");
        return false;
      }
      tab.f34 = tab.f12 + tab.lut_size;
      tab.f44 = tab.f12 + 2 * tab.lut_size;
    }
    // Ascending cosines spanning [-1, 1] and a CDF from 0 to 1 (metal_tables.py:421-425).
    const float *cs = t + (tab.cosine - base), *cd = t + (tab.cdf - base);
    if (fabsf(cs[0] + 1.0f) > 1e-6f || fabsf(cs[tab.lut_size - 1] - 1.0f) > 1e-6f ||
        fabsf(cd[0]) > 1e-6f || fabsf(cd[tab.lut_size - 1] - 1.0f) > 1e-6f) {
      set_error(error, size, "Invalid scattering angle grid or CDF");
      return false;
    }
    for (int a = 1; a < tab.lut_size; ++a) {
      if (!(cs[a] > cs[a - 1]) || cd[a] < cd[a - 1]) {
        set_error(error, size, "Invalid scattering angle grid or CDF");
        return false;
      }
    }
    const float *f11 = t + (tab.f11 - base), *f12 = t + (tab.f12 - base);
    double area = 0.0, rho = 1.0;
    for (int a = 1; a < tab.lut_size; ++a) {
      area += 0.5 * (double(cs[a]) - cs[a - 1]) * (double(f11[a]) + f11[a - 1]);
      const double b0 = double(f11[a - 1]) + (s.nmat > 1 ? fabs(double(f12[a - 1])) : 0.0);
      const double b1 = double(f11[a]) + (s.nmat > 1 ? fabs(double(f12[a])) : 0.0);
      if (b0 + b1 > 0.0) rho = fmax(rho, fmax(b0, b1) / (0.5 * (b0 + b1)));
    }
    if (!(area > 0.0)) {
      set_error(error, size, "Phase function integrates to zero");
      return false;
    }
    tab.norm = float(area / 2.0);
    tab.equal_angle = 1;
    for (int a = 0; a < tab.lut_size && tab.equal_angle; ++a) {
      const double expected = cos(3.14159265358979323846 * (1.0 - double(a) / double(tab.lut_size - 1)));
      if (fabs(double(cs[a]) - expected) > 2e-6) tab.equal_angle = 0;
    }
    tab.rho = float(rho * (1.0 + 1e-6));
  }
  for (int v = 0; v < s.views; ++v) {
    // TOA views point up, the views of a surface detector down; none along the horizon.
    if (!(s.detector_surface ? cosf(s.view[v]) < 0.0f : cosf(s.view[v]) > 0.0f)) {
      set_error(error, size, "Viewing angles must stay off the horizon, on the detector's side");
      return false;
    }
  }
  if (s.scats && s.paths) {
    set_error(error, size, "The scattering decomposition runs without layer paths");
    return false;
  }
  s.stokes_count = s.nmat * s.nx * s.ny * s.views;
  s.path_count = s.nx * s.ny * s.nz * s.views;
  s.tally_count = s.stokes_count + (s.paths ? s.path_count + s.nx * s.ny * s.views : 0);
  s.record_stride = ((s.paths ? s.nz + s.nmat : s.nmat) + 7) & ~7;
  size_t records = size_t(s.nx) * s.ny * s.views;
  if (s.scats) {
    // Stokes, single, multiple; single and multiple contributions per event cell.
    s.tally_count = 3 * s.stokes_count + 2 * s.nmat * s.path_count;
    s.record_stride = 8;
    records *= size_t(1 + s.nz);
  }
  // The tally addresses records by 32-bit byte offsets within one shard bank (float64 build).
  if (records * s.record_stride * sizeof(double) >= (size_t(1) << 31)) {
    set_error(error, size, "Tally bank of one shard exceeds 2 GiB");
    return false;
  }
  s.bank_count = int(records * s.record_stride);
  return true;
}

// Everything a group of rounds runs on the device, in stream order: clear the banks, the
// failure flags and the photon counters; transport; per scene, the float64 shard reduction
// and the assembly in output order. No memset nodes, so that a batch round can be captured.
template <typename R, bool TRACE, int MODE>
bool enqueue(Session *h, uint32_t rounds, uint32_t grid, int block, size_t lds, size_t record_count,
             monki_net::Net net, monki_net::EventLog log, char *error, size_t error_size) {
  const uint32_t scene_count = uint32_t(h->scenes.size());
  const size_t words = sizeof(R) / sizeof(uint32_t);
  if (scene_count == 1) {
    const Scene &s = h->scenes[0];
    launch_fill(static_cast<uint32_t *>(h->banks), size_t(s.bank_count) * s.shards * rounds * words, h->stream);
  } else {
    launch_fill(static_cast<uint32_t *>(h->banks), h->bank_elements * words, h->stream);
  }
  launch_fill(h->failure, 1 + scene_count, h->stream);
  R *banks = static_cast<R *>(h->banks);
  // A kWide grid never exceeds the blocks whose lanes have a row in global memory (persistent
  // blocks loop over the photons, so fewer blocks only take longer).
  if (MODE == kWide && grid > h->upper_blocks) grid = h->upper_blocks;
  if (grid) {
    hipLaunchKernelGGL((transport_kernel<R, TRACE, MODE>), dim3(grid), dim3(block), lds, h->stream,
                       h->d_scenes, h->d_controls, scene_count, h->row, banks, h->failure, h->work,
                       h->counters, TRACE ? h->records : nullptr, uint32_t(TRACE ? record_count : 0), net, log,
                       MODE == kGuided || MODE == kDirected ? h->guidance : unsteered(),
                       MODE == kWide ? reinterpret_cast<uint32_t *>(h->upper) : h->splits,
                       MODE == kWide ? h->upper_stride : kSplitWords + h->row, MODE == kWide ? h->wide : 0);
  }
  if (!check(hipGetLastError(), error, error_size, "transport launch")) return false;
  for (uint32_t i = 0; i < scene_count; ++i) {
    const Scene &s = h->scenes[i];
    const size_t summed = size_t(s.bank_count) * rounds;
    hipLaunchKernelGGL(reduce_shards<R>, dim3(unsigned((summed + 255) / 256)), dim3(256), 0, h->stream,
                       banks + s.bank_base, h->reduced + h->reduced_base[i], s.bank_count, s.shards, int(rounds));
    hipLaunchKernelGGL(assemble_tallies, dim3(unsigned((summed + 255) / 256)), dim3(256), 0, h->stream, s,
                       h->reduced + h->reduced_base[i], h->assembled + h->assembled_base[i], int(rounds));
  }
  return check(hipGetLastError(), error, error_size, "shard reduction launch");
}

// Persistent grid: enough resident blocks to fill every CU, never more than the work.
template <typename R, bool TRACE, int MODE>
uint32_t persistent_grid(const Session *h, uint64_t photons, int block, size_t lds) {
  int resident = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, transport_kernel<R, TRACE, MODE>, block,
                                                  lds) != hipSuccess || resident < 1)
    resident = 1;
  const uint64_t wanted = (photons + block - 1) / block;
  const uint64_t filled = uint64_t(resident) * uint64_t(h->compute_units > 0 ? h->compute_units : 1);
  return uint32_t(wanted < filled ? wanted : filled);
}

// Lanes per block: the session's choice, at most 128 in float64 (register file).
template <typename R>
int block_size(const Session *h) {
  return sizeof(R) == 8 && h->block > kBlock64 ? kBlock64 : h->block;
}

template <typename R>
size_t lds_bytes_for(int row, int block, bool v4) {
  // History rows and pending local estimates or open records per lane; per-wave layer tables
  // (4 rows of floats per wave) except with the full Stokes vector.
  return (size_t(row) * sizeof(R) + pending_bytes<R>(v4)) * block +
         (v4 ? 0 : size_t(block / 64) * 4 * row * sizeof(float));
}

template <typename R>
size_t lds_bytes(const Session *h, int block, bool v4 = false) {
  return lds_bytes_for<R>(h->row, block, v4);
}

// kWide: history rows of `wide` values, the per-wave layer tables still `row` long.
size_t lds_bytes_wide(int wide, int row, int block) {
  // Records of two layers: the five-field record area.
  return (size_t(wide) * sizeof(float) + pending_bytes<float>(true)) * block + size_t(block / 64) * 4 * row * sizeof(float);
}

// LDS of a launch in mode MODE.
template <typename R, int MODE>
size_t launch_lds(const Session *h, int block) {
  return MODE == kWide ? lds_bytes_wide(h->wide, h->row, block) : lds_bytes<R>(h, block, full_stokes(MODE));
}

// Blocks of `block` lanes a CU holds for the plain float32 tracer with `lds` bytes of LDS per
// block (the occupancy calculator: registers, LDS and its allocation granule).
int resident_blocks(bool wide, int block, size_t lds) {
  int blocks = 0;
  const hipError_t status =
      wide ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, transport_kernel<float, false, kWide>, block, lds)
           : hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, transport_kernel<float, false, kPlain>, block, lds);
  return status == hipSuccess ? blocks : 0;
}

// Whether `blocks` blocks with `lds` bytes each fit a CU's LDS with 2 KB of it to spare. Measured
// on gfx950 (160 KB per CU): six 128-lane blocks of 26,016, 26,528 or 26,720 B and three 256-lane
// blocks of 53,056 B run together; six of 27,040 B and three of 54,080 B, which the occupancy
// calculator also places together, ran a fifth and a half slower.
bool lds_fits(size_t lds, int blocks, size_t lds_per_cu) {
  return size_t(blocks) * lds + 2048 <= lds_per_cu;
}

// Plain float32 launches of scenes whose whole layer rows leave fewer than three waves per SIMD
// (the budget of the kernel's registers) run as kWide when that restores three waves: the
// longest odd LDS row with which both the occupancy calculator and the measured LDS limit give
// three waves per SIMD, at least the highest varying layer of every scene (so that the cell walks
// stay in LDS). Forward TOA scenes of nmat 1 or 3 only. Returns the row, or 0.
int wide_row(const Session *h, int block, size_t lds_per_cu) {
  int uniform_top = 0;
  for (const Scene &s : h->scenes) {
    if (s.nmat == 4 || s.backward || s.scats || s.detector_surface) return 0;
    uniform_top = s.uniform_top > uniform_top ? s.uniform_top : uniform_top;
  }
  const int needed = (12 * 64 + block - 1) / block;  // blocks per CU for three waves per SIMD
  const size_t whole = lds_bytes_for<float>(h->row, block, false);
  if (resident_blocks(false, block, whole) >= needed && lds_fits(whole, needed, lds_per_cu)) return 0;
  for (int wide = h->row - 2; wide >= uniform_top && wide >= 1; wide -= 2) {
    const size_t lds = lds_bytes_wide(wide, h->row, block);
    if (resident_blocks(true, block, lds) >= needed && lds_fits(lds, needed, lds_per_cu)) return wide;
  }
  return 0;
}

bool finish(Session *h, char *error, size_t error_size) {
  unsigned flag = 0;
  if (!check(hipMemcpyAsync(&flag, h->failure, sizeof(unsigned), hipMemcpyDeviceToHost, h->stream), error,
             error_size, "copy failure flag") ||
      !check(hipStreamSynchronize(h->stream), error, error_size, "transport"))
    return false;
  if (flag) {
    if (error && error_size)
      snprintf(error, error_size, "HIP transport failed (flag %u); round rejected", flag);
    return false;
  }
  return true;
}

// Room for the weight-window stacks of every lane of a steered launch (kept for later launches).
bool reserve_splits(Session *h, uint32_t grid, int block, char *error, size_t error_size) {
  const size_t words = size_t(grid) * block * kSplitDepth * (kSplitWords + h->row);
  if (words <= h->split_words) return true;
  (void)hipFree(h->splits);
  h->splits = nullptr;
  h->split_words = 0;
  if (!check(hipMalloc(&h->splits, words * sizeof(uint32_t)), error, error_size, "allocate weight-window stacks"))
    return false;
  h->split_words = words;
  return true;
}

// One group of rounds of the session's single scene (the classic entry points).
template <typename R, bool TRACE, int MODE = kPlain>
int launch(Session *h, const uint32_t *controls, double *out, float *records, size_t record_count,
           char *error, size_t error_size, void *proposal = nullptr) {
  if (!h || !controls || !out || h->scenes.size() != 1) {
    set_error(error, error_size, "Invalid HIP session or arguments");
    return 1;
  }
  if constexpr (MODE != kStokes4 && !backward(MODE) && !decomposition(MODE) && !downward(MODE)) {
    // The full Stokes vector and backward transport have instantiations of their own of the
    // plain tracer.
    const Scene &first = h->scenes[0];
    if (first.nmat == 4 || first.backward || first.scats || first.detector_surface) {
      if (MODE != kPlain) {
        set_error(error, error_size, "The steered and logging tracers take forward nmat 1 or 3 TOA scenes");
        return 1;
      }
      if (first.detector_surface && (first.backward || first.scats)) {
        set_error(error, error_size, "A ground detector runs forward without the scattering decomposition");
        return 1;
      }
      if (first.detector_surface && first.nmat == 4)
        return launch<R, TRACE, kSurface4>(h, controls, out, records, record_count, error, error_size, proposal);
      if (first.detector_surface)
        return launch<R, TRACE, kSurface>(h, controls, out, records, record_count, error, error_size, proposal);
      if (first.scats && first.backward)
        return first.nmat == 4
                   ? launch<R, TRACE, kBackScats4>(h, controls, out, records, record_count, error, error_size, proposal)
                   : launch<R, TRACE, kBackScats>(h, controls, out, records, record_count, error, error_size, proposal);
      if (first.scats && first.nmat == 4)
        return launch<R, TRACE, kScats4>(h, controls, out, records, record_count, error, error_size, proposal);
      if (first.scats)
        return launch<R, TRACE, kScats>(h, controls, out, records, record_count, error, error_size, proposal);
      if (first.backward && first.nmat == 4)
        return launch<R, TRACE, kBackward4>(h, controls, out, records, record_count, error, error_size, proposal);
      if (first.backward)
        return launch<R, TRACE, kBackward>(h, controls, out, records, record_count, error, error_size, proposal);
      return launch<R, TRACE, kStokes4>(h, controls, out, records, record_count, error, error_size, proposal);
    }
  }
  if constexpr (MODE == kPlain && sizeof(R) == 4) {
    if (h->wide) return launch<R, TRACE, kWide>(h, controls, out, records, record_count, error, error_size, proposal);
  }
  Controls c{controls[0], controls[1], controls[2], controls[3],
             controls[4], controls[5], controls[6], controls[7]};
  const Scene &s = h->scenes[0];
  if (!c.per_round || !c.rounds || c.rounds > h->round_capacity || !c.max_steps ||
      c.count > (1u << 31) ||
      uint64_t(c.first) + c.count > uint64_t(c.per_round) * c.rounds) {
    set_error(error, error_size, "Invalid HIP round group");
    return 1;
  }
  if (!check(hipSetDevice(h->device), error, error_size, "hipSetDevice") ||
      !check(hipMemcpy(h->d_controls, &c, sizeof(Controls), hipMemcpyHostToDevice), error, error_size,
             "upload controls"))
    return 1;
  if (TRACE && record_count > h->record_capacity) {
    (void)hipFree(h->records);
    h->records = nullptr;
    h->record_capacity = 0;
    if (!check(hipMalloc(&h->records, record_count * kTraceWords * sizeof(float)), error, error_size,
               "allocate trace records"))
      return 1;
    h->record_capacity = record_count;
  }
  monki_net::Net net{nullptr, nullptr, nullptr};
  monki_net::EventLog log{};
  uint32_t logged_before = 0;
  if (MODE == kLog || MODE == kGuided) {
    if (!proposal || monki_net::proposal_device(proposal) != h->device) {
      set_error(error, error_size, "A proposal on the session's device is required");
      return 1;
    }
    net = monki_net::proposal_net(proposal);
    if (MODE == kLog) {
      log = monki_net::proposal_log(proposal);
      if (!log.features || c.count > log.photons) {
        set_error(error, error_size, "Reserve proposal samples for this dispatch first");
        return 1;
      }
      if (log.layer_weight && log.layer_count < uint32_t(s.nz)) {
        set_error(error, error_size, "The path target needs a weight for every layer of the scene");
        return 1;
      }
      logged_before = monki_net::proposal_samples(proposal);
    }
  }
  const int block = block_size<R>(h);
  const size_t lds = launch_lds<R, MODE>(h, block);
  const uint32_t grid = persistent_grid<R, TRACE, MODE>(h, c.count, block, lds);
  if ((MODE == kGuided || MODE == kDirected) && !reserve_splits(h, grid, block, error, error_size)) return 1;
  if (MODE == kLog) {
    // Counting pass: the same histories (their streams are fixed by photon id), each photon's
    // number of logged events into starts[id]; then starts become the photons' first slots.
    log.counting = 1u;
    if (!enqueue<R, TRACE, MODE>(h, c.rounds, grid, block, lds, record_count, net, log, error, error_size) ||
        !finish(h, error, error_size))
      return 1;
    std::vector<uint32_t> starts(c.count);
    if (!check(hipMemcpy(starts.data(), log.starts, c.count * sizeof(uint32_t), hipMemcpyDeviceToHost), error,
               error_size, "copy logged counts"))
      return 1;
    uint64_t next = logged_before;
    for (uint32_t i = 0; i < c.count; ++i) {
      const uint64_t events = starts[i];
      starts[i] = uint32_t(next < log.capacity ? next : log.capacity);
      next += events;
    }
    const unsigned logged = unsigned(next < log.capacity ? next : log.capacity);
    if (!check(hipMemcpy(log.starts, starts.data(), c.count * sizeof(uint32_t), hipMemcpyHostToDevice), error,
               error_size, "upload log starts") ||
        !check(hipMemcpy(log.count, &logged, sizeof(unsigned), hipMemcpyHostToDevice), error, error_size,
               "set log count"))
      return 1;
    log.counting = 0u;
  }
  (void)hipEventRecord(h->start, h->stream);
  if (!enqueue<R, TRACE, MODE>(h, c.rounds, grid, block, lds, record_count, net, log, error, error_size)) return 1;
  (void)hipEventRecord(h->stop, h->stream);
  if (!finish(h, error, error_size)) return 1;
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, h->start, h->stop);
  h->kernel_seconds = ms * 1e-3;
  if (c.profile) h->gpu_seconds += ms * 1e-3;
  if (!check(hipMemcpy(out, h->assembled, size_t(s.tally_count) * c.rounds * sizeof(double),
                       hipMemcpyDeviceToHost),
             error, error_size, "copy tallies"))
    return 1;
  if (MODE == kLog && !monki_net::proposal_finish_log(proposal, logged_before, h->stream, error, error_size))
    return 1;
  if (TRACE && records && record_count) {
    const size_t n = record_count < c.count ? record_count : c.count;
    if (!check(hipMemcpy(records, h->records, n * kTraceWords * sizeof(float), hipMemcpyDeviceToHost),
               error, error_size, "copy trace records"))
      return 1;
  }
  return 0;
}

// Parse, upload and lay out the scenes of a session. tables holds the packed buffers of all
// scenes back to back; offsets[i] .. offsets[i + 1] is scene i (float32 words).
Session *create_session(uint32_t device, const float *tables, const size_t *offsets, uint32_t scene_count,
                        uint32_t round_capacity, uint32_t block, char *error, size_t error_size) {
  if (!tables || !offsets || !scene_count || scene_count > 4096 || !round_capacity ||
      (block != 64 && block != 128 && block != 256)) {
    set_error(error, error_size, "Invalid HIP create arguments");
    return nullptr;
  }
  int devices = 0;
  if (hipGetDeviceCount(&devices) != hipSuccess || device >= uint32_t(devices)) {
    set_error(error, error_size, "Requested HIP device is unavailable");
    return nullptr;
  }
  std::vector<Scene> parsed(scene_count);
  for (uint32_t i = 0; i < scene_count; ++i) {
    if (offsets[i + 1] <= offsets[i] ||
        !parse_tables(tables + offsets[i], offsets[i + 1] - offsets[i], nullptr, parsed[i], error, error_size))
      return nullptr;
  }
  Session *h = new Session();
  h->device = int(device);
  h->round_capacity = round_capacity;
  h->block = int(block);
  size_t banks = 0, reduced = 0, assembled = 0;
  int nz = 1;
  for (uint32_t i = 0; i < scene_count; ++i) {
    parsed[i].bank_base = banks;
    h->reduced_base.push_back(reduced);
    h->assembled_base.push_back(assembled);
    banks += size_t(parsed[i].bank_count) * parsed[i].shards * round_capacity;
    reduced += size_t(parsed[i].bank_count) * round_capacity;
    assembled += size_t(parsed[i].tally_count) * round_capacity;
    nz = parsed[i].nz > nz ? parsed[i].nz : nz;
  }
  h->bank_elements = banks;
  h->assembled_total = assembled;
  h->row = nz | 1;
  const size_t table_count = offsets[scene_count];
  hipDeviceProp_t prop{};
  if (!check(hipSetDevice(h->device), error, error_size, "hipSetDevice") ||
      !check(hipGetDeviceProperties(&prop, h->device), error, error_size, "device properties") ||
      !check(hipStreamCreate(&h->stream), error, error_size, "stream") ||
      !check(hipEventCreate(&h->start), error, error_size, "event") ||
      !check(hipEventCreate(&h->stop), error, error_size, "event") ||
      !check(hipMalloc(&h->tables, table_count * sizeof(float)), error, error_size, "allocate tables") ||
      !check(hipMalloc(&h->banks, banks * sizeof(double)), error, error_size, "allocate tally banks") ||
      !check(hipMalloc(&h->reduced, reduced * sizeof(double)), error, error_size, "allocate reduced tallies") ||
      !check(hipMalloc(&h->assembled, assembled * sizeof(double)), error, error_size, "allocate output tallies") ||
      !check(hipMalloc(&h->failure, (1 + scene_count) * sizeof(unsigned)), error, error_size, "allocate flags") ||
      !check(hipMalloc(&h->counters, kCounters * sizeof(unsigned long long)), error, error_size,
             "allocate counters") ||
      !check(hipMalloc(&h->d_scenes, scene_count * sizeof(Scene)), error, error_size, "allocate scenes") ||
      !check(hipMalloc(&h->d_controls, scene_count * sizeof(Controls)), error, error_size, "allocate controls") ||
      !check(hipMemcpy(h->tables, tables, table_count * sizeof(float), hipMemcpyHostToDevice), error,
             error_size, "upload tables") ||
      !check(hipMemset(h->counters, 0, kCounters * sizeof(unsigned long long)), error, error_size,
             "clear counters")) {
    release(h);
    return nullptr;
  }
  h->work = h->failure + 1;
  // CDF bucket tables (host, from the packed CDFs), one block per tabulated species.
  std::vector<uint16_t> buckets;
  std::vector<size_t> bucket_base;
  for (uint32_t i = 0; i < scene_count; ++i) {
    const float *t = tables + offsets[i];
    for (int sp = 0; sp < parsed[i].species; ++sp) {
      const SpeciesTables &tab = parsed[i].sp[sp];
      const float *cdf = t + (tab.cdf - static_cast<const float *>(nullptr));
      const int n = tab.lut_size;
      bucket_base.push_back(buckets.size());
      int low = 0;
      for (int b = 0; b <= kCdfBuckets; ++b) {
        const float edge = float(b) / float(kCdfBuckets);
        while (low < n - 2 && cdf[low + 1] <= edge) ++low;
        buckets.push_back(uint16_t(low));
      }
    }
  }
  if (!buckets.empty() &&
      (!check(hipMalloc(&h->buckets, buckets.size() * sizeof(uint16_t)), error, error_size, "allocate buckets") ||
       !check(hipMemcpy(h->buckets, buckets.data(), buckets.size() * sizeof(uint16_t), hipMemcpyHostToDevice),
              error, error_size, "upload buckets"))) {
    release(h);
    return nullptr;
  }
  size_t next_bucket = 0;
  for (uint32_t i = 0; i < scene_count; ++i) {
    Scene view{};
    parse_tables(tables + offsets[i], offsets[i + 1] - offsets[i], h->tables + offsets[i], view, nullptr, 0);
    view.bank_base = parsed[i].bank_base;
    for (int sp = 0; sp < view.species; ++sp) view.sp[sp].cdf_start = h->buckets + bucket_base[next_bucket++];
    // A species field equal to an earlier scene's (one cloud at several wavelengths): read that
    // scene's copy, so that the batch keeps one copy of the field in the caches.
    if (view.species_beta) {
      const size_t cells = size_t(view.nx) * view.ny * view.nz;
      const float *mine = tables + (view.species_beta - h->tables);
      for (uint32_t j = 0; j < i; ++j) {
        const Scene &other = h->scenes[j];
        if (!other.species_beta || other.nx != view.nx || other.ny != view.ny || other.nz != view.nz) continue;
        if (!memcmp(tables + (other.species_beta - h->tables), mine, cells * sizeof(float))) {
          view.species_beta = other.species_beta;
          break;
        }
      }
    }
    h->scenes.push_back(view);
  }
  if (!check(hipMemcpy(h->d_scenes, h->scenes.data(), scene_count * sizeof(Scene), hipMemcpyHostToDevice), error,
             error_size, "upload scenes")) {
    release(h);
    return nullptr;
  }
  // Some drivers report an empty marketing name: fall back to the architecture and CU count.
  if (prop.name[0])
    snprintf(h->name, sizeof(h->name), "%s (%s, %d CUs)", prop.name, prop.gcnArchName,
             prop.multiProcessorCount);
  else
    snprintf(h->name, sizeof(h->name), "%s, %d CUs", prop.gcnArchName, prop.multiProcessorCount);
  h->compute_units = prop.multiProcessorCount;
  // Scenes too tall for whole LDS rows at three waves per SIMD: the kWide split and the lanes'
  // global rows, enough for every block the LDS lets a CU hold.
  const size_t lds_per_cu = prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor : 65536;
  h->wide = wide_row(h, h->block, lds_per_cu);
  if (h->wide) {
    h->upper_stride = nz - h->wide;
    const size_t resident = size_t(resident_blocks(true, h->block, lds_bytes_wide(h->wide, h->row, h->block)));
    h->upper_blocks = uint32_t(size_t(h->compute_units > 0 ? h->compute_units : 1) * (resident ? resident : 1));
    const size_t floats = size_t(h->upper_blocks) * size_t(h->block) * size_t(h->upper_stride);
    if (!check(hipMalloc(&h->upper, floats * sizeof(float)), error, error_size, "allocate upper-layer rows")) {
      release(h);
      return nullptr;
    }
  }
  return h;
}

// Single-purpose probes of the device helpers, for tests against golden vectors.
// Case: uo, u1, phio, phi1, depol, I, Q, U -> cos_sca, C1, S1, C2, S2, F11, F22, F33, F12 (/4),
// and P*S (I, Q, U), evaluated in float32 exactly as transport does for a Rayleigh event.
__global__ void probe_scatter_kernel(const double *in, double *out, int cases) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= cases) return;
  const double *c = in + 8 * n;
  const float uo = float(c[0]), u1 = float(c[1]), phio = float(c[2]), phi1 = float(c[3]);
  const float cosine = clamp1(uo * u1 + sqrtf(fmaxf(0.0f, 1.0f - uo * uo)) *
                                            sqrtf(fmaxf(0.0f, 1.0f - u1 * u1)) * cosf(phio - phi1));
  Scene s{};
  s.depol = float(c[4]);
  s.nmat = 3;
  uint32_t reads = 0;
  const Elements<float> f = elements<float>(s, -1, cosine, 1.0f - cosine, 1.0f + cosine, reads);
  const Rotation<float> r = rotation<float>(cosine, uo, u1, phio, phi1);
  const Stokes<float> p = scatter<float>(f, r, Stokes<float>{float(c[5]), float(c[6]), float(c[7])});
  double *o = out + 12 * n;
  const float values[12] = {cosine, r.c1, r.s1, r.c2, r.s2, f.f11, f.f22, f.f33, f.f12, p.i, p.q, p.u};
  for (int i = 0; i < 12; ++i) o[i] = double(values[i]);
}

__global__ void probe_philox_kernel(const uint32_t *in, uint32_t *out, int cases) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= cases) return;
  const uint32_t *c = in + 6 * n;
  const Words w = philox(Words{c[0], c[1], c[2], c[3]}, c[4], c[5]);
  out[4 * n + 0] = w.x;
  out[4 * n + 1] = w.y;
  out[4 * n + 2] = w.z;
  out[4 * n + 3] = w.w;
}

template <typename T>
int run_probe(uint32_t device, void (*kernel)(const T *, T *, int), const T *in, size_t in_words,
              T *out, size_t out_words, int cases) {
  int devices = 0;
  if (hipGetDeviceCount(&devices) != hipSuccess || device >= uint32_t(devices)) return 1;
  T *din = nullptr, *dout = nullptr;
  int status = 1;
  if (hipSetDevice(int(device)) == hipSuccess && hipMalloc(&din, in_words * sizeof(T)) == hipSuccess &&
      hipMalloc(&dout, out_words * sizeof(T)) == hipSuccess &&
      hipMemcpy(din, in, in_words * sizeof(T), hipMemcpyHostToDevice) == hipSuccess) {
    hipLaunchKernelGGL(kernel, dim3((cases + 63) / 64), dim3(64), 0, 0, din, dout, cases);
    if (hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
        hipMemcpy(out, dout, out_words * sizeof(T), hipMemcpyDeviceToHost) == hipSuccess)
      status = 0;
  }
  (void)hipFree(din);
  (void)hipFree(dout);
  return status;
}

}  // namespace

// The device work of an accumulating batch round after the transport: per scene with layer
// paths, the sums and moments of the round's output, then the round counts.
void enqueue_accumulate(Session *h) {
  const uint32_t scene_count = uint32_t(h->scenes.size());
  for (uint32_t i = 0; i < scene_count; ++i) {
    const Scene &s = h->scenes[i];
    if (!s.paths) continue;
    const size_t paths = size_t(s.path_count), weights = size_t(s.nx) * s.ny * s.views;
    double *acc = h->moments + h->moments_base[i];
    const double *round = h->assembled + h->assembled_base[i];
    hipLaunchKernelGGL(accumulate_paths, dim3(unsigned((paths + 255) / 256)), dim3(256), 0, h->stream, s,
                       h->d_controls + i, h->moment_counts + i, round, acc, h->failure);
    hipLaunchKernelGGL(accumulate_weights, dim3(unsigned((weights + 255) / 256)), dim3(256), 0, h->stream, s,
                       h->d_controls + i, h->moment_counts + i, round, acc);
  }
  hipLaunchKernelGGL(count_rounds, dim3((scene_count + 255) / 256), dim3(256), 0, h->stream, h->d_controls,
                     h->moment_counts, scene_count);
}

// One round of every scene of a batch session. Without accumulation, all output tallies are
// copied to `out`; with it, only the Stokes part of each scene (its first stokes_count values)
// is, and the sums and moments stay on the device (scenes with zero photons are skipped).
int batch_round(Session *h, uint32_t round, uint32_t seed, uint32_t max_steps, const uint32_t *photons,
                double *out, bool accumulate, int use_graph, char *error, size_t error_size) {
  if (!h || !photons || !out || !round || !max_steps) {
    set_error(error, error_size, "Invalid HIP batch round");
    return 1;
  }
  const uint32_t scene_count = uint32_t(h->scenes.size());
  std::vector<Controls> controls(scene_count);
  uint64_t total = 0;
  for (uint32_t i = 0; i < scene_count; ++i) {
    if ((!photons[i] && !accumulate) || photons[i] > (1u << 31)) {
      set_error(error, error_size, "Invalid photon count in a batch round");
      return 1;
    }
    controls[i] = Controls{photons[i], 0u, round, seed, max_steps, 0u, photons[i] ? photons[i] : 1u, 1u};
    total += photons[i];
  }
  if (!check(hipSetDevice(h->device), error, error_size, "hipSetDevice")) return 1;
  if (accumulate && !h->moment_counts) {
    size_t doubles = 0;
    h->moments_base.clear();
    for (const Scene &s : h->scenes) {
      h->moments_base.push_back(doubles);
      doubles += s.paths ? moment_layout(s).total : 0;
    }
    if (!check(hipMalloc(&h->moments, (doubles ? doubles : 1) * sizeof(double)), error, error_size,
               "allocate round moments") ||
        !check(hipMalloc(&h->moment_counts, scene_count * sizeof(unsigned)), error, error_size,
               "allocate round counts"))
      return 1;
    launch_fill(reinterpret_cast<uint32_t *>(h->moments), (doubles ? doubles : 1) * 2, h->stream);
    launch_fill(h->moment_counts, scene_count, h->stream);
  }
  if (!check(hipMemcpy(h->d_controls, controls.data(), scene_count * sizeof(Controls), hipMemcpyHostToDevice),
             error, error_size, "upload controls"))
    return 1;
  const int block = block_size<float>(h);
  const bool wide = h->steer == kPlain && h->wide;
  const size_t lds = wide ? lds_bytes_wide(h->wide, h->row, block) : lds_bytes<float>(h, block);
  if (h->steer == kGuided && (!h->proposal || monki_net::proposal_device(h->proposal) != h->device)) {
    set_error(error, error_size, "A proposal on the session's device is required");
    return 1;
  }
  const uint32_t grid =
      !total                   ? 0u
      : h->steer == kGuided   ? persistent_grid<float, false, kGuided>(h, total, block, lds)
      : h->steer == kDirected ? persistent_grid<float, false, kDirected>(h, total, block, lds)
      : wide                  ? persistent_grid<float, false, kWide>(h, total, block, lds)
                              : persistent_grid<float, false, kPlain>(h, total, block, lds);
  // Stacks for a full-occupancy grid: a captured round keeps their address.
  if (h->steer != kPlain &&
      !reserve_splits(h,
                      h->steer == kGuided ? persistent_grid<float, false, kGuided>(h, uint64_t(1) << 40, block, lds)
                                          : persistent_grid<float, false, kDirected>(h, uint64_t(1) << 40, block, lds),
                      block, error, error_size))
    return 1;
  hipGraph_t &graph = accumulate ? h->accumulate_graph : h->graph;
  hipGraphExec_t &exec = accumulate ? h->accumulate_exec : h->graph_exec;
  // The round's device work: clears, transport in the session's sampling mode, reductions.
  auto queue = [&]() {
    const monki_net::Net net = h->steer == kGuided ? monki_net::proposal_net(h->proposal)
                                                   : monki_net::Net{nullptr, nullptr, nullptr};
    const monki_net::EventLog log{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0};
    const bool queued =
        h->steer == kGuided   ? enqueue<float, false, kGuided>(h, 1, grid, block, lds, 0, net, log, error, error_size)
        : h->steer == kDirected ? enqueue<float, false, kDirected>(h, 1, grid, block, lds, 0, net, log, error,
                                                                   error_size)
        : wide                  ? enqueue<float, false, kWide>(h, 1, grid, block, lds, 0, net, log, error, error_size)
                                : enqueue<float, false, kPlain>(h, 1, grid, block, lds, 0, net, log, error, error_size);
    if (queued && accumulate) enqueue_accumulate(h);
    return queued;
  };
  (void)hipEventRecord(h->start, h->stream);
  if (use_graph && total) {
    if (!exec) {
      // Capture the whole round once; replays read the per-scene controls from device memory.
      // The grid is the persistent one for the captured round's work, sized for a full CU load.
      if (!check(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal), error, error_size,
                 "begin capture"))
        return 1;
      const bool queued = queue();
      hipGraph_t captured = nullptr;
      const hipError_t ended = hipStreamEndCapture(h->stream, &captured);
      if (!queued || !check(ended, error, error_size, "end capture")) {
        if (captured) (void)hipGraphDestroy(captured);
        return 1;
      }
      graph = captured;
      if (!check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), error, error_size, "instantiate graph"))
        return 1;
    }
    if (!check(hipGraphLaunch(exec, h->stream), error, error_size, "graph launch")) return 1;
  } else {
    if (!queue()) return 1;
    if (!check(hipGetLastError(), error, error_size, "accumulation launch")) return 1;
  }
  (void)hipEventRecord(h->stop, h->stream);
  if (!finish(h, error, error_size)) return 1;
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, h->start, h->stop);
  h->kernel_seconds = ms * 1e-3;
  if (!accumulate)
    return check(hipMemcpy(out, h->assembled, h->assembled_total * sizeof(double), hipMemcpyDeviceToHost), error,
                 error_size, "copy tallies")
               ? 0
               : 1;
  size_t offset = 0;
  for (uint32_t i = 0; i < scene_count; ++i) {
    const size_t count = size_t(h->scenes[i].stokes_count);
    if (!check(hipMemcpy(out + offset, h->assembled + h->assembled_base[i], count * sizeof(double),
                         hipMemcpyDeviceToHost),
               error, error_size, "copy Stokes tallies"))
      return 1;
    offset += count;
  }
  return 0;
}

// ---- Rounds without allocation -------------------------------------------------------------
// monki_hip_run_static keeps everything a round group needs besides the caller's tables and
// output in one file-scope device array, `scratch`: scene constants, controls, flags and
// counters, the uploaded tables, CDF bucket tables, float32 tally banks, and the float64
// reduced and assembled tallies. Every device that the library runs on holds its own copy for
// the life of the process. Its address is looked up once per device with hipGetSymbolAddress
// (the array's name in host code is not a device address) and kept in static_scratch_at.
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int kScratchDevices = 64;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

void *static_scratch_at[kScratchDevices] = {nullptr};
int static_compute_units[kScratchDevices] = {0};

// Address of this device's scratch (the device must be current); null and error on failure.
unsigned char *static_scratch(int device, char *error, size_t error_size) {
  if (device < 0 || device >= kScratchDevices) {
    set_error(error, error_size, "Device index beyond the static scratch table");
    return nullptr;
  }
  if (!static_scratch_at[device]) {
    void *address = nullptr;
    if (!check(hipGetSymbolAddress(&address, HIP_SYMBOL(scratch)), error, error_size, "static scratch address"))
      return nullptr;
    int units = 0;
    if (!check(hipDeviceGetAttribute(&units, hipDeviceAttributeMultiprocessorCount, device), error, error_size,
               "compute units"))
      return nullptr;
    static_compute_units[device] = units;
    static_scratch_at[device] = address;
  }
  return static_cast<unsigned char *>(static_scratch_at[device]);
}

// The plain tracer's instantiation for a scene's tables (the choice monki_hip_run makes).
int plain_mode(const Scene &s) {
  if (s.detector_surface) return s.nmat == 4 ? kSurface4 : kSurface;
  if (s.scats && s.backward) return s.nmat == 4 ? kBackScats4 : kBackScats;
  if (s.scats) return s.nmat == 4 ? kScats4 : kScats;
  if (s.backward) return s.nmat == 4 ? kBackward4 : kBackward;
  return s.nmat == 4 ? kStokes4 : kPlain;
}

struct StaticRound {
  const Scene *host;  // scene with device pointers into scratch
  Scene *scene;
  Controls *controls;
  unsigned *flags;  // [failure, work]
  unsigned long long *counters;
  float *banks;
  double *reduced, *assembled;
  uint32_t rounds, photons;
  int compute_units;
};

template <int MODE>
bool static_enqueue(const StaticRound &r, char *error, size_t error_size) {
  const Scene &s = *r.host;
  const int block = kBlock32;
  const int row = s.nz | 1;
  const size_t lds = lds_bytes_for<float>(row, block, full_stokes(MODE));
  int resident = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, transport_kernel<float, false, MODE>, block, lds) !=
          hipSuccess ||
      resident < 1)
    resident = 1;
  const uint64_t wanted = (uint64_t(r.photons) + block - 1) / block;
  const uint64_t filled = uint64_t(resident) * uint64_t(r.compute_units > 0 ? r.compute_units : 1);
  const uint32_t grid = uint32_t(wanted < filled ? wanted : filled);
  const size_t bank_values = size_t(s.bank_count) * s.shards * r.rounds;
  launch_fill(reinterpret_cast<uint32_t *>(r.banks), bank_values, 0);
  launch_fill(r.flags, 2, 0);
  launch_fill(reinterpret_cast<uint32_t *>(r.counters), 2 * kCounters, 0);
  if (grid) {
    hipLaunchKernelGGL((transport_kernel<float, false, MODE>), dim3(grid), dim3(block), lds, 0, r.scene, r.controls,
                       1u, row, r.banks, r.flags, r.flags + 1, r.counters, nullptr, 0u,
                       monki_net::Net{nullptr, nullptr, nullptr},
                       monki_net::EventLog{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0},
                       unsteered(), nullptr, 0, 0);
  }
  const size_t summed = size_t(s.bank_count) * r.rounds;
  hipLaunchKernelGGL(reduce_shards<float>, dim3(unsigned((summed + 255) / 256)), dim3(256), 0, 0, r.banks, r.reduced,
                     s.bank_count, s.shards, int(r.rounds));
  hipLaunchKernelGGL(assemble_tallies, dim3(unsigned((summed + 255) / 256)), dim3(256), 0, 0, s, r.reduced,
                     r.assembled, int(r.rounds));
  return check(hipGetLastError(), error, error_size, "static round launch");
}

size_t aligned(size_t bytes) { return (bytes + 255) & ~size_t(255); }

int static_run(uint32_t device, const float *tables, size_t table_count, const uint32_t *controls, double *out,
               char *error, size_t error_size) {
  if (!tables || !controls || !out) {
    set_error(error, error_size, "Invalid static run arguments");
    return 1;
  }
  Scene parsed{};
  if (!parse_tables(tables, table_count, nullptr, parsed, error, error_size)) return 1;
  if (parsed.detector_surface && (parsed.backward || parsed.scats)) {
    set_error(error, error_size, "A ground detector runs forward without the scattering decomposition");
    return 1;
  }
  const Controls c{controls[0], controls[1], controls[2], controls[3],
                   controls[4], controls[5], controls[6], controls[7]};
  if (!c.per_round || !c.rounds || !c.max_steps || c.count > (1u << 31) ||
      uint64_t(c.first) + c.count > uint64_t(c.per_round) * c.rounds) {
    set_error(error, error_size, "Invalid HIP round group");
    return 1;
  }
  // Scratch layout, 256-byte aligned parts.
  const size_t at_scene = 0;
  const size_t at_controls = at_scene + aligned(sizeof(Scene));
  const size_t at_flags = at_controls + aligned(sizeof(Controls));
  const size_t at_counters = at_flags + aligned(2 * sizeof(unsigned));
  const size_t at_tables = at_counters + aligned(kCounters * sizeof(unsigned long long));
  const size_t at_buckets = at_tables + aligned(table_count * sizeof(float));
  const size_t at_banks = at_buckets + aligned(size_t(kMaxSpecies) * (kCdfBuckets + 1) * sizeof(uint16_t));
  const size_t at_reduced = at_banks + aligned(size_t(parsed.bank_count) * parsed.shards * c.rounds * sizeof(float));
  const size_t at_assembled = at_reduced + aligned(size_t(parsed.bank_count) * c.rounds * sizeof(double));
  const size_t end = at_assembled + aligned(size_t(parsed.tally_count) * c.rounds * sizeof(double));
  if (end > kScratchBytes) {
    char detail[160];
    snprintf(detail, sizeof(detail), "%zu bytes of device scratch needed, %zu available", end, kScratchBytes);
    set_error(error, error_size, "Round group too large for the static scratch: %s (fewer rounds or shards)",
              detail);
    return 1;
  }
  if (!check(hipSetDevice(int(device)), error, error_size, "hipSetDevice")) return 1;
  unsigned char *base = static_scratch(int(device), error, error_size);
  if (!base) return 1;
  // Scene view with device pointers into the uploaded tables, CDF buckets built here.
  const float *device_tables = reinterpret_cast<const float *>(base + at_tables);
  Scene view{};
  parse_tables(tables, table_count, device_tables, view, nullptr, 0);
  uint16_t buckets[kMaxSpecies][kCdfBuckets + 1];
  for (int sp = 0; sp < view.species; ++sp) {
    const float *cdf = tables + (parsed.sp[sp].cdf - static_cast<const float *>(nullptr));
    const int n = parsed.sp[sp].lut_size;
    int low = 0;
    for (int b = 0; b <= kCdfBuckets; ++b) {
      const float edge = float(b) / float(kCdfBuckets);
      while (low < n - 2 && cdf[low + 1] <= edge) ++low;
      buckets[sp][b] = uint16_t(low);
    }
    view.sp[sp].cdf_start = reinterpret_cast<const uint16_t *>(base + at_buckets) + size_t(sp) * (kCdfBuckets + 1);
  }
  if (!check(hipMemcpy(base + at_tables, tables, table_count * sizeof(float), hipMemcpyHostToDevice), error,
             error_size, "upload tables") ||
      !check(hipMemcpy(base + at_buckets, buckets, sizeof(buckets), hipMemcpyHostToDevice), error, error_size,
             "upload buckets") ||
      !check(hipMemcpy(base + at_scene, &view, sizeof(Scene), hipMemcpyHostToDevice), error, error_size,
             "upload scene") ||
      !check(hipMemcpy(base + at_controls, &c, sizeof(Controls), hipMemcpyHostToDevice), error, error_size,
             "upload controls"))
    return 1;
  const StaticRound r{&view,
                      reinterpret_cast<Scene *>(base + at_scene),
                      reinterpret_cast<Controls *>(base + at_controls),
                      reinterpret_cast<unsigned *>(base + at_flags),
                      reinterpret_cast<unsigned long long *>(base + at_counters),
                      reinterpret_cast<float *>(base + at_banks),
                      reinterpret_cast<double *>(base + at_reduced),
                      reinterpret_cast<double *>(base + at_assembled),
                      c.rounds,
                      c.count,
                      static_compute_units[device]};
  bool queued = false;
  switch (plain_mode(view)) {
    case kSurface4: queued = static_enqueue<kSurface4>(r, error, error_size); break;
    case kSurface: queued = static_enqueue<kSurface>(r, error, error_size); break;
    case kBackScats4: queued = static_enqueue<kBackScats4>(r, error, error_size); break;
    case kBackScats: queued = static_enqueue<kBackScats>(r, error, error_size); break;
    case kScats4: queued = static_enqueue<kScats4>(r, error, error_size); break;
    case kScats: queued = static_enqueue<kScats>(r, error, error_size); break;
    case kBackward4: queued = static_enqueue<kBackward4>(r, error, error_size); break;
    case kBackward: queued = static_enqueue<kBackward>(r, error, error_size); break;
    case kStokes4: queued = static_enqueue<kStokes4>(r, error, error_size); break;
    default: queued = static_enqueue<kPlain>(r, error, error_size); break;
  }
  if (!queued) return 1;
  unsigned flag = 0;
  if (!check(hipMemcpy(&flag, base + at_flags, sizeof(unsigned), hipMemcpyDeviceToHost), error, error_size,
             "copy failure flag"))
    return 1;
  if (flag) {
    char detail[32];
    snprintf(detail, sizeof(detail), "%u", flag);
    set_error(error, error_size, "HIP transport failed (flag %s); round rejected", detail);
    return 1;
  }
  return check(hipMemcpy(out, base + at_assembled, size_t(parsed.tally_count) * c.rounds * sizeof(double),
                         hipMemcpyDeviceToHost),
               error, error_size, "copy tallies")
             ? 0
             : 1;
}

}  // namespace

namespace {

// Captured batch rounds bake in the transport instantiation and guidance: drop them when the
// sampling changes.
void drop_graphs(Session *h) {
  if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  if (h->accumulate_exec) (void)hipGraphExecDestroy(h->accumulate_exec);
  if (h->accumulate_graph) (void)hipGraphDestroy(h->accumulate_graph);
  h->graph_exec = nullptr;
  h->graph = nullptr;
  h->accumulate_exec = nullptr;
  h->accumulate_graph = nullptr;
}

}  // namespace

extern "C" {

int monki_hip_abi_version(void) { return 2; }

void *monki_hip_create(uint32_t device, const float *tables, size_t table_count, size_t tally_count,
                       uint32_t round_capacity, char *error, size_t error_size) {
  if (!tables) {
    set_error(error, error_size, "Invalid HIP create arguments");
    return nullptr;
  }
  Scene parsed{};
  if (!parse_tables(tables, table_count, nullptr, parsed, error, error_size)) return nullptr;
  if (tally_count != size_t(parsed.tally_count) * size_t(parsed.shards)) {
    set_error(error, error_size, "Tally count does not match the packed tables");
    return nullptr;
  }
  const size_t offsets[2] = {0, table_count};
  return create_session(device, tables, offsets, 1, round_capacity, kBlock32, error, error_size);
}

void *monki_hip_batch_create(uint32_t device, const float *tables, const size_t *offsets, uint32_t scene_count,
                             uint32_t block, char *error, size_t error_size) {
  Session *h = create_session(device, tables, offsets, scene_count, 1, block, error, error_size);
  for (size_t i = 0; h && i < h->scenes.size(); ++i) {
    const Scene &sc = h->scenes[i];
    if (sc.nmat == 4 || sc.backward || sc.scats || sc.detector_surface || sc.register_opposite) {
      release(h);
      set_error(error, error_size, "Batch sessions take forward nmat 1 or 3 scenes (run others one at a time)");
      return nullptr;
    }
  }
  return h;
}

int monki_hip_batch_run(void *handle, uint32_t round, uint32_t seed, uint32_t max_steps, const uint32_t *photons,
                        double *out, int use_graph, char *error, size_t error_size) {
  return batch_round(static_cast<Session *>(handle), round, seed, max_steps, photons, out, false, use_graph, error,
                     error_size);
}

int monki_hip_batch_accumulate(void *handle, uint32_t round, uint32_t seed, uint32_t max_steps,
                               const uint32_t *photons, double *stokes, int use_graph, char *error,
                               size_t error_size) {
  return batch_round(static_cast<Session *>(handle), round, seed, max_steps, photons, stokes, true, use_graph,
                     error, error_size);
}

int monki_hip_batch_moments(void *handle, uint32_t scene, double *out, size_t count, char *error,
                            size_t error_size) {
  Session *h = static_cast<Session *>(handle);
  if (!h || !out || scene >= h->scenes.size() || !h->scenes[scene].paths) {
    set_error(error, error_size, "Invalid HIP batch moments request");
    return 1;
  }
  const MomentLayout m = moment_layout(h->scenes[scene]);
  if (count != 1 + m.total) {
    set_error(error, error_size, "Moments buffer size does not match the scene");
    return 1;
  }
  if (!h->moment_counts) {  // no accumulating round yet
    for (size_t i = 0; i < count; ++i) out[i] = 0.0;
    return 0;
  }
  unsigned rounds = 0;
  if (!check(hipSetDevice(h->device), error, error_size, "hipSetDevice") ||
      !check(hipMemcpy(&rounds, h->moment_counts + scene, sizeof(unsigned), hipMemcpyDeviceToHost), error,
             error_size, "copy round count") ||
      !check(hipMemcpy(out + 1, h->moments + h->moments_base[scene], m.total * sizeof(double),
                       hipMemcpyDeviceToHost),
             error, error_size, "copy moments"))
    return 1;
  out[0] = double(rounds);
  return 0;
}

int monki_hip_run(void *handle, const uint32_t *controls, double *out, char *error, size_t error_size) {
  return launch<float, false>(static_cast<Session *>(handle), controls, out, nullptr, 0, error,
                              error_size);
}

int monki_hip_run_static(uint32_t device, const float *tables, size_t table_count, const uint32_t *controls,
                         double *out, char *error, size_t error_size) {
  return static_run(device, tables, table_count, controls, out, error, error_size);
}

int monki_hip_run_fp64(void *handle, const uint32_t *controls, double *out, char *error,
                       size_t error_size) {
  return launch<double, false>(static_cast<Session *>(handle), controls, out, nullptr, 0, error,
                               error_size);
}

int monki_hip_trace(void *handle, const uint32_t *controls, uint32_t precision, double *out,
                    float *records, size_t record_count, char *error, size_t error_size) {
  if (!records || !record_count) {
    set_error(error, error_size, "Trace needs a record buffer");
    return 1;
  }
  if (precision == 32)
    return launch<float, true>(static_cast<Session *>(handle), controls, out, records, record_count,
                               error, error_size);
  if (precision == 64)
    return launch<double, true>(static_cast<Session *>(handle), controls, out, records,
                                record_count, error, error_size);
  set_error(error, error_size, "Trace precision must be 32 or 64");
  return 1;
}

int monki_hip_attach_proposal(void *handle, void *proposal, float alpha, uint32_t species_mask,
                              uint32_t max_order, char *error, size_t error_size) {
  Session *h = static_cast<Session *>(handle);
  if (!h || !(alpha > 0.0f && alpha <= 1.0f)) {
    set_error(error, error_size, "Invalid session or defensive fraction (0 < alpha <= 1)");
    return 1;
  }
  if (proposal && monki_net::proposal_device(proposal) != h->device) {
    set_error(error, error_size, "The proposal lives on another device");
    return 1;
  }
  h->proposal = proposal;
  h->guidance.alpha = alpha;
  h->guidance.species_mask = species_mask;
  h->guidance.max_order = max_order;
  h->steer = proposal ? kGuided : kPlain;
  drop_graphs(h);
  return 0;
}

int monki_hip_set_network_orders(void *handle, uint32_t orders, char *error, size_t error_size) {
  Session *h = static_cast<Session *>(handle);
  if (!h || !orders) {
    set_error(error, error_size, "Invalid session or network orders (at least 1)");
    return 1;
  }
  h->guidance.network_orders = orders;
  drop_graphs(h);
  return 0;
}

int monki_hip_run_guided(void *handle, const uint32_t *controls, double *out, char *error,
                         size_t error_size) {
  Session *h = static_cast<Session *>(handle);
  return launch<float, false, kGuided>(h, controls, out, nullptr, 0, error, error_size,
                                       h ? h->proposal : nullptr);
}

int monki_hip_set_directed(void *handle, float alpha, float kappa, uint32_t species_mask, uint32_t max_order,
                           char *error, size_t error_size) {
  Session *h = static_cast<Session *>(handle);
  if (!h || !(alpha > 0.0f && alpha <= 1.0f) || !(kappa >= 1.0f && kappa <= 1.0e6f)) {
    set_error(error, error_size, "Invalid session, defensive fraction (0 < alpha <= 1) or kappa (1 to 1e6)");
    return 1;
  }
  h->guidance.alpha = alpha;
  h->guidance.species_mask = species_mask;
  h->guidance.max_order = max_order;
  h->guidance.kappa = kappa;
  h->steer = kDirected;
  drop_graphs(h);
  return 0;
}

int monki_hip_set_steering(void *handle, int lobe_species, float lobe_fraction, int surface, float split_above,
                           float roulette_below, float roulette_weight, char *error, size_t error_size) {
  Session *h = static_cast<Session *>(handle);
  if (!h) {
    set_error(error, error_size, "Invalid HIP session");
    return 1;
  }
  for (const Scene &sc : h->scenes) {
    if (lobe_species >= sc.species) {
      set_error(error, error_size, "The lobe species is not a tabulated species of every scene");
      return 1;
    }
  }
  const bool window_ok = split_above == 0.0f || (split_above > 1.0f && roulette_below >= 0.0f &&
                                                 roulette_below < split_above &&
                                                 (roulette_below == 0.0f ||
                                                  (roulette_weight > roulette_below && roulette_weight <= split_above)));
  if (lobe_species < -1 || !(lobe_fraction >= 0.0f && lobe_fraction <= 1.0f) || !window_ok) {
    set_error(error, error_size,
              "Invalid steering: lobe species -1 or a tabulated species, lobe fraction in [0, 1], weight window "
              "0 (off) or split above 1 > roulette below >= 0, roulette weight in (roulette below, split above]");
    return 1;
  }
  if (lobe_species < 0 && (lobe_fraction > 0.0f || surface) && !(h->guidance.kappa >= 1.0f)) {
    set_error(error, error_size, "vMF view lobes need a concentration: set it with monki_hip_set_directed");
    return 1;
  }
  h->guidance.lobe_species = lobe_species;
  h->guidance.lobe_fraction = lobe_fraction;
  h->guidance.surface = surface ? 1 : 0;
  h->guidance.split_above = split_above;
  h->guidance.roulette_below = roulette_below;
  h->guidance.roulette_weight = roulette_weight;
  drop_graphs(h);
  return 0;
}

int monki_hip_run_directed(void *handle, const uint32_t *controls, double *out, char *error, size_t error_size) {
  Session *h = static_cast<Session *>(handle);
  if (h && !(h->guidance.kappa >= 1.0f)) {
    set_error(error, error_size, "Set the detector-directed lobes (monki_hip_set_directed) first");
    return 1;
  }
  return launch<float, false, kDirected>(h, controls, out, nullptr, 0, error, error_size);
}

int monki_hip_trace_guided(void *handle, const uint32_t *controls, double *out, float *records,
                           size_t record_count, char *error, size_t error_size) {
  Session *h = static_cast<Session *>(handle);
  if (!records || !record_count) {
    set_error(error, error_size, "Trace needs a record buffer");
    return 1;
  }
  return launch<float, true, kGuided>(h, controls, out, records, record_count, error, error_size,
                                      h ? h->proposal : nullptr);
}

int monki_hip_run_logged(void *handle, const uint32_t *controls, double *out, void *proposal,
                         char *error, size_t error_size) {
  return launch<float, false, kLog>(static_cast<Session *>(handle), controls, out, nullptr, 0, error,
                                    error_size, proposal);
}

int monki_hip_counters(void *handle, uint64_t *out, size_t count) {
  Session *h = static_cast<Session *>(handle);
  if (!h || !out) return 1;
  unsigned long long values[kCounters] = {0};
  if (hipSetDevice(h->device) != hipSuccess ||
      hipMemcpy(values, h->counters, sizeof(values), hipMemcpyDeviceToHost) != hipSuccess)
    return 1;
  for (size_t i = 0; i < count && i < size_t(kCounters); ++i) out[i] = values[i];
  return 0;
}

double monki_hip_gpu_seconds(void *handle) {
  return handle ? static_cast<Session *>(handle)->gpu_seconds : 0.0;
}

double monki_hip_kernel_seconds(void *handle) {
  return handle ? static_cast<Session *>(handle)->kernel_seconds : 0.0;
}

int monki_hip_layer_rows(void *handle) {
  if (!handle) return 0;
  const Session *h = static_cast<Session *>(handle);
  return h->wide ? h->wide : h->row;
}

const char *monki_hip_device_name(void *handle) {
  return handle ? static_cast<Session *>(handle)->name : "";
}

int monki_hip_probe_scatter(uint32_t device, const double *cases, size_t count, double *out) {
  if (!cases || !out || count > (1u << 20)) return 1;
  return run_probe<double>(device, probe_scatter_kernel, cases, 8 * count, out, 12 * count, int(count));
}

int monki_hip_probe_philox(uint32_t device, const uint32_t *cases, size_t count, uint32_t *out) {
  if (!cases || !out || count > (1u << 20)) return 1;
  return run_probe<uint32_t>(device, probe_philox_kernel, cases, 6 * count, out, 4 * count, int(count));
}

void monki_hip_close(void *handle) { release(static_cast<Session *>(handle)); }

}  // extern "C"
