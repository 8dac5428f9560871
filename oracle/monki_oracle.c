/*
 * CPU oracle for MONKI photon transport -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is the checker, never the product: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it. It restates Victor Trees's Fortran transport in
 * float64 C, routine by routine, with the Philox4x32-10 draw order of the reference GPU kernel
 * (kernels/metal/transport.metal) so that histories can be compared photon by photon:
 *
 *   launch ............................ mc.f90:351-464
 *   free path, event vs face ........... mc.f90:476-677, ADVANCE_internal mc.f90:1531-1580
 *   Lambertian surface ................. mc.f90:681-1007, GROUND_LAMB2 general_mod:31-65
 *   species choice, absorption ......... mc.f90:1028-1074
 *   local estimate per view ............ mc.f90:1109-1207
 *   detector walk, register mapping .... mc_routines_cartesian_mod.f90:849-1222, ADVANCE :1233
 *   Rayleigh matrix .................... mc_routines_general_mod.f90:74-136
 *   tabulated matrix, interpolation .... mc_routines_general_mod.f90:257-388, interpol_sp.f:105
 *   direction sampling ................. mc_routines_general_mod.f90:492-932
 *   rotation matrices .................. mc_routines_general_mod.f90:942-1242
 *   full Stokes vector (nmat 4) ........ F34/F44 of mc_routines_general_mod.f90:74-136, 257-388,
 *                                        1184-1242; mc.f90:1169-1186, 1344-1353
 *   backward transport ................. weight matrix mc.f90:411-418, 754-755, 793, 894-903,
 *                                        948-955, 981, 1150-1151, 1324-1334, 1393-1395;
 *                                        sampling sign general_mod:583-587, 810; injection
 *                                        register cartesian_mod:1035-1039; no mu0/mu mc.f90:1456
 *   surface detector and grid .......... k_detector = k_register = 0 (main.f90:711-745): views
 *                                        point down, register mapping from the ground
 *                                        (cartesian_mod:1107-1160), no surface signal
 *                                        toward them (mc.f90:804-808)
 *   scattering decomposition ........... calc_radiances_and_scats cartesian_mod:405-836 (order
 *                                        counter mc.f90:433, 689, 1084; backward: the first
 *                                        scattering cell, cartesian_mod:509-521)
 *   mu0/mu scaling ..................... mc.f90:1456-1474
 *
 * Differences from the Fortran, all deliberate: Philox streams instead of RANDOM_NUMBER
 * (the Fortran's compiler RNG has no counter form); tabulated-element brackets by binary
 * search (equal to the equal-angle index estimate of mie_scat_matrix on equal-angle grids);
 * the sampling loop is bounded at 4096 attempts (transport.metal:184) instead of unbounded;
 * the azimuth change after scattering is the same spherical-triangle angle written as an
 * atan2 instead of an acos (well conditioned for small angles); rotation cosines are clamped
 * to [-1, 1].
 * The detector walk is always explicit, cell by cell, with the analytical pixel mapping
 * of mc_routines_cartesian_mod.f90:1076-1105: no uniform-top shortcut.
 *
 * Input is the packed float32 table buffer of the product (layout v2), widened to double. Its
 * cosine block carries 1 - cos and 1 + cos next to cos, so sin^2 of a sampled scattering angle
 * keeps float64 precision near the forward peak although the cosines are stored in float32.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MAX_LAYERS 64
#define MAX_SPECIES 4
#define MAX_VIEWS 3
#define HEADER 72
#define EPS 1.0e-16
#define PI 3.14159265358979323846
#define TRACE_WORDS 8
#define ATTEMPTS 4096

typedef struct {
  int lut;
  double ssalb;
  const float *cosine, *cdf, *f11, *f22, *f33, *f12;
  const float *f34, *f44;  /* nmat 4 only: the two arrays following F12 */
} species_t;

typedef struct {
  int nx, ny, nz, nmat, views, species, paths, backward, scats;
  int surface_detector; /* detector at the ground (views point down) */
  int register_opposite; /* result grid on the other level than the detector */
  double xunit, yunit, depol, theta0, phi0, phi;
  double view[MAX_VIEWS];
  const float *zunit, *albedo, *beta, *cumfrac, *ssalb_gas;
  species_t sp[MAX_SPECIES];
  int stokes_count, path_count, contrib_count, tally_count;
} scene_t;

/* ---- Philox4x32-10 (Random123 constants) ------------------------------------------- */
void oracle_philox(const uint32_t counter[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t x = counter[0], y = counter[1], z = counter[2], w = counter[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int i = 0; i < 10; ++i) {
    uint64_t p0 = (uint64_t)0xD2511F53u * x, p1 = (uint64_t)0xCD9E8D57u * z;
    uint32_t nx = (uint32_t)(p1 >> 32) ^ y ^ k0, ny = (uint32_t)p1;
    uint32_t nz = (uint32_t)(p0 >> 32) ^ w ^ k1, nw = (uint32_t)p0;
    x = nx, y = ny, z = nz, w = nw;
    k0 += 0x9E3779B9u, k1 += 0xBB67AE85u;
  }
  out[0] = x, out[1] = y, out[2] = z, out[3] = w;
}

typedef struct {
  uint32_t photon, round, seed, stream, block, words[4];
  int lane;
} rng_t;

static double draw(rng_t *r) {
  if (r->lane == 4) {
    uint32_t c[4] = {r->photon, r->round, r->block++, r->stream}, k[2] = {r->seed, 0u};
    oracle_philox(c, k, r->words);
    r->lane = 0;
  }
  return ((double)(r->words[r->lane++] >> 9) + 0.5) * 0x1p-23;
}

/* ---- Scattering matrices ------------------------------------------------------------ */
/* Elements F11, F22, F33, F12, F34, F44 divided by 4 (ray_scat_matrix): F34 = 0 and
 * F44 = Delta' F33 before the depolarization scaling, Delta' = (1 - 2 depol) / (1 - depol). */
static void rayleigh6(double depol, double c, double f[6]) {
  double s2 = 1.0 - c * c;
  if (s2 < EPS) s2 = 0.0;
  double delta = (1.0 - depol) / (1.0 + depol / 2.0);
  double delta_prime = (1.0 - 2.0 * depol) / (1.0 - depol);
  f[0] = (delta * 0.75 * (1.0 + c * c) + (1.0 - delta)) / 4.0;
  f[1] = delta * 0.75 * (1.0 + c * c) / 4.0;
  f[2] = delta * 1.5 * c / 4.0;
  f[3] = -delta * 0.75 * s2 / 4.0;
  f[4] = 0.0;
  f[5] = delta * (delta_prime * 1.5 * c) / 4.0;
}

/* Elements F11, F22, F33, F12 divided by 4; species < 0 is Rayleigh (ray_scat_matrix). */
void oracle_rayleigh(double depol, double c, double f[4]) {
  double g[6];
  rayleigh6(depol, c, g);
  for (int e = 0; e < 4; ++e) f[e] = g[e];
}

/* F44 / 4 of the Rayleigh matrix (nmat 4). */
double oracle_rayleigh_f44(double depol, double c) {
  double g[6];
  rayleigh6(depol, c, g);
  return g[5];
}

/* linpol_dp_single on the bracketing pair (interpol_sp.f:105-152). */
static double lerp2(double x0, double x1, double y0, double y1, double x) {
  if (x == x1) return y1;
  if (x == x0) return y0;
  return (y1 - y0) / (x1 - x0) * (x - x0) + y0;
}

static int bracket_cos(const float *axis, int n, double v) {
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    int mid = (lo + hi) / 2;
    if ((double)axis[mid] <= v)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

void oracle_tabulated(const float *cosine, const float *f11, const float *f22, const float *f33,
                      const float *f12, int n, double c, double f[4]) {
  if (c < -1.0) c = -1.0;
  if (c > 1.0) c = 1.0;
  int lo = bracket_cos(cosine, n, c);
  double x0 = cosine[lo], x1 = cosine[lo + 1];
  f[0] = lerp2(x0, x1, f11[lo], f11[lo + 1], c) / 4.0;
  f[1] = lerp2(x0, x1, f22[lo], f22[lo + 1], c) / 4.0;
  f[2] = lerp2(x0, x1, f33[lo], f33[lo + 1], c) / 4.0;
  f[3] = lerp2(x0, x1, f12[lo], f12[lo + 1], c) / 4.0;
}

/* F34 / 4 and F44 / 4 of a tabulated matrix at cosine c (mie_scat_matrix, nmat 4). */
void oracle_tabulated_v(const float *cosine, const float *f34, const float *f44, int n, double c, double f[2]) {
  if (c < -1.0) c = -1.0;
  if (c > 1.0) c = 1.0;
  int lo = bracket_cos(cosine, n, c);
  double x0 = cosine[lo], x1 = cosine[lo + 1];
  f[0] = lerp2(x0, x1, f34[lo], f34[lo + 1], c) / 4.0;
  f[1] = lerp2(x0, x1, f44[lo], f44[lo + 1], c) / 4.0;
}

/* Linear interpolation in cos (mie_scat_matrix) with the weight taken against the packed
 * 1 - cos (1 + cos) nodes near the forward (backward) peak: the same affine interpolation,
 * with node positions that keep float64 precision although cos is stored in float32.
 * one_minus / one_plus are 1 - c and 1 + c of the scattering angle. */
static void elements_at(const scene_t *s, int species, double c, double one_minus, double one_plus,
                        double f[6]) {
  if (species < 0) {
    if (c < -1.0) c = -1.0;
    if (c > 1.0) c = 1.0;
    rayleigh6(s->depol, c, f);
  } else {
    const species_t *t = &s->sp[species];
    if (c < -1.0) c = -1.0;
    if (c > 1.0) c = 1.0;
    int n = t->lut, lo = bracket_cos(t->cosine, n, c);
    double w;
    if (c > 0.5)
      w = ((double)t->cosine[n + lo] - one_minus) / ((double)t->cosine[n + lo] - t->cosine[n + lo + 1]);
    else if (c < -0.5)
      w = (one_plus - t->cosine[2 * n + lo]) / ((double)t->cosine[2 * n + lo + 1] - t->cosine[2 * n + lo]);
    else
      w = (c - t->cosine[lo]) / ((double)t->cosine[lo + 1] - t->cosine[lo]);
    const float *tab[6] = {t->f11, t->f22, t->f33, t->f12, t->f34, t->f44};
    const int count = s->nmat == 4 ? 6 : 4;
    for (int e = 0; e < count; ++e) f[e] = (tab[e][lo] + w * ((double)tab[e][lo + 1] - tab[e][lo])) / 4.0;
    if (count == 4) f[4] = 0.0, f[5] = f[2];
  }
  if (s->nmat == 1) f[3] = 0.0;
}

static void elements(const scene_t *s, int species, double c, double f[6]) {
  elements_at(s, species, c, 1.0 - c, 1.0 + c, f);
}

/* stokes_rot_mat: out = (C1, S1, C2, S2) = (L1(2,2), L1(3,2), L2(2,2), L2(3,2)). */
void oracle_rotation(double cos_sca, double uo, double u1, double phio, double phi1, double out[4]) {
  double dphi = fabs(phio - phi1), rem = fmod(dphi, PI);
  out[0] = 1.0, out[1] = 0.0, out[2] = 1.0, out[3] = 0.0;
  if (rem < EPS || fabs(rem - PI) < EPS) return;
  if (fabs(uo) > 1.0 - EPS && fabs(u1) > 1.0 - EPS) return;
  double fac = 1.0 - uo * uo, sino = fac < EPS ? 0.0 : sqrt(fac);
  fac = 1.0 - u1 * u1;
  double sin1 = fac < EPS ? 0.0 : sqrt(fac);
  double wo[2] = {sino * cos(phio), sino * sin(phio)}, w1[2] = {sin1 * cos(phi1), sin1 * sin(phi1)};
  double cosi1, cosi2, q;
  if (fabs(uo) > 1.0 - EPS || fabs(u1) > 1.0 - EPS) {
    double wop[2] = {cos(phio), sin(phio)}, w1p[2] = {cos(phi1), sin(phi1)};
    if (fabs(u1) > 1.0 - EPS) {
      cosi1 = u1 >= 0 ? 1.0 : -1.0;
      cosi2 = -cosi1 * (w1p[0] * wop[0] + w1p[1] * wop[1]);
      q = wo[1] * w1p[0] - wo[0] * w1p[1];
    } else {
      cosi2 = uo >= 0 ? 1.0 : -1.0;
      cosi1 = -cosi2 * (w1p[0] * wop[0] + w1p[1] * wop[1]);
      q = wop[1] * w1[0] - wop[0] * w1[1];
    }
  } else {
    double a = (1.0 - uo * uo) * (1.0 - cos_sca * cos_sca);
    double b = (1.0 - u1 * u1) * (1.0 - cos_sca * cos_sca);
    cosi1 = (u1 - uo * cos_sca) / sqrt(a > EPS ? a : EPS);
    cosi2 = (uo - u1 * cos_sca) / sqrt(b > EPS ? b : EPS);
    q = wo[1] * w1[0] - wo[0] * w1[1];
  }
  /* |cos i| can pass one when the scattering angle is near zero (the denominators above are
   * floored at 1e-16); clamped as transport.metal:158-159 does, where the Fortran warns. */
  cosi1 = cosi1 > 1.0 ? 1.0 : (cosi1 < -1.0 ? -1.0 : cosi1);
  cosi2 = cosi2 > 1.0 ? 1.0 : (cosi2 < -1.0 ? -1.0 : cosi2);
  double sign = q >= 0 ? 1.0 : -1.0;
  fac = 1.0 - cosi1 * cosi1;
  double sini1 = fac < EPS ? 0.0 : sqrt(fac) * sign;
  fac = 1.0 - cosi2 * cosi2;
  double sini2 = fac < EPS ? 0.0 : sqrt(fac) * sign;
  out[0] = cosi1 * cosi1 - sini1 * sini1;
  out[1] = 2.0 * sini1 * cosi1;
  out[2] = cosi2 * cosi2 - sini2 * sini2;
  out[3] = 2.0 * sini2 * cosi2;
}

/* matmul_stokes_rot then P * S for I, Q, U (mc_routines_general_mod.f90:1184-1242). */
void oracle_scatter(const double f[4], const double r[4], const double s[3], double out[3]) {
  double C1 = r[0], S1 = r[1], C2 = r[2], S2 = r[3];
  double F11 = f[0], F22 = f[1], F33 = f[2], F12 = f[3];
  double P[3][3] = {{F11, F12 * C1, -F12 * S1},
                    {C2 * F12, C1 * C2 * F22 - S1 * S2 * F33, -C2 * S1 * F22 - C1 * S2 * F33},
                    {S2 * F12, C1 * S2 * F22 + C2 * S1 * F33, C1 * C2 * F33 - S1 * S2 * F22}};
  for (int i = 0; i < 3; ++i) out[i] = P[i][0] * s[0] + P[i][1] * s[1] + P[i][2] * s[2];
}

/* matmul_stokes_rot then P * S for the full Stokes vector (nmat 4): f = F11, F22, F33, F12,
 * F34, F44 (/4), with F43 = -F34; V is not rotated. */
void oracle_scatter4(const double f[6], const double r[4], const double s[4], double out[4]) {
  double C1 = r[0], S1 = r[1], C2 = r[2], S2 = r[3];
  double F11 = f[0], F22 = f[1], F33 = f[2], F12 = f[3], F34 = f[4], F44 = f[5];
  double P[4][4] = {{F11, F12 * C1, -F12 * S1, 0.0},
                    {C2 * F12, C1 * C2 * F22 - S1 * S2 * F33, -C2 * S1 * F22 - C1 * S2 * F33, -S2 * F34},
                    {S2 * F12, C1 * S2 * F22 + C2 * S1 * F33, C1 * C2 * F33 - S1 * S2 * F22, C2 * F34},
                    {0.0, -S1 * F34, -C1 * F34, F44}};
  for (int i = 0; i < 4; ++i) out[i] = P[i][0] * s[0] + P[i][1] * s[1] + P[i][2] * s[2] + P[i][3] * s[3];
}

/* The rotated phase matrix Arot of matmul_stokes_rot, zero-padded to 4 x 4 (nmat 1: F11 only). */
static void phase_matrix(int nmat, const double f[6], const double r[4], double P[4][4]) {
  double C1 = r[0], S1 = r[1], C2 = r[2], S2 = r[3];
  double F11 = f[0], F22 = f[1], F33 = f[2], F12 = f[3], F34 = f[4], F44 = f[5];
  memset(P, 0, 16 * sizeof(double));
  P[0][0] = F11;
  if (nmat == 1) return;
  P[0][1] = F12 * C1, P[0][2] = -F12 * S1;
  P[1][0] = C2 * F12, P[1][1] = C1 * C2 * F22 - S1 * S2 * F33, P[1][2] = -C2 * S1 * F22 - C1 * S2 * F33;
  P[2][0] = S2 * F12, P[2][1] = C1 * S2 * F22 + C2 * S1 * F33, P[2][2] = C1 * C2 * F33 - S1 * S2 * F22;
  if (nmat == 4) {
    P[3][1] = -S1 * F34, P[3][2] = -C1 * F34, P[3][3] = F44;
    P[1][3] = -S2 * F34, P[2][3] = C2 * F34;
  }
}

/* W <- W A^T for the backward weight matrix (nmat x nmat, zero-padded). */
static void times_transpose(double W[4][4], const double A[4][4]) {
  double out[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double sum = 0.0;
      for (int k = 0; k < 4; ++k) sum += W[i][k] * A[j][k];
      out[i][j] = sum;
    }
  memcpy(W, out, sizeof(out));
}

/* The continuation Stokes vector of backward transport: the first row of W with U negated
 * (mc.f90:949-955), divided by its first element, and W divided by W11 (mc.f90:981, 1393). */
static void backward_state(int nmat, double W[4][4], double st[4]) {
  st[0] = W[0][0], st[1] = W[0][1], st[2] = -W[0][2], st[3] = nmat == 4 ? W[0][3] : 0.0;
  if (nmat == 1) st[1] = st[2] = 0.0;
  double norm = st[0];
  if (norm > EPS)
    for (int m = 0; m < 4; ++m) st[m] /= norm;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) W[i][j] /= norm;
}

/* P * S for nmat 1, 3 or 4 (S and out hold nmat values; nmat 1 uses F11 only). */
static void scatter_n(int nmat, const double f[6], const double r[4], const double *st, double *out) {
  if (nmat == 4) {
    oracle_scatter4(f, r, st, out);
  } else {
    double s3[3] = {st[0], nmat > 1 ? st[1] : 0.0, nmat > 1 ? st[2] : 0.0}, o3[3];
    oracle_scatter(f, r, s3, o3);
    for (int m = 0; m < 3; ++m) out[m] = o3[m];
  }
}

/* Spherical trigonometry of rayl_scatter_pol / mie_scatter_pol_bound after acceptance, with
 * sin^2 of the scattering angle given (the Fortran's 1 - cos^2, evaluated as (1-cos)(1+cos)). */
static void new_direction(double cos_in, double cos_sca, double sin2, double sigma, double *cos_out,
                          double *dphi) {
  double a = 1.0 - cos_in * cos_in;
  double sin_in = sqrt(a > EPS ? a : EPS), sin_sca = sqrt(sin2 > EPS ? sin2 : EPS);
  double co = cos_in * cos_sca + sin_in * sin_sca * cos(sigma);
  if (co > 1.0) co = 1.0;
  if (co < -1.0) co = -1.0;
  /* |dphi| = acos((cos_sca - cos_in cos_out) / (sin_in sin_out)) of the Fortran, evaluated as
   * the atan2 of the same angle's sine and cosine: acos near +-1 loses half the digits, and the
   * forward peak of a cloud amplifies that loss along a history. */
  double magnitude = atan2(sin_sca * fabs(sin(sigma)), cos_sca * sin_in - cos_in * sin_sca * cos(sigma));
  *cos_out = co;
  *dphi = sigma >= 0 ? -magnitude : magnitude;
}

void oracle_new_direction(double cos_in, double cos_sca, double sigma, double *cos_out, double *dphi) {
  new_direction(cos_in, cos_sca, 1.0 - cos_sca * cos_sca, sigma, cos_out, dphi);
}

/* Rejection sampling, forward sign -1. Returns 0 after ATTEMPTS rejections. */
static int sample(const scene_t *s, int species, const double *st, double sign, double cos_in, rng_t *r,
                  double *cos_out, double *dphi, double *cos_sca, double sampled[2],
                  uint64_t *attempts, double *sigma_out) {
  double bound_ray[6];
  rayleigh6(s->depol, 1.0, bound_ray);
  for (int attempt = 0; attempt < ATTEMPTS; ++attempt) {
    ++*attempts;
    /* sin^2 = (1 - c)(1 + c): exact for Rayleigh (c = 2u - 1); for tabulated species both
     * factors are interpolated from the 1 - cos and 1 + cos rows of the packed cosine block,
     * which keep the precision of the float64 angle grid near the peaks. */
    double c, sin2, one_minus, one_plus;
    if (species < 0) {
      double u = draw(r);
      c = 2.0 * u - 1.0;
      one_minus = 2.0 - 2.0 * u;
      one_plus = 2.0 * u;
    } else {
      const species_t *t = &s->sp[species];
      double u = draw(r);
      int lo = bracket_cos(t->cdf, t->lut, u);
      const float *omc = t->cosine + t->lut, *opc = t->cosine + 2 * t->lut;
      c = lerp2(t->cdf[lo], t->cdf[lo + 1], t->cosine[lo], t->cosine[lo + 1], u);
      one_minus = lerp2(t->cdf[lo], t->cdf[lo + 1], omc[lo], omc[lo + 1], u);
      one_plus = lerp2(t->cdf[lo], t->cdf[lo + 1], opc[lo], opc[lo + 1], u);
    }
    sin2 = one_minus * one_plus;
    double sigma = 2.0 * PI * draw(r) - PI;
    double f[6];
    elements_at(s, species, c, one_minus, one_plus, f);
    /* sign -1 forward, +1 backward (rayl_scatter_pol / mie_scatter_pol_bound). */
    double pdf = f[0] + f[3] * (st[1] * cos(2.0 * sigma) + sign * st[2] * sin(2.0 * sigma));
    double bound = species < 0 ? bound_ray[0] : f[0] + fabs(f[3]);
    if (draw(r) * bound > pdf) continue;
    *cos_sca = c;
    if (sigma_out) *sigma_out = sigma;
    sampled[0] = one_minus, sampled[1] = one_plus;
    new_direction(cos_in, c, sin2, sigma, cos_out, dphi);
    return 1;
  }
  return 0;
}

/* ---- Geometry ----------------------------------------------------------------------- */
typedef struct {
  int i, j, k;
  double x, y, z, a, b, c;
} pos_t;

static double advance(const scene_t *s, pos_t *p) {
  double aa = fabs(p->a), ab = fabs(p->b), ac = fabs(p->c);
  double dx = p->x / (aa > EPS ? aa : EPS), dy = p->y / (ab > EPS ? ab : EPS);
  double dz = p->z / (ac > EPS ? ac : EPS);
  double d = fmin(dx, fmin(dy, dz));
  if (dx == d) {
    p->i = ((p->i + (p->a >= 0 ? 1 : -1)) % s->nx + s->nx) % s->nx;
    p->x = s->xunit;
    p->y = fmax(0.0, p->y - ab * d);
    p->z = fmax(0.0, p->z - ac * d);
  } else if (dy == d) {
    p->j = ((p->j + (p->b >= 0 ? 1 : -1)) % s->ny + s->ny) % s->ny;
    p->x = fmax(0.0, p->x - aa * d);
    p->y = s->yunit;
    p->z = fmax(0.0, p->z - ac * d);
  } else {
    p->k += p->c >= 0 ? 1 : -1;
    p->x = fmax(0.0, p->x - aa * d);
    p->y = fmax(0.0, p->y - ab * d);
    if (p->k >= 0 && p->k < s->nz) p->z = s->zunit[p->k];
  }
  return d;
}

static int cell(const scene_t *s, int i, int j, int k) { return i + s->nx * (j + s->ny * k); }

typedef struct {
  double *tallies;
  double *record;
  uint64_t *counters;
} sink_t;

/* calc_radiances_and_paths_pol for one event and all views. Returns 0 on a walk limit. */
/* With the decomposition (s->scats) the tallies are [Stokes | Stokes_sing | Stokes_mult |
 * contrib_sing | contrib_mult]: nscat is the event's scattering order (surface reflections
 * count), and the contribution also goes to the event cell, level 0 for a surface event. */
static int detect(const scene_t *s, const pos_t *ev, double prob[MAX_VIEWS][4],
                  const double *history, uint32_t max_steps, int column, int nscat, int *first_cell,
                  sink_t *out) {
  /* Cell of the 3-D contribution: the event's; in backward mode the first scattering event's. */
  int level = (ev->k == 0 && ev->z < EPS) ? 0 : ev->k + 1;
  int event_cell = ev->i + s->nx * (ev->j + s->ny * level);
  if (s->backward) {
    if (nscat == 1) *first_cell = event_cell;
    event_cell = *first_cell;
  }
  for (int v = 0; v < s->views; ++v) {
    if (!(prob[v][0] > EPS)) continue;
    double th = s->view[v];
    double va = cos(s->phi) * sin(th), vb = sin(s->phi) * sin(th), vc = cos(th);
    pos_t r = *ev;
    r.a = va, r.b = vb, r.c = vc;
    if (va * ev->a < 0) r.x = fmax(0.0, s->xunit - r.x);
    if (vb * ev->b < 0) r.y = fmax(0.0, s->yunit - r.y);
    if (vc * ev->c < 0) r.z = fmax(0.0, s->zunit[r.k] - r.z);
    double x_old = r.x, y_old = r.y, z_old = r.z;
    double tau = 0.0, lengths[MAX_LAYERS] = {0};
    uint32_t steps = 0;
    while (r.k >= 0 && r.k < s->nz - 1) {
      if (++steps > max_steps) return 0;
      int k_old = r.k;
      double beta = s->beta[cell(s, r.i, r.j, r.k)];
      double d = advance(s, &r);
      tau += d * beta;
      lengths[k_old] += d;
      ++out->counters[4];
    }
    /* Analytical register mapping (mc_routines_cartesian_mod.f90:1041-1105): the grid at TOA
     * (k_register = nz) or at the ground; direction -1 when the grid is not at the detector. */
    double height = 0.0;
    for (int k = 0; k < ev->k; ++k) height += s->zunit[k];
    height += (vc > 0 ? s->zunit[ev->k] - z_old : z_old);
    double top = 0.0;
    for (int k = 0; k < s->nz - 1; ++k) top += s->zunit[k];
    int register_toa = s->surface_detector == s->register_opposite;
    int direction = s->register_opposite ? -1 : 1;
    if (register_toa) height = top - height; /* k_register = nz: measured from TOA */
    double dxh = fabs(va) * height / fabs(vc), dyh = fabs(vb) * height / fabs(vc);
    int di = direction > 0 ? (int)((s->xunit - x_old + dxh) / s->xunit) : (int)((x_old + dxh) / s->xunit);
    int dj = direction > 0 ? (int)((s->yunit - y_old + dyh) / s->yunit) : (int)((y_old + dyh) / s->yunit);
    int si = (va >= 0 ? 1 : -1) * direction, sj = (vb >= 0 ? 1 : -1) * direction;
    int pi = (((ev->i + si * di) % s->nx) + s->nx) % s->nx;
    int pj = (((ev->j + sj * dj) % s->ny) + s->ny) % s->ny;
    if (s->backward) pi = column % s->nx, pj = column / s->nx; /* injection column */

    double att = exp(-tau), mu0mu = s->backward ? 1.0 : fabs(cos(s->theta0)) / fabs(vc);
    double c1 = prob[v][0] * att;
    int pixel = pi + s->nx * (pj + s->ny * v);
    for (int m = 0; m < s->nmat; ++m) out->tallies[s->nmat * pixel + m] += prob[v][m] * att * mu0mu;
    if (s->scats) {
      int group = nscat == 1 ? 0 : 1;
      size_t cell = (size_t)event_cell + (size_t)s->nx * s->ny * s->nz * v;
      for (int m = 0; m < s->nmat; ++m) {
        double value = prob[v][m] * att * mu0mu;
        out->tallies[(size_t)(1 + group) * s->stokes_count + s->nmat * pixel + m] += value;
        out->tallies[3 * (size_t)s->stokes_count + (size_t)group * s->contrib_count + s->nmat * cell + m] += value;
      }
    }
    if (out->record) out->record[4] += c1 * mu0mu;
    if (s->paths) {
      double total = 0.0;
      for (int k = 0; k < s->nz; ++k) {
        double add = c1 * (history[k] + lengths[k]);
        out->tallies[s->stokes_count + pi + s->nx * (pj + s->ny * (k + s->nz * v))] += add;
        total += add;
      }
      out->tallies[s->stokes_count + s->path_count + pixel] += c1;
      if (out->record) out->record[5] += c1, out->record[6] += total;
    }
  }
  return 1;
}

enum { ESCAPED = 1, ABSORBED = 2, FAILED = 3 };

/* One photon history. Returns the end status; *flag receives failure bits. */
static int photon_history(const scene_t *s, uint32_t photon, uint32_t round, uint32_t seed,
                          uint32_t max_steps, sink_t *out, unsigned *flag) {
  rng_t r = {photon, round, seed, 0u, 0u, {0, 0, 0, 0}, 4};
  int column = (int)(photon % (uint32_t)(s->nx * s->ny));
  pos_t p;
  p.i = column % s->nx;
  p.j = column / s->nx;
  p.k = s->nz - 1;
  p.a = sin(s->theta0) * cos(s->phi0);
  p.b = sin(s->theta0) * sin(s->phi0);
  p.c = cos(s->theta0);
  double phi_in = s->phi0;
  p.x = s->xunit * draw(&r);
  p.y = s->yunit * draw(&r);
  p.z = s->zunit[p.k];
  double st[4] = {1.0, 0.0, 0.0, 0.0}, weight = 1.0;
  double W[4][4] = {{1.0, 0.0, 0.0, 0.0}, {0.0, 1.0, 0.0, 0.0}, {0.0, 0.0, 1.0, 0.0}, {0.0, 0.0, 0.0, 1.0}};
  double history[MAX_LAYERS] = {0};
  double remains = -log(fmax(draw(&r), EPS));
  ++out->counters[0];
  double *rec = out->record;
  int nscat = 0; /* events so far, surface reflections included (mc.f90:689, 1084) */
  int first_cell = 0;

  for (uint32_t step = 0; step < max_steps; ++step) {
    double beta = (p.k < s->nz - 1 && p.k >= 0) ? s->beta[cell(s, p.i, p.j, p.k)] : 0.0;
    int surface = 0;
    int event = 0;
    if (beta > EPS) {
      double dist = remains / beta;
      double ax = fabs(p.a), ay = fabs(p.b), az = fabs(p.c);
      double xc = p.x / (ax > EPS ? ax : EPS), yc = p.y / (ay > EPS ? ay : EPS);
      double zc = p.z / (az > EPS ? az : EPS);
      if (dist < xc && dist < yc && dist < zc) {
        p.x = fmax(0.0, p.x - ax * dist);
        p.y = fmax(0.0, p.y - ay * dist);
        p.z = fmax(0.0, p.z - az * dist);
        history[p.k] += dist;
        event = 1;
      }
    }
    if (!event) {
      int k_old = p.k;
      double d = advance(s, &p);
      ++out->counters[3];
      history[k_old] += d;
      remains -= d * beta;
      if (p.k >= s->nz) return ESCAPED;
      if (p.k >= 0) continue;
      surface = 1;
      p.k = 0;
      p.z = 0.0;
    }

    int species = -1;
    if (surface) {
      weight *= s->albedo[p.i + s->nx * p.j];
      ++out->counters[2];
      ++nscat;
      if (rec) rec[1] += 1.0;
    } else {
      double u = draw(&r);
      int c = cell(s, p.i, p.j, p.k), cells = s->nx * s->ny * s->nz;
      for (int t = 0; t < s->species; ++t) {
        if (u < s->cumfrac[t * cells + c]) {
          species = t;
          break;
        }
      }
      weight *= species < 0 ? s->ssalb_gas[p.k] : s->sp[species].ssalb;
      ++out->counters[1];
      ++nscat;
      if (rec) rec[0] += 1.0;
    }
    if (weight < EPS) return ABSORBED;

    double prob[MAX_VIEWS][4];
    for (int v = 0; v < s->views; ++v) {
      double vc = cos(s->view[v]);
      if (surface && s->backward) {
        /* W As S0: the first column of W. */
        for (int m = 0; m < 4; ++m) prob[v][m] = vc > 0 ? W[m][0] * vc * weight : 0.0;
      } else if (surface) {
        /* Surface reflections only travel upward (mc.f90:804-808). */
        prob[v][0] = vc > 0 ? st[0] * vc * weight : 0.0, prob[v][1] = 0.0, prob[v][2] = 0.0, prob[v][3] = 0.0;
      } else {
        double a = 1.0 - p.c * p.c, b = 1.0 - vc * vc;
        double cs = p.c * vc + sqrt(a > EPS ? a : EPS) * sqrt(b > EPS ? b : EPS) * cos(phi_in - s->phi);
        double f[6], rot[4];
        elements(s, species, cs, f);
        if (s->nmat > 1)
          oracle_rotation(cs, p.c, vc, phi_in, s->phi, rot);
        else
          rot[0] = 1.0, rot[1] = 0.0, rot[2] = 1.0, rot[3] = 0.0;
        prob[v][3] = 0.0;
        if (s->backward) {
          /* S_back = W P^T S0 = W times the first row of P. */
          double P[4][4];
          phase_matrix(s->nmat, f, rot, P);
          for (int m = 0; m < 4; ++m)
            prob[v][m] = W[m][0] * P[0][0] + W[m][1] * P[0][1] + W[m][2] * P[0][2] + W[m][3] * P[0][3];
        } else {
          scatter_n(s->nmat, f, rot, st, prob[v]);
        }
        for (int m = 0; m < 4; ++m) prob[v][m] *= weight;
        if (prob[v][0] < 1.0e-100) prob[v][0] = prob[v][1] = prob[v][2] = prob[v][3] = 0.0;
      }
      if (!isfinite(prob[v][0]) || !isfinite(prob[v][1]) || !isfinite(prob[v][2]) || !isfinite(prob[v][3])) {
        *flag |= 1u;
        return FAILED;
      }
    }
    if (!detect(s, &p, prob, history, max_steps, column, nscat, &first_cell, out)) {
      *flag |= 2u;
      return FAILED;
    }

    double cos_out, phi_out;
    if (surface) {
      cos_out = sqrt(draw(&r));
      phi_out = 2.0 * PI * draw(&r);
      st[0] = 1.0, st[1] = 0.0, st[2] = 0.0, st[3] = 0.0;
      if (s->backward) {
        /* W <- W As keeps the first column (As(1,1) = 1 only); then normalize. */
        for (int i = 0; i < 4; ++i) W[i][1] = W[i][2] = W[i][3] = 0.0;
        backward_state(s->nmat, W, st);
      }
    } else {
      double dphi, cos_sca, sampled[2];
      if (!sample(s, species, st, s->backward ? 1.0 : -1.0, p.c, &r, &cos_out, &dphi, &cos_sca, sampled,
                  &out->counters[6], NULL)) {
        *flag |= 4u;
        return FAILED;
      }
      phi_out = phi_in + dphi;
      if (s->nmat > 1) {
        double f[6], rot[4], ns[4] = {0.0, 0.0, 0.0, 0.0};
        elements_at(s, species, cos_sca, sampled[0], sampled[1], f);
        oracle_rotation(cos_sca, p.c, cos_out, phi_in, phi_out, rot);
        if (s->backward) {
          double P[4][4];
          phase_matrix(s->nmat, f, rot, P);
          times_transpose(W, P);
          backward_state(s->nmat, W, st);
          ns[0] = 1.0, ns[1] = st[1], ns[2] = st[2], ns[3] = st[3];
        } else {
          scatter_n(s->nmat, f, rot, st, ns);
        }
        st[0] = 1.0, st[1] = ns[1] / ns[0], st[2] = ns[2] / ns[0], st[3] = ns[3] / ns[0];
        if (!isfinite(st[1]) || !isfinite(st[2]) || !isfinite(st[3]) ||
            sqrt(st[1] * st[1] + st[2] * st[2] + st[3] * st[3]) > 1.01) {
          *flag |= 8u;
          return FAILED;
        }
      }
      remains = -log(fmax(draw(&r), EPS));
    }
    double sine = sqrt(fmax(0.0, 1.0 - cos_out * cos_out));
    double a = cos(phi_out) * sine, b = sin(phi_out) * sine, c = cos_out;
    if (a * p.a < 0) p.x = fmax(0.0, s->xunit - p.x);
    if (b * p.b < 0) p.y = fmax(0.0, s->yunit - p.y);
    if (c * p.c < 0) p.z = fmax(0.0, s->zunit[p.k] - p.z);
    p.a = a, p.b = b, p.c = c;
    phi_in = phi_out;
    if (rec) rec[3] = weight, rec[7] = p.k;
  }
  *flag |= 16u;
  return FAILED;
}

static int word(const float *t, int index) { return (int)t[index]; }

/* Parse the packed float32 tables (layout v2; the oracle reads the expanded extinction and
   fraction tables, not the factored parts of words 64-65). Returns 0 on success. */
static int parse(const float *t, size_t count, scene_t *s) {
  if (count < HEADER || t[0] != 2.0f || (size_t)t[26] != count) return 1;
  s->nx = word(t, 1), s->ny = word(t, 2), s->nz = word(t, 3), s->nmat = word(t, 4);
  s->views = word(t, 5), s->species = word(t, 6), s->paths = word(t, 7), s->backward = word(t, 28);
  s->scats = word(t, 29);
  s->surface_detector = word(t, 30);
  s->register_opposite = word(t, 31);
  if (s->scats) s->paths = 0; /* calc_radiances_and_scats replaces the path tallies */
  if (s->nz > MAX_LAYERS || s->views > MAX_VIEWS || s->species > MAX_SPECIES) return 1;
  s->xunit = t[10], s->yunit = t[11], s->depol = t[12];
  s->theta0 = t[13], s->phi0 = t[14], s->phi = t[15];
  for (int v = 0; v < s->views; ++v) s->view[v] = t[16 + v];
  s->zunit = t + word(t, 19), s->albedo = t + word(t, 20), s->beta = t + word(t, 21);
  s->cumfrac = t + word(t, 22), s->ssalb_gas = t + word(t, 23);
  for (int i = 0; i < s->species; ++i) {
    int h = 32 + 8 * i;
    species_t *p = &s->sp[i];
    p->lut = word(t, h), p->ssalb = t[h + 1];
    p->cosine = t + word(t, h + 2), p->cdf = t + word(t, h + 3), p->f11 = t + word(t, h + 4);
    p->f22 = t + word(t, h + 5), p->f33 = t + word(t, h + 6), p->f12 = t + word(t, h + 7);
    p->f34 = s->nmat == 4 ? p->f12 + p->lut : NULL;
    p->f44 = s->nmat == 4 ? p->f12 + 2 * p->lut : NULL;
  }
  s->stokes_count = s->nmat * s->nx * s->ny * s->views;
  s->path_count = s->nx * s->ny * s->nz * s->views;
  s->contrib_count = s->nmat * s->nx * s->ny * s->nz * s->views;
  s->tally_count = s->stokes_count + (s->paths ? s->path_count + s->nx * s->ny * s->views : 0);
  if (s->scats) s->tally_count = 3 * s->stokes_count + 2 * s->contrib_count;
  return 0;
}

/*
 * Draw `count` scattering directions for species `species` (-1: gas) of the packed tables
 * with incident Stokes (1, q, u) and incident direction cosine cos_in, sample i on the Philox
 * stream (photon i, round 0, seed). out receives (cos_sca, sigma, cos_out, dphi) per sample.
 * Returns 0, or -1 on a table or sampling failure.
 */
int oracle_sample(const float *tables, size_t table_count, int species, double q, double u,
                  double cos_in, uint32_t seed, uint32_t count, double *out) {
  scene_t s;
  if (parse(tables, table_count, &s) || species >= s.species) return -1;
  const double st[4] = {1.0, q, u, 0.0};
  for (uint32_t i = 0; i < count; ++i) {
    rng_t r = {i, 0u, seed, 0u, 0u, {0, 0, 0, 0}, 4};
    double cos_out, dphi, cos_sca, sigma, sampled[2];
    uint64_t attempts = 0;
    if (!sample(&s, species, st, -1.0, cos_in, &r, &cos_out, &dphi, &cos_sca, sampled, &attempts, &sigma))
      return -1;
    double *o = out + 4 * (size_t)i;
    o[0] = cos_sca, o[1] = sigma, o[2] = cos_out, o[3] = dphi;
  }
  return 0;
}

int oracle_tally_count(const float *tables, size_t count) {
  scene_t s;
  return parse(tables, count, &s) ? -1 : s.tally_count;
}

/*
 * Run photons [first, first + count) of a round group with the product's control words
 * (count, first, round, seed, max_steps, -, photons per round, rounds). tallies receives
 * rounds x tally_count doubles (zeroed here). records (optional) receives TRACE_WORDS
 * doubles for the first record_count photons of the range. counters[8] are accumulated.
 * Returns the failure flags (0 on success).
 */
unsigned oracle_run(const float *tables, size_t table_count, const uint32_t *controls,
                    double *tallies, double *records, size_t record_count, uint64_t *counters) {
  scene_t s;
  if (parse(tables, table_count, &s)) return 0x80000000u;
  uint32_t count = controls[0], first = controls[1], round = controls[2], seed = controls[3];
  uint32_t max_steps = controls[4], per_round = controls[6], rounds = controls[7];
  memset(tallies, 0, sizeof(double) * (size_t)s.tally_count * rounds);
  unsigned flag = 0;
  for (uint32_t lane = 0; lane < count && !flag; ++lane) {
    uint32_t identity = first + lane;
    uint32_t offset = identity / per_round, photon = identity % per_round;
    double *rec = (records && lane < record_count) ? records + (size_t)lane * TRACE_WORDS : NULL;
    if (rec) {
      memset(rec, 0, sizeof(double) * TRACE_WORDS);
      rec[3] = 1.0;
    }
    sink_t out = {tallies + (size_t)offset * s.tally_count, rec, counters};
    int status = photon_history(&s, photon, round + offset, seed, max_steps, &out, &flag);
    if (rec) rec[2] = status;
  }
  return flag;
}
