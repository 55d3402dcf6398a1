// host_init.cpp -- host mirror of the reference's seeded initial particle
// state (emg_host_inipcl in include/emgpu.h).
//
// The reference draws from one MT19937 stream (lib/mtarm/mt19937arm.c) through
// RANU0 / RANN0 (src/wrapper/wrapper.F90:50-345): uniforms are genrand_res53
// values; normals are Box-Muller pairs over an array of uniforms.  inital
// seeds the stream with (myid+1)*10 and consumes one RANU0 and one RANN0 draw
// in its self test (src/main/inital.F90:703-737) before inipcl runs.
#include <cmath>
#include <cstdint>
#include <vector>

#include "emgpu.h"

namespace {

// MT19937 (Matsumoto & Nishimura 1998): 624-word state, twist + tempering
class Mt19937 {
 public:
  explicit Mt19937(uint32_t seed) {
    st_[0] = seed;
    for (uint32_t i = 1; i < 624; i++) st_[i] = 1812433253u * (st_[i - 1] ^ (st_[i - 1] >> 30)) + i;
    pos_ = 624;
  }
  uint32_t next() {
    if (pos_ == 624) regenerate();
    uint32_t y = st_[pos_++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    return y ^ (y >> 18);
  }
  // 53-bit uniform in [0, 1): genrand_res53
  double res53() {
    uint32_t hi = next() >> 5, lo = next() >> 6;
    return ((double)hi * 67108864.0 + (double)lo) * (1.0 / 9007199254740992.0);
  }

 private:
  void regenerate() {
    for (int i = 0; i < 624; i++) {
      uint32_t bits = (st_[i] & 0x80000000u) | (st_[(i + 1) % 624] & 0x7fffffffu);
      uint32_t nv = st_[(i + 397) % 624] ^ (bits >> 1);
      st_[i] = (bits & 1u) ? (nv ^ 0x9908b0dfu) : nv;
    }
    pos_ = 0;
  }
  uint32_t st_[624];
  int pos_;
};

void ranu(Mt19937 &g, double *out, int64_t n) {
  for (int64_t i = 0; i < n; i++) out[i] = g.res53();
}

// RANN0 array form: pairs (cos, sin) over n - n%2 uniforms, odd tail from two
// fresh uniforms (wrapper.F90:223-231 scalar form, 238-272 array form)
void rann(Mt19937 &g, double am, double sd, double *out, int64_t n) {
  const double twopi = 6.28318530717958646;
  if (n == 1) {
    double a = g.res53(), b = g.res53();
    out[0] = am + sd * std::sqrt(-2.0 * std::log(a)) * std::cos(twopi * b);
    return;
  }
  int64_t odd = n & 1, even = n - odd;
  ranu(g, out, even);
  for (int64_t i = 0; i < even; i += 2) {
    double rad = sd * std::sqrt(-2.0 * std::log(out[i]));
    double ang = twopi * out[i + 1];
    out[i] = am + rad * std::cos(ang);
    out[i + 1] = am + rad * std::sin(ang);
  }
  if (odd) {
    double a = g.res53(), b = g.res53();
    out[n - 1] = am + sd * std::sqrt(-2.0 * std::log(a)) * std::cos(twopi * b);
  }
}

}  // namespace

extern "C" int emg_host_inipcl(const emg_host_species *spec, int32_t nspec, const int32_t n[3],
                               const double t[9], uint32_t seed, emg_particle *out) {
  if (!spec || !n || !t || nspec < 1 || nspec > EMG_MAX_SPECIES) return EMG_ERR_ARG;
  int64_t total = 0;
  for (int s = 0; s < nspec; s++) {
    if (spec[s].npin < 0) return EMG_ERR_ARG;
    total += spec[s].npin;
  }
  if (total > 0 && !out) return EMG_ERR_ARG;
  Mt19937 g(seed);
  double scratch;
  ranu(g, &scratch, 1);            // inital: RANU0 self test
  rann(g, 0.0, 1.0, &scratch, 1);  // inital: RANN0 self test
  const double pi = std::acos(-1.0);
  emg_particle *p = out;
  std::vector<double> u;
  for (int s = 0; s < nspec; s++) {
    const emg_host_species &sp = spec[s];
    int64_t np = sp.npin;
    u.resize((size_t)(np > 0 ? np : 1));
    // three uniform arrays -> positions n * u (type_rdist = 1, xl = 0)
    ranu(g, u.data(), np);
    for (int64_t m = 0; m < np; m++) p[m].x = (double)n[0] * u[m] + 0.0;
    ranu(g, u.data(), np);
    for (int64_t m = 0; m < np; m++) p[m].y = (double)n[1] * u[m] + 0.0;
    ranu(g, u.data(), np);
    for (int64_t m = 0; m < np; m++) p[m].z = (double)n[2] * u[m] + 0.0;
    rann(g, sp.spe * std::cos(sp.speth / 180.0 * pi), sp.peth, u.data(), np);
    for (int64_t m = 0; m < np; m++) p[m].vx = u[m];
    rann(g, sp.spe * std::sin(sp.speth / 180.0 * pi), sp.peth, u.data(), np);
    for (int64_t m = 0; m < np; m++) p[m].vy = u[m];
    rann(g, sp.spa, sp.path, u.data(), np);
    for (int64_t m = 0; m < np; m++) p[m].vz = u[m];
    double dx = sp.vdri * std::sin(sp.vdthz / 180.0 * pi) * std::cos(sp.vdthxy / 180.0 * pi);
    double dy = sp.vdri * std::sin(sp.vdthz / 180.0 * pi) * std::sin(sp.vdthxy / 180.0 * pi);
    double dz = sp.vdri * std::cos(sp.vdthz / 180.0 * pi);
    for (int64_t m = 0; m < np; m++) {
      double vx = p[m].vx, vy = p[m].vy, vz = p[m].vz;
      double rx = vx * t[0] + vy * t[1] + vz * t[2];
      double ry = vx * t[3] + vy * t[4] + vz * t[5];
      double rz = vx * t[6] + vy * t[7] + vz * t[8];
      rx = rx + dx;
      ry = ry + dy;
      rz = rz + dz;
      p[m].vx = rx + sp.vdx;
      p[m].vy = ry + sp.vdy;
      p[m].vz = rz + sp.vdz;
      p[m].pid = 0;
      p[m].preside = 0;
      p[m].nid = 0;
      p[m].spec = s + 1;
    }
    p += np;
  }
  return EMG_OK;
}
