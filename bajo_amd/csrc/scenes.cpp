// Seeded synthetic scenes for the benchmark workloads (host code).
//
// bajo has no generator for these scenes (its fixtures are a regular grid,
// bajo/benchmark/bvh_fixtures.mojo:11-58, and OBJ assets), so they are
// defined here:
//   - PCG32 XSH-RR (O'Neill 2014, pcg32_srandom / pcg32_random), floats as
//     (u32 >> 8) * 2^-24, unit directions by rejection in [-1,1]^3;
//   - random triangle soup: v0 ~ U[0,1]^3, two edges of length
//     U[edge_min, edge_max] along random unit directions;
//   - geodesic icosphere of frequency f (20 f^2 triangles, face-major),
//     radially displaced by `amplitude` * (2 * value_noise(p * noise_frequency) - 1).
// Shared lattice points are computed from corner weights summed in corner
// index order in double precision, so neighbouring faces produce
// bit-identical vertices (the mesh is crack-free).
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "../../include/bajo_hip.h"

namespace {

struct Pcg32 {
  uint64_t state = 0, inc = 0;
  Pcg32(uint64_t initstate, uint64_t initseq) {
    inc = (initseq << 1u) | 1u;
    next();
    state += initstate;
    next();
  }
  uint32_t next() {
    uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((-rot) & 31));
  }
  float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
  void direction(double* out) {
    for (;;) {
      double x = 2.0 * uniform() - 1.0, y = 2.0 * uniform() - 1.0, z = 2.0 * uniform() - 1.0;
      double l2 = x * x + y * y + z * z;
      if (l2 > 1e-6 && l2 <= 1.0) {
        double inv = 1.0 / sqrt(l2);
        out[0] = x * inv;
        out[1] = y * inv;
        out[2] = z * inv;
        return;
      }
    }
  }
};

struct ValueNoise {
  uint8_t perm[256];
  float value[256];
  explicit ValueNoise(uint64_t seed) {
    Pcg32 rng(seed, 7);
    for (int i = 0; i < 256; ++i) perm[i] = (uint8_t)i;
    for (int i = 255; i > 0; --i) {
      int j = (int)(rng.next() % (uint32_t)(i + 1));
      uint8_t t = perm[i];
      perm[i] = perm[j];
      perm[j] = t;
    }
    for (int i = 0; i < 256; ++i) value[i] = rng.uniform();
  }
  double lattice(long x, long y, long z) const {
    int h = perm[(perm[(perm[x & 255] + (int)(y & 255)) & 255] + (int)(z & 255)) & 255];
    return value[h];
  }
  static double smooth(double t) { return t * t * (3.0 - 2.0 * t); }
  double eval(double x, double y, double z) const {
    double fx = floor(x), fy = floor(y), fz = floor(z);
    long ix = (long)fx, iy = (long)fy, iz = (long)fz;
    double tx = smooth(x - fx), ty = smooth(y - fy), tz = smooth(z - fz);
    double c[2][2][2];
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int d = 0; d < 2; ++d) c[a][b][d] = lattice(ix + a, iy + b, iz + d);
    double x00 = c[0][0][0] + (c[1][0][0] - c[0][0][0]) * tx;
    double x10 = c[0][1][0] + (c[1][1][0] - c[0][1][0]) * tx;
    double x01 = c[0][0][1] + (c[1][0][1] - c[0][0][1]) * tx;
    double x11 = c[0][1][1] + (c[1][1][1] - c[0][1][1]) * tx;
    double y0 = x00 + (x10 - x00) * ty, y1 = x01 + (x11 - x01) * ty;
    return y0 + (y1 - y0) * tz;
  }
};

const double kIcoT = 1.6180339887498948482;
const double kIcoV[12][3] = {{-1, kIcoT, 0}, {1, kIcoT, 0}, {-1, -kIcoT, 0}, {1, -kIcoT, 0},
                             {0, -1, kIcoT}, {0, 1, kIcoT}, {0, -1, -kIcoT}, {0, 1, -kIcoT},
                             {kIcoT, 0, -1}, {kIcoT, 0, 1}, {-kIcoT, 0, -1}, {-kIcoT, 0, 1}};
const int kIcoF[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11}, {1, 5, 9},  {5, 11, 4},
                          {11, 10, 2}, {10, 7, 6}, {7, 1, 8},   {3, 9, 4},  {3, 4, 2},   {3, 2, 6},  {3, 6, 8},
                          {3, 8, 9},  {4, 9, 5},  {2, 4, 11},  {6, 2, 10}, {8, 6, 7},   {9, 8, 1}};

}  // namespace

extern "C" {

int32_t bajo_scene_soup(uint64_t initstate, uint64_t initseq, uint64_t n, float edge_min, float edge_max,
                        float* out) {
  if (!out && n) return BAJO_ERR_INVALID_ARGUMENT;
  Pcg32 rng(initstate, initseq);
  for (uint64_t i = 0; i < n; ++i) {
    float* t = out + 9 * i;
    float v0[3] = {rng.uniform(), rng.uniform(), rng.uniform()};
    t[0] = v0[0];
    t[1] = v0[1];
    t[2] = v0[2];
    for (int e = 0; e < 2; ++e) {
      double len = (double)edge_min + (double)(edge_max - edge_min) * (double)rng.uniform();
      double dir[3];
      rng.direction(dir);
      for (int k = 0; k < 3; ++k) t[3 + 3 * e + k] = (float)((double)v0[k] + len * dir[k]);
    }
  }
  return BAJO_OK;
}

int32_t bajo_scene_icosphere(uint32_t f, uint64_t n, uint64_t noise_seed, float amplitude, float noise_frequency,
                             float* out) {
  if ((!out && n) || f == 0) return BAJO_ERR_INVALID_ARGUMENT;
  if (n > 20ull * f * f) return BAJO_ERR_INVALID_ARGUMENT;
  ValueNoise noise(noise_seed);
  double corner[12][3];
  for (int i = 0; i < 12; ++i) {
    double l = sqrt(kIcoV[i][0] * kIcoV[i][0] + kIcoV[i][1] * kIcoV[i][1] + kIcoV[i][2] * kIcoV[i][2]);
    for (int k = 0; k < 3; ++k) corner[i][k] = kIcoV[i][k] / l;
  }
  uint64_t written = 0;
  for (int face = 0; face < 20 && written < n; ++face) {
    const int ca = kIcoF[face][0], cb = kIcoF[face][1], cc = kIcoF[face][2];
    auto point = [&](long i, long j, float* dst) {
      // weights of corners (a, b, c) = (f - i - j, i, j), summed in corner-index order
      struct W { int c; double w; } w[3] = {{ca, (double)((long)f - i - j)}, {cb, (double)i}, {cc, (double)j}};
      for (int x = 0; x < 3; ++x)
        for (int y = x + 1; y < 3; ++y)
          if (w[y].c < w[x].c) { W tmp = w[x]; w[x] = w[y]; w[y] = tmp; }
      double p[3];
      for (int k = 0; k < 3; ++k)
        p[k] = (w[0].w * corner[w[0].c][k] + w[1].w * corner[w[1].c][k]) + w[2].w * corner[w[2].c][k];
      double l = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
      double u[3] = {p[0] / l, p[1] / l, p[2] / l};
      double nz = noise.eval(u[0] * noise_frequency, u[1] * noise_frequency, u[2] * noise_frequency);
      double r = 1.0 + (double)amplitude * (2.0 * nz - 1.0);
      for (int k = 0; k < 3; ++k) dst[k] = (float)(u[k] * r);
    };
    for (long i = 0; i < (long)f && written < n; ++i) {
      for (long j = 0; j + i < (long)f && written < n; ++j) {
        float* t = out + 9 * written;
        point(i, j, t);
        point(i + 1, j, t + 3);
        point(i, j + 1, t + 6);
        ++written;
        if (j + i + 1 < (long)f && written < n) {
          t = out + 9 * written;
          point(i + 1, j, t);
          point(i + 1, j + 1, t + 3);
          point(i, j + 1, t + 6);
          ++written;
        }
      }
    }
  }
  return BAJO_OK;
}

}  // extern "C"
