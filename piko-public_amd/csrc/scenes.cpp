// scenes.cpp -- deterministic synthetic scene generators for the Piko raster
// pipeline workloads (pretransformed triangle soups, a world-space height
// field with per-vertex normals, batched multi-scene specs, camera matrices,
// Bezier patches tiling a sphere for the Reyes pipeline).
//
// The reference draws its scenes from .scene/.obj files through assimp and
// builds its camera with gluPerspective/gluLookAt (rasterPipelineFixPt/
// main.cpp:268-297, 336-400); none of that is available here, so scenes are
// generated from a counter-based splitmix64 stream per (seed, primitive), so
// any subset of a scene can be regenerated independently and the output does
// not depend on thread count.  All geometry maths runs in double and is
// rounded once to float32.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/piko_scene.h"

namespace {

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// splitmix64 stream keyed by (seed, index): state = mix(seed) ^ index*golden.
struct Stream {
  uint64_t s;
  Stream(uint64_t seed, uint64_t idx) : s(mix64(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull) ^ (idx * 0xD1B54A32D192ED03ull)) {}
  uint64_t next() { s += 0x9E3779B97F4A7C15ull; return mix64(s); }
  double u01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
  float u01f() { return (float)(next() >> 40) * (1.0f / 16777216.0f); }       // [0,1)
  double normal() {  // Box-Muller, first output only
    double u1 = u01(), u2 = u01();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

// toABGRHost (rasterPipelineFixPt/main.cpp:330-334): truncating pack.
inline uint32_t to_abgr_host(float r, float g, float b) {
  return (255u << 24) | ((unsigned)(b * 255.0f) << 16) | ((unsigned)(g * 255.0f) << 8) |
         (unsigned)(r * 255.0f);
}

// Value noise lattice hash in [-1, 1].
inline double lattice(uint64_t seed, int octave, int64_t ix, int64_t iz) {
  uint64_t h = mix64(seed ^ mix64((uint64_t)octave * 0x9E3779B97F4A7C15ull ^
                                  mix64((uint64_t)ix * 0xC2B2AE3D27D4EB4Full ^
                                        (uint64_t)iz * 0x165667B19E3779F9ull)));
  return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

inline double fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }

double value_noise(uint64_t seed, int octaves, double base_freq, double x, double z) {
  double sum = 0.0, amp = 1.0, norm = 0.0, freq = base_freq;
  for (int o = 0; o < octaves; ++o) {
    double fx = x * freq, fz = z * freq;
    double flx = std::floor(fx), flz = std::floor(fz);
    int64_t ix = (int64_t)flx, iz = (int64_t)flz;
    double tx = fade(fx - flx), tz = fade(fz - flz);
    double a = lattice(seed, o, ix, iz), b = lattice(seed, o, ix + 1, iz);
    double c = lattice(seed, o, ix, iz + 1), d = lattice(seed, o, ix + 1, iz + 1);
    double ab = a + (b - a) * tx, cd = c + (d - c) * tx;
    sum += amp * (ab + (cd - ab) * tz);
    norm += amp;
    amp *= 0.5;
    freq *= 2.0;
  }
  return sum / norm;
}

// f(begin, end) over [0, n) split into contiguous ranges on the host's
// threads.  Every generator writes each element from its own counter-based
// stream, so the output does not depend on the split.
template <typename F>
void parallel_for(long long n, F f) {
  const long long min_per = 1 << 14;
  unsigned hw = std::thread::hardware_concurrency();
  long long nt = std::min<long long>(hw ? hw : 1, (n + min_per - 1) / min_per);
  if (nt <= 1) {
    f(0ll, n);
    return;
  }
  std::vector<std::thread> th;
  th.reserve((size_t)nt);
  for (long long k = 0; k < nt; ++k) {
    const long long b = n * k / nt, e = n * (k + 1) / nt;
    th.emplace_back([=] { f(b, e); });
  }
  for (auto &t : th) t.join();
}

}  // namespace

extern "C" {

int piko_scene_tris_uniform(uint64_t seed, int n, int W, int H, float area_median,
                            float sigma, float z_lo, float z_hi, float *pos, uint32_t *col) {
  if (n < 0 || W <= 0 || H <= 0 || !pos || !col) return -1;
  const double k = std::sqrt(4.0 / std::sqrt(3.0));  // side = k*sqrt(area)
  parallel_for(n, [&](long long i0, long long i1) {
  for (long long i = i0; i < i1; ++i) {
    Stream r(seed, (uint64_t)i);
    const double cx = r.u01() * W, cy = r.u01() * H;
    const double area = (double)area_median * std::exp((double)sigma * r.normal());
    const double theta = r.u01() * 6.283185307179586;
    const double side = k * std::sqrt(area);
    const double rad = side / std::sqrt(3.0);  // circumradius
    float *p = pos + 9 * (size_t)i;
    for (int v = 0; v < 3; ++v) {  // angles increase: counter-clockwise
      const double a = theta + v * (6.283185307179586 / 3.0);
      p[3 * v + 0] = (float)(cx + rad * std::cos(a));
      p[3 * v + 1] = (float)(cy + rad * std::sin(a));
      p[3 * v + 2] = z_lo + (z_hi - z_lo) * r.u01f();
    }
    for (int v = 0; v < 3; ++v) {
      const float cr = r.u01f(), cg = r.u01f(), cb = r.u01f();
      col[3 * (size_t)i + v] = to_abgr_host(cr, cg, cb);
    }
  }
  });
  return 0;
}

long long piko_scene_heightfield_count(int grid) { return 2ll * grid * grid; }

int piko_scene_heightfield(uint64_t seed, int grid, float amplitude, int octaves, float *pos,
                           float *nrm) {
  return piko_scene_heightfield_freq(seed, grid, amplitude, octaves, 4.0f, pos, nrm);
}

int piko_scene_heightfield_freq(uint64_t seed, int grid, float amplitude, int octaves, float base_freq,
                                float *pos, float *nrm) {
  if (grid <= 0 || !pos || !nrm || !(base_freq > 0.0f)) return -1;
  const int nv = grid + 1;
  const double step = 2.0 / grid;
  // heights on an (nv+2)^2 lattice with a one-vertex apron for central differences
  const int na = nv + 2;
  double *h = new double[(size_t)na * na];
  parallel_for(na, [&](long long j0, long long j1) {
  for (long long j = j0; j < j1; ++j)
    for (int i = 0; i < na; ++i) {
      const double x = -1.0 + (i - 1) * step, z = -1.0 + (j - 1) * step;
      h[(size_t)j * na + i] = (double)amplitude * value_noise(seed, octaves, base_freq, x, z);
    }
  });
  float *vp = new float[(size_t)nv * nv * 3];
  float *vn = new float[(size_t)nv * nv * 3];
  parallel_for(nv, [&](long long j0, long long j1) {
  for (long long j = j0; j < j1; ++j)
    for (int i = 0; i < nv; ++i) {
      const size_t a = (size_t)(j + 1) * na + (i + 1);
      const double dhx = (h[a + 1] - h[a - 1]) / (2.0 * step);
      const double dhz = (h[a + na] - h[a - na]) / (2.0 * step);
      double nx = -dhx, ny = 1.0, nz = -dhz;
      const double l = std::sqrt(nx * nx + ny * ny + nz * nz);
      const size_t v = (size_t)j * nv + i;
      vp[3 * v + 0] = (float)(-1.0 + i * step);
      vp[3 * v + 1] = (float)h[a];
      vp[3 * v + 2] = (float)(-1.0 + j * step);
      vn[3 * v + 0] = (float)(nx / l);
      vn[3 * v + 1] = (float)(ny / l);
      vn[3 * v + 2] = (float)(nz / l);
    }
  });
  // two triangles per quad, row-major; winding gives a +y geometric normal,
  // i.e. counter-clockwise on screen for a camera above the surface.
  parallel_for(grid, [&](long long j0, long long j1) {
  for (long long j = j0; j < j1; ++j)
    for (int i = 0; i < grid; ++i) {
      const size_t v00 = (size_t)j * nv + i, v10 = v00 + 1, v01 = v00 + nv, v11 = v01 + 1;
      const size_t tris[2][3] = {{v00, v01, v10}, {v10, v01, v11}};
      size_t t = 2 * ((size_t)j * grid + i);
      for (int q = 0; q < 2; ++q, ++t)
        for (int k = 0; k < 3; ++k) {
          std::memcpy(pos + 9 * t + 3 * k, vp + 3 * tris[q][k], 12);
          std::memcpy(nrm + 9 * t + 3 * k, vn + 3 * tris[q][k], 12);
        }
    }
  });
  delete[] h;
  delete[] vp;
  delete[] vn;
  return 0;
}

int piko_camera_view_proj(const float *eye, const float *target, const float *up,
                          float fovy_deg, float aspect, float znear, float zfar, float *m) {
  if (!eye || !target || !up || !m || !(znear > 0) || !(zfar > znear)) return -1;
  // right-handed look-at (camera looks down -z in eye space)
  double f[3] = {(double)target[0] - eye[0], (double)target[1] - eye[1], (double)target[2] - eye[2]};
  double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  for (double &c : f) c /= fl;
  double s[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
  double sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  for (double &c : s) c /= sl;
  double u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
  double V[4][4] = {{s[0], s[1], s[2], 0}, {u[0], u[1], u[2], 0}, {-f[0], -f[1], -f[2], 0}, {0, 0, 0, 1}};
  V[0][3] = -(s[0] * eye[0] + s[1] * eye[1] + s[2] * eye[2]);
  V[1][3] = -(u[0] * eye[0] + u[1] * eye[1] + u[2] * eye[2]);
  V[2][3] = (f[0] * eye[0] + f[1] * eye[1] + f[2] * eye[2]);
  // perspective with depth mapped to [0,1] (near -> 0, far -> 1), w = -z_eye
  const double ct = 1.0 / std::tan(fovy_deg * 3.141592653589793 / 360.0);
  double P[4][4] = {{ct / aspect, 0, 0, 0},
                    {0, ct, 0, 0},
                    {0, 0, zfar / (znear - zfar), znear * zfar / (znear - zfar)},
                    {0, 0, -1, 0}};
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      double acc = 0;
      for (int k = 0; k < 4; ++k) acc += P[r][k] * V[k][c];
      m[c * 4 + r] = (float)acc;  // column-major, as matmultfloat4 indexes it
    }
  return 0;
}

int piko_scene_soup_world(uint64_t seed, int n, float extent, float size_median, float sigma, float *pos,
                          float *nrm) {
  if (n < 0 || !pos || !nrm) return -1;
  parallel_for(n, [&](long long i0, long long i1) {
  for (long long i = i0; i < i1; ++i) {
    Stream r(seed ^ 0x5EEDull, (uint64_t)i);
    double c[3], nv[3];
    for (double &x : c) x = (2.0 * r.u01() - 1.0) * extent;
    // uniform random unit normal
    const double zc = 2.0 * r.u01() - 1.0, ph = r.u01() * 6.283185307179586;
    const double rs = std::sqrt(std::max(0.0, 1.0 - zc * zc));
    nv[0] = rs * std::cos(ph);
    nv[1] = rs * std::sin(ph);
    nv[2] = zc;
    // tangent frame (t1, t2) with t1 x t2 = n, so vertices at increasing angle
    // are counter-clockwise seen from the side the normal points to
    double a[3] = {std::fabs(nv[0]) < 0.9 ? 1.0 : 0.0, std::fabs(nv[0]) < 0.9 ? 0.0 : 1.0, 0.0};
    double t1[3] = {a[1] * nv[2] - a[2] * nv[1], a[2] * nv[0] - a[0] * nv[2], a[0] * nv[1] - a[1] * nv[0]};
    const double l1 = std::sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
    for (double &x : t1) x /= l1;
    double t2[3] = {nv[1] * t1[2] - nv[2] * t1[1], nv[2] * t1[0] - nv[0] * t1[2], nv[0] * t1[1] - nv[1] * t1[0]};
    const double rad = (double)size_median * std::exp((double)sigma * r.normal());
    const double th = r.u01() * 6.283185307179586;
    for (int v = 0; v < 3; ++v) {
      const double ang = th + v * (6.283185307179586 / 3.0);
      const double ca = std::cos(ang) * rad, sa = std::sin(ang) * rad;
      for (int k = 0; k < 3; ++k) {
        pos[9 * (size_t)i + 3 * v + k] = (float)(c[k] + ca * t1[k] + sa * t2[k]);
        nrm[9 * (size_t)i + 3 * v + k] = (float)nv[k];
      }
    }
  }
  });
  return 0;
}

int piko_scene_batch_spec(uint64_t base_seed, int nscenes, int n_lo, int n_hi, int px_lo_edge,
                          int px_hi_edge, int *n_tris, int *W, int *H) {
  if (nscenes < 0 || !n_tris || !W || !H || n_lo <= 0 || n_hi < n_lo) return -1;
  static const double aspects[3] = {1.0, 4.0 / 3.0, 16.0 / 9.0};
  for (int s = 0; s < nscenes; ++s) {
    Stream r(base_seed + (uint64_t)s, 0xB47C4ull);
    const double ln = std::log((double)n_lo) + r.u01() * (std::log((double)n_hi) - std::log((double)n_lo));
    n_tris[s] = (int)std::floor(std::exp(ln));
    const double lp_lo = 2.0 * std::log((double)px_lo_edge), lp_hi = 2.0 * std::log((double)px_hi_edge);
    const double px = std::exp(lp_lo + r.u01() * (lp_hi - lp_lo));
    const double asp = aspects[(int)(r.u01() * 3.0) % 3];
    int w = (int)std::lround(std::sqrt(px * asp));
    int h = (int)std::lround(std::sqrt(px / asp));
    W[s] = w < 16 ? 16 : w;
    H[s] = h < 16 ? 16 : h;
  }
  return 0;
}

// Bicubic Bezier patches tiling the unit sphere on a lat-long grid: a global
// control-point lattice of (3*nlat+1) rows (theta = pi*i/(3*nlat), pole to
// pole) by 3*nlon columns (phi = 2*pi*j/(3*nlon), periodic), each lattice
// point pushed radially by U[-noise, noise] (one draw per lattice point), so
// neighbouring patches share their boundary rows of control points.  The two
// pole rows sit on rings of angular radius kPoleRing instead of collapsing to
// a point: a collapsed row makes dP/du vanish there and the dice stage's
// normal (cross product, then normalise) would divide 0 by 0.  Patch (a, b)
// takes lattice rows 3a..3a+3 and columns 3b..3b+3; its control point
// CP[4*r + c] (row r along v = phi, column c along u = theta) is x,y,z,w
// with w = 1, the piko_patch CP layout (reyesPipe/pikoTypes.h:65-73).  With
// u southward and v eastward, dP/du x dP/dv points into the sphere, which the
// shade stage's normal inversion (shade.pikostage:70-73) turns outward.
// Patches are stored row-major in (a, b).
int piko_scene_sphere_patches(uint64_t seed, int nlat, int nlon, float noise, float *cp) {
  if (nlat <= 0 || nlon <= 0 || !cp || !(noise >= 0.0f)) return -1;
  const int rows = 3 * nlat + 1, cols = 3 * nlon;
  const double pi = 3.141592653589793, kPoleRing = 1e-4;
  for (int a = 0; a < nlat; ++a)
    for (int b = 0; b < nlon; ++b) {
      float *P = cp + ((size_t)a * nlon + b) * 64;
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
          const int i = 3 * a + c, j = (3 * b + r) % cols;
          Stream st(seed, (uint64_t)i * cols + j);
          const double rad = 1.0 + (double)noise * (2.0 * st.u01() - 1.0);
          const double th = std::min(std::max(pi * i / (rows - 1), kPoleRing), pi - kPoleRing);
          const double ph = 2.0 * pi * j / cols;
          float *q = P + 4 * (4 * r + c);
          q[0] = (float)(rad * std::sin(th) * std::cos(ph));
          q[1] = (float)(rad * std::cos(th));
          q[2] = (float)(rad * std::sin(th) * std::sin(ph));
          q[3] = 1.0f;
        }
    }
  return 0;
}

}  // extern "C"
