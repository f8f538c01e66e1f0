// omr_scenegen.cpp -- procedural benchmark scenes (seeded, deterministic).
//
//  kind 0  Cornell box: five walls and an area light (2 triangles each), a
//          mirror and a glass sphere (positions of CornellBoxScene,
//          src/scenes/CornellBoxScene.cpp:30-31).
//  kind 1  kind 0 + a level-6 icosphere (81,920 triangles) displaced by
//          4-octave value noise; each triangle Lambert (random albedo) or glass.
//  kind 2  1000 x 500 heightfield terrain (997,002 triangles, 6-octave fBm)
//          + 200 spheres (diffuse / mirror / glass) under a 2048 x 1024 HDR
//          sky with a Gaussian sun lobe.
//  kind 3  (2,3) torus knot tube (~300k triangles), noise displaced, on a
//          floor under a 512 x 256 sky.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/omr_capi.h"

namespace {

struct V {
  double x, y, z;
};
V operator+(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V operator-(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V operator*(V a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dotv(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V crossv(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V normv(V a) { double l = std::sqrt(dotv(a, a)); return l > 0 ? a * (1.0 / l) : a; }

uint64_t splitmix(uint64_t &s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
double uni(uint64_t &s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

// Value noise: hashed lattice values in [-1,1], smoothstep-trilinear.
double lattice(int64_t x, int64_t y, int64_t z, uint64_t seed) {
  uint64_t h = seed ^ ((uint64_t)x * 0x8da6b343ull) ^ ((uint64_t)y * 0xd8163841ull) ^ ((uint64_t)z * 0xcb1ab31full);
  uint64_t s = h;
  return uni(s) * 2.0 - 1.0;
}
double vnoise(double x, double y, double z, uint64_t seed) {
  double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
  int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
  double tx = x - fx, ty = y - fy, tz = z - fz;
  auto sm = [](double t) { return t * t * (3.0 - 2.0 * t); };
  tx = sm(tx); ty = sm(ty); tz = sm(tz);
  double c[2][2][2];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int d = 0; d < 2; d++) c[a][b][d] = lattice(ix + a, iy + b, iz + d, seed);
  double x00 = c[0][0][0] + (c[1][0][0] - c[0][0][0]) * tx, x10 = c[0][1][0] + (c[1][1][0] - c[0][1][0]) * tx;
  double x01 = c[0][0][1] + (c[1][0][1] - c[0][0][1]) * tx, x11 = c[0][1][1] + (c[1][1][1] - c[0][1][1]) * tx;
  double y0 = x00 + (x10 - x00) * ty, y1 = x01 + (x11 - x01) * ty;
  return y0 + (y1 - y0) * tz;
}
double fbm(V p, int octaves, uint64_t seed) {
  double sum = 0, amp = 1, norm = 0, f = 1;
  for (int o = 0; o < octaves; o++) {
    sum += amp * vnoise(p.x * f, p.y * f, p.z * f, seed + (uint64_t)o * 1013);
    norm += amp;
    amp *= 0.5;
    f *= 2.0;
  }
  return sum / norm;
}

struct Builder {
  std::vector<float> tri, nrm, sph;
  std::vector<uint32_t> pmat_tri, pmat_sph;
  std::vector<omr_material> mats;
  std::vector<float> env;
  uint32_t env_w = 0, env_h = 0;

  uint32_t material(uint32_t type, V color, V emission, double ior) {
    omr_material m;
    std::memset(&m, 0, sizeof(m));
    m.type = type;
    m.color[0] = (float)color.x; m.color[1] = (float)color.y; m.color[2] = (float)color.z;
    m.emission[0] = (float)emission.x; m.emission[1] = (float)emission.y; m.emission[2] = (float)emission.z;
    m.ior = (float)ior;
    m.rate = 1.0f;
    mats.push_back(m);
    return (uint32_t)mats.size() - 1;
  }
  // Triangle with its winding chosen so (v1-v0)x(v2-v0) faces `want`
  // (single-sided test: rays arriving from the `want` side hit it).
  void triangle(V a, V b, V c, V want, uint32_t mat) {
    V n = crossv(b - a, c - a);
    if (dotv(n, want) < 0) { V t = b; b = c; c = t; n = n * -1.0; }
    n = normv(n);
    V e1 = b - a, e2 = c - a;
    float t9[9] = {(float)a.x, (float)a.y, (float)a.z, (float)e1.x, (float)e1.y, (float)e1.z,
                   (float)e2.x, (float)e2.y, (float)e2.z};
    float n9[9] = {(float)n.x, (float)n.y, (float)n.z, 0, 0, 0, 0, 0, 0};
    tri.insert(tri.end(), t9, t9 + 9);
    nrm.insert(nrm.end(), n9, n9 + 9);
    pmat_tri.push_back(mat);
  }
  void quad(V a, V b, V c, V d, V want, uint32_t mat) {
    triangle(a, b, c, want, mat);
    triangle(a, c, d, want, mat);
  }
  void sphere(V c, double r, uint32_t mat) {
    sph.push_back((float)c.x); sph.push_back((float)c.y); sph.push_back((float)c.z); sph.push_back((float)r);
    pmat_sph.push_back(mat);
  }
};

void cornell(Builder &B, bool spheres) {
  uint32_t red = B.material(OMR_MAT_LAMBERT, {0.75, 0.25, 0.25}, {0, 0, 0}, 0);
  uint32_t blue = B.material(OMR_MAT_LAMBERT, {0.25, 0.25, 0.75}, {0, 0, 0}, 0);
  uint32_t grey = B.material(OMR_MAT_LAMBERT, {0.75, 0.75, 0.75}, {0, 0, 0}, 0);
  uint32_t light = B.material(OMR_MAT_LAMBERT, {0, 0, 0}, {12, 12, 12}, 0);
  const double s = 100;
  B.quad({0, 0, 0}, {0, 0, s}, {0, s, s}, {0, s, 0}, {1, 0, 0}, red);        // left x = 0
  B.quad({s, 0, 0}, {s, s, 0}, {s, s, s}, {s, 0, s}, {-1, 0, 0}, blue);      // right x = 100
  B.quad({0, 0, 0}, {s, 0, 0}, {s, 0, s}, {0, 0, s}, {0, 1, 0}, grey);       // floor
  B.quad({0, s, 0}, {0, s, s}, {s, s, s}, {s, s, 0}, {0, -1, 0}, grey);      // ceiling
  B.quad({0, 0, 0}, {0, s, 0}, {s, s, 0}, {s, 0, 0}, {0, 0, 1}, grey);       // back z = 0
  B.quad({35, 99.9, 35}, {35, 99.9, 65}, {65, 99.9, 65}, {65, 99.9, 35}, {0, -1, 0}, light);
  if (spheres) {
    uint32_t mirror = B.material(OMR_MAT_SPECULAR, {0.99, 0.99, 0.99}, {0, 0, 0}, 0);
    uint32_t glass = B.material(OMR_MAT_REFRACTION, {0.99, 0.99, 0.99}, {0, 0, 0}, 1.5);
    B.sphere({27, 16.5, 47}, 16.5, mirror);
    B.sphere({73, 16.5, 78}, 16.5, glass);
  }
}

void icosphere(Builder &B, V center, double radius, int level, uint64_t seed, double amp) {
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  std::vector<V> vs = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                       {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
  for (auto &v : vs) v = normv(v);
  std::vector<uint32_t> f = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
                             3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
  for (int l = 0; l < level; l++) {
    std::vector<uint32_t> nf;
    nf.reserve(f.size() * 4);
    std::vector<std::pair<uint64_t, uint32_t>> cache;
    auto mid = [&](uint32_t a, uint32_t b) {
      uint64_t key = a < b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a;
      for (auto it = cache.end(); it != cache.begin();) {  // recent edges first
        --it;
        if (it->first == key) return it->second;
        if (cache.end() - it > 64) break;
      }
      vs.push_back(normv((vs[a] + vs[b]) * 0.5));
      cache.push_back({key, (uint32_t)vs.size() - 1});
      return (uint32_t)vs.size() - 1;
    };
    // exact midpoint dedup through a map keyed by edge
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> adj(vs.size() + f.size());
    auto mid2 = [&](uint32_t a, uint32_t b) {
      uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
      for (auto &e : adj[lo])
        if (e.first == hi) return e.second;
      vs.push_back(normv((vs[a] + vs[b]) * 0.5));
      uint32_t id = (uint32_t)vs.size() - 1;
      adj[lo].push_back({hi, id});
      return id;
    };
    (void)mid;
    for (size_t i = 0; i < f.size(); i += 3) {
      uint32_t a = f[i], b = f[i + 1], c = f[i + 2];
      uint32_t ab = mid2(a, b), bc = mid2(b, c), ca = mid2(c, a);
      uint32_t add[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
      nf.insert(nf.end(), add, add + 12);
    }
    f.swap(nf);
  }
  std::vector<V> pos(vs.size());
  for (size_t i = 0; i < vs.size(); i++) {
    double d = fbm(vs[i] * 2.5, 4, seed);
    pos[i] = center + vs[i] * (radius * (1.0 + amp * d));
  }
  uint64_t rs = seed ^ 0xabcdefull;
  uint32_t glass = B.material(OMR_MAT_REFRACTION, {0.99, 0.99, 0.99}, {0, 0, 0}, 1.5);
  for (size_t i = 0; i < f.size(); i += 3) {
    V a = pos[f[i]], b = pos[f[i + 1]], c = pos[f[i + 2]];
    V outward = (a + b + c) * (1.0 / 3.0) - center;
    uint32_t m;
    if (uni(rs) < 0.5) {
      m = B.material(OMR_MAT_LAMBERT, {0.2 + 0.6 * uni(rs), 0.2 + 0.6 * uni(rs), 0.2 + 0.6 * uni(rs)}, {0, 0, 0}, 0);
    } else {
      m = glass;
    }
    B.triangle(a, b, c, outward, m);
  }
}

// Sky: luminance 0.2 at the horizon to 1.0 at the zenith (dimmer below), plus
// a Gaussian sun lobe 5e4 * exp(-g^2 / (2 sigma^2)), sigma = 1 degree.
void sky(Builder &B, uint32_t w, uint32_t h, uint64_t seed) {
  uint64_t s = seed ^ 0x5u;
  double elev = (10.0 + 50.0 * uni(s)) * M_PI / 180.0, az = 2.0 * M_PI * uni(s);
  V sun = {std::cos(elev) * std::cos(az), std::sin(elev), std::cos(elev) * std::sin(az)};
  double sigma = M_PI / 180.0;
  B.env.assign((size_t)w * h * 3, 0.0f);
  B.env_w = w;
  B.env_h = h;
  for (uint32_t iy = 0; iy < h; iy++) {
    double theta = M_PI * (iy + 0.5) / h;
    for (uint32_t ix = 0; ix < w; ix++) {
      double phi = 2.0 * M_PI * (ix + 0.5) / w;
      // inverse of IBL::Sample's mapping (src/renderer/IBL.h:30-45)
      V d = {std::sin(theta) * std::cos(phi), std::cos(theta), std::sin(theta) * std::sin(phi)};
      double lum = d.y >= 0 ? 0.2 + 0.8 * d.y : 0.2 * (1.0 + d.y) + 0.02;
      double g = std::acos(std::fmax(-1.0, std::fmin(1.0, dotv(d, sun))));
      double sl = 5.0e4 * std::exp(-g * g / (2.0 * sigma * sigma));
      float *p = &B.env[3 * ((size_t)iy * w + ix)];
      p[0] = (float)(lum * 0.75 + sl);
      p[1] = (float)(lum * 0.85 + sl * 0.95);
      p[2] = (float)(lum * 1.00 + sl * 0.85);
    }
  }
}

double terrain_h(double x, double z, uint64_t seed) { return 40.0 * fbm({x / 180.0, 0.37, z / 180.0}, 6, seed); }

void terrain(Builder &B, uint32_t nx, uint32_t nz, uint64_t seed) {
  uint32_t ground = B.material(OMR_MAT_LAMBERT, {0.45, 0.42, 0.35}, {0, 0, 0}, 0);
  std::vector<V> p((size_t)nx * nz);
  for (uint32_t z = 0; z < nz; z++)
    for (uint32_t x = 0; x < nx; x++) p[(size_t)z * nx + x] = {(double)x, terrain_h(x, z, seed), (double)z};
  B.tri.reserve(B.tri.size() + (size_t)(nx - 1) * (nz - 1) * 18);
  for (uint32_t z = 0; z + 1 < nz; z++)
    for (uint32_t x = 0; x + 1 < nx; x++) {
      V a = p[(size_t)z * nx + x], b = p[(size_t)z * nx + x + 1], c = p[(size_t)(z + 1) * nx + x + 1],
        d = p[(size_t)(z + 1) * nx + x];
      B.triangle(a, b, c, {0, 1, 0}, ground);
      B.triangle(a, c, d, {0, 1, 0}, ground);
    }
  uint64_t s = seed ^ 0x77ull;
  for (int i = 0; i < 200; i++) {
    double r = 0.5 + 2.5 * uni(s);
    double x = 20 + (nx - 40) * uni(s), z = 20 + (nz - 40) * uni(s);
    uint32_t m;
    if (i % 3 == 0) m = B.material(OMR_MAT_LAMBERT, {0.2 + 0.6 * uni(s), 0.2 + 0.6 * uni(s), 0.2 + 0.6 * uni(s)}, {0, 0, 0}, 0);
    else if (i % 3 == 1) m = B.material(OMR_MAT_SPECULAR, {0.95, 0.95, 0.95}, {0, 0, 0}, 0);
    else m = B.material(OMR_MAT_REFRACTION, {0.99, 0.99, 0.99}, {0, 0, 0}, 1.5);
    B.sphere({x, terrain_h(x, z, seed) + r, z}, r, m);
  }
}

void torus_knot(Builder &B, V center, double scale, uint32_t nu, uint32_t nv, uint64_t seed) {
  auto curve = [&](double t) {
    double p = 2, q = 3, r = std::cos(q * t) + 2.0;
    return V{r * std::cos(p * t), -std::sin(q * t), r * std::sin(p * t)};
  };
  std::vector<V> pos((size_t)nu * nv), cen(nu);
  for (uint32_t i = 0; i < nu; i++) {
    double t = 2.0 * M_PI * i / nu;
    V c = curve(t), c2 = curve(t + 1e-3);
    V T = normv(c2 - c);
    V N = normv(crossv(T, V{0, 1, 0}));
    if (dotv(N, N) < 0.5) N = V{1, 0, 0};
    V Bn = crossv(T, N);
    cen[i] = c;
    for (uint32_t j = 0; j < nv; j++) {
      double a = 2.0 * M_PI * j / nv;
      V dir = N * std::cos(a) + Bn * std::sin(a);
      V base = c + dir * 0.45;
      double disp = 0.08 * fbm(base * 3.0, 4, seed);
      pos[(size_t)i * nv + j] = center + (c + dir * (0.45 * (1.0 + disp))) * scale;
    }
  }
  uint32_t m0 = B.material(OMR_MAT_LAMBERT, {0.7, 0.5, 0.3}, {0, 0, 0}, 0);
  uint32_t m1 = B.material(OMR_MAT_REFRACTION, {0.99, 0.99, 0.99}, {0, 0, 0}, 1.5);
  for (uint32_t i = 0; i < nu; i++)
    for (uint32_t j = 0; j < nv; j++) {
      uint32_t i1 = (i + 1) % nu, j1 = (j + 1) % nv;
      V a = pos[(size_t)i * nv + j], b = pos[(size_t)i1 * nv + j], c = pos[(size_t)i1 * nv + j1],
        d = pos[(size_t)i * nv + j1];
      V out = (a + c) * 0.5 - (center + cen[i] * scale);
      uint32_t m = ((i / 50) % 2) ? m1 : m0;
      B.triangle(a, b, c, out, m);
      B.triangle(a, c, d, out, m);
    }
}

struct Owned {
  Builder b;
  std::vector<uint32_t> pmat;
};

}  // namespace

extern "C" void omr_camera_from_lookat(uint32_t width, uint32_t height, const double pos[3], const double dir[3],
                                       const double up[3], double screen_height, double dist_to_screen,
                                       double focal_distance, double aperture_radius, omr_camera *out) {
  // Camera::Camera (src/renderer/Camera.h:12-43), double precision
  V P = {pos[0], pos[1], pos[2]}, D = normv({dir[0], dir[1], dir[2]}), U = normv({up[0], up[1], up[2]});
  double sh = screen_height, sw = screen_height * (double)width / (double)height;
  V C = P + D * dist_to_screen;
  V X = normv(crossv(D, U)) * sw;
  V Y = normv(crossv(X, D)) * sh;
  std::memset(out, 0, sizeof(*out));
  out->width = width;
  out->height = height;
  double v[4][3] = {{P.x, P.y, P.z}, {C.x, C.y, C.z}, {X.x, X.y, X.z}, {Y.x, Y.y, Y.z}};
  for (int a = 0; a < 3; a++) {
    out->position[a] = (float)v[0][a];
    out->screen_center[a] = (float)v[1][a];
    out->screen_x[a] = (float)v[2][a];
    out->screen_y[a] = (float)v[3][a];
  }
  out->aperture_radius = (float)aperture_radius;
  out->focal_distance = (float)focal_distance;
}

namespace {
// Camera preset of each procedural kind: position, direction, screen height.
bool preset(int kind, double cp[3], double cd[3], double *sh) {
  *sh = 30;
  switch (kind) {
    case 0:
    case 1:
      cp[0] = 50; cp[1] = 50; cp[2] = 260;
      cd[0] = 0; cd[1] = 0; cd[2] = -1;
      return true;
    case 2:
      cp[0] = 500; cp[1] = 140; cp[2] = -120;
      cd[0] = 0; cd[1] = -0.32; cd[2] = 1;
      *sh = 24;
      return true;
    case 3:
      cp[0] = 0; cp[1] = 60; cp[2] = 130;
      cd[0] = 0; cd[1] = -0.28; cd[2] = -1;
      return true;
    case 4:
      cp[0] = 1000; cp[1] = 200; cp[2] = -200;
      cd[0] = 0; cd[1] = -0.3; cd[2] = 1;
      *sh = 24;
      return true;
    default:
      return false;
  }
}
}  // namespace

extern "C" int omr_procedural_camera(int kind, uint32_t width, uint32_t height, omr_camera *out) {
  double cp[3], cd[3], up[3] = {0, 1, 0}, sh;
  if (!out || width == 0 || height == 0 || !preset(kind, cp, cd, &sh)) return OMR_ERR_INVALID;
  omr_camera_from_lookat(width, height, cp, cd, up, sh, 40.0, 200.0, 0.0, out);
  return OMR_OK;
}

extern "C" int omr_procedural_scene(int kind, uint64_t seed, uint32_t width, uint32_t height, omr_procedural *out) {
  if (!out || width == 0 || height == 0) return OMR_ERR_INVALID;
  std::memset(out, 0, sizeof(*out));
  Owned *o = new Owned();
  Builder &B = o->b;
  switch (kind) {
    case 0:
    case 1:
      cornell(B, true);
      if (kind == 1) icosphere(B, {50, 62, 45}, 18.0, 6, seed, 0.08);
      break;
    case 2:
      terrain(B, 1000, 500, seed);
      sky(B, 2048, 1024, seed);
      break;
    case 3: {
      torus_knot(B, {0, 30, 0}, 10.0, 1000, 150, seed);
      uint32_t fl = B.material(OMR_MAT_LAMBERT, {0.6, 0.6, 0.6}, {0, 0, 0}, 0);
      B.quad({-200, 0, -200}, {-200, 0, 200}, {200, 0, 200}, {200, 0, -200}, {0, 1, 0}, fl);
      sky(B, 512, 256, seed + 11);
      break;
    }
    case 4:
      terrain(B, 2000, 1000, seed);
      sky(B, 2048, 1024, seed + 4);
      break;
    default:
      delete o;
      return OMR_ERR_INVALID;
  }
  o->pmat = B.pmat_tri;
  o->pmat.insert(o->pmat.end(), B.pmat_sph.begin(), B.pmat_sph.end());
  omr_scene_desc &d = out->desc;
  d.n_triangles = B.pmat_tri.size();
  d.triangles = B.tri.data();
  d.tri_normals = B.nrm.data();
  d.n_spheres = B.pmat_sph.size();
  d.spheres = B.sph.empty() ? nullptr : B.sph.data();
  d.prim_material = o->pmat.data();
  d.prim_material_count = nullptr;
  d.n_materials = B.mats.size();
  d.materials = B.mats.data();
  d.env_width = B.env_w;
  d.env_height = B.env_h;
  d.env_rgb = B.env.empty() ? nullptr : B.env.data();
  omr_procedural_camera(kind, width, height, &out->camera);
  out->internal = o;
  return OMR_OK;
}

extern "C" void omr_procedural_free(omr_procedural *p) {
  if (!p) return;
  delete (Owned *)p->internal;
  std::memset(p, 0, sizeof(*p));
}
