// Synthetic, seed-reproducible scenes for tests and benchmarks.
//
// The reference loads meshes from disk (lib/io/Load.cpp:122-139); these
// generators produce the same Mesh data (vec3 vertices + u32[3] faces)
// procedurally so every run builds from identical bytes.  All arithmetic is
// done in f64 and rounded once to f32, with counter-based hashing
// (splitmix64) so the output does not depend on iteration order or threads.
#include "types.hpp"

#include <cmath>
#include <cstring>
#include <map>
#include <utility>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Counter-based uniform double in [0, 1) from (seed, stream, index).
inline double hash01(uint64_t seed, uint64_t stream, uint64_t index) {
  uint64_t h = splitmix64(seed * 0x100000001B3ull ^ splitmix64(stream * 0x9E3779B97F4A7C15ull + index));
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

inline double smooth(double t) { return t * t * (3.0 - 2.0 * t); }

// 4-octave value noise in [0, 1]: lattice frequency 4 doubling per octave,
// amplitude halving per octave, smoothstep interpolation of hashed lattice values.
double valueNoise(uint64_t seed, double x, double z) {
  double sum = 0.0, norm = 0.0, amp = 1.0, freq = 4.0;
  for (int o = 0; o < 4; ++o) {
    double fx = x * freq, fz = z * freq;
    double ix = std::floor(fx), iz = std::floor(fz);
    double tx = smooth(fx - ix), tz = smooth(fz - iz);
    auto lat = [&](double a, double b) {
      uint64_t ia = (uint64_t)(int64_t)a & 0xFFFFFFFFull, ib = (uint64_t)(int64_t)b & 0xFFFFFFFFull;
      return hash01(seed, 1000 + o, (ia << 32) | ib);
    };
    double v00 = lat(ix, iz), v10 = lat(ix + 1, iz), v01 = lat(ix, iz + 1), v11 = lat(ix + 1, iz + 1);
    double v0 = v00 + (v10 - v00) * tx;
    double v1 = v01 + (v11 - v01) * tx;
    sum += amp * (v0 + (v1 - v0) * tz);
    norm += amp;
    amp *= 0.5;
    freq *= 2.0;
  }
  return sum / norm;
}

struct DVec {
  double x, y, z;
};

// Unit icosphere with `subdiv` midpoint subdivisions, faces wound
// counter-clockwise seen from outside (outward geometric normals).
struct Icosphere {
  std::vector<DVec> v;
  std::vector<bvht_Triangle> f;
};

Icosphere makeIcosphere(int subdiv) {
  Icosphere s;
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  const double base[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                              {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
  for (auto &b : base) {
    double l = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    s.v.push_back({b[0] / l, b[1] / l, b[2] / l});
  }
  const uint32_t ico[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                               {11, 10, 2}, {10, 7, 6}, {7, 1, 8},   {3, 9, 4},  {3, 4, 2},   {3, 2, 6}, {3, 6, 8},
                               {3, 8, 9},  {4, 9, 5},  {2, 4, 11},  {6, 2, 10}, {8, 6, 7},   {9, 8, 1}};
  for (auto &f : ico) s.f.push_back({f[0], f[1], f[2]});
  for (int it = 0; it < subdiv; ++it) {
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> mid;
    auto midpoint = [&](uint32_t a, uint32_t b) {
      auto key = a < b ? std::make_pair(a, b) : std::make_pair(b, a);
      auto itm = mid.find(key);
      if (itm != mid.end()) return itm->second;
      DVec p = {(s.v[a].x + s.v[b].x) * 0.5, (s.v[a].y + s.v[b].y) * 0.5, (s.v[a].z + s.v[b].z) * 0.5};
      double l = std::sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
      s.v.push_back({p.x / l, p.y / l, p.z / l});
      uint32_t id = (uint32_t)s.v.size() - 1;
      mid.emplace(key, id);
      return id;
    };
    std::vector<bvht_Triangle> nf;
    nf.reserve(s.f.size() * 4);
    for (auto &f : s.f) {
      uint32_t a = midpoint(f.v1, f.v2), b = midpoint(f.v2, f.v3), c = midpoint(f.v3, f.v1);
      nf.push_back({f.v1, a, c});
      nf.push_back({f.v2, b, a});
      nf.push_back({f.v3, c, b});
      nf.push_back({a, b, c});
    }
    s.f.swap(nf);
  }
  return s;
}

inline bvht_vec3 tof(double x, double y, double z) { return {(float)x, (float)y, (float)z}; }

}  // namespace

extern "C" {

void bvht_sceneHeightfield(uint32_t gx, uint32_t gz, uint32_t seed, float amplitude, bvht_vec3 *vert,
                           bvht_Triangle *faces, uint32_t *numVert, uint32_t *numFaces) {
  // Grid spacing 1/gx on both axes, so a gx x gz grid spans [0,1] x [0,gz/gx].
  uint32_t nv = (gx + 1) * (gz + 1), nf = 2 * gx * gz;
  if (numVert) *numVert = nv;
  if (numFaces) *numFaces = nf;
  if (!vert || !faces) return;
  const double h = 1.0 / (double)gx;
  for (uint32_t j = 0; j <= gz; ++j)
    for (uint32_t i = 0; i <= gx; ++i) {
      double x = i * h, z = j * h;
      vert[j * (gx + 1) + i] = tof(x, (double)amplitude * valueNoise(seed, x, z), z);
    }
  uint32_t k = 0;
  for (uint32_t j = 0; j < gz; ++j)
    for (uint32_t i = 0; i < gx; ++i) {
      uint32_t v00 = j * (gx + 1) + i, v10 = v00 + 1, v01 = v00 + (gx + 1), v11 = v01 + 1;
      // wound so that (b - a) x (c - a) points to +y
      faces[k++] = {v00, v01, v10};
      faces[k++] = {v10, v01, v11};
    }
}

void bvht_sceneSoup(uint32_t numTris, uint32_t seed, float size, bvht_vec3 *vert, bvht_Triangle *faces,
                    uint32_t *numVert, uint32_t *numFaces) {
  if (numVert) *numVert = 3 * numTris;
  if (numFaces) *numFaces = numTris;
  if (!vert || !faces) return;
  for (uint32_t t = 0; t < numTris; ++t) {
    double cx = hash01(seed, 1, 4 * (uint64_t)t), cy = hash01(seed, 1, 4 * (uint64_t)t + 1),
           cz = hash01(seed, 1, 4 * (uint64_t)t + 2);
    for (uint32_t k = 0; k < 3; ++k) {
      uint64_t idx = 9 * (uint64_t)t + 3 * k;
      double dx = (2.0 * hash01(seed, 2, idx) - 1.0) * size;
      double dy = (2.0 * hash01(seed, 2, idx + 1) - 1.0) * size;
      double dz = (2.0 * hash01(seed, 2, idx + 2) - 1.0) * size;
      vert[3 * t + k] = tof(cx + dx, cy + dy, cz + dz);
    }
    faces[t] = {3 * t, 3 * t + 1, 3 * t + 2};
  }
}

void bvht_sceneSpheresSoup(uint32_t numSpheres, uint32_t numSoup, uint32_t seed, bvht_vec3 *vert,
                           bvht_Triangle *faces, uint32_t *numVert, uint32_t *numFaces) {
  // Subdivision level s in {2,3,4,5} with weights {0.60, 0.25, 0.12, 0.03};
  // radius log-uniform in [1e-3, 5e-2]; centre uniform in the unit cube.
  static const Icosphere unit[4] = {makeIcosphere(2), makeIcosphere(3), makeIcosphere(4), makeIcosphere(5)};
  std::vector<int> level(numSpheres);
  uint64_t nv = 0, nf = 0;
  for (uint32_t s = 0; s < numSpheres; ++s) {
    double u = hash01(seed, 10, s);
    int l = u < 0.60 ? 0 : (u < 0.85 ? 1 : (u < 0.97 ? 2 : 3));
    level[s] = l;
    nv += unit[l].v.size();
    nf += unit[l].f.size();
  }
  if (numVert) *numVert = (uint32_t)(nv + 3ull * numSoup);
  if (numFaces) *numFaces = (uint32_t)(nf + numSoup);
  if (!vert || !faces) return;
  uint32_t vo = 0, fo = 0;
  const double lr0 = std::log(1e-3), lr1 = std::log(5e-2);
  for (uint32_t s = 0; s < numSpheres; ++s) {
    const Icosphere &ico = unit[level[s]];
    double r = std::exp(lr0 + (lr1 - lr0) * hash01(seed, 11, s));
    double cx = hash01(seed, 12, 3 * (uint64_t)s), cy = hash01(seed, 12, 3 * (uint64_t)s + 1),
           cz = hash01(seed, 12, 3 * (uint64_t)s + 2);
    for (size_t i = 0; i < ico.v.size(); ++i) vert[vo + i] = tof(cx + r * ico.v[i].x, cy + r * ico.v[i].y, cz + r * ico.v[i].z);
    for (size_t i = 0; i < ico.f.size(); ++i) faces[fo + i] = {ico.f[i].v1 + vo, ico.f[i].v2 + vo, ico.f[i].v3 + vo};
    vo += (uint32_t)ico.v.size();
    fo += (uint32_t)ico.f.size();
  }
  bvht_sceneSoup(numSoup, seed * 7919u + 17u, 0.01f, vert + vo, faces + fo, nullptr, nullptr);
  for (uint32_t t = 0; t < numSoup; ++t) {
    faces[fo + t].v1 += vo;
    faces[fo + t].v2 += vo;
    faces[fo + t].v3 += vo;
  }
}

}  // extern "C"

extern "C" {

void bvht_sceneBatch(uint32_t numMeshes, uint32_t seed, uint32_t minTris, uint32_t maxTris, uint32_t *faceCounts,
                     uint32_t *vertCounts, bvht_vec3 *vert, bvht_Triangle *faces, uint64_t *totalVert,
                     uint64_t *totalFaces) {
  // mesh i: seed + i; triangle target exp(U(ln minTris, ln maxTris)); kind i % 3
  static const Icosphere ico2 = makeIcosphere(2);  // 320 faces, 162 vertices
  const double l0 = std::log((double)minTris), l1 = std::log((double)maxTris);
  uint64_t tv = 0, tf = 0;
  for (uint32_t i = 0; i < numMeshes; ++i) {
    const uint32_t s = seed + i;
    const double target = std::exp(l0 + (l1 - l0) * hash01(s, 40, 0));
    uint32_t nv = 0, nf = 0, gx = 1, gz = 1, spheres = 1;
    switch (i % 3) {
      case 0:  // heightfield with about `target` triangles (2 per cell)
        gz = (uint32_t)std::max(1.0, std::round(std::sqrt(target / 4.0)));
        gx = (uint32_t)std::max(1.0, std::round(target / (2.0 * gz)));
        nv = (gx + 1) * (gz + 1);
        nf = 2 * gx * gz;
        break;
      case 1:  // cluster of subdivision-2 icospheres
        spheres = (uint32_t)std::max(1.0, std::round(target / 320.0));
        nv = spheres * (uint32_t)ico2.v.size();
        nf = spheres * (uint32_t)ico2.f.size();
        break;
      default:  // soup
        nf = (uint32_t)std::max(1.0, std::round(target));
        nv = 3 * nf;
        break;
    }
    if (faceCounts) faceCounts[i] = nf;
    if (vertCounts) vertCounts[i] = nv;
    if (vert && faces) {
      bvht_vec3 *v = vert + tv;
      bvht_Triangle *f = faces + tf;
      if (i % 3 == 0) {
        bvht_sceneHeightfield(gx, gz, s, 0.2f, v, f, nullptr, nullptr);
      } else if (i % 3 == 1) {
        for (uint32_t k = 0; k < spheres; ++k) {
          const double r = 0.02 + 0.08 * hash01(s, 41, k);
          const double cx = hash01(s, 42, 3 * k), cy = hash01(s, 42, 3 * k + 1), cz = hash01(s, 42, 3 * k + 2);
          const uint32_t vo = k * (uint32_t)ico2.v.size(), fo = k * (uint32_t)ico2.f.size();
          for (size_t q = 0; q < ico2.v.size(); ++q)
            v[vo + q] = tof(cx + r * ico2.v[q].x, cy + r * ico2.v[q].y, cz + r * ico2.v[q].z);
          for (size_t q = 0; q < ico2.f.size(); ++q)
            f[fo + q] = {ico2.f[q].v1 + vo, ico2.f[q].v2 + vo, ico2.f[q].v3 + vo};
        }
      } else {
        bvht_sceneSoup(nf, s, 0.01f, v, f, nullptr, nullptr);
      }
    }
    tv += nv;
    tf += nf;
  }
  if (totalVert) *totalVert = tv;
  if (totalFaces) *totalFaces = tf;
}

}  // extern "C"
