// Seeded scenario generators (include/ftl_scenario.h).  Host only.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#include "ftl.h"
#include "ftl_internal.h"
#include "ftl_scenario.h"

namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1p-53; }  // [0, 1)
  double normal() {
    const double u1 = 1.0 - uniform(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

struct D3 {
  double x, y, z;
};
D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
D3 operator*(double s, D3 a) { return {s * a.x, s * a.y, s * a.z}; }
double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
D3 normalize(D3 a) {
  const double n = std::sqrt(dot(a, a));
  return n > 0 ? (1.0 / n) * a : D3{0, 1, 0};
}

uint64_t spread10(uint64_t v) {  // 10 bits -> every third bit
  v &= 0x3ff;
  v = (v | (v << 16)) & 0x030000ff;
  v = (v | (v << 8)) & 0x0300f00f;
  v = (v | (v << 4)) & 0x030c30c3;
  v = (v | (v << 2)) & 0x09249249;
  return v;
}

uint64_t morton(D3 p, D3 lo, double ext) {
  auto q = [&](double v, double l) {
    double t = (v - l) / ext;
    t = std::min(std::max(t, 0.0), 1.0);
    return (uint64_t)std::min(1023.0, std::floor(t * 1024.0));
  };
  return spread10(q(p.x, lo.x)) | (spread10(q(p.y, lo.y)) << 1) | (spread10(q(p.z, lo.z)) << 2);
}

// Writes one straight strand: root + i * seg * dir, w = seg (0 at the root).
void write_strand(float* pos, float* vel, int npoints, D3 root, D3 dir, double length) {
  const double seg = length / (npoints - 1);
  const float segf = (float)seg;
  for (int i = 0; i < npoints; ++i) {
    const D3 p = root + (i * seg) * dir;
    pos[4 * i + 0] = (float)p.x;
    pos[4 * i + 1] = (float)p.y;
    pos[4 * i + 2] = (float)p.z;
    pos[4 * i + 3] = i == 0 ? 0.0f : segf;
    vel[4 * i + 0] = vel[4 * i + 1] = vel[4 * i + 2] = vel[4 * i + 3] = 0.0f;
  }
}

}  // namespace

extern "C" ftl_status ftl_scn_sphere_groom(const ftl_scn_sphere* d, float* pos, float* vel, float* rest_roots) {
  if (!d || !pos || !vel) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  if (d->nstrands <= 0 || d->npoints < 2 || d->npoints > FTL_MAX_POINTS || !(d->radius > 0) || !(d->length > 0))
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "bad sphere groom description");
  const int S = d->nstrands, N = d->npoints;
  const D3 c{d->center[0], d->center[1], d->center[2]};
  std::vector<D3> roots(S), dirs(S);
  Rng rng(d->seed);
  for (int s = 0; s < S; ++s) {
    const double z = 1.0 - 2.0 * rng.uniform();
    const double rxy = std::sqrt(std::max(0.0, 1.0 - z * z));
    const double phi = 6.283185307179586 * rng.uniform();
    const D3 n{rxy * std::cos(phi), rxy * std::sin(phi), z};
    const D3 j{rng.normal(), rng.normal(), rng.normal()};
    roots[s] = c + (double)d->radius * n;
    dirs[s] = normalize(n + (double)d->jitter * j);
  }
  std::vector<int> order(S);
  std::iota(order.begin(), order.end(), 0);
  if (d->coherent) {
    const double r = d->radius;
    const D3 lo{c.x - r, c.y - r, c.z - r};
    std::vector<uint64_t> key(S);
    for (int s = 0; s < S; ++s) key[s] = morton(roots[s], lo, 2 * r);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
  }
  for (int s = 0; s < S; ++s) {
    const int o = order[s];
    write_strand(pos + (size_t)4 * N * s, vel + (size_t)4 * N * s, N, roots[o], dirs[o], d->length);
    if (rest_roots) {
      rest_roots[4 * (size_t)s + 0] = (float)roots[o].x;
      rest_roots[4 * (size_t)s + 1] = (float)roots[o].y;
      rest_roots[4 * (size_t)s + 2] = (float)roots[o].z;
      rest_roots[4 * (size_t)s + 3] = 0.0f;
    }
  }
  return FTL_OK;
}

extern "C" ftl_status ftl_scn_reference_groom(int nroots, const float* roots, const float* normals, int npoints,
                                              float maxlength, uint64_t seed, float* pos, float* vel) {
  if (nroots <= 0 || !roots || !normals || !pos || !vel || npoints < 2 || npoints > FTL_MAX_POINTS)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "bad reference groom arguments");
  Rng rng(seed);
  // float arithmetic as in hair.cc:258-288
  const float scaleOffset = maxlength / static_cast<float>(npoints);
  for (int j = 0; j < nroots; ++j) {
    const float* v = roots + 3 * (size_t)j;
    const float* n = normals + 3 * (size_t)j;
    const double u = rng.uniform();
    const float random_value = static_cast<float>(1.0 + 0.1 * (1.0 - 2.0 * u));
    float lastOffset = 0.0f;
    for (int i = 0; i < npoints; ++i) {
      const size_t idx = (size_t)j * npoints + i;
      const float offset = static_cast<float>(i) * scaleOffset * random_value;
      for (int c = 0; c < 3; ++c) pos[4 * idx + c] = v[c] + offset * n[c];
      pos[4 * idx + 3] = offset - lastOffset;
      vel[4 * idx + 0] = vel[4 * idx + 1] = vel[4 * idx + 2] = vel[4 * idx + 3] = 0.0f;
      lastOffset = offset;
    }
  }
  return FTL_OK;
}

// --- skinned emitter -------------------------------------------------------

static void grid_dims(int ntris, int* nu, int* nv) {
  // 2 * nu * nv triangles on a (nu+1) x (nv+1) latitude/longitude grid.
  int k = (int)std::lround(std::sqrt(std::max(2, ntris) / 2.0));
  k = std::max(k, 2);
  *nu = k;
  *nv = k;
}

extern "C" ftl_status ftl_scn_skinned_sizes(const ftl_scn_skinned* d, int* nverts, int* ntris) {
  if (!d || !nverts || !ntris) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  int nu, nv;
  grid_dims(d->ntris, &nu, &nv);
  *nverts = (nu + 1) * (nv + 1);
  *ntris = 2 * nu * nv;
  return FTL_OK;
}

static std::vector<D3> bone_pivots(const ftl_scn_skinned* d) {
  Rng rng(d->seed ^ 0xb0b0b0b0ull);
  std::vector<D3> piv(d->nbones);
  for (int b = 0; b < d->nbones; ++b) {
    const double z = 1.0 - 2.0 * rng.uniform();
    const double rxy = std::sqrt(std::max(0.0, 1.0 - z * z));
    const double phi = 6.283185307179586 * rng.uniform();
    const double r = d->radius * (0.3 + 0.5 * rng.uniform());
    piv[b] = D3{r * rxy * std::cos(phi), r * rxy * std::sin(phi), r * z};
  }
  return piv;
}

extern "C" ftl_status ftl_scn_skinned_mesh(const ftl_scn_skinned* d, float* rest, int* joints, float* weights,
                                           int* tris) {
  if (!d || !rest || !joints || !weights || !tris) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  if (d->nbones < 4 || !(d->radius > 0)) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "need >= 4 bones and radius > 0");
  int nu, nv;
  grid_dims(d->ntris, &nu, &nv);
  const std::vector<D3> piv = bone_pivots(d);
  int vi = 0;
  for (int a = 0; a <= nv; ++a) {
    const double theta = 3.141592653589793 * a / nv;
    for (int b = 0; b <= nu; ++b, ++vi) {
      const double phi = 6.283185307179586 * b / nu;
      const D3 p{d->radius * std::sin(theta) * std::cos(phi), d->radius * std::cos(theta),
                 d->radius * std::sin(theta) * std::sin(phi)};
      rest[4 * vi + 0] = (float)p.x;
      rest[4 * vi + 1] = (float)p.y;
      rest[4 * vi + 2] = (float)p.z;
      rest[4 * vi + 3] = 1.0f;
      // 4 nearest pivots, inverse-square weights, largest first
      int best[4] = {-1, -1, -1, -1};
      double bd[4] = {1e300, 1e300, 1e300, 1e300};
      for (int k = 0; k < d->nbones; ++k) {
        const D3 q = p - piv[k];
        const double dd = dot(q, q);
        for (int s = 0; s < 4; ++s)
          if (dd < bd[s]) {
            for (int t = 3; t > s; --t) {
              bd[t] = bd[t - 1];
              best[t] = best[t - 1];
            }
            bd[s] = dd;
            best[s] = k;
            break;
          }
      }
      double w[4], sum = 0;
      for (int s = 0; s < 4; ++s) sum += (w[s] = 1.0 / (bd[s] + 1e-4));
      for (int s = 0; s < 4; ++s) {
        joints[4 * vi + s] = best[s];
        weights[4 * vi + s] = (float)(w[s] / sum);
      }
    }
  }
  int ti = 0;
  for (int a = 0; a < nv; ++a)
    for (int b = 0; b < nu; ++b) {
      const int v00 = a * (nu + 1) + b, v01 = v00 + 1, v10 = v00 + (nu + 1), v11 = v10 + 1;
      tris[3 * ti + 0] = v00;
      tris[3 * ti + 1] = v10;
      tris[3 * ti + 2] = v01;
      ++ti;
      tris[3 * ti + 0] = v01;
      tris[3 * ti + 1] = v10;
      tris[3 * ti + 2] = v11;
      ++ti;
    }
  return FTL_OK;
}

extern "C" ftl_status ftl_scn_skinned_groom(const ftl_scn_skinned* d, const float* rest, const int* tris, int* tri,
                                            float* bary, float* pos, float* vel) {
  if (!d || !rest || !tris || !tri || !bary || !pos || !vel) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  if (d->nstrands <= 0 || d->npoints < 2 || d->npoints > FTL_MAX_POINTS)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "bad skinned groom description");
  int nverts, ntris;
  ftl_scn_skinned_sizes(d, &nverts, &ntris);
  auto V = [&](int i) { return D3{rest[4 * i], rest[4 * i + 1], rest[4 * i + 2]}; };
  std::vector<double> cdf(ntris);
  double acc = 0;
  for (int t = 0; t < ntris; ++t) {
    const D3 a = V(tris[3 * t]), b = V(tris[3 * t + 1]), c = V(tris[3 * t + 2]);
    acc += 0.5 * std::sqrt(dot(cross(b - a, c - a), cross(b - a, c - a)));
    cdf[t] = acc;
  }
  Rng rng(d->seed ^ 0x5eedull);
  const int N = d->npoints;
  for (int s = 0; s < d->nstrands; ++s) {
    const double u = rng.uniform() * acc;
    int t = (int)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
    t = std::min(t, ntris - 1);
    double r1 = rng.uniform(), r2 = rng.uniform();
    if (r1 + r2 > 1.0) {
      r1 = 1.0 - r1;
      r2 = 1.0 - r2;
    }
    const float w1 = (float)r1, w2 = (float)r2;
    tri[s] = t;
    bary[2 * s] = w1;
    bary[2 * s + 1] = w2;
    const D3 a = V(tris[3 * t]), b = V(tris[3 * t + 1]), c = V(tris[3 * t + 2]);
    const double w0 = 1.0 - (double)w1 - (double)w2;
    const D3 root = w0 * a + (double)w1 * b + (double)w2 * c;
    const D3 n = normalize(root);
    const D3 j{rng.normal(), rng.normal(), rng.normal()};
    write_strand(pos + (size_t)4 * N * s, vel + (size_t)4 * N * s, N, root, normalize(n + (double)d->jitter * j),
                 d->length);
  }
  return FTL_OK;
}

extern "C" ftl_status ftl_scn_bones(const ftl_scn_skinned* d, double t, float* dq) {
  if (!d || !dq) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  const std::vector<D3> piv = bone_pivots(d);
  Rng rng(d->seed ^ 0xa11e5ull);
  for (int b = 0; b < d->nbones; ++b) {
    const double z = 1.0 - 2.0 * rng.uniform();
    const double rxy = std::sqrt(std::max(0.0, 1.0 - z * z));
    const double phi = 6.283185307179586 * rng.uniform();
    const D3 axis{rxy * std::cos(phi), rxy * std::sin(phi), z};
    const double w = 0.5 + 1.5 * rng.uniform();
    const double ph = 6.283185307179586 * rng.uniform();
    const double ang = d->max_angle * std::sin(w * t + ph);
    const double sh = std::sin(0.5 * ang), ch = std::cos(0.5 * ang);
    const D3 qv = sh * axis;
    const double qw = ch;
    // rotation about the pivot: x' = R x + (c - R c)
    const D3 c = piv[b];
    const D3 tcv = cross(qv, c);
    const D3 rc = c + 2.0 * cross(qv, tcv + qw * c);
    const D3 tr = c - rc;
    // dual = 0.5 * (tr, 0) * q
    const D3 dv = 0.5 * (qw * tr + cross(tr, qv));
    const double dw = -0.5 * dot(tr, qv);
    float* o = dq + 8 * (size_t)b;
    o[0] = (float)qv.x; o[1] = (float)qv.y; o[2] = (float)qv.z; o[3] = (float)qw;
    o[4] = (float)dv.x; o[5] = (float)dv.y; o[6] = (float)dv.z; o[7] = (float)dw;
  }
  return FTL_OK;
}

extern "C" ftl_status ftl_scn_strand_tangents(const float* pos, int nstrands, int npoints, float* tangents) {
  if (!pos || !tangents || nstrands <= 0 || npoints < 2)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "bad strand tangent arguments");
  for (int s = 0; s < nstrands; ++s) {
    const float* p = pos + (size_t)4 * npoints * s;
    float* t = tangents + (size_t)4 * npoints * s;
    for (int i = 0; i < npoints; ++i) {
      const int a = i > 0 ? i - 1 : 0, b = i < npoints - 1 ? i + 1 : npoints - 1;
      const float k = (b - a == 2) ? 0.5f : 1.0f;
      for (int c = 0; c < 3; ++c) t[4 * i + c] = (p[4 * b + c] - p[4 * a + c]) * k;
      t[4 * i + 3] = 0.0f;
    }
  }
  return FTL_OK;
}

extern "C" ftl_status ftl_scn_random_table(uint64_t seed, int n, float* out) {
  if (!out || n <= 0) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "bad random table arguments");
  Rng rng(seed);
  for (int i = 0; i < n; ++i) {
    for (int c = 0; c < 3; ++c) out[4 * (size_t)i + c] = (float)(rng.next() >> 40) * 0x1p-24f;  // [0, 1)
    out[4 * (size_t)i + 3] = 0.0f;
  }
  return FTL_OK;
}

