// Obstacle faces and broad-phase grids on the device (obstacles.hpp).
#include "obstacles.hpp"

#include <algorithm>
#include <cmath>

namespace fph {

namespace {

// Obstacle faces and their broad-phase grids (see ObsView).  Faces whose dropped normal component
// c is at least kTau are binned in 3D; those between kThin and kTau in a 2D grid of their dropped
// axis (the containment test ignores that coordinate, and a plane nearly parallel to it accepts
// crossings far along it); the rest are tested for every vertex.  Face and segment boxes are grown
// by the margin: a crossing point computed with an error of e from the plane is accepted at most
// e / c from the face along the plane, and e is a few ulps of the coordinates (at most ~5e-7 of
// the largest obstacle coordinate M), so two margins must exceed 5e-7 M / kTau = 2.5e-5 M; the
// margin is max(kMinMargin, kRelMargin M), eight times that.
constexpr float kMinMargin = 1.0e-2f;
constexpr float kRelMargin = 1.0e-4f;
constexpr float kTau = 0.02f;
constexpr float kThin = 1.0e-3f;

// Whether triangle (a, b, c) meets the box of centre m and half-sizes h: no separating axis among
// the box's three axes, the triangle's normal and the nine cross products of its edges with the box
// axes (the separating-axis theorem for a triangle and a box), in double precision.
bool triangle_meets_box(const double a[3], const double b[3], const double c[3], const double m[3], const double h[3]) {
  double v[3][3];
  for (int k = 0; k < 3; ++k) {
    v[0][k] = a[k] - m[k];
    v[1][k] = b[k] - m[k];
    v[2][k] = c[k] - m[k];
  }
  auto separated = [&](const double ax[3]) {  // projections of the triangle and of the box on ax
    double lo = INFINITY, hi = -INFINITY;
    for (int i = 0; i < 3; ++i) {
      const double p = v[i][0] * ax[0] + v[i][1] * ax[1] + v[i][2] * ax[2];
      lo = std::min(lo, p);
      hi = std::max(hi, p);
    }
    const double r = h[0] * std::fabs(ax[0]) + h[1] * std::fabs(ax[1]) + h[2] * std::fabs(ax[2]);
    return lo > r || hi < -r;
  };
  double e[3][3];
  for (int k = 0; k < 3; ++k) {
    e[0][k] = v[1][k] - v[0][k];
    e[1][k] = v[2][k] - v[1][k];
    e[2][k] = v[0][k] - v[2][k];
  }
  for (int k = 0; k < 3; ++k) {
    const double ax[3] = {k == 0 ? 1.0 : 0.0, k == 1 ? 1.0 : 0.0, k == 2 ? 1.0 : 0.0};
    if (separated(ax)) return false;
  }
  const double n[3] = {e[0][1] * e[1][2] - e[0][2] * e[1][1], e[0][2] * e[1][0] - e[0][0] * e[1][2],
                       e[0][0] * e[1][1] - e[0][1] * e[1][0]};
  if (separated(n)) return false;
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) {
      // e_i x unit_k
      const double ax[3] = {k == 0 ? 0.0 : (k == 1 ? -e[i][2] : e[i][1]), k == 1 ? 0.0 : (k == 0 ? e[i][2] : -e[i][0]),
                            k == 2 ? 0.0 : (k == 0 ? -e[i][1] : e[i][0])};
      if (separated(ax)) return false;
    }
  return true;
}

int build_grid(ObstacleSet& s, int g, const std::vector<fp::ObsFace>& faces, const std::vector<uint32_t>& ids,
               int collapse) {
  fp::GridView& G = s.view.grid[g];
  G = fp::GridView{};
  if (ids.empty()) return FP_OK;
  const float margin = s.view.margin;
  auto box = [&](const fp::ObsFace& q, int k, float& l, float& h) {
    l = std::min(q.v0[k], std::min(q.v1[k], q.v2[k])) - margin;
    h = std::max(q.v0[k], std::max(q.v1[k], q.v2[k])) + margin;
  };
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  double ext_sum = 0.0;
  int n_axes = 0;
  for (uint32_t f : ids)
    for (int k = 0; k < 3; ++k) {
      if (k == collapse) continue;
      float l, h;
      box(faces[f], k, l, h);
      lo[k] = std::min(lo[k], l);
      hi[k] = std::max(hi[k], h);
      ext_sum += (double)(h - l);
    }
  for (int k = 0; k < 3; ++k) n_axes += k != collapse;
  const double mean_ext = ext_sum / ((double)n_axes * (double)ids.size());
  for (int k = 0; k < 3; ++k) {
    if (k == collapse) {
      G.lo[k] = 0.0f;
      G.inv_cell[k] = 0.0f;
      G.dims[k] = 1;
      continue;
    }
    const double ext = std::max((double)(hi[k] - lo[k]), 1e-6);
    // cells of half the mean face extent: a segment of one step still falls in one or two cells,
    // each holding fewer faces than at the full extent (on a 320-face sphere under a draped cloth,
    // the search cost per step fell by a third; at a quarter it rose again)
    const double cell = std::max(mean_ext / 2.0, ext / 128.0);
    G.dims[k] = (int)std::min(128.0, std::max(1.0, std::ceil(ext / cell)));
    G.inv_cell[k] = (float)((double)G.dims[k] / ext);
    G.lo[k] = lo[k];
  }
  const uint64_t n_cells = (uint64_t)G.dims[0] * G.dims[1] * G.dims[2];
  // the kernel's mapping from a coordinate to a cell, with the same f32 operations
  auto range = [&](const fp::ObsFace& q, int k, int& c0, int& c1) {
    float l, h;
    box(q, k, l, h);
    const float fl = (l - G.lo[k]) * G.inv_cell[k];
    const float fh = (h - G.lo[k]) * G.inv_cell[k];
    c0 = fl < 0.0f ? 0 : std::min((int)fl, G.dims[k] - 1);
    c1 = fh >= (float)G.dims[k] ? G.dims[k] - 1 : std::max(0, (int)fh);
  };
  // In the 3D grid a face goes only to the cells its triangle meets once each cell is grown by the
  // margin (and by a millionth of a cell against the rounding of the kernel's f32 cell mapping): a
  // crossing point accepted within the margin of the triangle lies in such a cell, which the
  // segment's grown box then reaches.  A 2D grid keeps the box test (its collapsed axis is open).
  auto meets = [&](const fp::ObsFace& q, int x, int y, int z) {
    if (collapse >= 0) return true;
    const int idx[3] = {x, y, z};
    double m[3], h[3];
    for (int k = 0; k < 3; ++k) {
      const double w = 1.0 / (double)G.inv_cell[k];
      m[k] = (double)G.lo[k] + ((double)idx[k] + 0.5) * w;
      h[k] = 0.5 * w * (1.0 + 1e-6) + (double)margin;
    }
    const double a[3] = {q.v0[0], q.v0[1], q.v0[2]}, b[3] = {q.v1[0], q.v1[1], q.v1[2]},
                 c[3] = {q.v2[0], q.v2[1], q.v2[2]};
    return triangle_meets_box(a, b, c, m, h);
  };
  std::vector<uint32_t> start(n_cells + 1, 0);
  auto visit = [&](auto&& fn) {
    for (uint32_t f : ids) {  // ascending face ids -> ascending cell lists
      int r[3][2];
      for (int k = 0; k < 3; ++k) range(faces[f], k, r[k][0], r[k][1]);
      for (int z = r[2][0]; z <= r[2][1]; ++z)
        for (int y = r[1][0]; y <= r[1][1]; ++y)
          for (int x = r[0][0]; x <= r[0][1]; ++x)
            if (meets(faces[f], x, y, z)) fn(((uint64_t)z * G.dims[1] + y) * G.dims[0] + x, f);
    }
  };
  visit([&](uint64_t c, uint32_t) { start[c + 1]++; });
  for (uint64_t c = 0; c < n_cells; ++c) start[c + 1] += start[c];
  std::vector<uint32_t> fill(start.begin(), start.end() - 1), lists(start[n_cells]);
  visit([&](uint64_t c, uint32_t f) { lists[fill[c]++] = f; });
  HIP_TRY(s.cell_start[g].upload(start.data(), start.size()));
  HIP_TRY(s.cell_faces[g].upload(lists.data(), lists.size()));
  G.cell_start = s.cell_start[g].p;
  G.cell_faces = s.cell_faces[g].p;
  return FP_OK;
}

}  // namespace

int upload_obstacles(ObstacleSet& s, const std::vector<fp::HostObstacle>& obstacles) {
  std::vector<fp::ObsFace> faces;
  for (auto& o : obstacles) {
    s.n_vertices += o.vertices.size() / 3;
    for (uint64_t t = 0; t < o.n_faces(); ++t) {
      const float* q = &o.faces[9 * t];
      fp::ObsFace f{};
      for (int k = 0; k < 3; ++k) { f.v0[k] = q[k]; f.v1[k] = q[3 + k]; f.v2[k] = q[6 + k]; }
      fp::f3 a{q[0], q[1], q[2]}, b{q[3], q[4], q[5]}, c{q[6], q[7], q[8]};
      fp::f3 n = fp::h_normalize(fp::cross(fp::sub(b, a), fp::sub(c, a)));  // Face::normal (collidable_mesh.rs:41-43)
      f.n[0] = n.x; f.n[1] = n.y; f.n[2] = n.z;
      // collidable_mesh.rs:133-154: signed comparisons, x first, then y, else z
      f.axis = (n.x >= n.y && n.x >= n.z) ? 0 : ((n.y >= n.x && n.y >= n.z) ? 1 : 2);
      faces.push_back(f);
    }
  }
  s.n_faces = faces.size();
  HIP_TRY(s.faces.upload(faces.data(), faces.size()));
  fp::ObsView& O = s.view;
  O = fp::ObsView{};
  O.faces = s.faces.p;
  O.n_faces = (uint32_t)faces.size();
  float coord = 0.0f;  // the largest obstacle coordinate, for the margin
  for (const auto& f : faces)
    for (int k = 0; k < 3; ++k)
      coord = std::max(coord, std::max(std::fabs(f.v0[k]), std::max(std::fabs(f.v1[k]), std::fabs(f.v2[k]))));
  O.margin = std::max(kMinMargin, kRelMargin * (std::isfinite(coord) ? coord : 0.0f));
  for (int k = 0; k < 3; ++k) {  // no obstacles: an empty box
    s.hit_box.lo[k] = INFINITY;
    s.hit_box.hi[k] = -INFINITY;
  }
  if (faces.empty()) return FP_OK;
  for (int k = 0; k < 3; ++k) {  // until shown otherwise: everything
    s.hit_box.lo[k] = -INFINITY;
    s.hit_box.hi[k] = INFINITY;
  }
  if (faces.size() <= 8) return FP_OK;  // scan every face in order

  std::vector<uint32_t> always, ids[4];
  for (uint32_t f = 0; f < faces.size(); ++f) {
    const fp::ObsFace& q = faces[f];
    const float an = std::fabs(q.n[q.axis]);
    bool finite = true;
    for (int k = 0; k < 3; ++k)
      finite = finite && std::isfinite(q.v0[k]) && std::isfinite(q.v1[k]) && std::isfinite(q.v2[k]);
    if (!finite || !(an >= kThin)) always.push_back(f);
    else if (an >= kTau) ids[0].push_back(f);
    else ids[1 + q.axis].push_back(f);
  }
  O.use_grid = 1;
  s.n_always = always.size();
  for (int g = 0; g < 4; ++g) s.n_binned[g] = ids[g].size();
  HIP_TRY(s.always.upload(always.data(), always.size()));
  O.always = s.always.p;
  O.n_always = (uint32_t)always.size();
  for (int g = 0; g < 4; ++g)
    if (int e = build_grid(s, g, faces, ids[g], g == 0 ? -1 : g - 1)) return e;
  if (always.empty()) {
    // the union of the binned faces' boxes, grown twice by the margin (once as the grids grow them,
    // once more against the rounding of the grids' cell mapping); a thin face's dropped axis unbounded
    for (int k = 0; k < 3; ++k) {
      s.hit_box.lo[k] = INFINITY;
      s.hit_box.hi[k] = -INFINITY;
    }
    for (int g = 0; g < 4; ++g)
      for (uint32_t f : ids[g]) {
        const fp::ObsFace& q = faces[f];
        for (int k = 0; k < 3; ++k) {
          const bool open = g > 0 && k == g - 1;
          const float l = open ? -INFINITY : std::min(q.v0[k], std::min(q.v1[k], q.v2[k])) - 2.0f * O.margin;
          const float h = open ? INFINITY : std::max(q.v0[k], std::max(q.v1[k], q.v2[k])) + 2.0f * O.margin;
          s.hit_box.lo[k] = std::min(s.hit_box.lo[k], l);
          s.hit_box.hi[k] = std::max(s.hit_box.hi[k], h);
        }
      }
  }
  return FP_OK;
}

}  // namespace fph
