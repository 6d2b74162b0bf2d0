// Host mesh setup.  Reference: Mesh.cpp:14-32, 244-528, 692-925, 1013-1167.
#include "mesh.hpp"

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace corefl {

namespace {

// Coordinate array view with G = ngg+1 ghost layers.
struct Coord {
  std::vector<double> &v;
  int mx, my, mz, G;
  double &operator()(int i, int j, int k) {
    const long p2 = mx + 2L * G, p1 = (my + 2L * G) * p2;
    return v[(k + G) * p1 + (j + G) * p2 + (i + G)];
  }
};

void mirror_ghosts(HostBlock &b) {
  const int mx = b.mx, my = b.my, mz = b.mz, g = b.ngg, G = g + 1;
  Coord c[3] = {{b.x, mx, my, mz, G}, {b.y, mx, my, mz, G}, {b.z, mx, my, mz, G}};
  for (int q = 0; q < 3; ++q) {
    Coord &B = c[q];
    // reflected through the boundary node (Mesh.cpp:1018-1079)
    for (int k = 0; k < mz; ++k)
      for (int j = 0; j < my; ++j)
        for (int s = 1; s <= G; ++s) {
          B(-s, j, k) = 2 * B(0, j, k) - B(s, j, k);
          B(mx - 1 + s, j, k) = 2 * B(mx - 1, j, k) - B(mx - 1 - s, j, k);
        }
    for (int k = 0; k < mz; ++k)
      for (int i = 0; i < mx; ++i)
        for (int s = 1; s <= G; ++s) {
          B(i, -s, k) = 2 * B(i, 0, k) - B(i, s, k);
          B(i, my - 1 + s, k) = 2 * B(i, my - 1, k) - B(i, my - 1 - s, k);
        }
    for (int j = 0; j < my; ++j)
      for (int i = 0; i < mx; ++i)
        for (int s = 1; s <= G; ++s) {
          if (b.dim == 3) {
            B(i, j, -s) = 2 * B(i, j, 0) - B(i, j, s);
            B(i, j, mz - 1 + s) = 2 * B(i, j, mz - 1) - B(i, j, mz - 1 - s);
          } else { // 2D: copies, z shifted by the layer index
            B(i, j, -s) = B(i, j, 0) - (q == 2 ? s : 0);
            B(i, j, mz - 1 + s) = B(i, j, mz - 1) + (q == 2 ? s : 0);
          }
        }
  }
  // edges and corners by the parallelogram rule (Mesh.cpp:1088-1160)
  for (int q = 0; q < 3; ++q) {
    Coord &B = c[q];
    for (int a = 1; a <= G; ++a)
      for (int bb = 1; bb <= G; ++bb)
        for (int k = 0; k < mz; ++k) {
          const int i1 = -a, j1 = -bb, i2 = mx - 1 + a, j2 = my - 1 + bb;
          B(i1, j1, k) = B(i1, 0, k) + B(0, j1, k) - B(0, 0, k);
          B(i2, j2, k) = B(i2, my - 1, k) + B(mx - 1, j2, k) - B(mx - 1, my - 1, k);
          B(i1, j2, k) = B(i1, my - 1, k) + B(0, j2, k) - B(0, my - 1, k);
          B(i2, j1, k) = B(i2, 0, k) + B(mx - 1, j1, k) - B(mx - 1, 0, k);
        }
    for (int a = 1; a <= G; ++a)
      for (int j = 0; j < my; ++j)
        for (int cc = 1; cc <= G; ++cc) {
          const int i1 = -a, k1 = -cc, i2 = mx - 1 + a, k2 = mz - 1 + cc;
          B(i1, j, k1) = B(i1, j, 0) + B(0, j, k1) - B(0, j, 0);
          B(i2, j, k2) = B(i2, j, mz - 1) + B(mx - 1, j, k2) - B(mx - 1, j, mz - 1);
          B(i1, j, k2) = B(i1, j, mz - 1) + B(0, j, k2) - B(0, j, mz - 1);
          B(i2, j, k1) = B(i2, j, 0) + B(mx - 1, j, k1) - B(mx - 1, j, 0);
        }
    for (int i = -G; i <= mx - 1 + G; ++i)
      for (int a = 1; a <= G; ++a)
        for (int cc = 1; cc <= G; ++cc) {
          const int j1 = -a, k1 = -cc, j2 = my - 1 + a, k2 = mz - 1 + cc;
          B(i, j1, k1) = B(i, j1, 0) + B(i, 0, k1) - B(i, 0, 0);
          B(i, j2, k2) = B(i, j2, mz - 1) + B(i, my - 1, k2) - B(i, my - 1, mz - 1);
          B(i, j1, k2) = B(i, j1, mz - 1) + B(i, 0, k2) - B(i, 0, mz - 1);
          B(i, j2, k1) = B(i, j2, 0) + B(i, my - 1, k1) - B(i, my - 1, 0);
        }
  }
}

} // namespace

HostBlock make_block(int mx, int my, int mz, int ngg, const double *x, const double *y, const double *z,
                     double grid_scale) {
  HostBlock b;
  b.mx = mx; b.my = my; b.mz = mz; b.ngg = ngg;
  b.dim = mz == 1 ? 2 : 3;
  const int G = ngg + 1;
  const size_t n = (size_t)(mx + 2 * G) * (my + 2 * G) * (mz + 2 * G);
  b.x.assign(n, 0.0);
  b.y.assign(n, 0.0);
  b.z.assign(n, 0.0);
  Coord cx{b.x, mx, my, mz, G}, cy{b.y, mx, my, mz, G}, cz{b.z, mx, my, mz, G};
  const bool scale = std::abs(grid_scale - 1.0) > 1e-15; // Mesh.cpp:657
  for (int k = 0; k < mz; ++k)
    for (int j = 0; j < my; ++j)
      for (int i = 0; i < mx; ++i) {
        const size_t s = ((size_t)k * my + j) * mx + i;
        cx(i, j, k) = scale ? x[s] * grid_scale : x[s];
        cy(i, j, k) = scale ? y[s] * grid_scale : y[s];
        cz(i, j, k) = scale ? z[s] * grid_scale : z[s];
      }
  return b;
}

std::vector<HostBlock> read_grid_binary(const std::string &path, int ngg, double grid_scale) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open grid file " + path);
  int32_t nb = 0;
  f.read(reinterpret_cast<char *>(&nb), 4);
  if (nb <= 0 || nb > 100000) throw std::runtime_error("bad block count in " + path);
  std::vector<int32_t> dims(3 * nb);
  f.read(reinterpret_cast<char *>(dims.data()), 12L * nb);
  std::vector<HostBlock> out;
  for (int b = 0; b < nb; ++b) {
    const int mx = dims[3 * b], my = dims[3 * b + 1], mz = dims[3 * b + 2];
    const size_t n = (size_t)mx * my * mz;
    std::vector<double> x(n), y(n), z(n);
    f.read(reinterpret_cast<char *>(x.data()), 8 * n);
    f.read(reinterpret_cast<char *>(y.data()), 8 * n);
    f.read(reinterpret_cast<char *>(z.data()), 8 * n);
    if (!f) throw std::runtime_error("grid file truncated: " + path);
    out.push_back(make_block(mx, my, mz, ngg, x.data(), y.data(), z.data(), grid_scale));
  }
  return out;
}

std::vector<HostBlock> read_grid_ascii(const std::string &path, int ngg, double grid_scale) {
  // grid%4d.grd (Mesh.cpp:739-788): the block count on the first line, one "mx my mz" line per block, then per
  // block all x, all y and all z values (i fastest) separated by white space
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open grid file " + path);
  std::string line;
  if (!std::getline(f, line)) throw std::runtime_error("empty grid file " + path);
  const int nb = std::stoi(line);
  if (nb <= 0 || nb > 100000) throw std::runtime_error("bad block count in " + path);
  std::vector<int> dims(3 * nb);
  for (int b = 0; b < nb; ++b) {
    if (!std::getline(f, line)) throw std::runtime_error("grid file truncated: " + path);
    std::istringstream ls(line);
    ls >> dims[3 * b] >> dims[3 * b + 1] >> dims[3 * b + 2];
    if (!ls || dims[3 * b] < 1 || dims[3 * b + 1] < 1 || dims[3 * b + 2] < 1)
      throw std::runtime_error("bad block extents in " + path);
  }
  std::vector<HostBlock> out;
  for (int b = 0; b < nb; ++b) {
    const size_t n = (size_t)dims[3 * b] * dims[3 * b + 1] * dims[3 * b + 2];
    std::vector<double> x(n), y(n), z(n);
    for (auto *a : {&x, &y, &z})
      for (size_t q = 0; q < n; ++q)
        if (!(f >> (*a)[q])) throw std::runtime_error("grid file truncated: " + path);
    out.push_back(make_block(dims[3 * b], dims[3 * b + 1], dims[3 * b + 2], ngg, x.data(), y.data(), z.data(),
                             grid_scale));
  }
  return out;
}

void parse_boundaries(const std::string &text, std::vector<HostBlock> &blocks, int periodic_label) {
  std::istringstream in(text);
  for (auto &b : blocks) {
    std::string line;
    do {
      if (!std::getline(in, line)) throw std::runtime_error("boundary file: missing block entry");
    } while (line.find_first_not_of(" \t\r") == std::string::npos);
    const int nf = std::stoi(line);
    for (int f = 0; f < nf; ++f) {
      if (!std::getline(in, line)) throw std::runtime_error("boundary file: truncated");
      std::istringstream ls(line);
      int r[7];
      for (int &v : r) ls >> v;
      Face F;
      F.start[0] = r[0]; F.end[0] = r[1];
      F.start[1] = r[2]; F.end[1] = r[3];
      F.start[2] = r[4]; F.end[2] = r[5];
      F.type = r[6];
      F.face = 0;
      F.direction = 1;
      // register_boundary with ngg = 0 (Mesh.cpp:14-32)
      for (int d = 0; d < b.dim; ++d) {
        if (F.start[d] == F.end[d]) {
          F.face = d;
          if (F.start[d] == 0) F.direction = -1;
          break;
        }
      }
      for (int d = 0; d < b.dim; ++d)
        if (d != F.face && F.end[d] < F.start[d]) std::swap(F.start[d], F.end[d]);
      if (b.dim == 2) { F.start[2] = 0; F.end[2] = 0; }
      // one-sided viscous closures on every non-periodic face (Mesh.cpp:857-916)
      if (F.type != periodic_label) b.btype[2 * F.face + (F.direction > 0 ? 1 : 0)] = F.type;
      b.faces.push_back(F);
    }
  }
}

void mirror_ghost_grid(HostBlock &b) { mirror_ghosts(b); }

namespace {
int sgn(int v) { return (v > 0) - (v < 0); }

// InnerFace::register_boundary (Mesh.cpp:39-173): s, t = (x1, x2, y1, y2, z1, z2) of this block and of the
// target, 1-based; the normal direction has equal ends, the first in-face direction is marked negative in 3D.
InnerFace register_inner(const int *s, const int *t, int target_id, int dim, int ngg) {
  InnerFace f{};
  int rs[3] = {s[0], s[2], s[4]}, re[3] = {s[1], s[3], s[5]};
  int ts[3] = {t[0], t[2], t[4]}, te[3] = {t[1], t[3], t[5]};
  f.target_block = target_id - 1;
  f.face = -1;
  f.direction = 1;
  for (int d = 0; d < dim; ++d)
    if (rs[d] == re[d]) {
      f.face = d;
      if (rs[d] == 1) f.direction = -1;
      --rs[d];
      --re[d];
      break;
    }
  f.target_face = -1;
  f.target_direction = 1;
  for (int d = 0; d < dim; ++d)
    if (ts[d] == te[d]) {
      f.target_face = d;
      if (ts[d] == 1) f.target_direction = -1;
      --ts[d];
      --te[d];
      break;
    }
  if (f.face < 0 || f.target_face < 0) throw std::runtime_error("inner face without a constant index direction");
  int st[3] = {0, 0, 0}; // signed, 1-based source direction of each target direction
  st[f.target_face] = -(f.direction * f.target_direction) * (f.face + 1);
  auto pair = [&](int a, int ta) {
    st[ta] = sgn((te[ta] - ts[ta]) * (re[a] - rs[a])) * (a + 1);
    if (st[ta] == 0) throw std::runtime_error("inner face with a single-node in-face range");
  };
  if (dim == 2) {
    const int a = 1 - f.face, ta = 1 - f.target_face;
    rs[a] = std::abs(rs[a]) - 1;
    re[a] = std::abs(re[a]) - 1;
    ts[ta] = std::abs(ts[ta]) - 1;
    te[ta] = std::abs(te[ta]) - 1;
    pair(a, ta);
    st[2] = 3;
    rs[2] = re[2] = ts[2] = te[2] = 0;
  } else {
    int a = -1, ta = -1;
    for (int d = 0; d < 3 && a < 0; ++d)
      if (rs[d] < 0) a = d;
    for (int d = 0; d < 3 && ta < 0; ++d)
      if (ts[d] < 0) ta = d;
    if (a < 0 || ta < 0 || a == f.face || ta == f.target_face)
      throw std::runtime_error("inner face: the first in-face direction must be marked negative on both sides");
    rs[a] = -rs[a] - 1;
    re[a] = -re[a] - 1;
    ts[ta] = -ts[ta] - 1;
    te[ta] = -te[ta] - 1;
    pair(a, ta);
    const int b = 3 - f.face - a, tb = 3 - f.target_face - ta;
    --rs[b];
    --re[b];
    --ts[tb];
    --te[tb];
    pair(b, tb);
  }
  // the normal range becomes the coincident layer plus the ngg ghost layers
  f.face_layer = f.direction > 0 ? rs[f.face] : re[f.face];
  if (f.direction < 0) {
    rs[f.face] = -ngg;
    re[f.face] = 0;
  } else {
    rs[f.face] = re[f.face];
    re[f.face] = re[f.face] + ngg;
  }
  const int sg = sgn(st[f.target_face]);
  if (f.target_direction > 0) ts[f.target_face] = sg > 0 ? te[f.target_face] - ngg : te[f.target_face];
  else ts[f.target_face] = sg > 0 ? 0 : ngg;
  for (int d = 0; d < 3; ++d) {
    f.start[d] = rs[d];
    f.end[d] = re[d];
    f.target_start[d] = ts[d];
    f.loop_dir[d] = rs[d] <= re[d] ? 1 : -1;
    f.target_loop_dir[d] = sgn(st[d]);
    f.src_tar[d] = std::abs(st[d]) - 1;
    f.n_point[d] = std::abs(rs[d] - re[d]) + 1;
  }
  return f;
}
} // namespace

void parse_inner_faces(const std::string &text, std::vector<HostBlock> &blocks) {
  std::istringstream in(text);
  std::string line;
  if (!std::getline(in, line)) return; // total count (unused)
  std::vector<int> n_face(blocks.size());
  for (auto &n : n_face)
    if (!(in >> n) || n < 0) throw std::runtime_error("inner face file: bad face count");
  std::getline(in, line); // rest of the count line
  std::getline(in, line); // column header
  for (size_t b = 0; b < blocks.size(); ++b)
    for (int l = 0; l < n_face[b]; ++l) {
      int s[7], t[7];
      for (int *r : {s, t}) {
        if (!std::getline(in, line)) throw std::runtime_error("inner face file: truncated");
        std::istringstream ls(line);
        for (int q = 0; q < 7; ++q)
          if (!(ls >> r[q])) throw std::runtime_error("inner face file: bad record");
      }
      if (t[6] < 1 || t[6] > (int)blocks.size()) throw std::runtime_error("inner face file: bad target block");
      blocks[b].inner.push_back(register_inner(s, t, t[6], blocks[b].dim, blocks[b].ngg));
    }
}

void fill_inner_ghost_grid(std::vector<HostBlock> &blocks) {
  for (auto &B : blocks) {
    const int G = B.ngg + 1;
    for (const auto &f : B.inner) {
      HostBlock &T = blocks[f.target_block];
      Coord bx{B.x, B.mx, B.my, B.mz, G}, by{B.y, B.mx, B.my, B.mz, G}, bz{B.z, B.mx, B.my, B.mz, G};
      Coord tx{T.x, T.mx, T.my, T.mz, G}, ty{T.y, T.mx, T.my, T.mz, G}, tz{T.z, T.mx, T.my, T.mz, G};
      // the ngg + 1 ghost layers beyond the coincident layer, from the target's interior (init_inner_ghost_grid)
      int lo[3] = {f.start[0], f.start[1], f.start[2]}, tl[3] = {f.target_start[0], f.target_start[1],
                                                                 f.target_start[2]};
      lo[f.face] += f.direction < 0 ? -1 : 1;
      tl[f.target_face] += f.target_direction > 0 ? -1 : 1;
      const int T_ext[3] = {T.mx, T.my, T.mz};
      for (int n2 = 0; n2 < f.n_point[2]; ++n2)
        for (int n1 = 0; n1 < f.n_point[1]; ++n1)
          for (int n0 = 0; n0 < f.n_point[0]; ++n0) {
            const int d[3] = {f.loop_dir[0] * n0, f.loop_dir[1] * n1, f.loop_dir[2] * n2};
            int a[3], t[3];
            for (int c = 0; c < 3; ++c) a[c] = lo[c] + d[c];
            const int B_ext[3] = {B.mx, B.my, B.mz};
            for (int c = 0; c < 3; ++c)
              if (a[c] < -G || a[c] >= B_ext[c] + G) throw std::runtime_error("inner face outside the ghost grid");
            for (int c = 0; c < 3; ++c) t[c] = tl[c] + f.target_loop_dir[c] * d[f.src_tar[c]];
            for (int c = 0; c < 3; ++c)
              if (t[c] < 0 || t[c] >= T_ext[c]) throw std::runtime_error("inner face maps outside the target block");
            bx(a[0], a[1], a[2]) = tx(t[0], t[1], t[2]);
            by(a[0], a[1], a[2]) = ty(t[0], t[1], t[2]);
            bz(a[0], a[1], a[2]) = tz(t[0], t[1], t[2]);
          }
    }
  }
}

bool compute_metrics(HostBlock &b, std::string &err, bool mirror) {
  if (mirror) mirror_ghosts(b);
  const int g = b.ngg, G = g + 1;
  const long sz = b.size();
  b.jac.assign(sz, 0.0);
  b.metric.assign(9 * sz, 0.0);
  Coord X{b.x, b.mx, b.my, b.mz, G}, Y{b.y, b.mx, b.my, b.mz, G}, Z{b.z, b.mx, b.my, b.mz, G};
  // Second-order central metric of one node (Mesh.cpp:255-293); false if the Jacobian is not positive.
  auto node = [&](int i, int j, int k) -> bool {
    const double a1 = (X(i + 1, j, k) - X(i - 1, j, k)) * 0.5, b1 = (Y(i + 1, j, k) - Y(i - 1, j, k)) * 0.5,
                 c1 = (Z(i + 1, j, k) - Z(i - 1, j, k)) * 0.5;
    const double a2 = (X(i, j + 1, k) - X(i, j - 1, k)) * 0.5, b2 = (Y(i, j + 1, k) - Y(i, j - 1, k)) * 0.5,
                 c2 = (Z(i, j + 1, k) - Z(i, j - 1, k)) * 0.5;
    const double a3 = (X(i, j, k + 1) - X(i, j, k - 1)) * 0.5, b3 = (Y(i, j, k + 1) - Y(i, j, k - 1)) * 0.5,
                 c3 = (Z(i, j, k + 1) - Z(i, j, k - 1)) * 0.5;
    const double J = a1 * b2 * c3 + a2 * b3 * c1 + a3 * b1 * c2 - a1 * b3 * c2 - a2 * b1 * c3 - a3 * b2 * c1;
    if (J <= 0) return false;
    const long id = b.idx(i, j, k);
    b.jac[id] = J;
    double *m = b.metric.data() + id;
    m[0 * sz] = (b2 * c3 - b3 * c2) / J;
    m[1 * sz] = (a3 * c2 - a2 * c3) / J;
    m[2 * sz] = (a2 * b3 - a3 * b2) / J;
    m[3 * sz] = (b3 * c1 - b1 * c3) / J;
    m[4 * sz] = (a1 * c3 - a3 * c1) / J;
    m[5 * sz] = (a3 * b1 - a1 * b3) / J;
    m[6 * sz] = (b1 * c2 - b2 * c1) / J;
    m[7 * sz] = (a2 * c1 - a1 * c2) / J;
    m[8 * sz] = (a1 * b2 - a2 * b1) / J;
    return true;
  };
  auto copy_from = [&](int i, int j, int k, int si, int sj, int sk) {
    const long d = b.idx(i, j, k), s = b.idx(si, sj, sk);
    b.jac[d] = b.jac[s];
    for (int l = 0; l < 9; ++l) b.metric[l * sz + d] = b.metric[l * sz + s];
  };
  const int mx = b.mx, my = b.my, mz = b.mz;
  const int k0 = b.dim == 2 ? 0 : 1, k1 = b.dim == 2 ? 1 : mz - 1;
  bool ok = true;
  for (int i = 1; i < mx - 1 && ok; ++i)
    for (int j = 1; j < my - 1 && ok; ++j)
      for (int k = k0; k < k1; ++k)
        if (!node(i, j, k)) {
          err = "negative Jacobian at interior node (" + std::to_string(i) + "," + std::to_string(j) + "," +
                std::to_string(k) + ")";
          ok = false;
          break;
        }
  if (!ok) return false;
  // Boundary and ghost layers in the reference's order; a non-positive Jacobian takes the
  // neighbour's values towards the interior (Mesh.cpp:300-528).
  for (int i = 1; i < mx - 1; ++i)
    for (int j = 1; j < my - 1; ++j)
      for (int k = 0; k >= -g; --k)
        if (!node(i, j, k)) copy_from(i, j, k, i, j, k + 1);
  for (int i = 1; i < mx - 1; ++i)
    for (int j = 1; j < my - 1; ++j)
      for (int k = mz - 1; k < mz + g; ++k)
        if (!node(i, j, k)) copy_from(i, j, k, i, j, k - 1);
  for (int i = 1; i < mx - 1; ++i)
    for (int j = 0; j >= -g; --j)
      for (int k = -g; k < mz + g; ++k)
        if (!node(i, j, k)) copy_from(i, j, k, i, j + 1, k);
  for (int i = 1; i < mx - 1; ++i)
    for (int j = my - 1; j < my + g; ++j)
      for (int k = -g; k < mz + g; ++k)
        if (!node(i, j, k)) copy_from(i, j, k, i, j - 1, k);
  for (int i = 0; i >= -g; --i)
    for (int j = -g; j < my + g; ++j)
      for (int k = -g; k < mz + g; ++k)
        if (!node(i, j, k)) copy_from(i, j, k, i + 1, j, k);
  for (int i = mx - 1; i < mx + g; ++i)
    for (int j = -g; j < my + g; ++j)
      for (int k = -g; k < mz + g; ++k)
        if (!node(i, j, k)) copy_from(i, j, k, i - 1, j, k);
  return true;
}

} // namespace corefl
