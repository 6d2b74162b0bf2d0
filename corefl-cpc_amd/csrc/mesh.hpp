// Host-side mesh: grid and boundary readers, mirrored ghost grid, 2nd-order metrics.
// Reference: Mesh.cpp:244-528 (metrics), 692-925 (readers), 1013-1167 (ghost grid).
#pragma once

#include <string>
#include <vector>

namespace corefl {

// One boundary face (Boundary, Mesh.h:8-31) after register_boundary (Mesh.cpp:14-32).
struct Face {
  int start[3], end[3];
  int type;      // 6 outflow, 8 periodic, ...
  int face;      // normal direction 0/1/2
  int direction; // -1 on the small-index face, +1 otherwise
};

// One block-to-block face of this block after InnerFace::register_boundary (Mesh.cpp:34-173).  Node
// (start + loop_dir * n) for n in [0, n_point) per direction covers the coincident layer and the ngg ghost
// layers beyond it; its partner in the target block is target_start[t] + target_loop_dir[t] * d[src_tar[t]],
// d being the signed offset (loop_dir * n) along each direction of this block.
struct InnerFace {
  int start[3], end[3];
  int face, direction;  // normal direction of the face; -1 on the small-index face
  int face_layer;       // index of the coincident layer along the normal
  int target_block;     // 0-based
  int target_start[3];
  int target_face, target_direction;
  int src_tar[3];
  int loop_dir[3], target_loop_dir[3];
  int n_point[3];
};

struct HostBlock {
  int mx, my, mz, ngg, dim;
  // coordinates with ngg+1 ghost layers (i fastest)
  std::vector<double> x, y, z;
  // metrics/Jacobian in the ngg-ghost DZone layout: metric[l*sz + idx], jac[idx]
  std::vector<double> jac, metric;
  std::vector<Face> faces;
  std::vector<InnerFace> inner; // in inner%4d.txt order
  // physical-boundary flags per side (il, ir, jl, jr, kl, kr) for one-sided viscous stencils
  int btype[6] = {0, 0, 0, 0, 0, 0};

  long pad2() const { return mx + 2L * ngg; }
  long pad1() const { return (my + 2L * ngg) * pad2(); }
  long size() const { return pad1() * (mz + 2L * ngg); }
  long idx(int i, int j, int k) const { return (k + ngg) * pad1() + (j + ngg) * pad2() + (i + ngg); }
};

// Parse grid%4d.dat content (int32 n_block; n_block*(mx,my,mz); x,y,z float64 per block, i fastest).
std::vector<HostBlock> read_grid_binary(const std::string &path, int ngg, double grid_scale);
// Parse grid%4d.grd (ASCII: block count, one "mx my mz" line per block, then x, y, z of each block, i fastest).
std::vector<HostBlock> read_grid_ascii(const std::string &path, int ngg, double grid_scale);
// Build blocks from in-memory coordinates (same ordering as the binary file).
HostBlock make_block(int mx, int my, int mz, int ngg, const double *x, const double *y, const double *z,
                     double grid_scale);
// Parse boundary%4d.txt content for all blocks (Mesh.cpp:837-925).
void parse_boundaries(const std::string &text, std::vector<HostBlock> &blocks, int periodic_label);
// Parse inner%4d.txt content (Mesh.cpp:927-960): for every block its count of block-to-block faces, then two
// lines per face (this block's range and id, the target's range and id, 1-based, with the sign convention of
// the grid tool's interface records), registered with ngg ghost layers.
void parse_inner_faces(const std::string &text, std::vector<HostBlock> &blocks);
// Mirrored ghost grid of one block (Mesh.cpp:1013-1167).
void mirror_ghost_grid(HostBlock &b);
// Ghost coordinates across block-to-block faces from the neighbouring block (Mesh.cpp:1169-1219); after
// mirror_ghost_grid of every block.
void fill_inner_ghost_grid(std::vector<HostBlock> &blocks);
// Metrics and Jacobian (mirror = true builds the mirrored ghost grid first); returns false on a
// non-positive interior Jacobian.
bool compute_metrics(HostBlock &b, std::string &err, bool mirror = true);
// Symmetric-conservative metrics of order 2, 4 or 6 (Mesh.cu:10-250) on the device from the block's ghost grid,
// overwriting jac and metric; false (with err) on a non-positive Jacobian at any node, a 2D block or a HIP error.
bool sc_metrics_device(HostBlock &b, int order, std::string &err);

} // namespace corefl
