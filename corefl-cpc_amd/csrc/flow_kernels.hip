// Flow kernels for CDNA4 (gfx950): characteristic WENO5/7 + Riemann interface fluxes, collocated
// viscous fluxes, fused residual + SSP-RK3 update with primitive recovery, ghost-layer boundary
// fills, and the CFL time-step reduction.  All arithmetic is FP64.
//
// Thread mapping: threadIdx.x always walks i (the unit-stride index of the DZone layout), so every
// global load and store is coalesced along a 64-lane wavefront whatever the sweep direction.
#include <cstdlib>
#include "flow.hpp"

#include <cmath>

#include "stencil.hpp"
#include "weno.hpp"

namespace corefl {

namespace {

using wenodev::rcp;
using wenodev::weno5_half;
using wenodev::weno7_half;

constexpr double kGamma = 1.4;          // gamma_air (Constants.h:21)
constexpr double kRu = 8314.462618;     // R_u (Constants.h:17)
constexpr double kMwAir = 28.9647;      // mw_air (Constants.h:26)
constexpr double kRair = kRu / kMwAir;  // R_air (Constants.h:34)

__device__ __forceinline__ long nid(const DevBlock &b, int i, int j, int k) {
  return (k + b.g) * b.d1 + (j + b.g) * b.d2 + (i + b.g);
}

// XCD-aware tile order.  Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// A 1-D launch of 8*P workgroups is remapped so that the workgroups landing on one XCD take a contiguous
// run of tiles: along the plane (consecutive rows, the i and j stencils) or, for the z sweep, along k, so that
// the nodes of a stencil are read by workgroups sharing an L2.  The map b -> (b % 8) * P + b / 8 is a
// bijection on [0, 8P).
// Within a plane, a workgroup of 256 threads takes 256 consecutive nodes of the launch's nx x ny nodes in row-major
// order (i fastest), so that only the last workgroup of a plane has idle lanes whatever nx is: a 64-wide tiling of
// 257 nodes leaves a fifth wave per row with one active lane, a quarter more waves than nodes need.
struct TileGrid {
  int gx, gy, gz; // workgroups per plane (gy = 1), and planes (or plane groups) along k
  int fast;       // 0: plane chunk then k (each XCD takes a contiguous run), 2: k fastest
  int z0;         // first k tile (a launch over a slab of planes)
  int nx, ny;     // nodes of a plane covered by the launch
};

struct Tile {
  int z;    // plane (or plane group) index
  int i, j; // this thread's node, before the kernel's own offsets
};

// A launch over the nx x ny nodes of gz planes (or plane groups) starting at z0.
inline TileGrid plane_grid(int nx, int ny, int gz, int fast, int z0) {
  return TileGrid{(nx * ny + 255) / 256, 1, gz, fast, z0, nx, ny};
}

__device__ __forceinline__ bool xcd_tile(const TileGrid &t, Tile &out) {
  const unsigned total = (unsigned)t.gx * t.gz;
  const unsigned per = (total + 7) / 8;
  const unsigned n = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (n >= total) return false;
  unsigned x;
  if (t.fast == 2) {
    out.z = n % t.gz;
    x = n / t.gz;
  } else {
    x = n % t.gx;
    out.z = n / t.gx;
  }
  out.z += t.z0;
  const unsigned p = x * 256u + threadIdx.y * 64u + threadIdx.x, nx = t.nx;
  out.j = (int)(p / nx);
  out.i = (int)(p - (unsigned)out.j * nx);
  return out.j < t.ny;
}

inline unsigned tile_blocks(const TileGrid &t) {
  const unsigned total = (unsigned)t.gx * t.gz;
  return 8 * ((total + 7) / 8);
}

// ---------------------------------------------------------------- Riemann solvers
// Primitive state: rho, u, v, w, p, E (E = total energy per unit volume).

// AUSM+ on Jacobian-weighted averaged metrics (InviscidScheme.cu:397-469).
__device__ __forceinline__ void flux_ausm(const double *L, const double *R, const double *ml, const double *mr,
                                          double jl, double jr, double *fc) {
  const double k1 = 0.5 * (jl * ml[0] + jr * mr[0]);
  const double k2 = 0.5 * (jl * ml[1] + jr * mr[1]);
  const double k3 = 0.5 * (jl * ml[2] + jr * mr[2]);
  // divisions as products with the hardware reciprocal refined to the rounding level (wenodev::rcp)
  const double gk = sqrt(k1 * k1 + k2 * k2 + k3 * k3), igk = wenodev::rcp(gk);
  const double ul = (k1 * L[1] + k2 * L[2] + k3 * L[3]) * igk;
  const double ur = (k1 * R[1] + k2 * R[2] + k3 * R[3]) * igk;
  const double il = wenodev::rcp(L[0]), ir = wenodev::rcp(R[0]);
  const double c = 0.5 * (sqrt(kGamma * L[4] * il) + sqrt(kGamma * R[4] * ir));
  const double ic = wenodev::rcp(c);
  const double Ml = ul * ic, Mr = ur * ic;
  constexpr double alpha = 3 / 16.0;
  double mlp, plp, mrn, prn;
  if (fabs(Ml) > 1) {
    mlp = 0.5 * (Ml + fabs(Ml));
    plp = mlp / Ml;
  } else {
    const double a = (Ml + 1) * (Ml + 1) * 0.25, bb = (Ml * Ml - 1) * (Ml * Ml - 1);
    mlp = a + 0.125 * bb;
    plp = a * (2 - Ml) + alpha * Ml * bb;
  }
  if (fabs(Mr) > 1) {
    mrn = 0.5 * (Mr - fabs(Mr));
    prn = mrn / Mr;
  } else {
    const double a = (Mr - 1) * (Mr - 1) * 0.25, bb = (Mr * Mr - 1) * (Mr * Mr - 1);
    mrn = -a - 0.125 * bb;
    prn = a * (2 + Mr) - alpha * Mr * bb;
  }
  const double pc = plp * L[4] + prn * R[4];
  const double mh = mlp + mrn;
  const double mass = c * (L[0] * 0.5 * (mh + fabs(mh)) + R[0] * 0.5 * (mh - fabs(mh)));
  const double co = mass * gk;
  const double *S = mass >= 0 ? L : R;
  fc[0] = co;
  fc[1] = co * S[1] + pc * k1;
  fc[2] = co * S[2] + pc * k2;
  fc[3] = co * S[3] + pc * k3;
  fc[4] = co * (S[5] + S[4]) * (mass >= 0 ? il : ir);
}

// HLLC with Roe-averaged wave-speed bounds (InviscidScheme.cu:471-599, single species).
__device__ __forceinline__ void flux_hllc(const double *L, const double *R, const double *ml, const double *mr,
                                          double jl, double jr, double *fc) {
  const double sql = sqrt(L[0]), sqr = sqrt(R[0]);
  const double wl = sql / (sql + sqr), wr = sqr / (sql + sqr);
  const double ut = L[1] * wl + R[1] * wr, vt = L[2] * wl + R[2] * wr, wt = L[3] * wl + R[3] * wr;
  const double ht = (L[5] + L[4]) / L[0] * wl + (R[5] + R[4]) / R[0] * wr;
  const double ct = sqrt((kGamma - 1) * (ht - 0.5 * (ut * ut + vt * vt + wt * wt)));
  const double kx = 0.5 * (ml[0] + mr[0]), ky = 0.5 * (ml[1] + mr[1]), kz = 0.5 * (ml[2] + mr[2]);
  const double gk = sqrt(kx * kx + ky * ky + kz * kz);
  const double Ut = (kx * ut + ky * vt + kz * wt) / gk;
  const double ja = 0.5 * (jl + jr);
  const double Ul = kx * L[1] + ky * L[2] + kz * L[3];
  const double Ur = kx * R[1] + ky * R[2] + kz * R[3];
  const double sl = fmin(Ul / gk - sqrt(kGamma * L[4] / L[0]), Ut - ct);
  const double sr = fmax(Ur / gk + sqrt(kGamma * R[4] / R[0]), Ut + ct);
  if (sl >= 0 || sr < 0) { // fully supersonic: upwind physical flux
    const double *S = sl >= 0 ? L : R;
    const double Us = sl >= 0 ? Ul : Ur;
    const double m = S[0] * Us;
    fc[0] = ja * m;
    fc[1] = ja * (m * S[1] + S[4] * kx);
    fc[2] = ja * (m * S[2] + S[4] * ky);
    fc[3] = ja * (m * S[3] + S[4] * kz);
    fc[4] = ja * ((S[5] + S[4]) * Us);
    return;
  }
  const double dl = sl - Ul / gk, dr = sr - Ur / gk;
  const double sm = ((R[0] * Ur * dr - L[0] * Ul * dl) / gk + L[4] - R[4]) / (R[0] * dr - L[0] * dl);
  const double pm = L[0] * dl * (sm - Ul / gk) + L[4];
  const bool left = sm >= 0;
  const double *S = left ? L : R;
  const double ss = left ? sl : sr, Us = left ? Ul : Ur;
  const double pco = 1.0 / (ss - sm);
  const double Qc = ja * pco * sm * (ss * gk - Us) * S[0];
  const double dP = (ss * pm - sm * S[4]) * pco * ja;
  fc[0] = Qc;
  fc[1] = Qc * S[1] + dP * kx;
  fc[2] = Qc * S[2] + dP * ky;
  fc[3] = Qc * S[3] + dP * kz;
  fc[4] = Qc * S[5] / S[0] + pco * ja * (ss * pm * sm * gk - sm * S[4] * Us);
}

// Roe flux with Harten's entropy fix (InviscidScheme.cu:601-765, single species).
__device__ __forceinline__ void flux_roe(const double *L, const double *R, const double *ml, const double *mr,
                                         double jl, double jr, double efl, double efr, double *fc) {
  {
    const double ax = jl * ml[0], ay = jl * ml[1], az = jl * ml[2];
    const double Uk = L[1] * ax + L[2] * ay + L[3] * az, co = Uk * L[0];
    fc[0] = 0.5 * co;
    fc[1] = 0.5 * (co * L[1] + L[4] * ax);
    fc[2] = 0.5 * (co * L[2] + L[4] * ay);
    fc[3] = 0.5 * (co * L[3] + L[4] * az);
    fc[4] = 0.5 * Uk * (L[4] + L[5]);
  }
  {
    const double ax = jr * mr[0], ay = jr * mr[1], az = jr * mr[2];
    const double Uk = R[1] * ax + R[2] * ay + R[3] * az, co = Uk * R[0];
    fc[0] += 0.5 * co;
    fc[1] += 0.5 * (co * R[1] + R[4] * ax);
    fc[2] += 0.5 * (co * R[2] + R[4] * ay);
    fc[3] += 0.5 * (co * R[3] + R[4] * az);
    fc[4] += 0.5 * Uk * (R[4] + R[5]);
  }
  const double dl = sqrt(L[0]), dr = sqrt(R[0]);
  const double iv = 1.0 / (dl + dr);
  const double u = (dl * L[1] + dr * R[1]) * iv, v = (dl * L[2] + dr * R[2]) * iv, w = (dl * L[3] + dr * R[3]) * iv;
  const double ek = 0.5 * (u * u + v * v + w * w);
  const double h = (dl * ((L[5] + L[4]) / L[0]) + dr * ((R[5] + R[4]) / R[0])) * iv;
  const double c = sqrt((kGamma - 1) * (h - ek));
  double kx = 0.5 * (ml[0] + mr[0]), ky = 0.5 * (ml[1] + mr[1]), kz = 0.5 * (ml[2] + mr[2]);
  const double gk = sqrt(kx * kx + ky * ky + kz * kz);
  double Uk = kx * u + ky * v + kz * w;
  const double efa = 0.5 * (efl + efr);
  double lam[3] = {fabs(Uk - gk * c), fabs(Uk), fabs(Uk + gk * c)};
#pragma unroll
  for (int m = 0; m < 3; ++m)
    if (lam[m] < efa) lam[m] = 0.5 * (lam[m] * lam[m] / efa + efa);
  kx /= gk;
  ky /= gk;
  kz /= gk;
  Uk /= gk;
  const double ja = 0.5 * (jl + jr);
  const double q0 = ja * (R[0] - L[0]);
  const double q1 = ja * (R[0] * R[1] - L[0] * L[1]);
  const double q2 = ja * (R[0] * R[2] - L[0] * L[2]);
  const double q3 = ja * (R[0] * R[3] - L[0] * L[3]);
  const double q4 = ja * (R[5] - L[5]);
  double c1 = (kGamma - 1) * (ek * q0 - u * q1 - v * q2 - w * q3 + q4) / (c * c);
  double c2 = (kx * q1 + ky * q2 + kz * q3 - Uk * q0) / c;
  const double c3 = q0 - c1;
  const double b0 = -lam[0] * 0.5 * (c1 - c2);
  const double b1 = -lam[1] * (kx * c3 - ((kz * v - ky * w) * q0 - kz * q2 + ky * q3) / c);
  const double b2 = -lam[1] * (ky * c3 - ((kx * w - kz * u) * q0 - kx * q3 + kz * q1) / c);
  const double b3 = -lam[1] * (kz * c3 - ((ky * u - kx * v) * q0 - ky * q1 + kx * q2) / c);
  const double b4 = -lam[2] * 0.5 * (c1 + c2);
  const double c0 = kx * b1 + ky * b2 + kz * b3;
  c1 = c0 + b0 + b4;
  c2 = c * (b4 - b0);
  fc[0] += 0.5 * c1;
  fc[1] += 0.5 * (u * c1 + kx * c2 - c * (kz * b2 - ky * b3));
  fc[2] += 0.5 * (v * c1 + ky * c2 - c * (kx * b3 - kz * b1));
  fc[3] += 0.5 * (w * c1 + kz * c2 - c * (ky * b1 - kx * b2));
  fc[4] += 0.5 * (h * c1 + Uk * c2 - c * c * c0 / (kGamma - 1) +
                  c * ((kz * v - ky * w) * b1 + (kx * w - kz * u) * b2 + (ky * u - kx * v) * b3));
}

// Local Lax-Friedrichs (InviscidScheme.cu:772-805).
__device__ __forceinline__ void flux_lf(const double *L, const double *R, const double *ml, const double *mr,
                                        double jl, double jr, double *fc) {
  const double kx = 0.5 * (ml[0] + mr[0]), ky = 0.5 * (ml[1] + mr[1]), kz = 0.5 * (ml[2] + mr[2]);
  const double gk = sqrt(kx * kx + ky * ky + kz * kz);
  const double Ul = L[1] * kx + L[2] * ky + L[3] * kz, Ur = R[1] * kx + R[2] * ky + R[3] * kz;
  const double sp = fmax(fabs(Ul) + sqrt(kGamma * L[4] / L[0]) * gk, fabs(Ur) + sqrt(kGamma * R[4] / R[0]) * gk);
  const double hj = 0.25 * (jl + jr);
  const double ml_ = L[0] * Ul, mr_ = R[0] * Ur;
  fc[0] = (ml_ + mr_ - sp * (R[0] - L[0])) * hj;
  fc[1] = (ml_ * L[1] + mr_ * R[1] + kx * (L[4] + R[4]) - sp * (R[1] * R[0] - L[1] * L[0])) * hj;
  fc[2] = (ml_ * L[2] + mr_ * R[2] + ky * (L[4] + R[4]) - sp * (R[2] * R[0] - L[2] * L[0])) * hj;
  fc[3] = (ml_ * L[3] + mr_ * R[3] + kz * (L[4] + R[4]) - sp * (R[3] * R[0] - L[3] * L[0])) * hj;
  fc[4] = ((L[5] + L[4]) * Ul + (R[5] + R[4]) * Ur - sp * (R[5] - L[5])) * hj;
}

__device__ __forceinline__ void to_primitive(const double *q, double *p) {
  const double ri = wenodev::rcp(q[0]);
  p[0] = q[0];
  p[1] = q[1] * ri;
  p[2] = q[2] * ri;
  p[3] = q[3] * ri;
  p[4] = (kGamma - 1) * (q[4] - 0.5 * (q[1] * p[1] + q[2] * p[2] + q[3] * p[3]));
  p[5] = q[4];
}

// ---------------------------------------------------------------- interface flux kernel
// One thread per interface t+1/2 along DIR.  The stencil q[0..2R-1] holds conservative states of
// nodes t-R+1 .. t+R.  Characteristic fields use the left eigenvectors of the Roe average of nodes
// t and t+1 (rows as in WENO.cu:73-124), reconstruction is one-sided WENO in those fields,
// back-projection is the explicit R of WENO.cu:229-243, then the chosen Riemann solver.
template<int DIR, int R, int RS>
__global__ __launch_bounds__(256) void k_inviscid(DevBlock b, Scheme sc, double *F, TileGrid tg) {
  Tile t;
  if (!xcd_tile(tg, t)) return;
  int i = t.i, j = t.j, k = t.z;
  if constexpr (DIR == 0) i -= 1;
  else if constexpr (DIR == 1) j -= 1;
  else k -= 1;
  if (i >= b.mx || j >= b.my) return;
  const long st = DIR == 0 ? 1 : (DIR == 1 ? b.d2 : b.d1);
  const long id = nid(b, i, j, k);
  const long sz = b.sz;
  constexpr int N = 2 * R;
  double w[N][5];
#pragma unroll
  for (int m = 0; m < N; ++m)
#pragma unroll
    for (int l = 0; l < 5; ++l) w[m][l] = b.cv[id + (m - R + 1) * st + l * sz];

  double ml[3], mr[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    ml[c] = b.met[id + (3 * DIR + c) * sz];
    mr[c] = b.met[id + st + (3 * DIR + c) * sz];
  }
  const double jl = b.jac[id], jr = b.jac[id + st];

  // Roe average of the two interface nodes
  const double *qa = w[R - 1], *qb = w[R];
  const double ra = qa[0], rb = qb[0];
  const double sab = sqrt(ra * rb);
  const double ca = wenodev::rcp(ra + sab), cb = wenodev::rcp(sab + rb);
  const double um = ca * qa[1] + cb * qb[1], vm = ca * qa[2] + cb * qb[2], wm = ca * qa[3] + cb * qb[3];
  const double pa = (kGamma - 1) * (qa[4] - 0.5 * (qa[1] * qa[1] + qa[2] * qa[2] + qa[3] * qa[3]) * wenodev::rcp(ra));
  const double pb = (kGamma - 1) * (qb[4] - 0.5 * (qb[1] * qb[1] + qb[2] * qb[2] + qb[3] * qb[3]) * wenodev::rcp(rb));
  const double H = ca * (qa[4] + pa) + cb * (qb[4] + pb);
  const double ek = 0.5 * (um * um + vm * vm + wm * wm);
  const double c = sqrt((kGamma - 1) * (H - ek));
  double kx = 0.5 * (ml[0] + mr[0]), ky = 0.5 * (ml[1] + mr[1]), kz = 0.5 * (ml[2] + mr[2]);
  {
    const double r = wenodev::rcp(sqrt(kx * kx + ky * ky + kz * kz));
    kx *= r; ky *= r; kz *= r;
  }
  double nx = 0, ny = 0, nz = 0;
  if (fabs(kx) > 0.8) {
    nx = -kz; nz = kx;
  } else {
    ny = kz; nz = -ky;
  }
  {
    const double r = wenodev::rcp(sqrt(nx * nx + ny * ny + nz * nz));
    nx *= r; ny *= r; nz *= r;
  }
  double qx = ky * nz - kz * ny, qy = kz * nx - kx * nz, qz = kx * ny - ky * nx;
  {
    const double r = wenodev::rcp(sqrt(qx * qx + qy * qy + qz * qz));
    qx *= r; qy *= r; qz *= r;
  }
  const double Uk = kx * um + ky * vm + kz * wm;
  const double Un = nx * um + ny * vm + nz * wm;
  const double Uq = qx * um + qy * vm + qz * wm;
  // project in place: w[m] <- L q[m].  The rows of L (WENO.cu:73-124) in factored form, with
  // P = u.(rho u), K = k.(rho u), X = ek rho + E - P, Y = Uk rho - K:
  //   row 0 = (gm1 X + c Y) / 2c^2, row 4 = (gm1 X - c Y) / 2c^2, row 3 = rho - gm1 X / c^2,
  //   row 1 = n.(rho u) - Un rho,   row 2 = q.(rho u) - Uq rho.
  const double gm1 = kGamma - 1, ib = wenodev::rcp(c * c);
  const double hg = 0.5 * gm1 * ib, hc = 0.5 * c * ib, g2 = -gm1 * ib;
#pragma unroll
  for (int m = 0; m < N; ++m) {
    const double r = w[m][0], mx = w[m][1], my = w[m][2], mz = w[m][3], e = w[m][4];
    const double P = fma(um, mx, fma(vm, my, wm * mz));
    const double K = fma(kx, mx, fma(ky, my, kz * mz));
    const double X = fma(ek, r, e) - P, hY = hc * fma(Uk, r, -K);
    w[m][0] = fma(hg, X, hY);
    w[m][1] = fma(-Un, r, fma(nx, mx, fma(ny, my, nz * mz)));
    w[m][2] = fma(-Uq, r, fma(qx, mx, fma(qy, my, qz * mz)));
    w[m][3] = fma(g2, X, r);
    w[m][4] = fma(hg, X, -hY);
  }
  double WL[5], WR[5];
  const double eps = sc.weno_eps;
#pragma unroll
  for (int l = 0; l < 5; ++l) {
    if constexpr (R == 4) {
      const double a[8] = {w[0][l], w[1][l], w[2][l], w[3][l], w[4][l], w[5][l], w[6][l], w[7][l]};
      wenodev::weno7_pair(a, eps, WL[l], WR[l]);
    } else {
      WL[l] = weno5_half(w[0][l], w[1][l], w[2][l], w[3][l], w[4][l], eps);
      WR[l] = weno5_half(w[5][l], w[4][l], w[3][l], w[2][l], w[1][l], eps);
    }
  }
  // back-projection Q = R W
  auto back = [&](const double *W, double *Q) {
    const double s = W[0] + W[3] + W[4], d = W[0] - W[4];
    Q[0] = s;
    Q[1] = um * s - kx * c * d + nx * W[1] + qx * W[2];
    Q[2] = vm * s - ky * c * d + ny * W[1] + qy * W[2];
    Q[3] = wm * s - kz * c * d + nz * W[1] + qz * W[2];
    Q[4] = H * s - Uk * c * d + Un * W[1] + Uq * W[2] - c * c * (1.0 / gm1) * W[3];
  };
  double QL[5], QR[5], PL[6], PR[6];
  back(WL, QL);
  back(WR, QR);
  to_primitive(QL, PL);
  to_primitive(QR, PR);
  if (!(PL[0] > 0 && PL[4] > 0 && PR[0] > 0 && PR[4] > 0)) {
    // positivity fallback: first-order states of the two interface nodes
    const double Qa[5] = {b.cv[id], b.cv[id + sz], b.cv[id + 2 * sz], b.cv[id + 3 * sz], b.cv[id + 4 * sz]};
    const long ib2 = id + st;
    const double Qb[5] = {b.cv[ib2], b.cv[ib2 + sz], b.cv[ib2 + 2 * sz], b.cv[ib2 + 3 * sz], b.cv[ib2 + 4 * sz]};
    to_primitive(Qa, PL);
    to_primitive(Qb, PR);
  }
  double fc[5];
  if constexpr (RS == 1) flux_lf(PL, PR, ml, mr, jl, jr, fc);
  else if constexpr (RS == 2) flux_roe(PL, PR, ml, mr, jl, jr, b.efd[id], b.efd[id + st], fc);
  else if constexpr (RS == 4) flux_hllc(PL, PR, ml, mr, jl, jr, fc);
  else flux_ausm(PL, PR, ml, mr, jl, jr, fc);
#pragma unroll
  for (int l = 0; l < 5; ++l) F[id + l * sz] = fc[l];
}

// ---------------------------------------------------------------- primitive reconstruction (air)
// minmod limiter (Limiter.cuh:8-41; limiter 0 and 1 are both minmod).
__device__ __forceinline__ double minmod(double a, double c) {
  const double sa = (a > 0) - (a < 0), sc = (c > 0) - (c < 0);
  return 0.5 * (sa + sc) * fmin(fabs(a), fabs(c));
}

// compute_convective_term_pv_1D<Air> and Roe_compute_inviscid_flux_1D<Air> (InviscidScheme.cu:61-169, 306-395):
// one thread per interface t+1/2 along DIR.  rho, u, v, w, p of nodes t-1 .. t+2 reconstructed to the interface
// (Reconstruction.cu:5-44: REC 1 first order, 2 MUSCL with kappa = 1/3, 3 NND2, both with minmod), the total
// energy of each side from its reconstructed state (InviscidScheme.cu:212-217), then the Riemann solver RS.
template<int DIR, int REC, int RS>
__global__ __launch_bounds__(256) void k_inviscid_pv(DevBlock b, Scheme sc, double *F, TileGrid tg) {
  Tile t;
  if (!xcd_tile(tg, t)) return;
  int i = t.i, j = t.j, k = t.z;
  if constexpr (DIR == 0) i -= 1;
  else if constexpr (DIR == 1) j -= 1;
  else k -= 1;
  if (i >= b.mx || j >= b.my) return;
  const long st = DIR == 0 ? 1 : (DIR == 1 ? b.d2 : b.d1);
  const long id = nid(b, i, j, k), sz = b.sz;
  double PL[6], PR[6];
#pragma unroll
  for (int l = 0; l < 5; ++l) {
    const double *f = b.bv + l * sz + id;
    const double v1 = f[0], v2 = f[st];
    if constexpr (REC == 1) {
      PL[l] = v1;
      PR[l] = v2;
    } else {
      const double v0 = f[-st], v3 = f[2 * st];
      const double d0 = v1 - v0, d1 = v2 - v1, d2 = v3 - v2;
      if constexpr (REC == 2) {
        constexpr double kappa = 1.0 / 3.0;
        PL[l] = v1 + 0.25 * ((1 - kappa) * minmod(d0, d1) + (1 + kappa) * minmod(d1, d0));
        PR[l] = v2 - 0.25 * ((1 - kappa) * minmod(d2, d1) + (1 + kappa) * minmod(d1, d2));
      } else {
        PL[l] = v1 + 0.5 * minmod(d0, d1);
        PR[l] = v2 - 0.5 * minmod(d1, d2);
      }
    }
  }
  PL[5] = 0.5 * (PL[1] * PL[1] + PL[2] * PL[2] + PL[3] * PL[3]) * PL[0] + PL[4] / (kGamma - 1);
  PR[5] = 0.5 * (PR[1] * PR[1] + PR[2] * PR[2] + PR[3] * PR[3]) * PR[0] + PR[4] / (kGamma - 1);
  double ml[3], mr[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    ml[c] = b.met[id + (3 * DIR + c) * sz];
    mr[c] = b.met[id + st + (3 * DIR + c) * sz];
  }
  const double jl = b.jac[id], jr = b.jac[id + st];
  double fc[5];
  if constexpr (RS == 1) flux_lf(PL, PR, ml, mr, jl, jr, fc);
  else if constexpr (RS == 2) flux_roe(PL, PR, ml, mr, jl, jr, b.efd[id], b.efd[id + st], fc);
  else if constexpr (RS == 4) flux_hllc(PL, PR, ml, mr, jl, jr, fc);
  else flux_ausm(PL, PR, ml, mr, jl, jr, fc);
#pragma unroll
  for (int l = 0; l < 5; ++l) F[id + l * sz] = fc[l];
}

// ---------------------------------------------------------------- LLF flux-splitting WENO (air)
// inviscid_scheme 51/71: component-wise WENO5/WENO7 of the locally split fluxes (WENO.cu:890-990);
// 52/72: characteristic WENO of the fluxes split with the largest spectral radius on the stencil
// (WENO.cu:10-244 with the stencil of WENO.cu:1096-1136).  One thread per interface t+1/2 along DIR; the
// stencil nodes are t-R+1 .. t+R (m = 0 .. 2R-1).  With the hybrid scheme on, interfaces whose nodes
// t-ngg+1 .. t+ngg all have a shock sensor at or below the threshold take the linear upwind weights (UD5/UD7);
// the others, and all of them when it is off, the nonlinear ones.  F receives the interface flux at node t.
template<int DIR, int R, bool CH>
__global__ __launch_bounds__(256, 2) void k_fsweno(DevBlock b, Scheme sc, double *F, TileGrid tg) {
  Tile t;
  if (!xcd_tile(tg, t)) return;
  int i = t.i, j = t.j, k = t.z;
  if constexpr (DIR == 0) i -= 1;
  else if constexpr (DIR == 1) j -= 1;
  else k -= 1;
  if (i >= b.mx || j >= b.my) return;
  const long st = DIR == 0 ? 1 : (DIR == 1 ? b.d2 : b.d1);
  const long id = nid(b, i, j, k);
  const long sz = b.sz;
  constexpr int N = 2 * R;
  constexpr double gm1 = kGamma - 1;
  const bool shock = wenodev::stencil_shock(b.sensor, sc.sensor_threshold, id, st, b.g);

  if constexpr (!CH) {
    // split fluxes F+- = (J/2)(k.F +- lambda q), lambda = |Uk| + |k| c of the node itself; fp holds
    // nodes 0..N-2 (the plus stencil), fm nodes 1..N-1 (the minus stencil)
    double fp[N - 1][5], fm[N - 1][5];
#pragma unroll
    for (int m = 0; m < N; ++m) {
      const long im = id + (m - R + 1) * st;
      const double q0 = b.cv[im], q1 = b.cv[im + sz], q2 = b.cv[im + 2 * sz], q3 = b.cv[im + 3 * sz],
                   q4 = b.cv[im + 4 * sz];
      const double k0 = b.met[im + (3 * DIR) * sz], k1 = b.met[im + (3 * DIR + 1) * sz],
                   k2 = b.met[im + (3 * DIR + 2) * sz];
      const double ri = 1.0 / q0;
      const double Uk = (q1 * k0 + q2 * k1 + q3 * k2) * ri;
      const double p = gm1 * (q4 - 0.5 * (q1 * q1 + q2 * q2 + q3 * q3) * ri);
      const double lam = fabs(Uk) + sqrt((k0 * k0 + k1 * k1 + k2 * k2) * (kGamma * p * ri));
      const double hj = 0.5 * b.jac[im];
      const double pk = p * hj;
      if (m < N - 1) {
        const double s = hj * (Uk + lam);
        fp[m][0] = s * q0;
        fp[m][1] = fma(s, q1, pk * k0);
        fp[m][2] = fma(s, q2, pk * k1);
        fp[m][3] = fma(s, q3, pk * k2);
        fp[m][4] = fma(s, q4, pk * Uk);
      }
      if (m > 0) {
        const double s = hj * (Uk - lam);
        fm[m - 1][0] = s * q0;
        fm[m - 1][1] = fma(s, q1, pk * k0);
        fm[m - 1][2] = fma(s, q2, pk * k1);
        fm[m - 1][3] = fma(s, q3, pk * k2);
        fm[m - 1][4] = fma(s, q4, pk * Uk);
      }
    }
    const double e0 = sc.weno_eps, e1 = e0 * sc.v_ref * sc.v_ref, e4 = e1 * sc.v_ref * sc.v_ref;
#pragma unroll
    for (int l = 0; l < 5; ++l) {
      const double eps = l == 0 ? e0 : (l == 4 ? e4 : e1);
      double f;
      if constexpr (R == 4) {
        const double a[7] = {fp[0][l], fp[1][l], fp[2][l], fp[3][l], fp[4][l], fp[5][l], fp[6][l]};
        const double c[7] = {fm[6][l], fm[5][l], fm[4][l], fm[3][l], fm[2][l], fm[1][l], fm[0][l]};
        f = shock ? weno7_half(a, eps) + weno7_half(c, eps) : wenodev::ud7_half(a) + wenodev::ud7_half(c);
      } else if (shock) {
        f = weno5_half(fp[0][l], fp[1][l], fp[2][l], fp[3][l], fp[4][l], eps) +
            weno5_half(fm[4][l], fm[3][l], fm[2][l], fm[1][l], fm[0][l], eps);
      } else {
        f = wenodev::ud5_half(fp[0][l], fp[1][l], fp[2][l], fp[3][l], fp[4][l]) +
            wenodev::ud5_half(fm[4][l], fm[3][l], fm[2][l], fm[1][l], fm[0][l]);
      }
      F[id + l * sz] = f;
    }
  } else {
    // Two passes over the stencil keep the register footprint within two waves per SIMD: the first finds
    // the largest spectral radius, the second projects each node's flux and state and splits them.
    double lmax = 0, pa = 0, pb = 0;
#pragma unroll
    for (int m = 0; m < N; ++m) {
      const long im = id + (m - R + 1) * st;
      const double q0 = b.cv[im], q1 = b.cv[im + sz], q2 = b.cv[im + 2 * sz], q3 = b.cv[im + 3 * sz],
                   q4 = b.cv[im + 4 * sz];
      const double k0 = b.met[im + (3 * DIR) * sz], k1 = b.met[im + (3 * DIR + 1) * sz],
                   k2 = b.met[im + (3 * DIR + 2) * sz];
      const double ri = 1.0 / q0;
      const double Uk = (q1 * k0 + q2 * k1 + q3 * k2) * ri;
      const double p = gm1 * (q4 - 0.5 * (q1 * q1 + q2 * q2 + q3 * q3) * ri);
      lmax = fmax(lmax, fabs(Uk) + sqrt((k0 * k0 + k1 * k1 + k2 * k2) * (kGamma * p * ri)));
      if (m == R - 1) pa = p;
      if (m == R) pb = p;
    }
    // Roe state of the interface nodes; sound speed from the Roe-weighted p/rho (WENO.cu:17-30)
    const long ia = id, ib = id + st;
    const double ra = b.cv[ia], rb = b.cv[ib];
    const double sab = sqrt(ra * rb);
    const double rlc = 1 / (ra + sab), rrc = 1 / (sab + rb);
    const double um = rlc * b.cv[ia + sz] + rrc * b.cv[ib + sz];
    const double vm = rlc * b.cv[ia + 2 * sz] + rrc * b.cv[ib + 2 * sz];
    const double wm = rlc * b.cv[ia + 3 * sz] + rrc * b.cv[ib + 3 * sz];
    const double H = rlc * (b.cv[ia + 4 * sz] + pa) + rrc * (b.cv[ib + 4 * sz] + pb);
    const double cm = sqrt(kGamma * (rlc * pa + rrc * pb));
    double kx = 0.5 * (b.met[ia + (3 * DIR) * sz] + b.met[ib + (3 * DIR) * sz]);
    double ky = 0.5 * (b.met[ia + (3 * DIR + 1) * sz] + b.met[ib + (3 * DIR + 1) * sz]);
    double kz = 0.5 * (b.met[ia + (3 * DIR + 2) * sz] + b.met[ib + (3 * DIR + 2) * sz]);
    {
      const double r = 1.0 / sqrt(kx * kx + ky * ky + kz * kz);
      kx *= r; ky *= r; kz *= r;
    }
    double nx = 0, ny = 0, nz = 0;
    if (fabs(kx) > 0.8) {
      nx = -kz; nz = kx;
    } else {
      ny = kz; nz = -ky;
    }
    {
      const double r = 1.0 / sqrt(nx * nx + ny * ny + nz * nz);
      nx *= r; ny *= r; nz *= r;
    }
    double qx = ky * nz - kz * ny, qy = kz * nx - kx * nz, qz = kx * ny - ky * nx;
    {
      const double r = 1.0 / sqrt(qx * qx + qy * qy + qz * qz);
      qx *= r; qy *= r; qz *= r;
    }
    const double Ub = kx * um + ky * vm + kz * wm;
    const double Un = nx * um + ny * vm + nz * wm, Uq = qx * um + qy * vm + qz * wm;
    const double hic2 = 0.5 / (cm * cm), al = 0.5 * gm1 * (um * um + vm * vm + wm * wm);
    // Left eigenvectors of WENO.cu:73-124 applied to x: rows 0 and 4 are A +- B, row 3 is x0 - 2A, with
    // A = (al x0 - gm1 (u . x1:3) + gm1 x4) / (2 c^2) and B = c (Ub x0 - k . x1:3) / (2 c^2).
    auto project = [&](const double *x, double *w) {
      const double A = hic2 * (al * x[0] - gm1 * (um * x[1] + vm * x[2] + wm * x[3]) + gm1 * x[4]);
      const double B = hic2 * cm * (Ub * x[0] - (kx * x[1] + ky * x[2] + kz * x[3]));
      w[0] = A + B;
      w[1] = nx * x[1] + ny * x[2] + nz * x[3] - Un * x[0];
      w[2] = qx * x[1] + qy * x[2] + qz * x[3] - Uq * x[0];
      w[3] = x[0] - 2 * A;
      w[4] = A - B;
    };
    double vp[5][N - 1], vn[5][N - 1]; // plus stencil nodes 0..N-2; minus stencil nodes N-1..1 (mirrored)
#pragma unroll
    for (int m = 0; m < N; ++m) {
      const long im = id + (m - R + 1) * st;
      const double q[5] = {b.cv[im], b.cv[im + sz], b.cv[im + 2 * sz], b.cv[im + 3 * sz], b.cv[im + 4 * sz]};
      const double k0 = b.met[im + (3 * DIR) * sz], k1 = b.met[im + (3 * DIR + 1) * sz],
                   k2 = b.met[im + (3 * DIR + 2) * sz];
      const double ri = 1.0 / q[0];
      const double Uk = (q[1] * k0 + q[2] * k1 + q[3] * k2) * ri;
      const double p = gm1 * (q[4] - 0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) * ri);
      const double fl[5] = {q[0] * Uk, fma(q[1], Uk, p * k0), fma(q[2], Uk, p * k1), fma(q[3], Uk, p * k2),
                            (q[4] + p) * Uk};
      const double hj = 0.5 * b.jac[im];
      double LF[5], LQ[5];
      project(fl, LF);
      project(q, LQ);
#pragma unroll
      for (int l = 0; l < 5; ++l) {
        if (m < N - 1) vp[l][m] = hj * (LF[l] + lmax * LQ[l]);
        if (m > 0) vn[l][N - 1 - m] = hj * (LF[l] - lmax * LQ[l]);
      }
    }
    const double eps = sc.weno_eps;
    double fch[5];
#pragma unroll
    for (int l = 0; l < 5; ++l) {
      if constexpr (R == 4) {
        fch[l] = shock ? weno7_half(vp[l], eps) + weno7_half(vn[l], eps)
                       : wenodev::ud7_half(vp[l]) + wenodev::ud7_half(vn[l]);
      } else if (shock) {
        fch[l] = weno5_half(vp[l][0], vp[l][1], vp[l][2], vp[l][3], vp[l][4], eps) +
                 weno5_half(vn[l][0], vn[l][1], vn[l][2], vn[l][3], vn[l][4], eps);
      } else {
        fch[l] = wenodev::ud5_half(vp[l][0], vp[l][1], vp[l][2], vp[l][3], vp[l][4]) +
                 wenodev::ud5_half(vn[l][0], vn[l][1], vn[l][2], vn[l][3], vn[l][4]);
      }
    }
    // back-projection with the Roe-averaged enthalpy (WENO.cu:229-243)
    const double s1 = fch[0] + fch[3] + fch[4], s3 = fch[0] - fch[4];
    F[id] = s1;
    F[id + sz] = um * s1 - kx * cm * s3 + nx * fch[1] + qx * fch[2];
    F[id + 2 * sz] = vm * s1 - ky * cm * s3 + ny * fch[1] + qy * fch[2];
    F[id + 3 * sz] = wm * s1 - kz * cm * s3 + nz * fch[1] + qz * fch[2];
    F[id + 4 * sz] = H * s1 - Ub * cm * s3 + Un * fch[1] + Uq * fch[2] - cm * cm / gm1 * fch[3];
  }
}

// ---------------------------------------------------------------- viscous terms
using stencil::dd;
using stencil::dd_closure;
using stencil::dd_mid;

// z-marching node kernels (k_viscous_flux, k_update): each thread walks its (i, j) column through kz
// consecutive planes and keeps the values its z stencil shares between neighbouring planes in registers, so
// every plane of the z-differenced arrays is read from memory once per column instead of once per stencil
// point.  kz trades that reuse against the number of waves in flight (COREFL_KZ overrides the default).

__device__ __forceinline__ double sutherland(double T) {
  // Transport.cu:6-9, with (T/288.16)^1.5 as x sqrt(x) and the divisions as products with reciprocals
  const double x = T * (1.0 / 288.16);
  return 1.7894e-5 * (x * sqrt(x)) * (288.16 + 110) * wenodev::rcp(T + 110);
}

// compute_fv/gv/hv_2nd_order<Air> (ViscousScheme.cu:614-1010): viscous flux on the face between node t and
// t+1 along DIR, stored at node t (components 1..4), for t = -1 .. m-1.  Metrics and metric * Jacobian are
// the two nodes' averages; the derivative along DIR is the nodes' difference, the transverse ones the
// average of the two nodes' central differences (none along z on 2D blocks); viscosity the nodes' average.
template<int DIR>
__global__ __launch_bounds__(256) void k_viscous_face(DevBlock b, Scheme sc, TileGrid tg) {
  Tile t;
  if (!xcd_tile(tg, t)) return;
  int i = t.i, j = t.j, k = t.z;
  if constexpr (DIR == 0) i -= 1;
  else if constexpr (DIR == 1) j -= 1;
  else k -= 1;
  if (i >= b.mx || j >= b.my) return;
  const long sz = b.sz;
  const long st[3] = {1, b.d2, b.d1};
  const long id = nid(b, i, j, k), ie = id + st[DIR];
  const double *M = b.met;
  double m[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) m[c] = 0.5 * (M[id + c * sz] + M[ie + c * sz]);
  constexpr int comp[4] = {1, 2, 3, 5};
  double g[4][3];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double *f = b.bv + comp[q] * sz;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (a == DIR) g[q][a] = f[ie] - f[id];
      else if (a == 2 && b.dim == 2) g[q][a] = 0;
      else g[q][a] = 0.25 * (f[id + st[a]] - f[id - st[a]] + f[ie + st[a]] - f[ie - st[a]]);
    }
  }
  double d[4][3];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int c = 0; c < 3; ++c) d[q][c] = g[q][0] * m[c] + g[q][1] * m[3 + c] + g[q][2] * m[6 + c];
  const double mu = 0.5 * (sutherland(b.bv[id + 5 * sz]) + sutherland(b.bv[ie + 5 * sz]));
  const double txx = mu * (4 * d[0][0] - 2 * d[1][1] - 2 * d[2][2]) * (1.0 / 3);
  const double tyy = mu * (4 * d[1][1] - 2 * d[0][0] - 2 * d[2][2]) * (1.0 / 3);
  const double tzz = mu * (4 * d[2][2] - 2 * d[0][0] - 2 * d[1][1]) * (1.0 / 3);
  const double txy = mu * (d[0][1] + d[1][0]), txz = mu * (d[0][2] + d[2][0]), tyz = mu * (d[1][2] + d[2][1]);
  double kj[3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
    kj[c] = 0.5 * (M[id + (3 * DIR + c) * sz] * b.jac[id] + M[ie + (3 * DIR + c) * sz] * b.jac[ie]);
  const double f1 = kj[0] * txx + kj[1] * txy + kj[2] * txz;
  const double f2 = kj[0] * txy + kj[1] * tyy + kj[2] * tyz;
  const double f3 = kj[0] * txz + kj[1] * tyz + kj[2] * tzz;
  const double um = 0.5 * (b.bv[id + sz] + b.bv[ie + sz]);
  const double vm = 0.5 * (b.bv[id + 2 * sz] + b.bv[ie + 2 * sz]);
  const double wm = 0.5 * (b.bv[id + 3 * sz] + b.bv[ie + 3 * sz]);
  constexpr double cp = kGamma * kRair / (kGamma - 1);
  const double kc = mu * (cp / sc.prandtl);
  double *F = DIR == 0 ? b.fv : (DIR == 1 ? b.gv : b.hv);
  F[id + sz] = f1;
  F[id + 2 * sz] = f2;
  F[id + 3 * sz] = f3;
  F[id + 4 * sz] = um * f1 + vm * f2 + wm * f3 + kc * (kj[0] * d[3][0] + kj[1] * d[3][1] + kj[2] * d[3][2]);
}

// compute_viscous_flux_collocated<Air, ORDER> (ViscousScheme.cu:309-408): node fluxes fv/gv/hv[1..4].
template<int ORDER>
__global__ __launch_bounds__(256) void k_viscous_flux(DevBlock b, Scheme sc, TileGrid tg, int kz) {
  Tile t;
  if (!xcd_tile(tg, t)) return;
  const int i = t.i, j = t.j;
  if (i >= b.mx || j >= b.my) return;
  const int k0 = t.z * kz, k1 = min(k0 + kz, b.mz);
  const long sz = b.sz, d1 = b.d1, id0 = nid(b, i, j, k0);
  const bool z3 = b.dim == 3;
  constexpr int H = ORDER / 2, W = ORDER + 1;
  const int comp[4] = {1, 2, 3, 5};
  double wz[4][W]; // u, v, w, T on planes k - H .. k + H
  if (z3) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int m = 1; m < W; ++m) wz[q][m] = b.bv[id0 + (m - 1 - H) * d1 + comp[q] * sz];
  }
  for (int k = k0; k < k1; ++k) {
    const long id = id0 + (long)(k - k0) * d1;
    if (z3) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
#pragma unroll
        for (int m = 0; m < W - 1; ++m) wz[q][m] = wz[q][m + 1];
        wz[q][W - 1] = b.bv[id + H * d1 + comp[q] * sz];
      }
    }
    const bool zc = z3 && dd_closure<ORDER>(k, b.mz, b.bt[4], b.bt[5]);
    double g[4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double *f = b.bv + comp[q] * sz;
      g[q][0] = dd<ORDER>(f, id, 1, i, b.mx, b.bt[0], b.bt[1]);
      g[q][1] = dd<ORDER>(f, id, b.d2, j, b.my, b.bt[2], b.bt[3]);
      g[q][2] = !z3 ? 0.0 : zc ? dd<ORDER>(f, id, d1, k, b.mz, b.bt[4], b.bt[5]) : dd_mid<ORDER>(wz[q]);
    }
    const double *M = b.met + id;
    const double xx = M[0], xy = M[sz], xz = M[2 * sz], ex = M[3 * sz], ey = M[4 * sz], ez = M[5 * sz];
    const double zx = M[6 * sz], zy = M[7 * sz], zz = M[8 * sz];
    double d[4][3]; // physical gradients of u, v, w, T
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      d[q][0] = g[q][0] * xx + g[q][1] * ex + g[q][2] * zx;
      d[q][1] = g[q][0] * xy + g[q][1] * ey + g[q][2] * zy;
      d[q][2] = g[q][0] * xz + g[q][1] * ez + g[q][2] * zz;
    }
    const double u = z3 ? wz[0][H] : b.bv[id + sz], v = z3 ? wz[1][H] : b.bv[id + 2 * sz];
    const double w = z3 ? wz[2][H] : b.bv[id + 3 * sz], T = z3 ? wz[3][H] : b.bv[id + 5 * sz];
    const double mu = sutherland(T);
    const double txx = mu * (4 * d[0][0] - 2 * d[1][1] - 2 * d[2][2]) * (1.0 / 3);
    const double tyy = mu * (4 * d[1][1] - 2 * d[0][0] - 2 * d[2][2]) * (1.0 / 3);
    const double tzz = mu * (4 * d[2][2] - 2 * d[0][0] - 2 * d[1][1]) * (1.0 / 3);
    const double txy = mu * (d[0][1] + d[1][0]), txz = mu * (d[0][2] + d[2][0]), tyz = mu * (d[1][2] + d[2][1]);
    constexpr double cp = kGamma * kRair / (kGamma - 1);
    const double kc = mu * (cp / sc.prandtl);
    const double Ex = u * txx + v * txy + w * txz + kc * d[3][0];
    const double Ey = u * txy + v * tyy + w * tyz + kc * d[3][1];
    const double Ez = u * txz + v * tyz + w * tzz + kc * d[3][2];
    double *fv = b.fv + id, *gv = b.gv + id, *hv = b.hv + id;
    fv[sz] = xx * txx + xy * txy + xz * txz;
    fv[2 * sz] = xx * txy + xy * tyy + xz * tyz;
    fv[3 * sz] = xx * txz + xy * tyz + xz * tzz;
    fv[4 * sz] = xx * Ex + xy * Ey + xz * Ez;
    gv[sz] = ex * txx + ey * txy + ez * txz;
    gv[2 * sz] = ex * txy + ey * tyy + ez * tyz;
    gv[3 * sz] = ex * txz + ey * tyz + ez * tzz;
    gv[4 * sz] = ex * Ex + ey * Ey + ez * Ez;
    hv[sz] = zx * txx + zy * txy + zz * txz;
    hv[2 * sz] = zx * txy + zy * tyy + zz * tyz;
    hv[3 * sz] = zx * txz + zy * tyz + zz * tzz;
    hv[4 * sz] = zx * Ex + zy * Ey + zz * Ez;
  }
}

// One work item = one (face node, ghost layer g) pair.  On i faces the ghost layers of a node are
// contiguous in memory, so g varies fastest across lanes there; on j and k faces the face nodes vary
// fastest (along i).  The g == 1 item also averages the coincident periodic face nodes.  Ghost copies
// read nodes at least one layer inside the opposite face, which no item of the same launch writes.
struct FaceItem {
  int i, j, k, g;
};

__device__ __forceinline__ bool face_item(const DevBlock &b, const DevFace &f, FaceItem &it) {
  const long n0 = f.end[0] - f.start[0] + 1, n1 = f.end[1] - f.start[1] + 1, n2 = f.end[2] - f.start[2] + 1;
  const long npt = n0 * n1 * n2;
  const long item = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= npt * b.g) return false;
  long pt;
  if (f.face == 0) {
    it.g = (int)(item % b.g) + 1;
    pt = item / b.g;
  } else {
    it.g = (int)(item / npt) + 1;
    pt = item % npt;
  }
  it.i = f.start[0] + (int)(pt % n0);
  it.j = f.start[1] + (int)((pt / n0) % n1);
  it.k = f.start[2] + (int)(pt / (n0 * n1));
  return true;
}

// periodic_exchange task 1 (DataCommunication.cu:628-683): average coincident face nodes, copy ghosts of
// the collocated viscous fluxes (components 1..nv-1; component 0 of the viscous fluxes is never used).
__global__ __launch_bounds__(256) void k_viscous_exchange(DevBlock b, DevFace f, int nv) {
  FaceItem it;
  if (!face_item(b, f, it)) return;
  int o[3] = {it.i, it.j, it.k};
  const int ext[3] = {b.mx, b.my, b.mz};
  o[f.face] = f.direction < 0 ? ext[f.face] - 1 : 0;
  int dv[3] = {0, 0, 0};
  dv[f.face] = f.direction;
  const long sz = b.sz;
  double *A[3] = {b.fv, b.gv, b.hv};
  if (it.g == 1) {
    const long id = nid(b, it.i, it.j, it.k), io = nid(b, o[0], o[1], o[2]);
    for (int l = 1; l < nv; ++l)
      for (int a = 0; a < 3; ++a) {
        const double av = 0.5 * (A[a][id + l * sz] + A[a][io + l * sz]);
        A[a][id + l * sz] = av;
        A[a][io + l * sz] = av;
      }
  }
  const int g = it.g;
  const long ig = nid(b, it.i + g * dv[0], it.j + g * dv[1], it.k + g * dv[2]);
  const long is = nid(b, o[0] + g * dv[0], o[1] + g * dv[1], o[2] + g * dv[2]);
  for (int l = 1; l < nv; ++l)
    for (int a = 0; a < 3; ++a) A[a][ig + l * sz] = A[a][is + l * sz];
}

// compute_entropy_fix_delta<Air> (InviscidScheme.cu:269-304) on the interior plus one layer.
__global__ __launch_bounds__(256) void k_entropy_fix(DevBlock b, Scheme sc) {
  const int i = blockIdx.x * 64 + threadIdx.x - 1, j = blockIdx.y * 4 + threadIdx.y - 1;
  const int k = b.dim == 3 ? (int)blockIdx.z - 1 : 0;
  if (i > b.mx || j > b.my) return;
  const long id = nid(b, i, j, k), sz = b.sz;
  const double *M = b.met + id;
  const double u = b.bv[id + sz], v = b.bv[id + 2 * sz], w = b.bv[id + 3 * sz];
  const double U = fabs(u * M[0] + v * M[sz] + w * M[2 * sz]);
  const double V = fabs(u * M[3 * sz] + v * M[4 * sz] + w * M[5 * sz]);
  const double W = fabs(u * M[6 * sz] + v * M[7 * sz] + w * M[8 * sz]);
  const double gx = sqrt(M[0] * M[0] + M[sz] * M[sz] + M[2 * sz] * M[2 * sz]);
  const double ge = sqrt(M[3 * sz] * M[3 * sz] + M[4 * sz] * M[4 * sz] + M[5 * sz] * M[5 * sz]);
  const double gz = sqrt(M[6 * sz] * M[6 * sz] + M[7 * sz] * M[7 * sz] + M[8 * sz] * M[8 * sz]);
  const double c = sqrt(kGamma * b.bv[id + 4 * sz] / b.bv[id]);
  b.efd[id] = b.dim == 2 ? sc.entropy_fix_factor * (U + V + c * 0.5 * (gx + ge))
                         : sc.entropy_fix_factor * (U + V + W + c * (gx + ge + gz) / 3.0);
}

// ---------------------------------------------------------------- residual + RK stage
__device__ __forceinline__ void cons_from_prim(const DevBlock &b, long id) {
  const long sz = b.sz;
  const double rho = b.bv[id], u = b.bv[id + sz], v = b.bv[id + 2 * sz], w = b.bv[id + 3 * sz];
  b.cv[id] = rho;
  b.cv[id + sz] = rho * u;
  b.cv[id + 2 * sz] = rho * v;
  b.cv[id + 3 * sz] = rho * w;
  b.cv[id + 4 * sz] = 0.5 * (u * u + v * v + w * w) * rho + b.bv[id + 4 * sz] / (kGamma - 1);
}

__constant__ double kRkA[3] = {1.0, 0.75, 1.0 / 3.0}; // SSPRK3 (RK.cuh:11-15)
__constant__ double kRkB[3] = {0.0, 0.25, 2.0 / 3.0};
__constant__ double kRkC[3] = {1.0, 0.25, 2.0 / 3.0};

// dq = J (dFv/dxi + dGv/deta + dHv/dzeta) - sum_dir (F(i+1/2) - F(i-1/2)), then
// update_cv_and_bv_rk<Air> (RK.cuh:224-280).  Stage 0 also saves U^n into qn.
template<int ORDER, bool RHS_ONLY>
__global__ __launch_bounds__(256) void k_update(DevBlock b, int rk, const double *__restrict__ dtp, TileGrid tg,
                                                int kz) {
  Tile t;
  if (!xcd_tile(tg, t)) return;
  const int i = t.i, j = t.j;
  if (i >= b.mx || j >= b.my) return;
  const int k0 = t.z * kz, k1 = min(k0 + kz, b.mz);
  const long sz = b.sz, d1 = b.d1, id0 = nid(b, i, j, k0);
  const bool z3 = b.dim == 3;
  // hv (components 1..4) on planes k - HB .. k + HA: the face flux below k and at k (ORDER 2), or the
  // collocated stencil (ORDER 4, 8)
  constexpr int HB = ORDER == 2 ? 1 : ORDER / 2, HA = ORDER == 2 ? 0 : ORDER / 2, W = HB + HA + 1;
  double wz[4][W];
  double fzm[5]; // fz on plane k - 1
  if (z3) {
    if constexpr (ORDER > 0) {
#pragma unroll
      for (int l = 0; l < 4; ++l)
#pragma unroll
        for (int m = 1; m < W; ++m) wz[l][m] = b.hv[id0 + (m - 1 - HB) * d1 + (l + 1) * sz];
    }
#pragma unroll
    for (int l = 0; l < 5; ++l) fzm[l] = b.fz[id0 - d1 + l * sz];
  }
  const double dt = RHS_ONLY ? 0.0 : dtp[0];
  for (int k = k0; k < k1; ++k) {
    const long id = id0 + (long)(k - k0) * d1;
    if constexpr (ORDER > 0) {
      if (z3) {
#pragma unroll
        for (int l = 0; l < 4; ++l) {
#pragma unroll
          for (int m = 0; m < W - 1; ++m) wz[l][m] = wz[l][m + 1];
          wz[l][W - 1] = b.hv[id + HA * d1 + (l + 1) * sz];
        }
      }
    }
    double r[5];
    r[0] = 0;
    if constexpr (ORDER == 2) {
      // face fluxes (compute_dFv_dx / dGv_dy / dHv_dz, ViscousScheme.cu:75-128)
#pragma unroll
      for (int l = 1; l < 5; ++l) {
        double s = b.fv[id + l * sz] - b.fv[id - 1 + l * sz];
        s += b.gv[id + l * sz] - b.gv[id - b.d2 + l * sz];
        if (z3) s += wz[l - 1][1] - wz[l - 1][0];
        r[l] = s;
      }
    } else if constexpr (ORDER > 0) {
      const double J = b.jac[id];
      const bool zc = z3 && dd_closure<ORDER>(k, b.mz, b.bt[4], b.bt[5]);
#pragma unroll
      for (int l = 1; l < 5; ++l) {
        double s = dd<ORDER>(b.fv + l * sz, id, 1, i, b.mx, b.bt[0], b.bt[1]) +
                   dd<ORDER>(b.gv + l * sz, id, b.d2, j, b.my, b.bt[2], b.bt[3]);
        if (z3) s += zc ? dd<ORDER>(b.hv + l * sz, id, d1, k, b.mz, b.bt[4], b.bt[5]) : dd_mid<ORDER>(wz[l - 1]);
        r[l] = s * J;
      }
    } else {
#pragma unroll
      for (int l = 1; l < 5; ++l) r[l] = 0;
    }
#pragma unroll
    for (int l = 0; l < 5; ++l) r[l] -= b.fx[id + l * sz] - b.fx[id - 1 + l * sz];
#pragma unroll
    for (int l = 0; l < 5; ++l) r[l] -= b.fy[id + l * sz] - b.fy[id - b.d2 + l * sz];
    if (z3) {
#pragma unroll
      for (int l = 0; l < 5; ++l) {
        const double f = b.fz[id + l * sz];
        r[l] -= f - fzm[l];
        fzm[l] = f;
      }
    }
    if (b.sponge_sigma) {
      // sponge layer: relaxation towards the running mean, -sigma (U - U_mean) per unit volume
      const long iq = ((long)k * b.my + j) * b.mx + i;
      const double sg = b.sponge_sigma[iq] * b.jac[id];
#pragma unroll
      for (int l = 0; l < 5; ++l) r[l] -= sg * (b.cv[id + l * sz] - b.sponge_mean[iq + l * b.n]);
    }
    if constexpr (RHS_ONLY) {
      const long iq = ((long)k * b.my + j) * b.mx + i;
#pragma unroll
      for (int l = 0; l < 5; ++l) b.dq[iq + l * b.n] = r[l];
      continue;
    }
    const double dj = dt * wenodev::rcp(b.jac[id]);
    double q[5];
    if (rk == 0) {
#pragma unroll
      for (int l = 0; l < 5; ++l) {
        const double u0 = b.cv[id + l * sz];
        b.qn[id + l * sz] = u0;
        q[l] = kRkA[0] * u0 + kRkB[0] * u0 + kRkC[0] * dj * r[l];
      }
    } else {
#pragma unroll
      for (int l = 0; l < 5; ++l)
        q[l] = kRkA[rk] * b.qn[id + l * sz] + kRkB[rk] * b.cv[id + l * sz] + kRkC[rk] * dj * r[l];
    }
    if (b.mz == 1) q[3] = 0;
#pragma unroll
    for (int l = 0; l < 5; ++l) b.cv[id + l * sz] = q[l];
    const double di = wenodev::rcp(q[0]);
    const double u = q[1] * di, v = q[2] * di, w = q[3] * di;
    const double p = (kGamma - 1) * (q[4] - 0.5 * q[0] * (u * u + v * v + w * w));
    b.bv[id] = q[0];
    b.bv[id + sz] = u;
    b.bv[id + 2 * sz] = v;
    b.bv[id + 3 * sz] = w;
    b.bv[id + 4 * sz] = p;
    b.bv[id + 5 * sz] = p * kMwAir * di * (1.0 / kRu);
  }
}

// ---------------------------------------------------------------- boundary conditions
// apply_outflow without back pressure (BoundCond.cu:2065-2188) and apply_periodic (BoundCond.cu:3033-3083).
__global__ __launch_bounds__(256) void k_boundary(DevBlock b, DevFace f) {
  FaceItem it;
  if (!face_item(b, f, it)) return;
  int dv[3] = {0, 0, 0};
  dv[f.face] = f.direction;
  const long id = nid(b, it.i, it.j, it.k), sz = b.sz;
  const int g = it.g;
  const long ig = nid(b, it.i + g * dv[0], it.j + g * dv[1], it.k + g * dv[2]);
  if (f.type == 8) {
    int o[3] = {it.i, it.j, it.k};
    const int ext[3] = {b.mx, b.my, b.mz};
    o[f.face] = f.direction < 0 ? ext[f.face] - 1 : 0;
    if (g == 1) {
      const long io = nid(b, o[0], o[1], o[2]);
      for (int l = 0; l < 6; ++l) {
        const double a = 0.5 * (b.bv[id + l * sz] + b.bv[io + l * sz]);
        b.bv[id + l * sz] = a;
        b.bv[io + l * sz] = a;
      }
      cons_from_prim(b, id);
    }
    const long is = nid(b, o[0] + g * dv[0], o[1] + g * dv[1], o[2] + g * dv[2]);
    for (int l = 0; l < 6; ++l) b.bv[ig + l * sz] = b.bv[is + l * sz];
  } else { // zero-gradient outflow
    for (int l = 0; l < 6; ++l) b.bv[ig + l * sz] = b.bv[id + l * sz];
  }
  cons_from_prim(b, ig);
}

// One node of a block-to-block face: its index in b and its partner's index in the target block.
__device__ __forceinline__ bool inner_item(const DevInner &f, int *idx, int *it) {
  const long n0 = f.n_point[0], n1 = f.n_point[1], n2 = f.n_point[2];
  const long item = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= n0 * n1 * n2) return false;
  const int n[3] = {(int)(item % n0), (int)((item / n0) % n1), (int)(item / (n0 * n1))};
  int d[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    d[c] = f.loop_dir[c] * n[c];
    idx[c] = f.start[c] + d[c];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) it[c] = f.target_start[c] + f.target_loop_dir[c] * d[f.src_tar[c]];
  return true;
}

// inner_communication<Air> (DataCommunication.cu:36-103): the coincident layer takes the average of both
// blocks (or the partner's state where only this side moves, as at a wall corner), the ghost layers the
// partner's interior state; cv follows bv.
__global__ __launch_bounds__(256) void k_inner_exchange(DevBlock b, DevBlock t, DevInner f) {
  int idx[3], it[3];
  if (!inner_item(f, idx, it)) return;
  const long id = nid(b, idx[0], idx[1], idx[2]), jt = nid(t, it[0], it[1], it[2]);
  const long sz = b.sz, tz = t.sz;
  if (idx[f.face] == f.face_layer) {
    const double us = sqrt(b.bv[id + sz] * b.bv[id + sz] + b.bv[id + 2 * sz] * b.bv[id + 2 * sz] +
                           b.bv[id + 3 * sz] * b.bv[id + 3 * sz]);
    if (us > 1e-15) {
      const double ut = sqrt(t.bv[jt + tz] * t.bv[jt + tz] + t.bv[jt + 2 * tz] * t.bv[jt + 2 * tz] +
                             t.bv[jt + 3 * tz] * t.bv[jt + 3 * tz]);
      if (ut < 1e-15) {
        for (int l = 0; l < 6; ++l) b.bv[id + l * sz] = t.bv[jt + l * tz];
      } else {
        for (int l = 0; l < 6; ++l) {
          const double a = 0.5 * (t.bv[jt + l * tz] + b.bv[id + l * sz]);
          b.bv[id + l * sz] = a;
          t.bv[jt + l * tz] = a;
        }
      }
    }
  } else {
    for (int l = 0; l < 6; ++l) b.bv[id + l * sz] = t.bv[jt + l * tz];
  }
  cons_from_prim(b, id);
}

// inner_exchange task 1 (DataCommunication.cu:458-495): collocated viscous fluxes (components 1..nv-1).  These
// are contravariant fluxes along each block's own index directions, so the partner's flux along direction
// vmap[a], signed by vsign[a], is this block's flux along a (the reference copies direction a to direction a,
// which is the same thing when both blocks' index directions agree).
__global__ __launch_bounds__(256) void k_inner_viscous_exchange(DevBlock b, DevBlock t, DevInner f, int nv) {
  int idx[3], it[3];
  if (!inner_item(f, idx, it)) return;
  const long id = nid(b, idx[0], idx[1], idx[2]), jt = nid(t, it[0], it[1], it[2]);
  const long sz = b.sz, tz = t.sz;
  double *A[3] = {b.fv, b.gv, b.hv}, *T[3] = {t.fv, t.gv, t.hv};
  const bool coincident = idx[f.face] == f.face_layer;
  for (int a = 0; a < 3; ++a) {
    double *P = T[f.vmap[a]];
    const double sg = f.vsign[a];
    for (int l = 1; l < nv; ++l) {
      if (coincident) {
        const double v = 0.5 * (A[a][id + l * sz] + sg * P[jt + l * tz]);
        A[a][id + l * sz] = v;
        P[jt + l * tz] = sg * v;
      } else {
        A[a][id + l * sz] = sg * P[jt + l * tz];
      }
    }
  }
}

__global__ void k_cv_from_bv(DevBlock b) {
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= b.sz) return;
  cons_from_prim(b, id);
}

// ---------------------------------------------------------------- CFL time step
// local_time_step<Air> (TimeAdvanceFunc.cuh:28-115) fused with a workgroup minimum.
__global__ __launch_bounds__(256) void k_local_dt(DevBlock b, Scheme sc, double *part) {
  const unsigned pn = blockIdx.x * 256u + threadIdx.y * 64u + threadIdx.x;
  const int j = (int)(pn / (unsigned)b.mx), i = (int)(pn - (unsigned)j * (unsigned)b.mx), k = blockIdx.z;
  double dt = 1e+6;
  if (j < b.my) {
    const long id = nid(b, i, j, k), sz = b.sz;
    const double *M = b.met + id;
    const double gx = sqrt(M[0] * M[0] + M[sz] * M[sz] + M[2 * sz] * M[2 * sz]);
    const double ge = sqrt(M[3 * sz] * M[3 * sz] + M[4 * sz] * M[4 * sz] + M[5 * sz] * M[5 * sz]);
    const double gz = sqrt(M[6 * sz] * M[6 * sz] + M[7 * sz] * M[7 * sz] + M[8 * sz] * M[8 * sz]);
    const double u = b.bv[id + sz], v = b.bv[id + 2 * sz], w = b.bv[id + 3 * sz];
    const double T = b.bv[id + 5 * sz];
    const double c = sqrt(kGamma * kRair * T);
    double s = fmax(fabs(u * M[0] + v * M[sz] + w * M[2 * sz]) + c * gx,
                    fabs(u * M[3 * sz] + v * M[4 * sz] + w * M[5 * sz]) + c * ge);
    if (b.dim == 3) s = fmax(s, fabs(u * M[6 * sz] + v * M[7 * sz] + w * M[8 * sz]) + c * gz);
    if (sc.viscous_order > 0) {
      double ml = fmax(gx, ge);
      if (b.dim == 3) ml = fmax(ml, gz);
      const double dv = fmax(kGamma, 4.0 / 3.0) / b.bv[id] * (sutherland(T) / sc.prandtl);
      s = fmax(s, ml * ml * dv);
    }
    dt = sc.cfl / s;
  }
  // wave64 minimum via cross-lane shuffles, then across the 4 waves through LDS
  for (int off = 32; off > 0; off >>= 1) dt = fmin(dt, __shfl_xor(dt, off, 64));
  __shared__ double wmin[4];
  if (threadIdx.x == 0) wmin[threadIdx.y] = dt;
  __syncthreads();
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    const long p = ((long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    part[p] = fmin(fmin(wmin[0], wmin[1]), fmin(wmin[2], wmin[3]));
  }
}

__global__ __launch_bounds__(1024) void k_dt_reduce(const double *part, long n, double *dt, const double *time,
                                                    double t_end, int clip) {
  double m = 1e+6;
  for (long p = threadIdx.x; p < n; p += blockDim.x) m = fmin(m, part[p]);
  for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off, 64));
  __shared__ double wmin[16];
  const int wave = threadIdx.x / 64;
  if ((threadIdx.x & 63) == 0) wmin[wave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = wmin[0];
    for (int q = 1; q < (int)(blockDim.x / 64); ++q) r = fmin(r, wmin[q]);
    if (clip && time[0] + r > t_end) r = t_end - time[0];
    dt[0] = r;
  }
}

__global__ void k_time_advance(double *time, const double *dt) { time[0] += dt[0]; }

// one workgroup per 256 consecutive nodes of a plane (row-major), grid.z planes; the per-workgroup minima of the time
// step are indexed by workgroup in this grid
dim3 node_grid(const DevBlock &b) { return dim3((unsigned)((b.mx * b.my + 255) / 256), 1, b.mz); }

// Planes per column of the z-marching kernels.
int march_kz() {
  static const int kz = [] {
    const char *e = std::getenv("COREFL_KZ");
    const int v = e ? std::atoi(e) : 8;
    return v < 1 ? 1 : v > 256 ? 256 : v;
  }();
  return kz;
}

// Work-groups of the z-marching kernels: 256 consecutive (i, j) columns of a plane, kz planes each.
TileGrid march_tiles(const DevBlock &b, int kz) {
  return plane_grid(b.mx, b.my, (b.mz + kz - 1) / kz, 0, 0);
}

// compute_shock_sensor (FieldOperation.cu:74-179) at the block's nodes, air and mixtures alike (it reads bv and the
// metrics only): Ducros (0, the default): 0 where the fourth-order central divergence of the velocity is positive,
// else div^2 / (div^2 + |curl|^2 + eps); Jameson (1): sum over directions of |p+ + p- - 2p| / (p+ + p- + 2p);
// density/pressure jump (2): the largest relative forward jump of density or pressure.
__global__ __launch_bounds__(256) void k_shock_sensor(DevBlock b, int type, double eps) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= b.mx || j >= b.my) return;
  const long id = nid(b, i, j, k), sz = b.sz;
  const long st[3] = {1, b.d2, b.d1};
  double s;
  if (type == 1) {
    const double p2 = 2 * b.bv[id + 4 * sz];
    s = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double pp = b.bv[id + st[d] + 4 * sz] + b.bv[id - st[d] + 4 * sz];
      s += fabs(pp - p2) / (pp + p2);
    }
  } else if (type == 2) {
    const double r = b.bv[id], p = b.bv[id + 4 * sz];
    s = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) s = fmax(s, fabs(b.bv[id + st[d]] - r) / r);
#pragma unroll
    for (int d = 0; d < 3; ++d) s = fmax(s, fabs(b.bv[id + st[d] + 4 * sz] - p) / p);
  } else {
    double D[3][3]; // d(velocity c)/d(x_a)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double g[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double *f = b.bv + (1 + c) * sz + id;
        g[d] = (1.0 / 12) * f[-2 * st[d]] - (2.0 / 3) * f[-st[d]] + (2.0 / 3) * f[st[d]] - (1.0 / 12) * f[2 * st[d]];
      }
#pragma unroll
      for (int a = 0; a < 3; ++a)
        D[c][a] = g[0] * b.met[id + a * sz] + g[1] * b.met[id + (3 + a) * sz] + g[2] * b.met[id + (6 + a) * sz];
    }
    double div = D[0][0] + D[1][1] + D[2][2];
    if (div > 0) {
      s = 0;
    } else {
      div *= div;
      const double cx = D[2][1] - D[1][2], cy = D[0][2] - D[2][0], cz = D[1][0] - D[0][1];
      s = div / (div + cx * cx + cy * cy + cz * cz + eps);
    }
  }
  b.sensor[id] = s;
}

} // namespace

hipError_t launch_cv_from_bv_all(const DevBlock &b, hipStream_t s) {
  k_cv_from_bv<<<(unsigned)((b.sz + 255) / 256), 256, 0, s>>>(b);
  return hipGetLastError();
}

static unsigned face_blocks(const DevBlock &b, const DevFace &f) {
  const long n = (long)(f.end[0] - f.start[0] + 1) * (f.end[1] - f.start[1] + 1) * (f.end[2] - f.start[2] + 1) * b.g;
  return (unsigned)((n + 255) / 256);
}

hipError_t launch_boundary(const DevBlock &b, const DevFace &f, hipStream_t s) {
  k_boundary<<<face_blocks(b, f), 256, 0, s>>>(b, f);
  return hipGetLastError();
}

hipError_t launch_viscous_flux(const DevBlock &b, const Scheme &sc, hipStream_t s) {
  const dim3 tpb(64, 4, 1);
  if (sc.viscous_order == 2) {
    // one launch per face direction, over the faces -1/2 .. m-1/2 along it
    const TileGrid tx = plane_grid(b.mx + 1, b.my, b.mz, 0, 0);
    const TileGrid ty = plane_grid(b.mx, b.my + 1, b.mz, 0, 0);
    k_viscous_face<0><<<tile_blocks(tx), tpb, 0, s>>>(b, sc, tx);
    k_viscous_face<1><<<tile_blocks(ty), tpb, 0, s>>>(b, sc, ty);
    if (b.dim == 3) {
      const TileGrid tz = plane_grid(b.mx, b.my, b.mz + 1, 2, 0);
      k_viscous_face<2><<<tile_blocks(tz), tpb, 0, s>>>(b, sc, tz);
    }
    return hipGetLastError();
  }
  const int kz = march_kz();
  const TileGrid tm = march_tiles(b, kz);
  if (sc.viscous_order == 8) k_viscous_flux<8><<<tile_blocks(tm), tpb, 0, s>>>(b, sc, tm, kz);
  else k_viscous_flux<4><<<tile_blocks(tm), tpb, 0, s>>>(b, sc, tm, kz);
  return hipGetLastError();
}

hipError_t launch_viscous_exchange(const DevBlock &b, const DevFace &f, hipStream_t s, int nv) {
  k_viscous_exchange<<<face_blocks(b, f), 256, 0, s>>>(b, f, nv);
  return hipGetLastError();
}

static unsigned inner_blocks(const DevInner &f) {
  const long n = (long)f.n_point[0] * f.n_point[1] * f.n_point[2];
  return (unsigned)((n + 255) / 256);
}

hipError_t launch_inner_exchange(const DevBlock &b, const DevBlock &t, const DevInner &f, hipStream_t s) {
  k_inner_exchange<<<inner_blocks(f), 256, 0, s>>>(b, t, f);
  return hipGetLastError();
}

hipError_t launch_inner_viscous_exchange(const DevBlock &b, const DevBlock &t, const DevInner &f, hipStream_t s,
                                         int nv) {
  k_inner_viscous_exchange<<<inner_blocks(f), 256, 0, s>>>(b, t, f, nv);
  return hipGetLastError();
}

hipError_t launch_entropy_fix(const DevBlock &b, const Scheme &sc, hipStream_t s) {
  const dim3 grid((b.mx + 2 + 63) / 64, (b.my + 2 + 3) / 4, b.dim == 3 ? b.mz + 2 : 1);
  k_entropy_fix<<<grid, dim3(64, 4, 1), 0, s>>>(b, sc);
  return hipGetLastError();
}

template<int DIR, int R>
static void dispatch_rs(const DevBlock &b, const Scheme &sc, dim3 grid, double *F, hipStream_t s, int z0) {
  const dim3 tpb(64, 4, 1);
  const TileGrid tg = plane_grid((int)grid.x, (int)grid.y, (int)grid.z, DIR == 2 ? 2 : 0, z0);
  const unsigned nb = tile_blocks(tg);
  switch (sc.inviscid_scheme) {
    case 1: k_inviscid<DIR, R, 1><<<nb, tpb, 0, s>>>(b, sc, F, tg); break;
    case 2: k_inviscid<DIR, R, 2><<<nb, tpb, 0, s>>>(b, sc, F, tg); break;
    case 4: k_inviscid<DIR, R, 4><<<nb, tpb, 0, s>>>(b, sc, F, tg); break;
    default: k_inviscid<DIR, R, 3><<<nb, tpb, 0, s>>>(b, sc, F, tg); break;
  }
}

template<int DIR, int REC>
static void dispatch_pv(const DevBlock &b, const Scheme &sc, dim3 grid, double *F, hipStream_t s, int z0) {
  const dim3 tpb(64, 4, 1);
  const TileGrid tg = plane_grid((int)grid.x, (int)grid.y, (int)grid.z, DIR == 2 ? 2 : 0, z0);
  const unsigned nb = tile_blocks(tg);
  switch (sc.inviscid_scheme) {
    case 1: k_inviscid_pv<DIR, REC, 1><<<nb, tpb, 0, s>>>(b, sc, F, tg); break;
    case 2: k_inviscid_pv<DIR, REC, 2><<<nb, tpb, 0, s>>>(b, sc, F, tg); break;
    case 4: k_inviscid_pv<DIR, REC, 4><<<nb, tpb, 0, s>>>(b, sc, F, tg); break;
    default: k_inviscid_pv<DIR, REC, 3><<<nb, tpb, 0, s>>>(b, sc, F, tg); break;
  }
}

template<int DIR>
static void dispatch_pv_rec(const DevBlock &b, const Scheme &sc, dim3 grid, double *F, hipStream_t s, int z0) {
  if (sc.reconstruction == 1) dispatch_pv<DIR, 1>(b, sc, grid, F, s, z0);
  else if (sc.reconstruction == 2) dispatch_pv<DIR, 2>(b, sc, grid, F, s, z0);
  else dispatch_pv<DIR, 3>(b, sc, grid, F, s, z0);
}

template<int DIR>
static void dispatch_fs(const DevBlock &b, const Scheme &sc, dim3 grid, double *F, hipStream_t s, int z0) {
  const dim3 tpb(64, 4, 1);
  const TileGrid tg = plane_grid((int)grid.x, (int)grid.y, (int)grid.z, DIR == 2 ? 2 : 0, z0);
  const unsigned nb = tile_blocks(tg);
  switch (sc.inviscid_scheme) {
    case 51: k_fsweno<DIR, 3, false><<<nb, tpb, 0, s>>>(b, sc, F, tg); break;
    case 52: k_fsweno<DIR, 3, true><<<nb, tpb, 0, s>>>(b, sc, F, tg); break;
    case 71: k_fsweno<DIR, 4, false><<<nb, tpb, 0, s>>>(b, sc, F, tg); break;
    default: k_fsweno<DIR, 4, true><<<nb, tpb, 0, s>>>(b, sc, F, tg); break;
  }
}

hipError_t launch_inviscid_flux(const DevBlock &b, const Scheme &sc, int dir, hipStream_t s, int kb, int ke) {
  // the interfaces of the node planes [kb, ke): F(i+1/2) stored at node i for i = -1 .. mx-1 (x), likewise in y; in
  // z the interfaces k + 1/2 of planes kb .. ke-1, and also -1/2 when kb = 0 (tile z = k + 1)
  const int me = ke < b.mz ? ke : b.mz;
  dim3 grid; // nodes along i and j, planes along k
  double *F;
  int z0 = kb;
  if (dir == 0) {
    grid = dim3(b.mx + 1, b.my, me - kb);
    F = b.fx;
  } else if (dir == 1) {
    grid = dim3(b.mx, b.my + 1, me - kb);
    F = b.fy;
  } else {
    z0 = kb == 0 ? 0 : kb + 1;
    grid = dim3(b.mx, b.my, me + 1 - z0);
    F = b.fz;
  }
  if ((int)grid.z <= 0) return hipSuccess;
  if (sc.inviscid_scheme > 50) {
    if (dir == 0) dispatch_fs<0>(b, sc, grid, F, s, z0);
    else if (dir == 1) dispatch_fs<1>(b, sc, grid, F, s, z0);
    else dispatch_fs<2>(b, sc, grid, F, s, z0);
    return hipGetLastError();
  }
  if (sc.reconstruction >= 1 && sc.reconstruction <= 3) {
    if (dir == 0) dispatch_pv_rec<0>(b, sc, grid, F, s, z0);
    else if (dir == 1) dispatch_pv_rec<1>(b, sc, grid, F, s, z0);
    else dispatch_pv_rec<2>(b, sc, grid, F, s, z0);
    return hipGetLastError();
  }
  const bool w7 = sc.reconstruction == 5;
  if (dir == 0) w7 ? dispatch_rs<0, 4>(b, sc, grid, F, s, z0) : dispatch_rs<0, 3>(b, sc, grid, F, s, z0);
  else if (dir == 1) w7 ? dispatch_rs<1, 4>(b, sc, grid, F, s, z0) : dispatch_rs<1, 3>(b, sc, grid, F, s, z0);
  else w7 ? dispatch_rs<2, 4>(b, sc, grid, F, s, z0) : dispatch_rs<2, 3>(b, sc, grid, F, s, z0);
  return hipGetLastError();
}

hipError_t launch_update(const DevBlock &b, const Scheme &sc, int rk, const double *dt_dev, bool rhs_only,
                         hipStream_t s, int kb, int ke) {
  const dim3 tpb(64, 4, 1);
  const int kz = march_kz();
  TileGrid tg = march_tiles(b, kz);
  // the node planes [kb, ke) (kb a multiple of the march length)
  const int me = ke < b.mz ? ke : b.mz;
  tg.z0 = kb / kz;
  tg.gz = (me + kz - 1) / kz - tg.z0;
  if (tg.gz <= 0) return hipSuccess;
  const unsigned grid = tile_blocks(tg);
  const int o = sc.viscous_order;
  if (rhs_only) {
    if (o == 8) k_update<8, true><<<grid, tpb, 0, s>>>(b, rk, dt_dev, tg, kz);
    else if (o == 4) k_update<4, true><<<grid, tpb, 0, s>>>(b, rk, dt_dev, tg, kz);
    else if (o == 2) k_update<2, true><<<grid, tpb, 0, s>>>(b, rk, dt_dev, tg, kz);
    else k_update<0, true><<<grid, tpb, 0, s>>>(b, rk, dt_dev, tg, kz);
  } else {
    if (o == 8) k_update<8, false><<<grid, tpb, 0, s>>>(b, rk, dt_dev, tg, kz);
    else if (o == 4) k_update<4, false><<<grid, tpb, 0, s>>>(b, rk, dt_dev, tg, kz);
    else if (o == 2) k_update<2, false><<<grid, tpb, 0, s>>>(b, rk, dt_dev, tg, kz);
    else k_update<0, false><<<grid, tpb, 0, s>>>(b, rk, dt_dev, tg, kz);
  }
  return hipGetLastError();
}

int dt_partial_count(const DevBlock &b) {
  const dim3 g = node_grid(b);
  return (int)(g.x * g.y * g.z);
}

hipError_t launch_local_dt(const DevBlock &b, const Scheme &sc, double *part, hipStream_t s) {
  k_local_dt<<<node_grid(b), dim3(64, 4, 1), 0, s>>>(b, sc, part);
  return hipGetLastError();
}

hipError_t launch_dt_reduce(const double *part, long n, double *dt_dev, const double *time_dev, double t_end,
                            int clip, hipStream_t s) {
  k_dt_reduce<<<1, 1024, 0, s>>>(part, n, dt_dev, time_dev, t_end, clip);
  return hipGetLastError();
}

hipError_t launch_time_advance(double *time_dev, const double *dt_dev, hipStream_t s) {
  k_time_advance<<<1, 1, 0, s>>>(time_dev, dt_dev);
  return hipGetLastError();
}

hipError_t launch_shock_sensor(const DevBlock &b, const Scheme &sc, hipStream_t s) {
  if (!b.sensor) return hipSuccess;
  k_shock_sensor<<<dim3((b.mx + 63) / 64, (b.my + 3) / 4, b.mz), dim3(64, 4, 1), 0, s>>>(b, sc.shock_sensor, sc.sensor_eps);
  return hipGetLastError();
}

int march_planes() { return march_kz(); }

} // namespace corefl
