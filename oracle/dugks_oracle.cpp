// dugks_oracle.cpp -- TEST INFRASTRUCTURE, not product code.
//
// FP64 C++/OpenMP restatement of one DUGKS step exactly as the reference
// computes it, kept deliberately close to the reference's own structure:
// array-of-structures storage with a one-cell halo, a separate f-bar array,
// a ghost fill in x -> y -> z order, and every cell recomputing all six of its
// faces (so shared faces are evaluated twice, reference quirk Q5).  Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline may load it.
//
// Reference lines restated (paths relative to the reference repository):
//   lattice              src/lbm_d3q19.cu:10-21
//   coefficients         src/dugks.cu:101-125, dt from src/numerics.cu:101-105
//   f-bar+ at centres    calc_fb1, src/dugks.cu:1131-1196
//   ghost fill           extend_fb_x / extend_fb_x_wall / extend_fb_y,
//                        src/dugks.cu:1212-1273; z by exchange(),
//                        src/parallel.cu:185-234 (self-exchange = periodic)
//   trilinear faces      calc_fx_tri, src/dugks.cu:715-906; trilinear :964-978
//   gradient faces       calc_fx, src/dugks.cu:421-583
//   wall bounce-back     src/dugks.cu:807-823 (x); generalised to any axis
//   face collision       f2ruvwh + feq_s, src/dugks.cu:675-709, 1201-1207,
//                        applied at :912-943
//   update               src/dugks.cu:945-956
//   stretched axis       src/fortran/channel/src/dugks.cuf:114-121 (centres,
//                        ghost widths), :365-374 (wall extrapolation),
//                        :538-716 (weights), :833-834 (update)
//   macros               calc_macro, src/dugks.cu:1021-1064
//
// The adjoint (oracle_adjoint_step) has no reference counterpart; it is the
// reverse-mode derivative of exactly this forward step, stage by stage.
//
// Licence.  This file restates, line for line, algorithms of
// Kairzhan/DUGKS-GPU, which is distributed under the GNU General Public
// License version 3.  As a derivative of that code this file is licensed under
// the GNU General Public License, version 3 or (at your option) any later
// version: you may redistribute and/or modify it under those terms.  It is
// distributed WITHOUT ANY WARRANTY; without even the implied warranty of
// MERCHANTABILITY or FITNESS FOR A PARTICULAR PURPOSE.  See
// <https://www.gnu.org/licenses/> for the licence text.
//
// Compiled with -ffp-contract=off so every product and sum rounds as written.

#include "dugks_oracle.h"

#include <cmath>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int Q = 19;
// D3Q19 velocity ordering of src/lbm_d3q19.cu:10-21 (the restart-dump
// format depends on this order, so it is kept).
const int CV[3][Q] = {
    {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0},
    {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1},
    {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1}};
const int OPP[Q] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
const double RT = 1. / 3.;

double weight(int ip) { return ip == 0 ? 1.0 / 3.0 : (ip <= 6 ? 1.0 / 18.0 : 1.0 / 36.0); }

}  // namespace

struct oracle_ctx {
  int n[3];
  int periodic[3];
  int uniform[3];
  double width[3];
  double du[3];                 // uniform spacing
  std::vector<double> d[3];     // widths, ghost-inclusive, index 0..n+1
  std::vector<double> xc[3];    // centres, ghost-inclusive
  double dt, hdt, visc, tau;
  double bw0, bw1, bw2, fw0, fw1, fw2;
  double famb[3];
  int recon;
  bool cell_force = false;
  std::vector<double> f, fb;              // AoS with halo
  std::vector<double> fcell;              // [cell][3], halo-indexed

  size_t ncell_ext() const { return (size_t)(n[0] + 2) * (n[1] + 2) * (n[2] + 2); }
  size_t id(int i, int j, int k) const {
    return ((size_t)k * (n[1] + 2) + j) * (n[0] + 2) + i;
  }
  size_t ddf(int i, int j, int k) const {  // interior index, 1-based i,j,k
    return ((size_t)(k - 1) * n[1] + (j - 1)) * n[0] + (i - 1);
  }
};

namespace {

// Linear extrapolation weight for the wall ghost cell
// (src/fortran/channel/src/dugks.cuf:365-374; equals 2 on a uniform axis,
// the 2*fb(1)-fb(2) of extend_fb_x_wall, src/dugks.cu:1247-1248).
double ghost_q_lo(const oracle_ctx& o, int a) {
  const auto& x = o.xc[a];
  return (x[2] - x[0]) / (x[2] - x[1]);
}
double ghost_q_hi(const oracle_ctx& o, int a) {
  const auto& x = o.xc[a];
  int n = o.n[a];
  return (x[n + 1] - x[n - 1]) / (x[n] - x[n - 1]);
}

// calc_fb1 (src/dugks.cu:1131-1196)
void center_fbar(oracle_ctx& o) {
  const int nx = o.n[0], ny = o.n[1], nz = o.n[2];
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        const size_t c = o.id(i, j, k);
        const double* ff = &o.f[c * Q];
        double rr = 0, uu = 0, vv = 0, ww = 0;
        for (int ip = 0; ip < Q; ip++) {
          rr += ff[ip];
          uu += CV[0][ip] * ff[ip];
          vv += CV[1][ip] * ff[ip];
          ww += CV[2][ip] * ff[ip];
        }
        uu /= rr;
        vv /= rr;
        ww /= rr;
        double F[3];
        for (int a = 0; a < 3; a++) F[a] = o.cell_force ? o.fcell[c * 3 + a] : o.famb[a];
        uu += o.hdt * F[0];
        vv += o.hdt * F[1];
        ww += o.hdt * F[2];
        double* out = &o.fb[c * Q];
        for (int ip = 0; ip < Q; ip++) {
          double eu = (CV[0][ip] * uu + CV[1][ip] * vv + CV[2][ip] * ww) / RT;
          double uv = (uu * uu + vv * vv + ww * ww) / (RT);
          double feq = weight(ip) * rr * (1.0 + eu + 0.5 * (eu * eu - uv));
          double temp = ((CV[0][ip] - uu) * F[0] + (CV[1][ip] - vv) * F[1] +
                         (CV[2][ip] - ww) * F[2]) / RT;
          out[ip] = o.bw1 * ff[ip] + (o.bw2 + o.bw0 * temp) * feq;
        }
      }
}

// Ghost layers of an f-bar array, axis by axis in x -> y -> z order, each
// pass covering the whole ghost-inclusive extent of the other two axes so
// edges and corners come out consistent (src/dugks.cu:351-377).
void fill_ghosts(const oracle_ctx& o, double* fb) {
  for (int a = 0; a < 3; a++) {
    const int n = o.n[a];
    const int b = (a + 1) % 3, c = (a + 2) % 3;
    const double qlo = o.periodic[a] ? 0 : ghost_q_lo(o, a);
    const double qhi = o.periodic[a] ? 0 : ghost_q_hi(o, a);
    for (int s = 0; s <= o.n[c] + 1; s++)
      for (int r = 0; r <= o.n[b] + 1; r++) {
        int idx[3];
        auto at = [&](int m) {
          idx[a] = m;
          idx[b] = r;
          idx[c] = s;
          return o.id(idx[0], idx[1], idx[2]) * Q;
        };
        const size_t g0 = at(0), g1 = at(n + 1), i1 = at(1), i2 = at(2), in = at(n),
                     in1 = at(n - 1);
        for (int ip = 0; ip < Q; ip++) {
          if (o.periodic[a]) {
            fb[g0 + ip] = fb[in + ip];
            fb[g1 + ip] = fb[i1 + ip];
          } else if (o.uniform[a]) {
            fb[g0 + ip] = 2 * fb[i1 + ip] - fb[i2 + ip];
            fb[g1 + ip] = 2 * fb[in + ip] - fb[in1 + ip];
          } else {
            fb[g0 + ip] = qlo * fb[i1 + ip] + (1. - qlo) * fb[i2 + ip];
            fb[g1 + ip] = qhi * fb[in + ip] + (1. - qhi) * fb[in1 + ip];
          }
        }
      }
  }
}

// trilinear (src/dugks.cu:964-978)
inline double trilinear(double a, double b, double c, double d, double e, double f, double g,
                        double h, double mx, double my, double mz) {
  double x00 = (1 - mx) * a + mx * b;
  double x01 = (1 - mx) * c + mx * d;
  double x10 = (1 - mx) * e + mx * f;
  double x11 = (1 - mx) * g + mx * h;
  double y00 = (1 - my) * x00 + my * x01;
  double y11 = (1 - my) * x10 + my * x11;
  return (1 - mz) * y00 + mz * y11;
}

// Interpolation fraction along the face normal for the +/- face of cell m
// on axis a: fraction of the neighbour across the face
// (uniform: src/dugks.cu:768-772; stretched: dugks.cuf:541, 546).
double m_normal(const oracle_ctx& o, int a, int m, int side, int c) {
  const double h = o.hdt;
  if (o.uniform[a]) {
    const double dx = o.du[a];
    double m1 = (dx / 2. - h * c) / dx;
    return side > 0 ? m1 : 1 - m1;
  }
  const auto& d = o.d[a];
  if (side > 0) return (d[m] / 2 - h * c) / (0.5 * d[m] + 0.5 * d[m + 1]);
  return (d[m] / 2 + h * c) / (0.5 * d[m] + 0.5 * d[m - 1]);
}

// Tangential fraction toward the upwind cell m - c on axis b
// (uniform: src/dugks.cu:769-770; stretched: dugks.cuf:595, 658).
double m_tangent(const oracle_ctx& o, int b, int m, int c) {
  const double h = o.hdt;
  if (o.uniform[b]) return std::fabs((-h * c) / o.du[b]);
  const auto& d = o.d[b];
  return std::fabs((-h * c) / (0.5 * d[m] + 0.5 * d[m - c]));
}

// Six reconstructed faces of interior cell (i,j,k) from the ghost-filled
// f-bar array, trilinear scheme (src/dugks.cu:756-906).
void faces_tri(const oracle_ctx& o, const double* fb, int i, int j, int k, double ft[6][Q]) {
  const int p[3] = {i, j, k};
  auto FB = [&](int ip, int a, int b, int c) { return fb[o.id(a, b, c) * Q + ip]; };
  for (int a = 0; a < 3; a++) {
    int e[3] = {0, 0, 0};
    e[a] = 1;
    ft[2 * a + 1][0] = 0.5 * (FB(0, i, j, k) + FB(0, i + e[0], j + e[1], k + e[2]));
    ft[2 * a][0] = 0.5 * (FB(0, i, j, k) + FB(0, i - e[0], j - e[1], k - e[2]));
  }
  for (int a = 0; a < 3; a++) {
    for (int ip = 1; ip < Q; ip++) {
      for (int side = 1; side >= -1; side -= 2) {
        int off[3];
        double m[3];
        for (int b = 0; b < 3; b++) {
          if (b == a) {
            off[b] = side;
            m[b] = m_normal(o, a, p[a], side, CV[a][ip]);
          } else {
            off[b] = -CV[b][ip];
            m[b] = m_tangent(o, b, p[b], CV[b][ip]);
          }
        }
        const int ii = off[0], jj = off[1], kk = off[2];
        ft[2 * a + (side > 0 ? 1 : 0)][ip] =
            trilinear(FB(ip, i, j, k), FB(ip, i + ii, j, k), FB(ip, i, j + jj, k),
                      FB(ip, i + ii, j + jj, k), FB(ip, i, j, k + kk), FB(ip, i + ii, j, k + kk),
                      FB(ip, i, j + jj, k + kk), FB(ip, i + ii, j + jj, k + kk), m[0], m[1], m[2]);
      }
    }
  }
}

// Six faces, midpoint + central-difference gradient scheme
// (calc_fx, src/dugks.cu:459-583).  Uniform axes only.
void faces_grad(const oracle_ctx& o, const double* fb, int i, int j, int k, double ft[6][Q]) {
  auto FB = [&](int ip, int a, int b, int c) { return fb[o.id(a, b, c) * Q + ip]; };
  const double h = o.hdt, dx = o.du[0], dy = o.du[1], dz = o.du[2];
  for (int ip = 0; ip < Q; ip++) {
    const int cx = CV[0][ip], cy = CV[1][ip], cz = CV[2][ip];
    for (int side = 1; side >= -1; side -= 2) {
      // x faces
      {
        int i2 = i + side;
        double temp = 0.5 * (FB(ip, i2, j, k) + FB(ip, i, j, k));
        double gx = side > 0 ? (FB(ip, i2, j, k) - FB(ip, i, j, k)) / dx
                             : (FB(ip, i, j, k) - FB(ip, i2, j, k)) / dx;
        double aa = 0.5 * (FB(ip, i, j + 1, k) + FB(ip, i2, j + 1, k));
        double bb = 0.5 * (FB(ip, i, j - 1, k) + FB(ip, i2, j - 1, k));
        double gy = (aa - bb) / (2 * dy);
        aa = 0.5 * (FB(ip, i, j, k + 1) + FB(ip, i2, j, k + 1));
        bb = 0.5 * (FB(ip, i, j, k - 1) + FB(ip, i2, j, k - 1));
        double gz = (aa - bb) / (2 * dz);
        ft[side > 0 ? 1 : 0][ip] = temp - (h) * (cx * gx + cy * gy + cz * gz);
      }
      // y faces
      {
        int j2 = j + side;
        double temp = 0.5 * (FB(ip, i, j2, k) + FB(ip, i, j, k));
        double aa = 0.5 * (FB(ip, i + 1, j, k) + FB(ip, i + 1, j2, k));
        double bb = 0.5 * (FB(ip, i - 1, j, k) + FB(ip, i - 1, j2, k));
        double gx = (aa - bb) / (2 * dx);
        double gy = side > 0 ? (FB(ip, i, j2, k) - FB(ip, i, j, k)) / dy
                             : (FB(ip, i, j, k) - FB(ip, i, j2, k)) / dy;
        aa = 0.5 * (FB(ip, i, j, k + 1) + FB(ip, i, j2, k + 1));
        bb = 0.5 * (FB(ip, i, j, k - 1) + FB(ip, i, j2, k - 1));
        double gz = (aa - bb) / (2 * dz);
        ft[side > 0 ? 3 : 2][ip] = temp - (h) * (cx * gx + cy * gy + cz * gz);
      }
      // z faces
      {
        int k2 = k + side;
        double temp = 0.5 * (FB(ip, i, j, k2) + FB(ip, i, j, k));
        double aa = 0.5 * (FB(ip, i + 1, j, k) + FB(ip, i + 1, j, k2));
        double bb = 0.5 * (FB(ip, i - 1, j, k) + FB(ip, i - 1, j, k2));
        double gx = (aa - bb) / (2 * dx);
        aa = 0.5 * (FB(ip, i, j + 1, k) + FB(ip, i, j + 1, k2));
        bb = 0.5 * (FB(ip, i, j - 1, k) + FB(ip, i, j - 1, k2));
        double gy = (aa - bb) / (2 * dy);
        double gz = side > 0 ? (FB(ip, i, j, k2) - FB(ip, i, j, k)) / dz
                             : (FB(ip, i, j, k) - FB(ip, i, j, k2)) / dz;
        ft[side > 0 ? 5 : 4][ip] = temp - (h) * (cx * gx + cy * gy + cz * gz);
      }
    }
  }
}

// Half-way bounce-back at wall faces: incoming populations (c.n_inward > 0)
// take the value of their opposite (src/dugks.cu:807-823 for x).
void bounce_back(const oracle_ctx& o, int i, int j, int k, double ft[6][Q]) {
  const int p[3] = {i, j, k};
  for (int a = 0; a < 3; a++) {
    if (o.periodic[a]) continue;
    if (p[a] == 1)
      for (int ip = 1; ip < Q; ip++)
        if (CV[a][ip] > 0) ft[2 * a][ip] = ft[2 * a][OPP[ip]];
    if (p[a] == o.n[a])
      for (int ip = 1; ip < Q; ip++)
        if (CV[a][ip] < 0) ft[2 * a + 1][ip] = ft[2 * a + 1][OPP[ip]];
  }
}

// f2ruvwh + feq_s applied to one face (src/dugks.cu:675-709, 1201-1207, 912-943)
void face_collide(const oracle_ctx& o, double* ft) {
  double r = 0, u = 0, v = 0, w = 0;
  for (int ip = 0; ip < Q; ip++) {
    r += ft[ip];
    u += CV[0][ip] * ft[ip];
    v += CV[1][ip] * ft[ip];
    w += CV[2][ip] * ft[ip];
  }
  u /= r;
  v /= r;
  w /= r;
  u += 0.5 * o.hdt * o.famb[0];
  v += 0.5 * o.hdt * o.famb[1];
  w += 0.5 * o.hdt * o.famb[2];
  for (int ip = 0; ip < Q; ip++) {
    double force1 = (o.famb[0] * (CV[0][ip] - u) + o.famb[1] * (CV[1][ip] - v) +
                     o.famb[2] * (CV[2][ip] - w)) / RT;
    double eu = (CV[0][ip] * u + CV[1][ip] * v + CV[2][ip] * w) / RT;
    double uv = (u * u + v * v + w * w) / (RT);
    double feq = weight(ip) * r * (1.0 + eu + 0.5 * (eu * eu - uv));
    ft[ip] = o.fw1 * ft[ip] + (o.fw2 + o.fw0 * force1) * feq;
  }
}

double update_coef(const oracle_ctx& o, int a, int m) {
  return o.uniform[a] ? (o.dt) / (o.du[a]) : o.dt / o.d[a][m];
}

void flux_update(oracle_ctx& o) {
  const int nx = o.n[0], ny = o.n[1], nz = o.n[2];
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        double ft[6][Q];
        if (o.recon == 0)
          faces_tri(o, o.fb.data(), i, j, k, ft);
        else
          faces_grad(o, o.fb.data(), i, j, k, ft);
        bounce_back(o, i, j, k, ft);
        for (int s = 0; s < 6; s++) face_collide(o, ft[s]);
        const double cx = update_coef(o, 0, i), cy = update_coef(o, 1, j),
                     cz = update_coef(o, 2, k);
        const size_t c = o.id(i, j, k);
        for (int ip = 0; ip < Q; ip++) {
          double u = ft[1][ip] - ft[0][ip];
          double v = ft[3][ip] - ft[2][ip];
          double w = ft[5][ip] - ft[4][ip];
          double r = o.fb[c * Q + ip] * 4. / 3. - o.f[c * Q + ip] * 1. / 3.;
          o.f[c * Q + ip] = r - cx * CV[0][ip] * u - cy * CV[1][ip] * v - cz * CV[2][ip] * w;
        }
      }
}

}  // namespace

// ---------------------------------------------------------------------------
// Adjoint of one step.  Forward stages (per step):
//   S1 centre:   fb_c = bw1 f_c + (bw2 + bw0 T_i) feq_i(rho, u),
//                u = j/rho + h F_c, T_i = (c_i - u).F_c / RT
//   S2 ghosts:   linear (copy / extrapolation)
//   S3 faces:    linear reconstruction of fb, then bounce-back (a copy)
//   S4 collide:  ft = fw1 fbar + (fw2 + fw0 T'_i) feq_i(rho_f, u_f),
//                u_f = jbar/rho_f + (h/2) F_amb, T'_i = (c_i - u_f).F_amb / RT
//   S5 update:   f' = 4/3 fb_c - 1/3 f_c - sum_a cf_a c_ia (ft_a+ - ft_a-)
// The reverse pass runs S5^T .. S1^T; S3^T and S2^T are applied by
// differentiating the very same code paths (a scatter into the ghost-
// inclusive array, then the ghost fill reversed axis by axis z -> y -> x).
// ---------------------------------------------------------------------------
namespace {

// Derivative of g(x) = sum_i lam_i * [A f_i + (B + C T_i) feq_i(rho, u)]
// with respect to the moments, where rho = sum f, u = j/rho + shift, T_i =
// (c_i - u).F / RT.  Returns dg/drho (through feq's explicit rho) and
// dg/du (vector), holding F fixed.  Also returns dg/dF (explicit F in T_i).
void feq_term_adjoint(const double lam[Q], double B, double C, double rho, const double u[3],
                      const double F[3], double* g_rho, double g_u[3], double g_F[3]) {
  double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  *g_rho = 0;
  g_u[0] = g_u[1] = g_u[2] = 0;
  g_F[0] = g_F[1] = g_F[2] = 0;
  for (int ip = 0; ip < Q; ip++) {
    const double cu = (CV[0][ip] * u[0] + CV[1][ip] * u[1] + CV[2][ip] * u[2]);
    const double eu = cu / RT;
    const double poly = 1.0 + eu + 0.5 * (eu * eu - uu / RT);
    const double feq = weight(ip) * rho * poly;
    double T = 0;
    for (int a = 0; a < 3; a++) T += (CV[a][ip] - u[a]) * F[a];
    T /= RT;
    const double coef = B + C * T;
    // d feq / d rho = w poly ; d feq / d u_a = w rho (c_a/RT + eu c_a/RT - u_a/RT)
    *g_rho += lam[ip] * coef * weight(ip) * poly;
    for (int a = 0; a < 3; a++) {
      double dfeq = weight(ip) * rho * (CV[a][ip] / RT + eu * CV[a][ip] / RT - u[a] / RT);
      double dT_du = -F[a] / RT;
      g_u[a] += lam[ip] * (coef * dfeq + C * dT_du * feq);
      g_F[a] += lam[ip] * C * ((CV[a][ip] - u[a]) / RT) * feq;
    }
  }
}

// Reverse of "rho = sum f, u = j/rho + shift": adds to lam_f.
void moments_adjoint(const double f[Q], double g_rho, const double g_u[3], double lam_f[Q]) {
  double rho = 0, j[3] = {0, 0, 0};
  for (int ip = 0; ip < Q; ip++) {
    rho += f[ip];
    for (int a = 0; a < 3; a++) j[a] += CV[a][ip] * f[ip];
  }
  // u_a = j_a / rho (+ const): du_a/df_i = (c_ia - j_a/rho)/rho
  for (int ip = 0; ip < Q; ip++) {
    double s = g_rho;
    for (int a = 0; a < 3; a++) s += g_u[a] * (CV[a][ip] - j[a] / rho) / rho;
    lam_f[ip] += s;
  }
}

}  // namespace

extern "C" void oracle_adjoint_step(const oracle_ctx* oc, const double* lam_next,
                                    double* lam_prev, double* grad_force) {
  oracle_ctx& o = const_cast<oracle_ctx&>(*oc);
  const int nx = o.n[0], ny = o.n[1], nz = o.n[2];
  // Recompute the forward centre stage into a private ghost-inclusive array.
  std::vector<double> saved_fb = o.fb;
  center_fbar(o);
  fill_ghosts(o, o.fb.data());
  std::vector<double> fb = o.fb;  // forward f-bar with ghosts
  o.fb.swap(saved_fb);

  const size_t next = o.ncell_ext();
  std::vector<double> lam(next * Q, 0.0), lam_fb(next * Q, 0.0);
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++)
        for (int ip = 0; ip < Q; ip++)
          lam[o.id(i, j, k) * Q + ip] = lam_next[o.ddf(i, j, k) * Q + ip];

  // S5^T, S4^T, S3^T per cell: scatter into lam_fb (ghost-inclusive).
  // Serial: the reconstruction transpose scatters into neighbours.
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        const size_t c = o.id(i, j, k);
        const double* L = &lam[c * Q];
        double cf[3] = {update_coef(o, 0, i), update_coef(o, 1, j), update_coef(o, 2, k)};
        // update: f' = fb*4/3 - f/3 - sum_a cf_a c_ia (ft[2a+1] - ft[2a])
        for (int ip = 0; ip < Q; ip++) lam_fb[c * Q + ip] += L[ip] * 4. / 3.;
        double lam_ft[6][Q];
        for (int a = 0; a < 3; a++)
          for (int ip = 0; ip < Q; ip++) {
            lam_ft[2 * a + 1][ip] = -cf[a] * CV[a][ip] * L[ip];
            lam_ft[2 * a][ip] = cf[a] * CV[a][ip] * L[ip];
          }
        // forward faces (pre-collision, post bounce-back)
        double ft[6][Q];
        if (o.recon == 0)
          faces_tri(o, fb.data(), i, j, k, ft);
        else
          faces_grad(o, fb.data(), i, j, k, ft);
        bounce_back(o, i, j, k, ft);
        // S4^T: collision adjoint per face
        double lam_fbar[6][Q];
        for (int s = 0; s < 6; s++) {
          double r = 0, jj[3] = {0, 0, 0};
          for (int ip = 0; ip < Q; ip++) {
            r += ft[s][ip];
            for (int a = 0; a < 3; a++) jj[a] += CV[a][ip] * ft[s][ip];
          }
          double u[3];
          for (int a = 0; a < 3; a++) u[a] = jj[a] / r + 0.5 * o.hdt * o.famb[a];
          double g_rho, g_u[3], g_F[3];
          feq_term_adjoint(lam_ft[s], o.fw2, o.fw0, r, u, o.famb, &g_rho, g_u, g_F);
          for (int ip = 0; ip < Q; ip++) lam_fbar[s][ip] = o.fw1 * lam_ft[s][ip];
          moments_adjoint(ft[s], g_rho, g_u, lam_fbar[s]);
        }
        // bounce-back^T: incoming took the value of its opposite
        const int p[3] = {i, j, k};
        for (int a = 0; a < 3; a++) {
          if (o.periodic[a]) continue;
          if (p[a] == 1)
            for (int ip = 1; ip < Q; ip++)
              if (CV[a][ip] > 0) {
                lam_fbar[2 * a][OPP[ip]] += lam_fbar[2 * a][ip];
                lam_fbar[2 * a][ip] = 0;
              }
          if (p[a] == o.n[a])
            for (int ip = 1; ip < Q; ip++)
              if (CV[a][ip] < 0) {
                lam_fbar[2 * a + 1][OPP[ip]] += lam_fbar[2 * a + 1][ip];
                lam_fbar[2 * a + 1][ip] = 0;
              }
        }
        // S3^T: reconstruction transpose.  The reconstruction is linear in
        // fb, so its transpose is obtained from unit responses: perturbing
        // one (cell, ip) input at a time would be slow, instead re-run the
        // trilinear formulas with weights attached.
        for (int s = 0; s < 6; s++) {
          const int a = s / 2, side = (s % 2) ? 1 : -1;
          int e[3] = {0, 0, 0};
          e[a] = side;
          // rest population: plain average of the two cells
          lam_fb[c * Q + 0] += 0.5 * lam_fbar[s][0];
          lam_fb[o.id(i + e[0], j + e[1], k + e[2]) * Q + 0] += 0.5 * lam_fbar[s][0];
          for (int ip = 1; ip < Q; ip++) {
            const double Ls = lam_fbar[s][ip];
            if (Ls == 0) continue;
            if (o.recon == 0) {
              int off[3];
              double m[3];
              for (int b = 0; b < 3; b++) {
                if (b == a) {
                  off[b] = side;
                  m[b] = m_normal(o, a, p[a], side, CV[a][ip]);
                } else {
                  off[b] = -CV[b][ip];
                  m[b] = m_tangent(o, b, p[b], CV[b][ip]);
                }
              }
              for (int cz = 0; cz < 2; cz++)
                for (int cy = 0; cy < 2; cy++)
                  for (int cx = 0; cx < 2; cx++) {
                    double wgt = (cx ? m[0] : 1 - m[0]) * (cy ? m[1] : 1 - m[1]) *
                                 (cz ? m[2] : 1 - m[2]);
                    size_t t = o.id(i + cx * off[0], j + cy * off[1], k + cz * off[2]);
                    lam_fb[t * Q + ip] += wgt * Ls;
                  }
            } else {
              // gradient scheme: f = temp - h (c.grad); all linear in fb
              const double h = o.hdt;
              const double dd[3] = {o.du[0], o.du[1], o.du[2]};
              int q2[3] = {i, j, k};
              q2[a] += side;
              const size_t cc = c, cn = o.id(q2[0], q2[1], q2[2]);
              // temp and normal gradient
              const double gn = -h * CV[a][ip] / dd[a] * (side > 0 ? 1 : -1);
              lam_fb[cc * Q + ip] += Ls * (0.5 - gn);
              lam_fb[cn * Q + ip] += Ls * (0.5 + gn);
              for (int b = 0; b < 3; b++) {
                if (b == a || CV[b][ip] == 0) continue;
                const double g = -h * CV[b][ip] / (2 * dd[b]) * 0.5;
                int pp[3] = {i, j, k}, pn[3] = {q2[0], q2[1], q2[2]};
                pp[b] += 1;
                pn[b] += 1;
                lam_fb[o.id(pp[0], pp[1], pp[2]) * Q + ip] += Ls * g;
                lam_fb[o.id(pn[0], pn[1], pn[2]) * Q + ip] += Ls * g;
                pp[b] -= 2;
                pn[b] -= 2;
                lam_fb[o.id(pp[0], pp[1], pp[2]) * Q + ip] -= Ls * g;
                lam_fb[o.id(pn[0], pn[1], pn[2]) * Q + ip] -= Ls * g;
              }
            }
          }
        }
      }

  // S2^T: ghost fill reversed, axis order z -> y -> x.
  for (int a = 2; a >= 0; a--) {
    const int n = o.n[a];
    const int b = (a + 1) % 3, cc = (a + 2) % 3;
    const double qlo = o.periodic[a] ? 0 : (o.uniform[a] ? 2. : ghost_q_lo(o, a));
    const double qhi = o.periodic[a] ? 0 : (o.uniform[a] ? 2. : ghost_q_hi(o, a));
    for (int s = 0; s <= o.n[cc] + 1; s++)
      for (int r = 0; r <= o.n[b] + 1; r++) {
        int idx[3];
        auto at = [&](int m) {
          idx[a] = m;
          idx[b] = r;
          idx[cc] = s;
          return o.id(idx[0], idx[1], idx[2]) * Q;
        };
        const size_t g0 = at(0), g1 = at(n + 1), i1 = at(1), i2 = at(2), in = at(n),
                     in1 = at(n - 1);
        for (int ip = 0; ip < Q; ip++) {
          // reverse of: g1 = f(i1..), then g0 = f(in..)  (forward wrote g0 first)
          const double l1 = lam_fb[g1 + ip];
          lam_fb[g1 + ip] = 0;
          const double l0 = lam_fb[g0 + ip];
          lam_fb[g0 + ip] = 0;
          if (o.periodic[a]) {
            lam_fb[i1 + ip] += l1;
            lam_fb[in + ip] += l0;
          } else {
            lam_fb[in + ip] += qhi * l1;
            lam_fb[in1 + ip] += (1. - qhi) * l1;
            lam_fb[i1 + ip] += qlo * l0;
            lam_fb[i2 + ip] += (1. - qlo) * l0;
          }
        }
      }
  }

  // S1^T per interior cell, plus the -1/3 f term of S5.
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        const size_t c = o.id(i, j, k);
        const double* ff = &o.f[c * Q];
        const double* Lb = &lam_fb[c * Q];
        double out[Q];
        for (int ip = 0; ip < Q; ip++) out[ip] = -lam[c * Q + ip] * 1. / 3. + o.bw1 * Lb[ip];
        double r = 0, jj[3] = {0, 0, 0};
        for (int ip = 0; ip < Q; ip++) {
          r += ff[ip];
          for (int a = 0; a < 3; a++) jj[a] += CV[a][ip] * ff[ip];
        }
        double F[3];
        for (int a = 0; a < 3; a++) F[a] = o.cell_force ? o.fcell[c * 3 + a] : o.famb[a];
        double u[3];
        for (int a = 0; a < 3; a++) u[a] = jj[a] / r + o.hdt * F[a];
        double g_rho, g_u[3], g_F[3];
        feq_term_adjoint(Lb, o.bw2, o.bw0, r, u, F, &g_rho, g_u, g_F);
        moments_adjoint(ff, g_rho, g_u, out);
        const size_t cd = o.ddf(i, j, k);
        for (int ip = 0; ip < Q; ip++) lam_prev[cd * Q + ip] = out[ip];
        if (grad_force)
          for (int a = 0; a < 3; a++) grad_force[cd * 3 + a] += g_F[a] + g_u[a] * o.hdt;
      }
}

// ---------------------------------------------------------------------------
// C API
// ---------------------------------------------------------------------------
extern "C" {

oracle_ctx* oracle_create(int nx, int ny, int nz, const int periodic[3], const double width[3],
                          const double* sx, const double* sy, const double* sz, double dt,
                          double visc, const double famb[3], int recon) {
  oracle_ctx* o = new oracle_ctx;
  o->n[0] = nx;
  o->n[1] = ny;
  o->n[2] = nz;
  const double* sp[3] = {sx, sy, sz};
  for (int a = 0; a < 3; a++) {
    o->periodic[a] = periodic[a];
    o->width[a] = width[a];
    o->uniform[a] = sp[a] == nullptr;
    const int n = o->n[a];
    o->du[a] = width[a] / n;
    o->d[a].assign(n + 2, o->du[a]);
    o->xc[a].assign(n + 2, 0.0);
    if (!o->uniform[a])
      for (int m = 1; m <= n; m++) o->d[a][m] = sp[a][m - 1];
    // centres from face positions starting at 0 (dugks.cuf:114-121)
    double xf = 0;
    for (int m = 1; m <= n; m++) {
      o->xc[a][m] = xf + 0.5 * o->d[a][m];
      xf += o->d[a][m];
    }
    if (o->periodic[a]) {
      o->d[a][0] = o->d[a][n];
      o->d[a][n + 1] = o->d[a][1];
      o->xc[a][0] = o->xc[a][1] - 0.5 * (o->d[a][0] + o->d[a][1]);
      o->xc[a][n + 1] = o->xc[a][n] + 0.5 * (o->d[a][n] + o->d[a][n + 1]);
    } else {
      o->d[a][0] = o->d[a][1];
      o->d[a][n + 1] = o->d[a][n];
      o->xc[a][0] = -o->xc[a][1];
      o->xc[a][n + 1] = 2 * xf - o->xc[a][n];
    }
  }
  o->dt = dt;
  o->visc = visc;
  o->tau = visc / RT;
  o->hdt = dt / 2.;
  const double tau = o->tau, hdt = o->hdt;
  o->bw0 = (3. * tau * hdt) / (2. * tau + dt);
  o->bw1 = (2. * tau - hdt) / (2. * tau + dt);
  o->bw2 = 3. * hdt / (2. * tau + dt);
  o->fw0 = tau * hdt / (2. * tau + hdt);
  o->fw1 = 2. * tau / (2. * tau + hdt);
  o->fw2 = hdt / (2. * tau + hdt);
  for (int a = 0; a < 3; a++) o->famb[a] = famb[a];
  o->recon = recon;
  o->f.assign(o->ncell_ext() * Q, 0.0);
  o->fb.assign(o->ncell_ext() * Q, 0.0);
  return o;
}

void oracle_destroy(oracle_ctx* o) { delete o; }

void oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

void oracle_set_f(oracle_ctx* o, const double* f) {
  for (int k = 1; k <= o->n[2]; k++)
    for (int j = 1; j <= o->n[1]; j++)
      for (int i = 1; i <= o->n[0]; i++)
        std::memcpy(&o->f[o->id(i, j, k) * Q], &f[o->ddf(i, j, k) * Q], Q * sizeof(double));
}

void oracle_get_f(const oracle_ctx* o, double* f) {
  for (int k = 1; k <= o->n[2]; k++)
    for (int j = 1; j <= o->n[1]; j++)
      for (int i = 1; i <= o->n[0]; i++)
        std::memcpy(&f[o->ddf(i, j, k) * Q], &o->f[o->id(i, j, k) * Q], Q * sizeof(double));
}

void oracle_set_cell_force(oracle_ctx* o, const double* fx, const double* fy, const double* fz) {
  o->fcell.assign(o->ncell_ext() * 3, 0.0);
  for (int k = 1; k <= o->n[2]; k++)
    for (int j = 1; j <= o->n[1]; j++)
      for (int i = 1; i <= o->n[0]; i++) {
        size_t c = o->id(i, j, k), d = o->ddf(i, j, k);
        o->fcell[c * 3 + 0] = fx[d];
        o->fcell[c * 3 + 1] = fy[d];
        o->fcell[c * 3 + 2] = fz[d];
      }
  o->cell_force = true;
}

void oracle_set_ambient_force(oracle_ctx* o, const double famb[3]) {
  for (int a = 0; a < 3; a++) o->famb[a] = famb[a];
}

void oracle_step(oracle_ctx* o, int nsteps) {
  for (int s = 0; s < nsteps; s++) {
    center_fbar(*o);
    fill_ghosts(*o, o->fb.data());
    flux_update(*o);
  }
}

void oracle_macro(const oracle_ctx* o, double* rho, double* ux, double* uy, double* uz) {
  for (int k = 1; k <= o->n[2]; k++)
    for (int j = 1; j <= o->n[1]; j++)
      for (int i = 1; i <= o->n[0]; i++) {
        const double* ff = &o->f[o->id(i, j, k) * Q];
        double rr = 0, uu = 0, vv = 0, ww = 0;
        for (int ip = 0; ip < Q; ip++) {
          rr += ff[ip];
          uu += CV[0][ip] * ff[ip];
          vv += CV[1][ip] * ff[ip];
          ww += CV[2][ip] * ff[ip];
        }
        uu /= rr;
        vv /= rr;
        ww /= rr;
        size_t d = o->ddf(i, j, k);
        rho[d] = rr;
        ux[d] = uu + o->famb[0] * o->hdt;
        uy[d] = vv + o->famb[1] * o->hdt;
        uz[d] = ww + o->famb[2] * o->hdt;
      }
}

void oracle_coefficients(const oracle_ctx* o, double out[8]) {
  out[0] = o->tau;
  out[1] = o->hdt;
  out[2] = o->bw0;
  out[3] = o->bw1;
  out[4] = o->bw2;
  out[5] = o->fw0;
  out[6] = o->fw1;
  out[7] = o->fw2;
}

void oracle_reconstruct(const oracle_ctx* o, const double* fbar, double* faces) {
  std::vector<double> fb(o->ncell_ext() * Q, 0.0);
  for (int k = 1; k <= o->n[2]; k++)
    for (int j = 1; j <= o->n[1]; j++)
      for (int i = 1; i <= o->n[0]; i++)
        std::memcpy(&fb[o->id(i, j, k) * Q], &fbar[o->ddf(i, j, k) * Q], Q * sizeof(double));
  fill_ghosts(*o, fb.data());
  for (int k = 1; k <= o->n[2]; k++)
    for (int j = 1; j <= o->n[1]; j++)
      for (int i = 1; i <= o->n[0]; i++) {
        double ft[6][Q];
        if (o->recon == 0)
          faces_tri(*o, fb.data(), i, j, k, ft);
        else
          faces_grad(*o, fb.data(), i, j, k, ft);
        double* out = &faces[o->ddf(i, j, k) * 6 * Q];
        for (int s = 0; s < 6; s++)
          for (int ip = 0; ip < Q; ip++) out[s * Q + ip] = ft[s][ip];
      }
}

}  // extern "C"

/* Streak forcing of the reference, calc_perturb_force (src/dugks.cu:1086-1126),
 * on one rank (zoffset 0): the total centre force of every interior cell at
 * time step `step`; cells outside the forced band 3 < i < nx/4 + 2 (1-based i)
 * carry the ambient force.  d[] are the uniform cell widths (n->dx, dy, dz). */
extern "C" void oracle_perturb_force(int nx, int ny, int nz, const double width[3], const double d[3],
                                     double dt, double visc, int step, const double famb[3], double* fx,
                                     double* fy, double* fz) {
  const double two_pi = 2.0 * 3.14159265358979323846;
  const double beta = 3, Tpd = 2000, gamma = 2, Re_tau = 180;
  const double ustar = 2 * Re_tau * visc / width[0];
  const double FF = 2 * std::pow(ustar, 2) / width[0];
  const double A0 = 5 * 40.0 / nx * beta * std::sin(two_pi * step * dt / Tpd);
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        const size_t c = ((size_t)(k - 1) * ny + (j - 1)) * nx + (i - 1);
        if (3 < i && i < nx / 4 + 2) {
          const double xx = two_pi * (i - 3) * d[0] / (width[0] / 4.);
          const double yy = two_pi * (j - 0.5) * d[1] / width[1];
          const double zz = two_pi * ((k - 0.5) * d[2]) / width[2];
          fx[c] = 0.5 * FF * A0 * width[0] / 4 * (1 - std::cos(xx)) * std::cos(beta * yy) * std::cos(gamma * zz);
          fy[c] = 0.5 * FF * A0 * width[1] / gamma * std::sin(xx) * std::cos(beta * yy) * std::sin(gamma * zz);
          fz[c] = FF * (1 - A0 * width[2] / beta * std::sin(xx) * std::sin(beta * yy) * std::cos(gamma * zz));
        } else {
          fx[c] = famb[0];
          fy[c] = famb[1];
          fz[c] = famb[2];
        }
      }
}

