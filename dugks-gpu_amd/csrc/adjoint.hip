// Discrete adjoint of one DUGKS step (no counterpart in the reference).
//
// Forward stages of a step (src/dugks.cu:322-391 of the reference):
//   S1 centre    fb_c = bw1 f_c + (bw2 + bw0 T_i) feq_i(rho, u), u = j/rho + h F_c
//   S2 ghosts    periodic copy / wall extrapolation (linear)
//   S3 faces     trilinear reconstruction of fb (linear) + wall bounce-back
//   S4 collide   ft = fw1 fbar + (fw2 + fw0 T'_i) feq_i(rho_f, u_f), u_f = j_f/rho_f + (h/2) F
//   S5 update    f' = 4/3 fb_c - 1/3 f_c - sum_a (dt/dx_a) c_a (ft_a+ - ft_a-)
// Given lam' = dJ/df', the adjoint returns lam = (df'/df)^T lam' and
// accumulates dJ/dF_c.  Work split:
//   1. f-bar+ with ghosts is recomputed (the two-pass kernels);
//   2. k_adj_faces: every face once: S3 forward, S4^T driven by the face's
//      cotangent lam_ft_q = c_qa (dt/dx_R lam'_q(R) - dt/dx_L lam'_q(L)) of
//      its crossing populations.  The collision cotangent of every population
//      is lam_fbar_q = fw1 lam_ft_q + A + B.c_q; after the transpose of the
//      wall bounce-back the face stores the cotangent of its 19
//      reconstructed values (in the arithmetic type);
//   3. the gather: S3^T of those values over the faces whose stencil
//      contains each cell (the wall ghost extrapolation folded in), then S5^T
//      and S1^T.  k_adj_cells_z (32 x 8 column tiles)
//      stages each plane of face cotangents in LDS and marches along z;
//      k_adj_cells (every other case) runs one thread per cell over face
//      arrays with periodic ghost copies (k_face_ghosts).
// Deterministic (no atomics).  At most one wall axis.
#include "device_common.h"

namespace dugks {

namespace {

#include "fused_common.h"

__device__ __forceinline__ long long ext_index(const KParams& P, int m0, int m1, int m2) {
  return ((long long)m2 * P.e[1] + m1) * P.e[0] + m0;
}

// Wrap a ghost-inclusive index of a periodic axis into 1..n.
__device__ __forceinline__ int wrap(const KParams& P, int a, int m) {
  if (!P.periodic[a]) return m;
  const int n = P.n[a];
  return m < 1 ? m + n : (m > n ? m - n : m);
}

// lam' of population ip at a ghost-inclusive position (0 outside the domain).
__device__ __forceinline__ double lam_at(const KParams& P, const double* lam, int ip, int m0, int m1, int m2) {
  int m[3] = {wrap(P, 0, m0), wrap(P, 1, m1), wrap(P, 2, m2)};
  for (int a = 0; a < 3; a++)
    if (m[a] < 1 || m[a] > P.n[a]) return 0.0;
  const long long c = ((long long)(m[2] - 1) * P.n[1] + (m[1] - 1)) * P.n[0] + (m[0] - 1);
  return lam[ip * P.N + c];
}

// ---- kernel 2: one thread per face of axis A ----
template <class Real, int A>
__global__ void __launch_bounds__(256) k_adj_faces(KParams P, const Real* __restrict__ fbx,
                                                   const double* __restrict__ lam, Real* __restrict__ faces) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= P.eN) return;
  int m[3];
  m[0] = (int)(e % P.e[0]);
  m[1] = (int)((e / P.e[0]) % P.e[1]);
  m[2] = (int)(e / ((long long)P.e[0] * P.e[1]));
  for (int b = 0; b < 3; b++) {
    if (b == A) continue;
    if (m[b] < 1 || m[b] > P.n[b]) return;
  }
  const int nA = P.n[A];
  if (P.periodic[A] ? (m[A] < 1 || m[A] > nA) : (m[A] < 0 || m[A] > nA)) return;
  const long long st[3] = {1, P.e[0], (long long)P.e[0] * P.e[1]};
  const double* tab = P.tab[A] + m[A] * 9;
  // S3: reconstruction as the +A face of the left cell
  Real ft[Q];
  ft[0] = Real(0.5) * (fbx[e] + fbx[e + st[A]]);
  static_for<Q - 1>([&](auto I) {
    constexpr int ip = decltype(I)::value + 1;
    constexpr int c[3] = {CX[ip], CY[ip], CZ[ip]};
    const Real* p = fbx + ip * P.eN + e;
    long long off[3];
    Real w[3];
#pragma unroll
    for (int b = 0; b < 3; b++) {
      if (b == A) {
        off[b] = st[b];
        w[b] = (Real)tab[c[b] + 1];
      } else {
        off[b] = -c[b] * st[b];
        w[b] = c[b] == 0 ? Real(0) : (Real)P.tab[b][m[b] * 9 + (c[b] < 0 ? 6 : 7)];
      }
    }
    Real r = 0;
#pragma unroll
    for (int corner = 0; corner < 8; corner++) {
      const int bx = corner & 1, by = (corner >> 1) & 1, bz = (corner >> 2) & 1;
      const Real wt = (bx ? w[0] : Real(1) - w[0]) * (by ? w[1] : Real(1) - w[1]) * (bz ? w[2] : Real(1) - w[2]);
      if (wt != Real(0)) r += wt * p[bx * off[0] + by * off[1] + bz * off[2]];
    }
    ft[ip] = r;
  });
  // wall bounce-back (src/dugks.cu:807-823)
  const bool low = !P.periodic[A] && m[A] == 0, high = !P.periodic[A] && m[A] == nA;
  if (low || high) {
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      if constexpr (cv(A, ip) != 0) {
        if ((low && cv(A, ip) > 0) || (high && cv(A, ip) < 0)) ft[ip] = ft[OPP[ip]];
      }
    });
  }
  Real rho = 0, j[3] = {0, 0, 0};
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    rho += ft[ip];
    j[0] += Real(CX[ip]) * ft[ip];
    j[1] += Real(CY[ip]) * ft[ip];
    j[2] += Real(CZ[ip]) * ft[ip];
  });
  Real F[3], u[3], ut[3];
  for (int a = 0; a < 3; a++) {
    F[a] = (Real)P.famb_dev[a];
    ut[a] = j[a] / rho;
    u[a] = ut[a] + Real(0.5) * (Real)P.hdt * F[a];
  }
  // cotangent of the crossing populations (S5^T)
  int mR[3] = {m[0], m[1], m[2]};
  mR[A] += 1;
  const Real updL = (Real)tab[8], updR = (Real)P.tab[A][(m[A] + 1) * 9 + 8];
  Real lft[Q];
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    if constexpr (cv(A, ip) != 0) {
      const Real lL = (Real)lam_at(P, lam, ip, m[0], m[1], m[2]);
      const Real lR = (Real)lam_at(P, lam, ip, mR[0], mR[1], mR[2]);
      lft[ip] = Real(cv(A, ip)) * (updR * lR - updL * lL);
    } else {
      lft[ip] = 0;
    }
  });
  Real g_rho, g_u[3], g_F[3];
  feq_adjoint<Real>(lft, (Real)P.fw2, (Real)P.fw0, rho, u, F, g_rho, g_u, g_F, Crossing<A>{});
  const Real invr = Real(1) / rho;
  const Real Acoef = g_rho - (g_u[0] * ut[0] + g_u[1] * ut[1] + g_u[2] * ut[2]) * invr;
  const Real Bx = g_u[0] * invr, By = g_u[1] * invr, Bz = g_u[2] * invr;
  // collision cotangent of each (post-bounce) face value
  Real lpost[Q];
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    lpost[ip] = (Real)P.fw1 * lft[ip] + Acoef + Bx * Real(CX[ip]) + By * Real(CY[ip]) + Bz * Real(CZ[ip]);
  });
  // transpose of the bounce-back: an incoming population took its
  // opposite's reconstructed value
  Real* out = faces + (long long)A * Q * P.eN + e;
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    Real v = lpost[ip];
    if constexpr (cv(A, ip) != 0) {
      if (low || high) {
        const bool in = (low && cv(A, ip) > 0) || (high && cv(A, ip) < 0);
        v = in ? Real(0) : v + lpost[OPP[ip]];
      }
    }
    out[(long long)ip * P.eN] = v;
  });
}

// Face-slot index of a periodic axis' ghost position in the weight table
// (the table rows of ghost positions are only used for wall axes).
__device__ __forceinline__ int table_slot(const KParams& P, int a, int m) {
  if (!P.periodic[a]) return m;
  return m == 0 ? P.n[a] : (m == P.n[a] + 1 ? 1 : m);
}

// S3^T at ghost-inclusive position t (linear index et) for population ip: the
// sum over the faces of axis A whose reconstruction stencil contains t of
// the face cotangent times the stencil weight.  Face arrays hold periodic
// ghost copies (k_face_ghosts) and zeros at the ghost rows of wall axes, so
// every stencil face of an interior position is a plain offset from et.  A
// wall ghost position (GHOST) can reach outside the array and is checked.
template <class Real, int A, int ip, bool GHOST>
__device__ __forceinline__ Real gather_axis(const KParams& P, const Real* faces, const int t[3], long long et) {
  constexpr int c[3] = {CX[ip], CY[ip], CZ[ip]};
  constexpr int B1 = (A + 1) % 3, B2 = (A + 2) % 3;
  const long long st[3] = {1, P.e[0], (long long)P.e[0] * P.e[1]};
  const Real* fa = faces + ((long long)A * Q + ip) * P.eN + et;
  Real acc = 0;
  // normal: t is the left cell of the face in slot t_A, the right cell of slot t_A - 1
#pragma unroll
  for (int side = 0; side < 2; side++) {
    const int sl = t[A] - side;
    if (GHOST && (sl < 0 || sl > P.n[A])) continue;
    Real wn;
    if constexpr (ip == 0) {
      wn = Real(0.5);
    } else {
      const Real wR = tabv<Real>(P, A, table_slot(P, A, sl) * 9 + c[A] + 1);  // toward the right cell
      wn = side == 0 ? Real(1) - wR : wR;
    }
#pragma unroll
    for (int s1 = 0; s1 < 2; s1++) {
      if (s1 == 1 && (c[B1] == 0 || ip == 0)) continue;
      const int f1 = t[B1] + s1 * c[B1];
      if (GHOST && (f1 < 0 || f1 > P.n[B1] + 1)) continue;
      Real w1 = 1;
      if (c[B1] != 0 && ip != 0) {
        const Real m = tabv<Real>(P, B1, table_slot(P, B1, f1) * 9 + (c[B1] < 0 ? 6 : 7));
        w1 = s1 ? m : Real(1) - m;
      }
#pragma unroll
      for (int s2 = 0; s2 < 2; s2++) {
        if (s2 == 1 && (c[B2] == 0 || ip == 0)) continue;
        const int f2 = t[B2] + s2 * c[B2];
        if (GHOST && (f2 < 0 || f2 > P.n[B2] + 1)) continue;
        Real w2 = 1;
        if (c[B2] != 0 && ip != 0) {
          const Real m = tabv<Real>(P, B2, table_slot(P, B2, f2) * 9 + (c[B2] < 0 ? 6 : 7));
          w2 = s2 ? m : Real(1) - m;
        }
        const long long off = -side * st[A] + s1 * c[B1] * st[B1] + s2 * c[B2] * st[B2];
        acc += wn * w1 * w2 * fa[off];
      }
    }
  }
  return acc;
}

template <class Real, int ip, bool GHOST>
__device__ __forceinline__ Real gather(const KParams& P, const Real* faces, const int t[3]) {
  const long long et = ext_index(P, t[0], t[1], t[2]);
  return gather_axis<Real, 0, ip, GHOST>(P, faces, t, et) + gather_axis<Real, 1, ip, GHOST>(P, faces, t, et) +
         gather_axis<Real, 2, ip, GHOST>(P, faces, t, et);
}

// Periodic ghost copies of the face arrays along axis b (slot 0 <- n,
// n+1 <- 1), for all three face axes and 19 populations; axes in the order
// x, y, z so that edges and corners are consistent.
template <class Real>
__global__ void __launch_bounds__(256) k_face_ghosts(KParams P, Real* __restrict__ faces, int b) {
  const int b1 = (b + 1) % 3, b2 = (b + 2) % 3;
  const long long plane = (long long)P.e[b1] * P.e[b2];
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= plane * 3 * Q) return;
  const long long arr = id / plane, pos = id % plane;
  int m[3];
  m[b1] = (int)(pos % P.e[b1]);
  m[b2] = (int)(pos / P.e[b1]);
  Real* fa = faces + arr * P.eN;
  const int n = P.n[b];
  m[b] = n;
  const Real vn = fa[ext_index(P, m[0], m[1], m[2])];
  m[b] = 1;
  const Real v1 = fa[ext_index(P, m[0], m[1], m[2])];
  m[b] = 0;
  fa[ext_index(P, m[0], m[1], m[2])] = vn;
  m[b] = n + 1;
  fa[ext_index(P, m[0], m[1], m[2])] = v1;
}

// ---- kernel 3: one thread per interior cell ----
constexpr int kCX = 32, kCY = 8, kCZ = 1;  // cell tile of a 256-thread block
template <class Real, class Store>
__global__ void __launch_bounds__(256) k_adj_cells(KParams P, const Store* __restrict__ f,
                                                   const double* __restrict__ lam, const Real* __restrict__ faces,
                                                   double* __restrict__ lam_prev, double* __restrict__ grad,
                                                   int wall_axis) {
  // kCX x kCY x kCZ cell tiles, dealt to the XCDs in contiguous runs of
  // tiles (x fastest) so that neighbouring tiles share an L2: each face
  // value is gathered by cells of up to eight tiles
  const int gx = (P.n[0] + kCX - 1) / kCX, gy = (P.n[1] + kCY - 1) / kCY, gz = (P.n[2] + kCZ - 1) / kCZ;
  const int ntiles = gx * gy * gz;
  const int b = blockIdx.x;
  const int tt = (ntiles % 8 == 0) ? (b % 8) * (ntiles / 8) + b / 8 : b;
  const int ci = (tt % gx) * kCX + (int)threadIdx.x % kCX;
  const int cj = ((tt / gx) % gy) * kCY + ((int)threadIdx.x / kCX) % kCY;
  const int ck = (tt / (gx * gy)) * kCZ + (int)threadIdx.x / (kCX * kCY);
  if (ci >= P.n[0] || cj >= P.n[1] || ck >= P.n[2]) return;
  const long long c = ((long long)ck * P.n[1] + cj) * P.n[0] + ci;
  const int t[3] = {ci + 1, cj + 1, ck + 1};
  Real lfb[Q];
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    Real v = Real(4.0 / 3.0) * (Real)lam[ip * P.N + c] + gather<Real, ip, false>(P, faces, t);
    if (wall_axis >= 0) {
      // S2^T: the wall ghost g = q f(1) + (1-q) f(2) (and mirrored at the top)
      const int W = wall_axis, n = P.n[W];
      int g[3] = {t[0], t[1], t[2]};
      if (t[W] == 1 || t[W] == 2) {
        g[W] = 0;
        const Real q = (Real)P.qlo[W];
        v += (t[W] == 1 ? q : Real(1) - q) * gather<Real, ip, true>(P, faces, g);
      }
      if (t[W] == n || t[W] == n - 1) {
        g[W] = n + 1;
        const Real q = (Real)P.qhi[W];
        v += (t[W] == n ? q : Real(1) - q) * gather<Real, ip, true>(P, faces, g);
      }
    }
    lfb[ip] = v;
  });
  // S1^T
  Real ff[Q];
  load_cell<Real, Store>(f, c, P, ff);
  Real rho = 0, j[3] = {0, 0, 0};
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    rho += ff[ip];
    j[0] += Real(CX[ip]) * ff[ip];
    j[1] += Real(CY[ip]) * ff[ip];
    j[2] += Real(CZ[ip]) * ff[ip];
  });
  Real F[3], u[3], ut[3];
  for (int a = 0; a < 3; a++) {
    F[a] = (Real)P.famb_dev[a];
    if (P.fcell) F[a] += (Real)P.fcell[a * P.N + c];
    ut[a] = j[a] / rho;
    u[a] = ut[a] + (Real)P.hdt * F[a];
  }
  Real g_rho, g_u[3], g_F[3];
  feq_adjoint<Real>(lfb, (Real)P.bw2, (Real)P.bw0, rho, u, F, g_rho, g_u, g_F, AllPops{});
  const Real invr = Real(1) / rho;
  const Real Acoef = g_rho - (g_u[0] * ut[0] + g_u[1] * ut[1] + g_u[2] * ut[2]) * invr;
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    const Real v = -(Real)lam[ip * P.N + c] * Real(1.0 / 3.0) + (Real)P.bw1 * lfb[ip] + Acoef +
                   (g_u[0] * Real(CX[ip]) + g_u[1] * Real(CY[ip]) + g_u[2] * Real(CZ[ip])) * invr;
    lam_prev[ip * P.N + c] = (double)v;
  });
  if (grad) {
    for (int a = 0; a < 3; a++) grad[a * P.N + c] += (double)(g_F[a] + (Real)P.hdt * g_u[a]);
  }
}

// ---- kernel 3, z-marching form (32 x 8 column tiles) ----
// A workgroup owns a kGX x kGY tile of cell columns and marches along z over
// a chunk of planes.  Per plane it stages the x-, y- and z-face cotangents of
// the tile and a one-deep ring ((kGY+2) x (kGX+2) ghost-inclusive positions,
// periodic indices wrapped, so no ghost copies are needed) in LDS, with the
// next plane's loads in flight in registers; each cell then gathers its
// transposed in-plane lerps from LDS.  The transposed z-lerps are pushed
// into the f-bar+ cotangents of planes p-1, p and p+1 held in registers:
// the x- and y-faces of plane p (z-lerped toward p - c_z in the step) and
// the z-face p+1/2 (normal lerp between p and p+1).  Wall ghosts are folded
// at the in-plane level: a cell beside a wall adds q (or 1 - q) times the
// in-plane gather at the ghost position, which takes the same z transposes
// as its own.  Each face value is read from HBM about 1.3 times (the ring)
// instead of about 2.5 times.  The centre-stage adjoint needs only rho and j
// of the state: with the state stored in the arithmetic type it reads them
// from the face pass (16 or 32 bytes a cell) rather than the 19 values of f.
constexpr int kGX = 32, kGY = 8;
constexpr int kGW = kGX + 2;          // staged row length
constexpr int kGP = (kGY + 2) * kGW;  // staged positions per array and population
constexpr int kGT = kGX * kGY;        // lanes: one per cell column
static_assert(kGP <= 2 * kGT, "two staged positions per lane at most");

// mom: the region of rho, j behind the face records that holds fin's
// (storage in the arithmetic type).  FLOW (such storage, with FP32 lam and
// arithmetic or FP64 throughout; a sweep through the flow-rate
// controller): lam' still lacks the bulk velocity's cotangent ubar (dV/V)
// (c_ia - u_a)/rho of f(n+1), whose rho, j the previous pass left in the
// other region; it is added as lam' is used.
template <class Real, class Store, class Lam, bool FLOW = false>
__global__ void __launch_bounds__(kGT) k_adj_cells_z(KParams P, const Store* __restrict__ fin,
                                                     const Lam* __restrict__ lam, const Real* __restrict__ faces,
                                                     Lam* __restrict__ lam_prev, double* __restrict__ grad, int TZ,
                                                     int mom, const double* __restrict__ ubar, double inv_vol) {
  __shared__ Real S[3][Q][kGP];   // face cotangents of plane p: x-faces, y-faces, z-faces p+1/2
  __shared__ Real WXs[5][kGW];    // per staged column: wR for c = -1, 0, +1, then the shift weights
  __shared__ Real WYs[5][kGY + 2];  // per staged row

  const int nx = P.n[0], ny = P.n[1], nz = P.n[2];
  const int t = threadIdx.x;
  // XCD-aware tile order (workgroup b runs on XCD b % 8)
  const int gxn = nx / kGX, gyn = ny / kGY;
  const int ntiles = gxn * gyn * (int)gridDim.z;
  const int b = blockIdx.x + gxn * (blockIdx.y + gyn * blockIdx.z);
  const int tt = (ntiles % 8 == 0) ? (b % 8) * (ntiles / 8) + b / 8 : b;
  const int bx = tt % gxn, by = (tt / gxn) % gyn, bz = tt / (gxn * gyn);
  const int i0 = bx * kGX, j0 = by * kGY;
  const int z0 = bz * TZ, zend = min(z0 + TZ, nz);
  const int ex = P.e[0], ey = P.e[1];

  // the lane's cell, and its staged (ghost-inclusive, tile-local) position
  const int r = t / kGX, c = t % kGX;
  const int R = r + 1, Cc = c + 1;
  const unsigned plane = (unsigned)nx * (unsigned)ny;
  const unsigned colcell = (unsigned)(j0 + r) * (unsigned)nx + (unsigned)(i0 + c);

  // weights of the staged columns and rows (table_slot maps the periodic ghosts)
  for (int u = t; u < 5 * kGW + 5 * (kGY + 2); u += kGT) {
    const bool isx = u < 5 * kGW;
    const int v = isx ? u : u - 5 * kGW;
    const int n = isx ? kGW : kGY + 2;
    const int k = v / n, q = v % n;
    const int a = isx ? 0 : 1;
    const int slot = table_slot(P, a, (isx ? i0 : j0) + q);
    const Real w = tabv<Real>(P, a, slot * 9 + (k < 3 ? k : k + 3));
    if (isx)
      WXs[k][q] = w;
    else
      WYs[k][q] = w;
  }
  // z is periodic and uniform
  const Real mz6 = tabv<Real>(P, 2, 9 + 6), mz7 = tabv<Real>(P, 2, 9 + 7);
  const Real wRz[3] = {tabv<Real>(P, 2, 9 + 0), tabv<Real>(P, 2, 9 + 1), tabv<Real>(P, 2, 9 + 2)};

  Real Famb[3];
  Famb[0] = (Real)P.famb_dev[0];
  Famb[1] = (Real)P.famb_dev[1];
  Famb[2] = (Real)P.famb_dev[2];
  const Real hdt = (Real)P.hdt, bw0 = (Real)P.bw0, bw1 = (Real)P.bw1, bw2 = (Real)P.bw2;

  // staged positions of this lane: t and t + kGT
  unsigned exy[2];
  bool okp[2], okx[2], oky[2];
  for (int h = 0; h < 2; h++) {
    const int pp = t + h * kGT;
    okp[h] = pp < kGP;
    const int Rs = pp / kGW, Cs = pp % kGW;
    int x = i0 + Cs, y = j0 + Rs;
    if (P.periodic[0]) x = x == 0 ? nx : (x == nx + 1 ? 1 : x);
    if (P.periodic[1]) y = y == 0 ? ny : (y == ny + 1 ? 1 : y);
    exy[h] = (unsigned)y * (unsigned)ex + (unsigned)x;
    okx[h] = okp[h] && Cs <= kGX;  // x-face slots i0 .. i0+kGX
    oky[h] = okp[h] && Rs <= kGY;  // y-face slots j0 .. j0+kGY
  }
  // compact face records (launch_adjoint_faces_fused, compact_record)
  const unsigned pop_bytes = (unsigned)(P.eN * (long long)sizeof(Real));
  const RawBuf fx(faces, kCompact * pop_bytes), fy(faces + kCompact * P.eN, kCompact * pop_bytes),
      fz(faces + 2 * kCompact * P.eN, kCompact * pop_bytes);
  // wall faces among the staged slots: x-face slot 0 / nx, y-face slot 0 / ny
  bool wlo[2][2], whi[2][2];
  for (int h = 0; h < 2; h++) {
    const int pp = t + h * kGT;
    const int xs = i0 + pp % kGW, ys = j0 + pp / kGW;
    wlo[h][0] = !P.periodic[0] && xs == 0;
    whi[h][0] = !P.periodic[0] && xs == nx;
    wlo[h][1] = !P.periodic[1] && ys == 0;
    whi[h][1] = !P.periodic[1] && ys == ny;
  }
  const Real fw1 = (Real)P.fw1;
  Real pre[2][3][kCompact];
  // face records of plane p (ghost-inclusive z slot p + 1) into registers
  auto stage_load = [&](int p) {
    int zs = p + 1;
    zs = zs == 0 ? nz : (zs == nz + 1 ? 1 : zs);
    const unsigned zoff = (unsigned)zs * (unsigned)ey * (unsigned)ex;
    for (int h = 0; h < 2; h++) {
      const unsigned vb = (zoff + exy[h]) * (unsigned)sizeof(Real);
      const unsigned vx = okx[h] ? vb : kOutOfRange, vy = oky[h] ? vb : kOutOfRange, vz = okp[h] ? vb : kOutOfRange;
      static_for<kCompact>([&](auto K) {
        constexpr int k = decltype(K)::value;
        pre[h][0][k] = fx.ld<Real>(vx, k * pop_bytes);
        pre[h][1][k] = fy.ld<Real>(vy, k * pop_bytes);
        pre[h][2][k] = fz.ld<Real>(vz, k * pop_bytes);
      });
    }
  };
  // the 19 reconstructed-value cotangents of each staged face into LDS (a
  // missing face has an all-zero record and expands to zeros)
  auto expand_store = [&](auto AI, int h, int pp) {
    constexpr int A = decltype(AI)::value;
    Real rec[kCompact], lamf[Q];
    static_for<kCompact>([&](auto K) { rec[decltype(K)::value] = pre[h][A][decltype(K)::value]; });
    const bool lo = A < 2 && wlo[h][A < 2 ? A : 0], hi = A < 2 && whi[h][A < 2 ? A : 0];
    expand_record<Real, A>(rec, lo, hi, lamf);
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      S[A][ip][pp] = lamf[ip];
    });
  };
  auto stage_store = [&]() {
    for (int h = 0; h < 2; h++) {
      if (!okp[h]) continue;
      const int pp = t + h * kGT;
      expand_store(std::integral_constant<int, 0>{}, h, pp);
      expand_store(std::integral_constant<int, 1>{}, h, pp);
      expand_store(std::integral_constant<int, 2>{}, h, pp);
    }
  };

  // In-plane transposed lerps at staged position (Rp, Cp) for population ip:
  // the cotangent of Z(p) (x- and y-faces) and of Zn(p+1/2) (z-faces).
  // GUARD: the position is a wall ghost, whose stencil reaches outside the
  // staged region (nothing is there: 0).
  auto sget = [&](int A, int ip, int Rp, int Cp, bool guard) -> Real {
    if (guard && (Rp < 0 || Rp > kGY + 1 || Cp < 0 || Cp > kGX + 1)) return Real(0);
    return S[A][ip][Rp * kGW + Cp];
  };
  auto inplane = [&](auto I, int Rp, int Cp, bool guard, Real& hz_out) -> Real {
    constexpr int ip = decltype(I)::value;
    constexpr int cx = CX[ip], cy = CY[ip];
    // weights at a guarded position's outside neighbours only multiply 0;
    // clamp their index so that they stay table values
    auto WX = [&](int k, int Cq) { return WXs[k][guard ? min(max(Cq, 0), kGW - 1) : Cq]; };
    auto WY = [&](int k, int Rq) { return WYs[k][guard ? min(max(Rq, 0), kGY + 1) : Rq]; };
    if constexpr (ip == 0) {
      hz_out = sget(2, 0, Rp, Cp, guard);
      return Real(0.5) * (sget(0, 0, Rp, Cp, guard) + sget(0, 0, Rp, Cp - 1, guard)) +
             Real(0.5) * (sget(1, 0, Rp, Cp, guard) + sget(1, 0, Rp - 1, Cp, guard));
    } else {
      // x-faces: slot Cp (this column is the left cell), Cp-1 (right cell);
      // rows Rp and Rp+c_y
      const Real a0 = Real(1) - WX(cx + 1, Cp), a1 = WX(cx + 1, Cp - 1);
      Real hx = a0 * sget(0, ip, Rp, Cp, guard) + a1 * sget(0, ip, Rp, Cp - 1, guard);
      if constexpr (cy != 0) {
        const Real m0 = Real(1) - WY(cy < 0 ? 3 : 4, Rp), m1 = WY(cy < 0 ? 3 : 4, Rp + cy);
        hx = m0 * hx + m1 * (a0 * sget(0, ip, Rp + cy, Cp, guard) + a1 * sget(0, ip, Rp + cy, Cp - 1, guard));
      }
      // y-faces: slot Rp (this row is the lower cell), Rp-1 (upper cell);
      // columns Cp and Cp+c_x
      const Real b0 = Real(1) - WY(cy + 1, Rp), b1 = WY(cy + 1, Rp - 1);
      Real hy = b0 * sget(1, ip, Rp, Cp, guard) + b1 * sget(1, ip, Rp - 1, Cp, guard);
      if constexpr (cx != 0) {
        const Real m0 = Real(1) - WX(cx < 0 ? 3 : 4, Cp), m1 = WX(cx < 0 ? 3 : 4, Cp + cx);
        hy = m0 * hy + m1 * (b0 * sget(1, ip, Rp, Cp + cx, guard) + b1 * sget(1, ip, Rp - 1, Cp + cx, guard));
      }
      // z-face p+1/2: columns Cp, Cp+c_x; rows Rp, Rp+c_y
      Real hz = sget(2, ip, Rp, Cp, guard);
      Real hz2 = Real(0);
      if constexpr (cy != 0) hz2 = sget(2, ip, Rp + cy, Cp, guard);
      if constexpr (cx != 0) {
        const Real m0 = Real(1) - WX(cx < 0 ? 3 : 4, Cp), m1 = WX(cx < 0 ? 3 : 4, Cp + cx);
        hz = m0 * hz + m1 * sget(2, ip, Rp, Cp + cx, guard);
        if constexpr (cy != 0) hz2 = m0 * hz2 + m1 * sget(2, ip, Rp + cy, Cp + cx, guard);
      }
      if constexpr (cy != 0) {
        const Real m0 = Real(1) - WY(cy < 0 ? 3 : 4, Rp), m1 = WY(cy < 0 ? 3 : 4, Rp + cy);
        hz = m0 * hz + m1 * hz2;
      }
      hz_out = hz;
      return hx + hy;
    }
  };

  // wall ghost folded into this cell: position and weight (at most one wall axis)
  int gR = -1, gC = -1;
  Real gq = Real(0);
  if (!P.periodic[1]) {
    if (j0 == 0 && R <= 2) {
      gR = 0, gC = Cc, gq = R == 1 ? (Real)P.qlo[1] : Real(1) - (Real)P.qlo[1];
    } else if (j0 + kGY == ny && R >= kGY - 1) {
      gR = kGY + 1, gC = Cc, gq = R == kGY ? (Real)P.qhi[1] : Real(1) - (Real)P.qhi[1];
    }
  } else if (!P.periodic[0]) {
    if (i0 == 0 && Cc <= 2) {
      gR = R, gC = 0, gq = Cc == 1 ? (Real)P.qlo[0] : Real(1) - (Real)P.qlo[0];
    } else if (i0 + kGX == nx && Cc >= kGX - 1) {
      gR = R, gC = kGX + 1, gq = Cc == kGX ? (Real)P.qhi[0] : Real(1) - (Real)P.qhi[0];
    }
  }
  const bool ghost = gR >= 0;

  Real accP[Q], accC[Q];
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    accP[ip] = Real(0);
    accC[ip] = Real(0);
  });
  const StateBuf<Store> bin(fin, Q * (unsigned)(P.N * (long long)sizeof(Store)));
  const unsigned Nb = (unsigned)(P.N * (long long)sizeof(Store));
  const RawBuf lamb(lam, (unsigned)(Q * P.N * sizeof(Lam)));
  const unsigned lam_pop = (unsigned)(P.N * sizeof(Lam));
  // with the state stored in the arithmetic type the face pass has stored
  // rho and j of every cell behind the records: 4 values read per cell
  // instead of 19 of f
  constexpr bool kMoments = std::is_same<Store, Real>::value;
  const RawBuf mb(faces + 3LL * kCompact * P.eN + mom * 4LL * P.N, (unsigned)(4 * P.N * sizeof(Real)));
  const unsigned mom_pop = (unsigned)(P.N * sizeof(Real));
  // FLOW: rho and j_a of f(n+1) at the cell, and s0 = ubar dV / V
  static_assert(!FLOW || kMoments, "the controller term needs the moments of f(n+1)");
  const RawBuf mbn(faces + 3LL * kCompact * P.eN + (1 - mom) * 4LL * P.N, (unsigned)(4 * P.N * sizeof(Real)));
  const int fa = P.flow_axis;
  Real s0 = Real(0), rn = Real(1), jn = Real(0);
  if constexpr (FLOW) s0 = (Real)(ubar[0] * P.wd[0][i0 + c + 1] * P.wd[1][j0 + r + 1] * P.wd[2][1] * inv_vol);

  // inputs of the centre stage of one cell: lam', rho and j, the cell force,
  // the gradient so far
  Real lp[Q], rho = Real(1), j[3] = {0, 0, 0};
  double fc[3] = {0, 0, 0}, g0[3] = {0, 0, 0};
  auto load_centre = [&](unsigned cell) {
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      lp[ip] = (Real)lamb.ld<Lam>(cell * (unsigned)sizeof(Lam), ip * lam_pop);
    });
    if constexpr (kMoments) {
      rho = mb.ld<Real>(cell * (unsigned)sizeof(Real), 0);
      j[0] = mb.ld<Real>(cell * (unsigned)sizeof(Real), mom_pop);
      j[1] = mb.ld<Real>(cell * (unsigned)sizeof(Real), 2 * mom_pop);
      j[2] = mb.ld<Real>(cell * (unsigned)sizeof(Real), 3 * mom_pop);
      if constexpr (FLOW) {
        rn = mbn.ld<Real>(cell * (unsigned)sizeof(Real), 0);
        jn = mbn.ld<Real>(cell * (unsigned)sizeof(Real), (unsigned)(1 + fa) * mom_pop);
      }
    } else {
      Real ff[Q];
      static_for<Q>([&](auto I) {
        constexpr int ip = decltype(I)::value;
        ff[ip] = ld_state<Real, Store>(bin, cell * (unsigned)sizeof(Store), ip * Nb, P, ip);
      });
      cell_moments<Real>(ff, rho, j);
    }
    if (P.fcell)
      for (int a = 0; a < 3; a++) fc[a] = P.fcell[a * P.N + (long long)cell];
    if (grad)
      for (int a = 0; a < 3; a++) g0[a] = grad[a * P.N + (long long)cell];
  };
  // With FP32 lam and moments the registers allow these loads to be issued
  // before the plane's face records, so the staging and the gather hide them.
  constexpr bool kEarly = kMoments && std::is_same<Lam, float>::value;

  for (int p = z0 - 1; p <= zend; p++) {
    if constexpr (kEarly) {
      if (p - 1 >= z0) load_centre((unsigned)(p - 1) * plane + colcell);
    }
    stage_load(p);
    __syncthreads();  // the previous plane's gather is done with S
    stage_store();
    __syncthreads();
    // transposed reconstruction, merged with the shift of the cotangent
    // window after the plane p-1 is complete
    Real accN[Q];
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      constexpr int cz = CZ[ip];
      Real hz;
      Real hzr = inplane(I, R, Cc, false, hz);
      if (ghost) {
        Real gz;
        const Real g = inplane(I, gR, gC, true, gz);
        hzr += gq * g;
        hz += gq * gz;
      }
      Real nxt = Real(0);
      if constexpr (cz == 0) {
        accC[ip] += hzr;
      } else if constexpr (cz > 0) {
        accC[ip] += (Real(1) - mz7) * hzr;
        accP[ip] += mz7 * hzr;
      } else {
        accC[ip] += (Real(1) - mz6) * hzr;
        nxt = mz6 * hzr;
      }
      const Real wn = ip == 0 ? Real(0.5) : wRz[cz + 1];
      accC[ip] += (Real(1) - wn) * hz;
      accN[ip] = nxt + wn * hz;
    });
    // centre-stage adjoint of plane p-1 (as k_adj_cells)
    if (p - 1 >= z0) {
      const unsigned cell = (unsigned)(p - 1) * plane + colcell;
      const long long cl = (long long)cell;
      if constexpr (!kEarly) load_centre(cell);
      if constexpr (FLOW) {
        const Real inv = Real(1) / rn, sc = s0 * inv, t0 = -sc * (jn * inv);
        const Real ex = fa == 0 ? Real(1) : Real(0), ey = fa == 1 ? Real(1) : Real(0),
                   ez = fa == 2 ? Real(1) : Real(0);
        static_for<Q>([&](auto I) {
          constexpr int ip = decltype(I)::value;
          lp[ip] += fma_rn(Real(CX[ip]) * ex + Real(CY[ip]) * ey + Real(CZ[ip]) * ez, sc, t0);
        });
      }
      static_for<Q>([&](auto I) {
        constexpr int ip = decltype(I)::value;
        accP[ip] += Real(4.0 / 3.0) * lp[ip];
      });
      Real F[3], u[3], ut[3];
      const Real invr = Real(1) / rho;
      for (int a = 0; a < 3; a++) {
        F[a] = Famb[a];
        if (P.fcell) F[a] += (Real)fc[a];
        ut[a] = j[a] * invr;
        u[a] = ut[a] + hdt * F[a];
      }
      Real g_rho, g_u[3], g_F[3];
      feq_adjoint<Real>(accP, bw2, bw0, rho, u, F, g_rho, g_u, g_F, AllPops{});
      const Real Acoef = g_rho - (g_u[0] * ut[0] + g_u[1] * ut[1] + g_u[2] * ut[2]) * invr;
      static_for<Q>([&](auto I) {
        constexpr int ip = decltype(I)::value;
        const Real v = -lp[ip] * Real(1.0 / 3.0) + bw1 * accP[ip] + Acoef +
                       (g_u[0] * Real(CX[ip]) + g_u[1] * Real(CY[ip]) + g_u[2] * Real(CZ[ip])) * invr;
        lam_prev[ip * P.N + cl] = (Lam)v;
      });
      if (grad) {
        for (int a = 0; a < 3; a++) grad[a * P.N + cl] = g0[a] + (double)(g_F[a] + hdt * g_u[a]);
      }
    }
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      accP[ip] = accC[ip];
      accC[ip] = accN[ip];
    });
  }
}

// The z-marching gather covers FP32 and FP64 arithmetic on 32 x 8 column
// tiles with z periodic and uniform, at most one wall axis (x or y), and face
// arrays and lam' that 32-bit buffer offsets reach.
bool gather_z_supported(const Ctx& c) {
  const Setup& s = c.s;
  if (c.kernel == 0) return false;
  if (s.n[0] % kGX != 0 || s.n[1] % kGY != 0 || !s.periodic[2] || !s.uniform[2]) return false;
  if (!s.periodic[0] && !s.periodic[1]) return false;
  if (!s.periodic[0] && s.n[0] < 3) return false;
  if (!s.periodic[1] && s.n[1] < 3) return false;
  const double rb = c.precision == 0 ? 8.0 : 4.0;
  if ((double)Q * c.kp.eN * rb >= 2147483648.0 || (double)Q * c.kp.N * 8.0 >= 2147483648.0) return false;
  return true;
}

int gather_z_chunk(const KParams& P) {
  const long long cols = (long long)(P.n[0] / kGX) * (P.n[1] / kGY);
  int tz = 64;
  while (tz > 8 && cols * ((P.n[2] + tz - 1) / tz) < 4 * 256) tz /= 2;
  return tz;
}

// One adjoint step.  Lam = double: every configuration; Lam = float (FP32
// arithmetic with the z-marching gather only): lam' and lam in FP32.
template <class Real, class Store, class Lam>
hipError_t run_adjoint(Ctx& cx, const void* f, const Lam* lam_next, Lam* lam_prev, double* grad, int wall,
                       double* fpart, long long fpart_len, long long* nparts, int mom, const double* ubar) {
  const KParams& P = cx.kp;
  Real* faces = static_cast<Real*>(cx.adj_faces);
  constexpr bool kLam32 = std::is_same<Lam, float>::value;
  bool handled = false;
  // With the z-marching gather the fused adjoint-face pass writes the
  // compact face records, which the gather expands
  const int format = adjoint_record_format(cx, kLam32);
  const bool compact = format != kFacesFull;
  if (kLam32 && !compact) return hipErrorNotSupported;
  // Each format leaves some slots unwritten (ghost rows, missing faces) that
  // must read as 0: clear the arrays when the format changes.
  const int fmt = compact ? 1 : 0;
  if (cx.adj_faces_format != fmt) {
    hipError_t e0 = hipMemsetAsync(cx.adj_faces, 0, (size_t)P.eN * 3 * Q * sizeof(Real), cx.stream);
    if (e0 != hipSuccess) return e0;
    cx.adj_faces_format = fmt;
  }
  // the controller term rides on lam' where the face pass hands rho and j to
  // the gather (storage in the arithmetic type) and lam' is the sweep's:
  // FP32 with FP32 arithmetic, FP64 with FP64
  constexpr bool kFuse = std::is_same<Store, Real>::value && (kLam32 || std::is_same<Real, double>::value);
  if (ubar && !kFuse) return hipErrorNotSupported;
  hipError_t e = launch_adjoint_faces_fused(cx, f, lam_next, faces, format, &handled, fpart, fpart_len, nparts, mom,
                                            ubar);
  if (e != hipSuccess) return e;
  // the force cotangent of the face collisions comes from the fused pass only
  if (fpart && !handled) return hipErrorNotSupported;
  {
    if (compact) {
      const int TZ = gather_z_chunk(P);
      dim3 grid(P.n[0] / kGX, P.n[1] / kGY, (P.n[2] + TZ - 1) / TZ);
      const double inv_vol = 1.0 / (cx.s.width[0] * cx.s.width[1] * cx.s.width[2]);
      if constexpr (kFuse) {
        if (ubar) {
          hipLaunchKernelGGL((k_adj_cells_z<Real, Store, Lam, true>), grid, dim3(kGT), 0, cx.stream, P,
                             static_cast<const Store*>(f), lam_next, static_cast<const Real*>(cx.adj_faces),
                             lam_prev, grad, TZ, mom, ubar, inv_vol);
          return hipGetLastError();
        }
      }
      hipLaunchKernelGGL((k_adj_cells_z<Real, Store, Lam>), grid, dim3(kGT), 0, cx.stream, P,
                         static_cast<const Store*>(f), lam_next, static_cast<const Real*>(cx.adj_faces), lam_prev,
                         grad, TZ, mom, nullptr, inv_vol);
      return hipGetLastError();
    }
  }
  if constexpr (!kLam32) {
    if (!handled) {
      if (!cx.fbx) {
        e = cx.alloc(&cx.fbx, (size_t)P.eN * Q * sizeof(Real));
        if (e != hipSuccess) return e;
      }
      e = launch_fbar_ghosts(cx, f, cx.fbx);
      if (e != hipSuccess) return e;
      const unsigned nbe = (unsigned)((P.eN + 255) / 256);
      const Real* fbx = static_cast<const Real*>(cx.fbx);
      hipLaunchKernelGGL((k_adj_faces<Real, 0>), dim3(nbe), dim3(256), 0, cx.stream, P, fbx, lam_next, faces);
      hipLaunchKernelGGL((k_adj_faces<Real, 1>), dim3(nbe), dim3(256), 0, cx.stream, P, fbx, lam_next, faces);
      hipLaunchKernelGGL((k_adj_faces<Real, 2>), dim3(nbe), dim3(256), 0, cx.stream, P, fbx, lam_next, faces);
    }
    for (int b = 0; b < 3; b++) {
      if (!P.periodic[b]) continue;
      const long long cnt = (long long)P.e[(b + 1) % 3] * P.e[(b + 2) % 3] * 3 * Q;
      hipLaunchKernelGGL((k_face_ghosts<Real>), dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, cx.stream, P,
                         faces, b);
    }
    const unsigned nbc = (unsigned)((P.n[0] + kCX - 1) / kCX * ((P.n[1] + kCY - 1) / kCY) *
                                    ((P.n[2] + kCZ - 1) / kCZ));
    hipLaunchKernelGGL((k_adj_cells<Real, Store>), dim3(nbc), dim3(256), 0, cx.stream, P,
                       static_cast<const Store*>(f), lam_next, static_cast<const Real*>(cx.adj_faces), lam_prev, grad,
                       wall);
    return hipGetLastError();
  }
  return hipErrorNotSupported;
}

}  // namespace

namespace {

template <class Lam>
hipError_t adjoint_step(Ctx& c, const void* f, const Lam* lam_next, Lam* lam_prev, double* grad_force,
                        double* fpart, long long fpart_len, long long* nparts, int mom, const double* ubar) {
  const Setup& s = c.s;
  if (s.recon != 0) return hipErrorNotSupported;
  int walls = 0, wall = -1;
  for (int a = 0; a < 3; a++)
    if (!s.periodic[a]) {
      walls++;
      wall = a;
    }
  if (walls > 1) return hipErrorNotSupported;
  // FP16 non-equilibrium storage has no adjoint
  if (c.precision == 3) return hipErrorNotSupported;
  if constexpr (!std::is_same<Lam, double>::value) {
    if (c.precision == 0) return hipErrorNotSupported;
  }
  const size_t rb = c.precision == 0 ? 8 : 4;
  if (!c.adj_faces) {
    const size_t bytes = (size_t)c.kp.eN * 3 * Q * rb;
    hipError_t e = c.alloc(&c.adj_faces, bytes);
    if (e != hipSuccess) return e;
    c.adj_faces_format = -1;  // cleared before first use (run_adjoint)
  }
  switch (c.precision) {
    case 0:
      if constexpr (std::is_same<Lam, double>::value)
        return run_adjoint<double, double, Lam>(c, f, lam_next, lam_prev, grad_force, wall, fpart, fpart_len,
                                                nparts, mom, ubar);
      else
        return hipErrorNotSupported;
    case 1:
      return run_adjoint<float, float, Lam>(c, f, lam_next, lam_prev, grad_force, wall, fpart, fpart_len, nparts,
                                            mom, ubar);
    default:
      return run_adjoint<float, __half, Lam>(c, f, lam_next, lam_prev, grad_force, wall, fpart, fpart_len, nparts,
                                             mom, ubar);
  }
}

}  // namespace

hipError_t launch_adjoint(Ctx& c, const void* f, const double* lam_next, double* lam_prev, double* grad_force,
                          double* fpart, long long fpart_len, long long* nparts, int mom, const double* ubar) {
  return adjoint_step<double>(c, f, lam_next, lam_prev, grad_force, fpart, fpart_len, nparts, mom, ubar);
}

hipError_t launch_adjoint_lam32(Ctx& c, const void* f, const float* lam_next, float* lam_prev, double* grad_force,
                                double* fpart, long long fpart_len, long long* nparts, int mom,
                                const double* ubar) {
  return adjoint_step<float>(c, f, lam_next, lam_prev, grad_force, fpart, fpart_len, nparts, mom, ubar);
}

// The face-record format of an adjoint step: compact records for the
// z-marching gather wherever it and the fused face pass run (lam' in FP32
// with lam32), else the 19 cotangents of every face.
int adjoint_record_format(const Ctx& c, bool lam32) {
  if (!gather_z_supported(c) || !adjoint_faces_fused_covers(c)) return kFacesFull;
  return lam32 ? kFacesCompactLam32 : kFacesCompact;
}

bool adjoint_flow_fused(const Ctx& c) {
  if (c.precision == 0) return adjoint_record_format(c, false) == kFacesCompact;
  return c.precision == 1 && adjoint_lam32_supported(c);
}

bool adjoint_lam32_supported(const Ctx& c) {
  if (c.precision == 0 || c.precision == 3 || c.s.recon != 0) return false;
  return gather_z_supported(c) && adjoint_faces_fused_covers(c);
}

// The sweep's adjoint step: FP64 on the fused face pass (full records), or
// FP32 arithmetic with lam in FP32 (compact records); the forms
// launch_adjoint_faces_fused gives the force cotangent for.
bool adjoint_flow_supported(const Ctx& c) {
  if (c.s.recon != 0) return false;
  int walls = 0;
  for (int a = 0; a < 3; a++) walls += !c.s.periodic[a];
  if (walls > 1) return false;
  if (c.precision == 0) return adjoint_faces_fused_covers(c);
  return adjoint_lam32_supported(c);
}

}  // namespace dugks
