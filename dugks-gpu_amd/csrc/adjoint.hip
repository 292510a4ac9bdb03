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
//      is lam_fbar_q = fw1 lam_ft_q + A + B.c_q, so a face stores only the
//      four scalars (A, B);
//   3. k_adj_cells: per cell, S3^T as a gather over the faces whose stencil
//      contains the cell (bounce-back and the wall ghost extrapolation folded
//      in), then S5^T and S1^T.
// Deterministic (no atomics).  At most one wall axis.
#include "device_common.h"

namespace dugks {

namespace {

template <class Real>
__device__ __forceinline__ Real lerp_(Real a, Real b, Real m) {
  return (Real(1) - m) * a + m * b;
}

// d/d(rho, u, F) of sum_i lam_i (B + C T_i) feq_i(rho, u), T_i = (c_i - u).F/RT,
// over the populations in the compile-time set Sel (Sel(ip) true).
template <class Real, class Sel>
__device__ __forceinline__ void feq_adjoint(const Real (&lam)[Q], Real B, Real C, Real rho, const Real u[3],
                                            const Real F[3], Real& g_rho, Real g_u[3], Real g_F[3], Sel sel) {
  const Real uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const Real uF = u[0] * F[0] + u[1] * F[1] + u[2] * F[2];
  g_rho = 0;
  for (int a = 0; a < 3; a++) g_u[a] = g_F[a] = 0;
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    if constexpr (decltype(sel(I))::value) {
      const Real c[3] = {Real(CX[ip]), Real(CY[ip]), Real(CZ[ip])};
      const Real cu = c[0] * u[0] + c[1] * u[1] + c[2] * u[2];
      const Real eu = Real(3) * cu;
      const Real poly = Real(1) + eu + Real(0.5) * (eu * eu - Real(3) * uu);
      const Real w = Real(wgt(ip));
      const Real feq = w * rho * poly;
      const Real cF = c[0] * F[0] + c[1] * F[1] + c[2] * F[2];
      const Real T = Real(3) * (cF - uF);
      const Real coef = B + C * T;
      const Real l = lam[ip];
      g_rho += l * coef * w * poly;
      for (int a = 0; a < 3; a++) {
        const Real dfeq = w * rho * Real(3) * (c[a] + eu * c[a] - u[a]);
        g_u[a] += l * (coef * dfeq - C * Real(3) * F[a] * feq);
        g_F[a] += l * C * Real(3) * (c[a] - u[a]) * feq;
      }
    }
  });
}

struct AllPops {
  template <class I>
  __device__ constexpr std::true_type operator()(I) const { return {}; }
};
template <int A>
struct Crossing {
  template <class I>
  __device__ constexpr std::integral_constant<bool, cv(A, I::value) != 0> operator()(I) const { return {}; }
};

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

// Slot (left-cell index) of the face between m-1.. on axis a whose left cell
// is mL, or -1 if that face does not exist.  Periodic axes store faces
// 1..n (the face between 0 and 1 is the face between n and n+1); wall axes
// store 0..n.
__device__ __forceinline__ int face_slot(const KParams& P, int a, int mL) {
  const int n = P.n[a];
  if (P.periodic[a]) {
    if (mL == 0) return n;
    return (mL >= 1 && mL <= n) ? mL : -1;
  }
  return (mL >= 0 && mL <= n) ? mL : -1;
}

__device__ __forceinline__ int tangential_slot(const KParams& P, int b, int m) {
  m = wrap(P, b, m);
  return (m >= 1 && m <= P.n[b]) ? m : -1;
}

// ---- kernel 2: one thread per face of axis A ----
template <class Real, int A>
__global__ void __launch_bounds__(256) k_adj_faces(KParams P, const Real* __restrict__ fbx,
                                                   const double* __restrict__ lam, double* __restrict__ faces) {
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
  double* out = faces + (long long)A * 4 * P.eN;
  out[e] = (double)Acoef;
  out[P.eN + e] = (double)(g_u[0] * invr);
  out[2 * P.eN + e] = (double)(g_u[1] * invr);
  out[3 * P.eN + e] = (double)(g_u[2] * invr);
}

// Collision cotangent of population ip at the face of axis A with left cell
// slot s (ghost-inclusive coordinates mf), before bounce-back.
template <int A, int ip>
__device__ __forceinline__ double face_base(const KParams& P, const double* lam, const double* faces,
                                            const int mf[3], long long ef) {
  const double* fa = faces + (long long)A * 4 * P.eN;
  double v = fa[ef] + fa[P.eN + ef] * CX[ip] + fa[2 * P.eN + ef] * CY[ip] + fa[3 * P.eN + ef] * CZ[ip];
  if constexpr (cv(A, ip) != 0) {
    int mR[3] = {mf[0], mf[1], mf[2]};
    mR[A] += 1;
    const double updL = P.tab[A][mf[A] * 9 + 8], updR = P.tab[A][(mf[A] + 1) * 9 + 8];
    const double lL = lam_at(P, lam, ip, mf[0], mf[1], mf[2]);
    const double lR = lam_at(P, lam, ip, mR[0], mR[1], mR[2]);
    v += P.fw1 * cv(A, ip) * (updR * lR - updL * lL);
  }
  return v;
}

// ... after bounce-back at wall faces: incoming populations were replaced by
// their opposites, so their cotangent moves to the opposite.
template <int A, int ip>
__device__ __forceinline__ double face_lambda(const KParams& P, const double* lam, const double* faces,
                                              const int mf[3], long long ef) {
  const double base = face_base<A, ip>(P, lam, faces, mf, ef);
  if constexpr (cv(A, ip) == 0) {
    return base;
  } else {
    if (P.periodic[A]) return base;
    const bool low = mf[A] == 0, high = mf[A] == P.n[A];
    if (!low && !high) return base;
    const bool incoming = (low && cv(A, ip) > 0) || (high && cv(A, ip) < 0);
    if (incoming) return 0.0;
    return base + face_base<A, OPP[ip]>(P, lam, faces, mf, ef);
  }
}

// S3^T at ghost-inclusive position t for population ip: sum over the faces
// of axis A whose reconstruction stencil contains t, times the stencil weight.
template <int A, int ip>
__device__ __forceinline__ double gather_axis(const KParams& P, const double* lam, const double* faces, const int t[3]) {
  constexpr int c[3] = {CX[ip], CY[ip], CZ[ip]};
  constexpr int B1 = (A + 1) % 3, B2 = (A + 2) % 3;
  double acc = 0;
  // normal: t is the left cell of face mL = t_A, or the right cell of mL = t_A - 1
#pragma unroll
  for (int side = 0; side < 2; side++) {
    const int mLraw = t[A] - side;
    const int sl = face_slot(P, A, mLraw);
    if (sl < 0) continue;
    double wn;
    if constexpr (ip == 0) {
      wn = 0.5;
    } else {
      const double wR = P.tab[A][sl * 9 + c[A] + 1];  // fraction toward the right cell
      wn = side == 0 ? 1.0 - wR : wR;
    }
#pragma unroll
    for (int s1 = 0; s1 < 2; s1++) {
      if (c[B1] == 0 && s1 == 1) continue;
      if (ip == 0 && s1 == 1) continue;
      const int f1 = tangential_slot(P, B1, t[B1] + s1 * c[B1]);
      if (f1 < 0) continue;
      double w1 = 1.0;
      if (c[B1] != 0 && ip != 0) {
        const double m = P.tab[B1][f1 * 9 + (c[B1] < 0 ? 6 : 7)];
        w1 = s1 ? m : 1.0 - m;
      }
#pragma unroll
      for (int s2 = 0; s2 < 2; s2++) {
        if (c[B2] == 0 && s2 == 1) continue;
        if (ip == 0 && s2 == 1) continue;
        const int f2 = tangential_slot(P, B2, t[B2] + s2 * c[B2]);
        if (f2 < 0) continue;
        double w2 = 1.0;
        if (c[B2] != 0 && ip != 0) {
          const double m = P.tab[B2][f2 * 9 + (c[B2] < 0 ? 6 : 7)];
          w2 = s2 ? m : 1.0 - m;
        }
        int mf[3];
        mf[A] = sl;
        mf[B1] = f1;
        mf[B2] = f2;
        const long long ef = ext_index(P, mf[0], mf[1], mf[2]);
        acc += wn * w1 * w2 * face_lambda<A, ip>(P, lam, faces, mf, ef);
      }
    }
  }
  return acc;
}

template <int ip>
__device__ __forceinline__ double gather(const KParams& P, const double* lam, const double* faces, const int t[3]) {
  return gather_axis<0, ip>(P, lam, faces, t) + gather_axis<1, ip>(P, lam, faces, t) +
         gather_axis<2, ip>(P, lam, faces, t);
}

// ---- kernel 3: one thread per interior cell ----
template <class Real, class Store>
__global__ void __launch_bounds__(256) k_adj_cells(KParams P, const Store* __restrict__ f,
                                                   const double* __restrict__ lam, const double* __restrict__ faces,
                                                   double* __restrict__ lam_prev, double* __restrict__ grad,
                                                   int wall_axis) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P.N) return;
  const int t[3] = {(int)(c % P.n[0]) + 1, (int)((c / P.n[0]) % P.n[1]) + 1, (int)(c / ((long long)P.n[0] * P.n[1])) + 1};
  Real lfb[Q];
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    double v = 4.0 / 3.0 * lam[ip * P.N + c] + gather<ip>(P, lam, faces, t);
    if (wall_axis >= 0) {
      // S2^T: the wall ghost g = q f(1) + (1-q) f(2) (and mirrored at the top)
      const int W = wall_axis, n = P.n[W];
      int g[3] = {t[0], t[1], t[2]};
      if (t[W] == 1 || t[W] == 2) {
        g[W] = 0;
        const double q = P.qlo[W];
        v += (t[W] == 1 ? q : 1.0 - q) * gather<ip>(P, lam, faces, g);
      }
      if (t[W] == n || t[W] == n - 1) {
        g[W] = n + 1;
        const double q = P.qhi[W];
        v += (t[W] == n ? q : 1.0 - q) * gather<ip>(P, lam, faces, g);
      }
    }
    lfb[ip] = (Real)v;
  });
  // S1^T
  Real ff[Q];
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    ff[ip] = Codec<Store>::template load<Real>(f, ip * P.N + c, P, ip);
  });
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

template <class Real, class Store>
hipError_t run_adjoint(Ctx& cx, const void* f, const double* lam_next, double* lam_prev, double* grad, int wall) {
  const KParams& P = cx.kp;
  hipError_t e = launch_fbar_ghosts(cx, f, cx.fbx);
  if (e != hipSuccess) return e;
  const unsigned nbe = (unsigned)((P.eN + 255) / 256);
  const Real* fbx = static_cast<const Real*>(cx.fbx);
  hipLaunchKernelGGL((k_adj_faces<Real, 0>), dim3(nbe), dim3(256), 0, cx.stream, P, fbx, lam_next, cx.adj_faces);
  hipLaunchKernelGGL((k_adj_faces<Real, 1>), dim3(nbe), dim3(256), 0, cx.stream, P, fbx, lam_next, cx.adj_faces);
  hipLaunchKernelGGL((k_adj_faces<Real, 2>), dim3(nbe), dim3(256), 0, cx.stream, P, fbx, lam_next, cx.adj_faces);
  const unsigned nbc = (unsigned)((P.N + 255) / 256);
  hipLaunchKernelGGL((k_adj_cells<Real, Store>), dim3(nbc), dim3(256), 0, cx.stream, P,
                     static_cast<const Store*>(f), lam_next, cx.adj_faces, lam_prev, grad, wall);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_adjoint(Ctx& c, const void* f, const double* lam_next, double* lam_prev, double* grad_force) {
  const Setup& s = c.s;
  if (s.recon != 0) return hipErrorNotSupported;
  int walls = 0, wall = -1;
  for (int a = 0; a < 3; a++)
    if (!s.periodic[a]) {
      walls++;
      wall = a;
    }
  if (walls > 1) return hipErrorNotSupported;
  const size_t rb = c.precision == 0 ? 8 : 4;
  if (!c.fbx) {
    hipError_t e = c.alloc(&c.fbx, (size_t)c.kp.eN * Q * rb);
    if (e != hipSuccess) return e;
  }
  if (!c.adj_faces) {
    hipError_t e = c.alloc((void**)&c.adj_faces, (size_t)c.kp.eN * 12 * sizeof(double));
    if (e != hipSuccess) return e;
    e = hipMemsetAsync(c.adj_faces, 0, (size_t)c.kp.eN * 12 * sizeof(double), c.stream);
    if (e != hipSuccess) return e;
  }
  switch (c.precision) {
    case 0:
      return run_adjoint<double, double>(c, f, lam_next, lam_prev, grad_force, wall);
    case 1:
      return run_adjoint<float, float>(c, f, lam_next, lam_prev, grad_force, wall);
    default:
      return run_adjoint<float, __half>(c, f, lam_next, lam_prev, grad_force, wall);
  }
}

}  // namespace dugks
