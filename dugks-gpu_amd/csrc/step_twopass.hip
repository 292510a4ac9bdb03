// Two-pass DUGKS step: the reference's stage structure (src/dugks.cu:322-391)
// on structure-of-arrays storage.
//   k_fbar    f-bar+ of every interior cell into a ghost-inclusive array
//             (calc_fb1, src/dugks.cu:1131-1196)
//   k_ghost   ghost layers axis by axis, x -> y -> z: periodic copy or wall
//             extrapolation (extend_fb_x/_x_wall/_y, src/dugks.cu:1212-1273,
//             and the z self-exchange of src/parallel.cu:185-234)
//   k_flux    per cell: six faces (trilinear or gradient reconstruction),
//             wall bounce-back, face collision, update into the other buffer
//             (calc_fx_tri / calc_fx, src/dugks.cu:421-959)
// This path handles every configuration (any wall axes, stretched axes, both
// reconstructions) and is the cross-check for the fused kernel.
#include "device_common.h"

namespace dugks {

namespace {

template <class Real>
__device__ __forceinline__ void ambient(const KParams& P, Real F[3]) {
  if (P.famb_dev) {
    F[0] = (Real)P.famb_dev[0];
    F[1] = (Real)P.famb_dev[1];
    F[2] = (Real)P.famb_dev[2];
  } else {
    F[0] = (Real)P.famb[0];
    F[1] = (Real)P.famb[1];
    F[2] = (Real)P.famb[2];
  }
}

template <class Real, class Store>
__global__ void __launch_bounds__(256) k_fbar(KParams P, const Store* __restrict__ f,
                                              Real* __restrict__ fbx) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P.N) return;
  const int i = (int)(c % P.n[0]);
  const long long t = c / P.n[0];
  const int j = (int)(t % P.n[1]);
  const int k = (int)(t / P.n[1]);
  const long long e = ((long long)(k + 1) * P.e[1] + (j + 1)) * P.e[0] + (i + 1);
  Real ff[Q], fb[Q];
  load_cell<Real, Store>(f, c, P, ff);
  Real F[3];
  ambient(P, F);
  if (P.fcell) {
    F[0] += (Real)P.fcell[c];
    F[1] += (Real)P.fcell[P.N + c];
    F[2] += (Real)P.fcell[2 * P.N + c];
  }
  centre_fbar<Real>(ff, F, (Real)P.hdt, (Real)P.bw0, (Real)P.bw2, fb);
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    fbx[ip * P.eN + e] = fb[ip];
  });
}

// Ghost layer of axis a over the whole ghost-inclusive extent of the other
// two axes.
template <class Real>
__global__ void __launch_bounds__(256) k_ghost(KParams P, Real* __restrict__ fbx, int a) {
  const int b = (a + 1) % 3, cc = (a + 2) % 3;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nr = P.e[b];
  if (t >= nr * P.e[cc]) return;
  const int r = (int)(t % nr), s = (int)(t / nr);
  long long stride[3] = {1, P.e[0], (long long)P.e[0] * P.e[1]};
  const long long base = r * stride[b] + s * stride[cc];
  const int n = P.n[a];
  const long long sa = stride[a];
  const long long g0 = base, g1 = base + (n + 1) * sa, i1 = base + sa, i2 = base + 2 * sa,
                  in = base + n * sa, in1 = base + (n - 1) * sa;
  const Real qlo = (Real)P.qlo[a], qhi = (Real)P.qhi[a];
  for (int ip = 0; ip < Q; ip++) {
    Real* p = fbx + ip * P.eN;
    if (P.periodic[a]) {
      p[g0] = p[in];
      p[g1] = p[i1];
    } else {
      // q f(1) + (1 - q) f(2) as f(2) + q (f(1) - f(2)) (see lerp_)
      p[g0] = fma_rn(qlo, p[i1] - p[i2], p[i2]);
      p[g1] = fma_rn(qhi, p[in] - p[in1], p[in1]);
    }
  }
}

// (1 - m) a + m b as a + m (b - a): constants stay exactly constant whatever
// the rounding of m (see the fused kernel's lerp_)
template <class Real>
__device__ __forceinline__ Real lerp_(Real a, Real b, Real m) {
  return fma_rn(m, b - a, a);
}

// Face collision of the reference (f2ruvwh + feq_s, src/dugks.cu:675-709,
// 1201-1207): rho = sum fbar, u = j/rho + (h/2) F_amb,
// ft_i = fw1 fbar_i + (fw2 + fw0 (c_i - u).F/RT) feq_i.
template <class Real>
__device__ __forceinline__ void face_collide(Real (&ft)[Q], const Real F[3], Real hdt, Real fw0,
                                             Real fw1, Real fw2) {
  Real rho, j[3];
  cell_moments<Real>(ft, rho, j);
  const Real inv = rcp_real(rho);
  const Real hh = Real(0.5) * hdt;
  const EqTerm<Real> eq(rho, j[0] * inv + hh * F[0], j[1] * inv + hh * F[1], j[2] * inv + hh * F[2], F, fw2, fw0);
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    ft[ip] = eq.template relax<ip>(fw1, ft[ip], cdot<ip>(F[0], F[1], F[2]));
  });
}

// Reconstruction of the face of cell (m0,m1,m2) (ghost-inclusive indices) on
// axis A, side S (+1 / -1), into ft[0..18].
template <class Real, int A, int S>
__device__ __forceinline__ void reconstruct_tri(const KParams& P, const Real* __restrict__ fbx,
                                                long long e, const int m[3], Real (&ft)[Q]) {
  const long long st[3] = {1, P.e[0], (long long)P.e[0] * P.e[1]};
  const double* tabA = P.tab[A] + m[A] * 9;
  // rest population: average of the two cells (src/dugks.cu:756-763)
  ft[0] = Real(0.5) * (fbx[e] + fbx[e + S * st[A]]);
  static_for<Q - 1>([&](auto I) {
    constexpr int ip = decltype(I)::value + 1;
    constexpr int c[3] = {CX[ip], CY[ip], CZ[ip]};
    const Real* p = fbx + ip * P.eN + e;
    long long off[3];
    Real w[3];
#pragma unroll
    for (int b = 0; b < 3; b++) {
      if (b == A) {
        off[b] = S * st[b];
        w[b] = (Real)tabA[(S > 0 ? 0 : 3) + c[b] + 1];
      } else {
        off[b] = -c[b] * st[b];
        w[b] = c[b] == 0 ? Real(0) : (Real)P.tab[b][m[b] * 9 + (c[b] < 0 ? 6 : 7)];
      }
    }
    // trilinear of the reference (src/dugks.cu:964-978): x, then y, then z
    Real x00 = lerp_(p[0], p[off[0]], w[0]);
    Real y00 = x00;
    if (c[1] != 0 || A == 1) {
      Real x01 = lerp_(p[off[1]], p[off[0] + off[1]], w[0]);
      y00 = lerp_(x00, x01, w[1]);
    }
    Real r = y00;
    if (c[2] != 0 || A == 2) {
      Real x10 = lerp_(p[off[2]], p[off[0] + off[2]], w[0]);
      Real y11 = x10;
      if (c[1] != 0 || A == 1) {
        Real x11 = lerp_(p[off[1] + off[2]], p[off[0] + off[1] + off[2]], w[0]);
        y11 = lerp_(x10, x11, w[1]);
      }
      r = lerp_(y00, y11, w[2]);
    }
    ft[ip] = r;
  });
}

// Midpoint + central-gradient reconstruction (calc_fx, src/dugks.cu:459-583).
template <class Real, int A, int S>
__device__ __forceinline__ void reconstruct_grad(const KParams& P, const Real* __restrict__ fbx,
                                                 long long e, Real (&ft)[Q]) {
  const long long st[3] = {1, P.e[0], (long long)P.e[0] * P.e[1]};
  const Real h = (Real)P.hdt;
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    constexpr int c[3] = {CX[ip], CY[ip], CZ[ip]};
    const Real* p = fbx + ip * P.eN + e;
    const long long o = S * st[A];
    const Real temp = Real(0.5) * (p[o] + p[0]);
    Real g[3];
    g[A] = (S > 0 ? (p[o] - p[0]) : (p[0] - p[o])) / (Real)P.du[A];
#pragma unroll
    for (int b = 0; b < 3; b++) {
      if (b == A) continue;
      const Real aa = Real(0.5) * (p[st[b]] + p[o + st[b]]);
      const Real bb = Real(0.5) * (p[-st[b]] + p[o - st[b]]);
      g[b] = (aa - bb) / (Real(2) * (Real)P.du[b]);
    }
    ft[ip] = temp - h * (Real(c[0]) * g[0] + Real(c[1]) * g[1] + Real(c[2]) * g[2]);
  });
}

template <class Real, class Store>
__global__ void __launch_bounds__(256) k_flux(KParams P, const Store* __restrict__ fin,
                                              const Real* __restrict__ fbx, Store* __restrict__ fout) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P.N) return;
  const int i = (int)(c % P.n[0]);
  const long long t = c / P.n[0];
  const int j = (int)(t % P.n[1]);
  const int k = (int)(t / P.n[1]);
  const int m[3] = {i + 1, j + 1, k + 1};
  const long long e = ((long long)(k + 1) * P.e[1] + (j + 1)) * P.e[0] + (i + 1);
  Real F[3];
  ambient(P, F);
  const Real hdt = (Real)P.hdt, fw0 = (Real)P.fw0, fw1 = (Real)P.fw1, fw2 = (Real)P.fw2;
  Real acc[Q];
  // f~+ = 4/3 f-bar+ - 1/3 f~ (src/dugks.cu:954), as f-bar+ + (f-bar+ - f~)/3
  // (no mass bias from the rounding of 1/3)
  load_cell<Real, Store>(fin, c, P, acc);
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    const Real fb = fbx[ip * P.eN + e];
    acc[ip] = fb + (fb - acc[ip]) * Real(1.0 / 3.0);
  });
  auto face = [&](auto Ac, auto Sc) {
    constexpr int A = decltype(Ac)::value;
    constexpr int S = decltype(Sc)::value;
    Real ft[Q];
    if (P.recon == 0)
      reconstruct_tri<Real, A, S>(P, fbx, e, m, ft);
    else
      reconstruct_grad<Real, A, S>(P, fbx, e, ft);
    // half-way bounce-back at a wall face: incoming population takes the
    // value of its opposite (src/dugks.cu:807-823, generalised to any axis)
    if (!P.periodic[A] && ((S < 0 && m[A] == 1) || (S > 0 && m[A] == P.n[A]))) {
      static_for<Q>([&](auto I) {
        constexpr int ip = decltype(I)::value;
        if constexpr (cv(A, ip) * S < 0) ft[ip] = ft[OPP[ip]];
      });
    }
    face_collide<Real>(ft, F, hdt, fw0, fw1, fw2);
    const Real cf = (Real)P.tab[A][m[A] * 9 + 8];  // dt / width
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      if constexpr (cv(A, ip) != 0) {
        // update: - (dt/dx) c_a (F_+ - F_-)
        acc[ip] -= Real(S * cv(A, ip)) * cf * ft[ip];
      }
    });
  };
  face(std::integral_constant<int, 0>{}, std::integral_constant<int, -1>{});
  face(std::integral_constant<int, 0>{}, std::integral_constant<int, 1>{});
  face(std::integral_constant<int, 1>{}, std::integral_constant<int, -1>{});
  face(std::integral_constant<int, 1>{}, std::integral_constant<int, 1>{});
  face(std::integral_constant<int, 2>{}, std::integral_constant<int, -1>{});
  face(std::integral_constant<int, 2>{}, std::integral_constant<int, 1>{});
  store_cell<Real, Store>(fout, c, P, acc);
}

template <class Real, class Store>
hipError_t run_twopass(Ctx& cx, const void* fin, void* fout) {
  const KParams& P = cx.kp;
  const int bs = 256;
  const unsigned nb = (unsigned)((P.N + bs - 1) / bs);
  Real* fbx = static_cast<Real*>(cx.fbx);
  hipLaunchKernelGGL((k_fbar<Real, Store>), dim3(nb), dim3(bs), 0, cx.stream, P,
                     static_cast<const Store*>(fin), fbx);
  for (int a = 0; a < 3; a++) {
    const long long cnt = (long long)P.e[(a + 1) % 3] * P.e[(a + 2) % 3];
    hipLaunchKernelGGL((k_ghost<Real>), dim3((unsigned)((cnt + bs - 1) / bs)), dim3(bs), 0, cx.stream,
                       P, fbx, a);
  }
  if (cx.timing) hipEventRecord(cx.tev[0], cx.stream);
  hipLaunchKernelGGL((k_flux<Real, Store>), dim3(nb), dim3(bs), 0, cx.stream, P,
                     static_cast<const Store*>(fin), static_cast<const Real*>(fbx),
                     static_cast<Store*>(fout));
  if (cx.timing) hipEventRecord(cx.tev[1], cx.stream);
  cx.launches_per_step = 5;
  return hipGetLastError();
}

template <class Real, class Store>
hipError_t run_fbar_ghosts(Ctx& cx, const void* fin, void* fbx_) {
  const KParams& P = cx.kp;
  const int bs = 256;
  const unsigned nb = (unsigned)((P.N + bs - 1) / bs);
  Real* fbx = static_cast<Real*>(fbx_);
  hipLaunchKernelGGL((k_fbar<Real, Store>), dim3(nb), dim3(bs), 0, cx.stream, P,
                     static_cast<const Store*>(fin), fbx);
  for (int a = 0; a < 3; a++) {
    const long long cnt = (long long)P.e[(a + 1) % 3] * P.e[(a + 2) % 3];
    hipLaunchKernelGGL((k_ghost<Real>), dim3((unsigned)((cnt + bs - 1) / bs)), dim3(bs), 0, cx.stream,
                       P, fbx, a);
  }
  return hipGetLastError();
}

}  // namespace

// f-bar+ of every cell with the ghost layers filled (the first two stages of
// the two-pass step), into a ghost-inclusive SoA array of the arithmetic type.
hipError_t launch_fbar_ghosts(Ctx& c, const void* fin, void* fbx) {
  switch (c.precision) {
    case 0:
      return run_fbar_ghosts<double, double>(c, fin, fbx);
    case 1:
      return run_fbar_ghosts<float, float>(c, fin, fbx);
    case 2:
      return run_fbar_ghosts<float, __half>(c, fin, fbx);
    default:
      return run_fbar_ghosts<float, NeqHalf>(c, fin, fbx);
  }
}

hipError_t launch_step_twopass(Ctx& c, const void* fin, void* fout) {
  switch (c.precision) {
    case 0:
      return run_twopass<double, double>(c, fin, fout);
    case 1:
      return run_twopass<float, float>(c, fin, fout);
    case 2:
      return run_twopass<float, __half>(c, fin, fout);
    default:
      return run_twopass<float, NeqHalf>(c, fin, fout);
  }
}

}  // namespace dugks
