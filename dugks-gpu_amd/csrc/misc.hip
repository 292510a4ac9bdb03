// Utility kernels: macroscopic output, state export/import, FP64 global
// sums, constant-flow-rate controller, perturbation forcing, loss functionals.
#include <map>
#include <mutex>

#include "device_common.h"

namespace dugks {

namespace {

constexpr int kRedBlocks = 1024;  // fixed grid => fixed FP64 summation order
constexpr int kRedThreads = 256;

__device__ __forceinline__ void cell_coords(const KParams& P, long long c, int& i, int& j, int& k) {
  i = (int)(c % P.n[0]);
  const long long t = c / P.n[0];
  j = (int)(t % P.n[1]);
  k = (int)(t / P.n[1]);
}

__device__ __forceinline__ void ambient_d(const KParams& P, double F[3]) {
  for (int a = 0; a < 3; a++) F[a] = P.famb_dev ? P.famb_dev[a] : P.famb[a];
}

// calc_macro (src/dugks.cu:1021-1064): rho = sum f, u = j/rho + h F_amb.
template <class Store>
__global__ void __launch_bounds__(256) k_macro(KParams P, const Store* __restrict__ f, double* rho,
                                               double* ux, double* uy, double* uz) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P.N) return;
  double x[Q], r, j[3];
  load_cell<double, Store>(f, c, P, x);
  cell_moments<double>(x, r, j);
  double F[3];
  ambient_d(P, F);
  rho[c] = r;
  ux[c] = j[0] / r + F[0] * P.hdt;
  uy[c] = j[1] / r + F[1] * P.hdt;
  uz[c] = j[2] / r + F[2] * P.hdt;
}

// State -> FP64 SoA [19][N], one thread per cell.
template <class Store>
__global__ void __launch_bounds__(256) k_export(KParams P, const Store* __restrict__ f, double* dst) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P.N) return;
  double x[Q];
  load_cell<double, Store>(f, c, P, x);
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    dst[ip * P.N + c] = x[ip];
  });
}

// FP64 SoA [19][N] -> state, one thread per cell (a cell is encoded whole).
template <class Store>
__global__ void __launch_bounds__(256) k_import(KParams P, const double* __restrict__ src, Store* f) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P.N) return;
  double x[Q];
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    x[ip] = src[ip * P.N + c];
  });
  store_cell<double, Store>(f, c, P, x);
}

// Ops of block_reduce: NV sums.
template <int NV>
struct AllSums {
  int op[NV];
  __device__ constexpr AllSums() : op() {}
};

// Fixed-order block reduction of NV doubles per thread into out[blockIdx*NV..].
template <int NV>
__device__ __forceinline__ void block_reduce_store(double (&v)[NV], double* out) {
  constexpr AllSums<NV> sums;
  block_reduce<NV, kRedThreads>(v, sums.op);
  if (threadIdx.x == 0)
    for (int q = 0; q < NV; q++) out[blockIdx.x * NV + q] = v[q];
}

// Sums: mass, momentum (3), volume-weighted macro velocity (3), volume.
template <class Store>
__global__ void __launch_bounds__(kRedThreads) k_sums(KParams P, const Store* __restrict__ f,
                                                      double* partial) {
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  double F[3];
  ambient_d(P, F);
  for (long long c = (long long)blockIdx.x * kRedThreads + threadIdx.x; c < P.N;
       c += (long long)kRedBlocks * kRedThreads) {
    int i, j, k;
    cell_coords(P, c, i, j, k);
    const double dV = P.wd[0][i + 1] * P.wd[1][j + 1] * P.wd[2][k + 1];
    double x[Q], r, jm[3];
    load_cell<double, Store>(f, c, P, x);
    cell_moments<double>(x, r, jm);
    const double u = jm[0], v = jm[1], w = jm[2];
    acc[0] += r;
    acc[1] += u;
    acc[2] += v;
    acc[3] += w;
    acc[4] += (u / r + F[0] * P.hdt) * dV;
    acc[5] += (v / r + F[1] * P.hdt) * dV;
    acc[6] += (w / r + F[2] * P.hdt) * dV;
    acc[7] += dV;
  }
  block_reduce_store<8>(acc, partial);
}

template <int NV>
__global__ void __launch_bounds__(kRedThreads) k_reduce_final(const double* partial, int nparts,
                                                             double* out, int accumulate) {
  double acc[NV];
  for (int q = 0; q < NV; q++) acc[q] = 0;
  for (int b = threadIdx.x; b < nparts; b += kRedThreads)
    for (int q = 0; q < NV; q++) acc[q] += partial[b * NV + q];
  constexpr AllSums<NV> sums;
  block_reduce<NV, kRedThreads>(acc, sums.op);
  if (threadIdx.x == 0)
    for (int q = 0; q < NV; q++) out[q] = (accumulate ? out[q] : 0.0) + acc[q];
}

// Constant-flow-rate controller (single thread).  famb_dev layout:
// [0..2] ambient force, [3] previous bulk velocity, [4] previous-valid flag.
// The controlled quantity is the momentum bulk velocity U = <j/rho> (the
// output velocity minus h F, which would respond to F instantly and make the
// loop overshoot).  Per step U^{n+1} = U^n + dt (F^n - D) with D the wall
// drag; D is estimated from the last step's change and the new force closes
// a fraction `gain` of the gap: F^n = D + gain (U_target - U^n) / dt, so the
// error obeys e^{n+1} = (1 - gain) e^n for a steady drag (stable 0 < gain < 2).
__device__ __forceinline__ void flow_update(double* famb, int axis, double Ub, double target, double gain,
                                            double dt) {
  const double Fprev = famb[axis];
  double drag = Fprev;
  if (famb[4] != 0) drag = Fprev - (Ub - famb[3]) / dt;
  famb[axis] = drag + gain * (target - Ub) / dt;
  famb[3] = Ub;
  famb[4] = 1;
}

__global__ void k_flow_control(double* famb, const double* sums, int axis, double target, double gain,
                               double dt, double hdt) {
  // sums[4 + axis] / sums[7] is the bulk output velocity <j/rho + h F>
  const double Ub = sums[4 + axis] / sums[7] - hdt * famb[axis];
  flow_update(famb, axis, Ub, target, gain, dt);
}

// calc_perturb_force (src/dugks.cu:1086-1126) written as the additive
// centre force: streak forcing minus the ambient force inside the slab
// 3 < i < nx/4 + 2 (1-based), zero elsewhere.
__global__ void __launch_bounds__(256) k_perturb(KParams P, int step, double xwidth, double ywidth,
                                                 double zwidth, double visc, double* fcell) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P.N) return;
  int i0, j0, k0;
  cell_coords(P, c, i0, j0, k0);
  const int i = i0 + 1, j = j0 + 1, k = k0 + 1, nx = P.n[0];
  const double pi2 = 3.141592653589793 * 2;
  const double beta = 3, Tpd = 2000, gamma = 2, Re_tau = 180;
  const double ustar = 2 * Re_tau * visc / xwidth;
  const double FF = 2 * pow(ustar, 2) / xwidth;
  const double A0 = 5 * 40.0 / nx * beta * sin(pi2 * step * P.dt / Tpd);
  double F[3];
  ambient_d(P, F);
  double fx = 0, fy = 0, fz = 0;
  if (3 < i && i < nx / 4 + 2) {
    const double xx = pi2 * (i - 3) * P.du[0] / (xwidth / 4.);
    const double yy = pi2 * (j - 0.5) * P.du[1] / ywidth;
    const double zz = pi2 * ((k - 0.5) * P.du[2]) / zwidth;
    fx = 0.5 * FF * A0 * xwidth / 4 * (1 - cos(xx)) * cos(beta * yy) * cos(gamma * zz) - F[0];
    fy = 0.5 * FF * A0 * ywidth / gamma * sin(xx) * cos(beta * yy) * sin(gamma * zz) - F[1];
    fz = FF * (1 - A0 * zwidth / beta * sin(xx) * sin(beta * yy) * cos(gamma * zz)) - F[2];
  }
  fcell[c] = fx;
  fcell[P.N + c] = fy;
  fcell[2 * P.N + c] = fz;
}

// Loss functionals and their gradient with respect to f (SoA FP64).
//   kind 0: mean wall shear stress of the wall-normal axis aw along the
//           streamwise axis as: tau_w = nu rho u_s / delta at the first cell
//           off each wall, rho u_s = sum_i (c_is + h F_s) f_i (linear in f);
//           J = sum over both wall planes / (2 * cells per plane).
//   kind 1: kinetic energy J = sum 0.5 rho |u|^2 dV / V, u = j/rho + h F.
// The wall-shear loss only involves the two wall planes: unless the gradient
// array is being initialised (grad set, gacc == 0), the kernel visits those
// planes only (2 nx nz of the N cells on a y-wall channel).
// NV = 2: the second sum is dJ/dF_fa, the derivative with respect to the
// ambient force component fa (both losses read h F).
template <class Store, class G, int NV = 1>
__global__ void __launch_bounds__(kRedThreads) k_loss(KParams P, const Store* __restrict__ f, int kind,
                                                      int aw, int as, double visc, double vol,
                                                      double nplane, G* grad, double gacc,
                                                      double* partial, int fa = 0) {
  double acc[NV];
  for (int q = 0; q < NV; q++) acc[q] = 0;
  double F[3];
  ambient_d(P, F);
  const bool wall_only = kind == 0 && !(grad && gacc == 0.0);
  // in-plane axes of the wall planes, the faster-varying first, so that
  // consecutive threads read consecutive cells
  const int b1 = aw == 0 ? 1 : 0, b2 = aw == 2 ? 1 : 2;
  const long long npl = (long long)P.n[b1] * P.n[b2];
  const long long count = wall_only ? (P.n[aw] > 1 ? 2 : 1) * npl : P.N;
  for (long long w = (long long)blockIdx.x * kRedThreads + threadIdx.x; w < count;
       w += (long long)kRedBlocks * kRedThreads) {
    int idx[3];
    long long c = w;
    if (wall_only) {
      // w -> (side, in-plane position) -> cell of the low or high wall plane
      const long long q = w % npl;
      idx[aw] = w < npl ? 0 : P.n[aw] - 1;
      idx[b1] = (int)(q % P.n[b1]);
      idx[b2] = (int)(q / P.n[b1]);
      c = ((long long)idx[2] * P.n[1] + idx[1]) * P.n[0] + idx[0];
    } else {
      cell_coords(P, c, idx[0], idx[1], idx[2]);
    }
    double x[Q];
    load_cell<double, Store>(f, c, P, x);
    if (kind == 0) {
      const int m = idx[aw];
      const bool wall = m == 0 || m == P.n[aw] - 1;
      double coef = 0;
      if (wall) {
        const double delta = 0.5 * P.wd[aw][m + 1];
        coef = visc / (delta * 2.0 * nplane);
      }
      double val = 0;
      static_for<Q>([&](auto I) {
        constexpr int ip = decltype(I)::value;
        const double g = coef * (cv(0, ip) * (as == 0) + cv(1, ip) * (as == 1) + cv(2, ip) * (as == 2) +
                                 P.hdt * F[as]);
        val += g * x[ip];
        if (grad) {
          if (gacc == 0.0)
            grad[ip * P.N + c] = (G)g;
          else if (wall)
            grad[ip * P.N + c] = (G)((double)grad[ip * P.N + c] + gacc * g);
        }
      });
      acc[0] += val;
      if constexpr (NV == 2) {
        // d/dF_as of coef sum_i (c_is + h F_as) f_i is coef h rho
        if (fa == as) {
          double r = x[0];
          for (int ip = 1; ip < Q; ip++) r += x[ip];
          acc[1] += coef * P.hdt * r;
        }
      }
    } else {
      const double dV = P.wd[0][idx[0] + 1] * P.wd[1][idx[1] + 1] * P.wd[2][idx[2] + 1];
      double r, j[3];
      cell_moments<double>(x, r, j);
      double u[3], uu = 0;
      for (int a = 0; a < 3; a++) {
        u[a] = j[a] / r + P.hdt * F[a];
        uu += u[a] * u[a];
      }
      acc[0] += 0.5 * r * uu * dV / vol;
      if constexpr (NV == 2) acc[1] += r * u[fa] * P.hdt * dV / vol;
      if (grad) {
        static_for<Q>([&](auto I) {
          constexpr int ip = decltype(I)::value;
          const double cu = (CX[ip] + P.hdt * F[0]) * u[0] + (CY[ip] + P.hdt * F[1]) * u[1] +
                            (CZ[ip] + P.hdt * F[2]) * u[2];
          const double g = (dV / vol) * (cu - 0.5 * uu);
          if (gacc == 0.0)
            grad[ip * P.N + c] = (G)g;
          else
            grad[ip * P.N + c] = (G)((double)grad[ip * P.N + c] + gacc * g);
        });
      }
    }
  }
  block_reduce_store<NV>(acc, partial);
}

__global__ void __launch_bounds__(256) k_axpy(double a, const double* __restrict__ x, double* y, long long n) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] += a * x[t];
}

__global__ void __launch_bounds__(256) k_fill(double* __restrict__ y, double v, long long n) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] = v;
}

__global__ void __launch_bounds__(256) k_widen(const float* __restrict__ x, double* __restrict__ y, long long n) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] = (double)x[t];
}

// Planar statistics along axis pa (statistics_serial, src/output.cu:223-334),
// in FP64 on the device.  Grid (B, n_pa): block (b, m) sums over its share
// of row m of the u, v, w, uw, uv, vw, uu, vv, ww products (the profile
// sums) and of rho, u, v, w (the global means), and takes min/max of
// rho, u, v, w.  Fixed grid, fixed-order reductions: deterministic.
constexpr int kStatBlocks = 8;  // blocks per row
constexpr int kStatSums = 13;

template <class Store>
__global__ void __launch_bounds__(kRedThreads) k_stats(KParams P, const Store* __restrict__ f, int pa,
                                                       double* __restrict__ part, double* __restrict__ mm) {
  const int m = blockIdx.y, b = blockIdx.x;
  const int a1 = pa == 0 ? 1 : 0, a2 = pa == 2 ? 1 : 2;  // the other two axes
  const long long plane = (long long)P.n[a1] * P.n[a2];
  double acc[kStatSums];
  double mn[4], mx[4];
  for (int q = 0; q < kStatSums; q++) acc[q] = 0;
  for (int q = 0; q < 4; q++) {
    mn[q] = 1e300;
    mx[q] = -1e300;
  }
  double F[3];
  ambient_d(P, F);
  for (long long s = (long long)b * kRedThreads + threadIdx.x; s < plane; s += (long long)kStatBlocks * kRedThreads) {
    int idx[3];
    idx[pa] = m;
    idx[a1] = (int)(s % P.n[a1]);
    idx[a2] = (int)(s / P.n[a1]);
    const long long c = ((long long)idx[2] * P.n[1] + idx[1]) * P.n[0] + idx[0];
    double x[Q], r, j[3];
    load_cell<double, Store>(f, c, P, x);
    cell_moments<double>(x, r, j);
    const double u = j[0] / r + F[0] * P.hdt, v = j[1] / r + F[1] * P.hdt, w = j[2] / r + F[2] * P.hdt;
    const double val[kStatSums] = {u, v, w, u * w, u * v, v * w, u * u, v * v, w * w, r, u, v, w};
    for (int q = 0; q < kStatSums; q++) acc[q] += val[q];
    const double g[4] = {r, u, v, w};
    for (int q = 0; q < 4; q++) {
      mn[q] = fmin(mn[q], g[q]);
      mx[q] = fmax(mx[q], g[q]);
    }
  }
  // sums, then 4 minima, then 4 maxima
  double all[kStatSums + 8];
  int ops[kStatSums + 8];
  for (int q = 0; q < kStatSums; q++) {
    all[q] = acc[q];
    ops[q] = 0;
  }
  for (int q = 0; q < 4; q++) {
    all[kStatSums + q] = mn[q];
    ops[kStatSums + q] = 1;
    all[kStatSums + 4 + q] = mx[q];
    ops[kStatSums + 4 + q] = 2;
  }
  block_reduce<kStatSums + 8, kRedThreads>(all, ops);
  if (threadIdx.x == 0) {
    const long long slot = (long long)m * kStatBlocks + b;
    for (int q = 0; q < kStatSums; q++) part[slot * kStatSums + q] = all[q];
    for (int q = 0; q < 8; q++) mm[slot * 8 + q] = all[kStatSums + q];
  }
}

// Per row: the profile sums; globally: rho, u, v, w sums, min, max.  out:
// [np][9] row sums, then [4] global sums, [4] min, [4] max.
__global__ void __launch_bounds__(kRedThreads) k_stats_finish(const double* __restrict__ part,
                                                              const double* __restrict__ mm, int np,
                                                              double* __restrict__ out) {
  for (int m = threadIdx.x; m < np; m += kRedThreads)
    for (int q = 0; q < 9; q++) {
      double a = 0;
      for (int b = 0; b < kStatBlocks; b++) a += part[((long long)m * kStatBlocks + b) * kStatSums + q];
      out[(long long)m * 9 + q] = a;
    }
  if (threadIdx.x < 4) {
    const int q = threadIdx.x;
    double a = 0, lo = 1e300, hi = -1e300;
    for (long long s = 0; s < (long long)np * kStatBlocks; s++) {
      a += part[s * kStatSums + 9 + q];
      lo = fmin(lo, mm[s * 8 + q]);
      hi = fmax(hi, mm[s * 8 + 4 + q]);
    }
    out[(long long)np * 9 + q] = a;
    out[(long long)np * 9 + 4 + q] = lo;
    out[(long long)np * 9 + 8 + q] = hi;
  }
}

template <class Store>
hipError_t stats_t(Ctx& c, const void* f, int pa, double* scratch, double* out) {
  const int np = c.kp.n[pa];
  double* part = scratch;
  double* mm = scratch + (size_t)np * kStatBlocks * kStatSums;
  hipLaunchKernelGGL((k_stats<Store>), dim3(kStatBlocks, np), dim3(kRedThreads), 0, c.stream, c.kp,
                     static_cast<const Store*>(f), pa, part, mm);
  hipLaunchKernelGGL(k_stats_finish, dim3(1), dim3(kRedThreads), 0, c.stream, part, mm, np, out);
  return hipGetLastError();
}

template <class Store>
hipError_t macro_t(Ctx& c, const void* f, double* rho, double* ux, double* uy, double* uz) {
  const unsigned nb = (unsigned)((c.kp.N + 255) / 256);
  hipLaunchKernelGGL((k_macro<Store>), dim3(nb), dim3(256), 0, c.stream, c.kp,
                     static_cast<const Store*>(f), rho, ux, uy, uz);
  return hipGetLastError();
}

template <class Store>
hipError_t export_t(Ctx& c, const void* f, double* dst) {
  const unsigned nb = (unsigned)((c.kp.N + 255) / 256);
  hipLaunchKernelGGL((k_export<Store>), dim3(nb), dim3(256), 0, c.stream, c.kp,
                     static_cast<const Store*>(f), dst);
  return hipGetLastError();
}

template <class Store>
hipError_t import_t(Ctx& c, const double* src, void* f) {
  const unsigned nb = (unsigned)((c.kp.N + 255) / 256);
  hipLaunchKernelGGL((k_import<Store>), dim3(nb), dim3(256), 0, c.stream, c.kp, src,
                     static_cast<Store*>(f));
  return hipGetLastError();
}

template <class Store>
hipError_t sums_t(Ctx& c, const void* f, double* out) {
  hipLaunchKernelGGL((k_sums<Store>), dim3(kRedBlocks), dim3(kRedThreads), 0, c.stream, c.kp,
                     static_cast<const Store*>(f), c.red);
  hipLaunchKernelGGL((k_reduce_final<8>), dim3(1), dim3(kRedThreads), 0, c.stream, c.red, kRedBlocks,
                     out, 0);
  return hipGetLastError();
}

template <class Store, class G>
hipError_t loss_t(Ctx& c, const void* f, int kind, double* value, G* grad, double gacc, bool vacc, int fa) {
  const Setup& s = c.s;
  const int aw = s.wall_axis();
  int as = s.stream_axis();
  if (kind == 0 && aw < 0) return hipErrorInvalidValue;
  if (kind == 0 && as == aw) as = (aw + 1) % 3;
  if (fa > 2) return hipErrorInvalidValue;
  double vol = s.width[0] * s.width[1] * s.width[2];
  double nplane = aw >= 0 ? (double)s.ncell() / s.n[aw] : 1.0;
  if (fa < 0) {
    hipLaunchKernelGGL((k_loss<Store, G>), dim3(kRedBlocks), dim3(kRedThreads), 0, c.stream, c.kp,
                       static_cast<const Store*>(f), kind, aw < 0 ? 0 : aw, as, s.visc, vol, nplane, grad,
                       gacc, c.red, 0);
    hipLaunchKernelGGL((k_reduce_final<1>), dim3(1), dim3(kRedThreads), 0, c.stream, c.red, kRedBlocks,
                       value, vacc ? 1 : 0);
  } else {
    hipLaunchKernelGGL((k_loss<Store, G, 2>), dim3(kRedBlocks), dim3(kRedThreads), 0, c.stream, c.kp,
                       static_cast<const Store*>(f), kind, aw < 0 ? 0 : aw, as, s.visc, vol, nplane, grad,
                       gacc, c.red, fa);
    hipLaunchKernelGGL((k_reduce_final<2>), dim3(1), dim3(kRedThreads), 0, c.stream, c.red, kRedBlocks,
                       value, vacc ? 1 : 0);
  }
  return hipGetLastError();
}

// Adjoint of the bulk velocity U = sum (j_a/rho) dV / V that flow_update
// reads from the new state: lam_i += ubar dU/df_i = ubar (dV/V) (c_ia - u_a)/rho,
// u_a = j_a/rho, in every cell.
template <class Store, class G>
__global__ void __launch_bounds__(256) k_flow_cotangent(KParams P, const Store* __restrict__ f, int axis,
                                                        double vol, const double* __restrict__ ubar,
                                                        G* __restrict__ lam) {
  const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
  if (c >= P.N) return;
  int i, j, k;
  cell_coords(P, c, i, j, k);
  const double dV = P.wd[0][i + 1] * P.wd[1][j + 1] * P.wd[2][k + 1];
  double x[Q], r, jm[3];
  load_cell<double, Store>(f, c, P, x);
  cell_moments<double>(x, r, jm);
  const double ua = jm[axis] / r;
  const double coef = ubar[0] * dV / (vol * r);
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    const double ca = axis == 0 ? CX[ip] : (axis == 1 ? CY[ip] : CZ[ip]);
    lam[ip * P.N + c] = (G)((double)lam[ip * P.N + c] + coef * (ca - ua));
  });
}

// Per-block sums of v[axis N + cell] over the cells (fixed order).
__global__ void __launch_bounds__(kRedThreads) k_axis_sum(long long N, const double* __restrict__ v, int axis,
                                                          double* partial) {
  double acc[1] = {0};
  const double* a = v + (long long)axis * N;
  for (long long c = (long long)blockIdx.x * kRedThreads + threadIdx.x; c < N;
       c += (long long)kRedBlocks * kRedThreads)
    acc[0] += a[c];
  block_reduce_store<1>(acc, partial);
}

// One reverse step of the controller's scalar recurrence.  After step n,
// flow_update sets F^{n+1} = F^n - flag_n (U_{n+1} - U_n)/dt
// + g (target - U_{n+1})/dt (flag_n: a previous U exists; 1 after the first
// controlled step).  With nu_m = dJ/dF^m and phi_m the cotangent of F^m
// through step m (the face partials plus the change of the summed cell force
// gradient):
//   nu_{n+1} = phi_{n+1} + nu_{n+2} + dJ/dF^{n+1} of the loss at n+1,
//   ubar_{n+1} = -(flag_n + g)/dt nu_{n+1} + nu_{n+2}/dt.
// One block.
__global__ void __launch_bounds__(kRedThreads) k_flow_adjoint_scalar(double* ctl, const double* __restrict__ fpart,
                                                                     long long nparts, int step0, double gain,
                                                                     double dt, double invT) {
  double s[1] = {0};
  for (long long b = threadIdx.x; b < nparts; b += kRedThreads) s[0] += fpart[b];
  constexpr AllSums<1> sums;
  block_reduce<1, kRedThreads>(s, sums.op);
  if (threadIdx.x == 0) {
    const double phi = s[0] + (ctl[kFlowCtlSumNew] - ctl[kFlowCtlSum]);
    const double nu_next = ctl[kFlowCtlNuNext];
    const double nu = phi + nu_next + invT * ctl[kFlowCtlLoss + 1];
    const double flag = step0 ? ctl[kFlowCtlFlag0] : 1.0;
    ctl[kFlowCtlUbar] = -(flag + gain) / dt * nu + nu_next / dt;
    ctl[kFlowCtlNuNext] = nu;
    ctl[kFlowCtlSum] = ctl[kFlowCtlSumNew];
    ctl[kFlowCtlLoss + 1] = 0;
  }
}

}  // namespace

size_t stats_scratch_doubles(int np) { return (size_t)np * kStatBlocks * (kStatSums + 8) + (size_t)np * 9 + 12; }

// The scratch of dugks_get_statistics_noalloc: partial sums and final sums
// of launch_stats for up to kNoAllocStatRows rows (2.9 MB), a file-scope
// device array, so the entry point allocates nothing.  Each GPU holds its
// own copy for the life of the process; calls that overlap on one GPU share
// it and race.  k_stats writes every partial slot k_stats_finish reads, so
// nothing needs clearing between calls.
static __device__ __align__(256) double scratch[(size_t)kNoAllocStatRows * kStatBlocks * (kStatSums + 8) +
                                                (size_t)kNoAllocStatRows * 9 + 12];

hipError_t noalloc_stats_scratch(int device, double** ptr) {
  // the symbol's host-side name is not a device address: resolve it once
  // per device (with that device current) and keep the address
  static std::mutex mu;
  static std::map<int, double*> addr;
  std::lock_guard<std::mutex> lock(mu);
  auto it = addr.find(device);
  if (it != addr.end()) {
    *ptr = it->second;
    return hipSuccess;
  }
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  addr[device] = static_cast<double*>(p);
  *ptr = static_cast<double*>(p);
  return hipSuccess;
}

hipError_t launch_stats(Ctx& c, const void* f, int pa, double* scratch, double* out) {
  if (c.precision == 0) return stats_t<double>(c, f, pa, scratch, out);
  if (c.precision == 1) return stats_t<float>(c, f, pa, scratch, out);
  if (c.precision == 2) return stats_t<__half>(c, f, pa, scratch, out);
  return stats_t<NeqHalf>(c, f, pa, scratch, out);
}

hipError_t launch_macro(Ctx& c, const void* f, double* rho, double* ux, double* uy, double* uz) {
  if (c.precision == 0) return macro_t<double>(c, f, rho, ux, uy, uz);
  if (c.precision == 1) return macro_t<float>(c, f, rho, ux, uy, uz);
  if (c.precision == 2) return macro_t<__half>(c, f, rho, ux, uy, uz);
  return macro_t<NeqHalf>(c, f, rho, ux, uy, uz);
}

hipError_t launch_export(Ctx& c, const void* f, double* dst) {
  if (c.precision == 0) return export_t<double>(c, f, dst);
  if (c.precision == 1) return export_t<float>(c, f, dst);
  if (c.precision == 2) return export_t<__half>(c, f, dst);
  return export_t<NeqHalf>(c, f, dst);
}

hipError_t launch_import(Ctx& c, const double* src, void* f) {
  if (c.precision == 0) return import_t<double>(c, src, f);
  if (c.precision == 1) return import_t<float>(c, src, f);
  if (c.precision == 2) return import_t<__half>(c, src, f);
  return import_t<NeqHalf>(c, src, f);
}

hipError_t launch_global_sums(Ctx& c, const void* f, double* dev_out8) {
  if (c.precision == 0) return sums_t<double>(c, f, dev_out8);
  if (c.precision == 1) return sums_t<float>(c, f, dev_out8);
  if (c.precision == 2) return sums_t<__half>(c, f, dev_out8);
  return sums_t<NeqHalf>(c, f, dev_out8);
}

hipError_t launch_flow_control(Ctx& c, const void* f) {
  double* sums = c.red + kRedBlocks * 8;
  hipError_t e = launch_global_sums(c, f, sums);
  if (e != hipSuccess) return e;
  int axis = c.s.stream_axis();
  hipLaunchKernelGGL(k_flow_control, dim3(1), dim3(1), 0, c.stream, c.famb_dev, sums, axis,
                     c.s.flow_rate, c.s.flow_gain, c.s.dt, c.s.hdt);
  return hipGetLastError();
}

hipError_t launch_perturb_force(Ctx& c, int step) {
  const unsigned nb = (unsigned)((c.kp.N + 255) / 256);
  hipLaunchKernelGGL(k_perturb, dim3(nb), dim3(256), 0, c.stream, c.kp, step, c.s.width[0], c.s.width[1],
                     c.s.width[2], c.s.visc, c.fcell);
  return hipGetLastError();
}

hipError_t launch_loss(Ctx& c, const void* f, int kind, double* dev_value, double* dev_grad, double grad_acc,
                       bool value_acc, int force_axis) {
  const int fa = force_axis;
  if (c.precision == 0) return loss_t<double, double>(c, f, kind, dev_value, dev_grad, grad_acc, value_acc, fa);
  if (c.precision == 1) return loss_t<float, double>(c, f, kind, dev_value, dev_grad, grad_acc, value_acc, fa);
  if (c.precision == 2) return loss_t<__half, double>(c, f, kind, dev_value, dev_grad, grad_acc, value_acc, fa);
  return loss_t<NeqHalf, double>(c, f, kind, dev_value, dev_grad, grad_acc, value_acc, fa);
}

hipError_t launch_loss_lam32(Ctx& c, const void* f, int kind, double* dev_value, float* dev_grad, double grad_acc,
                             bool value_acc, int force_axis) {
  const int fa = force_axis;
  if (c.precision == 0) return hipErrorNotSupported;
  if (c.precision == 1) return loss_t<float, float>(c, f, kind, dev_value, dev_grad, grad_acc, value_acc, fa);
  if (c.precision == 2) return loss_t<__half, float>(c, f, kind, dev_value, dev_grad, grad_acc, value_acc, fa);
  return hipErrorNotSupported;
}

hipError_t launch_axpy(Ctx& c, double a, const double* x, double* y, long long n) {
  hipLaunchKernelGGL(k_axpy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, a, x, y, n);
  return hipGetLastError();
}

hipError_t launch_fill(Ctx& c, double* dst, double v, long long n) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, dst, v, n);
  return hipGetLastError();
}

hipError_t launch_widen(Ctx& c, const float* src, double* dst, long long n) {
  hipLaunchKernelGGL(k_widen, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, src, dst, n);
  return hipGetLastError();
}

template <class G>
hipError_t flow_cotangent_t(Ctx& c, const void* f, const double* ubar, G* lam) {
  const Setup& s = c.s;
  const double vol = s.width[0] * s.width[1] * s.width[2];
  const unsigned nb = (unsigned)((c.kp.N + 255) / 256);
  const int ax = s.stream_axis();
  switch (c.precision) {
    case 0:
      hipLaunchKernelGGL((k_flow_cotangent<double, G>), dim3(nb), dim3(256), 0, c.stream, c.kp,
                         static_cast<const double*>(f), ax, vol, ubar, lam);
      break;
    case 1:
      hipLaunchKernelGGL((k_flow_cotangent<float, G>), dim3(nb), dim3(256), 0, c.stream, c.kp,
                         static_cast<const float*>(f), ax, vol, ubar, lam);
      break;
    case 2:
      hipLaunchKernelGGL((k_flow_cotangent<__half, G>), dim3(nb), dim3(256), 0, c.stream, c.kp,
                         static_cast<const __half*>(f), ax, vol, ubar, lam);
      break;
    default:
      return hipErrorNotSupported;
  }
  return hipGetLastError();
}

hipError_t launch_flow_cotangent(Ctx& c, const void* f, const double* ubar, double* lam) {
  return flow_cotangent_t<double>(c, f, ubar, lam);
}

hipError_t launch_flow_cotangent_lam32(Ctx& c, const void* f, const double* ubar, float* lam) {
  return flow_cotangent_t<float>(c, f, ubar, lam);
}

hipError_t launch_axis_sum(Ctx& c, const double* v, int axis, double* out) {
  hipLaunchKernelGGL(k_axis_sum, dim3(kRedBlocks), dim3(kRedThreads), 0, c.stream, c.kp.N, v, axis, c.red);
  hipLaunchKernelGGL((k_reduce_final<1>), dim3(1), dim3(kRedThreads), 0, c.stream, c.red, kRedBlocks, out, 0);
  return hipGetLastError();
}

hipError_t launch_flow_adjoint_scalar(Ctx& c, double* ctl, const double* fpart, long long nparts, bool step0,
                                      double invT) {
  hipLaunchKernelGGL(k_flow_adjoint_scalar, dim3(1), dim3(kRedThreads), 0, c.stream, ctl, fpart, nparts,
                     step0 ? 1 : 0, c.s.flow_gain, c.s.dt, invT);
  return hipGetLastError();
}

size_t reduction_scratch_doubles() { return (size_t)kRedBlocks * 8 + 16; }

}  // namespace dugks
