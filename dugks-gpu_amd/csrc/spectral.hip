// Spectral statistics of decaying homogeneous isotropic turbulence on the
// device: stat_te of the reference's DHIT code (src/fortran/dhit/src/stat.f90:5-205),
// with hipFFT in place of FFTW and FP64 sums throughout.
//
//   te        0.5 <|u|^2> in physical space (the mean flow included)
//   E(k)      0.5 |u_hat|^2 of the fluctuation summed over shells
//             k = nint(|kappa|), kappa the integer wavenumber vector
//             (FFT normalised by N, the mean mode left out)
//   eps       2 nu sum |k|^2 0.5 |u_hat|^2 with k = 2 pi kappa / L
//   Kolmogorov, Taylor and large-eddy scales from etotal = sum E and eps
//   skewness, flatness of the longitudinal derivatives du/dx, dv/dy, dw/dz
//             (spectral derivatives, the Nyquist mode of the derivative
//             direction left out as the real part of the reference's complex
//             inverse transform leaves it out)
//
// Shell sums are deterministic: each one-wave block folds its modes into LDS
// bins lane group by lane group, and the block partials are summed per bin in
// block order.
#include <hipfft/hipfft.h>

#include <cmath>
#include <vector>

#include "device_common.h"

namespace dugks {

namespace {

constexpr int kSpecBlocks = 1024;  // fixed grid: fixed summation order
constexpr int kMaxBins = 1024;  // shells of grids up to about 1180 points per axis

// integer wavenumber of FFT index m of an axis of n points (the reference's
// convention: m < n/2 positive, m >= n/2 negative, the Nyquist index -n/2)
__device__ __forceinline__ int kappa(int m, int n) { return m < n / 2 ? m : m - n; }

// Adds 0.5 w |u_hat / N|^2 of one component into E (the half spectrum,
// nz x ny x (nx/2+1); w = 2 for the modes whose conjugate partner the half
// spectrum leaves out, else 1).  The mean mode is skipped.
__global__ void __launch_bounds__(256) k_spec_energy(const hipfftDoubleComplex* __restrict__ hat, int nx, int ny,
                                                     int nz, double invN, int first, double* __restrict__ E) {
  const int nh = nx / 2 + 1;
  const long long m = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long nm = (long long)nh * ny * nz;
  if (m >= nm) return;
  const int i = (int)(m % nh);
  const double w = (i == 0 || (nx % 2 == 0 && i == nx / 2)) ? 1.0 : 2.0;
  const double re = hat[m].x * invN, im = hat[m].y * invN;
  double e = m == 0 ? 0.0 : 0.5 * w * (re * re + im * im);
  E[m] = first ? e : E[m] + e;
}

// Shell sums of E and 2 nu k^2 E over the half spectrum into
// part[block][2][nb] (nb bins, all shells present in the spectrum).
__global__ void __launch_bounds__(64) k_spec_shells(const double* __restrict__ E, int nx, int ny, int nz,
                                                    double kunit, double visc, int nb, double* __restrict__ part) {
  __shared__ double sE[kMaxBins], sD[kMaxBins];
  const int lane = threadIdx.x;
  for (int b = lane; b < nb; b += 64) {
    sE[b] = 0;
    sD[b] = 0;
  }
  __syncthreads();
  const int nh = nx / 2 + 1;
  const long long nm = (long long)nh * ny * nz;
  const long long chunks = (nm + 63) / 64;
  const long long per = (chunks + gridDim.x - 1) / gridDim.x;
  const long long c0 = (long long)blockIdx.x * per, c1 = min(chunks, c0 + per);
  for (long long ch = c0; ch < c1; ch++) {
    const long long m = ch * 64 + lane;
    const bool valid = m < nm;
    int bin = 0;
    double e = 0, d = 0;
    if (valid) {
      const int i = (int)(m % nh);
      const long long t = m / nh;
      const int j = (int)(t % ny), k = (int)(t / ny);
      const int kx = kappa(i, nx), ky = kappa(j, ny), kz = kappa(k, nz);
      const double k2 = (double)kx * kx + (double)ky * ky + (double)kz * kz;
      bin = (int)llround(sqrt(k2));
      e = E[m];
      d = 2.0 * visc * k2 * kunit * kunit * e;
    }
    // fold the lanes bin by bin: each pass takes the bin of the first pending
    // lane and sums that bin's lanes with a butterfly
    unsigned long long pending = __ballot(valid);
    while (pending) {
      const int leader = __ffsll((long long)pending) - 1;
      const int b = __shfl(bin, leader, 64);
      const bool mine = valid && bin == b;
      const double se = wave_sum(mine ? e : 0.0), sd = wave_sum(mine ? d : 0.0);
      if (lane == 0) {
        sE[b] += se;
        sD[b] += sd;
      }
      pending &= ~__ballot(mine);
    }
  }
  __syncthreads();
  double* out = part + (long long)blockIdx.x * 2 * nb;
  for (int b = lane; b < nb; b += 64) {
    out[b] = sE[b];
    out[nb + b] = sD[b];
  }
}

// out[b] = sum over blocks of part[block][0][b], out[nb + b] likewise for D
__global__ void __launch_bounds__(256) k_spec_shells_final(const double* __restrict__ part, int nblocks, int nb,
                                                           double* __restrict__ out) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= 2 * nb) return;
  double s = 0;
  for (int q = 0; q < nblocks; q++) s += part[(long long)q * 2 * nb + b];
  out[b] = s;
}

// hat *= i k_a / N along axis a (k = 2 pi kappa / L), the derivative
// direction's Nyquist mode and the mean mode set to zero.
__global__ void __launch_bounds__(256) k_spec_deriv(hipfftDoubleComplex* __restrict__ hat, int nx, int ny, int nz,
                                                    int a, double kunit, double invN) {
  const int nh = nx / 2 + 1;
  const long long m = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long nm = (long long)nh * ny * nz;
  if (m >= nm) return;
  const int i = (int)(m % nh);
  const long long t = m / nh;
  const int j = (int)(t % ny), k = (int)(t / ny);
  const int idx[3] = {i, j, k}, n[3] = {nx, ny, nz};
  const bool nyq = n[a] % 2 == 0 && idx[a] == n[a] / 2;
  const double ka = nyq ? 0.0 : kunit * (a == 0 ? (double)i : (double)kappa(idx[a], n[a]));
  const double re = hat[m].x, im = hat[m].y;
  // i k (re + i im) = -k im + i k re
  hat[m].x = -ka * im * invN;
  hat[m].y = ka * re * invN;
}

// sums of d^2, d^3, d^4 of a real field (n values) into part[block][3]
__global__ void __launch_bounds__(256) k_moments234(const double* __restrict__ d, long long n,
                                                   double* __restrict__ part) {
  double v[3] = {0, 0, 0};
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long long)kSpecBlocks * 256) {
    const double x = d[q], x2 = x * x;
    v[0] += x2;
    v[1] += x2 * x;
    v[2] += x2 * x2;
  }
  constexpr int ops[3] = {0, 0, 0};
  block_reduce<3, 256>(v, ops);
  if (threadIdx.x == 0)
    for (int q = 0; q < 3; q++) part[blockIdx.x * 3 + q] = v[q];
}

// sums of u^2 + v^2 + w^2 of the three fields into part[block]
__global__ void __launch_bounds__(256) k_energy_phys(const double* __restrict__ ux, const double* __restrict__ uy,
                                                    const double* __restrict__ uz, long long n,
                                                    double* __restrict__ part) {
  double v[1] = {0};
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long long)kSpecBlocks * 256)
    v[0] += ux[q] * ux[q] + uy[q] * uy[q] + uz[q] * uz[q];
  constexpr int ops[1] = {0};
  block_reduce<1, 256>(v, ops);
  if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

// out[q] = sum over blocks of part[block][q], q < nv (one thread per value)
__global__ void __launch_bounds__(64) k_sum_parts(const double* __restrict__ part, int nblocks, int nv,
                                                  double* __restrict__ out) {
  const int q = threadIdx.x;
  if (q >= nv) return;
  double s = 0;
  for (int b = 0; b < nblocks; b++) s += part[(long long)b * nv + q];
  out[q] = s;
}

}  // namespace

int spectral_bins(const Ctx& c) {
  const double h[3] = {c.kp.n[0] / 2.0, c.kp.n[1] / 2.0, c.kp.n[2] / 2.0};
  return (int)std::llround(std::sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2])) + 1;
}

// Spectral statistics of the current state (see the file comment).  shells:
// [2][nb] (E, D) of all nb = spectral_bins(c) shells; sums[0..3]: physical
// 0.5 sum |u|^2 / N, then the sums of d^2, d^3, d^4 over the three
// longitudinal derivatives.  Host buffers; synchronises.
hipError_t spectral_statistics(Ctx& c, double* shells, double sums[4]) {
  const KParams& P = c.kp;
  const int nx = P.n[0], ny = P.n[1], nz = P.n[2];
  const long long N = P.N;
  const int nb = spectral_bins(c);
  if (nb > kMaxBins) return hipErrorInvalidValue;
  const long long nm = (long long)(nx / 2 + 1) * ny * nz;
  const double invN = 1.0 / (double)N;
  const double kunit = 2.0 * M_PI / c.s.width[0];
  hipError_t e = hipSuccess;
  double *macro = nullptr, *E = nullptr, *part = nullptr, *dev_out = nullptr;
  hipfftDoubleComplex* hat = nullptr;
  hipfftHandle fwd = 0, inv = 0;
  bool have_fwd = false, have_inv = false;
  auto cleanup = [&]() {
    if (have_fwd) hipfftDestroy(fwd);
    if (have_inv) hipfftDestroy(inv);
    hipFree(macro);
    hipFree(E);
    hipFree(part);
    hipFree(dev_out);
    hipFree(hat);
  };
  auto fft_ok = [&](hipfftResult r) { return r == HIPFFT_SUCCESS ? hipSuccess : hipErrorUnknown; };
  const size_t part_doubles = std::max((size_t)kSpecBlocks * 2 * nb, (size_t)kSpecBlocks * 3);
  if ((e = hipMalloc((void**)&macro, 4 * N * sizeof(double))) != hipSuccess ||
      (e = hipMalloc((void**)&E, nm * sizeof(double))) != hipSuccess ||
      (e = hipMalloc((void**)&part, part_doubles * sizeof(double))) != hipSuccess ||
      (e = hipMalloc((void**)&dev_out, (2 * nb + 16) * sizeof(double))) != hipSuccess ||
      (e = hipMalloc((void**)&hat, nm * sizeof(hipfftDoubleComplex))) != hipSuccess) {
    cleanup();
    return e;
  }
  // cuFFT ordering: the first size is the slowest axis ([k][j][i] storage)
  if ((e = fft_ok(hipfftPlan3d(&fwd, nz, ny, nx, HIPFFT_D2Z))) == hipSuccess) have_fwd = true;
  if (e == hipSuccess && (e = fft_ok(hipfftPlan3d(&inv, nz, ny, nx, HIPFFT_Z2D))) == hipSuccess) have_inv = true;
  if (e == hipSuccess) e = fft_ok(hipfftSetStream(fwd, c.stream));
  if (e == hipSuccess) e = fft_ok(hipfftSetStream(inv, c.stream));
  if (e != hipSuccess) {
    cleanup();
    return e;
  }
  double* u[3] = {macro + N, macro + 2 * N, macro + 3 * N};  // after rho
  e = launch_macro(c, c.f[c.cur], macro, u[0], u[1], u[2]);
  // physical kinetic energy
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_energy_phys, dim3(kSpecBlocks), dim3(256), 0, c.stream, u[0], u[1], u[2], N, part);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(64), 0, c.stream, part, kSpecBlocks, 1, dev_out + 2 * nb);
    e = hipGetLastError();
  }
  const unsigned mb = (unsigned)((nm + 255) / 256);
  // energy density of the three components, then the shell sums
  for (int a = 0; a < 3 && e == hipSuccess; a++) {
    e = fft_ok(hipfftExecD2Z(fwd, u[a], hat));
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_spec_energy, dim3(mb), dim3(256), 0, c.stream, hat, nx, ny, nz, invN, a == 0 ? 1 : 0, E);
      e = hipGetLastError();
    }
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_spec_shells, dim3(kSpecBlocks), dim3(64), 0, c.stream, E, nx, ny, nz, kunit, c.s.visc, nb,
                       part);
    hipLaunchKernelGGL(k_spec_shells_final, dim3((2 * nb + 255) / 256), dim3(256), 0, c.stream, part, kSpecBlocks,
                       nb, dev_out);
    e = hipGetLastError();
  }
  // longitudinal derivatives: sums of d^2, d^3, d^4 per component
  // (the field d overwrites the velocity component it came from)
  for (int a = 0; a < 3 && e == hipSuccess; a++) {
    e = fft_ok(hipfftExecD2Z(fwd, u[a], hat));
    if (e != hipSuccess) break;
    hipLaunchKernelGGL(k_spec_deriv, dim3(mb), dim3(256), 0, c.stream, hat, nx, ny, nz, a, kunit, invN);
    if ((e = hipGetLastError()) != hipSuccess) break;
    e = fft_ok(hipfftExecZ2D(inv, hat, u[a]));
    if (e != hipSuccess) break;
    hipLaunchKernelGGL(k_moments234, dim3(kSpecBlocks), dim3(256), 0, c.stream, u[a], N, part);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(64), 0, c.stream, part, kSpecBlocks, 3, dev_out + 2 * nb + 1 + 3 * a);
    e = hipGetLastError();
  }
  std::vector<double> host(2 * nb + 10);
  if (e == hipSuccess)
    e = hipMemcpyAsync(host.data(), dev_out, (2 * nb + 10) * sizeof(double), hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  cleanup();
  if (e != hipSuccess) return e;
  for (int q = 0; q < 2 * nb; q++) shells[q] = host[q];
  sums[0] = 0.5 * host[2 * nb] * invN;
  for (int q = 0; q < 3; q++) sums[1 + q] = host[2 * nb + 1 + q] + host[2 * nb + 4 + q] + host[2 * nb + 7 + q];
  return hipSuccess;
}

}  // namespace dugks
