// Device helpers shared by the step, adjoint and utility kernels.
#pragma once

#include <type_traits>
#include <utility>

#include "dugks_internal.h"

namespace dugks {

// Compile-time loop: f(std::integral_constant<int, I>) for I = 0..N-1, so that
// lattice components are constants in every unrolled body.
template <class F, int... Is>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, Is...>) {
  (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Storage codecs: Store is the element type of f in HBM, Real the
// arithmetic type.
template <class Store>
struct Codec;

template <>
struct Codec<double> {
  template <class Real>
  __device__ __forceinline__ static Real load(const double* p, long long i, const KParams&, int) {
    return (Real)p[i];
  }
  template <class Real>
  __device__ __forceinline__ static void store(double* p, long long i, Real v, const KParams&, int) {
    p[i] = (double)v;
  }
};

template <>
struct Codec<float> {
  template <class Real>
  __device__ __forceinline__ static Real load(const float* p, long long i, const KParams&, int) {
    return (Real)p[i];
  }
  template <class Real>
  __device__ __forceinline__ static void store(float* p, long long i, Real v, const KParams&, int) {
    p[i] = (float)v;
  }
};

// FP16 deviation storage: s_i = (f_i - w_i rho0) / w_i, f_i = w_i (rho0 + s_i).
template <>
struct Codec<__half> {
  template <class Real>
  __device__ __forceinline__ static Real load(const __half* p, long long i, const KParams& P, int ip) {
    return (Real)((float)wgt(ip) * (P.rho0f + __half2float(p[i])));
  }
  template <class Real>
  __device__ __forceinline__ static void store(__half* p, long long i, Real v, const KParams& P, int ip) {
    p[i] = __float2half((float)v * (float)(1.0 / wgt(ip)) - P.rho0f);
  }
};

// FP16 non-equilibrium storage (the element tag of that mode's state arrays).
// A state of N cells is one buffer: [4][N] float rho - 1, u_x, u_y, u_z
// (u = j/rho, the momentum velocity of f; the density as its difference from
// 1 keeps FP32 digits for the small density variations), then [Q][N] half
// s_i = n_i kNeqScale / w_i with n_i = f_i - feq_i(rho, u).  The non-equilibrium part is small against
// w_i, so FP16 rounds it far less than it would round f or f - w_i rho0; the
// scale keeps typical |n_i| / w_i (1e-6 .. 1e-2) in FP16's normal range.
struct NeqHalf {
  unsigned short bits;
};
constexpr float kNeqScale = 1024.0f;
// bytes of one state of N cells
__host__ __device__ constexpr unsigned long long neq_state_bytes(unsigned long long N) { return N * (16 + 2 * 19); }
// bytes of one state of N cells stored as Store
template <class Store>
__host__ __device__ constexpr unsigned long long state_bytes(unsigned long long N) {
  if constexpr (std::is_same<Store, NeqHalf>::value)
    return neq_state_bytes(N);
  else
    return N * 19 * sizeof(Store);
}

// Raw buffer access to a state array: one buffer resource in SGPRs, a 32-bit
// per-lane byte offset and a uniform byte offset per population, so 19
// population loads share one address register.  Requires the array to span
// fewer than 2^32 bytes; reads past the end return 0 and stores past it are
// dropped by the hardware range check.
template <class Store>
struct StateBuf {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ StateBuf(const Store* p, unsigned bytes)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<Store*>(p), 0, (int)bytes, 0x00020000)) {}
  __device__ __forceinline__ Store raw(unsigned vb, unsigned sb) const;
  __device__ __forceinline__ void put(Store v, unsigned vb, unsigned sb) const;
};
template <>
__device__ __forceinline__ double StateBuf<double>::raw(unsigned vb, unsigned sb) const {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)vb, (int)sb, 0));
}
template <>
__device__ __forceinline__ float StateBuf<float>::raw(unsigned vb, unsigned sb) const {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)vb, (int)sb, 0));
}
template <>
__device__ __forceinline__ __half StateBuf<__half>::raw(unsigned vb, unsigned sb) const {
  return __ushort_as_half(__builtin_amdgcn_raw_buffer_load_b16(r, (int)vb, (int)sb, 0));
}
// The 16 bits of an FP16 state value zero-extended to 32 (see RawCell).
__device__ __forceinline__ unsigned raw_b16_32(const __amdgpu_buffer_rsrc_t& r, unsigned vb, unsigned sb) {
  return (unsigned)__builtin_amdgcn_raw_buffer_load_b16(r, (int)vb, (int)sb, 0);
}
template <>
__device__ __forceinline__ void StateBuf<double>::put(double v, unsigned vb, unsigned sb) const {
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)vb, (int)sb, 0);
}
template <>
__device__ __forceinline__ void StateBuf<float>::put(float v, unsigned vb, unsigned sb) const {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)vb, (int)sb, 0);
}
template <>
__device__ __forceinline__ void StateBuf<__half>::put(__half v, unsigned vb, unsigned sb) const {
  __builtin_amdgcn_raw_buffer_store_b16(__half_as_ushort(v), r, (int)vb, (int)sb, 0);
}

template <>
struct StateBuf<NeqHalf> {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ StateBuf(const NeqHalf* p, unsigned bytes)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<NeqHalf*>(p), 0, (int)bytes, 0x00020000)) {}
  __device__ __forceinline__ __half raw(unsigned vb, unsigned sb) const {
    return __ushort_as_half(__builtin_amdgcn_raw_buffer_load_b16(r, (int)vb, (int)sb, 0));
  }
  // the 16 bits zero-extended to 32 (see RawCell)
  __device__ __forceinline__ unsigned raw32(unsigned vb, unsigned sb) const {
    return (unsigned)__builtin_amdgcn_raw_buffer_load_b16(r, (int)vb, (int)sb, 0);
  }
  __device__ __forceinline__ void put(__half v, unsigned vb, unsigned sb) const {
    __builtin_amdgcn_raw_buffer_store_b16(__half_as_ushort(v), r, (int)vb, (int)sb, 0);
  }
  __device__ __forceinline__ float rawf(unsigned vb, unsigned sb) const {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)vb, (int)sb, 0));
  }
  __device__ __forceinline__ void putf(float v, unsigned vb, unsigned sb) const {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)vb, (int)sb, 0);
  }
};

// Raw buffer view of a plain array (< 2 GiB): per-lane 32-bit byte offset plus
// a uniform byte offset; loads past the end return 0, stores past it are
// dropped.
struct RawBuf {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ RawBuf(const void* p, unsigned bytes)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, 0x00020000)) {}
  __device__ __forceinline__ double ld_f64(unsigned vb, unsigned sb) const {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)vb, (int)sb, 0));
  }
  template <class T>
  __device__ __forceinline__ T ld(unsigned vb, unsigned sb) const;
  __device__ __forceinline__ void st(float v, unsigned vb, unsigned sb) const {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)vb, (int)sb, 0);
  }
  __device__ __forceinline__ void st(double v, unsigned vb, unsigned sb) const {
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)vb, (int)sb, 0);
  }
};
template <>
__device__ __forceinline__ float RawBuf::ld<float>(unsigned vb, unsigned sb) const {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)vb, (int)sb, 0));
}
template <>
__device__ __forceinline__ double RawBuf::ld<double>(unsigned vb, unsigned sb) const {
  return ld_f64(vb, sb);
}
// A byte offset no RawBuf of this code base reaches: loads through it read 0.
constexpr unsigned kOutOfRange = 0x80000000u;

// Decode / encode one population through a StateBuf (the Codec rules).
template <class Real, class Store>
__device__ __forceinline__ Real ld_state(const StateBuf<Store>& b, unsigned vb, unsigned sb, const KParams& P,
                                         int ip) {
  if constexpr (std::is_same<Store, __half>::value)
    return (Real)((float)wgt(ip) * (P.rho0f + __half2float(b.raw(vb, sb))));
  else
    return (Real)b.raw(vb, sb);
}
template <class Real, class Store>
__device__ __forceinline__ void st_state(const StateBuf<Store>& b, unsigned vb, unsigned sb, Real v,
                                         const KParams& P, int ip) {
  if constexpr (std::is_same<Store, __half>::value)
    b.put(__float2half((float)v * (float)(1.0 / wgt(ip)) - P.rho0f), vb, sb);
  else
    b.put((Store)v, vb, sb);
}

// a * b + c rounded once
__device__ __forceinline__ float fma_rn(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_rn(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Sum over the 64 lanes of a wave by an xor butterfly: a fixed order, and
// every lane ends with the same bits (each level adds the same two values).
__device__ __forceinline__ double wave_sum(double v) {
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
  for (int m = 32; m >= 1; m >>= 1) v = fmin(v, __shfl_xor(v, m, 64));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, 64));
  return v;
}

// Block reduction of NV values per thread (block of NT threads, a multiple
// of 64): wave butterflies, then the wave results in wave order.  Op 0 sums,
// 1 takes the minimum, 2 the maximum (per value, from ops[q]).  The result
// is valid in thread 0.  Deterministic for a fixed block size.
template <int NV, int NT>
__device__ __forceinline__ void block_reduce(double (&v)[NV], const int (&ops)[NV]) {
  constexpr int NW = NT / 64;
  __shared__ double part[NV][NW];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int q = 0; q < NV; q++) {
    v[q] = ops[q] == 0 ? wave_sum(v[q]) : (ops[q] == 1 ? wave_min(v[q]) : wave_max(v[q]));
    if (lane == 0) part[q][w] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int q = 0; q < NV; q++) {
      double a = part[q][0];
      for (int k = 1; k < NW; k++)
        a = ops[q] == 0 ? a + part[q][k] : (ops[q] == 1 ? fmin(a, part[q][k]) : fmax(a, part[q][k]));
      v[q] = a;
    }
}

// 1/x: correctly rounded in FP64; the hardware reciprocal (1 ulp) in FP32.
template <class Real>
__device__ __forceinline__ Real rcp_real(Real x) {
  return Real(1) / x;
}
template <>
__device__ __forceinline__ float rcp_real<float>(float x) {
  return __builtin_amdgcn_rcpf(x);
}

// Per-axis weight table entry in the arithmetic type (FP32 kernels read a
// float copy of the FP64 tables).
template <class Real>
__device__ __forceinline__ Real tabv(const KParams& P, int a, int idx);
template <>
__device__ __forceinline__ double tabv<double>(const KParams& P, int a, int idx) {
  return P.tab[a][idx];
}
template <>
__device__ __forceinline__ float tabv<float>(const KParams& P, int a, int idx) {
  return P.tabf[a][idx];
}

// c_ip . (x, y, z) over the non-zero components only.  IEEE arithmetic
// cannot drop a product with a zero component (0 * inf, the sign of zero),
// so writing the dot product out would cost a multiply-add per zero
// component.  Zero for the rest population.
template <int ip, class Real>
__device__ __forceinline__ Real cdot(Real x, Real y, Real z) {
  constexpr int cx = CX[ip], cy = CY[ip], cz = CZ[ip];
  if constexpr (cx != 0) {
    Real s = cx > 0 ? x : -x;
    if constexpr (cy != 0) s = cy > 0 ? s + y : s - y;
    if constexpr (cz != 0) s = cz > 0 ? s + z : s - z;
    return s;
  } else if constexpr (cy != 0) {
    Real s = cy > 0 ? y : -y;
    if constexpr (cz != 0) s = cz > 0 ? s + z : s - z;
    return s;
  } else if constexpr (cz != 0) {
    return cz > 0 ? z : -z;
  } else {
    return Real(0);
  }
}

// a + c_ip . (x, y, z), the same way (a for the rest population).
template <int ip, class Real>
__device__ __forceinline__ Real plus_cdot(Real a, Real x, Real y, Real z) {
  if constexpr (ip == 0)
    return a;
  else
    return a + cdot<ip>(x, y, z);
}

// Weight class of a population: 0 rest (1/3), 1 axis (1/18), 2 diagonal (1/36).
__device__ constexpr int wclass(int ip) { return ip == 0 ? 0 : (ip <= 6 ? 1 : 2); }
__device__ constexpr double wclass_weight(int k) { return k == 0 ? W0 : (k == 1 ? W1 : W2); }

// Equilibrium-with-force term of the centre stage and the face collision,
//   G_i = (P + Kf (c_i - u).F / RT) feq_i(rho, u),
// feq_i = w_i rho (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u.u), held as the
// per-weight-class products w rho (P - 3 Kf u.F) and 3 Kf w rho, so that
// each population costs c.u, c.F and four fused multiply-adds.
// Tag of the constructors that form the per-class products in FP64 and round
// them to the arithmetic type once (see centre_fbar).
struct ExactProducts {};

template <class Real>
struct EqTerm {
  Real ux, uy, uz, base, A[3], B[3];
  __device__ __forceinline__ EqTerm(Real rho, Real ux_, Real uy_, Real uz_, const Real F[3], Real P, Real Kf)
      : ux(ux_), uy(uy_), uz(uz_) {
    base = Real(1) - Real(1.5) * (ux * ux + uy * uy + uz * uz);  // 1 - u.u/(2RT)
    const Real uF = ux * F[0] + uy * F[1] + uz * F[2];
    const Real P0 = P - Real(3) * Kf * uF;
    const Real K3 = Real(3) * Kf;
    for (int k = 0; k < 3; k++) {
      const Real wr = Real(wclass_weight(k)) * rho;
      A[k] = wr * P0;
      B[k] = wr * K3;
    }
  }
  __device__ __forceinline__ EqTerm(ExactProducts, Real rho, Real ux_, Real uy_, Real uz_, const Real F[3], Real P,
                                    Real Kf)
      : ux(ux_), uy(uy_), uz(uz_) {
    base = Real(1) - Real(1.5) * (ux * ux + uy * uy + uz * uz);
    const double uF = (double)ux * F[0] + (double)uy * F[1] + (double)uz * F[2];
    const double P0 = (double)P - 3.0 * (double)Kf * uF;
    const double K3 = 3.0 * (double)Kf;
    for (int k = 0; k < 3; k++) {
      const double wr = wclass_weight(k) * (double)rho;
      A[k] = (Real)(wr * P0);
      B[k] = (Real)(wr * K3);
    }
  }
  // G_ip given c_ip.F (cF, unused for the rest population)
  template <int ip>
  __device__ __forceinline__ Real term(Real cF) const {
    if constexpr (ip == 0) {
      return A[0] * base;
    } else {
      const Real cu = cdot<ip>(ux, uy, uz);
      const Real e = cu * (cu * Real(4.5) + Real(3)) + base;
      return (cF * B[wclass(ip)] + A[wclass(ip)]) * e;
    }
  }
  // f - b f + G_ip given nb = -b, with f - b f formed by one fused
  // multiply-add: the populations then sum to (1 - b) rho + sum G exactly up
  // to unbiased rounding, whereas b' f with b' = 1 - b rounded to the
  // arithmetic type would scale the mass by the rounding of b' on every
  // application (a steady drift in FP32)
  template <int ip>
  __device__ __forceinline__ Real relax_keep(Real nb, Real f, Real cF) const {
    const Real kept = fma_rn(nb, f, f);
    if constexpr (ip == 0) {
      return fma_rn(A[0], base, kept);
    } else {
      const Real cu = cdot<ip>(ux, uy, uz);
      const Real e = cu * (cu * Real(4.5) + Real(3)) + base;
      return fma_rn(cF * B[wclass(ip)] + A[wclass(ip)], e, kept);
    }
  }
  // s f + G_ip
  template <int ip>
  __device__ __forceinline__ Real relax(Real s, Real f, Real cF) const {
    if constexpr (ip == 0) {
      return A[0] * base + s * f;
    } else {
      const Real cu = cdot<ip>(ux, uy, uz);
      const Real e = cu * (cu * Real(4.5) + Real(3)) + base;
      return (cF * B[wclass(ip)] + A[wclass(ip)]) * e + s * f;
    }
  }
};

// ---- shifted populations (FP32 arithmetic of the fused step) ----
// The fused FP32 step carries h_i = f_i - w_i (the populations less the rest
// state of density 1) instead of f_i.  Its linear parts (the reconstruction,
// the update, the wall extrapolation and bounce-back) have weights that sum
// to one or take differences, so they act on h as on f, and only the
// equilibrium terms change form.  |h_i| is a few tenths of w_i, so every
// rounding is that much smaller than on f, and the rest state drops out of
// the fluxes exactly instead of to rounding: in FP32 the accumulated
// rounding of 10^4 steps otherwise reaches 1e-3 of the rms velocities.
//
// The relaxation f' = (1 - b) f + G_i, G_i = w_i rho e_i (b + K (c_i - u).F),
// e_i = 1 + eps_i, eps_i = 3 c.u + 4.5 (c.u)^2 - 1.5 u.u, K = 3 Kf, in shifted
// form (drho = rho - 1):
//   h' = (1 - b) h + C0 + B c.F + (P + B c.F) eps_i,
//   C0 = w (b drho - K rho u.F),  P = w rho (b - K u.F),  B = w K rho.
template <class Real>
struct EqShift {
  Real ux, uy, uz, nuu, C0[3], P[3], B[3];
  __device__ __forceinline__ EqShift(Real drho, Real ux_, Real uy_, Real uz_, const Real F[3], Real b, Real Kf)
      : ux(ux_), uy(uy_), uz(uz_) {
    const Real rho = Real(1) + drho;
    nuu = Real(-1.5) * (ux * ux + uy * uy + uz * uz);
    const Real uF = ux * F[0] + uy * F[1] + uz * F[2];
    const Real K3 = Real(3) * Kf;
    const Real c0 = b * drho - K3 * rho * uF;
    const Real p0 = rho * (b - K3 * uF);
    const Real b0 = K3 * rho;
    for (int k = 0; k < 3; k++) {
      const Real w = Real(wclass_weight(k));
      C0[k] = w * c0;
      P[k] = w * p0;
      B[k] = w * b0;
    }
  }
  __device__ __forceinline__ EqShift(ExactProducts, Real drho, Real ux_, Real uy_, Real uz_, const Real F[3], Real b,
                                     Real Kf)
      : ux(ux_), uy(uy_), uz(uz_) {
    const double dr = (double)drho, rho = 1.0 + dr;
    nuu = Real(-1.5) * (ux * ux + uy * uy + uz * uz);
    const double uF = (double)ux * F[0] + (double)uy * F[1] + (double)uz * F[2];
    const double K3 = 3.0 * (double)Kf;
    const double c0 = (double)b * dr - K3 * rho * uF;
    const double p0 = rho * ((double)b - K3 * uF);
    const double b0 = K3 * rho;
    for (int k = 0; k < 3; k++) {
      const double w = wclass_weight(k);
      C0[k] = (Real)(w * c0);
      P[k] = (Real)(w * p0);
      B[k] = (Real)(w * b0);
    }
  }
  // (1 - b) h + (G_ip - b w) given nb = -b, with h - b h one fused multiply-add
  template <int ip>
  __device__ __forceinline__ Real relax(Real nb, Real h, Real cF) const {
    const Real kept = fma_rn(nb, h, h);
    if constexpr (ip == 0) {
      return fma_rn(P[0], nuu, C0[0] + kept);
    } else {
      constexpr int k = wclass(ip);
      const Real cu = cdot<ip>(ux, uy, uz);
      const Real eps = cu * (cu * Real(4.5) + Real(3)) + nuu;
      return fma_rn(fma_rn(cF, B[k], P[k]), eps, fma_rn(cF, B[k], C0[k] + kept));
    }
  }
};

// The equilibrium part of a cell's shifted populations under FP16
// non-equilibrium storage: w_i rho eps_i + w_i drho, eps_i = 3 c.u +
// 4.5 (c.u)^2 - 1.5 u.u, from the stored drho and u (both codecs directions
// use it, so a stored cell decodes to what was encoded up to the FP16
// rounding of s).
template <class Real>
struct NeqEq {
  Real ux, uy, uz, nuu, Pw[3], Cw[3];
  __device__ __forceinline__ NeqEq(Real drho, const Real (&u)[3]) : ux(u[0]), uy(u[1]), uz(u[2]) {
    nuu = Real(-1.5) * (ux * ux + uy * uy + uz * uz);
    // each product rounded once (w (1 + drho) as one fused multiply-add)
    for (int k = 0; k < 3; k++) {
      const Real w = Real(wclass_weight(k));
      Pw[k] = fma_rn(w, drho, w);
      Cw[k] = w * drho;
    }
  }
  template <int ip>
  __device__ __forceinline__ Real eps() const {
    if constexpr (ip == 0) {
      return nuu;
    } else {
      const Real cu = cdot<ip>(ux, uy, uz);
      return cu * (cu * Real(4.5) + Real(3)) + nuu;
    }
  }
  // w_i rho eps_i + w_i drho
  template <int ip>
  __device__ __forceinline__ Real eq() const {
    return fma_rn(Pw[wclass(ip)], eps<ip>(), Cw[wclass(ip)]);
  }
  // h_i from the stored s_i
  template <int ip>
  __device__ __forceinline__ Real shift_of(Real s) const {
    return fma_rn(Pw[wclass(ip)], eps<ip>(), fma_rn(s, Real(wgt(ip) / (double)kNeqScale), Cw[wclass(ip)]));
  }
};

// Centre stage in shifted populations: hb_i = f-bar+_i - w_i from h_i, the
// cell's drho = rho - 1 and momentum velocity u0 (with FP16 non-equilibrium
// storage the stored ones, which the decoded populations match to the FP16
// rounding of their non-equilibrium parts).
template <class Real>
__device__ __forceinline__ void centre_fbar_shift(const Real (&h)[Q], Real drho, const Real (&u0)[3],
                                                  const Real F[3], Real hdt, Real bw0, Real bw2, Real (&hb)[Q]) {
  const EqShift<Real> eq(ExactProducts{}, drho, u0[0] + hdt * F[0], u0[1] + hdt * F[1], u0[2] + hdt * F[2], F, bw2,
                         bw0);
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    hb[ip] = eq.template relax<ip>(-bw2, h[ip], cdot<ip>(F[0], F[1], F[2]));
  });
}

// Centre-stage f-bar+ of one cell (calc_fb1 of the reference, src/dugks.cu:1131-1196):
//   rho = sum f, u = j/rho + h F, feq_i(rho, u),
//   fb_i = bw1 f_i + (bw2 + bw0 (c_i - u).F / RT) feq_i.
// This form takes rho and the momentum velocity u0 = j/rho of the cell (from
// cell_moments and moment_velocity below, or from FP16 non-equilibrium
// storage, which holds them).
// The stage conserves the cell's mass: sum fb = rho.  In FP32 the weights
// and the rounded coefficients bias that sum by ~1e-8 of rho per step
// (times bw2, which exceeds 1 when dt >> tau), a steady drift of the total
// mass; there the rest population takes what the moving ones leave of rho
// (the same value in exact arithmetic; rho must then be the sum of f), so
// the stage keeps the mass to one unbiased rounding.  bw1 = 1 - bw2 enters
// as f - bw2 f (EqTerm::relax_keep).
// The per-class products w rho (bw2 - 3 bw0 u.F) and 3 bw0 w rho are formed
// in FP64 and rounded once (ExactProducts): rounded twice in FP32 (w rho,
// then times bw2) they took 0.9 % of the Re = 1600 Taylor-Green vortex's
// dissipation away (0.22 % at Re = 400, 0.05 % at Re = 100: the same
// amount in absolute terms), against 0.01 % rounded once.  Carrying the
// weights to double-float precision, the equilibrium polynomial in FP64, or
// f - bw2 f and the final sum in FP64 each left the shortfall as it was.
template <class Real>
__device__ __forceinline__ void centre_fbar(const Real (&f)[Q], Real rho, const Real (&u0)[3], const Real F[3],
                                            Real hdt, Real bw0, Real bw2, Real (&fb)[Q]) {
  const EqTerm<Real> eq(ExactProducts{}, rho, u0[0] + hdt * F[0], u0[1] + hdt * F[1], u0[2] + hdt * F[2], F, bw2,
                        bw0);
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    if constexpr (ip != 0 || sizeof(Real) != 4)
      fb[ip] = eq.template relax_keep<ip>(-bw2, f[ip], cdot<ip>(F[0], F[1], F[2]));
  });
  if constexpr (sizeof(Real) == 4) {
    Real moving = fb[1];
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      if constexpr (ip >= 2) moving += fb[ip];
    });
    fb[0] = rho - moving;
  }
}

// u0 = j / rho
template <class Real>
__device__ __forceinline__ void moment_velocity(Real rho, const Real (&j)[3], Real (&u0)[3]) {
  const Real inv = rcp_real(rho);
  u0[0] = j[0] * inv;
  u0[1] = j[1] * inv;
  u0[2] = j[2] * inv;
}

// rho = sum_i f_i and j = sum_i c_i f_i of one cell, summed in population
// order (each sum starts from its first term).
template <class Real>
__device__ __forceinline__ void cell_moments(const Real (&ff)[Q], Real& rho, Real (&j)[3]) {
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    if constexpr (ip == 0) rho = ff[0]; else rho += ff[ip];
    // first populations with a non-zero x, y, z component: 1, 3, 5 (c = +1)
    if constexpr (ip == 1) j[0] = ff[1]; else if constexpr (CX[ip] != 0) j[0] += CX[ip] > 0 ? ff[ip] : -ff[ip];
    if constexpr (ip == 3) j[1] = ff[3]; else if constexpr (CY[ip] != 0) j[1] += CY[ip] > 0 ? ff[ip] : -ff[ip];
    if constexpr (ip == 5) j[2] = ff[5]; else if constexpr (CZ[ip] != 0) j[2] += CZ[ip] > 0 ? ff[ip] : -ff[ip];
  });
}

// rho and the momentum velocity u0 = j / rho of one cell; with `shifted`
// the populations are h = f - w and rho_out is sum h = rho - 1.  FP32
// populations are summed in FP64 and the results rounded to FP32 once: the
// rounding of FP32 sums (about 1e-8 of j, a quarter of the velocity a
// typical body force adds per step) repeats from step to step in a slowly
// evolving flow and acts as a steady source, which over 10^4 steps of the
// Re_tau = 180 channel moves the rms-velocity profiles by 1e-3 of their
// size (FP64 sums: 2e-4).
template <class Real>
__device__ __forceinline__ void cell_rho_u(const Real (&ff)[Q], Real& rho_out, Real (&u0)[3], bool shifted = false) {
  if constexpr (sizeof(Real) == 4) {
    double r, j[3];
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      const double v = (double)ff[ip];
      if constexpr (ip == 0) r = v; else r += v;
      if constexpr (ip == 1) j[0] = v; else if constexpr (CX[ip] != 0) j[0] += CX[ip] > 0 ? v : -v;
      if constexpr (ip == 3) j[1] = v; else if constexpr (CY[ip] != 0) j[1] += CY[ip] > 0 ? v : -v;
      if constexpr (ip == 5) j[2] = v; else if constexpr (CZ[ip] != 0) j[2] += CZ[ip] > 0 ? v : -v;
    });
    rho_out = (Real)r;
    const double rho = shifted ? 1.0 + r : r;
    // 1/rho: the hardware reciprocal and one Newton step (error ~1e-14)
    double x = __builtin_amdgcn_rcp(rho);
    x = __builtin_fma(x, __builtin_fma(-rho, x, 1.0), x);
    u0[0] = (Real)(j[0] * x);
    u0[1] = (Real)(j[1] * x);
    u0[2] = (Real)(j[2] * x);
  } else {
    Real j[3];
    cell_moments<Real>(ff, rho_out, j);
    moment_velocity<Real>(shifted ? Real(1) + rho_out : rho_out, j, u0);
  }
}

template <class Real>
__device__ __forceinline__ void centre_fbar(const Real (&f)[Q], const Real F[3], Real hdt, Real bw0, Real bw2,
                                            Real (&fb)[Q]) {
  Real rho, u0[3];
  cell_rho_u<Real>(f, rho, u0);
  centre_fbar<Real>(f, rho, u0, F, hdt, bw0, bw2, fb);
}

// ---- FP16 non-equilibrium storage (NeqHalf) ----

// e_i(u) = 1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u.u, given base = 1 - 1.5 u.u
template <int ip, class Real>
__device__ __forceinline__ Real eq_shape(Real base, Real ux, Real uy, Real uz) {
  if constexpr (ip == 0) {
    return base;
  } else {
    const Real cu = cdot<ip>(ux, uy, uz);
    return cu * (cu * Real(4.5) + Real(3)) + base;
  }
}

// f_i = w_i rho e_i(u) + (w_i / kNeqScale) s_i
template <class Real>
__device__ __forceinline__ void neq_decode(Real rho, const Real (&u)[3], const __half (&s)[Q], Real (&ff)[Q]) {
  const Real base = Real(1) - Real(1.5) * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  Real wr[3];
  for (int k = 0; k < 3; k++) wr[k] = Real(wclass_weight(k)) * rho;
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    const Real e = eq_shape<ip>(base, u[0], u[1], u[2]);
    ff[ip] = wr[wclass(ip)] * e + Real(wgt(ip) / (double)kNeqScale) * (Real)__half2float(s[ip]);
  });
}

// rho, u = j/rho of f and s_i = (f_i - w_i rho e_i(u)) kNeqScale / w_i in FP16
template <class Real>
__device__ __forceinline__ void neq_encode(const Real (&ff)[Q], Real& rho, Real (&u)[3], __half (&s)[Q]) {
  cell_rho_u<Real>(ff, rho, u);
  const Real base = Real(1) - Real(1.5) * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  const Real rs = rho * Real(kNeqScale);
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    const Real e = eq_shape<ip>(base, u[0], u[1], u[2]);
    s[ip] = __float2half((float)(ff[ip] * Real((double)kNeqScale / wgt(ip)) - rs * e));
  });
}

// f of cell c (of N) in the arithmetic type, any storage.
template <class Real, class Store>
__device__ __forceinline__ void load_cell(const Store* f, long long c, const KParams& P, Real (&ff)[Q]) {
  if constexpr (std::is_same<Store, NeqHalf>::value) {
    const float* m = reinterpret_cast<const float*>(f);
    const __half* s = reinterpret_cast<const __half*>(m + 4 * P.N);
    __half sv[Q];
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      sv[ip] = s[ip * P.N + c];
    });
    const Real u[3] = {(Real)m[P.N + c], (Real)m[2 * P.N + c], (Real)m[3 * P.N + c]};
    neq_decode<Real>(Real(1) + (Real)m[c], u, sv, ff);
  } else {
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      ff[ip] = Codec<Store>::template load<Real>(f, ip * P.N + c, P, ip);
    });
  }
}

// Stores f of cell c, any storage.
template <class Real, class Store>
__device__ __forceinline__ void store_cell(Store* f, long long c, const KParams& P, const Real (&ff)[Q]) {
  if constexpr (std::is_same<Store, NeqHalf>::value) {
    float* m = reinterpret_cast<float*>(f);
    __half* s = reinterpret_cast<__half*>(m + 4 * P.N);
    Real rho, u[3];
    __half sv[Q];
    neq_encode<Real>(ff, rho, u, sv);
    m[c] = (float)(rho - Real(1));
    m[P.N + c] = (float)u[0];
    m[2 * P.N + c] = (float)u[1];
    m[3 * P.N + c] = (float)u[2];
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      s[ip * P.N + c] = sv[ip];
    });
  } else {
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      Codec<Store>::template store<Real>(f, ip * P.N + c, ff[ip], P, ip);
    });
  }
}

// One cell's stored f as loaded through a StateBuf (the loads can be issued
// ahead and decoded at use): the 19 stored values, plus rho and u for FP16
// non-equilibrium storage.  Byte offsets: vb = cell index * element size,
// population ip at ip * Nb (Nb = N * element size; NeqHalf: the FP16 part
// follows the 16 N bytes of rho, u).
// FP16 values are held zero-extended in 32-bit registers: as 16-bit values
// the compiler packs two per register, which makes it wait for each pair of
// loads as they are issued and defeats the prefetch.
template <class Store>
struct RawSlot {
  using type = Store;
};
template <>
struct RawSlot<__half> {
  using type = unsigned;
};
template <class Real, class Store>
struct RawCell {
  typename RawSlot<Store>::type v[Q];
  __device__ __forceinline__ void load(const StateBuf<Store>& b, unsigned cell, unsigned N) {
    const unsigned vb = cell * (unsigned)sizeof(Store), Nb = N * (unsigned)sizeof(Store);
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      if constexpr (std::is_same<Store, __half>::value)
        v[ip] = raw_b16_32(b.r, vb, ip * Nb);
      else
        v[ip] = b.raw(vb, ip * Nb);
    });
  }
  // Waits here for the loads of load() (an empty asm that reads and
  // "writes" each slot): later reads of the slots then wait for nothing.
  // The vector-memory counter counts loads and stores together, in issue
  // order, so a wait for a load issued before a run of stores that the
  // compiler cannot see on every path (stores under a branch) is compiled as
  // a wait for those stores too; settling the slots before such stores
  // avoids it.
  __device__ __forceinline__ void settle() {
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      asm volatile("" : "+v"(v[ip]));
    });
  }
  // f, rho and the momentum velocity u0 = j/rho
  __device__ __forceinline__ void decode(const KParams& P, Real (&ff)[Q], Real& rho, Real (&u0)[3]) const {
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      if constexpr (std::is_same<Store, __half>::value)
        ff[ip] = (Real)((float)wgt(ip) * (P.rho0f + __half2float(__ushort_as_half((unsigned short)v[ip]))));
      else
        ff[ip] = (Real)v[ip];
    });
    cell_rho_u<Real>(ff, rho, u0);
  }
  // shifted populations h = f - w (see EqShift), drho = sum h and the
  // momentum velocity u0 = j / (1 + drho); f - w is exact for FP32 f within
  // a factor 2 of w
  __device__ __forceinline__ void decode_shift(const KParams& P, Real (&h)[Q], Real& drho, Real (&u0)[3]) const {
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      if constexpr (std::is_same<Store, __half>::value)
        h[ip] = Real(wgt(ip)) * ((Real)__half2float(__ushort_as_half((unsigned short)v[ip])) + (P.rho0f - 1.0f));
      else
        h[ip] = (Real)v[ip] - Real(wgt(ip));
    });
    cell_rho_u<Real>(h, drho, u0, true);
  }
  // the rho and j only (FP32 storage path of the adjoint-face pass)
  __device__ __forceinline__ void moments(const Real (&ff)[Q], Real& rho, Real (&j)[3]) const {
    cell_moments<Real>(ff, rho, j);
  }
};
template <class Real>
struct RawCell<Real, NeqHalf> {
  unsigned v[Q];  // FP16 bits, zero-extended (see RawSlot)
  float m[4];
  __device__ __forceinline__ void load(const StateBuf<NeqHalf>& b, unsigned cell, unsigned N) {
    const unsigned vf = cell * 4u, vh = cell * 2u;
    for (int q = 0; q < 4; q++) m[q] = b.rawf(vf, q * 4u * N);
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      v[ip] = b.raw32(vh, 16u * N + ip * 2u * N);
    });
  }
  // as RawCell::settle
  __device__ __forceinline__ void settle() {
    for (int q = 0; q < 4; q++) asm volatile("" : "+v"(m[q]));
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      asm volatile("" : "+v"(v[ip]));
    });
  }
  __device__ __forceinline__ void decode(const KParams& P, Real (&ff)[Q], Real& rho, Real (&u0)[3]) const {
    rho = Real(1) + (Real)m[0];
    u0[0] = (Real)m[1];
    u0[1] = (Real)m[2];
    u0[2] = (Real)m[3];
    __half s[Q];
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      s[ip] = __ushort_as_half((unsigned short)v[ip]);
    });
    neq_decode<Real>(rho, u0, s, ff);
  }
  // shifted populations h_i = f_i - w_i = w_i rho eps_i + w_i drho + w_i s_i / kNeqScale
  // straight from the stored drho, u and s (see EqShift), with the
  // per-class products w rho and w drho rounded once (NeqEq; as in the
  // centre stage, products rounded twice bias the flow's dissipation.  On
  // the Re = 1600 Taylor-Green vortex, against FP64: this decode and the
  // centre stage's EqShift products both rounded twice, 0.67 % less
  // dissipation; only EqShift's once, 0.54 % more; only the decode's once,
  // 1.2 % less; both once, 0.01 %)
  __device__ __forceinline__ void decode_shift(const KParams& P, Real (&h)[Q], Real& drho, Real (&u0)[3]) const {
    drho = (Real)m[0];
    u0[0] = (Real)m[1];
    u0[1] = (Real)m[2];
    u0[2] = (Real)m[3];
    const NeqEq<Real> eq(drho, u0);
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      const Real sv = (Real)__half2float(__ushort_as_half((unsigned short)v[ip]));
      h[ip] = eq.template shift_of<ip>(sv);
    });
  }
};

// Stores population ip of a cell given in shifted form h = f - w (FP32 f =
// h + w, or the FP16 deviation h / w + 1 - rho0); not FP16 non-equilibrium
// storage, which encodes whole cells (store_neq_shift).
template <class Real, class Store>
__device__ __forceinline__ void st_state_shift(const StateBuf<Store>& b, unsigned vb, unsigned sb, Real h,
                                               const KParams& P, int ip) {
  if constexpr (std::is_same<Store, __half>::value)
    b.put(__float2half((float)h * (float)(1.0 / wgt(ip)) + (1.0f - P.rho0f)), vb, sb);
  else
    b.put((Store)(h + Real(wgt(ip))), vb, sb);
}

// FP16 non-equilibrium encoding of a cell given in shifted form: drho = sum
// h, u0 = j / (1 + drho), s_i = (h_i / w_i - drho - rho eps_i) kNeqScale.
// Returns drho and u0.
template <class Real>
__device__ __forceinline__ void store_neq_shift(const StateBuf<NeqHalf>& b, unsigned cell, unsigned N,
                                                const Real (&h)[Q], Real& drho, Real (&u0)[3]) {
  cell_rho_u<Real>(h, drho, u0, true);
  const NeqEq<Real> eq(drho, u0);
  const unsigned vf = cell * 4u, vh = cell * 2u;
  b.putf((float)drho, vf, 0);
  for (int q = 0; q < 3; q++) b.putf((float)u0[q], vf, (q + 1) * 4u * N);
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    const Real n = (h[ip] - eq.template eq<ip>()) * Real((double)kNeqScale / wgt(ip));
    b.put(__float2half((float)n), vh, 16u * N + ip * 2u * N);
  });
}

// Stores the new f of a cell through a StateBuf; returns its rho and momentum
// velocity u0 (needed for the encoding, and by the flow-rate partials) when
// `want_u` or the storage needs them.
template <class Real, class Store>
__device__ __forceinline__ void store_raw_cell(const StateBuf<Store>& b, unsigned cell, unsigned N,
                                               const KParams& P, const Real (&ff)[Q], bool want_u, Real& rho,
                                               Real (&u0)[3]) {
  if constexpr (std::is_same<Store, NeqHalf>::value) {
    __half s[Q];
    neq_encode<Real>(ff, rho, u0, s);
    const unsigned vf = cell * 4u, vh = cell * 2u;
    b.putf((float)(rho - Real(1)), vf, 0);
    for (int q = 0; q < 3; q++) b.putf((float)u0[q], vf, (q + 1) * 4u * N);
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      b.put(s[ip], vh, 16u * N + ip * 2u * N);
    });
  } else {
    const unsigned vb = cell * (unsigned)sizeof(Store), Nb = N * (unsigned)sizeof(Store);
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      st_state<Real, Store>(b, vb, ip * Nb, ff[ip], P, ip);
    });
    if (want_u) cell_rho_u<Real>(ff, rho, u0);
  }
}

// Adjoint of the equilibrium-with-force term (used by the adjoint of the
// centre stage and of the face collision).
// G = sum_i lam_i (B + C T_i) feq_i(rho, u), T_i = 3 (c_i - u).F,
// feq_i = w_i rho P_i, P_i = 1 + e_i + e_i^2/2 - 3 u.u/2, e_i = 3 c_i.u, over
// the populations in the compile-time set Sel (Sel(ip) true).  With
// s_i = lam_i w_i (B + C T_i):
//   dG/drho = sum s_i P_i,
//   dG/du   = 3 rho (sum s_i (1 + e_i) c_i - u sum s_i) - 3 C F rho sum lam_i w_i P_i,
//   dG/dF   = 3 C rho (sum lam_i w_i P_i c_i - u sum lam_i w_i P_i),
// so each population only adds to a handful of sums.
template <class Real, class Sel>
__device__ __forceinline__ void feq_adjoint(const Real (&lam)[Q], Real B, Real C, Real rho, const Real u[3],
                                            const Real F[3], Real& g_rho, Real g_u[3], Real g_F[3], Sel sel) {
  const Real uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const Real uF = u[0] * F[0] + u[1] * F[1] + u[2] * F[2];
  const Real base = Real(1) - Real(1.5) * uu;
  Real S0 = 0, S1 = 0, S2 = 0, Sc[3] = {0, 0, 0}, Sp[3] = {0, 0, 0};
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    if constexpr (decltype(sel(I))::value) {
      constexpr int cx = CX[ip], cy = CY[ip], cz = CZ[ip];
      const Real cu = Real(cx) * u[0] + Real(cy) * u[1] + Real(cz) * u[2];
      const Real cF = Real(cx) * F[0] + Real(cy) * F[1] + Real(cz) * F[2];
      const Real e = Real(3) * cu;
      const Real Pq = base + e * (Real(1) + Real(0.5) * e);
      const Real lw = lam[ip] * Real(wgt(ip));
      const Real s = lw * (B + C * Real(3) * (cF - uF));
      const Real lp = lw * Pq;
      const Real se = s * (Real(1) + e);
      S0 += s;
      S1 += s * Pq;
      S2 += lp;
      if constexpr (cx != 0) { Sc[0] += cx > 0 ? se : -se; Sp[0] += cx > 0 ? lp : -lp; }
      if constexpr (cy != 0) { Sc[1] += cy > 0 ? se : -se; Sp[1] += cy > 0 ? lp : -lp; }
      if constexpr (cz != 0) { Sc[2] += cz > 0 ? se : -se; Sp[2] += cz > 0 ? lp : -lp; }
    }
  });
  g_rho = S1;
  const Real r3 = Real(3) * rho;
  for (int a = 0; a < 3; a++) {
    g_u[a] = r3 * (Sc[a] - u[a] * S0) - r3 * C * F[a] * S2;
    g_F[a] = r3 * C * (Sp[a] - u[a] * S2);
  }
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

}  // namespace dugks
