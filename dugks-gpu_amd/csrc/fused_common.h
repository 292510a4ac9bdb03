// Building blocks of the fused step kernel and its adjoint-face mode:
// per-axis direction sets, f-bar+ of a (possibly wall-ghost) column, and the
// face reconstruction record with its collision and the transpose of that
// collision.  Included inside namespace dugks { namespace { ... } } of the
// kernel files.
#pragma once

// populations with a non-zero component along each axis
constexpr int XD[10] = {1, 2, 7, 8, 9, 10, 11, 12, 13, 14};
constexpr int YD[10] = {3, 4, 7, 8, 9, 10, 15, 16, 17, 18};
constexpr int ZD[10] = {5, 6, 11, 12, 13, 14, 15, 16, 17, 18};
constexpr int ZP[5] = {5, 11, 12, 15, 16};  // c_z = +1

template <int A>
__device__ constexpr int dirs(int k) {
  return A == 0 ? XD[k] : (A == 1 ? YD[k] : ZD[k]);
}

// (1 - m) a + m b, evaluated as a + m (b - a): a constant stays exactly
// constant whatever the rounding of m.  In FP32, (1 - m) and m rounded apart
// do not sum to 1, which on a stretched axis biases every face value by a
// fixed fraction of its size; the flux divergence turns that into a steady
// source that shifts the mean flow (the 10^4-step channel's FP32 profiles
// drift by up to 3e-3 in the other form).
template <class Real>
__device__ __forceinline__ Real lerp_(Real a, Real b, Real m) {
  return fma_rn(m, b - a, a);
}

template <class Real>
__device__ __forceinline__ void ambient_force(const KParams& P, Real F[3]) {
  F[0] = (Real)P.famb_dev[0];
  F[1] = (Real)P.famb_dev[1];
  F[2] = (Real)P.famb_dev[2];
}

// f-bar+ of interior cell (i, j, k) (0-based); SHIFT: in shifted form
// f-bar+ - w (see EqShift).
template <class Real, class Store, bool SHIFT>
__device__ __forceinline__ void fbar_cell(const KParams& P, const StateBuf<Store>& f, unsigned Nb, int i, int j,
                                          int k, const Real Famb[3], Real (&fb)[Q]) {
  const unsigned c = ((unsigned)k * P.n[1] + j) * P.n[0] + i;
  RawCell<Real, Store> raw;
  raw.load(f, c, (unsigned)P.N);
  Real ff[Q], rho, u0[3];
  if constexpr (SHIFT)
    raw.decode_shift(P, ff, rho, u0);
  else
    raw.decode(P, ff, rho, u0);
  Real F[3] = {Famb[0], Famb[1], Famb[2]};
  if (P.fcell) {
    F[0] += (Real)P.fcell[c];
    F[1] += (Real)P.fcell[P.N + c];
    F[2] += (Real)P.fcell[2 * P.N + c];
  }
  if constexpr (SHIFT)
    centre_fbar_shift<Real>(ff, rho, u0, F, (Real)P.hdt, (Real)P.bw0, (Real)P.bw2, fb);
  else
    centre_fbar<Real>(ff, rho, u0, F, (Real)P.hdt, (Real)P.bw0, (Real)P.bw2, fb);
}

// f-bar+ of a column that may be a wall ghost in x and/or y: the ghost value
// is the linear extrapolation q f(1) + (1-q) f(2) from the wall side.
// gi, gj are 0-based; a wall ghost is -1 or n; periodic indices are wrapped
// by the caller.  SHIFT: in shifted form (the extrapolation weights sum to 1).
template <class Real, class Store, bool SHIFT>
__device__ void fbar_column(const KParams& P, const StateBuf<Store>& f, unsigned Nb, int gi, int gj, int kz,
                            const Real Famb[3], Real (&fb)[Q]) {
  const bool gx = gi < 0 || gi >= P.n[0];
  const bool gy = gj < 0 || gj >= P.n[1];
  if (!gx && !gy) {
    fbar_cell<Real, Store, SHIFT>(P, f, Nb, gi, gj, kz, Famb, fb);
    return;
  }
  // the two columns (and weights) of the extrapolation along each ghost axis;
  // a non-ghost axis uses the column itself with weight 1
  int x0 = gi, x1 = gi, y0 = gj, y1 = gj;
  Real qx = Real(1), qy = Real(1);
  if (gx) {
    const bool lo = gi < 0;
    qx = (Real)(lo ? P.qlo[0] : P.qhi[0]);
    x0 = lo ? 0 : P.n[0] - 1;
    x1 = lo ? 1 : P.n[0] - 2;
  }
  if (gy) {
    const bool lo = gj < 0;
    qy = (Real)(lo ? P.qlo[1] : P.qhi[1]);
    y0 = lo ? 0 : P.n[1] - 1;
    y1 = lo ? 1 : P.n[1] - 2;
  }
  // The extrapolation q f(1) + (1 - q) f(2) along a ghost axis is taken as
  // f(2) + q (f(1) - f(2)) (exact for constants, see lerp_); a corner ghost
  // of two wall axes takes the tensor product of the two (the x -> y ghost
  // fill of the reference gives the same linear combination).  f-bar+ has
  // one code copy per case in a rolled loop, and no indexed local arrays
  // (those would live in scratch memory): these lanes are few.
  if (!(gx && gy)) {
    const Real q = gx ? qx : qy;
#pragma unroll 1
    for (int k = 0; k < 2; k++) {
      // k = 0: the second column (x1 or y1), k = 1: the first
      const int ci = gx ? (k ? x0 : x1) : gi, cj = gy ? (k ? y0 : y1) : gj;
      Real t[Q];
      fbar_cell<Real, Store, SHIFT>(P, f, Nb, ci, cj, kz, Famb, t);
      static_for<Q>([&](auto I) {
        constexpr int ip = decltype(I)::value;
        fb[ip] = k ? fma_rn(q, t[ip] - fb[ip], fb[ip]) : t[ip];
      });
    }
    return;
  }
  bool first = true;
#pragma unroll 1
  for (int k = 0; k < 4; k++) {
    const int a = k & 1, b = k >> 1;
    const Real w = (a ? Real(1) - qx : qx) * (b ? Real(1) - qy : qy);
    Real t[Q];
    fbar_cell<Real, Store, SHIFT>(P, f, Nb, a ? x1 : x0, b ? y1 : y0, kz, Famb, t);
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      fb[ip] = first ? w * t[ip] : fb[ip] + w * t[ip];
    });
    first = false;
  }
}

// Index of population ip in the crossing set of axis A (-1 if it does not
// cross a face normal to A).
template <int A>
__device__ constexpr int cross_index(int ip) {
  for (int k = 0; k < 10; k++)
    if (dirs<A>(k) == ip) return k;
  return -1;
}

// Transpose of the wall bounce-back: a population that entered through a
// wall took its opposite's value, so its cotangent moves to that opposite.
template <class Real, int A>
__device__ __forceinline__ void bounce_transpose(const Real (&post)[Q], bool low_wall, bool high_wall,
                                                 Real (&lam)[Q]) {
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    Real v = post[ip];
    if constexpr (cv(A, ip) != 0) {
      const bool in = (low_wall && cv(A, ip) > 0) || (high_wall && cv(A, ip) < 0);
      const bool opp_in = (low_wall && cv(A, ip) < 0) || (high_wall && cv(A, ip) > 0);
      if (in) v = Real(0);
      if (opp_in) v += post[OPP[ip]];
    }
    lam[ip] = v;
  });
}

// Cotangents of the 19 reconstructed values of a face normal to axis A from
// the transposed collision's coefficients: fw1 mu_p + Ac + B.c_p, then the
// transposed bounce-back.
template <class Real, int A>
__device__ __forceinline__ void expand_cotangents(const Real (&mu)[10], Real Ac, const Real (&B)[3], Real fw1,
                                                  bool low_wall, bool high_wall, Real (&lam)[Q]) {
  Real post[Q];
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    constexpr int k = cross_index<A>(ip);
    Real v = plus_cdot<ip>(Ac, B[0], B[1], B[2]);
    if constexpr (k >= 0) v += fw1 * mu[k];
    post[ip] = v;
  });
  bounce_transpose<Real, A>(post, low_wall, high_wall, lam);
}

// Compact face record (kCompact values): the cotangents fw1 mu_k + Ac + B.c_p
// of the ten crossing populations, then Ac and the two components of B
// tangential to the face.  The nine other populations have c_A = 0, so B_A
// only enters the crossing ones.
template <int A>
__device__ constexpr int tangential(int t) {
  return t == 0 ? (A == 0 ? 1 : 0) : (A == 2 ? 1 : 2);
}
template <class Real, int A>
__device__ __forceinline__ void compact_record(const Real (&mu)[10], Real Ac, const Real (&B)[3], Real fw1,
                                               Real (&rec)[kCompact]) {
  static_for<10>([&](auto K) {
    constexpr int k = decltype(K)::value;
    constexpr int ip = dirs<A>(k);
    rec[k] = plus_cdot<ip>(Ac, B[0], B[1], B[2]) + fw1 * mu[k];
  });
  rec[10] = Ac;
  rec[11] = B[tangential<A>(0)];
  rec[12] = B[tangential<A>(1)];
}

// The 19 reconstructed-value cotangents of a compact record (as
// expand_cotangents of the coefficients it was made from).
template <class Real, int A>
__device__ __forceinline__ void expand_record(const Real (&rec)[kCompact], bool low_wall, bool high_wall,
                                              Real (&lam)[Q]) {
  Real B[3];
  B[A] = Real(0);
  B[tangential<A>(0)] = rec[11];
  B[tangential<A>(1)] = rec[12];
  const Real Ac = rec[10];
  Real post[Q];
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    constexpr int k = cross_index<A>(ip);
    if constexpr (k >= 0)
      post[ip] = rec[k];
    else
      post[ip] = plus_cdot<ip>(Ac, B[0], B[1], B[2]);
  });
  bounce_transpose<Real, A>(post, low_wall, high_wall, lam);
}

// First population that does not cross a face normal to axis A and has a
// non-zero component b (b != A); the moments of a face start from it.
template <int A>
__device__ constexpr int first_tangential(int b) {
  for (int ip = 0; ip < Q; ip++)
    if (cv(A, ip) == 0 && cv(b, ip) != 0) return ip;
  return -1;
}

// Reconstructed populations of one face normal to axis A, held as the ten
// values that cross the face plus the partial moments of the nine that do
// not (those only enter rho and j of the face collision).  The populations
// are put in order 0..18 (the rest population first); every moment starts
// from its first term, so no sum adds to a zero.
template <class Real, int A>
struct Face {
  Real cross[10];
  Real rho, j[3];
  template <int ip>
  __device__ __forceinline__ void put(Real v) {
    constexpr int k = cross_index<A>(ip);
    if constexpr (k >= 0) {
      cross[k] = v;
    } else {
      if constexpr (ip == 0) rho = v; else rho += v;
      static_for<3>([&](auto B) {
        constexpr int b = decltype(B)::value;
        if constexpr (b != A && cv(b, ip) != 0) {
          const Real s = cv(b, ip) > 0 ? v : -v;
          if constexpr (ip == first_tangential<A>(b)) j[b] = s; else j[b] += s;
        }
      });
    }
  }
  // rho and j of the face: the partial moments plus the crossing populations
  // (j_A starts from the first crossing population, c_A = +1)
  __device__ __forceinline__ void moments() {
    static_for<10>([&](auto K) {
      constexpr int k = decltype(K)::value;
      constexpr int ip = dirs<A>(k);
      const Real v = cross[k];
      rho += v;
      static_for<3>([&](auto B) {
        constexpr int b = decltype(B)::value;
        if constexpr (cv(b, ip) != 0) {
          const Real s = cv(b, ip) > 0 ? v : -v;
          if constexpr (b == A && k == 0) j[b] = s; else j[b] += s;
        }
      });
    });
  }
  // Half-way bounce-back: populations entering the domain through a wall
  // take the value of their opposite (src/dugks.cu:807-823).
  __device__ __forceinline__ void bounce(bool low_wall, bool high_wall) {
    if (low_wall) {
      static_for<10>([&](auto K) {
        constexpr int k = decltype(K)::value;
        constexpr int ip = dirs<A>(k);
        if constexpr (cv(A, ip) > 0) cross[k] = cross[cross_index<A>(OPP[ip])];
      });
    }
    if (high_wall) {
      static_for<10>([&](auto K) {
        constexpr int k = decltype(K)::value;
        constexpr int ip = dirs<A>(k);
        if constexpr (cv(A, ip) < 0) cross[k] = cross[cross_index<A>(OPP[ip])];
      });
    }
  }
  // Face collision (f2ruvwh + feq_s, src/dugks.cu:675-709, 1201-1207):
  // rho = sum fbar, u = j/rho + (h/2) F, out = fw1 fbar + (fw2 + fw0 (c-u).F/RT) feq
  // for the crossing populations.
  __device__ __forceinline__ void collide(const Real F[3], Real hdt, Real fw0, Real fw1, Real fw2,
                                          Real (&out)[10]) {
    moments();
    const Real inv = rcp_real(rho);
    const Real hh = Real(0.5) * hdt;
    const EqTerm<Real> eq(rho, j[0] * inv + hh * F[0], j[1] * inv + hh * F[1], j[2] * inv + hh * F[2], F, fw2,
                          fw0);
    static_for<10>([&](auto K) {
      constexpr int k = decltype(K)::value;
      constexpr int ip = dirs<A>(k);
      out[k] = eq.template relax<ip>(fw1, cross[k], cdot<ip>(F[0], F[1], F[2]));
    });
  }
  // collide in shifted populations (the face holds f-bar - w, its rho the
  // sum of those, drho): out - w for the crossing populations (see EqShift)
  __device__ __forceinline__ void collide_shift(const Real F[3], Real hdt, Real fw0, Real fw2, Real (&out)[10]) {
    moments();
    const Real inv = rcp_real(Real(1) + rho);
    const Real hh = Real(0.5) * hdt;
    const EqShift<Real> eq(rho, j[0] * inv + hh * F[0], j[1] * inv + hh * F[1], j[2] * inv + hh * F[2], F, fw2,
                           fw0);
    static_for<10>([&](auto K) {
      constexpr int k = decltype(K)::value;
      constexpr int ip = dirs<A>(k);
      out[k] = eq.template relax<ip>(-fw2, cross[k], cdot<ip>(F[0], F[1], F[2]));
    });
  }
  // The transposed collision up to its equilibrium coefficients: the
  // cotangent of reconstructed value p is fw1 mu_p + Ac + B.c_p (mu_p = 0
  // for the populations that do not cross), before the bounce transpose.
  // FORCE: also adds the cotangent of the face's force F to gF[0..2]
  // (through the equilibrium's explicit F and through u = j/rho + (h/2) F).
  template <bool FORCE = false>
  __device__ __forceinline__ void adjoint_coeffs(const Real F[3], Real hdt, Real fw0, Real fw2,
                                                 const Real (&mu)[10], Real& Ac, Real (&B)[3],
                                                 double* gF = nullptr) {
    moments();
    const Real inv = rcp_real(rho);
    const Real hh = Real(0.5) * hdt;
    const Real ut[3] = {j[0] * inv, j[1] * inv, j[2] * inv};
    const Real u[3] = {ut[0] + hh * F[0], ut[1] + hh * F[1], ut[2] + hh * F[2]};
    Real lq[Q];
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      constexpr int k = cross_index<A>(ip);
      if constexpr (k >= 0) lq[ip] = mu[k]; else lq[ip] = Real(0);
    });
    Real g_rho, g_u[3], g_F[3];
    feq_adjoint<Real>(lq, fw2, fw0, rho, u, F, g_rho, g_u, g_F, Crossing<A>{});
    Ac = g_rho - (g_u[0] * ut[0] + g_u[1] * ut[1] + g_u[2] * ut[2]) * inv;
    B[0] = g_u[0] * inv;
    B[1] = g_u[1] * inv;
    B[2] = g_u[2] * inv;
    if constexpr (FORCE) {
      gF[0] += (double)(g_F[0] + hh * g_u[0]);
      gF[1] += (double)(g_F[1] + hh * g_u[1]);
      gF[2] += (double)(g_F[2] + hh * g_u[2]);
    }
  }
  // Transpose of collide (after bounce) and bounce: given the cotangent mu
  // of the ten outputs, the cotangent of the 19 reconstructed values before
  // bounce-back.  d out_q / d fbar_p = fw1 delta_qp + (d/d rho + c_p.d/d j)
  // of the equilibrium term, so lam_p = fw1 mu_p + A + B.c_p; a population
  // that entered through a wall took its opposite's value, so its cotangent
  // moves to that opposite.
  template <bool FORCE = false>
  __device__ __forceinline__ void adjoint(const Real F[3], Real hdt, Real fw0, Real fw1, Real fw2,
                                          const Real (&mu)[10], bool low_wall, bool high_wall, Real (&lam)[Q],
                                          double* gF = nullptr) {
    Real Ac, B[3];
    adjoint_coeffs<FORCE>(F, hdt, fw0, fw2, mu, Ac, B, gF);
    expand_cotangents<Real, A>(mu, Ac, B, fw1, low_wall, high_wall, lam);
  }
};

