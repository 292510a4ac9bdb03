// One-pass adjoint of the fused DUGKS step.
//
// Given lam' = dJ/df~(n+1), one launch returns lam = (df~(n+1)/df~(n))^T lam'
// and adds dJ/dF_cell to the force gradient.  It covers the configurations of
// the fused step with 16 x 12 tiles, FP32 arithmetic (FP32 or FP16 state) and
// at most one wall axis (x or y; z is periodic and uniform).  The split path
// of adjoint.hip (adjoint-face pass, face ghost copies, gather) covers every
// other case and cross-checks this kernel.
//
// A workgroup owns a 16 x 12 column tile and marches along z as the step
// kernel does.  It recomputes f-bar+ on the tile plus a two-deep ring of
// columns (20 x 16 = 320 lanes, five waves), so that every face whose
// reconstruction stencil reaches a cell of the tile, or a wall-ghost column
// beside it, is reconstructed and transposed inside the workgroup.  Per plane:
//   phase 1: f-bar+(p+1) (calc_fb1, src/dugks.cu:1131-1196); Z(p), the
//            tangential z-lerp, and Zn(p+1/2), the normal z-lerp, into LDS as
//            in the step kernel; lam'(p) of the tile and its one-deep ring
//            into LDS;
//   phase 2: the x- and y-faces of plane p and the z-face p+1/2 of every
//            column of the tile and its one-deep ring: reconstruction
//            (src/dugks.cu:756-906, 964-978), the cotangent mu of the
//            crossing outputs from lam' of the face's two cells (update,
//            src/dugks.cu:945-956), then the transposed collision and
//            bounce-back (Face::adjoint; src/dugks.cu:675-709, 807-823); the
//            19 cotangents of each face go to LDS (0 for faces that do not
//            exist: wall ghost rows and columns, beyond the ghosts);
//   phase 3: each tile cell and wall-ghost column gathers the transposed
//            in-plane lerps from the LDS face planes, which gives the
//            cotangents of its Z(p) and Zn(p+1/2), and pushes the transposed
//            z-lerps into the f-bar+ cotangents of planes p-1, p and p+1 held
//            in registers.  The cotangent of f-bar+(p-1) is then complete:
//            wall-ghost columns fold theirs into the two interior columns
//            their extrapolation reads (the transposed ghost fill,
//            src/dugks.cu:1234-1251, src/fortran/channel/src/dugks.cuf:365-374),
//            and the centre stage is transposed into lam(p-1) and dJ/dF.
// A z-chunk [z0, zend) runs the faces of planes z0-1 .. zend, so that the
// cotangents of its own planes are complete; neighbouring chunks recompute
// those two planes of faces and write disjoint planes of lam.
#include "device_common.h"

namespace dugks {

bool fused_supported(const Ctx& c, int* tx, int* ty);

namespace {

#include "fused_common.h"

constexpr int kTX = 16, kTY = 12;
constexpr int kNP = (kTX + 4) * (kTY + 4);  // lanes: the tile and its two-deep ring
constexpr int kS2 = kTY + 5;                // column stride of the two-deep planes (odd: fewer bank conflicts)
constexpr int kP2 = (kTX + 4) * kS2;        // words per two-deep plane
constexpr int kR1 = kTY + 2;                // column stride of the one-deep plane
constexpr int kP1 = (kTX + 2) * kR1;        // words per one-deep plane
constexpr int kXcdCount = 8;                // accelerator complex dies of an MI355X

// LDS plane index of tile position (r, c): the two-deep region is
// r in [-1, kTY+2], c in [-2, kTX+1]; the one-deep region r in [0, kTY+1],
// c in [-1, kTX].  Column-major.
__device__ __forceinline__ int i2(int r, int c) { return (c + 2) * kS2 + (r + 1); }
__device__ __forceinline__ int i1(int r, int c) { return (c + 1) * kR1 + r; }

template <class Store>
__global__ void __launch_bounds__(kNP) k_adj_onepass(KParams P, const Store* __restrict__ fin,
                                                     const double* __restrict__ lam, double* __restrict__ lam_prev,
                                                     double* __restrict__ grad, int TZ) {
  using Real = float;
  __shared__ Real A[Q][kP2];    // Z(p): tangential z-lerp of f-bar+(p)
  __shared__ Real An[Q][kP2];   // Zn(p+1/2): normal z-lerp between p and p+1
  __shared__ Real FX[Q][kP2];   // cotangents of the x-face left of (r, c), plane p
  __shared__ Real FY[Q][kP2];   // cotangents of the y-face below (r, c), plane p
  __shared__ Real FZ[Q][kP2];   // cotangents of the z-face p+1/2 of column (r, c)
  __shared__ Real Lm[Q][kP1];   // lam'(p) of the tile and its one-deep ring (0 off the domain)
  __shared__ Real Gh[Q][32];    // f-bar+ cotangents of the wall-ghost columns, for the fold
  __shared__ Real WX[6][kTX + 4];
  __shared__ Real WY[6][kTY + 4];

  const int nx = P.n[0], ny = P.n[1], nz = P.n[2];
  const int t = threadIdx.x;
  // XCD-aware tile order, as in the step kernel
  const int gxn = nx / kTX, gyn = ny / kTY;
  const int ntiles = gxn * gyn * (int)gridDim.z;
  const int b = blockIdx.x + gxn * (blockIdx.y + gyn * blockIdx.z);
  const int tt = (ntiles % kXcdCount == 0) ? (b % kXcdCount) * (ntiles / kXcdCount) + b / kXcdCount : b;
  const int bx = tt % gxn, by = (tt / gxn) % gyn, bz = tt / (gxn * gyn);
  const int i0 = bx * kTX, j0 = by * kTY;
  const int z0 = bz * TZ;
  const int zend = min(z0 + TZ, nz);

  // Position of the lane: the tile's cells first (three whole waves), then
  // the one-deep ring (rows, then columns), then the two-deep ring.
  // y = j0 - 1 + r, x = i0 + c.
  int r, c;
  if (t < kTX * kTY) {
    r = t / kTX + 1;
    c = t % kTX;
  } else {
    const int u = t - kTX * kTY;
    constexpr int n1 = 2 * kTX, n2 = n1 + 2 * (kTY + 2), n3 = n2 + 2 * (kTX + 4);
    if (u < n1) {
      r = u < kTX ? 0 : kTY + 1;
      c = u % kTX;
    } else if (u < n2) {
      const int v = u - n1;
      c = v < kTY + 2 ? -1 : kTX;
      r = v % (kTY + 2);
    } else if (u < n3) {
      const int v = u - n2;
      r = v < kTX + 4 ? -1 : kTY + 2;
      c = v % (kTX + 4) - 2;
    } else {
      const int v = u - n3;
      c = v < kTY + 2 ? -2 : kTX + 1;
      r = v % (kTY + 2);
    }
  }
  const bool perx = P.periodic[0] != 0, pery = P.periodic[1] != 0;
  int gi = i0 + c, gj = j0 - 1 + r;
  if (perx) gi = (gi + nx) % nx;
  if (pery) gj = (gj + ny) % ny;
  const bool xin = gi >= 0 && gi < nx, yin = gj >= 0 && gj < ny;
  const bool xg = !xin && (gi == -1 || gi == nx), yg = !yin && (gj == -1 || gj == ny);
  const bool has_fb = (xin || xg) && (yin || yg);  // f-bar+ exists (cell or wall ghost)
  const bool own = xin && yin;                      // a cell of the domain
  const bool ring1 = r >= 0 && r <= kTY + 1 && c >= -1 && c <= kTX;
  const bool interior = t < kTX * kTY;
  // wall-ghost columns beside the tile's cells: their f-bar+ cotangent is
  // folded into the two cells their extrapolation reads
  const bool gx_lane = xg && yin && r >= 1 && r <= kTY;
  const bool gy_lane = yg && xin && c >= 0 && c < kTX;
  const bool gather = interior || gx_lane || gy_lane;
  const bool wall_tile = (!perx && (i0 == 0 || i0 + kTX == nx)) || (!pery && (j0 == 0 || j0 + kTY == ny));
  // faces this lane reconstructs: the x-face left of (r, c), the y-face
  // below it, the z-face above it; a face exists when its cells' row
  // (x-face), column (y-face) or both (z-face) lie in the domain
  const bool do_x = r >= 0 && r <= kTY + 1 && c >= 0 && c <= kTX && yin;
  const bool do_y = r >= 1 && r <= kTY + 1 && c >= -1 && c <= kTX && xin;
  const bool do_z = ring1 && own;

  // weight tables in LDS (layout documented on KParams); row k of WX/WY:
  // 0-2 the -face normal weight toward the column/row for c = -1, 0, +1,
  // 3-4 the tangential shift weights for c < 0 and c > 0, 5 dt/d.
  // Positions beyond the ghosts clamp to the table ends; they only weight
  // faces that do not exist.
  for (int u = t; u < 6 * (kTX + 4) + 6 * (kTY + 4); u += kNP) {
    const bool isx = u < 6 * (kTX + 4);
    const int v = isx ? u : u - 6 * (kTX + 4);
    const int n = isx ? kTX + 4 : kTY + 4;
    const int k = v / n, q = v % n;
    const int a = isx ? 0 : 1;
    const int na = P.n[a];
    int g = (isx ? i0 : j0) - 2 + q;
    if (P.periodic[a]) g = (g + na) % na;
    const int m = min(max(g + 1, 0), na + 1);
    const int idx = k < 3 ? max(m - 1, 0) * 9 + k : m * 9 + (k + 3);
    (isx ? WX[k][q] : WY[k][q]) = tabv<Real>(P, a, idx);
  }
  auto xn = [&](int k, int cc) { return WX[k][cc + 2]; };
  auto xt = [&](int s, int cc) { return WX[3 + s][cc + 2]; };
  auto upx = [&](int cc) { return WX[5][cc + 2]; };
  auto yn = [&](int k, int rr) { return WY[k][rr + 1]; };
  auto yt = [&](int s, int rr) { return WY[3 + s][rr + 1]; };
  auto upy = [&](int rr) { return WY[5][rr + 1]; };
  // uniform z weights (z is periodic and uniform)
  const Real wzt_m = tabv<Real>(P, 2, 9 + 6);  // tangential, toward p+1 (c_z = -1)
  const Real wzt_p = tabv<Real>(P, 2, 9 + 7);  // tangential, toward p-1 (c_z = +1)
  const Real wzn[3] = {tabv<Real>(P, 2, 9 + 0), tabv<Real>(P, 2, 9 + 1), tabv<Real>(P, 2, 9 + 2)};
  const Real updz = tabv<Real>(P, 2, 9 + 8);

  Real Famb[3];
  ambient_force<Real>(P, Famb);
  const Real hdt = (Real)P.hdt, fw0 = (Real)P.fw0, fw1 = (Real)P.fw1, fw2 = (Real)P.fw2;
  const Real bw0 = (Real)P.bw0, bw1 = (Real)P.bw1, bw2 = (Real)P.bw2;

  const unsigned plane = (unsigned)nx * (unsigned)ny;
  const unsigned colcell = own ? (unsigned)gj * (unsigned)nx + (unsigned)gi : 0u;
  const unsigned Nb = (unsigned)(P.N * (long long)sizeof(Store));  // bytes per population
  const StateBuf<Store> bin(fin, Q * Nb);
  const RawBuf lamb(lam, (unsigned)(Q * P.N * 8));  // < 2 GiB, checked at launch
  const unsigned lam_pop = (unsigned)(P.N * 8);

  // lam' of the lane's column at plane k (0 off the domain and on the
  // two-deep ring); the rest population never crosses a face
  const bool lam_lane = ring1 && own;
  auto load_lam = [&](int k, Real (&l)[Q]) {
    const unsigned vb = lam_lane ? ((unsigned)k * plane + colcell) * 8u : kOutOfRange;
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      if constexpr (ip != 0) l[ip] = (Real)lamb.ld_f64(vb, ip * lam_pop);
    });
  };
  // f-bar+ of the lane's column at plane k straight from f (0 beyond the
  // ghosts)
  auto fbar_direct = [&](int k, Real (&fb)[Q]) {
    if (has_fb) {
      fbar_column<Real, Store>(P, bin, Nb, gi, gj, k, Famb, fb);
    } else {
      static_for<Q>([&](auto I) { fb[decltype(I)::value] = Real(0); });
    }
  };
  // f of a cell column two planes ahead (FP16 storage loads at use)
  constexpr bool kPrefetch = !std::is_same<Store, __half>::value;
  const bool pre = kPrefetch && own;
  Real fpre[Q];
  auto prefetch = [&](int k) {
    const unsigned vb = ((unsigned)k * plane + colcell) * (unsigned)sizeof(Store);
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      fpre[ip] = ld_state<Real, Store>(bin, vb, ip * Nb, P, ip);
    });
  };

  // faces that this lane never reconstructs read as 0
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    FX[ip][i2(r, c)] = Real(0);
    FY[ip][i2(r, c)] = Real(0);
    FZ[ip][i2(r, c)] = Real(0);
  });

  // ---- pre-loop: f-bar+ at z0-2 and z0-1, lam'(z0-1) ----
  const int km2 = ((z0 - 2) % nz + nz) % nz, km1 = ((z0 - 1) % nz + nz) % nz;
  Real fbm[5], fbc[Q], fbn[Q];
  fbar_direct(km2, fbn);
  static_for<5>([&](auto K) {
    constexpr int k = decltype(K)::value;
    fbm[k] = fbn[ZP[k]];
  });
  fbar_direct(km1, fbc);
  if (pre) prefetch(z0 % nz);
  // f-bar+ cotangents of planes p-1 and p (push form)
  Real accP[Q], accC[Q];
  static_for<Q>([&](auto I) {
    constexpr int ip = decltype(I)::value;
    accP[ip] = Real(0);
    accC[ip] = Real(0);
  });

  for (int p = z0 - 1; p <= zend; p++) {
    // Re-derive the lane position each plane: the compiler would otherwise
    // keep every weight it reads from LDS in registers across the loop.
    asm volatile("" : "+v"(r), "+v"(c));
    const int kp = (p + nz) % nz;
    const int kn = (kp + 1) % nz;
    // lam'(p) for the LDS plane (loaded here, stored after the f-bar+ work)
    Real lamc[Q];
    lamc[0] = Real(0);
    load_lam(kp, lamc);
    // f-bar+(p+1)
    if (pre) {
      Real F[3] = {Famb[0], Famb[1], Famb[2]};
      if (P.fcell) {
        const unsigned cc = (unsigned)kn * plane + colcell;
        F[0] += (Real)P.fcell[cc];
        F[1] += (Real)P.fcell[P.N + cc];
        F[2] += (Real)P.fcell[2 * P.N + cc];
      }
      centre_fbar<Real>(fpre, F, hdt, bw0, bw1, bw2, fbn);
      if (p < zend) prefetch((kn + 1) % nz);
    } else {
      fbar_direct(kn, fbn);
    }
    // ---- phase 1: Z(p), Zn(p+1/2), lam'(p) ----
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      constexpr int cz = CZ[ip];
      if constexpr (cz == 0) {
        A[ip][i2(r, c)] = fbc[ip];
      } else if constexpr (cz > 0) {
        constexpr int zk = ip == 5 ? 0 : (ip == 11 ? 1 : (ip == 12 ? 2 : (ip == 15 ? 3 : 4)));
        A[ip][i2(r, c)] = lerp_(fbc[ip], fbm[zk], wzt_p);
      } else {
        A[ip][i2(r, c)] = lerp_(fbc[ip], fbn[ip], wzt_m);
      }
      if constexpr (ip == 0)
        An[0][i2(r, c)] = Real(0.5) * (fbc[0] + fbn[0]);
      else
        An[ip][i2(r, c)] = lerp_(fbc[ip], fbn[ip], wzn[cz + 1]);
      if constexpr (ip != 0)
        if (ring1) Lm[ip][i1(r, c)] = lamc[ip];
    });
    static_for<5>([&](auto K) {
      constexpr int k = decltype(K)::value;
      fbm[k] = fbc[ZP[k]];
    });
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      fbc[ip] = fbn[ip];
    });
    __syncthreads();

    // ---- phase 2: face cotangents ----
    if (do_x) {
      // x-face between columns c-1 and c
      Face<Real, 0> fc;
      static_for<Q>([&](auto I) {
        constexpr int ip = decltype(I)::value;
        constexpr int cx = CX[ip], cy = CY[ip];
        if constexpr (ip == 0) {
          fc.template put<0>(Real(0.5) * (A[0][i2(r, c - 1)] + A[0][i2(r, c)]));
        } else {
          const Real w = xn(cx + 1, c);  // fraction toward the right cell
          Real v = lerp_(A[ip][i2(r, c - 1)], A[ip][i2(r, c)], w);
          if constexpr (cy != 0) {
            const Real wy = yt(cy < 0 ? 0 : 1, r);
            const Real v2 = lerp_(A[ip][i2(r - cy, c - 1)], A[ip][i2(r - cy, c)], w);
            v = lerp_(v, v2, wy);
          }
          fc.template put<ip>(v);
        }
      });
      const bool lw = !perx && gi == 0, hw = !perx && gi == nx;
      if (!perx) fc.bounce(lw, hw);
      const Real uL = upx(c - 1), uR = upx(c);
      Real mu[10], lf[Q];
      static_for<10>([&](auto K) {
        constexpr int k = decltype(K)::value;
        constexpr int ip = XD[k];
        mu[k] = Real(CX[ip]) * (uR * Lm[ip][i1(r, c)] - uL * Lm[ip][i1(r, c - 1)]);
      });
      fc.adjoint(Famb, hdt, fw0, fw1, fw2, mu, lw, hw, lf);
      static_for<Q>([&](auto I) {
        constexpr int ip = decltype(I)::value;
        FX[ip][i2(r, c)] = lf[ip];
      });
    }
    if (do_y) {
      // y-face between rows r-1 and r
      Face<Real, 1> fc;
      static_for<Q>([&](auto I) {
        constexpr int ip = decltype(I)::value;
        constexpr int cx = CX[ip], cy = CY[ip];
        if constexpr (ip == 0) {
          fc.template put<0>(Real(0.5) * (A[0][i2(r - 1, c)] + A[0][i2(r, c)]));
        } else {
          const Real w = yn(cy + 1, r);  // fraction toward the upper cell
          Real lo = A[ip][i2(r - 1, c)], hi = A[ip][i2(r, c)];
          if constexpr (cx != 0) {
            const Real wx = xt(cx < 0 ? 0 : 1, c);
            lo = lerp_(lo, A[ip][i2(r - 1, c - cx)], wx);
            hi = lerp_(hi, A[ip][i2(r, c - cx)], wx);
          }
          fc.template put<ip>(lerp_(lo, hi, w));
        }
      });
      const bool lw = !pery && gj == 0, hw = !pery && gj == ny;
      if (!pery) fc.bounce(lw, hw);
      const Real uL = upy(r - 1), uR = upy(r);
      Real mu[10], lf[Q];
      static_for<10>([&](auto K) {
        constexpr int k = decltype(K)::value;
        constexpr int ip = YD[k];
        mu[k] = Real(CY[ip]) * (uR * Lm[ip][i1(r, c)] - uL * Lm[ip][i1(r - 1, c)]);
      });
      fc.adjoint(Famb, hdt, fw0, fw1, fw2, mu, lw, hw, lf);
      static_for<Q>([&](auto I) {
        constexpr int ip = decltype(I)::value;
        FY[ip][i2(r, c)] = lf[ip];
      });
    }
    if (do_z) {
      // z-face between planes p and p+1: bilinear over the upwind tangential
      // neighbours of Zn
      Face<Real, 2> fc;
      static_for<Q>([&](auto I) {
        constexpr int ip = decltype(I)::value;
        constexpr int cx = CX[ip], cy = CY[ip];
        Real v = An[ip][i2(r, c)];
        if constexpr (cx != 0) {
          const Real w = xt(cx < 0 ? 0 : 1, c);
          v = lerp_(v, An[ip][i2(r, c - cx)], w);
          if constexpr (cy != 0) {
            const Real wy = yt(cy < 0 ? 0 : 1, r);
            const Real v2 = lerp_(An[ip][i2(r - cy, c)], An[ip][i2(r - cy, c - cx)], w);
            v = lerp_(v, v2, wy);
          }
        } else if constexpr (cy != 0) {
          const Real wy = yt(cy < 0 ? 0 : 1, r);
          v = lerp_(v, An[ip][i2(r - cy, c)], wy);
        }
        fc.template put<ip>(v);
      });
      Real lamn[Q], mu[10], lf[Q];
      lamn[0] = Real(0);
      load_lam(kn, lamn);  // lam'(p+1) of the own column
      static_for<10>([&](auto K) {
        constexpr int k = decltype(K)::value;
        constexpr int ip = ZD[k];
        mu[k] = Real(CZ[ip]) * updz * (lamn[ip] - Lm[ip][i1(r, c)]);
      });
      fc.adjoint(Famb, hdt, fw0, fw1, fw2, mu, false, false, lf);
      static_for<Q>([&](auto I) {
        constexpr int ip = decltype(I)::value;
        FZ[ip][i2(r, c)] = lf[ip];
      });
    }
    __syncthreads();

    // ---- phase 3: transposed reconstruction into the f-bar+ cotangents ----
    // Cotangent of Z(p) at (r, c) for population ip: the transposed in-plane
    // lerps of the x-faces c (this column is their right cell) and c+1 (left
    // cell), rows r and r+c_y, and of the y-faces r (this row is their upper
    // cell) and r+1 (lower cell), columns c and c+c_x.
    auto gather_z = [&](auto I) -> Real {
      constexpr int ip = decltype(I)::value;
      constexpr int cx = CX[ip], cy = CY[ip];
      if constexpr (ip == 0) {
        return Real(0.5) * (FX[0][i2(r, c)] + FX[0][i2(r, c + 1)]) +
               Real(0.5) * (FY[0][i2(r, c)] + FY[0][i2(r + 1, c)]);
      } else {
        const Real a0 = xn(cx + 1, c), a1 = Real(1) - xn(cx + 1, c + 1);
        Real sx = a0 * FX[ip][i2(r, c)] + a1 * FX[ip][i2(r, c + 1)];
        if constexpr (cy != 0) {
          const Real t0 = yt(cy < 0 ? 0 : 1, r), t1 = yt(cy < 0 ? 0 : 1, r + cy);
          sx = (Real(1) - t0) * sx + t1 * (a0 * FX[ip][i2(r + cy, c)] + a1 * FX[ip][i2(r + cy, c + 1)]);
        }
        const Real b0 = yn(cy + 1, r), b1 = Real(1) - yn(cy + 1, r + 1);
        Real sy = b0 * FY[ip][i2(r, c)] + b1 * FY[ip][i2(r + 1, c)];
        if constexpr (cx != 0) {
          const Real s0 = xt(cx < 0 ? 0 : 1, c), s1 = xt(cx < 0 ? 0 : 1, c + cx);
          sy = (Real(1) - s0) * sy + s1 * (b0 * FY[ip][i2(r, c + cx)] + b1 * FY[ip][i2(r + 1, c + cx)]);
        }
        return sx + sy;
      }
    };
    // Cotangent of Zn(p+1/2) at (r, c): the transposed bilinear lerp of the
    // z-faces of columns (r, c), (r, c+c_x), (r+c_y, c), (r+c_y, c+c_x).
    auto gather_zn = [&](auto I) -> Real {
      constexpr int ip = decltype(I)::value;
      constexpr int cx = CX[ip], cy = CY[ip];
      Real sz = FZ[ip][i2(r, c)];
      Real sz2 = Real(0);
      if constexpr (cy != 0) sz2 = FZ[ip][i2(r + cy, c)];
      if constexpr (cx != 0) {
        const Real s0 = xt(cx < 0 ? 0 : 1, c), s1 = xt(cx < 0 ? 0 : 1, c + cx);
        sz = (Real(1) - s0) * sz + s1 * FZ[ip][i2(r, c + cx)];
        if constexpr (cy != 0) sz2 = (Real(1) - s0) * sz2 + s1 * FZ[ip][i2(r + cy, c + cx)];
      }
      if constexpr (cy != 0) {
        const Real t0 = yt(cy < 0 ? 0 : 1, r), t1 = yt(cy < 0 ? 0 : 1, r + cy);
        sz = (Real(1) - t0) * sz + t1 * sz2;
      }
      return sz;
    };
    // Transposed z-lerps.  Z(p) mixes f-bar+(p) with f-bar+(p - c_z) and
    // Zn(p+1/2) mixes f-bar+(p) with f-bar+(p+1).  First the populations with
    // c_z = +1, whose Z(p) is the last input of the f-bar+(p-1) cotangent.
    if (gather) {
      static_for<5>([&](auto K) {
        constexpr int ip = ZP[decltype(K)::value];
        const Real lz = gather_z(std::integral_constant<int, ip>{});
        accP[ip] += wzt_p * lz;
        accC[ip] += (Real(1) - wzt_p) * lz;
      });
    }
    // accP is now the f-bar+ cotangent of plane p-1: fold the wall ghosts
    if (wall_tile) {
      if (gx_lane || gy_lane) {
        const int side = (gx_lane ? c == kTX : r == kTY + 1) ? 1 : 0;
        const int pos = side * 16 + (gx_lane ? r : c);
        static_for<Q>([&](auto I) {
          constexpr int ip = decltype(I)::value;
          Gh[ip][pos] = accP[ip];
        });
      }
      __syncthreads();
      if (interior) {
        // ghost = q f(first) + (1 - q) f(second) from the wall side
        if (!perx) {
          const bool lo1 = i0 == 0 && c <= 1, hi1 = i0 + kTX == nx && c >= kTX - 2;
          if (lo1 || hi1) {
            const int pos = (lo1 ? 0 : 16) + r;
            const Real q = lo1 ? (Real)P.qlo[0] : (Real)P.qhi[0];
            const Real wq = (lo1 ? c == 0 : c == kTX - 1) ? q : Real(1) - q;
            static_for<Q>([&](auto I) {
              constexpr int ip = decltype(I)::value;
              accP[ip] += wq * Gh[ip][pos];
            });
          }
        }
        if (!pery) {
          const bool lo1 = j0 == 0 && r <= 2, hi1 = j0 + kTY == ny && r >= kTY - 1;
          if (lo1 || hi1) {
            const int pos = (lo1 ? 0 : 16) + c;
            const Real q = lo1 ? (Real)P.qlo[1] : (Real)P.qhi[1];
            const Real wq = (lo1 ? r == 1 : r == kTY) ? q : Real(1) - q;
            static_for<Q>([&](auto I) {
              constexpr int ip = decltype(I)::value;
              accP[ip] += wq * Gh[ip][pos];
            });
          }
        }
      }
    }
    // centre-stage adjoint of plane p-1 (as k_adj_cells)
    if (interior && p - 1 >= z0) {
      const unsigned cell = (unsigned)(p - 1) * plane + colcell;
      const long long cl = (long long)cell;
      Real ff[Q], lp[Q];
      static_for<Q>([&](auto I) {
        constexpr int ip = decltype(I)::value;
        lp[ip] = (Real)lamb.ld_f64(cell * 8u, ip * lam_pop);
        ff[ip] = ld_state<Real, Store>(bin, cell * (unsigned)sizeof(Store), ip * Nb, P, ip);
      });
      static_for<Q>([&](auto I) {
        constexpr int ip = decltype(I)::value;
        accP[ip] += Real(4.0 / 3.0) * lp[ip];
      });
      Real rho = 0, j[3] = {0, 0, 0};
      static_for<Q>([&](auto I) {
        constexpr int ip = decltype(I)::value;
        rho += ff[ip];
        if constexpr (CX[ip] != 0) j[0] += CX[ip] > 0 ? ff[ip] : -ff[ip];
        if constexpr (CY[ip] != 0) j[1] += CY[ip] > 0 ? ff[ip] : -ff[ip];
        if constexpr (CZ[ip] != 0) j[2] += CZ[ip] > 0 ? ff[ip] : -ff[ip];
      });
      Real F[3], u[3], ut[3];
      const Real invr = Real(1) / rho;
      for (int a = 0; a < 3; a++) {
        F[a] = Famb[a];
        if (P.fcell) F[a] += (Real)P.fcell[a * P.N + cl];
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
        lam_prev[ip * P.N + cl] = (double)v;
      });
      if (grad) {
        for (int a = 0; a < 3; a++) grad[a * P.N + cl] += (double)(g_F[a] + hdt * g_u[a]);
      }
    }
    // The other populations' Z(p) and every Zn(p+1/2), merged with the shift
    // of the cotangent window: accP <- plane p, accC <- plane p+1.
    if (gather) {
      static_for<Q>([&](auto I) {
        constexpr int ip = decltype(I)::value;
        constexpr int cz = CZ[ip];
        Real cur = accC[ip], nxt = Real(0);
        if constexpr (cz == 0) {
          cur += gather_z(I);
        } else if constexpr (cz < 0) {
          const Real lz = gather_z(I);
          cur += (Real(1) - wzt_m) * lz;
          nxt = wzt_m * lz;
        }
        const Real la = gather_zn(I);
        const Real wn = ip == 0 ? Real(0.5) : wzn[cz + 1];
        accP[ip] = cur + (Real(1) - wn) * la;
        accC[ip] = nxt + wn * la;
      });
    }
    // no barrier needed here: the next phase 1 writes A, An and Lm, which
    // phase 3 does not read, and the next phase 2 writes the face planes only
    // after the next phase-1 barrier; Gh is rewritten only after two more
    // barriers
  }
}

// z-chunk length: 64 planes unless that leaves fewer than 4 workgroups per
// CU (one resident workgroup of five waves per CU); halve it, down to 8.
int onepass_chunk(const KParams& P) {
  const long long cols = (long long)(P.n[0] / kTX) * (P.n[1] / kTY);
  int tz = 64;
  while (tz > 8 && cols * ((P.n[2] + tz - 1) / tz) < 4 * 256) tz /= 2;
  return tz;
}

}  // namespace

// One-pass adjoint step; handled = false when the configuration is not
// covered (the caller then runs the split path).
hipError_t launch_adjoint_onepass(Ctx& c, const void* f, const double* lam_next, double* lam_prev, double* grad,
                                  bool* handled) {
  *handled = false;
  if (c.adjoint_path != 2 || c.precision == 0 || c.kernel == 0) return hipSuccess;  // DUGKS_ADJOINT_ONEPASS
  int tx = 0, ty = 0;
  if (!fused_supported(c, &tx, &ty) || tx != kTX || ty != kTY) return hipSuccess;
  const Setup& s = c.s;
  if (!s.periodic[0] && !s.periodic[1]) return hipSuccess;  // one wall axis at most
  // lam' is addressed with 32-bit buffer offsets
  if ((double)Q * c.kp.N * 8.0 >= 2147483648.0) return hipSuccess;
  *handled = true;
  const KParams& P = c.kp;
  const int TZ = onepass_chunk(P);
  dim3 grid(P.n[0] / kTX, P.n[1] / kTY, (P.n[2] + TZ - 1) / TZ);
  if (c.precision == 1)
    hipLaunchKernelGGL((k_adj_onepass<float>), grid, dim3(kNP), 0, c.stream, P, static_cast<const float*>(f),
                       lam_next, lam_prev, grad, TZ);
  else
    hipLaunchKernelGGL((k_adj_onepass<__half>), grid, dim3(kNP), 0, c.stream, P, static_cast<const __half*>(f),
                       lam_next, lam_prev, grad, TZ);
  return hipGetLastError();
}

}  // namespace dugks
