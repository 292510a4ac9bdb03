// Fused DUGKS step: one kernel per time step, f read once and written once.
//
// Decomposition.  A workgroup owns a TX x TY tile of (x, y) columns and
// marches along z over a chunk of TZ planes.  Its threads are TY + 2 rows of
// TX lanes (the interior rows and the two y-halo rows) plus one wave for the
// two x-halo columns (corners included); every thread owns one column.  Each
// plane it:
//   - loads f of its column at plane p+1 and computes f-bar+ (calc_fb1 of
//     the reference, src/dugks.cu:1131-1196); halo columns recompute it, wall
//     ghost columns extrapolate it (src/dugks.cu:1234-1251 and the stretched
//     form, src/fortran/channel/src/dugks.cuf:365-374);
//   - reconstructs every face once (the reference computes each interior face
//     twice, quirk Q5): the trilinear interpolation of src/dugks.cu:756-906 is
//     evaluated z-lerp first, so the x- and y-faces of plane p read a single
//     LDS plane Z(p) and the z-face between p and p+1 reads Zn(p+1/2);
//   - applies wall bounce-back and the face collision (f2ruvwh + feq_s,
//     src/dugks.cu:675-709, 1201-1207) and keeps only the populations that
//     cross the face;
//   - updates plane p (src/dugks.cu:945-956) into the other f buffer.
// Two barriers per plane: [f-bar(p+1); write Z(p), Zn(p+1/2)] | [faces] |
// [update], the update of one plane overlapping the next plane's f-bar.
// z must be periodic; x and y may be periodic or walls, uniform or stretched.
// GRAD: the gradient reconstruction instead (calc_fx, src/dugks.cu:421-583:
// the mean of the two cells beside the face less h c.g, with the normal
// derivative from those two cells and the tangential ones by central
// differences of the two-cell means), on uniform axes; the LDS then holds a
// ring of the f-bar+ planes p-1, p and p+1, written as each is computed.
// Other configurations use the two-pass kernels.
#include "device_common.h"

namespace dugks {

namespace {

#include "fused_common.h"

constexpr int kXcds = 8;  // accelerator complex dies of an MI355X

// threads of a TX x TY tile's workgroup (Tile<TX, TY>::NT)
constexpr int Tile_nt(int TX, int TY) { return ((TY + 2) * TX + 2 * (TY + 2) + 63) / 64 * 64; }

template <int TX, int TY>
struct Tile {
  static constexpr int NROW = (TY + 2) * TX;    // the tile rows and the two y-halo rows
  static constexpr int NHALO = 2 * (TY + 2);     // the two x-halo columns
  static constexpr int NT = (NROW + NHALO + 63) / 64 * 64;
  static constexpr int AC = TX + 2;
  static_assert(TX % 16 == 0 && TX <= 64, "rows of 16, 32, 48 or 64 lanes");
};

// Inputs and outputs of the adjoint-face mode of the fused kernel.
struct AdjFaceIO {
  const void* lam;    // lam' = dJ/df(n+1), SoA [19][nz][ny][nx], FP64 or (LamT = float) FP32
  void* faces;        // out: [3][19][ghost-inclusive face slot], arithmetic type; COMPACT:
                      // [3][kCompact][slot], then (storage in the arithmetic type) two
                      // [4][N] regions of rho, j of every cell
  int mom = 0;        // COMPACT, such storage: the region the pass writes rho, j of fin into
  // sweep through the flow-rate controller: lam' still lacks the bulk
  // velocity's cotangent ubar (dV/V) (c_ia - u_a)/rho of f(n+1), whose rho, j
  // the previous pass left in the other region; added as lam' is loaded
  // (nullptr: lam' is complete)
  const double* ubar = nullptr;
  double inv_vol = 0;  // 1 / V
};

// Gradient reconstruction of one population at a face's foot point: the
// mean m of the two cells beside the face less h c.g (calc_fx,
// src/dugks.cu:421-583); the terms of the components c_a = 0 are left out
// (they add zero).
template <int cx, int cy, int cz, class Real>
__device__ __forceinline__ Real grad_foot(Real m, Real h, Real gx, Real gy, Real gz) {
  if constexpr (cx == 0 && cy == 0 && cz == 0) {
    return m;
  } else {
    Real s;
    if constexpr (cx != 0) s = Real(cx) * gx;
    if constexpr (cy != 0) {
      if constexpr (cx != 0) s += Real(cy) * gy; else s = Real(cy) * gy;
    }
    if constexpr (cz != 0) {
      if constexpr (cx != 0 || cy != 0) s += Real(cz) * gz; else s = Real(cz) * gz;
    }
    return m - h * s;
  }
}

// ADJ = false: one DUGKS step fin -> fout.
// ADJ = true: the adjoint-face pass of the step taken from fin: every face is
// reconstructed and its collision linearised as in the step, and the
// cotangents of its 19 reconstructed values, driven by lam' (adj.lam), are
// written to adj.faces; nothing else is written.  FADJ (with ADJ): also the
// cotangent of the ambient force along the flow axis through the face
// collisions, summed per block into partial[tile] (fixed order).
template <class Real, class Store, int TX, int TY, bool ADJ = false, bool COMPACT = false, class LamT = double,
          bool FADJ = false, bool GRAD = false>
// FP32 arithmetic: two blocks per CU (at most 256 registers per lane)
__global__ void __launch_bounds__(((TY + 2) * TX + 2 * (TY + 2) + 63) / 64 * 64, sizeof(Real) == 4 ? 2 : 1) k_fused(
    KParams P, const Store* __restrict__ fin,
                                                          Store* __restrict__ fout, int TZ,
                                                          double* __restrict__ partial, AdjFaceIO adj) {
  using T = Tile<TX, TY>;
  constexpr int AC = T::AC;
  static_assert(!(GRAD && ADJ), "the adjoint differentiates the trilinear reconstruction");
  // With FP16 non-equilibrium storage the step carries shifted populations
  // h = f - w (see EqShift): f-bar+, the LDS planes, the face values and the
  // update are all f - w, decoded from and encoded to the stored rho - 1, u
  // and non-equilibrium part without forming f (fewer operations than
  // through f).  The other storage modes and the adjoint-face pass work on f.
  constexpr bool kShift = std::is_same<Store, NeqHalf>::value && !ADJ;
  // LDS planes indexed (row, column).  16-wide rows put four rows in a wave:
  // column-major planes (column stride TY + 2, even) then keep the lanes of
  // a ds_read on distinct banks; 32-wide rows are conflict-free row-major.
  constexpr bool CM = TX == 16;
  auto ia = [](int rr, int cc) { return CM ? cc * (TY + 2) + rr : rr * AC + cc; };       // A, An, Lam
  auto ic = [](int rr, int cc) { return CM ? cc * TY + rr : rr * (TX + 1) + cc; };       // Cx (rows 0..TY-1)
  auto id = [](int rr, int cc) { return CM ? cc * (TY + 1) + rr : rr * TX + cc; };       // Dy (rows 0..TY)
  // GRAD: A holds the ring of f-bar+ planes, plane k in slot (k + 3) % 3
  // (population ip of slot s at A[s * Q + ip]); An is unused
  __shared__ Real A[GRAD ? 3 * Q : Q][(TY + 2) * AC];  // Z(p): tangential z-lerp of f-bar(p)
  __shared__ Real An[GRAD ? 1 : Q][(TY + 2) * AC];     // Zn(p+1/2): normal z-lerp between p and p+1
  __shared__ Real Cx[ADJ ? 1 : 10][TY * (TX + 1)];  // crossing populations of -x faces
  __shared__ Real Dy[ADJ ? 1 : 10][(TY + 1) * TX];  // crossing populations of -y faces
  __shared__ Real Lam[ADJ ? Q : 1][(TY + 2) * AC];  // adjoint-face pass: lam'(p) of the tile and halo
  __shared__ Real WX[6][TX + 2];           // weights per tile column
  __shared__ Real WY[6][TY + 2];           // weights per tile row

  const int nx = P.n[0], ny = P.n[1], nz = P.n[2];
  const int t = threadIdx.x;
  // XCD-aware tile order.  Workgroups are dealt round-robin to the 8 XCDs
  // (workgroup b runs on XCD b % 8), each with its own L2.  Tile tt =
  // (b % 8) * (ntiles / 8) + b / 8 gives every XCD a contiguous run of tiles
  // (x fastest), so the tiles whose halos overlap share one L2 and run at
  // about the same time; the plain order would put x-neighbours on
  // different XCDs and every x-halo line would come from HBM.
  const int gx = nx / TX, gy = ny / TY;
  const int ntiles = gx * gy * (int)gridDim.z;
  const int b = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const int tt = (ntiles % kXcds == 0) ? (b % kXcds) * (ntiles / kXcds) + b / kXcds : b;
  const int bx = tt % gx, by = (tt / gx) % gy, bz = tt / (gx * gy);
  const int i0 = bx * TX, j0 = by * TY;
  const int z0 = bz * TZ;
  const int zend = min(z0 + TZ, nz);
  const unsigned Nb = (unsigned)(P.N * (long long)sizeof(Store));  // bytes per population
  const unsigned span = (unsigned)state_bytes<Store>((unsigned long long)P.N);
  const StateBuf<Store> bin(fin, span), bout(fout, span);

  // column of this thread in tile coordinates: r = 0..TY+1 (y = j0-1+r),
  // c = -1..TX (x = i0+c).  Lanes: the TY interior rows first (whole waves
  // when TY*TX is a multiple of 64, so the interior-only phases run on as few
  // waves as possible), then the two y-halo rows, then the two x-halo columns
  // packed behind them; lanes past those idle.
  int r = 0, c = 0;
  bool active = true;
  if (t < TY * TX) {
    r = t / TX + 1;
    c = t % TX;
  } else if (t < T::NROW) {
    const int u = t - TY * TX;
    r = u < TX ? 0 : TY + 1;
    c = u % TX;
  } else if (t < T::NROW + T::NHALO) {
    const int u = t - T::NROW;
    r = u % (TY + 2);
    c = u < TY + 2 ? -1 : TX;
  } else {
    active = false;
  }
  const bool interior = t < T::NROW && r >= 1 && r <= TY;
  int gi = i0 + c, gj = j0 - 1 + r;
  if (P.periodic[0]) gi = (gi + nx) % nx;
  if (P.periodic[1]) gj = (gj + ny) % ny;
  // ghost-inclusive table indices of the column
  const int mi = gi + 1, mj = gj + 1;
  // weight tables in the arithmetic type (layout documented on KParams)
  auto tx = [&](int idx) { return tabv<Real>(P, 0, idx); };
  auto ty = [&](int idx) { return tabv<Real>(P, 1, idx); };
  auto tz = [&](int idx) { return tabv<Real>(P, 2, idx); };

  Real Famb[3];
  ambient_force<Real>(P, Famb);
  const Real hdt = (Real)P.hdt, fw0 = (Real)P.fw0, fw1 = (Real)P.fw1, fw2 = (Real)P.fw2;
  // uniform z weights (z is periodic and uniform here)
  const Real wzt_m = tz(9 + 6);  // tangential, toward p+1 (c_z = -1)
  const Real wzt_p = tz(9 + 7);  // tangential, toward p-1 (c_z = +1)
  const Real wzn[3] = {tz(9 + 0), tz(9 + 1), tz(9 + 2)};  // normal, toward p+1
  const Real updz = tz(9 + 8);
  // GRAD: 1/d and 1/(2d) per axis (every axis uniform)
  Real gin[3] = {Real(0), Real(0), Real(0)}, gtn[3] = {Real(0), Real(0), Real(0)};
  if constexpr (GRAD) {
    for (int a = 0; a < 3; a++) {
      gin[a] = (Real)(1.0 / P.du[a]);
      gtn[a] = (Real)(0.5 / P.du[a]);
    }
  }
  auto rslot = [](int k) { return (k + 3) % 3; };  // ring slot of plane k >= -1
  // Per-column and per-row weights of the tile, copied once into LDS: the
  // tables are not written by this kernel, but the compiler cannot prove that
  // against the f stores and would reload them (or hold them in registers)
  // across the plane loop.  Row k of WX/WY: 0-2 the -face normal weight
  // toward the upper cell for c = -1, 0, +1; 3-4 the tangential shift
  // weights for c < 0 and c > 0; 5 dt/d.  The -face index is clamped for the
  // ghost column/row, which never owns a -face.
  for (int u = t; u < 6 * (TX + 2) + 6 * (TY + 2); u += T::NT) {
    const bool isx = u < 6 * (TX + 2);
    const int v = isx ? u : u - 6 * (TX + 2);
    const int n = isx ? TX + 2 : TY + 2;
    const int k = v / n, q = v % n;
    int g = isx ? i0 - 1 + q : j0 - 1 + q;
    const int na = isx ? nx : ny;
    if (P.periodic[isx ? 0 : 1]) g = (g + na) % na;
    const int m = g + 1;
    const int idx = k < 3 ? max(m - 1, 0) * 9 + k : m * 9 + (k + 3);
    (isx ? WX[k][q] : WY[k][q]) = isx ? tx(idx) : ty(idx);
  }
  auto xn = [&](int k) { return WX[k][c + 1]; };
  auto xt = [&](int s) { return WX[3 + s][c + 1]; };
  auto yn = [&](int k) { return WY[k][r]; };
  auto yt = [&](int s) { return WY[3 + s][r]; };

  Real fbm[5], fbc[Q], fbn[Q], zm[10];
  double flow = 0;
  const int as = P.flow_axis;
  // FADJ: the face collisions' force cotangents of this thread's faces
  double fsum[3] = {0, 0, 0};

  // Zn(p+1/2) of this column into An: z-normal lerp between fb(p) and fb(p+1)
  auto store_zn = [&](const Real (&lo)[Q], const Real (&hi)[Q]) {
    if (!active) return;
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      if constexpr (ip == 0)
        An[0][ia(r, c + 1)] = Real(0.5) * (lo[0] + hi[0]);
      else
        An[ip][ia(r, c + 1)] = lerp_(lo[ip], hi[ip], wzn[CZ[ip] + 1]);
    });
  };
  // +z face of an interior column from Zn in A: bilinear over the upwind
  // tangential neighbours, then collision; out = crossing populations
  auto zface_rec = [&](Face<Real, 2>& fc) {
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      constexpr int cx = CX[ip], cy = CY[ip];
      Real v = An[ip][ia(r, c + 1)];
      if constexpr (cx != 0) {
        const Real w = xt(cx < 0 ? 0 : 1);
        v = lerp_(v, An[ip][ia(r, c + 1 - cx)], w);
        if constexpr (cy != 0) {
          const Real wy = yt(cy < 0 ? 0 : 1);
          const Real v2 = lerp_(An[ip][ia(r - cy, c + 1)], An[ip][ia(r - cy, c + 1 - cx)], w);
          v = lerp_(v, v2, wy);
        }
      } else if constexpr (cy != 0) {
        const Real wy = yt(cy < 0 ? 0 : 1);
        v = lerp_(v, An[ip][ia(r - cy, c + 1)], wy);
      }
      fc.template put<ip>(v);
    });
  };
  // GRAD: population ip of ring slot sl at LDS position (rr, cc)
  auto ga = [&](int sl, int ip, int rr, int cc) -> Real { return A[sl * Q + ip][ia(rr, cc)]; };
  // GRAD reconstructions from the ring (slots sm, s0, sp of planes p-1, p,
  // p+1): the x-face between tile columns c-1 and c of row r, the y-face
  // between rows r-1 and r of column c, both at plane p, and the z-face
  // between planes p (s0) and p+1 (sp) of the thread's column
  auto grad_face_x = [&](Face<Real, 0>& fc, int sm, int s0, int sp) {
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      constexpr int cx = CX[ip], cy = CY[ip], cz = CZ[ip];
      const Real fl = ga(s0, ip, r, c), fr = ga(s0, ip, r, c + 1);
      Real gy = Real(0), gz = Real(0);
      if constexpr (cy != 0)
        gy = (Real(0.5) * (ga(s0, ip, r + 1, c) + ga(s0, ip, r + 1, c + 1)) -
              Real(0.5) * (ga(s0, ip, r - 1, c) + ga(s0, ip, r - 1, c + 1))) * gtn[1];
      if constexpr (cz != 0)
        gz = (Real(0.5) * (ga(sp, ip, r, c) + ga(sp, ip, r, c + 1)) -
              Real(0.5) * (ga(sm, ip, r, c) + ga(sm, ip, r, c + 1))) * gtn[2];
      fc.template put<ip>(grad_foot<cx, cy, cz>(Real(0.5) * (fr + fl), hdt, (fr - fl) * gin[0], gy, gz));
    });
  };
  auto grad_face_y = [&](Face<Real, 1>& fc, int sm, int s0, int sp) {
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      constexpr int cx = CX[ip], cy = CY[ip], cz = CZ[ip];
      const Real fl = ga(s0, ip, r - 1, c + 1), fr = ga(s0, ip, r, c + 1);
      Real gx = Real(0), gz = Real(0);
      if constexpr (cx != 0)
        gx = (Real(0.5) * (ga(s0, ip, r - 1, c + 2) + ga(s0, ip, r, c + 2)) -
              Real(0.5) * (ga(s0, ip, r - 1, c) + ga(s0, ip, r, c))) * gtn[0];
      if constexpr (cz != 0)
        gz = (Real(0.5) * (ga(sp, ip, r - 1, c + 1) + ga(sp, ip, r, c + 1)) -
              Real(0.5) * (ga(sm, ip, r - 1, c + 1) + ga(sm, ip, r, c + 1))) * gtn[2];
      fc.template put<ip>(grad_foot<cx, cy, cz>(Real(0.5) * (fr + fl), hdt, gx, (fr - fl) * gin[1], gz));
    });
  };
  auto grad_face_z = [&](Face<Real, 2>& fc, int s0, int sp) {
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      constexpr int cx = CX[ip], cy = CY[ip], cz = CZ[ip];
      const Real fl = ga(s0, ip, r, c + 1), fr = ga(sp, ip, r, c + 1);
      Real gx = Real(0), gy = Real(0);
      if constexpr (cx != 0)
        gx = (Real(0.5) * (ga(s0, ip, r, c + 2) + ga(sp, ip, r, c + 2)) -
              Real(0.5) * (ga(s0, ip, r, c) + ga(sp, ip, r, c))) * gtn[0];
      if constexpr (cy != 0)
        gy = (Real(0.5) * (ga(s0, ip, r + 1, c + 1) + ga(sp, ip, r + 1, c + 1)) -
              Real(0.5) * (ga(s0, ip, r - 1, c + 1) + ga(sp, ip, r - 1, c + 1))) * gtn[1];
      fc.template put<ip>(grad_foot<cx, cy, cz>(Real(0.5) * (fr + fl), hdt, gx, gy, (fr - fl) * gin[2]));
    });
  };
  // the +z face of plane p (GRAD: from ring slots s0 and sp)
  auto zface = [&](Real (&out)[10], int s0, int sp) {
    Face<Real, 2> fc;
    if constexpr (GRAD)
      grad_face_z(fc, s0, sp);
    else
      zface_rec(fc);
    if constexpr (kShift)
      fc.collide_shift(Famb, hdt, fw0, fw2, out);
    else
      fc.collide(Famb, hdt, fw0, fw1, fw2, out);
  };
  // adjoint-face mode: lam' of population ip at an interior cell (0 for a
  // ghost, cell < 0) and the store of a face's cotangents
  // (both arrays are < 2 GiB, checked at launch: 32-bit buffer offsets)
  constexpr unsigned kLamBytes = sizeof(LamT);
  const RawBuf lamb(adj.lam, (unsigned)(Q * P.N * kLamBytes));
  const unsigned lam_pop = (unsigned)(P.N * kLamBytes);         // bytes per population
  const unsigned face_pop = (unsigned)(P.eN * sizeof(Real));    // bytes per population
  // lam' of the thread's own column at plane k (0 for a wall ghost column).
  // The pass computes the -x and -y faces of its cells (and a wall's +x, +y
  // face, whose outside column is a ghost), so the right halo column, the top
  // halo row and the bottom-left corner never need lam' and load none.
  const bool lam_cell = active && gi >= 0 && gi < nx && gj >= 0 && gj < ny && c < TX && r <= TY &&
                        !(c < 0 && r == 0);
  auto load_lam = [&](int k, Real (&l)[Q]) {
    const unsigned vb = lam_cell ? (unsigned)(((long long)k * ny + gj) * nx + gi) * kLamBytes : kOutOfRange;
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      if constexpr (ip != 0) l[ip] = (Real)lamb.ld<LamT>(vb, ip * lam_pop);
    });
  };
  // the flow-rate controller's term of lam' (AdjFaceIO::ubar): rho and j_a
  // of f(n+1) at the column's cell of plane k, loaded with lam', added before
  // lam' is used: lam'_i += s (c_ia - u_a), s = ubar (dV/V) / rho
  constexpr bool kFlowFuse = ADJ && FADJ && COMPACT && std::is_same<Store, Real>::value;
  const bool fuse = kFlowFuse && adj.ubar != nullptr;
  const RawBuf momn(static_cast<Real*>(adj.faces) + 3LL * kCompact * P.eN + (1 - adj.mom) * 4LL * P.N,
                    (unsigned)(4 * P.N * sizeof(Real)));
  Real s0 = Real(0), ex = Real(0), ey = Real(0), ez = Real(0);
  if constexpr (kFlowFuse) {
    if (fuse && lam_cell)
      s0 = (Real)(adj.ubar[0] * P.wd[0][mi] * P.wd[1][mj] * P.wd[2][1] * adj.inv_vol);
    ex = as == 0 ? Real(1) : Real(0);
    ey = as == 1 ? Real(1) : Real(0);
    ez = as == 2 ? Real(1) : Real(0);
  }
  auto load_flow = [&](int k, Real& rn, Real& jn) {
    if constexpr (kFlowFuse) {
      const unsigned vb = fuse && lam_cell ? (unsigned)(((long long)k * ny + gj) * nx + gi) * sizeof(Real)
                                           : kOutOfRange;
      rn = momn.ld<Real>(vb, 0);
      jn = momn.ld<Real>(vb, (unsigned)((1 + as) * P.N * sizeof(Real)));
    }
  };
  auto add_flow = [&](Real (&l)[Q], Real rn, Real jn) {
    if constexpr (kFlowFuse) {
      // rn is 0 where nothing was loaded (s0 is 0 there)
      const Real inv = Real(1) / (rn != Real(0) ? rn : Real(1));
      const Real s = s0 * inv, t0 = -s * (jn * inv);
      static_for<Q>([&](auto I) {
        constexpr int ip = decltype(I)::value;
        if constexpr (ip != 0) l[ip] += fma_rn(Real(CX[ip]) * ex + Real(CY[ip]) * ey + Real(CZ[ip]) * ez, s, t0);
      });
    }
  };
  Real lamc[Q], lamn[Q], rnn = Real(0), jnn = Real(0);
  if constexpr (ADJ) {
    load_lam(z0, lamc);
    if constexpr (kFlowFuse) {
      Real rn = Real(0), jn = Real(0);
      load_flow(z0, rn, jn);
      add_flow(lamc, rn, jn);
    }
  }
  // compact form (COMPACT): per face the kCompact values of compact_record
  auto store_compact = [&](int a, long long slot, const Real (&rec)[kCompact]) {
    const RawBuf fb(static_cast<Real*>(adj.faces) + (long long)a * kCompact * P.eN, kCompact * face_pop);
    const unsigned vb = (unsigned)slot * (unsigned)sizeof(Real);
    static_for<kCompact>([&](auto K) {
      constexpr int k = decltype(K)::value;
      fb.st(rec[k], vb, k * face_pop);
    });
  };
  auto store_face = [&](int a, long long slot, const Real (&lam)[Q]) {
    const RawBuf fb(static_cast<Real*>(adj.faces) + (long long)a * Q * P.eN, Q * face_pop);
    const unsigned vb = (unsigned)slot * (unsigned)sizeof(Real);
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      fb.st(lam[ip], vb, ip * face_pop);
    });
  };
  const long long e0 = P.e[0], e1 = P.e[1];

  // Columns inside the domain (tile and periodic halo) compute f-bar+ from f
  // of their own column; wall ghost columns extrapolate it (fbar_column).
  const bool own = active && gi >= 0 && gi < nx && gj >= 0 && gj < ny;
  // stored f of the own column at plane k (decoded at use)
  auto load_own = [&](int k, RawCell<Real, Store>& raw) { raw.load(bin, ((unsigned)k * ny + gj) * nx + gi, (unsigned)P.N); };
  // body force at plane k of the own column
  auto force_own = [&](int k, Real (&F)[3]) {
    F[0] = Famb[0];
    F[1] = Famb[1];
    F[2] = Famb[2];
    if (P.fcell) {
      const unsigned cc = ((unsigned)k * ny + gj) * nx + gi;
      F[0] += (Real)P.fcell[cc];
      F[1] += (Real)P.fcell[P.N + cc];
      F[2] += (Real)P.fcell[2 * P.N + cc];
    }
  };
  // The update of plane p needs 4/3 f-bar+(p) - 1/3 f(p); it is formed when
  // f(p) arrives for f-bar+(p), so f is read once per step: gc holds it for
  // the current plane, gn for the next.  Written f-bar+ + (f-bar+ - f)/3:
  // the two terms of the difference carry the same mass, so the rounding of
  // 1/3 biases nothing (the rounded 4/3 and 1/3 of the plain form add about
  // 3e-8 of the mass per step in FP32).
  Real gc[Q], gn[Q];
  auto form_g = [&](const Real (&fb)[Q], const Real (&ff)[Q], Real (&g)[Q]) {
    static_for<Q>([&](auto I) {
      constexpr int ip = decltype(I)::value;
      g[ip] = fb[ip] + (fb[ip] - ff[ip]) * Real(1.0 / 3.0);
    });
  };

  // f of the own column one plane ahead of its f-bar: the loads are issued
  // one iteration before their f-bar needs them.  The first (plane z0+1) is
  // issued ahead of the pre-loop's loads, so it arrives with them.
  constexpr bool kPrefetch = true;
  RawCell<Real, Store> fpre;
  if (kPrefetch && own) load_own((z0 + 1) % nz, fpre);

  // ---- pre-loop: f-bar at z0-1 and z0, and the -z face of plane z0 ----
  {
    const int km = (z0 - 1 + nz) % nz;
    if (active) {
      fbar_column<Real, Store, kShift>(P, bin, Nb, gi, gj, km, Famb, fbn);
      if (!ADJ && own) {
        RawCell<Real, Store> raw;
        Real ff[Q], F[3], rho, u0[3];
        load_own(z0, raw);
        force_own(z0, F);
        if constexpr (kShift) {
          raw.decode_shift(P, ff, rho, u0);
          centre_fbar_shift<Real>(ff, rho, u0, F, hdt, (Real)P.bw0, (Real)P.bw2, fbc);
        } else {
          raw.decode(P, ff, rho, u0);
          centre_fbar<Real>(ff, rho, u0, F, hdt, (Real)P.bw0, (Real)P.bw2, fbc);
        }
        if constexpr (!ADJ) form_g(fbc, ff, gc);
      } else {
        fbar_column<Real, Store, kShift>(P, bin, Nb, gi, gj, z0, Famb, fbc);
      }
      static_for<5>([&](auto K) {
        constexpr int k = decltype(K)::value;
        fbm[k] = fbn[ZP[k]];
      });
    }
    if constexpr (GRAD) {
      if (active) {
        static_for<Q>([&](auto I) {
          constexpr int ip = decltype(I)::value;
          A[rslot(z0 - 1) * Q + ip][ia(r, c + 1)] = fbn[ip];
          A[rslot(z0) * Q + ip][ia(r, c + 1)] = fbc[ip];
        });
      }
    } else {
      store_zn(fbn, fbc);
    }
    __syncthreads();
    if constexpr (!ADJ) {
      if (interior) zface(zm, rslot(z0 - 1), rslot(z0));
    }
    __syncthreads();
  }

  if constexpr (kPrefetch) fpre.settle();
  // compact adjoint-face records with the state stored in the arithmetic
  // type (FP32 or FP64): rho and j of every cell
  // for the gather, written by the interior lanes from the prefetched f.
  // A chunk writes its planes z0+1 .. zend (mod nz), so the chunks cover
  // every plane once.
  auto store_moments = [&](int k, Real rho, const Real (&jm)[3]) {
    const RawBuf mb(static_cast<Real*>(adj.faces) + 3LL * kCompact * P.eN + adj.mom * 4LL * P.N,
                    (unsigned)(4 * P.N * sizeof(Real)));
    const unsigned vb = (((unsigned)k * ny + gj) * nx + gi) * (unsigned)sizeof(Real);
    const unsigned pop = (unsigned)(P.N * sizeof(Real));
    mb.st(rho, vb, 0);
    mb.st(jm[0], vb, pop);
    mb.st(jm[1], vb, 2 * pop);
    mb.st(jm[2], vb, 3 * pop);
  };
  // One plane p: fbc / gc / lamc hold f-bar+, 4/3 f-bar+ - 1/3 f and lam' of
  // plane p, and the plane computes those of p + 1 into fbn / gn / lamn.  The
  // step's loop below runs the planes in pairs with the two register sets
  // swapped, so no copy moves plane p+1's values into plane p's registers.
  auto plane = [&](int p, Real (&fbc)[Q], Real (&fbn)[Q], Real (&gc)[Q], Real (&gn)[Q], Real (&lamc)[Q],
                   Real (&lamn)[Q]) {
    const int kn = (p + 1) % nz;
    if constexpr (ADJ) {
      load_lam(kn, lamn);
      load_flow(kn, rnn, jnn);
    }
    if (own) {
      if constexpr (!kPrefetch) load_own(kn, fpre);
      Real F[3];
      force_own(kn, F);
      Real ff[Q], rho, u0[3];
      if constexpr (kShift) {
        fpre.decode_shift(P, ff, rho, u0);
        centre_fbar_shift<Real>(ff, rho, u0, F, hdt, (Real)P.bw0, (Real)P.bw2, fbn);
      } else {
        fpre.decode(P, ff, rho, u0);
        if constexpr (ADJ && COMPACT && std::is_same<Store, Real>::value) {
          if (interior) {
            Real jm[3];
            cell_moments<Real>(ff, rho, jm);
            store_moments(kn, rho, jm);
          }
        }
        centre_fbar<Real>(ff, rho, u0, F, hdt, (Real)P.bw0, (Real)P.bw2, fbn);
      }
      if constexpr (!ADJ) {
        if (interior) form_g(fbn, ff, gn);
      }
      if (kPrefetch && p + 1 < zend) load_own((p + 2) % nz, fpre);
    } else if (active) {
      fbar_column<Real, Store, kShift>(P, bin, Nb, gi, gj, kn, Famb, fbn);
    }
    // ---- phase 1: Z(p) (tangential z-lerp of f-bar(p)) and Zn(p+1/2) ----
    // (GRAD: f-bar+(p+1) into its ring slot)
    const int sm = rslot(p - 1), s0 = rslot(p), sp = rslot(p + 1);
    if constexpr (GRAD) {
      if (active) {
        static_for<Q>([&](auto I) {
          constexpr int ip = decltype(I)::value;
          A[sp * Q + ip][ia(r, c + 1)] = fbn[ip];
        });
      }
    } else if (active) {
      static_for<Q>([&](auto I) {
        constexpr int ip = decltype(I)::value;
        constexpr int cz = CZ[ip];
        if constexpr (cz == 0) {
          A[ip][ia(r, c + 1)] = fbc[ip];
        } else if constexpr (cz > 0) {
          constexpr int zk = ip == 5 ? 0 : (ip == 11 ? 1 : (ip == 12 ? 2 : (ip == 15 ? 3 : 4)));
          A[ip][ia(r, c + 1)] = lerp_(fbc[ip], fbm[zk], wzt_p);
        } else {
          A[ip][ia(r, c + 1)] = lerp_(fbc[ip], fbn[ip], wzt_m);
        }
      });
    }
    if constexpr (!GRAD) store_zn(fbc, fbn);
    if constexpr (ADJ) {
      if (active) {
        static_for<Q>([&](auto I) {
          constexpr int ip = decltype(I)::value;
          if constexpr (ip != 0) Lam[ip][ia(r, c + 1)] = lamc[ip];
        });
      }
    }
    __syncthreads();
    // ---- phase 2: the faces of plane p (-x, -y from Z; +z from Zn) ----
    // the forward needs the tile's right and top boundary faces too; the
    // adjoint-face pass leaves them to the next tile except at a wall
    const bool xhalo = active && t >= T::NROW && c == TX && r >= 1 && r <= TY;
    const bool do_x = interior || (xhalo && (!ADJ || (!P.periodic[0] && gi == nx)));
    const bool do_y = t < T::NROW && r >= 1 && (!ADJ || r <= TY || (!P.periodic[1] && gj == ny));
    if (do_x) {
      // face between x-columns gi-1 (left, ghost-inclusive index mi-1) and gi
      Face<Real, 0> fc;
      if constexpr (GRAD) grad_face_x(fc, sm, s0, sp);
      else static_for<Q>([&](auto I) {
        constexpr int ip = decltype(I)::value;
        constexpr int cx = CX[ip], cy = CY[ip];
        if constexpr (ip == 0) {
          fc.template put<0>(Real(0.5) * (A[0][ia(r, c)] + A[0][ia(r, c + 1)]));
        } else {
          const Real w = xn(cx + 1);  // fraction toward the right cell
          Real v = lerp_(A[ip][ia(r, c)], A[ip][ia(r, c + 1)], w);
          if constexpr (cy != 0) {
            const Real wy = yt(cy < 0 ? 0 : 1);
            const Real v2 = lerp_(A[ip][ia(r - cy, c)], A[ip][ia(r - cy, c + 1)], w);
            v = lerp_(v, v2, wy);
          }
          fc.template put<ip>(v);
        }
      });
      const bool lw = !P.periodic[0] && gi == 0, hw = !P.periodic[0] && gi == nx;
      if (!P.periodic[0]) fc.bounce(lw, hw);
      if constexpr (ADJ) {
        const Real uL = WX[5][c], uR = WX[5][c + 1];
        Real mu[10], lam[Q];
        static_for<10>([&](auto K) {
          constexpr int k = decltype(K)::value;
          constexpr int ip = XD[k];
          mu[k] = Real(CX[ip]) * (uR * Lam[ip][ia(r, c + 1)] - uL * Lam[ip][ia(r, c)]);
        });
        // (no "cond ? nx : gi": the optimiser turns a select between two
        // captured variables into a select between their addresses, which
        // keeps them in scratch memory)
        const int slot = gi + ((gi == 0 && P.periodic[0]) ? nx : 0);
        const long long fs = ((long long)(p + 1) * e1 + (gj + 1)) * e0 + slot;
        if constexpr (COMPACT) {
          Real Ac, B[3], rec[kCompact];
          fc.template adjoint_coeffs<FADJ>(Famb, hdt, fw0, fw2, mu, Ac, B, fsum);
          compact_record<Real, 0>(mu, Ac, B, fw1, rec);
          store_compact(0, fs, rec);
        } else {
          fc.template adjoint<FADJ>(Famb, hdt, fw0, fw1, fw2, mu, lw, hw, lam, fsum);
          store_face(0, fs, lam);
        }
      } else {
        Real out[10];
        if constexpr (kShift)
          fc.collide_shift(Famb, hdt, fw0, fw2, out);
        else
          fc.collide(Famb, hdt, fw0, fw1, fw2, out);
        static_for<10>([&](auto K) {
          constexpr int k = decltype(K)::value;
          Cx[k][ic(r - 1, c)] = out[k];
        });
      }
    }
    if (do_y) {
      // face between y-rows gj-1 (below) and gj
      Face<Real, 1> fc;
      if constexpr (GRAD) grad_face_y(fc, sm, s0, sp);
      else static_for<Q>([&](auto I) {
        constexpr int ip = decltype(I)::value;
        constexpr int cx = CX[ip], cy = CY[ip];
        if constexpr (ip == 0) {
          fc.template put<0>(Real(0.5) * (A[0][ia(r - 1, c + 1)] + A[0][ia(r, c + 1)]));
        } else {
          const Real w = yn(cy + 1);  // fraction toward the upper cell
          Real lo = A[ip][ia(r - 1, c + 1)], hi = A[ip][ia(r, c + 1)];
          if constexpr (cx != 0) {
            const Real wx = xt(cx < 0 ? 0 : 1);
            lo = lerp_(lo, A[ip][ia(r - 1, c + 1 - cx)], wx);
            hi = lerp_(hi, A[ip][ia(r, c + 1 - cx)], wx);
          }
          fc.template put<ip>(lerp_(lo, hi, w));
        }
      });
      const bool lw = !P.periodic[1] && gj == 0, hw = !P.periodic[1] && gj == ny;
      if (!P.periodic[1]) fc.bounce(lw, hw);
      if constexpr (ADJ) {
        const Real uL = WY[5][r - 1], uR = WY[5][r];
        Real mu[10], lam[Q];
        static_for<10>([&](auto K) {
          constexpr int k = decltype(K)::value;
          constexpr int ip = YD[k];
          mu[k] = Real(CY[ip]) * (uR * Lam[ip][ia(r, c + 1)] - uL * Lam[ip][ia(r - 1, c + 1)]);
        });
        const int slot = gj + ((gj == 0 && P.periodic[1]) ? ny : 0);
        const long long fs = ((long long)(p + 1) * e1 + slot) * e0 + (gi + 1);
        if constexpr (COMPACT) {
          Real Ac, B[3], rec[kCompact];
          fc.template adjoint_coeffs<FADJ>(Famb, hdt, fw0, fw2, mu, Ac, B, fsum);
          compact_record<Real, 1>(mu, Ac, B, fw1, rec);
          store_compact(1, fs, rec);
        } else {
          fc.template adjoint<FADJ>(Famb, hdt, fw0, fw1, fw2, mu, lw, hw, lam, fsum);
          store_face(1, fs, lam);
        }
      } else {
        Real out[10];
        if constexpr (kShift)
          fc.collide_shift(Famb, hdt, fw0, fw2, out);
        else
          fc.collide(Famb, hdt, fw0, fw1, fw2, out);
        static_for<10>([&](auto K) {
          constexpr int k = decltype(K)::value;
          Dy[k][id(r - 1, c)] = out[k];
        });
      }
    }
    Real zp[10];
    // lam' of plane p+1 complete (its loads were issued at the top of the plane)
    if constexpr (kFlowFuse) add_flow(lamn, rnn, jnn);
    if constexpr (ADJ) {
      if (interior) {
        // +z face of the cell: between planes p and p+1 (z is periodic)
        Face<Real, 2> fc;
        zface_rec(fc);
        Real mu[10], lam[Q];
        static_for<10>([&](auto K) {
          constexpr int k = decltype(K)::value;
          constexpr int ip = ZD[k];
          mu[k] = Real(CZ[ip]) * updz * (lamn[ip] - lamc[ip]);
        });
        const long long fs = ((long long)(p + 1) * e1 + (gj + 1)) * e0 + (gi + 1);
        if constexpr (COMPACT) {
          Real Ac, B[3], rec[kCompact];
          fc.template adjoint_coeffs<FADJ>(Famb, hdt, fw0, fw2, mu, Ac, B, fsum);
          compact_record<Real, 2>(mu, Ac, B, fw1, rec);
          store_compact(2, fs, rec);
        } else {
          fc.template adjoint<FADJ>(Famb, hdt, fw0, fw1, fw2, mu, false, false, lam, fsum);
          store_face(2, fs, lam);
        }
      }
    } else {
      if (interior) zface(zp, s0, sp);
    }
    __syncthreads();
    // ---- phase 3: update of plane p ----
    // The f prefetched in phase 1 is taken before this plane's stores, so
    // that the next plane's decode does not wait for them (see
    // RawCell::settle).  FP32 storage takes it after the update's arithmetic,
    // which then overlaps the loads still in flight (0.831 against 0.850 ms
    // on the 256x192x256 channel); the other storage modes lost with that
    // order (FP64 2.005 against 1.919 ms, FP16 non-equilibrium 1.028
    // against 0.927 ms) and take it first.
    constexpr bool kLateSettle = std::is_same<Store, float>::value;
    if constexpr (!ADJ && kPrefetch && !kLateSettle) fpre.settle();
    if constexpr (!ADJ) {
      if (interior) {
        const unsigned cell = ((unsigned)p * ny + gj) * nx + gi;
        const unsigned vb = cell * (unsigned)sizeof(Store);
        const Real updx = WX[5][c + 1], updy = WY[5][r];
        // f~(p+1) = 4/3 f-bar+ - 1/3 f~ - sum_a (dt/dx_a) c_a (F_a+ - F_a-)
        constexpr bool kNeq = std::is_same<Store, NeqHalf>::value;
        Real fnew[Q];
        static_for<Q>([&](auto I) {
          constexpr int ip = decltype(I)::value;
          Real v = gc[ip];
          constexpr int kx = cross_index<0>(ip), ky = cross_index<1>(ip), kz = cross_index<2>(ip);
          if constexpr (kx >= 0) v -= Real(CX[ip]) * updx * (Cx[kx][ic(r - 1, c + 1)] - Cx[kx][ic(r - 1, c)]);
          if constexpr (ky >= 0) v -= Real(CY[ip]) * updy * (Dy[ky][id(r, c)] - Dy[ky][id(r - 1, c)]);
          if constexpr (kz >= 0) {
            v -= Real(CZ[ip]) * updz * (zp[kz] - zm[kz]);
            zm[kz] = zp[kz];
          }
          // per-population stores go out as each value is ready (FP32
          // storage: after the prefetch is taken, below); FP16
          // non-equilibrium storage encodes the whole cell below
          if constexpr (!kNeq && !kLateSettle) {
            if constexpr (kShift)
              st_state_shift<Real, Store>(bout, vb, ip * Nb, v, P, ip);
            else
              st_state<Real, Store>(bout, vb, ip * Nb, v, P, ip);
          }
          fnew[ip] = v;
        });
        if constexpr (kLateSettle) {
          if constexpr (kPrefetch) fpre.settle();
          static_for<Q>([&](auto I) {
            constexpr int ip = decltype(I)::value;
            st_state<Real, Store>(bout, vb, ip * Nb, fnew[ip], P, ip);
          });
        }
        Real rr, u0n[3];
        if constexpr (kNeq && kShift) {
          store_neq_shift<Real>(bout, cell, (unsigned)P.N, fnew, rr, u0n);
        } else if constexpr (kNeq) {
          store_raw_cell<Real, Store>(bout, cell, (unsigned)P.N, P, fnew, true, rr, u0n);
        } else if (partial) {
          // the controller's reading only: FP32 sums suffice (the moments
          // that feed the step are summed in FP64, cell_rho_u)
          Real jn[3];
          cell_moments<Real>(fnew, rr, jn);
          if constexpr (kShift) rr += Real(1);
          moment_velocity<Real>(rr, jn, u0n);
        }
        // bulk velocity of the new state for the flow-rate controller
        // (per-cell moments in the arithmetic type, the sum over cells in FP64)
        if (partial) flow += (double)(as == 0 ? u0n[0] : (as == 1 ? u0n[1] : u0n[2]));
      } else if constexpr (kPrefetch && kLateSettle) {
        fpre.settle();
      }
    }
    if (!GRAD && active) {
      static_for<5>([&](auto K) {
        constexpr int k = decltype(K)::value;
        fbm[k] = fbc[ZP[k]];
      });
    }
    // no barrier: the next phase 1 writes A/An (GRAD: the ring slot of p-1,
    // last read in this plane's phase 2), which phase 3 does not read, and
    // the next phase 2 writes Cx/Dy only after the next phase-1 barrier
  };
  // The step runs planes in pairs; the adjoint-face pass (more registers
  // live across the loop) one plane per iteration with the values of p+1
  // copied down.
  constexpr bool kPairs = !ADJ;
  if constexpr (kPairs) {
    for (int p = z0; p < zend; p += 2) {
      plane(p, fbc, fbn, gc, gn, lamc, lamn);
      if (p + 1 < zend) plane(p + 1, fbn, fbc, gn, gc, lamn, lamc);
    }
  } else {
    for (int p = z0; p < zend; p++) {
      plane(p, fbc, fbn, gc, gn, lamc, lamn);
      if (active) {
        static_for<Q>([&](auto I) {
          constexpr int ip = decltype(I)::value;
          fbc[ip] = fbn[ip];
          if constexpr (ADJ && ip != 0) lamc[ip] = lamn[ip];
          if constexpr (!ADJ) gc[ip] = gn[ip];
        });
      }
    }
  }

  if (!ADJ && partial) {
    // fixed-order block reduction of the flow sums; every column of the tile
    // has the same z width (z is uniform here)
    double v[1] = {interior ? flow * (P.wd[0][mi] * P.wd[1][mj] * P.wd[2][1]) : 0.0};
    constexpr int kSum[1] = {0};
    block_reduce<1, T::NT>(v, kSum);
    if (t == 0) partial[tt] = v[0];
  }
  if constexpr (ADJ && FADJ) {
    // the flow-axis component (assigned, not selected as an lvalue: a select
    // between array elements can keep the array in scratch memory)
    double v[1] = {fsum[0]};
    if (as == 1) v[0] = fsum[1];
    if (as == 2) v[0] = fsum[2];
    constexpr int kSum[1] = {0};
    block_reduce<1, T::NT>(v, kSum);
    if (t == 0) partial[tt] = v[0];
  }
}

// Final reduction of the per-block flow sums <j/rho> dV and the flow-rate
// controller update (the law of k_flow_control).
__global__ void __launch_bounds__(256) k_flow_finish(const double* partial, int nparts, double* famb, int axis,
                                                     double volume, double target, double gain, double dt) {
  double s[1] = {0};
  for (int b = threadIdx.x; b < nparts; b += 256) s[0] += partial[b];
  constexpr int kSum[1] = {0};
  block_reduce<1, 256>(s, kSum);
  if (threadIdx.x == 0) {
    const double Ub = s[0] / volume;
    const double Fprev = famb[axis];
    double drag = Fprev;
    if (famb[4] != 0) drag = Fprev - (Ub - famb[3]) / dt;
    famb[axis] = drag + gain * (target - Ub) / dt;
    famb[3] = Ub;
    famb[4] = 1;
  }
}

// z-chunk length.  A chunk recomputes f-bar+ of the two planes before it.
// With two resident workgroups per CU (FP32 arithmetic on 4-wave tiles) the
// other workgroup covers that prologue, and more rounds of workgroups even
// out the finish: 32 planes, on the 256x192x256 channel (256 columns of
// 16x12, 2048 tiles) FP32 0.823 ms against 0.833 with 64 planes and 0.833
// with 16, FP16 deviation storage 0.738 against 0.758 and 0.740 (128
// planes: 0.865 / 0.783); on 128^3 (128 columns of 32x4, 512 tiles) 32
// planes against 8: Taylor-Green FP32 0.148 against 0.166 ms, FP64 0.397
// against 0.444 ms.  With one workgroup per CU (FP64, and the 8-wave
// 32x12 tiles) nothing covers it and long chunks win for the step: 128
// planes, FP32 on 32x12 0.724 ms against 0.731 with 64 and 0.751 with 32,
// FP64 on 16x12 1.921 against 1.945 and 1.974.  The adjoint-face passes
// (adj) keep 32 planes (with 128 their adjoint steps were 0.7 % slower).
// Either start is halved, down to 8, while the grid has fewer tiles than one
// round of resident workgroups on the 256 CUs.
int blocks_per_cu(bool fp64, int TX, int TY) { return fp64 || Tile_nt(TX, TY) > 256 ? 1 : 2; }
int z_chunk(const Ctx& cx, int TX, int TY, bool adj = false) {
  if (cx.fused_tz > 0) return cx.fused_tz;
  const long long cols = (long long)(cx.kp.n[0] / TX) * (cx.kp.n[1] / TY);
  const int per_cu = blocks_per_cu(cx.precision == 0, TX, TY);
  int tz = per_cu == 1 && !adj ? 128 : 32;
  while (tz > 8 && cols * ((cx.kp.n[2] + tz - 1) / tz) < per_cu * 256) tz /= 2;
  return tz;
}

template <class Real, class Store, int TX, int TY, bool GRAD>
hipError_t run_fused(Ctx& cx, const void* fin, void* fout) {
  const KParams& P = cx.kp;
  using T = Tile<TX, TY>;
  const int TZ = z_chunk(cx, TX, TY);
  dim3 grid(P.n[0] / TX, P.n[1] / TY, (P.n[2] + TZ - 1) / TZ);
  const long long nblocks = (long long)grid.x * grid.y * grid.z;
  double* partial = nullptr;
  const bool flow = cx.s.flow_rate != 0 && !cx.flow_frozen;
  if (flow) {
    if ((size_t)nblocks > cx.flow_partial_len) return hipErrorInvalidValue;
    partial = cx.flow_partial;
  }
  if (cx.timing) hipEventRecord(cx.tev[0], cx.stream);
  hipLaunchKernelGGL((k_fused<Real, Store, TX, TY, false, false, double, false, GRAD>), grid, dim3(T::NT), 0,
                     cx.stream, P,
                     static_cast<const Store*>(fin), static_cast<Store*>(fout), TZ, partial, AdjFaceIO{});
  if (cx.timing) hipEventRecord(cx.tev[1], cx.stream);
  if (flow) {
    const Setup& s = cx.s;
    const double vol = s.width[0] * s.width[1] * s.width[2];
    hipLaunchKernelGGL(k_flow_finish, dim3(1), dim3(256), 0, cx.stream, partial, (int)nblocks, cx.famb_dev,
                       s.stream_axis(), vol, s.flow_rate, s.flow_gain, s.dt);
  }
  cx.launches_per_step = flow ? 2 : 1;
  return hipGetLastError();
}

template <int TX, int TY>
dim3 adj_faces_grid(const Ctx& cx, int* TZ) {
  const KParams& P = cx.kp;
  *TZ = z_chunk(cx, TX, TY, true);
  return dim3(P.n[0] / TX, P.n[1] / TY, (P.n[2] + *TZ - 1) / *TZ);
}

// FADJ: also the per-block force partials into fpart (nparts of them).
template <class Real, class Store, int TX, int TY, bool COMPACT = false, class LamT = double, bool FADJ = false>
hipError_t run_adj_faces(Ctx& cx, const void* f, const void* lam, void* faces, double* fpart = nullptr,
                         long long fpart_len = 0, long long* nparts = nullptr, int mom = 0,
                         const double* ubar = nullptr) {
  const KParams& P = cx.kp;
  using T = Tile<TX, TY>;
  int TZ = 0;
  const dim3 grid = adj_faces_grid<TX, TY>(cx, &TZ);
  if constexpr (FADJ) {
    const long long nb = (long long)grid.x * grid.y * grid.z;
    if (nb > fpart_len) return hipErrorInvalidValue;
    *nparts = nb;
  }
  hipLaunchKernelGGL((k_fused<Real, Store, TX, TY, true, COMPACT, LamT, FADJ>), grid, dim3(T::NT), 0, cx.stream, P,
                     static_cast<const Store*>(f), static_cast<Store*>(nullptr), TZ, FADJ ? fpart : nullptr,
                     AdjFaceIO{lam, faces, mom, ubar, 1.0 / (cx.s.width[0] * cx.s.width[1] * cx.s.width[2])});
  return hipGetLastError();
}

}  // namespace

// Either reconstruction (gradient reconstruction needs uniform axes, which
// the configuration enforces); the adjoint-face mode trilinear only.
bool fused_supported(const Ctx& c, int* tx, int* ty) {
  const Setup& s = c.s;
  if (!s.periodic[2] || !s.uniform[2]) return false;
  // buffer offsets are 32-bit: one state array must span < 4 GiB
  if ((double)c.state_bytes >= 4294967296.0) return false;
  // Tiles of 4 waves: 16 x 12 cells when the grid divides (1.31 f-bar
  // evaluations per cell, the x-halo lanes packed behind the rows), else
  // 32 x 4 (1.59 per cell, the last wave holding only the 12 x-halo lanes;
  // 16 x 8 tiles of 3 waves were no faster on 128^3).  Four waves per block
  // keep FP64 at one wave per SIMD without spills and fit two FP32 blocks
  // (at most 256 registers per lane) on a CU.
  if (s.n[0] % 16 == 0 && s.n[1] % 12 == 0) {
    *tx = 16;
    *ty = 12;
  } else {
    *tx = 32;
    *ty = 4;
  }
  if (s.n[0] % *tx != 0 || s.n[1] % *ty != 0) return false;
  // a wall axis needs two interior cells per side of the extrapolation
  if (!s.periodic[0] && s.n[0] < 3) return false;
  if (!s.periodic[1] && s.n[1] < 3) return false;
  return true;
}

namespace {
template <bool GRAD>
hipError_t launch_fused_recon(Ctx& c, const void* fin, void* fout, bool tall) {
  switch (c.precision) {
    case 0:
      return tall ? run_fused<double, double, 16, 12, GRAD>(c, fin, fout)
                  : run_fused<double, double, 32, 4, GRAD>(c, fin, fout);
    case 1:
      return tall ? run_fused<float, float, 16, 12, GRAD>(c, fin, fout)
                  : run_fused<float, float, 32, 4, GRAD>(c, fin, fout);
    case 2:
      return tall ? run_fused<float, __half, 16, 12, GRAD>(c, fin, fout)
                  : run_fused<float, __half, 32, 4, GRAD>(c, fin, fout);
    default:
      return tall ? run_fused<float, NeqHalf, 16, 12, GRAD>(c, fin, fout)
                  : run_fused<float, NeqHalf, 32, 4, GRAD>(c, fin, fout);
  }
}
}  // namespace

hipError_t launch_step_fused(Ctx& c, const void* fin, void* fout, bool* handled) {
  int tx = 32, ty = 4;
  *handled = false;
  if (!fused_supported(c, &tx, &ty)) return hipSuccess;
  *handled = true;
  // FP32 storage takes 32 x 12 tiles where the grid divides: one workgroup
  // of 8 waves per CU, rows of 32 floats that are whole 128-byte lines for
  // every load and store (16-wide rows store half lines), f-bar+ on 1.24x the
  // cells.  On the 256x192x256 channel 0.751 ms against 0.819 on 16 x 12;
  // the FP16 storage modes lost with them (non-equilibrium 1.080 against
  // 0.967 ms, deviation 0.809 against 0.771) and FP64 would need 210 KB of
  // LDS.
  if (c.precision == 1 && c.s.recon == 0 && c.kp.n[0] % 32 == 0 && c.kp.n[1] % 12 == 0)
    return run_fused<float, float, 32, 12, false>(c, fin, fout);
  return c.s.recon == 0 ? launch_fused_recon<false>(c, fin, fout, tx == 16)
                        : launch_fused_recon<true>(c, fin, fout, tx == 16);
}

bool adjoint_faces_fused_covers(const Ctx& c) {
  int tx = 32, ty = 4;
  // the adjoint does not cover FP16 non-equilibrium storage or the gradient
  // reconstruction
  if (c.kernel == 0 || c.precision == 3 || c.s.recon != 0 || !fused_supported(c, &tx, &ty)) return false;
  // lam' and each face array are addressed with 32-bit buffer offsets
  const double rb = c.precision == 0 ? 8.0 : 4.0;
  return (double)Q * c.kp.N * 8.0 < 2147483648.0 && (double)Q * c.kp.eN * rb < 2147483648.0;
}

namespace {
// Tiles of the adjoint-face pass: 0 = 32 x 4, 1 = 16 x 12, 2 = 32 x 12.
// Compact records go with the z-marching gather (nx % 32 == 0).  With FP32
// arithmetic their rows must be whole 128-byte lines, so 32-wide: 16 x 12
// tiles (half-line stores) took the adjoint step with FP32 lam to 2.92 ms on
// the 256x192x256 channel against 2.76 on 32 x 4, and 32 x 12 tiles (one
// workgroup of 8 waves per CU, f-bar+ on 1.24x the cells instead of 1.59x,
// 93 % of the lanes busy) bring it to 2.60 where ny % 12 == 0.  Both lam
// forms take the same tiles, so the same kernel code gives the same bits
// for a lam' that FP32 holds exactly.  With FP64 arithmetic 16 doubles are a
// whole line, and 16 x 12 tiles serve every format where the grid divides
// (FP64 adjoint step 5.88 against 6.61 ms on 32 x 4).  FP32 full records
// take the step's tiles.
int adj_tile(const Ctx& c, int format, int tx) {
  const bool compact32 = c.precision != 0 && format != kFacesFull;
  if (compact32) return c.kp.n[1] % 12 == 0 ? 2 : 0;
  return tx == 16 ? 1 : 0;
}
}  // namespace

hipError_t launch_adjoint_faces_fused(Ctx& c, const void* f, const void* lam, void* faces, int format,
                                      bool* handled, double* fpart, long long fpart_len, long long* nparts, int mom,
                                      const double* ubar) {
  int tx = 32, ty = 4;
  *handled = false;
  if (!adjoint_faces_fused_covers(c) || (format == kFacesCompactLam32 && c.precision == 0)) return hipSuccess;
  fused_supported(c, &tx, &ty);
  *handled = true;
  const int tile = adj_tile(c, format, tx);
  const bool tall = tile == 1, wide = tile == 2;
  if (fpart) {
    // with the ambient-force cotangent: the formats of the forward-adjoint
    // sweep (FP64, and FP32 arithmetic with lam in FP32)
    if (c.precision == 0 && format == kFacesFull)
      return tall ? run_adj_faces<double, double, 16, 12, false, double, true>(c, f, lam, faces, fpart, fpart_len,
                                                                                nparts)
                  : run_adj_faces<double, double, 32, 4, false, double, true>(c, f, lam, faces, fpart, fpart_len,
                                                                               nparts);
    if (c.precision == 0 && format == kFacesCompact)
      return tall ? run_adj_faces<double, double, 16, 12, true, double, true>(c, f, lam, faces, fpart, fpart_len,
                                                                               nparts, mom, ubar)
                  : run_adj_faces<double, double, 32, 4, true, double, true>(c, f, lam, faces, fpart, fpart_len,
                                                                              nparts, mom, ubar);
    if (c.precision == 1 && format == kFacesCompactLam32)
      return wide ? run_adj_faces<float, float, 32, 12, true, float, true>(c, f, lam, faces, fpart, fpart_len,
                                                                            nparts, mom, ubar)
                  : run_adj_faces<float, float, 32, 4, true, float, true>(c, f, lam, faces, fpart, fpart_len,
                                                                           nparts, mom, ubar);
    if (c.precision == 2 && format == kFacesCompactLam32)
      return wide ? run_adj_faces<float, __half, 32, 12, true, float, true>(c, f, lam, faces, fpart, fpart_len,
                                                                             nparts)
                  : run_adj_faces<float, __half, 32, 4, true, float, true>(c, f, lam, faces, fpart, fpart_len,
                                                                            nparts);
    return hipErrorNotSupported;
  }
  switch (c.precision) {
    case 0:
      if (format == kFacesCompact)
        return tall ? run_adj_faces<double, double, 16, 12, true>(c, f, lam, faces, nullptr, 0, nullptr, mom)
                    : run_adj_faces<double, double, 32, 4, true>(c, f, lam, faces, nullptr, 0, nullptr, mom);
      return tall ? run_adj_faces<double, double, 16, 12>(c, f, lam, faces)
                  : run_adj_faces<double, double, 32, 4>(c, f, lam, faces);
    case 1:
      if (format == kFacesCompactLam32)
        return wide ? run_adj_faces<float, float, 32, 12, true, float>(c, f, lam, faces, nullptr, 0, nullptr, mom)
                    : run_adj_faces<float, float, 32, 4, true, float>(c, f, lam, faces, nullptr, 0, nullptr, mom);
      if (format == kFacesCompact)
        return wide ? run_adj_faces<float, float, 32, 12, true>(c, f, lam, faces, nullptr, 0, nullptr, mom)
                    : run_adj_faces<float, float, 32, 4, true>(c, f, lam, faces, nullptr, 0, nullptr, mom);
      return tall ? run_adj_faces<float, float, 16, 12>(c, f, lam, faces)
                  : run_adj_faces<float, float, 32, 4>(c, f, lam, faces);
    default:
      if (format == kFacesCompactLam32)
        return wide ? run_adj_faces<float, __half, 32, 12, true, float>(c, f, lam, faces)
                    : run_adj_faces<float, __half, 32, 4, true, float>(c, f, lam, faces);
      if (format == kFacesCompact)
        return wide ? run_adj_faces<float, __half, 32, 12, true>(c, f, lam, faces)
                    : run_adj_faces<float, __half, 32, 4, true>(c, f, lam, faces);
      return tall ? run_adj_faces<float, __half, 16, 12>(c, f, lam, faces)
                  : run_adj_faces<float, __half, 32, 4>(c, f, lam, faces);
  }
}

long long adjoint_faces_fused_blocks(const Ctx& c, int format) {
  int tx = 32, ty = 4, TZ = 0;
  if (!fused_supported(c, &tx, &ty)) return 0;
  const int tile = adj_tile(c, format, tx);
  const dim3 g = tile == 1 ? adj_faces_grid<16, 12>(c, &TZ)
                           : (tile == 2 ? adj_faces_grid<32, 12>(c, &TZ) : adj_faces_grid<32, 4>(c, &TZ));
  return (long long)g.x * g.y * g.z;
}

}  // namespace dugks
