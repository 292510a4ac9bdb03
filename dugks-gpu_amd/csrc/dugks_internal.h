// Internal declarations shared by the HIP kernels and the C ABI.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "lattice.h"
#include "setup.h"

namespace dugks {

// Kernel parameters, passed by value to every step kernel.
//
// Device storage of f: structure of arrays f[ip][k][j][i] (direction stride
// N = nx*ny*nz), interior cells only; ghost cells are never stored, the
// kernels apply the periodic wrap / wall extrapolation on the fly.
//
// Per-axis weight tables tab[a], 9 doubles per ghost-inclusive index m =
// 0..n+1 (m = 1 is the first interior cell):
//   [0..2] fraction toward m+1 at the +face of cell m for c_a = -1, 0, +1
//   [3..5] fraction toward m-1 at the -face of cell m for c_a = -1, 0, +1
//   [6]    tangential fraction toward m+1 (population with c_a = -1)
//   [7]    tangential fraction toward m-1 (population with c_a = +1)
//   [8]    dt / width(m)   (update coefficient)
struct KParams {
  int n[3];
  int periodic[3];
  long long N;    // nx*ny*nz
  int e[3];       // n+2
  long long eN;   // ghost-inclusive cell count
  double dt, hdt, bw0, bw1, bw2, fw0, fw1, fw2;
  double famb[3];
  const double* famb_dev;  // device ambient force (3 doubles), overrides famb when set
  const double* fcell;     // additive centre force [3][N], or nullptr
  const double* tab[3];
  const float* tabf[3];    // float copies of tab for FP32 arithmetic
  const double* wd[3];     // ghost-inclusive cell widths, index 0..n+1
  double qlo[3], qhi[3];   // wall ghost = q*f(1) + (1-q)*f(2) (and mirrored at the top)
  int uniform[3];
  double du[3];
  int recon;
  int flow_axis;           // streamwise axis of the flow-rate controller
  // FP16 deviation codec: f_i = w_i (rho0 + s_i), s_i = f_i / w_i - rho0
  float rho0f;
};

struct Ctx;

// Launchers (implemented in the .hip files; all stream-ordered).
hipError_t launch_step_twopass(Ctx& c, const void* fin, void* fout);
hipError_t launch_fbar_ghosts(Ctx& c, const void* fin, void* fbx);
hipError_t launch_step_fused(Ctx& c, const void* fin, void* fout, bool* handled);
// Adjoint-face pass of the fused kernel (see adjoint.hip); handled = false
// when the fused kernel does not cover the configuration.  Face record
// formats: kFacesFull, the 19 reconstructed-value cotangents per face, lam'
// in FP64; kFacesCompact (the z-marching gather's configurations, FP32 or
// FP64 arithmetic; adjoint_record_format), kCompact values per face:
// the reconstructed-value cotangents of the ten populations crossing the face
// (fw1 mu + Ac + B.c), then Ac and the two tangential components of B, from
// which the nine others follow; kFacesCompactLam32, the same with lam' in FP32.
// With the state stored in the arithmetic type (FP32 or FP64 storage) the
// compact formats also carry, after the three record arrays (at 3 kCompact eN
// values), two regions of rho, j_x, j_y, j_z of every cell ([4][N] each),
// which the gather uses instead of reading f.
constexpr int kCompact = 13;
constexpr int kFacesFull = 0, kFacesCompact = 1, kFacesCompactLam32 = 2;
// With fpart set, the pass also writes per block the cotangent of the
// ambient force along the flow axis through the face collisions (nparts
// values, at most fpart_len, else hipErrorInvalidValue); only FP64 records
// and FP32-lam compact records have this form, else hipErrorNotSupported.
// mom (compact records, storage in the arithmetic type): which of the two
// rho, j regions behind the records the pass writes; ubar (a sweep through
// the flow-rate controller on those records, with fpart): the bulk-velocity
// cotangent to add to lam' from the other region (see AdjFaceIO).
hipError_t launch_adjoint_faces_fused(Ctx& c, const void* f, const void* lam, void* faces, int format,
                                      bool* handled, double* fpart = nullptr, long long fpart_len = 0,
                                      long long* nparts = nullptr, int mom = 0, const double* ubar = nullptr);
// Whether launch_adjoint_faces_fused covers the configuration.
bool adjoint_faces_fused_covers(const Ctx& c);
// Number of per-block force partials the pass writes for a format.
long long adjoint_faces_fused_blocks(const Ctx& c, int format);
// Planar statistics along axis pa into out (device): [n_pa][9] row sums of
// u, v, w, uw, uv, vw, uu, vv, ww, then global sums, min and max of
// rho, u, v, w.  scratch holds stats_scratch_doubles(n_pa) doubles.
size_t stats_scratch_doubles(int np);
// Statistics scratch held by the library for dugks_get_statistics_noalloc:
// one array per GPU for the life of the process, sized for at most
// kNoAllocStatRows wall-normal rows.  noalloc_stats_scratch resolves its
// address on `device` (which must be current), once per device.
constexpr int kNoAllocStatRows = 2048;
hipError_t noalloc_stats_scratch(int device, double** ptr);
hipError_t launch_stats(Ctx& c, const void* f, int pa, double* scratch, double* out);
hipError_t launch_macro(Ctx& c, const void* f, double* rho, double* ux, double* uy, double* uz);
hipError_t launch_export(Ctx& c, const void* f, double* dst);
hipError_t launch_import(Ctx& c, const double* src, void* f);
hipError_t launch_global_sums(Ctx& c, const void* f, double* dev_out8);
hipError_t launch_flow_control(Ctx& c, const void* f);
hipError_t launch_perturb_force(Ctx& c, int step);
// fpart: as launch_adjoint_faces_fused (the face part of the ambient-force
// cotangent); hipErrorNotSupported where the fused face pass does not run.
hipError_t launch_adjoint(Ctx& c, const void* f, const double* lam_next, double* lam_prev,
                          double* grad_force, double* fpart = nullptr, long long fpart_len = 0,
                          long long* nparts = nullptr, int mom = 0, const double* ubar = nullptr);
// The same with lam' and lam in FP32; only FP32 arithmetic on the z-marching
// gather's configurations (adjoint_lam32_supported), else hipErrorNotSupported.
// mom, ubar: as launch_adjoint_faces_fused; with ubar the gather adds the
// same controller term (adjoint_flow_fused).
hipError_t launch_adjoint_lam32(Ctx& c, const void* f, const float* lam_next, float* lam_prev,
                                double* grad_force, double* fpart = nullptr, long long fpart_len = 0,
                                long long* nparts = nullptr, int mom = 0, const double* ubar = nullptr);
// Whether the sweep's adjoint step adds the flow-rate controller's term to
// lam' as it reads it (compact records with FP32 storage and FP32 lam, or
// FP64 throughout) instead of a pass of its own.
bool adjoint_flow_fused(const Ctx& c);
bool adjoint_lam32_supported(const Ctx& c);
// The face-record format (kFaces*) of an adjoint step of this context, lam'
// in FP32 (lam32) or FP64.
int adjoint_record_format(const Ctx& c, bool lam32);
// Whether the adjoint covers the flow-rate controller (the ambient-force
// cotangent of the face collisions is available).
bool adjoint_flow_supported(const Ctx& c);
// Loss value into dev_value (added to it when value_acc) and, when dev_grad
// is set, its gradient dJ/df written into dev_grad (grad_acc == 0) or added
// to it times grad_acc.
// force_axis >= 0: dev_value[1] also receives dJ/dF of the ambient force
// component force_axis (the loss reads u = j/rho + h F), under the same
// value_acc rule.
hipError_t launch_loss(Ctx& c, const void* f, int kind, double* dev_value, double* dev_grad, double grad_acc = 0,
                       bool value_acc = false, int force_axis = -1);
// The same with the gradient in FP32 (FP32 arithmetic only).
hipError_t launch_loss_lam32(Ctx& c, const void* f, int kind, double* dev_value, float* dev_grad,
                             double grad_acc = 0, bool value_acc = false, int force_axis = -1);
// Adjoint of the flow-rate controller.  The controller reads the bulk
// velocity U = sum (j_a/rho) dV / V of the new state; its cotangent ubar
// (device scalar) adds ubar (dV/V) (c_ia - u_a)/rho to lam of every
// population of every cell.
hipError_t launch_flow_cotangent(Ctx& c, const void* f, const double* ubar, double* lam);
hipError_t launch_flow_cotangent_lam32(Ctx& c, const void* f, const double* ubar, float* lam);
// out[0] = sum over cells of v[axis N + cell] (fixed order).
hipError_t launch_axis_sum(Ctx& c, const double* v, int axis, double* out);
// Device scalars of the controller's reverse recurrence (SweepPool::ctl):
// nu = dJ/dF of the force after the current one (dJ/dF^0 when the replay
// ends), the summed cell force gradient after the last adjoint step and its
// fresh value, the bulk-velocity cotangent to apply, the controller's
// previous-valid flag when the record began, and the loss value (discarded)
// followed by the loss's dJ/dF.
constexpr int kFlowCtlNuNext = 0, kFlowCtlSum = 1, kFlowCtlSumNew = 2, kFlowCtlUbar = 3, kFlowCtlFlag0 = 4,
              kFlowCtlLoss = 5, kFlowCtlLen = 8;
// One reverse step of the controller's scalar recurrence over ctl; step0:
// the bulk velocity being differentiated is that of step 1 (the flag then
// comes from the recorded state).
hipError_t launch_flow_adjoint_scalar(Ctx& c, double* ctl, const double* fpart, long long nparts, bool step0,
                                      double invT);
// dst[i] = src[i] (FP32 to FP64), n values.
hipError_t launch_widen(Ctx& c, const float* src, double* dst, long long n);
hipError_t launch_axpy(Ctx& c, double a, const double* x, double* y, long long n);

// dst[i] = v for i < n (a fill kernel: safe inside captured graphs, where
// memset nodes are avoided).
hipError_t launch_fill(Ctx& c, double* dst, double v, long long n);

// Checkpoint pool of the sqrt(T) forward-adjoint sweep, kept on the context
// between calls (dugks_forward_record fills it, dugks_adjoint_replay reads
// it): checkpoints cps[q] = f^{q K} (K = interval), segment states seg[s] =
// f^{n0+1+s} of the segment being replayed, the final state fT, and two
// adjoint buffers.  Buffers are allocated on demand and reused.
struct SweepPool {
  std::vector<void*> cps, seg;
  void* fT = nullptr;
  void* lam[2] = {nullptr, nullptr};
  size_t lam_bytes = 0;
  double bytes = 0;  // device bytes held by the pool
  // the recorded forward sweep
  bool recorded = false;
  int istep0 = 0, T = 0, loss_kind = 0, interval = 0, ncp = 0;
  // with the flow-rate controller: the ambient force along the flow axis
  // before each step (frec[n] drove step n, frec[T] is the final one), the
  // controller's scalars and the per-block face partials of the force
  // cotangent
  bool flow = false;
  bool replayed = false;  // the reverse sweep of the record has run (ctl holds dJ/dF^0)
  double* frec = nullptr;
  size_t frec_len = 0;
  double* ctl = nullptr;
  double* fpart = nullptr;
  size_t fpart_len = 0;
};

// Library context.
struct Ctx {
  Setup s;
  KParams kp;
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int precision = 0;
  int kernel = 0;
  size_t state_bytes = 0;     // bytes of one stored state (all cells)
  void* f[2] = {nullptr, nullptr};
  int cur = 0;                // f[cur] holds the current state
  void* fbx = nullptr;        // ghost-inclusive f-bar (two-pass kernels, adjoint)
  double* stats = nullptr;      // planar statistics: partial sums and results
  double* stats_noalloc = nullptr;  // this device's library-held statistics scratch
  size_t stats_len = 0;
  void* adj_faces = nullptr;    // adjoint face records (19 cotangents, or kCompact values), 3 axes, ghost-inclusive
  int adj_faces_format = -1;    // what adj_faces holds: -1 nothing yet, 0 full records, 1 compact records
  double* tab[3] = {nullptr, nullptr, nullptr};
  double* wd[3] = {nullptr, nullptr, nullptr};
  float* tabf[3] = {nullptr, nullptr, nullptr};
  double* famb_dev = nullptr; // 3 doubles + controller state
  double* fcell = nullptr;    // [3][N] additive centre force
  double* red = nullptr;      // reduction scratch
  size_t red_len = 0;
  std::string err;
  double bytes_allocated = 0;
  // optional timing of the main step kernel (dugks_time_steps)
  hipEvent_t tev[2] = {nullptr, nullptr};
  bool timing = false;
  std::vector<float> kernel_ms;
  int launches_per_step = 0;
  // fused kernel: z-chunk length and the per-block flow-rate partial sums
  int fused_tz = 0;
  double* flow_partial = nullptr;
  size_t flow_partial_len = 0;
  // the step kernels leave the flow-rate controller alone (the adjoint's
  // replay, which sets the recorded force itself)
  bool flow_frozen = false;
  SweepPool sweep;

  hipError_t alloc(void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess) bytes_allocated += (double)bytes;
    return e;
  }
};

// Spectral (DHIT) statistics of the current state, spectral.hip: the number
// of wavenumber shells of the grid, and shells [2][nb] (E, 2 nu k^2 E) plus
// sums[4] (physical 0.5 <|u|^2>; sums of d^2, d^3, d^4 of the longitudinal
// derivatives over the three components) into host buffers.
int spectral_bins(const Ctx& c);
hipError_t spectral_statistics(Ctx& c, double* shells, double sums[4]);

// Builds the kernel parameter block and the device tables for a setup.
hipError_t build_params(Ctx& c);

}  // namespace dugks
