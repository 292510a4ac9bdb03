// C ABI of the engine (include/dugks.h).  Host C++ compiled with hipcc.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/dugks.h"
#include "dugks_internal.h"
#include "io.h"

namespace dugks {
size_t reduction_scratch_doubles();
}

using namespace dugks;

struct dugks_ctx {
  Ctx c;
};

namespace {

thread_local std::string g_init_err;

void release_sweep(Ctx& c);

int fail(dugks_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->c.err = msg;
  return code;
}

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      return fail(ctx, DUGKS_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

#define CHECK_CTX()                        \
  do {                                     \
    if (!ctx) return DUGKS_EARG;           \
    hipSetDevice(ctx->c.device);           \
  } while (0)

// bytes of one stored state of N cells: 19 values per cell (FP64, FP32, FP16
// deviations), or rho, u in FP32 and 19 FP16 values (FP16 non-equilibrium)
size_t state_bytes_of(int precision, size_t N) {
  switch (precision) {
    case DUGKS_PREC_FP64: return N * Q * 8;
    case DUGKS_PREC_FP32: return N * Q * 4;
    case DUGKS_PREC_FP16S: return N * Q * 2;
    default: return N * (4 * 4 + Q * 2);
  }
}
size_t real_bytes(int precision) { return precision == 0 ? 8 : 4; }

// Weight tables of one axis (layout documented on KParams).
std::vector<double> axis_table(const Setup& s, int a) {
  const int n = s.n[a];
  const double h = s.hdt, dt = s.dt;
  std::vector<double> t((size_t)(n + 2) * 9, 0.0);
  const auto& d = s.d[a];
  for (int m = 0; m <= n + 1; m++) {
    double* r = &t[(size_t)m * 9];
    const double dm = d[m];
    const double dp1 = d[std::min(m + 1, n + 1)], dm1 = d[std::max(m - 1, 0)];
    for (int c = -1; c <= 1; c++) {
      if (s.uniform[a]) {
        // src/dugks.cu:768-772
        const double dx = s.du[a];
        const double m1 = (dx / 2. - h * c) / dx;
        r[c + 1] = m1;
        r[3 + c + 1] = 1 - m1;
      } else {
        // src/fortran/channel/src/dugks.cuf:541, 546
        r[c + 1] = (dm / 2 - h * c) / (0.5 * dm + 0.5 * dp1);
        r[3 + c + 1] = (dm / 2 + h * c) / (0.5 * dm + 0.5 * dm1);
      }
    }
    if (s.uniform[a]) {
      r[6] = std::fabs((-h * -1) / s.du[a]);
      r[7] = std::fabs((-h * 1) / s.du[a]);
      r[8] = dt / s.du[a];
    } else {
      // src/fortran/channel/src/dugks.cuf:595, 658, 833
      r[6] = std::fabs((-h * -1) / (0.5 * dm + 0.5 * dp1));
      r[7] = std::fabs((-h * 1) / (0.5 * dm + 0.5 * dm1));
      r[8] = dt / dm;
    }
  }
  return t;
}

int create(const Setup& s, const dugks_opts* opts, dugks_ctx** out) {
  std::unique_ptr<dugks_ctx> holder(new dugks_ctx);
  dugks_ctx* ctx = holder.get();
  Ctx& c = ctx->c;
  c.s = s;
  dugks_opts o;
  dugks_default_opts(&o);
  if (opts) o = *opts;
  c.precision = o.precision >= 0 ? o.precision : s.precision;
  if (c.precision < 0 || c.precision > 3) return fail(nullptr, DUGKS_EARG, "bad precision");
  c.s.precision = c.precision;
  c.kernel = o.kernel >= 0 ? o.kernel : s.kernel;
  c.state_bytes = state_bytes_of(c.precision, s.ncell());
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    g_init_err = "no HIP device available";
    return DUGKS_EHIP;
  }
  if (o.device >= 0) {
    if (o.device >= ndev) {
      g_init_err = "device ordinal out of range";
      return DUGKS_EARG;
    }
    c.device = o.device;
  } else {
    hipGetDevice(&c.device);
  }
  hipSetDevice(c.device);
  // address of this device's statistics scratch (dugks_get_statistics_noalloc)
  if (noalloc_stats_scratch(c.device, &c.stats_noalloc) != hipSuccess) {
    g_init_err = "cannot resolve the statistics scratch on this device";
    return DUGKS_EHIP;
  }
  if (o.stream) {
    c.stream = static_cast<hipStream_t>(o.stream);
  } else {
    if (hipStreamCreate(&c.stream) != hipSuccess) {
      g_init_err = "hipStreamCreate failed";
      return DUGKS_EHIP;
    }
    c.own_stream = true;
  }
  if (build_params(c) != hipSuccess) {
    g_init_err = "device allocation failed: " + c.err;
    dugks_free(holder.release());
    return DUGKS_ENOMEM;
  }
  const size_t N = c.s.ncell();
  for (int b = 0; b < 2; b++)
    if (c.alloc(&c.f[b], c.state_bytes) != hipSuccess) {
      g_init_err = "cannot allocate f (" + std::to_string(c.state_bytes) + " bytes)";
      dugks_free(holder.release());
      return DUGKS_ENOMEM;
    }
  // initial state
  std::vector<double> f;
  try {
    if (c.s.restart) {
      f.resize(N * Q);
      read_dump(join_path(c.s.base_dir, dump_name(c.s.nrestart)), f.data(), N * Q, 8);
    } else {
      initial_state(c.s, f);
    }
  } catch (const ConfigError& e) {
    g_init_err = e.what();
    dugks_free(holder.release());
    return e.code;
  }
  int rc = dugks_set_f(ctx, f.data());
  if (rc != DUGKS_OK) {
    g_init_err = c.err;
    dugks_free(holder.release());
    return rc;
  }
  if (c.s.flow_auto) {
    double sums[8];
    if ((rc = dugks_global_sums(ctx, sums)) != DUGKS_OK) {
      g_init_err = c.err;
      dugks_free(holder.release());
      return rc;
    }
    // target of the controller: the momentum bulk velocity <j/rho> of the
    // initial state (the output velocity minus h F)
    c.s.flow_rate = sums[4 + c.s.stream_axis()] / sums[7] - c.s.hdt * c.s.famb[c.s.stream_axis()];
  }
  if (o.verbose) {
    std::printf("dugks: grid %d x %d x %d, precision %d, kernel %d, dt=%.9g tau=%.9g\n", s.n[0], s.n[1],
                s.n[2], c.precision, c.kernel, s.dt, s.tau);
    std::printf("dugks: bw0=%.9g bw1=%.9g bw2=%.9g fw0=%.9g fw1=%.9g fw2=%.9g\n", s.bw0, s.bw1, s.bw2,
                s.fw0, s.fw1, s.fw2);
  }
  *out = holder.release();
  return DUGKS_OK;
}

int ensure_fcell(dugks_ctx* ctx) {
  Ctx& c = ctx->c;
  if (!c.fcell) {
    HIPCHK(c.alloc((void**)&c.fcell, c.s.ncell() * 3 * sizeof(double)));
    HIPCHK(hipMemsetAsync(c.fcell, 0, c.s.ncell() * 3 * sizeof(double), c.stream));
  }
  c.kp.fcell = c.fcell;
  return DUGKS_OK;
}

// One step of the current state into the other buffer.
int step_once(dugks_ctx* ctx, int istep) {
  Ctx& c = ctx->c;
  if (c.s.perturb) {
    int rc = ensure_fcell(ctx);
    if (rc) return rc;
    HIPCHK(launch_perturb_force(c, istep));
  }
  const void* fin = c.f[c.cur];
  void* fout = c.f[1 - c.cur];
  bool handled = false;
  if (c.kernel != DUGKS_KERNEL_TWOPASS) HIPCHK(launch_step_fused(c, fin, fout, &handled));
  if (handled) {
    c.cur = 1 - c.cur;  // the fused launch also ran the flow-rate controller
    return DUGKS_OK;
  }
  {
    if (c.kernel == DUGKS_KERNEL_FUSED)
      return fail(ctx, DUGKS_EUNSUPPORTED, "the fused kernel does not cover this configuration");
    if (!c.fbx) {
      const size_t eN = (size_t)c.kp.eN;
      HIPCHK(c.alloc(&c.fbx, eN * Q * real_bytes(c.precision)));
      HIPCHK(hipMemsetAsync(c.fbx, 0, eN * Q * real_bytes(c.precision), c.stream));
    }
    HIPCHK(launch_step_twopass(c, fin, fout));
  }
  c.cur = 1 - c.cur;
  if (c.s.flow_rate != 0) HIPCHK(launch_flow_control(c, c.f[c.cur]));
  return DUGKS_OK;
}

}  // namespace

namespace dugks {

hipError_t build_params(Ctx& c) {
  const Setup& s = c.s;
  KParams& P = c.kp;
  std::memset(&P, 0, sizeof P);
  for (int a = 0; a < 3; a++) {
    P.n[a] = s.n[a];
    P.periodic[a] = s.periodic[a];
    P.e[a] = s.n[a] + 2;
    P.uniform[a] = s.uniform[a];
    P.du[a] = s.du[a];
    P.famb[a] = s.famb[a];
    if (s.periodic[a]) {
      P.qlo[a] = P.qhi[a] = 0;
    } else if (s.uniform[a]) {
      P.qlo[a] = P.qhi[a] = 2.0;  // 2 fb(1) - fb(2), src/dugks.cu:1247-1248
    } else {
      const auto& x = s.xc[a];
      const int n = s.n[a];
      P.qlo[a] = (x[2] - x[0]) / (x[2] - x[1]);  // dugks.cuf:368
      P.qhi[a] = (x[n + 1] - x[n - 1]) / (x[n] - x[n - 1]);  // dugks.cuf:372
    }
  }
  P.N = (long long)s.n[0] * s.n[1] * s.n[2];
  P.eN = (long long)P.e[0] * P.e[1] * P.e[2];
  P.dt = s.dt;
  P.hdt = s.hdt;
  P.bw0 = s.bw0;
  P.bw1 = s.bw1;
  P.bw2 = s.bw2;
  P.fw0 = s.fw0;
  P.fw1 = s.fw1;
  P.fw2 = s.fw2;
  P.recon = s.recon;
  P.flow_axis = s.stream_axis();
  P.rho0f = (float)s.rho0;
  hipError_t e;
  for (int a = 0; a < 3; a++) {
    std::vector<double> t = axis_table(s, a);
    if ((e = c.alloc((void**)&c.tab[a], t.size() * sizeof(double))) != hipSuccess) return e;
    if ((e = hipMemcpy(c.tab[a], t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
      return e;
    P.tab[a] = c.tab[a];
    std::vector<float> tf(t.begin(), t.end());
    if ((e = c.alloc((void**)&c.tabf[a], tf.size() * sizeof(float))) != hipSuccess) return e;
    if ((e = hipMemcpy(c.tabf[a], tf.data(), tf.size() * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess)
      return e;
    P.tabf[a] = c.tabf[a];
    if ((e = c.alloc((void**)&c.wd[a], s.d[a].size() * sizeof(double))) != hipSuccess) return e;
    if ((e = hipMemcpy(c.wd[a], s.d[a].data(), s.d[a].size() * sizeof(double), hipMemcpyHostToDevice)) !=
        hipSuccess)
      return e;
    P.wd[a] = c.wd[a];
  }
  if ((e = c.alloc((void**)&c.famb_dev, 8 * sizeof(double))) != hipSuccess) return e;
  double fa[8] = {s.famb[0], s.famb[1], s.famb[2], 0, 0, 0, 0, 0};
  if ((e = hipMemcpy(c.famb_dev, fa, sizeof fa, hipMemcpyHostToDevice)) != hipSuccess) return e;
  P.famb_dev = c.famb_dev;
  c.red_len = reduction_scratch_doubles() + 64;
  if ((e = c.alloc((void**)&c.red, c.red_len * sizeof(double))) != hipSuccess) return e;
  // per-block flow sums of the fused kernel: at most (nx/32)(ny/8) blocks
  // per z-chunk, z-chunks of at least 1 plane
  c.flow_partial_len = (size_t)(s.n[0] / 32 + 1) * (s.n[1] / 8 + 1) * (s.n[2] + 1);
  if ((e = c.alloc((void**)&c.flow_partial, c.flow_partial_len * sizeof(double))) != hipSuccess) return e;
  return hipSuccess;
}

}  // namespace dugks

namespace {
template <class F>
int host_guard(F&& f) {
  g_init_err.clear();
  try {
    return f();
  } catch (const ConfigError& e) {
    g_init_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_init_err = e.what();
    return DUGKS_ECONFIG;
  }
}
}  // namespace

extern "C" {

void dugks_default_opts(dugks_opts* o) {
  if (!o) return;
  o->precision = DUGKS_PREC_CONFIG;
  o->device = -1;
  o->stream = nullptr;
  o->kernel = DUGKS_KERNEL_AUTO;
  o->verbose = 0;
}

const char* dugks_init_error(void) { return g_init_err.c_str(); }

int dugks_init_from_string(const char* text, const char* base_dir, const dugks_opts* opts,
                           dugks_ctx** out) {
  g_init_err.clear();
  if (!text || !out) {
    g_init_err = "null argument";
    return DUGKS_EARG;
  }
  *out = nullptr;
  try {
    Config cfg = Config::from_text(text);
    Setup s = make_setup(cfg, base_dir ? base_dir : "");
    return create(s, opts, out);
  } catch (const ConfigError& e) {
    g_init_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_init_err = e.what();
    return DUGKS_ECONFIG;
  }
}

int dugks_init(const char* config_path, const dugks_opts* opts, dugks_ctx** out) {
  g_init_err.clear();
  if (!config_path || !out) {
    g_init_err = "null argument";
    return DUGKS_EARG;
  }
  *out = nullptr;
  try {
    Config cfg = Config::from_file(config_path);
    std::string p(config_path);
    size_t sl = p.find_last_of('/');
    std::string dir = sl == std::string::npos ? "" : p.substr(0, sl);
    Setup s = make_setup(cfg, dir);
    return create(s, opts, out);
  } catch (const ConfigError& e) {
    g_init_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_init_err = e.what();
    return DUGKS_ECONFIG;
  }
}

void dugks_free(dugks_ctx* ctx) {
  if (!ctx) return;
  Ctx& c = ctx->c;
  hipSetDevice(c.device);
  if (c.stream) hipStreamSynchronize(c.stream);
  for (int b = 0; b < 2; b++) hipFree(c.f[b]);
  hipFree(c.fbx);
  hipFree(c.adj_faces);
  hipFree(c.stats);
  for (int a = 0; a < 3; a++) {
    hipFree(c.tab[a]);
    hipFree(c.tabf[a]);
    hipFree(c.wd[a]);
  }
  hipFree(c.famb_dev);
  hipFree(c.fcell);
  hipFree(c.red);
  hipFree(c.flow_partial);
  release_sweep(c);
  for (int q = 0; q < 2; q++)
    if (c.tev[q]) hipEventDestroy(c.tev[q]);
  if (c.own_stream && c.stream) hipStreamDestroy(c.stream);
  delete ctx;
}

const char* dugks_last_error(const dugks_ctx* ctx) { return ctx ? ctx->c.err.c_str() : "null context"; }

int dugks_step(dugks_ctx* ctx, int istep0, int nsteps) {
  CHECK_CTX();
  if (nsteps < 0) return fail(ctx, DUGKS_EARG, "nsteps < 0");
  for (int s = 0; s < nsteps; s++) {
    int rc = step_once(ctx, istep0 + s);
    if (rc) return rc;
  }
  return DUGKS_OK;
}

int dugks_time_steps(dugks_ctx* ctx, int istep0, int nsteps, double out[5]) {
  CHECK_CTX();
  if (nsteps <= 0 || !out) return fail(ctx, DUGKS_EARG, "bad timing arguments");
  Ctx& c = ctx->c;
  hipEvent_t all[2];
  for (int q = 0; q < 2; q++) {
    if (!c.tev[q]) HIPCHK(hipEventCreate(&c.tev[q]));
    HIPCHK(hipEventCreate(&all[q]));
  }
  double sum = 0, mn = 1e300, mx = 0;
  HIPCHK(hipEventRecord(all[0], c.stream));
  c.timing = true;
  int rc = DUGKS_OK;
  for (int s = 0; s < nsteps && rc == DUGKS_OK; s++) {
    rc = step_once(ctx, istep0 + s);
    if (rc) break;
    hipError_t e = hipEventSynchronize(c.tev[1]);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, c.tev[0], c.tev[1]);
    if (e != hipSuccess) rc = fail(ctx, DUGKS_EHIP, std::string("timing: ") + hipGetErrorString(e));
    sum += ms;
    mn = std::min(mn, (double)ms);
    mx = std::max(mx, (double)ms);
  }
  c.timing = false;
  float total = 0;
  if (rc == DUGKS_OK) {
    HIPCHK(hipEventRecord(all[1], c.stream));
    HIPCHK(hipEventSynchronize(all[1]));
    HIPCHK(hipEventElapsedTime(&total, all[0], all[1]));
  }
  hipEventDestroy(all[0]);
  hipEventDestroy(all[1]);
  if (rc) return rc;
  out[0] = sum / nsteps;
  out[1] = mn;
  out[2] = mx;
  out[3] = total;
  out[4] = c.launches_per_step;
  return DUGKS_OK;
}

int dugks_sync(dugks_ctx* ctx) {
  CHECK_CTX();
  HIPCHK(hipStreamSynchronize(ctx->c.stream));
  return DUGKS_OK;
}

int dugks_get_dims(const dugks_ctx* ctx, int dims[3]) {
  if (!ctx || !dims) return DUGKS_EARG;
  for (int a = 0; a < 3; a++) dims[a] = ctx->c.s.n[a];
  return DUGKS_OK;
}

int dugks_get_device(const dugks_ctx* ctx, int* device) {
  if (!ctx || !device) return DUGKS_EARG;
  *device = ctx->c.device;
  return DUGKS_OK;
}

int dugks_get_params(const dugks_ctx* ctx, double out[16]) {
  if (!ctx || !out) return DUGKS_EARG;
  const Setup& s = ctx->c.s;
  double mins[3];
  for (int a = 0; a < 3; a++) {
    mins[a] = s.du[a];
    if (!s.uniform[a])
      for (int m = 1; m <= s.n[a]; m++) mins[a] = std::min(mins[a], s.d[a][m]);
  }
  double v[16] = {s.dt,  s.hdt, s.tau, s.visc, s.bw0,   s.bw1,   s.bw2,   s.fw0,
                  s.fw1, s.fw2, mins[0], mins[1], mins[2], s.rho0, (double)ctx->c.precision,
                  (double)ctx->c.kernel};
  std::memcpy(out, v, sizeof v);
  return DUGKS_OK;
}

int dugks_get_schedule(const dugks_ctx* ctx, int out[6]) {
  if (!ctx || !out) return DUGKS_EARG;
  const Setup& s = ctx->c.s;
  int v[6] = {s.istep0, s.nsteps, s.noutput2, s.noutput3, s.ndump, s.nprint};
  std::memcpy(out, v, sizeof v);
  return DUGKS_OK;
}

int dugks_get_statistics(dugks_ctx* ctx, double* profiles, double diag[12], int* rows) {
  CHECK_CTX();
  if (!profiles || !diag) return fail(ctx, DUGKS_EARG, "null output");
  Ctx& c = ctx->c;
  const int pa = stats_axis(c.s);
  const int np = c.s.n[pa];
  if (rows) *rows = np;
  const size_t need = stats_scratch_doubles(np);
  if (c.stats_len < need) {
    hipFree(c.stats);
    c.stats = nullptr;
    c.stats_len = 0;
    HIPCHK(c.alloc((void**)&c.stats, need * sizeof(double)));
    c.stats_len = need;
  }
  double* out = c.stats + need - ((size_t)np * 9 + 12);
  HIPCHK(launch_stats(c, c.f[c.cur], pa, c.stats, out));
  std::vector<double> sums((size_t)np * 9 + 12);
  HIPCHK(hipMemcpyAsync(sums.data(), out, sums.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  finish_statistics(c.s, sums.data(), sums.data() + (size_t)np * 9, profiles, diag);
  return DUGKS_OK;
}

int dugks_get_statistics_noalloc(dugks_ctx* ctx, double* profiles, double diag[12], int* rows) {
  CHECK_CTX();
  if (!profiles || !diag) return fail(ctx, DUGKS_EARG, "null output");
  Ctx& c = ctx->c;
  const int pa = stats_axis(c.s);
  const int np = c.s.n[pa];
  if (rows) *rows = np;
  if (np > kNoAllocStatRows)
    return fail(ctx, DUGKS_EUNSUPPORTED,
                "dugks_get_statistics_noalloc: " + std::to_string(np) + " wall-normal rows exceed its limit of " +
                    std::to_string(kNoAllocStatRows));
  if (!c.stats_noalloc) return fail(ctx, DUGKS_ESTATE, "no statistics scratch on this device");
  // the library's per-device scratch (see noalloc_stats_scratch): partial
  // sums first, the final sums behind them
  double* out = c.stats_noalloc + stats_scratch_doubles(np) - ((size_t)np * 9 + 12);
  HIPCHK(launch_stats(c, c.f[c.cur], pa, c.stats_noalloc, out));
  // the final sums go straight into the caller's buffers, finished in place
  HIPCHK(hipMemcpyAsync(profiles, out, (size_t)np * 9 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipMemcpyAsync(diag, out + (size_t)np * 9, 12 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  finish_statistics(c.s, profiles, diag, profiles, diag);
  return DUGKS_OK;
}

int dugks_spectral_statistics(dugks_ctx* ctx, double* spectrum, int nbins, double out[13]) {
  CHECK_CTX();
  if (!out || nbins < 0 || (nbins > 0 && !spectrum)) return fail(ctx, DUGKS_EARG, "bad spectral statistics arguments");
  Ctx& c = ctx->c;
  const Setup& s = c.s;
  for (int a = 0; a < 3; a++)
    if (!s.periodic[a] || !s.uniform[a])
      return fail(ctx, DUGKS_EUNSUPPORTED, "spectral statistics need periodic uniform axes");
  if (s.width[1] != s.width[0] || s.width[2] != s.width[0])
    return fail(ctx, DUGKS_EUNSUPPORTED, "spectral statistics need a cubic domain (equal widths)");
  const int nb = spectral_bins(c);
  std::vector<double> shells(2 * (size_t)nb);
  double sums[4];
  hipError_t e = spectral_statistics(c, shells.data(), sums);
  if (e == hipErrorInvalidValue) return fail(ctx, DUGKS_EUNSUPPORTED, "grid too large for the shell sums");
  if (e != hipSuccess) return fail(ctx, DUGKS_EHIP, std::string("spectral statistics: ") + hipGetErrorString(e));
  // stat_te, src/fortran/dhit/src/stat.f90:170-197
  double etotal = 0, eps = 0;
  for (int b = 0; b < nb; b++) {
    etotal += shells[b];
    eps += shells[nb + b];
  }
  const double nu = s.visc, L = s.width[0], dx = L / s.n[0];
  const double N3 = 3.0 * (double)s.ncell();
  const double uprm = std::sqrt(2.0 / 3.0 * etotal);
  const double tmse = std::sqrt(15 * nu * uprm * uprm / eps);
  const double bottom = sums[1] / N3;
  out[0] = sums[0];                                      // te
  out[1] = etotal;                                       // sum of E(k)
  out[2] = eps;                                          // 2 nu sum k^2 E(k)
  out[3] = std::pow(nu * nu * nu / eps, 0.25);           // Kolmogorov length
  out[4] = std::pow(nu * eps, 0.25);                     // Kolmogorov velocity
  out[5] = std::sqrt(nu / eps);                          // Kolmogorov time
  out[6] = uprm;                                         // u' = sqrt(2 etotal / 3)
  out[7] = tmse;                                         // Taylor microscale
  out[8] = uprm * tmse / nu;                             // Taylor-scale Reynolds number
  out[9] = uprm * uprm * uprm / eps;                     // large-eddy length
  out[10] = (L / 2 - 1.5 * dx) * 2 * M_PI / L * out[3];  // k_max eta
  out[11] = (sums[2] / N3) / std::pow(bottom, 1.5);      // derivative skewness
  out[12] = (sums[3] / N3) / (bottom * bottom);          // derivative flatness
  for (int b = 0; b < nbins; b++) spectrum[b] = b < nb ? shells[b] : 0.0;
  return DUGKS_OK;
}

int dugks_write_spectral_outputs(dugks_ctx* ctx, int step, const char* dir) {
  CHECK_CTX();
  const Setup& s = ctx->c.s;
  const int nspec = s.n[0] + 1;  // shells 0..nx, as the reference writes them
  std::vector<double> spec(nspec);
  double st[13];
  int rc = dugks_spectral_statistics(ctx, spec.data(), nspec, st);
  if (rc) return rc;
  try {
    write_spectral_outputs(step, s.dt, st, spec.data(), nspec, dir ? dir : "");
  } catch (const ConfigError& e) {
    return fail(ctx, e.code, e.what());
  }
  return DUGKS_OK;
}

int dugks_write_outputs(dugks_ctx* ctx, int step, const char* dir) {
  CHECK_CTX();
  const Setup& s = ctx->c.s;
  const size_t N = s.ncell();
  std::vector<double> prof((size_t)s.n[stats_axis(s)] * 9);
  double diag[12];
  int rc = dugks_get_statistics(ctx, prof.data(), diag, nullptr);
  if (rc) return rc;
  std::vector<double> m(4 * N);
  rc = dugks_get_macro(ctx, m.data(), m.data() + N, m.data() + 2 * N, m.data() + 3 * N);
  if (rc) return rc;
  try {
    write_outputs(s, step, m.data(), m.data() + N, m.data() + 2 * N, m.data() + 3 * N, prof.data(), diag,
                  dir ? dir : "");
  } catch (const ConfigError& e) {
    return fail(ctx, e.code, e.what());
  }
  return DUGKS_OK;
}

int dugks_get_spacing(const dugks_ctx* ctx, int axis, double* widths) {
  if (!ctx || !widths || axis < 0 || axis > 2) return DUGKS_EARG;
  const Setup& s = ctx->c.s;
  for (int m = 1; m <= s.n[axis]; m++) widths[m - 1] = s.d[axis][m];
  return DUGKS_OK;
}

int dugks_get_f_device(dugks_ctx* ctx, double* dev_f) {
  CHECK_CTX();
  if (!dev_f) return fail(ctx, DUGKS_EARG, "null device pointer");
  HIPCHK(launch_export(ctx->c, ctx->c.f[ctx->c.cur], dev_f));
  return DUGKS_OK;
}

int dugks_set_f_device(dugks_ctx* ctx, const double* dev_f) {
  CHECK_CTX();
  if (!dev_f) return fail(ctx, DUGKS_EARG, "null device pointer");
  Ctx& c = ctx->c;
  HIPCHK(launch_import(c, dev_f, c.f[c.cur]));
  // the flow-rate controller restarts from the new state (stream-ordered)
  HIPCHK(launch_fill(c, c.famb_dev + 3, 0.0, 2));
  return DUGKS_OK;
}

int dugks_get_f(dugks_ctx* ctx, double* host_f) {
  CHECK_CTX();
  if (!host_f) return fail(ctx, DUGKS_EARG, "null host pointer");
  Ctx& c = ctx->c;
  const size_t N = c.s.ncell();
  double* tmp = nullptr;
  HIPCHK(hipMalloc((void**)&tmp, N * Q * sizeof(double)));
  hipError_t e = launch_export(c, c.f[c.cur], tmp);
  std::vector<double> soa;
  if (e == hipSuccess) {
    soa.resize(N * Q);
    e = hipMemcpyAsync(soa.data(), tmp, N * Q * sizeof(double), hipMemcpyDeviceToHost, c.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  }
  hipFree(tmp);
  if (e != hipSuccess) return fail(ctx, DUGKS_EHIP, std::string("get_f: ") + hipGetErrorString(e));
  for (size_t cell = 0; cell < N; cell++)
    for (int ip = 0; ip < Q; ip++) host_f[cell * Q + ip] = soa[(size_t)ip * N + cell];
  return DUGKS_OK;
}

int dugks_set_f(dugks_ctx* ctx, const double* host_f) {
  CHECK_CTX();
  if (!host_f) return fail(ctx, DUGKS_EARG, "null host pointer");
  Ctx& c = ctx->c;
  const size_t N = c.s.ncell();
  std::vector<double> soa(N * Q);
  for (size_t cell = 0; cell < N; cell++)
    for (int ip = 0; ip < Q; ip++) {
      const double v = host_f[cell * Q + ip];
      if (!std::isfinite(v)) return fail(ctx, DUGKS_ENUMERIC, "non-finite value in f");
      soa[(size_t)ip * N + cell] = v;
    }
  double* tmp = nullptr;
  HIPCHK(hipMalloc((void**)&tmp, N * Q * sizeof(double)));
  hipError_t e = hipMemcpyAsync(tmp, soa.data(), N * Q * sizeof(double), hipMemcpyHostToDevice, c.stream);
  if (e == hipSuccess) e = launch_import(c, tmp, c.f[c.cur]);
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  hipFree(tmp);
  if (e != hipSuccess) return fail(ctx, DUGKS_EHIP, std::string("set_f: ") + hipGetErrorString(e));
  // the flow-rate controller restarts from the new state
  double z[2] = {0, 0};
  HIPCHK(hipMemcpy(c.famb_dev + 3, z, sizeof z, hipMemcpyHostToDevice));
  return DUGKS_OK;
}

int dugks_get_macro(dugks_ctx* ctx, double* rho, double* ux, double* uy, double* uz) {
  CHECK_CTX();
  if (!rho || !ux || !uy || !uz) return fail(ctx, DUGKS_EARG, "null host pointer");
  Ctx& c = ctx->c;
  const size_t N = c.s.ncell();
  double* tmp = nullptr;
  HIPCHK(hipMalloc((void**)&tmp, 4 * N * sizeof(double)));
  hipError_t e = launch_macro(c, c.f[c.cur], tmp, tmp + N, tmp + 2 * N, tmp + 3 * N);
  double* dst[4] = {rho, ux, uy, uz};
  for (int q = 0; q < 4 && e == hipSuccess; q++)
    e = hipMemcpyAsync(dst[q], tmp + q * N, N * sizeof(double), hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  hipFree(tmp);
  if (e != hipSuccess) return fail(ctx, DUGKS_EHIP, std::string("get_macro: ") + hipGetErrorString(e));
  return DUGKS_OK;
}

int dugks_set_ambient_force(dugks_ctx* ctx, const double force[3]) {
  CHECK_CTX();
  if (!force) return fail(ctx, DUGKS_EARG, "null force");
  Ctx& c = ctx->c;
  for (int a = 0; a < 3; a++) c.s.famb[a] = c.kp.famb[a] = force[a];
  double fa[5] = {force[0], force[1], force[2], 0, 0};
  HIPCHK(hipMemcpyAsync(c.famb_dev, fa, sizeof fa, hipMemcpyHostToDevice, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return DUGKS_OK;
}

int dugks_get_ambient_force(dugks_ctx* ctx, double force[3]) {
  CHECK_CTX();
  if (!force) return fail(ctx, DUGKS_EARG, "null force");
  HIPCHK(hipMemcpyAsync(force, ctx->c.famb_dev, 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->c.stream));
  HIPCHK(hipStreamSynchronize(ctx->c.stream));
  return DUGKS_OK;
}

int dugks_set_cell_force(dugks_ctx* ctx, const double* fx, const double* fy, const double* fz) {
  CHECK_CTX();
  Ctx& c = ctx->c;
  if (!fx && !fy && !fz) {
    c.kp.fcell = nullptr;
    return DUGKS_OK;
  }
  if (!fx || !fy || !fz) return fail(ctx, DUGKS_EARG, "give all three components or none");
  int rc = ensure_fcell(ctx);
  if (rc) return rc;
  const size_t N = c.s.ncell();
  const double* src[3] = {fx, fy, fz};
  for (int a = 0; a < 3; a++)
    HIPCHK(hipMemcpyAsync(c.fcell + a * N, src[a], N * sizeof(double), hipMemcpyHostToDevice, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return DUGKS_OK;
}

int dugks_set_cell_force_device(dugks_ctx* ctx, const double* dev_force) {
  CHECK_CTX();
  Ctx& c = ctx->c;
  if (!dev_force) {
    c.kp.fcell = nullptr;
    return DUGKS_OK;
  }
  int rc = ensure_fcell(ctx);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(c.fcell, dev_force, 3 * c.s.ncell() * sizeof(double), hipMemcpyDeviceToDevice,
                        c.stream));
  return DUGKS_OK;
}

int dugks_write_dump(dugks_ctx* ctx, int step, const char* dir, int rb) {
  CHECK_CTX();
  if (rb != 4 && rb != 8) return fail(ctx, DUGKS_EARG, "real_bytes must be 4 or 8");
  const size_t count = ctx->c.s.ncell() * Q;
  std::vector<double> f(count);
  int rc = dugks_get_f(ctx, f.data());
  if (rc) return rc;
  try {
    write_dump(join_path(dir ? dir : "", dump_name(step)), f.data(), count, rb);
  } catch (const ConfigError& e) {
    return fail(ctx, e.code, e.what());
  }
  return DUGKS_OK;
}

int dugks_read_dump(dugks_ctx* ctx, int step, const char* dir, int rb) {
  CHECK_CTX();
  if (rb != 4 && rb != 8) return fail(ctx, DUGKS_EARG, "real_bytes must be 4 or 8");
  const size_t count = ctx->c.s.ncell() * Q;
  std::vector<double> f(count);
  try {
    read_dump(join_path(dir ? dir : "", dump_name(step)), f.data(), count, rb);
  } catch (const ConfigError& e) {
    return fail(ctx, e.code, e.what());
  }
  return dugks_set_f(ctx, f.data());
}

int dugks_global_sums(dugks_ctx* ctx, double out[8]) {
  CHECK_CTX();
  if (!out) return fail(ctx, DUGKS_EARG, "null output");
  Ctx& c = ctx->c;
  double* dev = c.red + reduction_scratch_doubles();
  HIPCHK(launch_global_sums(c, c.f[c.cur], dev));
  HIPCHK(hipMemcpyAsync(out, dev, 8 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return DUGKS_OK;
}

int dugks_loss(dugks_ctx* ctx, int kind, double* value, double* dev_grad) {
  CHECK_CTX();
  if (!value || (kind != 0 && kind != 1)) return fail(ctx, DUGKS_EARG, "bad loss arguments");
  Ctx& c = ctx->c;
  if (kind == 0 && c.s.wall_axis() < 0) return fail(ctx, DUGKS_EUNSUPPORTED, "wall shear needs a wall axis");
  double* dev = c.red + reduction_scratch_doubles() + 8;
  HIPCHK(launch_loss(c, c.f[c.cur], kind, dev, dev_grad));
  HIPCHK(hipMemcpyAsync(value, dev, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return DUGKS_OK;
}

int dugks_adjoint_step(dugks_ctx* ctx, const double* lam_next, double* lam_prev, double* grad_force) {
  CHECK_CTX();
  if (!lam_next || !lam_prev) return fail(ctx, DUGKS_EARG, "null adjoint buffer");
  Ctx& c = ctx->c;
  const size_t bytes = (size_t)Q * c.s.ncell() * sizeof(double);
  const char *a = reinterpret_cast<const char*>(lam_next), *b = reinterpret_cast<const char*>(lam_prev);
  if (a < b + bytes && b < a + bytes) return fail(ctx, DUGKS_EARG, "lam_prev overlaps lam_next");
  HIPCHK(launch_adjoint(c, c.f[c.cur], lam_next, lam_prev, grad_force));
  return DUGKS_OK;
}

int dugks_adjoint_step_f32(dugks_ctx* ctx, const float* lam_next, float* lam_prev, double* grad_force) {
  CHECK_CTX();
  if (!lam_next || !lam_prev) return fail(ctx, DUGKS_EARG, "null adjoint buffer");
  Ctx& c = ctx->c;
  const size_t bytes = (size_t)Q * c.s.ncell() * sizeof(float);
  const char *a = reinterpret_cast<const char*>(lam_next), *b = reinterpret_cast<const char*>(lam_prev);
  if (a < b + bytes && b < a + bytes) return fail(ctx, DUGKS_EARG, "lam_prev overlaps lam_next");
  if (!adjoint_lam32_supported(c))
    return fail(ctx, DUGKS_EUNSUPPORTED, "FP32 lam needs FP32 arithmetic, nx % 32 == 0, ny % 8 == 0, "
                                         "periodic uniform z, the fused step and at most one wall axis");
  HIPCHK(launch_adjoint_lam32(c, c.f[c.cur], lam_next, lam_prev, grad_force));
  return DUGKS_OK;
}

namespace {

void release_sweep(Ctx& c) {
  SweepPool& w = c.sweep;
  for (void* p : w.cps) hipFree(p);
  for (void* p : w.seg) hipFree(p);
  hipFree(w.fT);
  hipFree(w.lam[0]);
  hipFree(w.lam[1]);
  hipFree(w.frec);
  hipFree(w.ctl);
  hipFree(w.fpart);
  c.bytes_allocated -= w.bytes;
  w = SweepPool();
}

// A device buffer of the pool of at least n doubles (reallocated when short).
hipError_t pool_doubles(Ctx& c, double*& p, size_t& len, size_t n) {
  if (p && len >= n) return hipSuccess;
  if (p) {
    hipFree(p);
    c.sweep.bytes -= (double)len * sizeof(double);
    c.bytes_allocated -= (double)len * sizeof(double);
    p = nullptr;
    len = 0;
  }
  hipError_t e = c.alloc((void**)&p, n * sizeof(double));
  if (e != hipSuccess) return e;
  len = n;
  c.sweep.bytes += (double)n * sizeof(double);
  return hipSuccess;
}

// While it lives, the step kernels leave the flow-rate controller alone;
// when it ends (also on an error return) the ambient force along the flow
// axis is set back to the one the recorded forward ended with, in stream
// order.
struct FrozenFlow {
  Ctx& c;
  int axis;
  FrozenFlow(Ctx& cx, int ax) : c(cx), axis(ax) { c.flow_frozen = true; }
  ~FrozenFlow() {
    c.flow_frozen = false;
    hipMemcpyAsync(c.famb_dev + axis, c.sweep.frec + c.sweep.T, sizeof(double), hipMemcpyDeviceToDevice,
                   c.stream);
  }
};

// Grow a list of state buffers to n entries (existing ones are kept).
hipError_t pool_states(Ctx& c, std::vector<void*>& v, size_t n, size_t bytes) {
  while (v.size() < n) {
    void* p = nullptr;
    hipError_t e = c.alloc(&p, bytes);
    if (e != hipSuccess) return e;
    c.sweep.bytes += (double)bytes;
    v.push_back(p);
  }
  return hipSuccess;
}

// lam in FP32 where the FP32-arithmetic adjoint supports it: its values are
// FP32 results either way, so this halves the lam traffic of the sweep
bool sweep_lam32(const Ctx& c) { return adjoint_lam32_supported(c); }

}  // namespace

int dugks_forward_record(dugks_ctx* ctx, int istep0, int T, int loss_kind, int interval, double* J) {
  CHECK_CTX();
  if (T <= 0 || !J || (loss_kind != 0 && loss_kind != 1)) return fail(ctx, DUGKS_EARG, "bad forward_record arguments");
  Ctx& c = ctx->c;
  if (loss_kind == 0 && c.s.wall_axis() < 0) return fail(ctx, DUGKS_EUNSUPPORTED, "wall shear needs a wall axis");
  if (c.s.perturb) return fail(ctx, DUGKS_EUNSUPPORTED, "the adjoint does not cover perturb (streak) forcing");
  if (c.precision == DUGKS_PREC_FP16N)
    return fail(ctx, DUGKS_EUNSUPPORTED, "the adjoint does not cover FP16 non-equilibrium storage");
  const bool flow = c.s.flow_rate != 0;
  if (flow && !adjoint_flow_supported(c))
    return fail(ctx, DUGKS_EUNSUPPORTED,
                "the adjoint of the flow-rate controller needs FP64, or FP32 arithmetic with lam in FP32, on the "
                "fused face pass and at most one wall axis");
  SweepPool& w = c.sweep;
  w.recorded = false;
  const size_t nf = c.s.ncell() * Q;
  const size_t sb = c.state_bytes;
  if (interval <= 0) interval = (int)std::ceil(std::sqrt((double)T));
  interval = std::min(interval, T);
  const int ncp = (T + interval - 1) / interval;  // checkpoints at steps 0, K, 2K, ...
  auto chk = [&](hipError_t e, const char* what) {
    if (e != hipSuccess)
      return fail(ctx, e == hipErrorOutOfMemory ? DUGKS_ENOMEM : DUGKS_EHIP,
                  std::string(what) + ": " + hipGetErrorString(e));
    return DUGKS_OK;
  };
  int rc;
  if ((rc = chk(pool_states(c, w.cps, (size_t)ncp, sb), "checkpoint alloc"))) return rc;
  if (!w.fT) {
    if ((rc = chk(c.alloc(&w.fT, sb), "final-state alloc"))) return rc;
    w.bytes += (double)sb;
  }
  // with the controller: the force before each step (frec[n] drives step n)
  // and the controller's previous-valid flag when the record begins
  const int ax = c.s.stream_axis();
  if (flow) {
    const int fmt = adjoint_record_format(c, c.precision != 0);
    if ((rc = chk(pool_doubles(c, w.frec, w.frec_len, (size_t)T + 1), "force record alloc"))) return rc;
    size_t ctl_len = w.ctl ? (size_t)kFlowCtlLen : 0;
    if ((rc = chk(pool_doubles(c, w.ctl, ctl_len, kFlowCtlLen), "controller scalars alloc"))) return rc;
    const long long nb = adjoint_faces_fused_blocks(c, fmt);
    if ((rc = chk(pool_doubles(c, w.fpart, w.fpart_len, (size_t)std::max(nb, 1LL)), "force partials alloc")))
      return rc;
    if ((rc = chk(hipMemcpyAsync(w.ctl + kFlowCtlFlag0, c.famb_dev + 4, sizeof(double), hipMemcpyDeviceToDevice,
                                 c.stream),
                  "flag copy")))
      return rc;
    if ((rc = chk(hipMemcpyAsync(w.frec, c.famb_dev + ax, sizeof(double), hipMemcpyDeviceToDevice, c.stream),
                  "force copy")))
      return rc;
  }
  // dval[0]: the loss summed over the forward sweep on the device, read once
  double* dval = c.red + reduction_scratch_doubles() + 9;
  if ((rc = chk(launch_fill(c, dval, 0.0, 1), "clear"))) return rc;
  for (int n = 0; n < T; n++) {
    if (n % interval == 0)
      if ((rc = chk(hipMemcpyAsync(w.cps[n / interval], c.f[c.cur], sb, hipMemcpyDeviceToDevice, c.stream),
                    "checkpoint copy")))
        return rc;
    if ((rc = step_once(ctx, istep0 + n))) return rc;
    // the loss of f^{n+1} reads the force the controller has just set
    if (flow)
      if ((rc = chk(hipMemcpyAsync(w.frec + n + 1, c.famb_dev + ax, sizeof(double), hipMemcpyDeviceToDevice,
                                   c.stream),
                    "force copy")))
        return rc;
    if ((rc = chk(launch_loss(c, c.f[c.cur], loss_kind, dval, nullptr, 0, true), "loss"))) return rc;
  }
  if ((rc = chk(hipMemcpyAsync(w.fT, c.f[c.cur], sb, hipMemcpyDeviceToDevice, c.stream), "final copy")))
    return rc;
  double Jsum = 0;
  if ((rc = chk(hipMemcpyAsync(&Jsum, dval, sizeof(double), hipMemcpyDeviceToHost, c.stream), "loss copy")))
    return rc;
  if ((rc = chk(hipStreamSynchronize(c.stream), "sync"))) return rc;
  *J = Jsum / T;
  w.istep0 = istep0;
  w.T = T;
  w.loss_kind = loss_kind;
  w.interval = interval;
  w.ncp = ncp;
  w.flow = flow;
  w.replayed = false;
  w.recorded = true;
  return DUGKS_OK;
}

int dugks_adjoint_replay(dugks_ctx* ctx, double* dev_grad_force, double* dev_grad_f0, double* peak_bytes) {
  CHECK_CTX();
  if (!dev_grad_force) return fail(ctx, DUGKS_EARG, "null gradient buffer");
  Ctx& c = ctx->c;
  SweepPool& w = c.sweep;
  if (!w.recorded) return fail(ctx, DUGKS_ESTATE, "no recorded forward sweep (dugks_forward_record)");
  const size_t N = c.s.ncell();
  const size_t nf = N * Q;
  const size_t sb = c.state_bytes;
  const bool lam32 = sweep_lam32(c);
  const size_t lb = lam32 ? sizeof(float) : sizeof(double);
  const int T = w.T, K = w.interval;
  auto chk = [&](hipError_t e, const char* what) {
    if (e != hipSuccess)
      return fail(ctx, e == hipErrorOutOfMemory ? DUGKS_ENOMEM : DUGKS_EHIP,
                  std::string(what) + ": " + hipGetErrorString(e));
    return DUGKS_OK;
  };
  int rc;
  // segment states f^{n0+1} .. f^{n0+K-1}; f^{n0} is the checkpoint itself
  if ((rc = chk(pool_states(c, w.seg, (size_t)(K - 1), sb), "segment alloc"))) return rc;
  if (w.lam_bytes < nf * lb) {
    for (int q = 0; q < 2; q++) {
      hipFree(w.lam[q]);
      w.lam[q] = nullptr;
    }
    w.bytes -= 2.0 * w.lam_bytes;
    c.bytes_allocated -= 2.0 * w.lam_bytes;
    w.lam_bytes = 0;
    for (int q = 0; q < 2; q++)
      if ((rc = chk(c.alloc(&w.lam[q], nf * lb), "lambda alloc"))) return rc;
    w.lam_bytes = nf * lb;
    w.bytes += 2.0 * w.lam_bytes;
  }
  void *lam = w.lam[0], *lam2 = w.lam[1];
  // dval[1]: loss values of the reverse sweep (discarded)
  double* dval = c.red + reduction_scratch_doubles() + 9;
  if (lam32) {
    if ((rc = chk(hipMemsetAsync(lam, 0, nf * lb, c.stream), "clear"))) return rc;
  } else {
    if ((rc = chk(launch_fill(c, static_cast<double*>(lam), 0.0, (long long)nf), "clear"))) return rc;
  }
  if ((rc = chk(launch_fill(c, dev_grad_force, 0.0, (long long)(3 * N)), "clear"))) return rc;
  const double invT = 1.0 / T;
  // With the flow-rate controller each step runs with its recorded force
  // and the controller is differentiated through its scalar recurrence
  // (k_flow_adjoint_scalar): before the adjoint of step n, lam = dJ/df^{n+1}
  // receives the bulk-velocity cotangent ubar_{n+1} dU/df^{n+1}; the adjoint
  // step also yields the face part of the force cotangent phi_n, and the
  // change of the summed cell force gradient gives the rest.
  const bool flow = w.flow;
  const bool fused = flow && adjoint_flow_fused(c);
  const int ax = c.s.stream_axis();
  std::unique_ptr<FrozenFlow> frozen;
  long long nparts = 0;
  auto set_force = [&](int n) {
    return hipMemcpyAsync(c.famb_dev + ax, w.frec + n, sizeof(double), hipMemcpyDeviceToDevice, c.stream);
  };
  w.replayed = false;
  if (flow) {
    frozen.reset(new FrozenFlow(c, ax));
    if ((rc = chk(launch_fill(c, w.ctl, 0.0, kFlowCtlFlag0), "clear"))) return rc;
    if ((rc = chk(launch_fill(c, w.ctl + kFlowCtlLoss, 0.0, kFlowCtlLen - kFlowCtlLoss), "clear"))) return rc;
  }
  // Recompute each segment from its checkpoint, then step back through it.
  for (int q = w.ncp - 1; q >= 0; q--) {
    const int n0 = q * K, n1 = std::min(T, n0 + K);
    auto state = [&](int n) -> const void* {  // f^n for n0 <= n <= n1
      if (n == n0) return w.cps[q];
      if (n < n1) return w.seg[n - n0 - 1];
      return n1 == T ? w.fT : w.cps[q + 1];
    };
    for (int n = n0 + 1; n < n1; n++) {
      bool handled = false;
      hipError_t e = flow ? set_force(n - 1) : hipSuccess;
      void* out = w.seg[n - n0 - 1];
      if (e == hipSuccess && c.kernel != DUGKS_KERNEL_TWOPASS) e = launch_step_fused(c, state(n - 1), out, &handled);
      if (e == hipSuccess && !handled) e = launch_step_twopass(c, state(n - 1), out);
      if ((rc = chk(e, "segment recompute"))) return rc;
    }
    for (int n = n1 - 1; n >= n0; n--) {
      // lam = dJ/df^{n+1}: add the loss term of step n+1 (and the
      // controller's), then lam2 = (df^{n+1}/df^n)^T lam about f^n
      if (flow) {
        hipError_t e = hipSuccess;
        if (n < T - 1) e = launch_axis_sum(c, dev_grad_force, ax, w.ctl + kFlowCtlSumNew);
        if (e == hipSuccess) e = set_force(n + 1);
        if (e == hipSuccess)
          e = lam32 ? launch_loss_lam32(c, state(n + 1), w.loss_kind, w.ctl + kFlowCtlLoss, static_cast<float*>(lam),
                                        invT, false, ax)
                    : launch_loss(c, state(n + 1), w.loss_kind, w.ctl + kFlowCtlLoss, static_cast<double*>(lam), invT,
                                  false, ax);
        if (e == hipSuccess)
          e = launch_flow_adjoint_scalar(c, w.ctl, w.fpart, n < T - 1 ? nparts : 0, n == 0, invT);
        // the bulk-velocity term: added to lam as the adjoint step reads it
        // where the step holds rho, j of f^{n+1} (from the step n+1 before
        // it, in the other of the two moment regions), else a pass of its own
        const bool ride = fused && n < T - 1;
        if (e == hipSuccess && !ride)
          e = lam32 ? launch_flow_cotangent_lam32(c, state(n + 1), w.ctl + kFlowCtlUbar, static_cast<float*>(lam))
                    : launch_flow_cotangent(c, state(n + 1), w.ctl + kFlowCtlUbar, static_cast<double*>(lam));
        if (e == hipSuccess) e = set_force(n);
        if (e == hipSuccess)
          e = lam32 ? launch_adjoint_lam32(c, state(n), static_cast<const float*>(lam), static_cast<float*>(lam2),
                                           dev_grad_force, w.fpart, (long long)w.fpart_len, &nparts,
                                           fused ? (n & 1) : 0, ride ? w.ctl + kFlowCtlUbar : nullptr)
                    : launch_adjoint(c, state(n), static_cast<const double*>(lam), static_cast<double*>(lam2),
                                     dev_grad_force, w.fpart, (long long)w.fpart_len, &nparts,
                                     fused ? (n & 1) : 0, ride ? w.ctl + kFlowCtlUbar : nullptr);
        if ((rc = chk(e, "controller adjoint"))) return rc;
      } else if (lam32) {
        float *l = static_cast<float*>(lam), *l2 = static_cast<float*>(lam2);
        if ((rc = chk(launch_loss_lam32(c, state(n + 1), w.loss_kind, dval + 1, l, invT), "loss grad"))) return rc;
        if ((rc = chk(launch_adjoint_lam32(c, state(n), l, l2, dev_grad_force), "adjoint"))) return rc;
      } else {
        double *l = static_cast<double*>(lam), *l2 = static_cast<double*>(lam2);
        if ((rc = chk(launch_loss(c, state(n + 1), w.loss_kind, dval + 1, l, invT), "loss grad"))) return rc;
        if ((rc = chk(launch_adjoint(c, state(n), l, l2, dev_grad_force), "adjoint"))) return rc;
      }
      std::swap(lam, lam2);
    }
  }
  if (flow) {
    // nu_0 = phi_0 + nu_1 = dJ/dF^0; the force the forward ended with is set
    // back as `frozen` ends, ahead of the copies below in stream order
    hipError_t e = launch_axis_sum(c, dev_grad_force, ax, w.ctl + kFlowCtlSumNew);
    if (e == hipSuccess) e = launch_flow_adjoint_scalar(c, w.ctl, w.fpart, nparts, true, invT);
    if ((rc = chk(e, "controller adjoint"))) return rc;
    frozen.reset();
  }
  if (dev_grad_f0)
    if ((rc = chk(lam32 ? launch_widen(c, static_cast<const float*>(lam), dev_grad_f0, (long long)nf)
                        : hipMemcpyAsync(dev_grad_f0, lam, nf * sizeof(double), hipMemcpyDeviceToDevice, c.stream),
                  "grad copy")))
      return rc;
  if ((rc = chk(hipStreamSynchronize(c.stream), "sync"))) return rc;
  w.replayed = true;
  if (peak_bytes) {
    // device memory in use at the end of the sweep, when every buffer it
    // needs (pool, adjoint face records, state) is allocated
    size_t free_b = 0, total_b = 0;
    if ((rc = chk(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo"))) return rc;
    peak_bytes[0] = (double)(total_b - free_b);
  }
  return DUGKS_OK;
}

int dugks_sweep_force_gradient(dugks_ctx* ctx, double* dJdF0) {
  CHECK_CTX();
  if (!dJdF0) return fail(ctx, DUGKS_EARG, "null output");
  Ctx& c = ctx->c;
  const SweepPool& w = c.sweep;
  if (!w.recorded || !w.replayed || !w.flow)
    return fail(ctx, DUGKS_ESTATE, "no replayed sweep with the flow-rate controller");
  HIPCHK(hipMemcpyAsync(dJdF0, w.ctl + kFlowCtlNuNext, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return DUGKS_OK;
}

int dugks_release_sweep(dugks_ctx* ctx) {
  CHECK_CTX();
  release_sweep(ctx->c);
  return DUGKS_OK;
}

int dugks_sweep_bytes(dugks_ctx* ctx, double* bytes) {
  if (!ctx || !bytes) return DUGKS_EARG;
  *bytes = ctx->c.sweep.bytes;
  return DUGKS_OK;
}

int dugks_forward_adjoint(dugks_ctx* ctx, int istep0, int T, int loss_kind, int interval, double* J,
                          double* dev_grad_force, double* dev_grad_f0, double* peak_bytes) {
  CHECK_CTX();
  if (!dev_grad_force) return fail(ctx, DUGKS_EARG, "bad forward_adjoint arguments");
  int rc = dugks_forward_record(ctx, istep0, T, loss_kind, interval, J);
  if (rc) return rc;
  return dugks_adjoint_replay(ctx, dev_grad_force, dev_grad_f0, peak_bytes);
}

// ---- host-only helpers ----


int dugks_config_derive(const char* text, const char* base_dir, int dims[3], double params[16]) {
  if (!text || !dims || !params) return DUGKS_EARG;
  return host_guard([&]() {
    Setup s = make_setup(Config::from_text(text), base_dir ? base_dir : "");
    double mins[3];
    for (int a = 0; a < 3; a++) {
      dims[a] = s.n[a];
      mins[a] = s.du[a];
      if (!s.uniform[a])
        for (int m = 1; m <= s.n[a]; m++) mins[a] = std::min(mins[a], s.d[a][m]);
    }
    double v[16] = {s.dt,  s.hdt, s.tau,   s.visc,  s.bw0,   s.bw1,  s.bw2,          s.fw0,
                    s.fw1, s.fw2, mins[0], mins[1], mins[2], s.rho0, (double)s.precision, (double)s.kernel};
    std::memcpy(params, v, sizeof v);
    return DUGKS_OK;
  });
}

int dugks_config_spacing(const char* text, const char* base_dir, int axis, double* widths) {
  if (!text || !widths || axis < 0 || axis > 2) return DUGKS_EARG;
  return host_guard([&]() {
    Setup s = make_setup(Config::from_text(text), base_dir ? base_dir : "");
    for (int m = 1; m <= s.n[axis]; m++) widths[m - 1] = s.d[axis][m];
    return DUGKS_OK;
  });
}

int dugks_config_initial_state(const char* text, const char* base_dir, double* host_f) {
  if (!text || !host_f) return DUGKS_EARG;
  return host_guard([&]() {
    Setup s = make_setup(Config::from_text(text), base_dir ? base_dir : "");
    if (s.restart) throw ConfigError(DUGKS_ESTATE, "restart=1: the initial state comes from a dump");
    std::vector<double> f;
    initial_state(s, f);
    std::memcpy(host_f, f.data(), f.size() * sizeof(double));
    return DUGKS_OK;
  });
}

int dugks_mesh_tanh(int n, double alpha, double width, double* faces) {
  if (n < 1 || !(alpha > 0 && alpha < 1) || !(width > 0) || !faces) return DUGKS_EARG;
  std::vector<double> z = tanh_faces(n, alpha, width);
  std::memcpy(faces, z.data(), z.size() * sizeof(double));
  return DUGKS_OK;
}

int dugks_dump_write(const char* path, const double* f, long long count, int rb) {
  if (!path || !f || count < 0) return DUGKS_EARG;
  return host_guard([&]() {
    write_dump(path, f, (size_t)count, rb);
    return DUGKS_OK;
  });
}

int dugks_velmean(const char* profiles_path, int n, int ignore_before, int mean_col, int stress_col,
                  const char* out_path) {
  if (!profiles_path || !out_path || n <= 0 || mean_col < 0 || mean_col > 2 || stress_col < 3 || stress_col > 5)
    return DUGKS_EARG;
  return host_guard([&]() {
    velmean(profiles_path, n, ignore_before, mean_col, stress_col, out_path);
    return DUGKS_OK;
  });
}

int dugks_dump_read(const char* path, double* f, long long count, int rb) {
  if (!path || !f || count < 0) return DUGKS_EARG;
  return host_guard([&]() {
    read_dump(path, f, (size_t)count, rb);
    return DUGKS_OK;
  });
}

}  // extern "C"
