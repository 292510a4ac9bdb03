// Host-side setup: config.txt parsing, derived numerics, mesh, initial state.
//
// Reference behaviour restated:
//   Config           src/config.h:17-74
//   init_numerics    src/numerics.cu:62-106 (dx = width/n, dt = CFL dx/sqrt(6RT))
//   periods          src/parallel.cu:95-97
//   coefficients     src/dugks.cu:101-125
//   initial states   src/dugks.cu:221-281, feq :304-309
//   tanh mesh        src/fortran/channel/run/mesh.py:21-25, centres and ghost
//                    widths src/fortran/channel/src/dugks.cuf:114-121,
//                    dt from the first wall cell :138
#include "setup.h"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <sstream>

#include "../../include/dugks.h"
#include "lattice.h"

namespace dugks {

namespace {

std::string trim(const std::string& s) {
  size_t b = s.find_first_not_of(" \t\r\n");
  if (b == std::string::npos) return "";
  size_t e = s.find_last_not_of(" \t\r\n");
  return s.substr(b, e - b + 1);
}

// splitmix64: small deterministic generator so seeded initial states do not
// depend on the C++ library's distributions.
struct SplitMix {
  uint64_t s;
  explicit SplitMix(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  int range(int lo, int hi) { return lo + (int)(uniform() * (hi - lo + 1)); }
};

const double kPi = 3.141592653589793;  // src/constants.h:6
const double kPi2 = kPi * 2;

}  // namespace

// ---------------------------------------------------------------- Config --

Config Config::from_text(const std::string& text) {
  Config c;
  std::istringstream in(text);
  std::string line;
  while (std::getline(in, line, '\n')) {
    size_t p = line.find('=');
    if (p == std::string::npos) continue;
    size_t q = line.find('=', p + 1);
    std::string key = trim(line.substr(0, p));
    std::string val = trim(q == std::string::npos ? line.substr(p + 1) : line.substr(p + 1, q - p - 1));
    if (key.empty()) continue;
    c.kv_.insert({key, val});  // first occurrence wins (std::map::insert)
  }
  if (c.has("parse_real")) {
    std::string m = c.str("parse_real");
    if (m == "double")
      c.real_as_float_ = false;
    else if (m != "float")
      throw ConfigError(DUGKS_ECONFIG, "parse_real must be float or double, got '" + m + "'");
  }
  return c;
}

Config Config::from_file(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw ConfigError(DUGKS_EIO, "cannot open config file '" + path + "'");
  std::stringstream ss;
  ss << f.rdbuf();
  return from_text(ss.str());
}

std::string Config::str(const std::string& k) const {
  auto it = kv_.find(k);
  if (it == kv_.end()) throw ConfigError(DUGKS_ECONFIG, "config key not found: " + k);
  return it->second;
}

std::string Config::str(const std::string& k, const std::string& dflt) const {
  return has(k) ? str(k) : dflt;
}

int Config::integer(const std::string& k) const {
  std::string v = str(k);
  errno = 0;
  char* end = nullptr;
  long x = std::strtol(v.c_str(), &end, 10);
  if (end == v.c_str() || errno == ERANGE)
    throw ConfigError(DUGKS_ECONFIG, "config key " + k + ": not an integer: '" + v + "'");
  return (int)x;
}

int Config::integer(const std::string& k, int dflt) const { return has(k) ? integer(k) : dflt; }

double Config::real(const std::string& k) const {
  std::string v = str(k);
  char* end = nullptr;
  double x;
  if (real_as_float_)
    x = (double)std::strtof(v.c_str(), &end);
  else
    x = std::strtod(v.c_str(), &end);
  if (end == v.c_str())
    throw ConfigError(DUGKS_ECONFIG, "config key " + k + ": not a number: '" + v + "'");
  return x;
}

double Config::real(const std::string& k, double dflt) const { return has(k) ? real(k) : dflt; }

// ----------------------------------------------------------------- Setup --

int Setup::wall_axis() const {
  for (int a = 0; a < 3; a++)
    if (!periodic[a]) return a;
  return -1;
}

int Setup::stream_axis() const {
  if (flow_axis >= 0) return flow_axis;
  int best = -1;
  double m = 0;
  for (int a = 0; a < 3; a++)
    if (std::fabs(famb[a]) > m) {
      m = std::fabs(famb[a]);
      best = a;
    }
  if (best >= 0) return best;
  for (int a = 0; a < 3; a++)
    if (periodic[a]) return a;
  return 0;
}

std::vector<double> tanh_faces(int n, double alpha, double w) {
  std::vector<double> z(n + 1);
  for (int k = 0; k <= n; k++) {
    double eta = -1. + 2. * ((k + 0) - 0.0) / n;
    z[k] = w / (2. * alpha) * std::tanh(eta * 0.5 * (std::log(1. + alpha) - std::log(1. - alpha))) +
           0.5 * w;
  }
  return z;
}

namespace {

std::vector<double> read_faces(const std::string& path, int n, double w) {
  std::ifstream f(path);
  if (!f) throw ConfigError(DUGKS_EIO, "cannot open mesh file '" + path + "'");
  std::vector<double> z;
  double x;
  while ((int)z.size() < n + 1 && (f >> x)) z.push_back(x);
  if ((int)z.size() != n + 1)
    throw ConfigError(DUGKS_ECONFIG, "mesh file '" + path + "' holds fewer than n+1 face coordinates");
  const double z0 = z[0], span = z[n] - z[0];
  if (!(span > 0)) throw ConfigError(DUGKS_ECONFIG, "mesh file '" + path + "' is not increasing");
  for (auto& v : z) v = (v - z0) * (w / span);
  return z;
}

void build_axis(Setup& s, int a, const std::vector<double>* faces) {
  const int n = s.n[a];
  s.d[a].assign(n + 2, s.du[a]);
  s.xc[a].assign(n + 2, 0.0);
  if (faces) {
    s.uniform[a] = 0;
    for (int m = 1; m <= n; m++) {
      s.d[a][m] = (*faces)[m] - (*faces)[m - 1];
      s.xc[a][m] = 0.5 * ((*faces)[m] + (*faces)[m - 1]);
      if (!(s.d[a][m] > 0))
        throw ConfigError(DUGKS_ECONFIG, "mesh faces must be strictly increasing");
    }
  } else {
    s.uniform[a] = 1;
    for (int m = 1; m <= n; m++) s.xc[a][m] = (m - 0.5) * s.du[a];
  }
  if (s.periodic[a]) {
    s.d[a][0] = s.d[a][n];
    s.d[a][n + 1] = s.d[a][1];
    s.xc[a][0] = s.xc[a][1] - 0.5 * (s.d[a][0] + s.d[a][1]);
    s.xc[a][n + 1] = s.xc[a][n] + 0.5 * (s.d[a][n] + s.d[a][n + 1]);
  } else {
    s.d[a][0] = s.d[a][1];
    s.d[a][n + 1] = s.d[a][n];
    s.xc[a][0] = -s.xc[a][1];
    s.xc[a][n + 1] = 2 * s.width[a] - s.xc[a][n];
  }
}

}  // namespace

Setup make_setup(const Config& cfg, const std::string& base_dir) {
  Setup s;
  s.base_dir = base_dir;
  s.case_name = cfg.str("case");
  s.CFL = cfg.real("CFL");
  const char* wk[3] = {"xwidth", "ywidth", "zwidth"};
  const char* nk[3] = {"nx", "ny", "nz"};
  const char* pk[3] = {"periods_x", "periods_y", "periods_z"};
  const char* mk[3] = {"mesh_x", "mesh_y", "mesh_z"};
  const char* fk[3] = {"ambient_fx", "ambient_fy", "ambient_fz"};
  for (int a = 0; a < 3; a++) {
    s.width[a] = cfg.real(wk[a]);
    s.n[a] = cfg.integer(nk[a]);
    s.periodic[a] = cfg.integer(pk[a], 1) != 0;
    s.famb[a] = cfg.real(fk[a], 0.0);
    if (s.n[a] < 1) throw ConfigError(DUGKS_ECONFIG, std::string(nk[a]) + " must be >= 1");
    if (!s.periodic[a] && s.n[a] < 2)
      throw ConfigError(DUGKS_ECONFIG, std::string("a wall axis needs ") + nk[a] + " >= 2");
    if (!(s.width[a] > 0)) throw ConfigError(DUGKS_ECONFIG, std::string(wk[a]) + " must be > 0");
    s.du[a] = s.width[a] / s.n[a];
  }
  s.istep0 = cfg.integer("istep0", 1);
  s.nsteps = cfg.integer("nsteps", 0);
  s.noutput2 = cfg.integer("noutput2", 1 << 30);
  s.noutput3 = cfg.integer("noutput3", 1 << 30);
  s.ndump = cfg.integer("ndump", 1 << 30);
  s.restart = cfg.integer("restart", 0);
  s.nrestart = cfg.integer("nrestart", 0);
  s.perturb = cfg.integer("perturb", 0);
  s.nprint = cfg.integer("nprint", 1);
  if (s.noutput2 <= 0 || s.ndump <= 0 || s.nprint <= 0)
    throw ConfigError(DUGKS_ECONFIG, "noutput2, ndump and nprint must be positive");
  s.visc = cfg.real("visc");
  if (!(s.visc > 0)) throw ConfigError(DUGKS_ECONFIG, "visc must be > 0");

  std::string prec = cfg.str("precision", "fp64");
  if (prec == "fp64" || prec == "double")
    s.precision = DUGKS_PREC_FP64;
  else if (prec == "fp32" || prec == "float")
    s.precision = DUGKS_PREC_FP32;
  else if (prec == "fp16" || prec == "fp16s")
    s.precision = DUGKS_PREC_FP16S;
  else if (prec == "fp16n")
    s.precision = DUGKS_PREC_FP16N;
  else
    throw ConfigError(DUGKS_ECONFIG, "precision must be fp64, fp32, fp16 or fp16n");
  std::string rc = cfg.str("recon", "tri");
  if (rc == "tri")
    s.recon = 0;
  else if (rc == "grad")
    s.recon = 1;
  else
    throw ConfigError(DUGKS_ECONFIG, "recon must be tri or grad");
  std::string kn = cfg.str("kernel", "auto");
  if (kn == "auto")
    s.kernel = DUGKS_KERNEL_AUTO;
  else if (kn == "twopass")
    s.kernel = DUGKS_KERNEL_TWOPASS;
  else if (kn == "fused")
    s.kernel = DUGKS_KERNEL_FUSED;
  else
    throw ConfigError(DUGKS_ECONFIG, "kernel must be auto, twopass or fused");

  s.u0 = cfg.real("u0", 0.1);
  s.rho0 = cfg.real("rho0", 1.0);
  if (cfg.str("flow_rate", "") == "auto") {
    s.flow_auto = true;
    s.flow_rate = 1.0;  // placeholder, replaced by the initial bulk velocity
  } else {
    s.flow_rate = cfg.real("flow_rate", 0.0);
  }
  s.flow_axis = cfg.integer("flow_axis", -1);
  s.flow_gain = cfg.real("flow_gain", 1.0);
  if (s.flow_axis < -1 || s.flow_axis > 2) throw ConfigError(DUGKS_ECONFIG, "flow_axis must be 0, 1 or 2");
  s.retau = cfg.real("retau", 180.0);
  {
    std::string mp = cfg.str("mean_profile", "loglaw");
    if (mp == "laminar")
      s.laminar_profile = 1;
    else if (mp != "loglaw")
      throw ConfigError(DUGKS_ECONFIG, "mean_profile must be loglaw or laminar");
  }
  s.pert_amp = cfg.real("pert_amp", 0.1);
  s.kmin = cfg.integer("kmin", 1);
  s.kmax = cfg.integer("kmax", 8);
  s.nmodes = cfg.integer("nmodes", 24);
  s.seed = (unsigned long long)cfg.integer("seed", 12345);

  // mesh
  for (int a = 0; a < 3; a++) {
    std::string m = cfg.str(mk[a], "uniform");
    if (m == "uniform") {
      build_axis(s, a, nullptr);
      continue;
    }
    if (s.periodic[a])
      throw ConfigError(DUGKS_EUNSUPPORTED, std::string(mk[a]) + ": a stretched axis must be a wall axis");
    std::vector<double> faces;
    if (m.rfind("tanh:", 0) == 0) {
      double alpha = std::strtod(m.c_str() + 5, nullptr);
      if (!(alpha > 0 && alpha < 1)) throw ConfigError(DUGKS_ECONFIG, "tanh mesh needs 0 < alpha < 1");
      faces = tanh_faces(s.n[a], alpha, s.width[a]);
    } else {
      std::string path = (m.size() && m[0] == '/') ? m : (base_dir.empty() ? m : base_dir + "/" + m);
      faces = read_faces(path, s.n[a], s.width[a]);
    }
    build_axis(s, a, &faces);
  }
  bool any_stretched = !(s.uniform[0] && s.uniform[1] && s.uniform[2]);
  if (s.recon == 1 && any_stretched)
    throw ConfigError(DUGKS_EUNSUPPORTED, "recon=grad needs uniform axes");

  // time step: numerics.cu:105 uses dx only (quirk Q2); the stretched
  // channel code uses its first wall cell (dugks.cuf:138).
  std::string rule = cfg.str("dt_rule", "auto");
  if (rule == "auto") rule = any_stretched ? "min" : "dx";
  double h = 0;
  if (rule == "dx") {
    h = s.du[0];
  } else if (rule == "min") {
    h = 1e300;
    for (int a = 0; a < 3; a++)
      for (int m = 1; m <= s.n[a]; m++) h = std::min(h, s.d[a][m]);
  } else {
    throw ConfigError(DUGKS_ECONFIG, "dt_rule must be auto, dx or min");
  }
  s.dt = s.CFL * h / std::sqrt(6 * RT);
  if (cfg.has("dt")) s.dt = cfg.real("dt");
  if (!(s.dt > 0)) throw ConfigError(DUGKS_ECONFIG, "dt must be > 0");

  // coefficients (src/dugks.cu:101-125)
  const double tau = s.visc / RT, dt = s.dt, hdt = dt / 2.;
  s.tau = tau;
  s.hdt = hdt;
  s.bw0 = (3. * tau * hdt) / (2. * tau + dt);
  s.bw1 = (2. * tau - hdt) / (2. * tau + dt);
  s.bw2 = 3. * hdt / (2. * tau + dt);
  s.fw0 = tau * hdt / (2. * tau + hdt);
  s.fw1 = 2. * tau / (2. * tau + hdt);
  s.fw2 = hdt / (2. * tau + hdt);

  // initial state kind
  if (s.case_name == "tg_eq")
    s.ic = IC_TG_EQ;
  else if (s.case_name == "channel")
    s.ic = IC_CHANNEL;
  else if (s.case_name == "general")
    s.ic = IC_GENERAL;
  else if (s.case_name == "rest")
    s.ic = IC_REST;
  else if (s.case_name == "shear_wave")
    s.ic = IC_SHEAR_WAVE;
  else if (s.case_name == "turb_channel")
    s.ic = IC_TURB_CHANNEL;
  else
    throw ConfigError(DUGKS_ECONFIG, "unknown case '" + s.case_name + "'");
  if (s.ic == IC_GENERAL && !s.restart)
    throw ConfigError(DUGKS_ECONFIG, "case=general needs restart=1");
  if (s.ic == IC_TURB_CHANNEL && s.wall_axis() < 0)
    throw ConfigError(DUGKS_ECONFIG, "case=turb_channel needs a wall axis");
  if (s.flow_rate != 0 && s.wall_axis() < 0)
    throw ConfigError(DUGKS_ECONFIG, "flow_rate forcing needs a wall axis");
  return s;
}

double feq_host(int ip, double r, double u, double v, double w) {
  double eu = (CX[ip] * u + CY[ip] * v + CZ[ip] * w) / RT;
  double uv = (u * u + v * v + w * w) / RT;
  return wgt(ip) * r * (1.0 + eu + 0.5 * (eu * eu - uv));
}

namespace {

// Mean profile of the synthetic channel state: linear sublayer joined to the
// log law U+ = ln(y+)/0.40 + 5.6 (examples/channel/log_plot.plt:18-19) where
// the two cross.
double uplus(double yp) {
  static double yc = -1;
  if (yc < 0) {
    double lo = 5, hi = 30;
    for (int it = 0; it < 200; it++) {
      double mid = 0.5 * (lo + hi);
      if (mid - (std::log(mid) / 0.40 + 5.6) < 0) lo = mid; else hi = mid;
    }
    yc = 0.5 * (lo + hi);
  }
  return yp <= yc ? yp : std::log(yp) / 0.40 + 5.6;
}

void turb_channel_velocity(const Setup& s, std::vector<double>& U) {
  const int nx = s.n[0], ny = s.n[1], nz = s.n[2];
  const int aw = s.wall_axis();
  int as = s.stream_axis();
  if (as == aw) as = (aw + 1) % 3;
  const int ap = 3 - aw - as;
  const double H = s.width[aw] / 2;
  const double utau = s.retau * s.visc / H;
  const double Uc = s.laminar_profile ? s.u0 : utau * uplus(s.retau);
  const int ns = s.n[as], np = s.n[ap];
  // Perturbation: u' = curl(psi), psi_c = W(eta) P_c(s, p),
  // W = (1 - eta^2)^2, P_c = sum_m a_mc sin(2 pi (kx_m s/Ls + kz_m p/Lp) + phi_mc).
  SplitMix rng(s.seed);
  struct Mode { int kx, kz; double a[3], ph[3]; };
  std::vector<Mode> modes(std::max(0, s.nmodes));
  for (auto& m : modes) {
    m.kx = rng.range(s.kmin, s.kmax);
    m.kz = rng.range(s.kmin, s.kmax);
    for (int c = 0; c < 3; c++) {
      m.a[c] = 2 * rng.uniform() - 1;
      m.ph[c] = kPi2 * rng.uniform();
    }
  }
  // P_c, dP_c/ds, dP_c/dp on the (s, p) plane
  std::vector<double> P(3 * ns * np, 0.0), Ps(3 * ns * np, 0.0), Pp(3 * ns * np, 0.0);
  for (const auto& m : modes) {
    const double ks = kPi2 * m.kx / s.width[as], kp = kPi2 * m.kz / s.width[ap];
    for (int c = 0; c < 3; c++)
      for (int ip = 0; ip < np; ip++) {
        const double xp = s.xc[ap][ip + 1];
        for (int is = 0; is < ns; is++) {
          const double xs = s.xc[as][is + 1];
          const double arg = ks * xs + kp * xp + m.ph[c];
          const double sn = std::sin(arg), cs = std::cos(arg);
          const size_t q = ((size_t)c * np + ip) * ns + is;
          P[q] += m.a[c] * sn;
          Ps[q] += m.a[c] * ks * cs;
          Pp[q] += m.a[c] * kp * cs;
        }
      }
  }
  // local frame (s, w, p) components -> global
  U.assign((size_t)3 * nx * ny * nz, 0.0);
  double umax = 0;
  std::vector<double> up((size_t)3 * nx * ny * nz, 0.0);
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        const int idx[3] = {i, j, k};
        const double eta = (s.xc[aw][idx[aw] + 1] - H) / H;
        const double W = (1 - eta * eta) * (1 - eta * eta);
        const double Wd = -4 * eta * (1 - eta * eta) / H;  // dW/dw
        const int is = idx[as], ip = idx[ap];
        auto q = [&](int c) { return ((size_t)c * np + ip) * ns + is; };
        // components of psi in the local frame: c = 0 (s), 1 (w), 2 (p)
        const double dps_dw = Wd * P[q(0)], dpp_dw = Wd * P[q(2)];
        const double dpw_dp = W * Pp[q(1)], dps_dp = W * Pp[q(0)];
        const double dpp_ds = W * Ps[q(2)], dpw_ds = W * Ps[q(1)];
        const double us = dpp_dw - dpw_dp;
        const double uw = dps_dp - dpp_ds;
        const double uq = dpw_ds - dps_dw;
        const size_t cell = ((size_t)k * ny + j) * nx + i;
        up[cell * 3 + as] = us;
        up[cell * 3 + aw] = uw;
        up[cell * 3 + ap] = uq;
        umax = std::max(umax, std::sqrt(us * us + uw * uw + uq * uq));
        const double yw = H * (1 - std::fabs(eta));
        U[cell * 3 + as] = s.laminar_profile ? Uc * (1 - eta * eta) : utau * uplus(yw * utau / s.visc);
      }
  const double scale = umax > 0 ? s.pert_amp * Uc / umax : 0.0;
  for (size_t q = 0; q < U.size(); q++) U[q] += scale * up[q];
}

}  // namespace

void initial_state(const Setup& s, std::vector<double>& f) {
  const int nx = s.n[0], ny = s.n[1], nz = s.n[2];
  const size_t N = s.ncell();
  f.assign(N * Q, 0.0);
  if (s.ic == IC_GENERAL) return;
  std::vector<double> U;  // per-cell velocity for the channel state
  if (s.ic == IC_TURB_CHANNEL) turb_channel_velocity(s, U);
  const double Lx = s.width[0], Ly = s.width[1], Lz = s.width[2];
  const double vel_scale = s.u0;
  for (int k = 0; k < nz; k++) {
    const double z = s.xc[2][k + 1];
    for (int j = 0; j < ny; j++) {
      const double y = s.xc[1][j + 1];
      for (int i = 0; i < nx; i++) {
        const double x = s.xc[0][i + 1];
        const size_t cell = ((size_t)k * ny + j) * nx + i;
        double r = 1.0, u = 0, v = 0, w = 0;
        switch (s.ic) {
          case IC_TG_EQ:  // src/dugks.cu:222-240
            r = 1.0 + 1.0 * vel_scale * vel_scale / 16 *
                          (std::cos(2 * kPi2 / Lx * x) + std::cos(2 * kPi2 / Ly * y)) *
                          (std::cos(2 * kPi2 / Lz * z) + 2) / RT;
            u = vel_scale * std::sin(kPi2 / Lx * x) * std::cos(kPi2 / Ly * y) * std::cos(kPi2 / Lz * z);
            v = -vel_scale * std::cos(kPi2 / Lx * x) * std::sin(kPi2 / Ly * y) * std::cos(kPi2 / Lz * z);
            w = 0;
            break;
          case IC_CHANNEL:  // src/dugks.cu:241-259 (plug flow, quirk Q11)
            w = vel_scale;
            break;
          case IC_SHEAR_WAVE:
            u = vel_scale * std::sin(kPi2 / Ly * y);
            break;
          case IC_TURB_CHANNEL:
            u = U[cell * 3 + 0];
            v = U[cell * 3 + 1];
            w = U[cell * 3 + 2];
            break;
          default:
            break;
        }
        for (int ip = 0; ip < Q; ip++) f[cell * Q + ip] = feq_host(ip, r, u, v, w);
      }
    }
  }
  if (s.ic != IC_TURB_CHANNEL) return;
  // Chapman-Enskog non-equilibrium part: f~neq = (1 + dt/(2 tau)) f_neq,
  // f_neq_i = -tau w_i rho (c_i c_i - RT I) : S / RT, S = sym(grad u), with
  // second-order differences (walls: mirrored ghost, u = 0 at the wall).
  const double fac = (1 + s.dt / (2 * s.tau)) * (-s.tau) / RT;
  auto vel = [&](int i, int j, int k, int c) -> double {
    int idx[3] = {i, j, k};
    double sign = 1;
    for (int a = 0; a < 3; a++) {
      if (idx[a] < 0 || idx[a] >= s.n[a]) {
        if (s.periodic[a])
          idx[a] = (idx[a] + s.n[a]) % s.n[a];
        else {
          idx[a] = idx[a] < 0 ? 0 : s.n[a] - 1;
          sign = -sign;
        }
      }
    }
    return sign * U[(((size_t)idx[2] * ny + idx[1]) * nx + idx[0]) * 3 + c];
  };
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        const int idx[3] = {i, j, k};
        double G[3][3];  // G[a][c] = d u_c / d x_a
        for (int a = 0; a < 3; a++) {
          int lo[3] = {i, j, k}, hi[3] = {i, j, k};
          lo[a]--;
          hi[a]++;
          const double dist = s.xc[a][idx[a] + 2] - s.xc[a][idx[a]];
          for (int c = 0; c < 3; c++)
            G[a][c] = (vel(hi[0], hi[1], hi[2], c) - vel(lo[0], lo[1], lo[2], c)) / dist;
        }
        const size_t cell = ((size_t)k * ny + j) * nx + i;
        for (int ip = 0; ip < Q; ip++) {
          double qs = 0;
          for (int a = 0; a < 3; a++)
            for (int c = 0; c < 3; c++) {
              double Sac = 0.5 * (G[a][c] + G[c][a]);
              double Qac = cv(a, ip) * cv(c, ip) - (a == c ? RT : 0.0);
              qs += Qac * Sac;
            }
          f[cell * Q + ip] += fac * wgt(ip) * 1.0 * qs;
        }
      }
}

}  // namespace dugks
