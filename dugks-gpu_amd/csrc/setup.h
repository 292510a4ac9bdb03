// Host-side setup: config.txt parsing, derived numerics, mesh, initial state.
// Plain C++ (no HIP), shared by the library and the driver.
#pragma once

#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace dugks {

struct ConfigError : std::runtime_error {
  int code;
  ConfigError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// key=value reader with the reference's rules (src/config.h:17-74): lines
// without '=' are ignored, the text before the first '=' is the key and the
// text up to the next '=' the value; the first occurrence of a key wins.
// Keys and values are trimmed of blanks and CR.  A missing key is an error
// (the reference inserted "" and crashed in stoi, quirk Q10).
class Config {
 public:
  static Config from_text(const std::string& text);
  static Config from_file(const std::string& path);
  bool has(const std::string& k) const { return kv_.count(k) != 0; }
  std::string str(const std::string& k) const;
  std::string str(const std::string& k, const std::string& dflt) const;
  int integer(const std::string& k) const;
  int integer(const std::string& k, int dflt) const;
  // Reals go through float like std::stof (src/numerics.cu:74-99, quirk Q3)
  // unless parse_real=double.
  double real(const std::string& k) const;
  double real(const std::string& k, double dflt) const;
  const std::map<std::string, std::string>& all() const { return kv_; }

 private:
  std::map<std::string, std::string> kv_;
  bool real_as_float_ = true;
};

enum IcKind { IC_GENERAL = 0, IC_TG_EQ, IC_CHANNEL, IC_REST, IC_SHEAR_WAVE, IC_TURB_CHANNEL };

struct Setup {
  std::string case_name;
  std::string base_dir;
  int n[3] = {0, 0, 0};
  int periodic[3] = {1, 1, 1};
  double width[3] = {0, 0, 0};
  int uniform[3] = {1, 1, 1};
  std::vector<double> d[3];   // cell widths, ghost-inclusive (index 0..n+1)
  std::vector<double> xc[3];  // cell centres, ghost-inclusive
  double du[3] = {0, 0, 0};   // width/n
  double CFL = 0, dt = 0, visc = 0, tau = 0, hdt = 0;
  double bw0 = 0, bw1 = 0, bw2 = 0, fw0 = 0, fw1 = 0, fw2 = 0;
  double famb[3] = {0, 0, 0};
  int istep0 = 1, nsteps = 0, noutput2 = 1, noutput3 = 1, ndump = 1;
  int restart = 0, nrestart = 0, perturb = 0;
  int recon = 0;      // 0 tri, 1 grad
  int precision = 0;  // DUGKS_PREC_*
  int kernel = -1;    // DUGKS_KERNEL_*
  int nprint = 1;     // driver prints "istep" every nprint steps
  int ic = IC_GENERAL;
  double u0 = 0.1;    // velocity scale of the analytic initial states
  double rho0 = 1.0;  // reference density of the FP16 deviation codec
  // constant-flow-rate forcing: target bulk velocity along flow_axis
  double flow_rate = 0.0;
  bool flow_auto = false;  // flow_rate=auto: target = bulk velocity of the initial state
  int flow_axis = -1;
  double flow_gain = 1.0;
  // turbulent-channel initial state
  double retau = 180.0, pert_amp = 0.1;
  int laminar_profile = 0;  // mean_profile=laminar: U = u0 (1 - eta^2) instead of the log law
  int kmin = 1, kmax = 8, nmodes = 24;
  unsigned long long seed = 12345;

  size_t ncell() const { return (size_t)n[0] * n[1] * n[2]; }
  int wall_axis() const;        // first non-periodic axis, -1 if none
  int stream_axis() const;      // axis of the largest |famb| (or flow_axis)
};

// Builds a Setup from config text; base_dir locates mesh and dump files.
Setup make_setup(const Config& cfg, const std::string& base_dir);

// Face coordinates 0..n of the reference's tanh mesh generator
// (src/fortran/channel/run/mesh.py:21-25) for a wall-bounded axis of width w.
std::vector<double> tanh_faces(int n, double alpha, double w);

// Equilibrium of the host IC (feq, src/dugks.cu:304-309).
double feq_host(int ip, double r, double u, double v, double w);

// Initial f~ in the restart-dump layout (nx*ny*nz*19 doubles) for the case
// selected in the setup (src/dugks.cu:221-281 plus the new cases).
void initial_state(const Setup& s, std::vector<double>& f);

}  // namespace dugks
