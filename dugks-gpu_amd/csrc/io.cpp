// Restart dump and output writers (host C++).
//   dump / read_dump     src/output.cu:22-89 (raw real values, f[k][j][i][ip])
//   tecplot slices       src/output.cu:152-218
//   diag.dat, profiles   src/output.cu:223-334
#include "io.h"

#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <vector>

#include "../../include/dugks.h"

namespace dugks {

std::string dump_name(int step) {
  char buf[32];
  std::snprintf(buf, sizeof buf, "ddf%8.8d.dat", step);
  return buf;
}

std::string join_path(const std::string& dir, const std::string& name) {
  if (dir.empty()) return name;
  if (dir.back() == '/') return dir + name;
  return dir + "/" + name;
}

void write_dump(const std::string& path, const double* f, size_t count, int real_bytes) {
  if (real_bytes != 8 && real_bytes != 4) throw ConfigError(DUGKS_EARG, "dump real_bytes must be 4 or 8");
  std::ofstream o(path, std::ios::binary);
  if (!o) throw ConfigError(DUGKS_EIO, "cannot open '" + path + "' for writing");
  if (real_bytes == 8) {
    o.write(reinterpret_cast<const char*>(f), count * sizeof(double));
  } else {
    const size_t chunk = 1 << 20;
    std::vector<float> buf;
    for (size_t b = 0; b < count; b += chunk) {
      size_t m = std::min(chunk, count - b);
      buf.resize(m);
      for (size_t q = 0; q < m; q++) buf[q] = (float)f[b + q];
      o.write(reinterpret_cast<const char*>(buf.data()), m * sizeof(float));
    }
  }
  if (!o) throw ConfigError(DUGKS_EIO, "write failed on '" + path + "'");
}

void read_dump(const std::string& path, double* f, size_t count, int real_bytes) {
  if (real_bytes != 8 && real_bytes != 4) throw ConfigError(DUGKS_EARG, "dump real_bytes must be 4 or 8");
  std::ifstream in(path, std::ios::binary | std::ios::ate);
  if (!in) throw ConfigError(DUGKS_EIO, "cannot open dump '" + path + "'");
  const std::streamoff size = in.tellg();
  if ((size_t)size < count * (size_t)real_bytes)
    throw ConfigError(DUGKS_EIO, "dump '" + path + "' is shorter than nx*ny*nz*19 values");
  in.seekg(0);
  if (real_bytes == 8) {
    in.read(reinterpret_cast<char*>(f), count * sizeof(double));
  } else {
    const size_t chunk = 1 << 20;
    std::vector<float> buf;
    for (size_t b = 0; b < count; b += chunk) {
      size_t m = std::min(chunk, count - b);
      buf.resize(m);
      in.read(reinterpret_cast<char*>(buf.data()), m * sizeof(float));
      for (size_t q = 0; q < m; q++) f[b + q] = buf[q];
    }
  }
  if (!in) throw ConfigError(DUGKS_EIO, "read failed on '" + path + "'");
}

namespace {

void slice(const Setup& s, const char* tag, int fixed_axis, int step, const double* rho,
           const double* ux, const double* uy, const double* uz, const std::string& dir) {
  const int nx = s.n[0], ny = s.n[1], nz = s.n[2];
  char name[32];
  std::snprintf(name, sizeof name, "s2%s%8.8d.tec", tag, step);
  std::ofstream file(join_path(dir, name));
  if (!file) throw ConfigError(DUGKS_EIO, std::string("cannot write ") + name);
  int dims[3] = {nx, ny, nz};
  dims[fixed_axis] = 1;
  file << "TITLE = \"Example: Simple 2D-Slice Data\"" << std::endl;
  file << "VARIABLES = \"X\", \"Y\", \"Z\", \"RHO\", \"UX\", \"UY\", \"UZ\"" << std::endl;
  file << "ZONE I=" << dims[0] << ", J=" << dims[1] << ", K=" << dims[2] << " F=POINT" << std::endl;
  auto coord = [&](int a, int m) { return s.uniform[a] ? (m + 0.5) * s.du[a] : s.xc[a][m + 1]; };
  const int fixed = s.n[fixed_axis] / 2;
  for (int k = 0; k < (fixed_axis == 2 ? 1 : nz); k++)
    for (int j = 0; j < (fixed_axis == 1 ? 1 : ny); j++)
      for (int i = 0; i < (fixed_axis == 0 ? 1 : nx); i++) {
        int idx[3] = {i, j, k};
        idx[fixed_axis] = fixed;
        size_t c = ((size_t)idx[2] * ny + idx[1]) * nx + idx[0];
        file << coord(0, idx[0]) << " " << coord(1, idx[1]) << " " << coord(2, idx[2]) << " "
             << rho[c] << " " << ux[c] << " " << uy[c] << " " << uz[c] << std::endl;
      }
}

}  // namespace

void write_outputs(const Setup& s, int step, const double* rho, const double* ux,
                   const double* uy, const double* uz, const double* profiles, const double diag[12],
                   const std::string& dir) {
  slice(s, "X", 0, step, rho, ux, uy, uz, dir);
  slice(s, "Y", 1, step, rho, ux, uy, uz, dir);
  slice(s, "Z", 2, step, rho, ux, uy, uz, dir);
  FILE* file = std::fopen(join_path(dir, "diag.dat").c_str(), "a");
  if (!file) throw ConfigError(DUGKS_EIO, "cannot append diag.dat");
  std::fprintf(file,
               "%8.8d  % 7.7e % 7.7e % 7.7e % 7.7e  % 7.7e % 7.7e % 7.7e % 7.7e  % 7.7e % 7.7e "
               "% 7.7e % 7.7e\n",
               step, diag[0], diag[1], diag[2], diag[3], diag[4], diag[5], diag[6], diag[7], diag[8], diag[9],
               diag[10], diag[11]);
  std::fclose(file);
  // per wall-normal index: means, Reynolds stresses xz/xy/yz, variances
  const int np = s.n[stats_axis(s)];
  file = std::fopen(join_path(dir, "profiles-2.dat").c_str(), "a");
  if (!file) throw ConfigError(DUGKS_EIO, "cannot append profiles-2.dat");
  std::fprintf(file, "%8.8d\n", step);
  for (int m = 0; m < np; m++) {
    const double* a = profiles + (size_t)m * 9;
    std::fprintf(file, "% 7.7e % 7.7e % 7.7e   % 7.7e % 7.7e % 7.7e   % 7.7e % 7.7e % 7.7e \n", a[0], a[1],
                 a[2], a[3], a[4], a[5], a[6], a[7], a[8]);
  }
  std::fclose(file);
}

int stats_axis(const Setup& s) {
  const int pa = s.wall_axis();
  return pa < 0 ? 0 : pa;
}

void finish_statistics(const Setup& s, const double* rows, const double* glob, double* profiles, double diag[12]) {
  const int pa = stats_axis(s);
  const int np = s.n[pa];
  const double cnt = (double)s.ncell() / np;
  for (int m = 0; m < np; m++) {
    const double* a = rows + (size_t)m * 9;
    double* o = profiles + (size_t)m * 9;
    for (int q = 0; q < 9; q++) o[q] = a[q] / cnt;
    o[6] -= o[0] * o[0];
    o[7] -= o[1] * o[1];
    o[8] -= o[2] * o[2];
    o[3] -= o[0] * o[2];
    o[4] -= o[0] * o[1];
    o[5] -= o[1] * o[2];
  }
  const double* g = glob;
  for (int q = 0; q < 4; q++) {
    diag[q] = g[q] / (double)s.ncell();
    diag[4 + q] = g[4 + q];
    diag[8 + q] = g[8 + q];
  }
}

void write_spectral_outputs(int step, double dt, const double st[13], const double* spectrum, int nspec,
                            const std::string& dir) {
  // Fortran formats I7, F16.8, 1pe15.6 and I4 with 1X separators
  auto open_append = [&](const char* name) {
    FILE* f = std::fopen(join_path(dir, name).c_str(), "a");
    if (!f) throw ConfigError(DUGKS_EIO, std::string("cannot append ") + name);
    return f;
  };
  const double t = step * dt;
  FILE* f = open_append("stat_te.dat");
  std::fprintf(f, "%7d %16.8f %15.6E\n", step, t, st[0]);
  std::fclose(f);
  // etotal, ediss_total0, ediss_total (written as 0 by the reference), eta,
  // vk, tk, uprm, tmse, Retms, xl, kmxeta
  f = open_append("stat_te_spectral.dat");
  std::fprintf(f, "%7d %16.8f %15.6E %15.6E %15.6E %15.6E %15.6E %15.6E %15.6E %15.6E %15.6E %15.6E %15.6E \n", step,
               t, st[1], st[2], 0.0, st[3], st[4], st[5], st[6], st[7], st[8], st[9], st[10]);
  std::fclose(f);
  char name[32];
  std::snprintf(name, sizeof name, "spe%9.9d.dat", step);
  f = std::fopen(join_path(dir, name).c_str(), "w");
  if (!f) throw ConfigError(DUGKS_EIO, std::string("cannot write ") + name);
  for (int i = 0; i < nspec; i++) std::fprintf(f, "%4d %15.6E\n", i, spectrum[i]);
  std::fclose(f);
  f = open_append("skew-flat.dat");
  std::fprintf(f, "%7d %15.6E %15.6E %15.6E\n", step, t, st[11], st[12]);
  std::fclose(f);
}

// A double as Python's repr() writes it (the VELMEAN.dat rows are Python
// f-strings): the shortest digit string that reads back to the same double,
// in positional notation with at least one fractional digit ("1.0",
// "0.0001") when the decimal exponent e of the leading digit satisfies
// -4 <= e < 16, else as d[.ddd]e+XX with at least two exponent digits
// ("1e-05", "1.5e+16").  The shortest round-trip digits come from
// std::to_chars in scientific form.
std::string py_repr(double v) {
  if (std::isnan(v)) return "nan";
  if (std::isinf(v)) return v < 0 ? "-inf" : "inf";
  char buf[64];
  auto r = std::to_chars(buf, buf + sizeof buf, v, std::chars_format::scientific);
  std::string s(buf, r.ptr);
  std::string sign;
  if (s[0] == '-') {
    sign = "-";
    s = s.substr(1);
  }
  const size_t epos = s.find('e');
  const int e = std::stoi(s.substr(epos + 1));
  std::string digits;
  for (size_t q = 0; q < epos; q++)
    if (s[q] != '.') digits += s[q];
  const int nd = (int)digits.size();
  if (e < -4 || e >= 16) {
    std::string m = digits.substr(0, 1);
    if (nd > 1) m += "." + digits.substr(1);
    char eb[16];
    std::snprintf(eb, sizeof eb, "e%c%02d", e < 0 ? '-' : '+', e < 0 ? -e : e);
    return sign + m + eb;
  }
  const int decpt = e + 1;  // digits before the decimal point
  if (decpt <= 0) return sign + "0." + std::string((size_t)-decpt, '0') + digits;
  if (decpt < nd) return sign + digits.substr(0, (size_t)decpt) + "." + digits.substr((size_t)decpt);
  return sign + digits + std::string((size_t)(decpt - nd), '0') + ".0";
}

// Time average and half-channel symmetrisation of profiles-2.dat into
// VELMEAN.dat (examples/channel/ProcessProfiles.py:17-79 of the reference):
// the blocks whose step exceeds ignore_before are averaged; the mean and the
// rms columns are symmetrised as (a + reverse a) / 2, the shear stress as
// (s - reverse s) / 2.  Output rows: index (1-based), mean, stress, and the
// three variance columns, each written as Python's repr() writes it.  The
// step of each block is the one in the file (the reference script counts
// blocks on its params schedule, which gives the same steps for a file
// written on that schedule).
void velmean(const std::string& in_path, int n, int ignore_before, int mean_col, int stress_col,
             const std::string& out_path) {
  FILE* f = std::fopen(in_path.c_str(), "r");
  if (!f) throw ConfigError(DUGKS_EIO, "cannot open " + in_path);
  std::vector<double> mean(n, 0.0), stress(n, 0.0), var[3];
  for (auto& v : var) v.assign(n, 0.0);
  int count = 0, step = 0;
  while (std::fscanf(f, "%d", &step) == 1) {
    const bool use = step > ignore_before;
    if (use) count++;
    for (int k = 0; k < n; k++) {
      double row[9];
      for (double& x : row)
        if (std::fscanf(f, "%lf", &x) != 1) {
          std::fclose(f);
          throw ConfigError(DUGKS_EIO, "truncated block in " + in_path);
        }
      if (use) {
        mean[k] += row[mean_col];
        stress[k] += row[stress_col];
        for (int q = 0; q < 3; q++) var[q][k] += row[6 + q];
      }
    }
  }
  std::fclose(f);
  if (count == 0) throw ConfigError(DUGKS_EIO, "no profile block after the ignored steps in " + in_path);
  auto sym = [&](std::vector<double>& a, double sign) {
    std::vector<double> r(n);
    for (int k = 0; k < n; k++) r[k] = 0.5 * (a[k] / count + sign * a[n - 1 - k] / count);
    a = r;
  };
  sym(mean, 1.0);
  sym(stress, -1.0);
  for (auto& v : var) sym(v, 1.0);
  FILE* g = std::fopen(out_path.c_str(), "w");
  if (!g) throw ConfigError(DUGKS_EIO, "cannot write " + out_path);
  for (int k = 0; k < n; k++) {
    const std::string row = py_repr((double)(k + 1)) + " " + py_repr(mean[k]) + " " + py_repr(stress[k]) + " " +
                            py_repr(var[0][k]) + " " + py_repr(var[1][k]) + " " + py_repr(var[2][k]) + "\n";
    std::fputs(row.c_str(), g);
  }
  std::fclose(g);
}

}  // namespace dugks
