// Host file formats kept from the reference: the raw restart dump and the
// Tecplot / diag.dat / profiles-2.dat outputs.
#pragma once

#include <string>

#include "setup.h"

namespace dugks {

// "ddf%08d.dat" (src/output.cu:294-295)
std::string dump_name(int step);
std::string join_path(const std::string& dir, const std::string& name);

// Raw dump of count doubles as FP64 (real_bytes = 8) or FP32 (4), no header
// (src/output.cu:22-41); read is the inverse (src/output.cu:46-89).
void write_dump(const std::string& path, const double* f, size_t count, int real_bytes);
void read_dump(const std::string& path, double* f, size_t count, int real_bytes);

// save_tecplot_serial (src/output.cu:152-218) and statistics_serial
// (src/output.cu:223-334): the slices from the macroscopic fields, the
// statistics from the device's FP64 sums (finish_statistics).  Profiles are
// taken along the wall-normal axis (x in the reference), averaging over the
// other two.
void write_outputs(const Setup& s, int step, const double* rho, const double* ux,
                   const double* uy, const double* uz, const double* profiles, const double diag[12],
                   const std::string& dir);
// Axis of the planar statistics: the wall axis, else x (statistics_serial
// averages over the planes normal to x).
int stats_axis(const Setup& s);
// From the device sums, rows [n][9] and glob [4] sums + [4] min + [4] max,
// to the reference's profile columns (means, stresses uw/uv/vw, variances)
// and diag (means, min, max of rho, u, v, w).  profiles may be rows and diag
// may be glob (in place).
void finish_statistics(const Setup& s, const double* rows, const double* glob, double* profiles, double diag[12]);

// VELMEAN.dat from profiles-2.dat (examples/channel/ProcessProfiles.py:17-79).
// DHIT statistics files of stat_te (src/fortran/dhit/src/stat.f90:58-198),
// appended / written in dir: stat_te.dat, stat_te_spectral.dat,
// spe%09d.dat (shells 0..nspec-1) and skew-flat.dat.  stats is the
// dugks_spectral_statistics vector.
void write_spectral_outputs(int step, double dt, const double stats[13], const double* spectrum, int nspec,
                            const std::string& dir);

void velmean(const std::string& in_path, int n, int ignore_before, int mean_col, int stress_col,
             const std::string& out_path);

}  // namespace dugks
