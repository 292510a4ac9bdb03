// dugks_main: the host driver, src/main.cu:13-56 of the reference over the
// C ABI: init -> time loop (step, periodic output_2d, periodic dump) ->
// "Main loop time".  Usage: dugks_main [config.txt] [output_dir]
#include <chrono>
#include <cstdio>
#include <string>

#include "../../include/dugks.h"

int main(int argc, char** argv) {
  const char* cfg = argc > 1 ? argv[1] : "config.txt";
  const char* out = argc > 2 ? argv[2] : ".";
  dugks_opts o;
  dugks_default_opts(&o);
  o.verbose = 1;
  dugks_ctx* ctx = nullptr;
  int rc = dugks_init(cfg, &o, &ctx);
  if (rc != DUGKS_OK) {
    std::fprintf(stderr, "dugks_init failed (%d): %s\n", rc, dugks_init_error());
    return 1;
  }
  int sch[6], dims[3];
  dugks_get_schedule(ctx, sch);
  dugks_get_dims(ctx, dims);
  const int istep0 = sch[0], nsteps = sch[1], noutput2 = sch[2], ndump = sch[4], nprint = sch[5];
  auto t0 = std::chrono::steady_clock::now();
  for (int istep = istep0; istep <= istep0 + nsteps - 1; istep++) {
    if (istep % nprint == 0) std::printf("istep = %8d\n", istep);
    if ((rc = dugks_step(ctx, istep, 1)) != DUGKS_OK) break;
    if (istep == istep0 || istep % noutput2 == 0)
      if ((rc = dugks_write_outputs(ctx, istep, out)) != DUGKS_OK) break;
    if (istep % ndump == 0)
      if ((rc = dugks_write_dump(ctx, istep, out, 8)) != DUGKS_OK) break;
  }
  if (rc == DUGKS_OK) rc = dugks_sync(ctx);
  if (rc != DUGKS_OK) {
    std::fprintf(stderr, "dugks error (%d): %s\n", rc, dugks_last_error(ctx));
    dugks_free(ctx);
    return 1;
  }
  double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  double cells = (double)dims[0] * dims[1] * dims[2];
  std::printf("Main loop time: %g seconds.\n", secs);
  if (nsteps > 0 && secs > 0) std::printf("MLUPS: %g\n", cells * nsteps / secs / 1e6);
  dugks_free(ctx);
  return 0;
}
