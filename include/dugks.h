/*
 * dugks.h -- C ABI of the MI355X DUGKS engine.
 *
 * The reference (Kairzhan/DUGKS-GPU) has no FFI; its solver boundary is the
 * set of global-state C++ free functions declared in src/dugks.h:1-9
 * (init_dugks, dugks_step, upload_macro, copy_vel, copy_ddf) with inputs in
 * globals (src/numerics.h:8-33, src/parallel.h:10-25) and errors handled by
 * std::exit.  Each entry point below names the reference function it
 * replaces.  Conventions:
 *   - every function returns 0 on success or a negative DUGKS_E* code, and
 *     dugks_last_error() gives the message; nothing calls exit();
 *   - the context owns all device memory, the caller owns host buffers;
 *   - one context per host thread (thread-compatible, not thread-safe);
 *   - work is ordered on the context's HIP stream; dugks_step returns without
 *     a host synchronisation, every host-buffer transfer synchronises;
 *   - host f arrays use the reference restart-dump layout
 *     f[((k*ny + j)*nx + i)*19 + ip] (src/output.cu:22-41), FP64;
 *   - device f arrays use structure-of-arrays f[ip][k][j][i], FP64.
 */
#ifndef DUGKS_H
#define DUGKS_H

#ifdef __cplusplus
extern "C" {
#endif

#define DUGKS_OK 0
#define DUGKS_EARG (-1)        /* bad argument */
#define DUGKS_ECONFIG (-2)     /* config.txt missing key / bad value */
#define DUGKS_EIO (-3)         /* file I/O */
#define DUGKS_EHIP (-4)        /* HIP runtime error (incl. no device) */
#define DUGKS_ESTATE (-5)      /* call not valid in the current state */
#define DUGKS_EUNSUPPORTED (-6)/* feature combination not supported */
#define DUGKS_ENOMEM (-7)      /* device or host allocation failed */
#define DUGKS_ENUMERIC (-8)    /* non-finite or non-positive density found */

/* Precision modes (the reference has one compile-time switch, PREC4,
 * src/precision.h:4-8). */
#define DUGKS_PREC_CONFIG (-1) /* take "precision" from config.txt (default fp64) */
#define DUGKS_PREC_FP64 0      /* FP64 arithmetic, FP64 storage of f */
#define DUGKS_PREC_FP32 1      /* FP32 arithmetic, FP32 storage of f */
#define DUGKS_PREC_FP16S 2     /* FP32 arithmetic, FP16 storage of (f - w_i rho0)/w_i */
#define DUGKS_PREC_FP16N 3     /* FP32 arithmetic, FP32 rho and u = j/rho plus FP16 storage of
                                  (f - feq(rho, u)) * 1024 / w_i; no adjoint */

/* Step kernels. */
#define DUGKS_KERNEL_AUTO (-1)
#define DUGKS_KERNEL_TWOPASS 0 /* f-bar kernel + ghost fill + face/update kernel */
#define DUGKS_KERNEL_FUSED 1   /* one fused kernel per step: z periodic and uniform, and nx % 16 == 0,
                                  ny % 12 == 0 or nx % 32 == 0, ny % 4 == 0, one state array
                                  < 4 GiB; either reconstruction
                                  (config key recon = tri | grad); DUGKS_KERNEL_AUTO takes it
                                  whenever these hold */

typedef struct dugks_ctx dugks_ctx;

typedef struct dugks_opts {
  int precision; /* DUGKS_PREC_* */
  int device;    /* HIP device ordinal, -1 = current device */
  void* stream;  /* hipStream_t to order work on, NULL = context-owned stream */
  int kernel;    /* DUGKS_KERNEL_* */
  int verbose;   /* 0 = silent, 1 = parameter summary on init */
} dugks_opts;

void dugks_default_opts(dugks_opts* opts);

/* Replaces Numerics::init_numerics (src/numerics.cu:62-106), init_parallel's
 * periods (src/parallel.cu:95-97) and init_dugks (src/dugks.cu:95-302):
 * reads config_path ("key=value" lines, src/config.h:17-74), derives the mesh
 * and coefficients, allocates device memory and builds the initial state
 * (case IC, or read_dump of ddf<nrestart>.dat from the config's directory when
 * restart=1, src/output.cu:46-89). */
int dugks_init(const char* config_path, const dugks_opts* opts, dugks_ctx** out);

/* Same, from the text of a config file; restart dumps are looked up in
 * base_dir (NULL = current directory). */
int dugks_init_from_string(const char* config_text, const char* base_dir,
                           const dugks_opts* opts, dugks_ctx** out);

void dugks_free(dugks_ctx* ctx);
const char* dugks_last_error(const dugks_ctx* ctx);
/* Message of the last failed dugks_init* on this thread (ctx not created). */
const char* dugks_init_error(void);

/* Replaces dugks_step(int) (src/dugks.cu:322-391), called once per step by
 * src/main.cu:26-44: advances nsteps steps, istep0 being the index of the
 * first (used by time-dependent forcing).  Stream-ordered, no host sync. */
int dugks_step(dugks_ctx* ctx, int istep0, int nsteps);

/* dugks_step with a pair of hipEvents recorded on the context's stream
 * around every step kernel launch (the fused kernel, or the face/update
 * kernel of the two-pass path): out[0] = mean kernel time (ms), out[1] =
 * min, out[2] = max, out[3] = wall time of the whole loop on the stream (ms),
 * out[4] = launches per step.  Synchronises.  The reference only prints the
 * main-loop wall time (src/main.cu:47-51). */
int dugks_time_steps(dugks_ctx* ctx, int istep0, int nsteps, double out[5]);

/* Waits for all work queued on the context's stream. */
int dugks_sync(dugks_ctx* ctx);

/* Grid size nx, ny, nz (Numerics::nx/ny/nz). */
int dugks_get_dims(const dugks_ctx* ctx, int dims[3]);

/* HIP device ordinal the context runs on (the device chosen by init_cuda,
 * src/cuda.cu:13-33; every device pointer passed to this context must be
 * on it). */
int dugks_get_device(const dugks_ctx* ctx, int* device);

/* Derived parameters, in this order: dt, hdt, tau, visc, bw0, bw1, bw2, fw0,
 * fw1, fw2, dx, dy, dz (minimum width on a stretched axis), rho0,
 * precision, kernel (the test_params printout, src/dugks.cu:983-1016). */
int dugks_get_params(const dugks_ctx* ctx, double out[16]);

/* Run schedule from the config (Numerics::istep0, nsteps, noutput2,
 * noutput3, ndump, plus nprint, the cadence of the "istep =" line that
 * src/main.cu:27-29 prints every step). */
int dugks_get_schedule(const dugks_ctx* ctx, int out[6]);

/* Replaces output_2d (src/output.cu:118-147): Tecplot slices s2X/s2Y/s2Z,
 * and diag.dat / profiles-2.dat appended (statistics_serial,
 * src/output.cu:223-334, FP64 accumulators) in dir (NULL = current). */
int dugks_write_outputs(dugks_ctx* ctx, int step, const char* dir);

/* Planar statistics of the current state along the wall-normal axis (x when
 * there is no wall), computed on the device with FP64 sums: statistics_serial,
 * src/output.cu:223-334.  profiles [n][9] with n = *rows (the wall-normal
 * size; the buffer must hold max(nx, ny, nz) * 9 doubles): mean u, v, w;
 * stresses <u'w'>, <u'v'>, <v'w'>; variances <u'u'>, <v'v'>, <w'w'> (the
 * profiles-2.dat columns).  diag[12]: mean, min, max of rho, ux, uy, uz (the
 * diag.dat columns).  Host buffers; synchronises the stream. */
int dugks_get_statistics(dugks_ctx* ctx, double* profiles, double diag[12], int* rows);

/* dugks_get_statistics without allocating: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results (the same kernels, bit for bit); its scratch is a
 * library-held device array, one per GPU for the life of the process, so it
 * calls no hipMalloc or hipFree.  Limit: at most 2048 rows along the
 * wall-normal axis (DUGKS_EUNSUPPORTED beyond).  Calls that overlap on one
 * GPU share that scratch and race. */
int dugks_get_statistics_noalloc(dugks_ctx* ctx, double* profiles, double diag[12], int* rows);

/* Spectral statistics of decaying homogeneous isotropic turbulence, stat_te
 * of the reference's DHIT code (src/fortran/dhit/src/stat.f90:5-205), of the
 * current velocity field (the output velocity of dugks_get_macro), on the
 * device with hipFFT and FP64 sums.  Needs periodic uniform axes and a cubic
 * domain.  The fluctuation u' = u - <u> is transformed (FFT normalised by
 * N); shell k holds the modes with nint(|kappa|) = k, kappa the integer
 * wavenumber vector (wavenumber 2 pi kappa / L).  out[13]:
 *   te = 0.5 <|u|^2> (physical space, mean included), etotal = sum_k E(k),
 *   eps = 2 nu sum |k|^2 E, Kolmogorov length, velocity and time,
 *   u' = sqrt(2 etotal / 3), Taylor microscale sqrt(15 nu u'^2 / eps),
 *   Re_lambda, large-eddy length u'^3 / eps, k_max eta with
 *   k_max = (L/2 - 1.5 dx) 2 pi / L, and the skewness and flatness of the
 *   longitudinal derivatives du/dx, dv/dy, dw/dz (spectral derivatives).
 * spectrum: E(k) of shells 0..nbins-1 (may be NULL when nbins = 0).
 * Synchronises.  Differences from the reference: exact pi (it uses
 * 3.14159), FP64 sums (it sums in FP32). */
int dugks_spectral_statistics(dugks_ctx* ctx, double* spectrum, int nbins, double out[13]);
/* The stat_te files in dir (NULL = current): appends to stat_te.dat,
 * stat_te_spectral.dat and skew-flat.dat and writes spe%09d.dat (shells
 * 0..nx), in the reference's formats (stat.f90:58-198). */
int dugks_write_spectral_outputs(dugks_ctx* ctx, int step, const char* dir);

/* Cell widths of one axis (n values); centres are cumulative midpoints from
 * the low wall/edge at coordinate 0. */
int dugks_get_spacing(const dugks_ctx* ctx, int axis, double* widths);

/* Replaces copy_ddf (src/dugks.cu:407-416) + collect_f: whole-domain f in the
 * restart-dump layout, FP64 (decoded from the storage precision). */
int dugks_get_f(dugks_ctx* ctx, double* host_f);
/* Inverse: load f from the restart-dump layout (read_dump's scatter). */
int dugks_set_f(dugks_ctx* ctx, const double* host_f);

/* Device-side exchange in SoA layout f[ip][k][j][i], FP64, for callers that
 * keep state on the GPU (the torch wrapper).  Stream-ordered, no host sync.
 * dugks_set_f_device restarts the flow-rate controller like dugks_set_f; it
 * does not check the values for finiteness (dugks_set_f does). */
int dugks_get_f_device(dugks_ctx* ctx, double* dev_f);
int dugks_set_f_device(dugks_ctx* ctx, const double* dev_f);

/* Replaces upload_macro + calc_macro + copy_vel (src/dugks.cu:311-320,
 * 1021-1064, 393-405): rho = sum f, u = sum c f / rho + h F_amb, arrays
 * [k][j][i] of nx*ny*nz doubles (no halo). */
int dugks_get_macro(dugks_ctx* ctx, double* rho, double* ux, double* uy, double* uz);

/* Ambient body force (ambient_fx/fy/fz, HydroParameters, src/dugks.cu:48-50,
 * 211-215).  Used at cell centres (calc_force) and at faces (f2ruvwh). */
int dugks_set_ambient_force(dugks_ctx* ctx, const double force[3]);
int dugks_get_ambient_force(dugks_ctx* ctx, double force[3]);

/* Per-cell centre force added to the ambient force in the f-bar stage only,
 * as the FORCEX/Y/Z arrays of calc_fb1 (src/dugks.cu:1131-1196; faces keep
 * the ambient force, quirk Q4).  Arrays [k][j][i]; NULL pointers clear the
 * field.  This is the body-force control of the adjoint. */
int dugks_set_cell_force(dugks_ctx* ctx, const double* fx, const double* fy, const double* fz);
/* Device variant: dev_force is [3][nz][ny][nx] FP64 on the context's device. */
int dugks_set_cell_force_device(dugks_ctx* ctx, const double* dev_force);

/* Replaces dump / read_dump (src/output.cu:22-89): ddf%08d.dat in dir (NULL =
 * current directory), raw FP64 (real_bytes=8) or FP32 (real_bytes=4, the
 * PREC4 build's files), no header. */
int dugks_write_dump(dugks_ctx* ctx, int step, const char* dir, int real_bytes);
int dugks_read_dump(dugks_ctx* ctx, int step, const char* dir, int real_bytes);

/* Global sums in FP64 with a fixed reduction order: out[0] = mass sum f,
 * out[1..3] = momentum sum c f, out[4..6] = sum(u dV) of the output velocity
 * u of dugks_get_macro (the bulk velocity is out[4+a]/out[7]), out[7] =
 * domain volume. */
int dugks_global_sums(dugks_ctx* ctx, double out[8]);

/* ---- Adjoint (no reference counterpart: the reference has no adjoint) ---- */

/* Discrete adjoint of one step about the context's current state f^n:
 * lam_prev = (df^{n+1}/df^n)^T lam_next and, when grad_force is not NULL,
 * grad_force += (df^{n+1}/dF_cell)^T lam_next.  All three are device arrays:
 * lam_* SoA [19][nz][ny][nx] FP64, grad_force [3][nz][ny][nx] FP64;
 * lam_prev must not overlap lam_next.  The state is not advanced. */
int dugks_adjoint_step(dugks_ctx* ctx, const double* lam_next, double* lam_prev,
                       double* grad_force);

/* dugks_adjoint_step with lam_next and lam_prev in FP32 (device SoA
 * [19][nz][ny][nx] float; grad_force stays FP64).  For the FP32 and
 * FP16-store modes on the configurations of the z-marching gather (nx % 32 ==
 * 0, ny % 8 == 0, periodic uniform z, the fused step, at most one wall axis),
 * else DUGKS_EUNSUPPORTED.  Those modes compute lam in FP32 arithmetic, so
 * FP32 storage loses nothing and halves the lam traffic; dugks_forward_adjoint
 * uses it internally where it applies. */
int dugks_adjoint_step_f32(dugks_ctx* ctx, const float* lam_next, float* lam_prev, double* grad_force);

/* Loss functionals J(f) = sum_cells g(f) evaluated on the current state and
 * their gradient dJ/df (device SoA FP64, written, not accumulated).
 * kind 0: mean wall shear stress (walls of the wall-normal axis, streamwise
 *         axis = axis of the largest ambient force component);
 * kind 1: kinetic energy 0.5 * sum rho |u|^2 dV / V. */
#define DUGKS_LOSS_WALL_SHEAR 0
#define DUGKS_LOSS_KINETIC_ENERGY 1
int dugks_loss(dugks_ctx* ctx, int kind, double* value, double* dev_grad);

/* The sqrt(T)-checkpointed sweep in two calls.
 *
 * dugks_forward_record: T steps from the current state accumulating
 * J = (1/T) sum_{n=1..T} loss(f^n) on the device, keeping f every `interval`
 * steps (interval <= 0: ceil(sqrt(T)); clamped to T) and f^T in a checkpoint
 * pool on the context.  On return the state is f^T and *J holds the loss.
 *
 * dugks_adjoint_replay: the reverse sweep of the recorded forward, each
 * segment recomputed from its checkpoint.  dev_grad_force ([3][nz][ny][nx]
 * FP64) gets dJ/dF_cell, dev_grad_f0 ([19][nz][ny][nx] FP64, may be NULL)
 * gets dJ/df^0.  It uses the context's cell force, which must be the one the
 * forward was recorded with; the context's state is not used or changed, so
 * other work may run between the two calls.  peak_bytes (may be NULL) gets
 * the device memory in use (hipMemGetInfo) at the end of the sweep, when all
 * of its buffers are allocated.  The pool stays allocated for the next
 * sweep until dugks_release_sweep or dugks_free.
 *
 * With the flow-rate controller on (flow_rate != 0) the forward also records
 * the ambient force before each step, and the reverse sweep differentiates
 * the controller: each step's force is a function of the bulk velocities
 * the controller read, so the gradients include the controller's response.
 * This needs FP64, or FP32 arithmetic whose adjoint keeps lam in FP32, on
 * the fused face pass (DUGKS_EUNSUPPORTED otherwise); streak forcing
 * (perturb) is not covered.  dugks_sweep_force_gradient then gives dJ/dF^0,
 * the derivative with respect to the ambient force along the flow axis when
 * the record began.
 *
 * dugks_forward_adjoint is the two in one call. */
int dugks_forward_record(dugks_ctx* ctx, int istep0, int T, int loss_kind, int interval, double* J);
int dugks_adjoint_replay(dugks_ctx* ctx, double* dev_grad_force, double* dev_grad_f0, double* peak_bytes);
int dugks_forward_adjoint(dugks_ctx* ctx, int istep0, int T, int loss_kind, int interval,
                          double* J, double* dev_grad_force, double* dev_grad_f0,
                          double* peak_bytes);
int dugks_sweep_force_gradient(dugks_ctx* ctx, double* dJdF0);
/* Frees the checkpoint pool; *bytes of dugks_sweep_bytes is its current size. */
int dugks_release_sweep(dugks_ctx* ctx);
int dugks_sweep_bytes(dugks_ctx* ctx, double* bytes);

/* ---- Host-only helpers: no device is touched; messages via dugks_init_error ---- */

/* init_numerics without a device (src/numerics.cu:62-106): grid size and the
 * dugks_get_params vector derived from a config text. */
int dugks_config_derive(const char* config_text, const char* base_dir, int dims[3],
                        double params[16]);
/* Cell widths of one axis derived from a config text. */
int dugks_config_spacing(const char* config_text, const char* base_dir, int axis,
                         double* widths);
/* Initial f of the config's case in the dump layout (host part of init_dugks,
 * src/dugks.cu:221-281). */
int dugks_config_initial_state(const char* config_text, const char* base_dir, double* host_f);
/* Face coordinates 0..n of the tanh wall-normal mesh
 * (src/fortran/channel/run/mesh.py:21-25) for a channel of the given width. */
int dugks_mesh_tanh(int n, double alpha, double width, double* faces);
/* Raw dump I/O of host arrays (src/output.cu:22-89). */
int dugks_dump_write(const char* path, const double* f, long long count, int real_bytes);
int dugks_dump_read(const char* path, double* f, long long count, int real_bytes);

/* Post-processing of profiles-2.dat into VELMEAN.dat, the reference's
 * examples/channel/ProcessProfiles.py:17-79 (host only): the profile blocks
 * of n rows whose step exceeds ignore_before are averaged, then the two
 * channel halves are merged -- mean and variances symmetric, the shear
 * stress antisymmetric.  mean_col / stress_col select the profiles-2.dat
 * columns of the streamwise mean (0-2: u, v, w) and of the shear stress
 * (3: <u'w'>, 4: <u'v'>, 5: <v'w'>).  Rows written: index (1-based), mean,
 * stress, and the three variances. */
int dugks_velmean(const char* profiles_path, int n, int ignore_before, int mean_col, int stress_col,
                  const char* out_path);

#ifdef __cplusplus
}
#endif
#endif /* DUGKS_H */
