/*
 * dugks_oracle.h -- TEST INFRASTRUCTURE, not product code.
 *
 * FP64 CPU restatement of the reference DUGKS step (Kairzhan/DUGKS-GPU),
 * used only as the correctness checker by tests/, by __graft_entry__.smoke()
 * and by bench.py's cpu_baseline leg.  The product library never links it.
 *
 * Layout used by every array argument below (the reference restart-dump
 * layout, src/output.cu:22-41): f[((k*ny + j)*nx + i)*19 + ip], ip fastest,
 * interior cells only.  Per-cell scalar fields use [(k*ny + j)*nx + i].
 */
#ifndef DUGKS_ORACLE_H
#define DUGKS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_ctx oracle_ctx;

/* recon: 0 = trilinear faces (calc_fx_tri, src/dugks.cu:715-959, the launched
 *            kernel), 1 = midpoint + central gradient (calc_fx, :421-649).
 * periodic[a]: 1 = periodic axis, 0 = no-slip walls at both ends of axis a.
 * spacing[a]: NULL for a uniform axis of width width[a] (d = width/n), or an
 *            array of n[a] cell widths (a stretched wall-normal axis, the
 *            non-uniform variant of src/fortran/channel/src/dugks.cuf:538-716).
 * dt, visc: time step and kinematic viscosity (tau = visc/RT).
 */
oracle_ctx* oracle_create(int nx, int ny, int nz, const int periodic[3],
                          const double width[3], const double* spacing_x,
                          const double* spacing_y, const double* spacing_z,
                          double dt, double visc, const double famb[3], int recon);
void oracle_destroy(oracle_ctx*);
void oracle_set_threads(int nthreads);

void oracle_set_f(oracle_ctx*, const double* f);
void oracle_get_f(const oracle_ctx*, double* f);
/* Per-cell body force used at cell centres (FORCEX/Y/Z of calc_fb1).  When it
 * is never set, every cell carries the ambient force (calc_force,
 * src/dugks.cu:1066-1084). */
void oracle_set_cell_force(oracle_ctx*, const double* fx, const double* fy, const double* fz);
void oracle_set_ambient_force(oracle_ctx*, const double famb[3]);

void oracle_step(oracle_ctx*, int nsteps);

/* calc_macro (src/dugks.cu:1021-1064): rho = sum f, u = sum c f / rho + h F_amb */
void oracle_macro(const oracle_ctx*, double* rho, double* ux, double* uy, double* uz);

/* Coefficients the oracle derived from (dt, visc), in this order:
 * tau, hdt, bw0, bw1, bw2, fw0, fw1, fw2 (src/dugks.cu:102-125). */
void oracle_coefficients(const oracle_ctx*, double out[8]);

/* Face reconstruction on an explicit f-bar field (interior cells, ddf
 * layout): fills the ghost layer exactly as the step does, then writes the
 * 6 x 19 reconstructed face values of every interior cell (before wall
 * bounce-back and before the face collision) to faces[cell*114 + s*19 + ip],
 * s = 0:-x 1:+x 2:-y 3:+y 4:-z 5:+z.  Used to check reconstruction alone. */
void oracle_reconstruct(const oracle_ctx*, const double* fbar, double* faces);

/* Discrete adjoint of one step, linearised about the state currently held
 * (f^n): given lam_next = dJ/df^{n+1} (ddf layout) returns
 * lam_prev = (df^{n+1}/df^n)^T lam_next and accumulates
 * grad_force[c*3 + a] += (df^{n+1}/dF_cell[a])^T lam_next, where F_cell is
 * the per-cell centre force.  The held state is not advanced. */
void oracle_adjoint_step(const oracle_ctx*, const double* lam_next, double* lam_prev,
                         double* grad_force);

/* Streak forcing of the reference (calc_perturb_force, src/dugks.cu:1086-1126):
 * total centre force per interior cell at a time step, [k][j][i] arrays;
 * ambient outside the forced band 3 < i < nx/4 + 2 (1-based). */
void oracle_perturb_force(int nx, int ny, int nz, const double width[3], const double d[3], double dt,
                          double visc, int step, const double famb[3], double* fx, double* fy, double* fz);

#ifdef __cplusplus
}
#endif
#endif
