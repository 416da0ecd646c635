/* libotfwi_cpml.so: the second-order C-PML pressure scheme
 * (simulator/src/regression_CPML.cu, prec_deriv = 4) on gfx950, forward only.
 *
 * Section 1 is the reference's own C ABI (regression_CPML.cu:349-501) so a
 * caller of cuda_regression_CPML.so can load this library instead.  Section 2
 * is a handle API that runs a batch of shots per launch and reports errors. */
#ifndef OTFWI_CPML_H
#define OTFWI_CPML_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------- */
/* 1. Reference ABI (one process-wide model, void, errors on stderr).      */
/* ---------------------------------------------------------------------- */

typedef double otfwi_cpml_adj_fn(float *, float *, float *, float *, int, int);

/* regression_CPML.cu:349-426.  Grid x by z, t samples, ns sources, nm
 * receivers, ppe pixels per element (1: cuda_simulate2's idx selects one
 * source, else every source fires), positions, dx, dz, dt, dt^2, the source
 * scale (4 pi in the reference's callers), c^2, kappa, rho and rho at the half
 * points in x and z ([z][x] each), the C-PML profiles d, K, alpha, a, b and
 * their half-point versions along x ([x]) and z ([z]), the wavelets
 * (ns x t), the initial state (2 planes: P at t = 0, 1) and the address of
 * the recording (nm x t, library-owned pinned memory).  rho, d and alpha are
 * accepted and unused, as in the reference's kernels. */
void init_memory_sim(int x, int z, int t, int ns, int nm, int ppe, int *ps_x, int *ps_z,
                     int *pm_x, int *pm_z, float dx_h, float dz_h, float dt_h, float dt2_h,
                     float qtpi_h, float *cq, float *kp_unr, float *rho_h, float *rho_hx_h,
                     float *rho_hz_h, float *d_x_h, float *K_x_h, float *alpha_x_h,
                     float *a_x_h, float *b_x_h, float *d_x_half_h, float *K_x_half_h,
                     float *alpha_x_half_h, float *a_x_half_h, float *b_x_half_h, float *d_z_h,
                     float *K_z_h, float *alpha_z_h, float *a_z_h, float *b_z_h,
                     float *d_z_half_h, float *K_z_half_h, float *alpha_z_half_h,
                     float *a_z_half_h, float *b_z_half_h, float *src, float *init, float **rec);
/* regression_CPML.cu:429-433: new c^2 ([z][x]). */
void setCquad(float *cq);
/* regression_CPML.cu:436-443: new source positions (the wavelets stay). */
void set_source(int ns, int *ps_x, int *ps_z);
/* regression_CPML.cu:446-501: one run into the recording; en_out (video
 * frames) is accepted and ignored. */
void cuda_simulate2(int en_out, int idx_source);
/* regression_CPML.cu:40-45: stored, unused (the scheme has no adjoint). */
void def_python_adj(otfwi_cpml_adj_fn *func);

/* ---------------------------------------------------------------------- */
/* 2. Handle API (0 = success, < 0 = error; message in last_error).         */
/* ---------------------------------------------------------------------- */

typedef struct otfwi_cpml otfwi_cpml;

const char *otfwi_cpml_last_error(void);

/* consts = {dx, dz, dt, dt^2, source scale}; fields = {c^2, kappa, rho_half_x,
 * rho_half_z} ([z][x] each); prof = {K_x, a_x, b_x, K_x_half, a_x_half,
 * b_x_half} ([x] each) then the same six along z ([z]); initial: 2 planes or
 * NULL (zero).  No reference counterpart (replaces the allocation half of
 * init_memory_sim). */
int otfwi_cpml_create(otfwi_cpml **out, int X, int Z, int T, int ppe, const float *consts,
                      const float *const *fields, const float *const *prof, const float *initial);
int otfwi_cpml_destroy(otfwi_cpml *h);
int otfwi_cpml_set_cquad(otfwi_cpml *h, const float *cq);
/* Sources (positions, wavelets n_src x T) and receivers; every point must lie
 * in the grid. */
int otfwi_cpml_set_points(otfwi_cpml *h, int n_src, const int *sx, const int *sz,
                          const float *src, int n_rx, const int *rx, const int *rz);
/* n_shots runs advanced together: shot s fires source shot_idx[s] (-1: all;
 * every source when ppe != 1); rec_out: n_shots x n_rx x T (host). */
int otfwi_cpml_forward(otfwi_cpml *h, int n_shots, const int *shot_idx, float *rec_out);

#ifdef __cplusplus
}
#endif
#endif /* OTFWI_CPML_H */
