/*
 * libotfwi_pml.so -- MI355X (gfx950) split-field-PML acoustic propagator,
 * boundary-saving adjoint gradient and L2 / Wasserstein-2 misfits.
 *
 * Two C ABIs are exported:
 *
 *  1. The reference's own six symbols, with identical signatures and
 *     semantics, so the reference's ctypes wrapper (cuda_interface_pml.py:9-30)
 *     can load this library in place of simulator/bin/cuda_regression_PML_single.so.
 *
 *  2. A handle-based API with integer error codes (0 = success, otherwise a
 *     negative OTFWI_E_* code; otfwi_last_error() gives the message), batched
 *     shots, and device-resident misfits.  The package's Python wrapper uses
 *     this one.
 *
 * Conventions shared by both (the reference's): fields and models are
 * row-major (Z, X), x fastest; traces are (n_rx, T) per shot; point lists are
 * given as separate x and z arrays in the caller's order; units are normalized
 * (c = v dt/dx, constvec = {dx, dz, dt}, usually {1, 1, 1}).  Sources,
 * receivers and ring points must lie in the updated interior
 * 4 <= x < X-4, 4 <= z < Z-4 and must not repeat within one list.
 * All calls are synchronous with respect to the calling thread.
 */
#ifndef OTFWI_PML_H
#define OTFWI_PML_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------- */
/* 1. Reference ABI (regression_PML_single.cu).  Process-global state, FP32.  */
/* ---------------------------------------------------------------------- */

/* Host callback: computes the adjoint source from the simulated traces.
 * Both buffers hold n_rx*T floats and belong to the library.
 * Replaces the adj_func typedef, regression_PML_single.cu:19. */
typedef void (*otfwi_adj_fn)(float *simulated, float *adjoint_source);

/* regression_PML_single.cu:102-153.  Allocates the model, copies c^2, the
 * dense (Z,X) damping arrays, sources (n_src*T) and receivers, and hands back
 * a library-owned pinned buffer of n_rx*T floats through *rec.  `init` is
 * accepted and ignored (the reference never allocates its target). */
void init_memory_sim(int X, int Z, int T, float *cquad, float *const_vec, float *d_x,
                     float *d_z, int n_src, int *src_x, int *src_z, int n_rx, int *rx_x,
                     int *rx_z, float *src, float *init, float **rec);

/* regression_PML_single.cu:229-250.  Sets the boundary ring (n_rev points)
 * and hands back a library-owned pinned buffer of Z*X floats through *grad.
 * `obs` is not used (the reference copies it but never reads it). */
void init_memory_reg(float *obs, float **grad, int *rev_x, int *rev_z, int n_rev);

/* regression_PML_single.cu:253-257.  New c^2 (Z*X floats). */
void setCquad(float *cquad);

/* regression_PML_single.cu:618-702.  One forward run firing source
 * idx_source (-1: all sources at once); the traces land in the buffer given
 * by init_memory_sim (sample 0 is 0).  en_out (video frames) is ignored. */
void cuda_simulate(int en_out, int idx_source);

/* regression_PML_single.cu:704-846.  Forward run with ring save, host
 * callback for the adjoint source, reverse reconstruction + adjoint run +
 * imaging; the gradient lands in the buffer given by init_memory_reg.
 * *mse is set to 0, as in the reference. */
void cuda_grad_ext(float *mse, otfwi_adj_fn adj, int idx_source);

/* regression_PML_single.cu:39-43.  Stores a callback pointer; unused, as in
 * the reference. */
void def_python_adj(void *fn);

/* ---------------------------------------------------------------------- */
/* 2. Handle API.                                                          */
/* ---------------------------------------------------------------------- */

enum {
    OTFWI_OK = 0,
    OTFWI_E_ARG = -1,       /* bad argument (shape, range, missing setup) */
    OTFWI_E_HIP = -2,       /* HIP runtime error */
    OTFWI_E_MEM = -3,       /* device allocation failed */
    OTFWI_E_STATE = -4,     /* call order violated (e.g. adjoint before forward) */
    OTFWI_E_NUMERIC = -5    /* non-finite or unsupported numeric input */
};

/* Misfit kinds of the gradient, objective and misfit calls: L2 and W2
 * (adjoint_functions.L2, .wasserstein, adjoint_functions.py:8-11, 25-67) and
 * the squared-envelope misfit (adjoint_functions.quad_envelope,
 * adjoint_functions.py:14-22: E = |hilbert(sim)|^2 - |hilbert(obs)|^2,
 * misfit sum E^2, adjoint source sim E - Im(hilbert(E Im(hilbert(sim)))),
 * hilbert as scipy.signal.hilbert forms it; FP64 on the device). */
enum { OTFWI_MISFIT_L2 = 0, OTFWI_MISFIT_W2 = 1, OTFWI_MISFIT_ENV = 2 };

/* The four further Wasserstein-type misfits of adjoint_functions.py, same
 * values and adjoint sources (FP64 on the device, one wavefront per trace):
 * wasserstein_quad (:70-104, densities sim^2 and obs^2), wasserstein_experimental
 * (:107-141), wasserstein_pm (:144-223, positive parts; the reference builds its
 * negative-part CDFs from the positive parts) and wasserstein_yang (:226-266).
 * Accepted wherever a misfit kind is (gradient, objective, otfwi_misfit); not
 * by the exact-derivative calls. */
enum {
    OTFWI_MISFIT_W2_QUAD = 3,
    OTFWI_MISFIT_W2_EXPERIMENTAL = 4,
    OTFWI_MISFIT_W2_PM = 5,
    OTFWI_MISFIT_W2_YANG = 6
};

typedef struct otfwi_model otfwi_model;

/* Message of the last failed call on this thread ("" if none). */
const char *otfwi_last_error(void);

/* Library / device information: writes a short description (device name,
 * gfx arch, CU count) into buf. */
int otfwi_device_info(char *buf, int len);

/* New model of X x Z cells and T time samples.  precision: 4 = FP32 state and
 * arithmetic (production), 8 = FP64 (validation build).  constvec = {dx, dz,
 * dt}.  Replaces the allocation half of init_memory_sim
 * (regression_PML_single.cu:102-153). */
int otfwi_create(otfwi_model **out, int X, int Z, int T, int precision, const float *constvec);
int otfwi_destroy(otfwi_model *m);

/* Dense (Z,X) damping arrays as Cuda_interface_PML._calcPML produces them
 * (cuda_interface_pml.py:116-132).  d_x must be constant along z and d_z
 * along x (separable profiles); the kernels read the reference's one-cell
 * shifted values d_x[z][x+1] and d_z[z+1][x]. */
int otfwi_set_damping(otfwi_model *m, const float *d_x, const float *d_z);

/* c^2, (Z,X) FP32 host array (setCquad, regression_PML_single.cu:253-257). */
int otfwi_set_cquad(otfwi_model *m, const float *cquad);
/* c^2 from device memory (FP32, Z*X contiguous), read on the model's stream
 * (the caller's stream must have finished writing it).  In the FP32 build the
 * model array is formed on the device and only a validity flag comes back;
 * a value that is not finite or is negative fails the call (OTFWI_E_NUMERIC)
 * and leaves the model without c^2 until the next set. */
int otfwi_set_cquad_device(otfwi_model *m, const float *cquad_dev);

/* Sources: positions and wavelets (n_src*T FP32); the kernels inject the
 * running sum of each wavelet (somaFonteIntegral, regression_PML_single.cu:276-292). */
int otfwi_set_sources(otfwi_model *m, int n_src, const int *src_x, const int *src_z,
                      const float *wavelets);
int otfwi_set_receivers(otfwi_model *m, int n_rx, const int *rx_x, const int *rx_z);
/* Boundary-saving ring (init_regression, cuda_interface_pml.py:134-159). */
int otfwi_set_ring(otfwi_model *m, int n_ring, const int *ring_x, const int *ring_z);
/* Observed traces for n_shots shots, (n_shots, n_rx, T) FP32. */
int otfwi_set_observed(otfwi_model *m, int n_shots, const float *obs);

/* Tile heights of the forward / adjoint step kernels (16, 32 or 64; 0 keeps
 * the current choice) and the largest number of shots resident at once
 * (0 = all). */
int otfwi_set_tuning(otfwi_model *m, int th_fwd, int th_adj, int max_batch);
int otfwi_get_tuning(otfwi_model *m, int *th_fwd, int *th_adj, int *max_batch);

/* Launch style of the forward / adjoint step kernels: 0 = one work-group per
 * tile, 1 = persistent work-groups (a CU-filling grid that walks the tiles and
 * prefetches the next tile while computing the current one).  Both give
 * bitwise-identical results.  No reference counterpart (the reference
 * launches fixed 32x32 blocks, regression_PML_single.cu:120-135). */
int otfwi_set_launch(otfwi_model *m, int persistent_fwd, int persistent_adj);
int otfwi_get_launch(otfwi_model *m, int *persistent_fwd, int *persistent_adj);

/* Cache blocking of the step loops: the forward (adjoint) loop runs
 * steps_fwd (steps_adj) consecutive time steps on a group of shots_fwd
 * (shots_adj) shots before moving to the next group (0 shots = all at once,
 * 1 step = plain step-major order), so a group's state can stay in the
 * Infinity Cache between its steps.  Results are bitwise independent of the
 * blocking.  otfwi_get_blocking writes {steps_fwd, shots_fwd, steps_adj,
 * shots_adj}.  No reference counterpart (regression_PML_single.cu:745, 793
 * step every field of one shot per launch). */
int otfwi_set_blocking(otfwi_model *m, int steps_fwd, int shots_fwd, int steps_adj,
                       int shots_adj);
int otfwi_get_blocking(otfwi_model *m, int *out);

/* Forward time steps per launch: 1, 2, or 0 (default: 2 when a two-step
 * launch has at least 2 work-groups per CU, 1 otherwise).  With 2, the
 * forward loop advances pairs of steps in one launch of a z-marching kernel
 * that keeps the intermediate state on chip (the first steps of a range and
 * its last step stay single launches, so the buffers of the adjoint's seeds
 * are where a step-per-launch loop leaves them).  Results are bitwise
 * identical for every setting.  Pairs apply to forward loops without
 * illumination or cache blocking.  No reference counterpart
 * (regression_PML_single.cu:745 launches one forward step at a time). */
int otfwi_set_fwd_steps(otfwi_model *m, int steps);
int otfwi_get_fwd_steps(otfwi_model *m, int *steps);

/* Adjoint time steps per launch pair: 1, 2, or 0 (default: 2 when a two-step
 * launch has at least 1.5 work-groups per CU, 1 otherwise).  With 2, the
 * adjoint loop advances pairs of steps with two z-marching kernels: one
 * reconstructs the forward field over both steps (boundary ring written
 * back) and writes the Laplacian of both reconstructed states, the next
 * advances the adjoint field over both steps and adds both gradient terms.
 * Results are bitwise identical for every setting; two extra planes per shot
 * hold the Laplacians.  It applies to adjoint loops without cache blocking.  No
 * reference counterpart (regression_PML_single.cu:793 launches one adjoint
 * step at a time). */
int otfwi_set_adj_steps(otfwi_model *m, int steps);
int otfwi_get_adj_steps(otfwi_model *m, int *steps);

/* Work-group width of the two-step kernels: 256 or 384 threads (owning 224
 * or 352 columns each), or 0 (default, 256).  Results do not depend on it;
 * 384 was slower on MI355X at 1024^2.  No reference counterpart. */
int otfwi_set_pair_width(otfwi_model *m, int nw);

/* Streams the two-step phase of a loop runs on: 1, 2, 4, or 0 (default: 2
 * when each half of the batch's shots still launches 2 work-groups per CU,
 * else 1).  With k > 1 the shots are split into k groups of consecutive
 * shots, each advancing on the model's stream or one of the library's extra
 * streams, so each group's launches fill the others' launch tails; results
 * are bitwise identical.  With graphs on (otfwi_set_graph) each group's
 * two-step launches are captured on that group's own stream and replayed
 * there between a fork and a join recorded outside the graphs.  No reference
 * counterpart. */
int otfwi_set_pair_streams(otfwi_model *m, int k);

/* The steps per launch (1 or 2) the forward and adjoint loops use for a
 * batch of n_shots shots under the current settings (without illumination
 * or cache blocking, which keep the forward at one step).  No reference
 * counterpart. */
int otfwi_steps_per_launch(otfwi_model *m, int n_shots, int *fwd, int *adj);

/* How the forward and adjoint time loops of a batch of n_shots shots run
 * under the current settings: out[0] / out[2] = streams of the forward /
 * adjoint two-step phase (1 when the loop runs one step per launch or does
 * not split), out[1] / out[3] = 1 when that loop is replayed from captured
 * hipGraphs (one per stream group for a split two-step phase, whose one to
 * three single steps before and after it launch directly), 0 when it is
 * launched directly.  No reference counterpart. */
int otfwi_loop_modes(otfwi_model *m, int n_shots, int *out);

/* Replay the per-step launch loops as captured hipGraphs (default 1); a
 * graph is re-captured only when shots, tiles, launch style or device buffers
 * change.  A loop whose two-step phase splits over streams
 * (otfwi_set_pair_streams) is captured per stream group.  No reference counterpart (regression_PML_single.cu:661-678
 * launches and synchronizes four kernels per step). */
int otfwi_set_graph(otfwi_model *m, int use_graph);

/* Autotuner: times `steps` launches of every tile height and launch style of
 * the forward and adjoint step kernels on a batch of n_shots shots of this
 * model and keeps the fastest; then sets the forward and adjoint steps per
 * launch (otfwi_set_fwd_steps / otfwi_set_adj_steps) to 2 where the
 * two-step kernels take less time per step.  choice[0..3] (may be NULL) =
 * th_fwd, persistent_fwd, th_adj, persistent_adj; ms[0..1] (may be NULL) =
 * best average launch time of the one-step forward / adjoint kernels in
 * milliseconds.  No reference counterpart (fixed 32x32 blocks). */
int otfwi_autotune(otfwi_model *m, int n_shots, int steps, int *choice, float *ms);

/* Device time per launch (HIP events on the model's stream, mean of `reps`
 * launches after a warm-up launch) of the step kernels on a batch of n_shots
 * shots: ms[0] one forward step, ms[1] one adjoint step, ms[2] a forward
 * pair, ms[3] a reconstruction pair, ms[4] a lambda pair (-1 where a kernel
 * does not apply).  Leaves the batch's field buffers in an unspecified state.
 * No reference counterpart. */
int otfwi_kernel_times(otfwi_model *m, int n_shots, int reps, float *ms);

/* Device time of the last otfwi_gradient / otfwi_gradient_device call,
 * from HIP events on the model's stream: out[0] forward time loops (ms),
 * out[1] misfit kernels (ms), out[2] adjoint time loops (ms), out[3] number
 * of forward-step launches, out[4] number of adjoint-step launches.  No
 * reference counterpart (the reference prints wall time only,
 * cuda_interface_pml.py:233). */
int otfwi_get_timing(otfwi_model *m, double *out);

/* Forward runs of n_shots shots, shot k firing source shot_src[k] (-1: all
 * sources).  rec_out (may be NULL) receives (n_shots, n_rx, T) FP32 traces.
 * Replaces the per-shot loop of Cuda_interface_PML.simulate
 * (cuda_interface_pml.py:185-206) over cuda_simulate. */
int otfwi_forward(otfwi_model *m, int n_shots, const int *shot_src, float *rec_out);

/* Whole gradient with the misfit on the device: forward with ring save,
 * misfit + adjoint source per trace, reverse/adjoint run, FP64 mean over
 * shots.  misfit_out (may be NULL): per-shot misfit (n_shots FP64);
 * grad_out: (Z,X) FP32 mean gradient (may be NULL); rec_out may be NULL.
 * Replaces Cuda_interface_PML.calc_grad (cuda_interface_pml.py:208-236)
 * with adj_fun = adjoint_functions.L2 or .wasserstein. */
int otfwi_gradient(otfwi_model *m, int n_shots, const int *shot_src, int misfit_kind,
                   double *misfit_out, float *grad_out, float *rec_out);

/* Same, writing the FP32 (Z,X) mean gradient to device memory. */
int otfwi_gradient_device(otfwi_model *m, int n_shots, const int *shot_src, int misfit_kind,
                          double *misfit_out, float *grad_out_dev);

/* Gradients of several independent models (uneven grids, T and shot
 * counts) at once: n_workers host threads (<= 0: 4) take the models in
 * longest-processing-time order (cost X*Z*T*n_shots) and run each one's
 * otfwi_gradient on that model's own stream, so kernels of different models
 * share the GPU.  Per model i: n_shots[i], shot_src[i] (may be NULL = all
 * fire -1), misfit_kind[i], misfit_out[i] / grad_out[i] (host, may be NULL).
 * Results equal those of separate otfwi_gradient calls.  A model must not be
 * used by another thread meanwhile.  No reference counterpart (one model
 * per process, regression_PML_single.cu:21-31). */
int otfwi_gradient_multi(int n_models, otfwi_model *const *models, const int *n_shots,
                         const int *const *shot_src, const int *misfit_kind,
                         double *const *misfit_out, float *const *grad_out, int n_workers);

/* Gradients of several independent FP32 models (uneven grids, T and shot
 * counts) with every two-step launch covering all of them: per pair index
 * one launch whose work-groups look up their model in a device table, over
 * the models whose loops still have that pair; the forward and the adjoint
 * two-step phases are each captured into one hipGraph on the first model's
 * stream, and each model's single steps, misfit and reductions run on its
 * own stream, joined with events.  The shots of every model are split into
 * two stream groups, each group's launches captured on its own stream, when
 * half of the first launch still makes 2 work-groups per CU, or as the
 * first model's otfwi_set_pair_streams says (1: one group; 2 or 4: two).
 * Arguments as otfwi_gradient_multi.  Every
 * model must run two steps per launch of 256 threads for its batch (the
 * default at a few work-groups per CU and more), hold its shots in one
 * batch and collect no illumination; otherwise OTFWI_E_ARG names the model.
 * Results equal those of separate otfwi_gradient calls.  No reference
 * counterpart (one model per process, regression_PML_single.cu:21-31). */
int otfwi_gradient_fused(int n_models, otfwi_model *const *models, const int *n_shots,
                         const int *const *shot_src, const int *misfit_kind,
                         double *const *misfit_out, float *const *grad_out);

/* Two-stage gradient for a host-side adjoint source (an arbitrary adj_fun,
 * cuda_interface_pml.py:238-249): otfwi_forward_for_grad runs the forward
 * with ring save and returns the traces; otfwi_adjoint_from_host takes the
 * (n_shots, n_rx, T) FP32 adjoint source and returns the mean gradient.
 * At most max_batch shots (otfwi_set_tuning) per pair of calls; any other
 * run on the model in between invalidates the first stage (the second then
 * fails with OTFWI_E_STATE). */
int otfwi_forward_for_grad(otfwi_model *m, int n_shots, const int *shot_src, float *rec_out);
int otfwi_adjoint_from_host(otfwi_model *m, const float *adj_src, float *grad_out);

/* Exact discrete adjoint gradient with snapshot checkpointing: the
 * transpose of the forward recursion (not the reference's imaging
 * condition), for J = mean over shots of the misfit value (L2: sum of
 * (sim - obs)^2; W2: the reference W2 value, whose exact trace derivative is
 * taken).  The forward state is saved every snap_every steps (<= 0:
 * round(sqrt(T-1))) and each segment's A fields are recomputed before the
 * adjoint runs back through it.  grad_out (FP32) and/or grad64_out (FP64)
 * receive dJ/d(c^2) on the (Z,X) grid; misfit_out (may be NULL) the per-shot
 * misfit.  No reference counterpart (the reference reconstructs the forward
 * field by boundary saving, regression_PML_single.cu:758, 805, and images
 * with the 9-point Laplacian, :390-445). */
int otfwi_gradient_exact(otfwi_model *m, int n_shots, const int *shot_src, int misfit_kind,
                         int snap_every, double *misfit_out, float *grad_out, double *grad64_out);

/* Same, writing the FP32 (Z,X) mean gradient to device memory (the FP64
 * accumulation scaled and rounded once, as grad_out of otfwi_gradient_exact),
 * for callers whose model and gradient stay on the GPU (the autograd
 * wrapper).  Replaces the same reference interfaces as otfwi_gradient_exact. */
int otfwi_gradient_exact_device(otfwi_model *m, int n_shots, const int *shot_src, int misfit_kind,
                                int snap_every, double *misfit_out, float *grad_out_dev);

/* c^2 in FP64 (for the FP64 build and finite-difference checks): the model
 * is formed from these values directly instead of their FP32 rounding.
 * Replaces setCquad (regression_PML_single.cu:253-257) when FP64 input is
 * wanted. */
int otfwi_set_cquad64(otfwi_model *m, const double *cquad);

/* Misfit of the current model without the adjoint run: forward runs (no
 * ring save) and the device misfit; misfit_out receives n_shots FP64
 * per-shot sums.  Used by the misfit-decrease acceptance test of the
 * inversion driver; the reference evaluates the objective only through the
 * full gradient (regression.py:45-49 over cuda_interface_pml.py:208-236). */
int otfwi_objective(otfwi_model *m, int n_shots, const int *shot_src, int misfit_kind,
                    double *misfit_out);

/* Source-side illumination collected by the next gradient calls: every
 * stride-th forward step adds sum over the call's shots of P(t)^2 (times
 * stride) to a (Z,X) FP32 map (stride 0 = off).  otfwi_get_illum copies the
 * map of the last gradient call to host memory, or to device memory when
 * out_on_device is nonzero.  No reference counterpart (the reference
 * preconditions with a fixed mask and scale, regression.py:45-49). */
int otfwi_set_illum(otfwi_model *m, int stride);
int otfwi_get_illum(otfwi_model *m, float *out, int out_on_device);

/* Stand-alone misfit on host arrays (n_traces of T samples, FP64 in and
 * out, as adj_wrapper hands them over): per-trace misfit and adjoint source.
 * Replaces adjoint_functions.L2 / .wasserstein / .quad_envelope
 * (adjoint_functions.py:8-11, 25-67, 14-22). */
int otfwi_misfit(int misfit_kind, int n_traces, int T, const double *sim, const double *obs,
                 double *misfit_per_trace, double *adj_out);

/* Exact derivative of the misfit value with respect to the simulated traces,
 * on host arrays (FP64): L2 2 (sim - obs); W2 the derivative of the
 * reference W2 value (adjoint_functions.py:25-67) through its interpolated
 * inverse CDF, trapezoid CDF and normalization (needs sim + c >= 0).  The
 * trace derivative used by otfwi_gradient_exact (L2 and W2 only); no
 * reference counterpart. */
int otfwi_misfit_exact(int misfit_kind, int n_traces, int T, const double *sim,
                       const double *obs, double *misfit_per_trace, double *rho_out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): otfwi_misfit on device arrays that allocates nothing.
 * sim, obs, adj_out: n_traces x T FP64 device arrays; misfit_out: n_traces
 * FP64 device array; stream: the hipStream_t the work is enqueued on (NULL =
 * the null stream); the call returns after that stream has finished it.
 * Kinds L2 and W2 (the envelope misfit is refused: use otfwi_misfit).  The
 * W2 kernels' work arrays live in a fixed 32 MiB device array of the library
 * (one per GPU, held until the process ends): W2 needs T <= 10240 and runs
 * the traces in chunks of (32 MiB - 256) / (24 T + 4) at a time (349 traces
 * at T = 4000).  Every call on one GPU shares that array, so calls that
 * overlap on one GPU (several host threads or streams) race; serialize them. */
int otfwi_misfit_noalloc(int misfit_kind, int n_traces, int T, const double *sim,
                         const double *obs, double *misfit_out, double *adj_out, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* OTFWI_PML_H */
