/*
 * amb.h -- C ABI of the MI355X Active Model B+ integrator (libamb.so).
 *
 * Plain C types only: doubles, 64-bit sizes and opaque handles.  Every call
 * returns AMB_OK (0) or an error code; amb_last_error() gives a thread-local
 * message for the last failure on the calling thread.  No C++ exception crosses
 * this boundary.  A handle must not be used from two threads at once; separate
 * handles are independent (all parameters travel as kernel arguments, there is
 * no process-global device state).
 *
 * Field layout everywhere: row-major (ny, nx) float64, linear index col + row*nx,
 * periodic in both axes.  Vector quantities are 2*ny*nx doubles laid out
 * [component y | component x] (component 0 = rows, the reference's convention).
 */
#ifndef AMB_H
#define AMB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes -------------------------------------------------------- */
#define AMB_OK 0
#define AMB_E_ARG 1        /* invalid argument (shape, null pointer, parameter) */
#define AMB_E_HIP 2        /* HIP runtime error (message carries hipGetErrorString) */
#define AMB_E_NONFINITE 3  /* field became NaN/Inf during run (reference: CudaError) */
#define AMB_E_NODEVICE 4   /* no HIP device visible */
#define AMB_E_STATE 5      /* call not valid in the handle's current state */

/* ---- parameter structs --------------------------------------------------- */

/* Model parameters a, b, c, kappa, lambda, zeta, temperature.
 * Replaces ActiveModelBPlusParams (src/parameters.h:53-77). */
typedef struct {
    double a, b, c, kappa, lamb, zeta, T;
} amb_model;

/* Time step and grid spacings.  Replaces HostStencilParams (src/parameters.h:23-41). */
typedef struct {
    double dt, dx, dy;
} amb_stencil;

#define AMB_F64 64
#define AMB_F32 32

/* Options with no reference counterpart (the reference is FP64 only and seeds its
 * noise from std::random_device, src/integrator.cu:320-323).  Passing NULL means
 * {precision = AMB_F64, seed = 0, timestep = 0, tile = 0}. */
typedef struct {
    int precision;     /* AMB_F64 or AMB_F32: storage and arithmetic of the field */
    uint64_t seed;     /* Philox key of the conserved noise */
    int64_t timestep;  /* initial step counter (exact resume: field, timestep, seed) */
    int tile;          /* 0 = default tile; k > 0 selects tile index k-1 (see amb_tile_count) */
} amb_options;

typedef struct amb_handle amb_handle;

/* ---- single field: the drop-in for class Integrator ---------------------- */

/* Copies `field` (ny*nx doubles) to the device.
 * Replaces Integrator::Integrator(const HostFieldRef&, Stencil, Model)
 * (src/integrator.h:34, src/integrator.cu:288-333). */
int amb_create(const double* field, int64_t ny, int64_t nx, const amb_stencil* stencil,
               const amb_model* model, const amb_options* options, amb_handle** out);

/* Frees every device resource.  Replaces Integrator::~Integrator (src/integrator.cu:353-366).
 * NULL is accepted. */
int amb_destroy(amb_handle* h);

/* Advances nsteps explicit-Euler steps (nsteps >= 0), then checks the field for
 * NaN/Inf and returns AMB_E_NONFINITE if any site is not finite.
 * Replaces Integrator::run(int) (src/integrator.cu:497-538). */
int amb_run(amb_handle* h, int64_t nsteps);

/* Copies the current field (ny*nx doubles) to `out`.
 * Replaces Integrator::get_field (src/integrator.cu:378-383). */
int amb_get_field(amb_handle* h, double* out);

/* Observables at the current step, i.e. the quantities the next step will use
 * (noise included), as the reference's getters define them
 * (src/integrator.cu:385-465, lazily computed at src/integrator.cu:477-495).
 * Scalars fill ny*nx doubles, currents 2*ny*nx doubles. */
#define AMB_OBS_MU 0                 /* get_chemical_potential (scalar) */
#define AMB_OBS_MU_PASSIVE 1         /* get_passive_chemical_potential (scalar) */
#define AMB_OBS_MU_ACTIVE 2          /* get_active_chemical_potential (scalar) */
#define AMB_OBS_J 3                  /* get_current */
#define AMB_OBS_J_DETERMINISTIC 4    /* get_deterministic_current */
#define AMB_OBS_J_PASSIVE 5          /* get_passive_current */
#define AMB_OBS_J_ACTIVE 6           /* get_active_current */
#define AMB_OBS_J_LAMBDA 7           /* get_lambda_current */
#define AMB_OBS_J_CIRCULATING 8      /* get_circulating_current */
#define AMB_OBS_J_RANDOM 9           /* get_random_current */
int amb_get_observable(amb_handle* h, int which, double* out);

/* Replaces the current field (ny*nx doubles); the step counter is unchanged.  No
 * reference counterpart (the reference can only construct a new Integrator). */
int amb_set_field(amb_handle* h, const double* field);

/* Step counter and parameters.  Replace Integrator::get_timestep, get_stencil,
 * get_model (src/integrator.h:38-39,52-55). */
int amb_get_timestep(const amb_handle* h, int64_t* timestep);
int amb_get_params(const amb_handle* h, amb_stencil* stencil, amb_model* model);
int amb_get_shape(const amb_handle* h, int64_t* ny, int64_t* nx);
int amb_get_options(const amb_handle* h, amb_options* options);

/* Sum of the field accumulated in FP64 and the number of non-finite sites.
 * Replaces kernel::check_finite (src/integrator.cu:265-272) and adds the mass. */
int amb_mass(amb_handle* h, double* sum, int64_t* nonfinite);

/* The same two numbers as found by the last amb_run / amb_run_timed: with nsteps > 0
 * they come from the run's last step launch, where every tile adds up its new values
 * (FP64, fixed order) and counts the non-finite ones; one small fixed-order kernel then
 * sums the tiles' partials.  Replaces kernel::check_finite's separate pass over the
 * field after the step loop (src/integrator.cu:264-272,525-537). */
int amb_last_mass(amb_handle* h, double* sum, int64_t* nonfinite);

/* Test-only: explicit unscaled normals for the next `nsteps` steps
 * (nsteps * 2 * ny * nx doubles, step-major, each step [xi_y | xi_x]) replace the
 * Philox stream until they are used up; the caller's buffer is copied.  Lets a
 * CPU scheme and the GPU read one shared noise buffer.  No reference counterpart. */
int amb_set_noise(amb_handle* h, const double* noise, int64_t nsteps);

/* Like amb_run, and returns the device time of the steps, measured with HIP
 * events recorded on the handle's own stream (the stream the step kernels are
 * launched on).  No reference counterpart; used by the benchmark. */
int amb_run_timed(amb_handle* h, int64_t nsteps, double* device_ms);

/* Tile shapes compiled for the step kernel of a precision (autotuner search space). */
int amb_tile_count(int precision, int* count);
int amb_tile_shape(int precision, int index, int* tile_cols, int* tile_rows, int* threads);
int amb_set_tile(amb_handle* h, int index);

/* Autotuner for one field: times `reps` steps of every compiled tile shape on a
 * copy of the current state (field and step counter untouched), selects the
 * fastest for this handle, records it in the tuning table and returns its index in
 * `chosen` (nullable).  reps < 0: take the tuning table's entry for this precision,
 * grid and noise mode when there is one (no timing), else time |reps| steps. */
int amb_autotune(amb_handle* h, int reps, int* chosen);

/* Tuning table (process-wide, no reference counterpart): the autotuners' choices per
 * (precision, ny, nx, noisy).  amb_create and amb_batch_create start from its entries
 * (tile option 0), before the occupancy rule.  Entries are plain data, so a caller can
 * save them and load them into a later process (activemodelbplus.tuning does, as JSON).
 * amb_tuning_get gives tile = -1 when there is no entry. */
int amb_tuning_set(int precision, int64_t ny, int64_t nx, int noisy, int tile);
int amb_tuning_get(int precision, int64_t ny, int64_t nx, int noisy, int* tile);
int amb_tuning_count(int64_t* count);
int amb_tuning_entry(int64_t index, int* precision, int64_t* ny, int64_t* nx, int* noisy, int* tile);
int amb_tuning_clear(void);

/* Thread-local message of the last error on this thread ("" if none). */
const char* amb_last_error(void);

/* Number of visible HIP devices (0 without a GPU; never fails). */
int amb_device_count(int* count);

/* ---- allocation-free run ---------------------------------------------------- */

/* amb_run's steps on caller-owned device memory, allocating nothing: no counterpart in
 * the reference; it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  `field` and
 * `work` are two distinct device buffers of ny*nx elements of `precision` (float for
 * AMB_F32, double for AMB_F64); the field advances nsteps explicit-Euler steps with the
 * noise counter starting at `timestep` and key `seed` (exactly amb_run's results for a
 * handle with the same options) and ends in `field`; `work` is overwritten.  `stream`
 * is a hipStream_t (NULL: the default stream); the call returns after synchronising it.
 * mass / nonfinite (nullable) receive the FP64 sum and the non-finite count of the new
 * field; AMB_E_NONFINITE is returned when that count is not zero.  The library's
 * scratch (step counter and the fused check's per-tile partials) is one static device
 * array of 1 MiB + 256 B per GPU, shared by every call on that GPU: fields with more
 * than 65536 tiles of the default tile shape are rejected (AMB_E_ARG), and calls that
 * overlap on one GPU race. */
int amb_run_noalloc(void* field, void* work, int64_t ny, int64_t nx, int precision, const amb_stencil* stencil,
                    const amb_model* model, uint64_t seed, int64_t timestep, int64_t nsteps, void* stream,
                    double* mass, int64_t* nonfinite);

/* ---- batched ensemble (no reference counterpart) ------------------------- */

/* One independent realisation: its own grid, stencil, model and noise seed. */
typedef struct {
    int64_t ny, nx;
    amb_stencil stencil;
    amb_model model;
    uint64_t seed;
} amb_realisation;

typedef struct amb_batch amb_batch;

/* fields[i] points at ny_i*nx_i doubles.  precision: AMB_F32 or AMB_F64. */
int amb_batch_create(const amb_realisation* desc, int64_t count, const double* const* fields,
                     int precision, amb_batch** out);
int amb_batch_destroy(amb_batch* b);
/* Advances every realisation by nsteps in hipGraph-captured launches. */
int amb_batch_run(amb_batch* b, int64_t nsteps);
int amb_batch_run_timed(amb_batch* b, int64_t nsteps, double* device_ms);
int amb_batch_get_field(amb_batch* b, int64_t index, double* out);
int amb_batch_mass(amb_batch* b, int64_t index, double* sum, int64_t* nonfinite);
/* Total FP64 sum over all realisations and non-finite count found by the last
 * amb_batch_run, fused into its last step launch as for amb_last_mass. */
int amb_batch_last_mass(amb_batch* b, double* sum, int64_t* nonfinite);

/* Tile shape (index from amb_tile_count) per realisation; realisations sharing a
 * shape share one launch per step. */
int amb_batch_set_tiles(amb_batch* b, const int* tile_per_realisation);
int amb_batch_get_tiles(amb_batch* b, int* tile_per_realisation);

/* Resident members: with enable != 0 (the default) every realisation whose two field
 * buffers and one tile's scratch fit in 80 KiB of LDS (FP32 up to 64x64, FP64 up to about
 * 44x44 sites) is advanced by one workgroup that keeps it in LDS: one launch per run
 * chunk advances all of them by many steps, on a graph branch beside the tiled members'
 * per-step launches, with no inter-workgroup synchronisation.  Results are the same bits
 * either way.  count receives the number of resident members. */
int amb_batch_set_resident(amb_batch* b, int enable);
int amb_batch_resident_count(amb_batch* b, int64_t* count);

/* Autotuner: for every distinct grid size and noise mode, times `reps` steps of every
 * compiled tile shape and keeps the fastest (state and step counter are restored),
 * recording it in the tuning table; reps < 0 takes the table's entries where they
 * exist.  `chosen` (nullable) receives the tile index per realisation. */
int amb_batch_autotune(amb_batch* b, int reps, int* chosen);

/* ---- structure factor ------------------------------------------------------ */

/* Radially binned structure factor of the handle's current field, on the device;
 * replaces the host analysis of examples/structureFactor.ipynb (cell 7):
 * phi_k = FFT(phi - mean) dx dy, |k| = 2 pi sqrt(fx^2 + fy^2) on numpy's fftfreq
 * grid.  For each digitize index i = #{edges < |k|} (numpy.digitize(|k|, edges,
 * right=True)), i = 0 .. nedges: sums[i] = sum of |phi_k|^2 (FP64) and counts[i] =
 * the number of modes.  edges increasing, 1 <= nedges <= 1024. */
int amb_structure_factor(amb_handle* h, const double* edges, int nedges, double* sums, int64_t* counts);

/* ---- learned flux-form surrogate (no reference counterpart) --------------- */

/* A periodic U-Net (3 levels, 32/64/128 channels, 3x3 kernels, GELU; BF16 weights
 * and activations on matrix cores, FP32 accumulation) mapping (phi, lambda, zeta,
 * phi0) to a current J (its FP32 head scaled by 1e-3; phi enters as bfloat16 plus its
 * rounding remainder, ~16 bits); one surrogate step is phi <- phi - div J with the
 * backward-difference divergence, so sum(phi) is conserved exactly.  Grid sides must
 * be multiples of 64.  Parameters live in one FP32 vector (amb_surrogate_param_count
 * floats); training uses Adam with optional gradient-norm clipping. */
typedef struct amb_surrogate amb_surrogate;

/* max_batch in [1, 256] with max_batch * ny * nx <= 2^26. */
int amb_surrogate_create(int64_t ny, int64_t nx, int max_batch, uint64_t seed, amb_surrogate** out);
int amb_surrogate_destroy(amb_surrogate* s);
int amb_surrogate_param_count(const amb_surrogate* s, int64_t* n);
int amb_surrogate_get_params(amb_surrogate* s, float* out);
int amb_surrogate_set_params(amb_surrogate* s, const float* in);

/* Loss (mean over the batch of |phi - div J - target|^2 / (|target - phi|^2 + 1e-8 ny nx),
 * the floor keeping nearly stationary samples from dominating) and its
 * gradient with respect to the parameter vector (grad nullable), for host inputs of
 * batch fields (ny*nx each) and batch parameter triples (lambda, zeta, phi0). */
int amb_surrogate_loss_grad(amb_surrogate* s, const double* phi, const double* params, const double* target,
                            int batch, double* loss, float* grad);

/* Training pairs (count fields each of phi_in and phi_out, FP32, and count triples),
 * copied to the device; then `iterations` Adam steps on random mini-batches. */
int amb_surrogate_set_dataset(amb_surrogate* s, const float* phi_in, const float* phi_out, const float* params,
                              int64_t count);
int amb_surrogate_train(amb_surrogate* s, int64_t iterations, int batch, double lr, double clip, double* mean_loss);

/* Training-input noise for amb_surrogate_train: each sample's input field is
 * perturbed by the backward-difference divergence of Gaussian noise (so its sum is
 * unchanged) of about rel times the sample's RMS increment, drawn from a
 * counter-based stream keyed by seed; the target is unchanged, so the network
 * learns to damp its own errors.  rel = 0 switches it off (the default). */
int amb_surrogate_set_train_noise(amb_surrogate* s, double rel, uint64_t seed);

/* One surrogate step on host fields (current: nullable, 2*ny*nx per field [y, x]). */
int amb_surrogate_predict(amb_surrogate* s, const double* phi, const double* params, int batch, double* phi_next,
                          double* current);

/* nsteps surrogate steps in place on an FP32 integrator's field (same grid); its step
 * counter advances by nsteps * euler_steps_per_step.  device_ms: HIP-event time. */
int amb_surrogate_advance(amb_surrogate* s, amb_handle* h, int64_t nsteps, const double* params3,
                          int64_t euler_steps_per_step, double* device_ms);

/* Audit of one surrogate step against the explicit-Euler path, on the device.  From
 * the current field of `main`, `judge` (a second FP32 integrator on the same grid)
 * runs euler_steps Euler steps and the surrogate takes one step on `main`;
 * *rel_err = |phi_S - phi_E|_2 / |phi_E|_2 (FP64 sums).  If rel_err <= tol `main`
 * keeps the surrogate field (*accepted = 1), otherwise, and whenever either field is
 * not finite, it takes the Euler field (*accepted = 0).  main's step counter advances
 * by euler_steps either way. */
int amb_surrogate_audit(amb_surrogate* s, amb_handle* main, amb_handle* judge, const double* params3,
                        int64_t euler_steps, double tol, double* rel_err, int* accepted);

/* The same audit overlapped with further work.  _begin enqueues the judge's
 * euler_steps Euler steps on the judge's stream, takes the surrogate step on `main`
 * (kept aside) and advances main's step counter; the caller may then advance `main`
 * further (speculatively) while the judge runs.  _end waits for the judge and
 * compares; if rel_err > tol (or not finite) `main` is rolled back to the Euler field
 * at the audited step (counter = start + euler_steps), discarding the later steps.
 * One audit may be pending per surrogate; neither integrator may be destroyed
 * while it is. */
int amb_surrogate_audit_begin(amb_surrogate* s, amb_handle* main, amb_handle* judge, const double* params3,
                              int64_t euler_steps);
int amb_surrogate_audit_end(amb_surrogate* s, double tol, double* rel_err, int* accepted);

#ifdef __cplusplus
}
#endif

#endif /* AMB_H */
