/*
 * halhip.h -- C ABI of the MI355X (gfx950) compute backend for HAL-CPP's hot
 * path: explicit 5-point diffusion, ADI diffusion and cell-bucket agent forces,
 * plus batched wells and adjoint ADI.
 *
 * Conventions
 *   - Fields are flat row-major arrays, index x*yDim + y (HAL's layout,
 *     include/Hal/Grids/GridGeometry2D.h:12).
 *   - Every function returns an int status (HALHIP_OK == 0).  On failure the
 *     calling thread's halhip_last_error() holds a message.  No C++ exception
 *     crosses this boundary.
 *   - Handles are opaque.  One thread per handle; calls on a handle are
 *     synchronous (they return after the device work finished) unless the name
 *     says "_async".
 *   - "_f64" entry points take the reference interfaces' FP64 host arrays and
 *     compute in FP32 on the device, exactly like the reference's GPU adapters.
 *     Resident entry points keep FP32 state on the device between calls.
 */
#ifndef HALHIP_H
#define HALHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define HALHIP_OK 0
#define HALHIP_ERR_INVALID_ARGUMENT 1 /* maps to std::invalid_argument */
#define HALHIP_ERR_NO_DEVICE 2        /* no usable GPU: callers may skip */
#define HALHIP_ERR_OUT_OF_MEMORY 3    /* device allocation failed */
#define HALHIP_ERR_HIP 4              /* any other HIP runtime error */
#define HALHIP_ERR_UNSUPPORTED 5      /* shape outside the documented limits */

/* Axis boundary kinds.  A wrapped axis is periodic regardless of bc. */
#define HALHIP_BC_NONE 0      /* zero flux on non-wrapped axes */
#define HALHIP_BC_DIRICHLET 1 /* ghost-cell Dirichlet on non-wrapped axes */

/* Dirichlet boundary description.  The reference's BCFn(x, y) is only ever
 * evaluated at ghost cells, so four edge arrays capture any BCFn:
 *   xlo[y] = bc(-1, y), xhi[y] = bc(xDim, y)   (y in [0, yDim))
 *   ylo[x] = bc(x, -1), yhi[x] = bc(x, yDim)   (x in [0, xDim))
 * A null edge array means "the constant `value`".  Reference:
 * include/Hal/Pde/ADIequations.h:339-378, include/Hal/Pde/PDEequations.h:264-287. */
typedef struct halhip_bc
{
    int32_t kind;
    double value;
    const double *xlo;
    const double *xhi;
    const double *ylo;
    const double *yhi;
} halhip_bc;

typedef struct halhip_context halhip_context;
typedef struct halhip_diffusion2d halhip_diffusion2d;
typedef struct halhip_adi2d halhip_adi2d;
typedef struct halhip_agents2d halhip_agents2d;
typedef struct halhip_batch halhip_batch;
typedef struct halhip_adjoint2d halhip_adjoint2d;

/* ---- library / context -------------------------------------------------- */

/* Message of the last failed call on this thread ("" if none). */
const char *halhip_last_error(void);

/* Number of visible GPUs (0 without a GPU; never fails on a GPU-less host). */
int halhip_device_count(int32_t *count);

/* Replaces Hal::Gpu::ComputeContext (include/Hal/Gpu/ComputeContext.h:18-49):
 * binds a device.  HALHIP_ERR_NO_DEVICE when there is none, which callers
 * treat as "skip" like tests/GpuTestSupport.cpp:9-20. */
int halhip_context_create(int32_t device, halhip_context **out);
int halhip_context_destroy(halhip_context *ctx);

/* Waits for all work on the context's device (hipDeviceSynchronize). */
int halhip_context_synchronize(halhip_context *ctx);

/* Measurement aid: `repeats` launches of a plain copy kernel moving `bytes`
 * (read `bytes`, write `bytes`); msPerCopy gets the mean device time.  Used to
 * calibrate rocprofv3's FETCH_SIZE / WRITE_SIZE counters. */
int halhip_calibration_copy(halhip_context *ctx, int64_t bytes, int32_t repeats, float *msPerCopy);

/* Test hook: the next `count` device allocations made through ctx fail with
 * HALHIP_ERR_OUT_OF_MEMORY, so error paths can be exercised from the host. */
int halhip_context_fail_next_allocs(halhip_context *ctx, int32_t count);

/* ---- host-side seeded inputs (Hal::Rng, src/Rng.cpp:13-101) -------------- */

/* n draws of Rng(seed).NextDouble(). */
int halhip_rng_doubles(uint64_t seed, int64_t n, double *out);

/* The generator as an explicit state of four uint64 words, for checkpoints:
 * the words in order are the 32 bytes Hal::Rng::Save writes and
 * Hal::Rng::Load reads (src/Rng.cpp:122-130, raw native-endian
 * BinaryWriter::WriteArray), so a state saved by HAL resumes here and back.
 * seed_state is Hal::Rng(seed)'s state (src/Rng.cpp:37-45); next_doubles
 * draws n NextDouble values (src/Rng.cpp:56-69, 91-95) and advances state. */
int halhip_rng_seed_state(uint64_t seed, uint64_t *state);
int halhip_rng_next_doubles(uint64_t *state, int64_t n, double *out);

/* ---- explicit diffusion -------------------------------------------------- */

/* Replaces Hal::Gpu::DiffusionKernel2D / GpuDiffusionBackend2D
 * (include/Hal/Gpu/DiffusionKernel2D.h:16-20, src/Gpu/GpuDiffusionBackend2D.cpp:22-52).
 * A plan is sized at creation (like the reference's kernels) and owns its
 * device buffers and one HIP stream. */
int halhip_diffusion2d_create(halhip_context *ctx, int32_t xDim, int32_t yDim, halhip_diffusion2d **out);
int halhip_diffusion2d_destroy(halhip_diffusion2d *plan);

/* IDiffusionBackend2D::Diffuse (include/Hal/Pde/IDiffusionBackend2D.h:26-32):
 * deltas[i] += one explicit step's change of field[i]; n must equal
 * xDim*yDim of the plan.  bc may be null. */
int halhip_diffusion2d_diffuse_f64(halhip_diffusion2d *plan, const double *field, double *deltas, int64_t n,
                                   double diffRate, int32_t wrapX, int32_t wrapY, const halhip_bc *bc);

/* Per-cell rates (src/Pde/PDEequations.cpp:199-228): deltas += step with
 * x-face rates ratesX and y-face rates ratesY (each n values). */
int halhip_diffusion2d_diffuse_rates_f64(halhip_diffusion2d *plan, const double *field, double *deltas,
                                         const double *ratesX, const double *ratesY, int64_t n, int32_t wrapX,
                                         int32_t wrapY);

/* Resident state: FP32 field on the device. */
/* deltas += first-order upwind advection with the uniform velocity
 * (xVel, yVel): Hal::Internal::Advection2 (src/Pde/PDEequations.cpp:380-413;
 * boundary form include/Hal/Pde/PDEequations.h:413-447), as called by
 * PDEGrid2D::Advection (include/Hal/Grids/PDEGrid2D.h:349-375).  Off-grid
 * neighbours follow bc as in diffusion. */
int halhip_diffusion2d_advect_f64(halhip_diffusion2d *plan, const double *field, double *deltas, int64_t n,
                                  double xVel, double yVel, int32_t wrapX, int32_t wrapY, const halhip_bc *bc);

/* deltas += conservative upwind advection with per-cell velocities
 * (xVels, yVels: n values each): Hal::Internal::Advection2
 * (src/Pde/PDEequations.cpp:535-573; boundary form
 * include/Hal/Pde/PDEequations.h:721-783; PDEGrid2D::Advection with vectors,
 * include/Hal/Grids/PDEGrid2D.h:377-399).  Without bc, off-grid values and
 * velocities are 0 (walls); with a Dirichlet bc the field's off-grid values
 * are its edges and velBc (optional) gives the off-grid velocities: its x
 * edges the x velocity at x = -1 / xDim, its y edges the y velocity at
 * y = -1 / yDim. */
int halhip_diffusion2d_advect_cells_f64(halhip_diffusion2d *plan, const double *field, double *deltas,
                                        const double *xVels, const double *yVels, int64_t n, int32_t wrapX,
                                        int32_t wrapY, const halhip_bc *bc, const halhip_bc *velBc);

/* Uniform advection velocity included in every resident tick
 * (Diffusion + Advection + MulAll + Add + Update); 0, 0 (the default)
 * disables it.  Rejects |xVel| + |yVel| > 0.5 like
 * PDEGrid2D::CheckAdvectionStable (include/Hal/Grids/PDEGrid2D.h:444-451). */
int halhip_diffusion2d_set_advection(halhip_diffusion2d *plan, double xVel, double yVel);

/* Resident ticks of grids with yDim <= 1024 run several ticks per launch
 * (temporal blocking): each workgroup owns rowsPerGroup rows and advances
 * ticksPerLaunch ticks in LDS on rowsPerGroup + 2 ticksPerLaunch rows.
 * set_multi_shape pins (rowsPerGroup, ticksPerLaunch) for tests and sweeps
 * (0, 0: automatic); results do not depend on it.  get_multi_shape reports
 * the shape in use ((0, 0): one launch per tick). */
int halhip_diffusion2d_set_multi_shape(halhip_diffusion2d *plan, int32_t rowsPerGroup, int32_t ticksPerLaunch);
int halhip_diffusion2d_get_multi_shape(halhip_diffusion2d *plan, int32_t *rowsPerGroup, int32_t *ticksPerLaunch);

int halhip_diffusion2d_upload(halhip_diffusion2d *plan, const float *field, int64_t n);
int halhip_diffusion2d_download(halhip_diffusion2d *plan, float *field, int64_t n);
/* The resident field saved as / loaded from a HAL PDEGrid2D file
 * (Hal::SavePDEGrid2D / LoadPDEGrid2D, src/Grids/PDEGridSerialization.cpp:33-61:
 * int32 xDim, int32 yDim, bool wrapX, bool wrapY, int32 updateCt, FP64
 * field, FP64 deltas).  A resident plan holds committed state: saving writes
 * zero deltas, loading refuses a file with pending deltas or other
 * dimensions, and returns the file's wrap flags and update count. */
int halhip_diffusion2d_save_grid(halhip_diffusion2d *plan, const char *path, int32_t wrapX, int32_t wrapY,
                              int32_t updateCt);
int halhip_diffusion2d_load_grid(halhip_diffusion2d *plan, const char *path, int32_t *wrapX, int32_t *wrapY,
                              int32_t *updateCt);

/* Point sources added every tick (PDEGrid2D::Add, include/Hal/Grids/PDEGrid2D.h:110-114).
 * Indices are flat cell indices; a repeated index adds twice. */
int halhip_diffusion2d_set_sources(halhip_diffusion2d *plan, int32_t count, const int64_t *index,
                                   const float *value);

/* `steps` model ticks on the resident field, each
 *   u += diffRate*laplacian(u) - decay*u + sources
 * (Diffusion + MulAll(-decay) + Add + Update, PDEGrid2D.h:103-218;
 * DiffusionKernel2D::RunSteps, src/Gpu/DiffusionKernel2D.cpp:244-304). */
int halhip_diffusion2d_run_steps(halhip_diffusion2d *plan, int32_t steps, double diffRate, double decay,
                                 int32_t wrapX, int32_t wrapY, const halhip_bc *bc);

/* Same ticks enqueued on the plan's stream; returns at once.  Pair with
 * halhip_diffusion2d_sync. elapsedMs (may be null) receives the device time of
 * the enqueued work measured with HIP events once synced. */
int halhip_diffusion2d_run_steps_async(halhip_diffusion2d *plan, int32_t steps, double diffRate, double decay,
                                       int32_t wrapX, int32_t wrapY, const halhip_bc *bc);
int halhip_diffusion2d_sync(halhip_diffusion2d *plan, float *elapsedMs);
/* Builds every graph a later run_steps with the same arguments would replay
 * (for either ping-pong buffer) and runs no tick; resident state is
 * unchanged.  No counterpart in the reference (DiffusionKernel2D::RunSteps
 * records its command buffer per call, src/Gpu/DiffusionKernel2D.cpp:244-304). */
int halhip_diffusion2d_prepare_steps(halhip_diffusion2d *plan, int32_t steps, double diffRate, double decay,
                                     int32_t wrapX, int32_t wrapY, const halhip_bc *bc);
/* ---- PDEGrid2D deltas on the resident field ----
 * Hal::PDEGrid2D's delta calls (include/Hal/Grids/PDEGrid2D.h:89-245,
 * 289-399) applied to the plan's resident (committed) field u: each call
 * changes the device deltas d, and delta_update commits u += d, d = 0
 * (PDEGrid2D::Update, :178-186).  d starts at zero and upload clears it.
 * FP32 on the device.  Point ops run in list order:
 *   HALHIP_DELTA_SET    d[i] = v - u[i]        (Set)
 *   HALHIP_DELTA_ADD    d[i] += v              (Add)
 *   HALHIP_DELTA_MUL    d[i] += u[i] * v       (Mul)
 *   HALHIP_DELTA_SCALE  d[i] += u[i] * (v - 1) (Scale)
 * delta_all applies one of them to every cell (SetAll, AddAll, MulAll, a
 * field-wide Scale) or HALHIP_DELTA_SET_NON_NEGATIVE (SetNonNegative:
 * d[i] = -u[i] where u[i] + d[i] < 0); delta_set_values is SetAll(pVals).
 * The stencil calls add the interface deltas of the committed field:
 * Diffusion (with bc), Diffusion(ratesX, ratesY), Advection (uniform, with
 * bc), Advection(xVels, yVels[, bc, velBc]) and, through an ADI plan of the
 * same dimensions (whose resident field it replaces), DiffusionADI.
 * delta_reduce: MaxDelta, MaxDeltaScaled(offset), GetMax, GetMin and the
 * field's sum (GetAvg * n). */
#define HALHIP_DELTA_SET 0
#define HALHIP_DELTA_ADD 1
#define HALHIP_DELTA_MUL 2
#define HALHIP_DELTA_SCALE 3
#define HALHIP_DELTA_SET_NON_NEGATIVE 4
#define HALHIP_REDUCE_MAX_DELTA 0
#define HALHIP_REDUCE_MAX_DELTA_SCALED 1
#define HALHIP_REDUCE_FIELD_MAX 2
#define HALHIP_REDUCE_FIELD_MIN 3
#define HALHIP_REDUCE_FIELD_SUM 4
int halhip_diffusion2d_delta_point_ops(halhip_diffusion2d *plan, int32_t count, const int32_t *ops,
                                       const int64_t *index, const double *values);
int halhip_diffusion2d_delta_all(halhip_diffusion2d *plan, int32_t op, double value);
int halhip_diffusion2d_delta_set_values(halhip_diffusion2d *plan, const double *values, int64_t n);
int halhip_diffusion2d_delta_diffuse(halhip_diffusion2d *plan, double diffRate, int32_t wrapX, int32_t wrapY,
                                     const halhip_bc *bc);
int halhip_diffusion2d_delta_diffuse_rates(halhip_diffusion2d *plan, const double *ratesX, const double *ratesY,
                                           int64_t n, int32_t wrapX, int32_t wrapY);
int halhip_diffusion2d_delta_advect(halhip_diffusion2d *plan, double xVel, double yVel, int32_t wrapX,
                                    int32_t wrapY, const halhip_bc *bc);
int halhip_diffusion2d_delta_advect_cells(halhip_diffusion2d *plan, const double *xVels, const double *yVels,
                                          int64_t n, int32_t wrapX, int32_t wrapY, const halhip_bc *bc,
                                          const halhip_bc *velBc);
int halhip_diffusion2d_delta_adi(halhip_diffusion2d *plan, halhip_adi2d *adi, double diffRate, int32_t wrapX,
                                 int32_t wrapY, const halhip_bc *bc);
int halhip_diffusion2d_delta_reduce(halhip_diffusion2d *plan, int32_t what, double offset, double *out);
int halhip_diffusion2d_delta_update(halhip_diffusion2d *plan);
int halhip_diffusion2d_download_deltas(halhip_diffusion2d *plan, float *deltas, int64_t n);

int halhip_diffusion2d_set_profiling(halhip_diffusion2d *plan, int32_t on);
int halhip_diffusion2d_profile(halhip_diffusion2d *plan, int32_t maxKernels, int32_t *nKernels, char *names,
                               int32_t namesLen, double *totalMs, int64_t *launches);

/* ---- ADI diffusion ------------------------------------------------------- */

/* Replaces Hal::Gpu::AdiKernel2D / GpuAdiDiffusionBackend2D
 * (include/Hal/Gpu/AdiKernel2D.h:33-41, src/Gpu/GpuAdiDiffusionBackend2D.cpp:24-50).
 * Limits: xDim * yDim <= 2^31; no cap on either axis (the reference's GPU
 * path caps both at 1024, its CPU path has none,
 * include/Hal/Pde/TdmaSolver.h:123-130).  Rows of up to HALHIP_ADI_MAX_Y
 * columns are solved inside one workgroup; longer rows are cut into
 * 4096-column segments across workgroups and joined exactly
 * (HALHIP_ADI_ROWS_SEGMENTED), at about twice the traffic per tick. */
#define HALHIP_ADI_MAX_Y 8192
int halhip_adi2d_create(halhip_context *ctx, int32_t xDim, int32_t yDim, halhip_adi2d **out);
int halhip_adi2d_destroy(halhip_adi2d *plan);

/* IAdiDiffusionBackend2D::DiffuseADI (include/Hal/Pde/IAdiDiffusionBackend2D.h:25-31):
 * deltas[i] += ADI(field)[i] - field[i].  bc may be null (zero flux /
 * periodic); a Dirichlet bc gives PDEGrid2D::DiffusionADI(rate, bc)
 * (include/Hal/Grids/PDEGrid2D.h:328-347).  The call works on a field
 * buffer of its own (allocated on the first call): the plan's resident field
 * is left as it was. */
int halhip_adi2d_diffuse_f64(halhip_adi2d *plan, const double *field, double *deltas, int64_t n, double diffRate,
                             int32_t wrapX, int32_t wrapY, const halhip_bc *bc);

int halhip_adi2d_upload(halhip_adi2d *plan, const float *field, int64_t n);
int halhip_adi2d_download(halhip_adi2d *plan, float *field, int64_t n);
/* The resident field saved as / loaded from a HAL PDEGrid2D file
 * (Hal::SavePDEGrid2D / LoadPDEGrid2D, src/Grids/PDEGridSerialization.cpp:33-61:
 * int32 xDim, int32 yDim, bool wrapX, bool wrapY, int32 updateCt, FP64
 * field, FP64 deltas).  A resident plan holds committed state: saving writes
 * zero deltas, loading refuses a file with pending deltas or other
 * dimensions, and returns the file's wrap flags and update count. */
int halhip_adi2d_save_grid(halhip_adi2d *plan, const char *path, int32_t wrapX, int32_t wrapY,
                              int32_t updateCt);
int halhip_adi2d_load_grid(halhip_adi2d *plan, const char *path, int32_t *wrapX, int32_t *wrapY,
                              int32_t *updateCt);
int halhip_adi2d_set_sources(halhip_adi2d *plan, int32_t count, const int64_t *index, const float *value);

/* `steps` model ticks u <- ADI(u) - decay*u + sources on the resident field
 * (DiffusionADI + MulAll(-decay) + Add + Update; AdiKernel2D::RunSteps). */
int halhip_adi2d_run_steps(halhip_adi2d *plan, int32_t steps, double diffRate, double decay, int32_t wrapX,
                           int32_t wrapY, const halhip_bc *bc);
int halhip_adi2d_run_steps_async(halhip_adi2d *plan, int32_t steps, double diffRate, double decay, int32_t wrapX,
                                 int32_t wrapY, const halhip_bc *bc);
int halhip_adi2d_sync(halhip_adi2d *plan, float *elapsedMs);
/* Builds (captures and instantiates) every graph a later run_steps with the
 * same arguments would replay, from either reduction half, and runs no tick:
 * a serving loop or a timed run then pays no capture cost.  Resident state is
 * unchanged.  Runs of fewer than 20 steps replay no graph.  No counterpart in
 * the reference (its Vulkan RunSteps records its command buffer per call,
 * src/Gpu/AdiKernel2D.cpp:296-368). */
int halhip_adi2d_prepare_steps(halhip_adi2d *plan, int32_t steps, double diffRate, double decay, int32_t wrapX,
                               int32_t wrapY, const halhip_bc *bc);

/* Strip decomposition of the plan: each workgroup holds `stripRows` full rows
 * with `columnsPerThread` adjacent columns per thread (1, 2 or 4 with 4, 8 or
 * 16 rows; 8 with 4 or 8 rows); `vector` moves rows as 16-byte vectors and
 * needs yDim % columnsPerThread == 0.  Segmented rows take 4 columns per
 * thread with 4 or 8 rows.  Results do not depend on the shape beyond FP32
 * rounding.  The default is chosen from the grid size. */
int halhip_adi2d_set_strip_shape(halhip_adi2d *plan, int32_t columnsPerThread, int32_t stripRows, int32_t vector);
int halhip_adi2d_get_strip_shape(halhip_adi2d *plan, int32_t *columnsPerThread, int32_t *stripRows, int32_t *vector);

/* Methods of the resident ADI tick, for testing and benchmarking; both are
 * HALHIP_ADI_AUTO by default.
 *   rowSolver: HALHIP_ADI_ROWS_WORKGROUP_SCAN solves each contiguous line as a
 *     scan across the workgroup's registers; HALHIP_ADI_ROWS_LDS stages the
 *     strip in LDS and solves each line with constant Thomas factors plus an
 *     exact rank-2 correction for its ends (or, on a periodic y axis, its
 *     corners); supported shapes only.  HALHIP_ADI_ROWS_SEGMENTED cuts the
 *     rows into segments across workgroups, each solved with zero carries,
 *     then joins them with the segments' exact responses to their carries
 *     (4096-column segments; forced on rows of up to HALHIP_ADI_MAX_Y it
 *     uses 256-column segments, for testing).  Rows longer than
 *     HALHIP_ADI_MAX_Y are always segmented, with scan carries; asking them
 *     for another row solver or windowed carries is HALHIP_ERR_UNSUPPORTED.
 *   carries: HALHIP_ADI_CARRIES_SCAN runs a cross-strip scan kernel per tick;
 *     HALHIP_ADI_CARRIES_WINDOW composes each strip's carries from the K
 *     strips on either side when the line factors decay fast enough that the
 *     dropped terms are below 2^-28 (LDS rows only; a periodic xDim also
 *     forms its column corrections from the first and last strips).
 * A method that does not apply falls back to the other one;
 * halhip_adi2d_get_methods reports what the last plain run used and the
 * window K (0 without windowed carries).  No counterpart in the reference. */
#define HALHIP_ADI_AUTO 0
#define HALHIP_ADI_ROWS_WORKGROUP_SCAN 1
#define HALHIP_ADI_ROWS_LDS 2
#define HALHIP_ADI_ROWS_SEGMENTED 3
#define HALHIP_ADI_CARRIES_SCAN 1
#define HALHIP_ADI_CARRIES_WINDOW 2
int halhip_adi2d_set_methods(halhip_adi2d *plan, int32_t rowSolver, int32_t carries);
int halhip_adi2d_get_methods(halhip_adi2d *plan, int32_t *rowSolver, int32_t *carries, int32_t *window);

/* Per-kernel device time of the last run (HIP events around each launch when
 * profiling is on; profiling is off by default and does not change results).
 * names: comma-separated kernel names; ms/count arrays of length *nKernels. */
int halhip_adi2d_set_profiling(halhip_adi2d *plan, int32_t on);
int halhip_adi2d_profile(halhip_adi2d *plan, int32_t maxKernels, int32_t *nKernels, char *names, int32_t namesLen,
                         double *totalMs, int64_t *launches);

/* ---- agent forces -------------------------------------------------------- */

/* Replaces Hal::Gpu::AgentForceKernel2D / GpuAgentForceBackend2D
 * (include/Hal/Gpu/AgentForceKernel2D.h, src/Gpu/GpuAgentForceBackend2D.cpp:24-68).
 * Sized at creation: grid xDim*yDim bins of size 1 and maxAgents capacity. */
int halhip_agents2d_create(halhip_context *ctx, int32_t xDim, int32_t yDim, int32_t maxAgents,
                           halhip_agents2d **out);
int halhip_agents2d_destroy(halhip_agents2d *plan);

/* IAgentForceBackend2D::AccumulateForces (include/Hal/Agents/IAgentForceBackend2D.h:28-36):
 * positions (2*nAgents, x0,y0,x1,...), radii (nAgents); velocityDeltas
 * (2*nAgents) is overwritten. */
int halhip_agents2d_forces_f64(halhip_agents2d *plan, const double *positions, const double *radii,
                               int32_t nAgents, double *velocityDeltas, double interactionRad,
                               double springConstant, int32_t wrapX, int32_t wrapY);

/* Resident agents: FP32 positions and radii on the device. */
int halhip_agents2d_upload(halhip_agents2d *plan, const float *positions, const float *radii, int32_t nAgents);
int halhip_agents2d_download_positions(halhip_agents2d *plan, float *positions, int32_t nAgents);
/* Velocity deltas (2*nAgents) of the last tick of the last run (or of the
 * last forces-only pass). */
int halhip_agents2d_download_velocities(halhip_agents2d *plan, float *velocityDeltas, int32_t nAgents);
/* Bins of the last tick: offsets (xDim*yDim+1 int32, exclusive prefix of
 * per-bin counts) and bucket (nAgents int32, agent indices in bin order,
 * ascending index within a bin). */
int halhip_agents2d_download_bins(halhip_agents2d *plan, int32_t *offsets, int32_t *bucket);
/* Bound on the polls the one-pass bin scan makes while it waits for an
 * earlier tile's published total (default 2^20).  A scan that reaches it
 * stops waiting and finishes; the next download or per-call forces call then
 * fails with HALHIP_ERR_HIP and the bins must be recomputed (upload again).
 * A limit of 0 lets tests force that path.  No counterpart in the
 * reference. */
int halhip_agents2d_set_scan_spin_limit(halhip_agents2d *plan, int32_t limit);

/* Velocity carried across ticks, as SphericalAgent2D keeps xVel/yVel
 * (include/Hal/Agents/SphericalAgent2D.h:100-156): each tick the velocity
 * is carry + the tick's force delta, its norm is capped at maxVel
 * (CapVelocity, :139-150; maxVel == 0 means no cap), the move uses it
 * (ForceMove), and friction * velocity is carried (ApplyFriction, :133-137).
 * friction in [0, 1].  friction == 0 with no cap (the default) carries
 * nothing: the move uses the force delta alone.  Enabling starts every
 * agent at rest, and so does an upload.  upload/download_motion move the
 * carried velocities (2*nAgents, agent-index order) for checkpoints. */
int halhip_agents2d_set_motion(halhip_agents2d *plan, double friction, double maxVel);
int halhip_agents2d_upload_motion(halhip_agents2d *plan, const float *velocities, int32_t nAgents);
int halhip_agents2d_download_motion(halhip_agents2d *plan, float *velocities, int32_t nAgents);

/* `ticks` ticks of bin -> scan -> stable scatter -> force -> move.  The move
 * is p += dt*v with AgentPT2D::MoveSafePT semantics
 * (include/Hal/Agents/AgentPT2D.h:157-186): an axis whose target leaves a
 * non-wrapped grid keeps its old coordinate; a wrapped axis wraps.  ticks == 0
 * with dt == 0 computes bins and velocities only. */
int halhip_agents2d_run_ticks(halhip_agents2d *plan, int32_t ticks, double interactionRad, double springConstant,
                              double dt, int32_t wrapX, int32_t wrapY);
int halhip_agents2d_run_ticks_async(halhip_agents2d *plan, int32_t ticks, double interactionRad,
                                    double springConstant, double dt, int32_t wrapX, int32_t wrapY);
int halhip_agents2d_sync(halhip_agents2d *plan, float *elapsedMs);
/* Builds every graph a later run_ticks with the same arguments and agent
 * count would replay (from either buffer set) and runs no tick; resident
 * state is unchanged.  Runs of fewer than 10 ticks replay no graph.  No
 * counterpart in the reference (AgentForceKernel2D::Dispatch records its
 * command buffer per call, src/Gpu/AgentForceKernel2D.cpp:189-290). */
int halhip_agents2d_prepare_ticks(halhip_agents2d *plan, int32_t ticks, double interactionRad,
                                  double springConstant, double dt, int32_t wrapX, int32_t wrapY);
int halhip_agents2d_set_profiling(halhip_agents2d *plan, int32_t on);
int halhip_agents2d_profile(halhip_agents2d *plan, int32_t maxKernels, int32_t *nKernels, char *names,
                            int32_t namesLen, double *totalMs, int64_t *launches);

/* Bins without allocation.  The same bins as download_bins after a binning
 * of these positions (offsets = exclusive prefix of per-bin counts, xDim*yDim
 * + 1 values; bucket = agent indices in bin order, ascending within a bin),
 * for callers that cannot allocate device memory: code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc.  No counterpart in
 * the reference.  positions (2*nAgents floats, x then y per agent), offsets
 * and bucket are DEVICE pointers owned by the caller.  The work is enqueued
 * on `stream` (a hipStream_t; null is the null stream) and the call returns
 * without waiting; the results are ready when the stream reaches that point.
 * Nothing is allocated or freed: the working memory is one fixed array per
 * GPU (12 MiB, held until the process ends), so nAgents is at most
 * HALHIP_AGENTS2D_NOALLOC_MAX_AGENTS and xDim*yDim at most
 * HALHIP_AGENTS2D_NOALLOC_MAX_BINS (larger inputs are rejected with
 * HALHIP_ERR_INVALID_ARGUMENT).  Every call on one GPU shares that array:
 * calls that overlap in time on one GPU (other streams, other threads) race
 * and must be serialised by the caller. */
#define HALHIP_AGENTS2D_NOALLOC_MAX_AGENTS 1048576
#define HALHIP_AGENTS2D_NOALLOC_MAX_BINS 1048576
int halhip_agents2d_bins_noalloc(halhip_context *ctx, const float *positions, int32_t nAgents, int32_t xDim,
                                 int32_t yDim, int32_t *offsets, int32_t *bucket, void *stream);

/* ---- batched wells --------------------------------------------------------- */

/* Many independent ADI grids ("wells") of uneven size ticked together, each
 * tick u_w <- ADI(u_w; rate_w) - decay_w*u_w + sources_w, i.e. per well
 * PDEGrid2D::DiffusionADI + MulAll(-decay) + Add + Update
 * (include/Hal/Grids/PDEGrid2D.h:103-218, 328-347).  Replaces the host-side
 * MultiWellExperiment::Step / StepMultiThread loop
 * (include/Hal/Tools/MultiWellExperiment.h:82-103).  Well sides <= 1024.
 * Fields are uploaded and downloaded concatenated in well order, each well in
 * HAL's x*yDim+y layout. */
int halhip_batch_create(halhip_context *ctx, int32_t nWells, const int32_t *xDims, const int32_t *yDims,
                        halhip_batch **out);
int halhip_batch_destroy(halhip_batch *batch);
/* Per-well diffusion rate and decay (decays may be null: 0). */
int halhip_batch_set_wells(halhip_batch *batch, const double *rates, const double *decays);
/* Point sources added every tick: well index, flat cell index in that well,
 * value.  Repeated cells add in list order. */
int halhip_batch_set_sources(halhip_batch *batch, int32_t count, const int32_t *well, const int64_t *cell,
                             const float *value);
int halhip_batch_upload(halhip_batch *batch, const float *fields, int64_t n);
int halhip_batch_download(halhip_batch *batch, float *fields, int64_t n);
/* `ticks` ticks of every well: two batched launches per tick, replayed from a
 * hipGraph.  bc: null, or a constant Dirichlet value for every well. */
int halhip_batch_run_ticks(halhip_batch *batch, int32_t ticks, int32_t wrapX, int32_t wrapY, const halhip_bc *bc);
int halhip_batch_run_ticks_async(halhip_batch *batch, int32_t ticks, int32_t wrapX, int32_t wrapY,
                                 const halhip_bc *bc);
/* The same ticks as a per-instance launch loop (the same kernels, one well at
 * a time): the baseline the batched launch is measured against. */
int halhip_batch_run_ticks_per_instance(halhip_batch *batch, int32_t ticks, int32_t wrapX, int32_t wrapY,
                                        const halhip_bc *bc);
int halhip_batch_run_ticks_per_instance_async(halhip_batch *batch, int32_t ticks, int32_t wrapX, int32_t wrapY,
                                              const halhip_bc *bc);
int halhip_batch_sync(halhip_batch *batch, float *elapsedMs);
/* Builds every graph a later run_ticks with the same arguments would replay
 * (from either reduction half) and runs no tick; the wells are unchanged.
 * Runs of fewer than 10 ticks replay no graph.  No counterpart in the
 * reference (MultiWellExperiment steps wells on the host). */
int halhip_batch_prepare_ticks(halhip_batch *batch, int32_t ticks, int32_t wrapX, int32_t wrapY,
                               const halhip_bc *bc);
/* Strip height of the batched kernels: 4, 8 or 16 rows (16 only with a
 * workgroup shape of at most 2 columns per thread).  set_strip_rows sets
 * every width class (see halhip_batch_set_shapes); set_class_rows sets one
 * height per class; get_strip_rows reports the widest class's. */
int halhip_batch_set_strip_rows(halhip_batch *batch, int32_t rows);
int halhip_batch_get_strip_rows(halhip_batch *batch, int32_t *rows);
int halhip_batch_set_class_rows(halhip_batch *batch, const int32_t *rows);
int halhip_batch_get_class_rows(halhip_batch *batch, int32_t *rows);
/* Wells are grouped by width (<= 256, <= 512, <= 1024 columns) and each
 * width class runs as a launch of its own, with a workgroup shape picked from
 * three options (threads x columns per thread):
 *   <= 256:  0: 256 x 1   1: 128 x 2   2:   64 x 4
 *   <= 512:  0: 256 x 2   1: 512 x 1   2:  128 x 4
 *   <= 1024: 0: 256 x 4   1: 512 x 2   2: 1024 x 1
 * set_shapes takes one option index per class (0 is the default);
 * get_shapes reports them with their threads and columns per thread (any
 * output may be null).  No counterpart in the reference. */
int halhip_batch_set_shapes(halhip_batch *batch, const int32_t *shapes);
int halhip_batch_get_shapes(halhip_batch *batch, int32_t *shapes, int32_t *threads, int32_t *columnsPerThread);
/* Times carries {scan, windowed} x, per width class, every workgroup shape
 * with every strip height its registers allow, on `ticks` ticks (state is
 * restored afterwards; each class's option is timed on its own launches and
 * chosen independently), and keeps the fastest.  With a cache file, results
 * are stored per batch signature (well shapes, gfx target) as
 * "windowed:shape.rows,shape.rows,shape.rows" and reused.  chosenRows gets
 * the widest class's strip height. */
int halhip_batch_autotune(halhip_batch *batch, int32_t ticks, int32_t wrapX, int32_t wrapY, const halhip_bc *bc,
                          const char *cachePath, int32_t *chosenRows);
/* Carry method of the batch's sweep A across strips: HALHIP_ADI_CARRIES_SCAN
 * (the default: a scan launch per tick) or HALHIP_ADI_CARRIES_WINDOW (each
 * strip composes its carries from its neighbours' reductions, as in
 * halhip_adi2d_set_methods; used only when every well allows it).
 * halhip_batch_get_carries reports the method the last build used;
 * halhip_batch_autotune times both.  No counterpart in the reference. */
int halhip_batch_set_carries(halhip_batch *batch, int32_t method);
int halhip_batch_get_carries(halhip_batch *batch, int32_t *method);

int halhip_batch_set_profiling(halhip_batch *batch, int32_t on);
int halhip_batch_profile(halhip_batch *batch, int32_t maxKernels, int32_t *nKernels, char *names, int32_t namesLen,
                         double *totalMs, int64_t *launches);

/* ---- adjoint ADI --------------------------------------------------------------- */

/* Per-cell-diffusivity ADI on an xDim x yDim grid, its discrete adjoint and
 * gradients.  No reference counterpart: the forward tick generalises
 * PDEGrid2D::DiffusionADI (include/Hal/Pde/ADIequations.h:306-378) to a
 * per-cell rate D with the face convention of the per-cell explicit stencil
 * (src/Pde/PDEequations.cpp:199-228: the face between cells i and i+1 carries
 * D_i, the wrap face D_{n-1}) and appends MulAll(-decay) + Update.  Sides
 * <= 4096. */
int halhip_adjoint2d_create(halhip_context *ctx, int32_t xDim, int32_t yDim, halhip_adjoint2d **out);
int halhip_adjoint2d_destroy(halhip_adjoint2d *plan);
/* Diffusivity field (xDim*yDim, HAL layout), decay rate, number of ticks
 * and checkpoint interval (the backward pass recomputes each interval from
 * its checkpoint). */
int halhip_adjoint2d_set_problem(halhip_adjoint2d *plan, const double *diffusivity, double decay, int32_t ticks,
                                 int32_t checkpointEvery);
/* The same with boundaries, as the reference's ADI takes them
 * (include/Hal/Pde/ADIequations.h:306-378): a wrapped axis is periodic; a
 * Dirichlet bc (constant or edge arrays) applies on the non-wrapped axes,
 * its ghost face carrying the boundary cell's own D; otherwise zero flux.
 * set_problem is this call with no wrap and a null bc. */
int halhip_adjoint2d_set_problem_bc(halhip_adjoint2d *plan, const double *diffusivity, double decay, int32_t ticks,
                                    int32_t checkpointEvery, int32_t wrapX, int32_t wrapY, const halhip_bc *bc);
/* u_T from u0 (host FP32 arrays of xDim*yDim). */
int halhip_adjoint2d_forward(halhip_adjoint2d *plan, const float *u0, float *uT);
/* J = 1/2 |u_T - target|^2 and its gradients with respect to every cell's
 * diffusivity (gradDiffusivity, xDim*yDim) and to the decay rate. */
int halhip_adjoint2d_gradient(halhip_adjoint2d *plan, const float *u0, const float *target, double *loss,
                              float *gradDiffusivity, double *gradDecay);
int halhip_adjoint2d_set_profiling(halhip_adjoint2d *plan, int32_t on);
int halhip_adjoint2d_profile(halhip_adjoint2d *plan, int32_t maxKernels, int32_t *nKernels, char *names,
                             int32_t namesLen, double *totalMs, int64_t *launches);

#ifdef __cplusplus
}
#endif

#endif /* HALHIP_H */
