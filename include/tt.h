/*
 * tt.h -- C ABI of libtt.so, the MI355X (gfx950) kernels behind the
 * turbulence_tracing drop-in.
 *
 * The reference (jdhare/turbulence_tracing) is pure Python and has no FFI;
 * each entry point below replaces the numpy/scipy work inside one reference
 * function, cited as file:line relative to the reference root.  The Python
 * modules of turbulence_tracing_amd bind these with ctypes.
 *
 * Conventions
 *  - Every pointer argument named as device memory must be device memory
 *    (hipMalloc / torch CUDA tensor); structs are host memory.
 *  - Every call is asynchronous on `stream` (a hipStream_t, NULL = the
 *    default stream) and returns TT_OK or an error code; tt_last_error()
 *    returns a thread-local message for the last failure.
 *  - Ray arrays are SoA rows of length ld (>= n): s0/sf are 6 x ld float64
 *    [x, y, z, vx, vy, vz] in m and m/s; rf is 4 x ld float64
 *    [x, theta, y, phi] in m and rad (mm after the optics, like the
 *    reference's m_to_mm).
 *  - Grid fields are stored x-fastest, z-slowest: index (iz*ny + iy)*nx + ix.
 *  - A ray-parallel call with n = 0 (an empty ray set) checks its sizes and
 *    host structs only, touches no device memory and returns TT_OK, so its
 *    device pointers may be NULL (an empty torch tensor has none).
 *  - The caller owns every device buffer.  The library allocates only inside
 *    an FFT plan (its twiddle table), released by tt_fft_plan_destroy.
 */
#ifndef TT_H
#define TT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TT_OK 0
#define TT_ERR_ARG 1
#define TT_ERR_HIP 2
#define TT_ERR_UNSUPPORTED 3

int tt_version(void);
const char *tt_last_error(void);

/* Uniform grid: node (ix, iy, iz) sits at (x0 + ix dx, y0 + iy dy, z0 + iz dz).
 * Mirrors the x, y, z coordinate arrays of ElectronCube
 * (particle_tracking/particle_tracker.py:125-145). */
typedef struct {
    int32_t nx, ny, nz;
    double x0, y0, z0;
    double dx, dy, dz;
} tt_grid;

/* ------------------------------------------------------------------ field
 * Replaces gaussian3D_FFT (gaussian_fields/turboGen.py:488-538): centred-grid
 * Gaussian noise, Hermitian symmetrisation, sqrt(P(k)) shaping, zero mode,
 * inverse 3-D FFT, real part.  Output `out` is float32 [nz][ny][nx] and holds
 * the reference's array f[ix][iy][iz] (axis 0 = x).  When `sumsq` (device,
 * one double) is non-NULL the sum of squares of the output is ADDED to it.
 * Any nx, ny, nz >= 2 and <= 4096 is accepted (mixed radix 2/3/4/5/8 plus
 * direct odd-prime passes).
 */
typedef struct tt_fft_plan_s *tt_fft_plan;
int tt_fft_plan_create(int32_t nx, int32_t ny, int32_t nz, tt_fft_plan *plan);
int tt_fft_plan_destroy(tt_fft_plan plan);
/* bytes of device workspace the generate calls need (the half spectrum) */
int tt_fft_workspace_bytes(tt_fft_plan plan, size_t *bytes);

/* Power-law spectrum P = amp * Ka^beta on kmin <= Ka <= kmax, zero elsewhere,
 * Ka = sqrt(kx^2 + ky^2 + (kz/aniso)^2) in cycles/sample (the units of
 * np.fft.fftfreq used by turboGen.py:516). */
typedef struct {
    double beta, aniso, kmin, kmax, amp;
    uint64_t seed;
} tt_spectrum;

/* Native generator: Philox4x32-10 noise drawn in-kernel. */
int tt_field_generate(tt_fft_plan plan, const tt_spectrum *spec, void *workspace, float *out,
                      double *sumsq, void *stream);
/* Drop-in generator: the caller's centred-grid noise wr, wi and sqrt(P)
 * (float32 device arrays, C order [cx][cy][cz], as turboGen.py:522-530 builds
 * them on the host). */
int tt_field_from_noise(tt_fft_plan plan, const float *wr, const float *wi, const float *sqrt_psd,
                        void *workspace, float *out, double *sumsq, void *stream);

/* Radial power spectrum of a real field, replacing spectrum_3D_scalar
 * (gaussian_fields/calculate_spectrum_3d.py:3-59) and compute3Dspectrum
 * (gaussian_fields/cmpspec.py:147-194).  field: float32 C-order array of
 * the reference's shape (M0, M1, M2), read with the plan created as
 * tt_fft_plan_create(M2, M1, M0); workspace as for tt_field_generate.  The
 * bins are ADDED to sums / counts (float64, nbins each, caller-zeroed).
 * mode 0 (spectrum_3D_scalar): |fftn(f)|^2 summed, and points counted, in
 *   bins b <= nbins - 2 of K = sqrt((k0^2 + k1^2) + k2^2), k_i =
 *   fftfreq(M_i, dx): bin_w b <= K < bin_w (b + 1), bin_w = K.max() / nbins
 *   computed by the caller as the reference does (C3:45);
 * mode 1 (compute3Dspectrum): |fftn(f) / nt|^2 summed into shells
 *   round(sqrt(v0^2 + v1^2 + v2^2)) < nbins over the reference's loop
 *   values v_i in range(-M_i // 2, M_i // 2 - 1) (CS:175-180); bin_w unused. */
int tt_power_spectrum(tt_fft_plan plan, const float *field, int32_t mode, int32_t nbins, double dx, double bin_w,
                      void *workspace, double *sums, double *counts, void *stream);

/* ---------------------------------------------------------------- gradient
 * Replaces calc_dndr (particle_tracking/particle_tracker.py:220-241):
 * ne/nc = alpha + beta * f, clipped at ne_max, then np.gradient semantics
 * (2nd-order central interior, 1st-order one-sided edges) along each axis.
 * `field` is float32 (dtype 0) or float64 (dtype 1) with element strides
 * (sx, sy, sz).  Writes float4 out4[iz][iy][ix] =
 *   { -0.5 dz d(ne/nc)/dx, -0.5 dz d(ne/nc)/dy, -0.5 dz d(ne/nc)/dz,
 *     k0 dz (sqrt(1 - ne/nc) - 1) }
 * i.e. the reference's dndx/dndy/dndz divided by c^2 and multiplied by dz,
 * plus the phase rate per z-cell.  When sumsq (device, one double, e.g. from
 * tt_field_generate) is non-NULL, beta is divided by sqrt(*sumsq / count):
 * the caller-side unit-RMS normalisation of the field, applied on the device. */
int tt_grad_stencil(const void *field, int32_t dtype, int64_t sx, int64_t sy, int64_t sz,
                    const tt_grid *grid, double alpha, double beta, double ne_max, double k0,
                    const double *sumsq, int64_t count, float *out4, void *stream);

/* Plane forms of a node field for the tracer's cell cache (the trilinear
 * interpolation of calc_dndr's gradients, PT:239-241, reorganised per z-plane
 * cell): forms[iz][iy][ix] = 4 float4 {A, B, C, E} with A = f(ix, iy),
 * B = f(ix+1, iy) - A, C = f(ix, iy+1) - A, E = (f(ix+1, iy+1) - f(ix, iy+1))
 * - B on plane iz, so the field on that plane cell is A + B tx + C ty +
 * E tx ty (B = C = E = 0 on the last column and row).  64 bytes per node;
 * pass the result as tt_trace_opts.forms. */
int tt_plane_forms(const float *field4, const tt_grid *grid, float *forms, void *stream);

/* ------------------------------------------------------------------ tracer
 * Replaces solve + ray_at_exit + the dndr/dsdt calls inside solve_ivp
 * (particle_tracking/particle_tracker.py:243-256, 312-380, 398-419).
 * Fixed-step RK4 with z (in cells) as the independent variable, `h` = 1/m
 * cells per step, from each ray's start plane to the last grid plane; the
 * interferometric phase k0 int (n-1) ds is accumulated in float64.
 * order (nullable, int32 device) lists the rays to process in that order
 * (e.g. from tt_ray_sort); outputs are written at each ray's own index.
 * sf (nullable) receives the state at c t = cT, continued straight past the
 * exit plane (the field is zero outside the grid, PT:239-241).
 * status (nullable, int32): 0 ok; 1 handed over by the z-march: W = vz / c
 * <= 0.97 at launch or at the end of a step (a slope above ~0.25, or a ray
 * slowing towards a turning point, where a fixed z-step is not accurate to
 * 1e-3 cells), or a side face (x or y) of the grid crossed (the field jumps to
 * zero there); its outputs are NaN and tt_trace_time traces it in time;
 * 2 NaN input (or a ray tt_trace_time gave up on); 3 traced in time by
 * tt_trace_time.  Without status, handed-over rays just get NaN outputs.
 * variant 0: each lane caches its current cell as a trilinear polynomial in
 * registers (the fast path); variant 1: eight float4 gathers per RK4 stage.
 * With h = 1/2 and variant 0 the phase is summed by the composite Simpson
 * rule over the lattice points (the ray's states on the z-planes and
 * half-planes), elsewhere by the RK4 weights.  forms (nullable, device,
 * 16-byte aligned): the plane forms tt_plane_forms writes for field4; with
 * h = 1/2, variant 0 and no sf, the cell cache then loads them instead of
 * forming them from field4 (same results, bit for bit, fewer instructions). */
typedef struct {
    double h;
    double cT;
    int32_t variant;
    const float *forms;
} tt_trace_opts;
int tt_trace_rk4(const float *field4, const tt_grid *grid, const tt_trace_opts *opts,
                 const double *s0, int64_t ld_s0, int64_t n, const int32_t *order,
                 double *rf, int64_t ld_rf, double *phase, double *sf, int64_t ld_sf,
                 int32_t *status, void *stream);

/* Rays in time, as the reference integrates them (particle_tracking/
 * particle_tracker.py:312-331, 398-419, then ray_at_exit PT:333-380), for the
 * rays the z-march cannot follow: per ray, the Dormand-Prince 5(4) pair of
 * the reference's RK45 in c t with its own step control at a tight tolerance
 * (1e-10 of a cell in position, of c in velocity; fp64 state and trilinear
 * weights, steps at most half the smallest cell edge, field zero outside the
 * grid) from the launch state to c t = opts->cT, then the back-projection to
 * the plane z = z.max() of ray_at_exit (so a ray that turned back is reported
 * as the reference reports it).  With status non-NULL only the rays whose
 * status equals only_status are traced (e.g. 1, the rays tt_trace_rk4 hands
 * over), and their status becomes 3 (2 for NaN input,
 * or for a ray given up after 2^18 attempted steps, with NaN outputs); with
 * status NULL every ray is traced.  order, rf, phase, sf as in
 * tt_trace_rk4; opts->h and variant are unused. */
int tt_trace_time(const float *field4, const tt_grid *grid, const tt_trace_opts *opts, const double *s0,
                  int64_t ld_s0, int64_t n, const int32_t *order, double *rf, int64_t ld_rf, double *phase,
                  double *sf, int64_t ld_sf, int32_t *status, int32_t only_status, void *stream);

/* tt_trace_rk4 over a ray list whose length is in device memory: the rays
 * order[0 .. *n_dev) are traced (n_dev: device int32, read by the kernel;
 * n_cap >= *n_dev bounds the launch), with no host synchronisation between
 * the producer of the list and the trace.  No sf output; status (or NULL) as
 * in tt_trace_rk4, written for the listed rays only, so that tt_trace_time
 * over all rays with only_status 1 then traces the listed rays the z-march
 * handed over (the other rays' entries must not be 1).  No reference
 * counterpart as a call: it serves the surrogate's fallback, which re-traces
 * the rays tt_surrogate_apply flags (the full solve, PT:312-380, on those
 * rays). */
int tt_trace_rk4_counted(const float *field4, const tt_grid *grid, const tt_trace_opts *opts, const double *s0,
                         int64_t ld_s0, int64_t n_cap, const int32_t *order, const int32_t *n_dev, double *rf,
                         int64_t ld_rf, double *phase, int32_t *status, void *stream);

/* Ensembles: the rays of many realisations (uneven grids and ray counts)
 * traced in ONE persistent launch.  No reference counterpart as a call: it
 * replaces running ElectronCube.solve once per realisation, as the reference's
 * drivers do in a loop or across MPI ranks (particle_tracking/
 * example_multiprocess.py:41-62, example_MPI.py:82-149).  Each item is one
 * tt_trace_rk4 call's arguments (no sf).  With a status array, the rays the
 * z-march hands over (status 1) are then traced in time, as tt_trace_time
 * does after tt_trace_rk4 (light-travel budget cT, m, as in tt_trace_opts),
 * in a second launch of the same call; without one they are NaN.
 * tt_batch_prepare uploads the
 * items and a chunk table (chunk rays per entry, items with the most z-steps
 * first) into `workspace` with synchronous copies and fills `info`;
 * tt_batch_trace is stream-ordered and allocation-free, so it can be captured
 * in a graph: it resets the chunk counter with a kernel and launches
 * blocks_per_cu (0 = the occupancy limit) x CUs blocks that take chunks until
 * none is left.  Results equal tt_trace_rk4's for the same variant. */
typedef struct {
    const float *field4;
    tt_grid grid;
    const double *s0;
    int64_t ld_s0, n;
    const int32_t *order;
    double *rf;
    int64_t ld_rf;
    double *phase;
    int32_t *status; /* per ray as tt_trace_rk4's, or NULL */
    double cT;       /* used with status: > 0 */
} tt_batch_item;

typedef struct {
    int32_t nitems, nchunks, chunk;
    int32_t steps_per_cell; /* m = 1/h */
} tt_batch_info;

typedef struct {
    int32_t block;         /* threads per block: 64, 128 or 256 */
    int32_t blocks_per_cu; /* persistent blocks per CU, 0 = occupancy limit */
    int32_t variant;       /* 0 cell-cached, 1 per-stage gathers (as tt_trace_opts) */
} tt_batch_tuning;

int tt_batch_workspace_bytes(const tt_batch_item *items, int32_t nitems, int32_t chunk, size_t *bytes);
int tt_batch_prepare(const tt_batch_item *items, int32_t nitems, double h, int32_t chunk, void *workspace,
                     size_t bytes, tt_batch_info *info);
int tt_batch_trace(void *workspace, const tt_batch_info *info, const tt_batch_tuning *tune, void *stream);

/* Trilinear samples of the node field at n points pts (3 x ld float64, m):
 * out (4 x ld_out float64) = the four field channels, zero strictly outside
 * the grid and NaN for NaN points -- the RegularGridInterpolator semantics of
 * ElectronCube.dndr (particle_tracking/particle_tracker.py:239-256). */
int tt_sample_field(const float *field4, const tt_grid *grid, const double *pts, int64_t ld, int64_t n,
                    double *out, int64_t ld_out, void *stream);

/* ------------------------------------------------------------------- beam
 * Replaces init_beam (particle_tracking/particle_tracker.py:258-310): same
 * distribution (position uniform in a disc of radius beam_size, polar angle
 * chi ~ N(0, divergence), azimuth phi ~ U[0, pi)), drawn from Philox4x32-10
 * with counter = ray index + offset.  axis: 0 = x, 1 = y, 2 = z. */
int tt_beam_sample(int64_t n, double beam_size, double divergence, double extent, int32_t axis,
                   uint64_t seed, uint64_t offset, double *s0, int64_t ld, void *stream);

/* Ray ordering for coalesced field gathers: a counting sort of the rays by
 * the Morton code of their entry column (cell >> brick_shift).  `scratch`
 * must hold tt_ray_sort_scratch_bytes bytes of device memory. */
int tt_ray_sort_scratch_bytes(const tt_grid *grid, int32_t brick_shift, size_t *bytes);
int tt_ray_sort(const double *s0, int64_t ld, int64_t n, const tt_grid *grid, int32_t brick_shift,
                void *scratch, int32_t *order, void *stream);
/* The same ordering with no scratch argument and no allocation: no
 * counterpart in the reference.  It is for callers that cannot allocate
 * device memory (code under graph capture, tight serving loops).  The column
 * counts live in a static device array of the library, 4 MiB per device
 * (2^20 entry-column bins: grids up to 1024 columns per axis after
 * brick_shift); larger grids are rejected with TT_ERR_ARG.  Every call on one
 * device shares that array, so calls that overlap on one device (on
 * different streams) race; calls queued on one stream do not. */
int tt_ray_sort_noalloc(const double *s0, int64_t ld, int64_t n, const tt_grid *grid, int32_t brick_shift,
                        int32_t *order, void *stream);
/* The same ordering, also moving the rays: out (6 x ldo float64) = s0's
 * columns in `order`.  staging (device, 8 n float64, 16-byte aligned) holds
 * one 64-byte record per ray in between, so the scattered writes are whole
 * records and the final rows are written coalesced. */
int tt_ray_sort_rays(const double *s0, int64_t ld, int64_t n, const tt_grid *grid, int32_t brick_shift,
                     void *scratch, int32_t *order, double *staging, double *out, int64_t ldo, void *stream);

/* ------------------------------------------------------------- detectors
 * Replaces the ray_transfer_matrix optics and Rays.histogram
 * (particle_tracking/ray_transfer_matrix.py:37-154, 173-191, 208-299).
 * An optical train is a list of ops applied in order to [x, theta, y, phi]
 * in mm/rad; a rejected ray becomes NaN (RTM:77-78). */
enum {
    TT_OP_DISTANCE = 1, /* a = d                          RTM:62-71   */
    TT_OP_LENS = 2,     /* a = f1, b = f2                 RTM:42-60   */
    TT_OP_APERTURE = 3, /* reject r^2 > a^2               RTM:73-79   */
    TT_OP_STOP = 4,     /* reject r^2 < a^2               RTM:81-87   */
    TT_OP_KNIFE = 5,    /* axis 0=x 2=y, a = offset, b = direction sign RTM:138-154 */
    TT_OP_RECT = 6,     /* reject x^2 > a^2 and y^2 > b^2 RTM:128-136 */
    TT_OP_ANNULUS = 7   /* reject a^2 < r^2 < b^2          RTM:89-111  */
};
#define TT_MAX_OPS 24
typedef struct {
    int32_t op, axis;
    double a, b;
} tt_optic;
typedef struct {
    int32_t n_ops;
    int32_t to_mm; /* 1: scale rows 0 and 2 by 1e3 first (m_to_mm, RTM:37-40) */
    tt_optic ops[TT_MAX_OPS];
} tt_train;

/* np.histogram2d over [xmin, xmax] x [ymin, ymax] with nbx x nby bins,
 * last bin closed, NaN dropped, stored transposed H[iy][ix] (RTM:188-191). */
typedef struct {
    int32_t nbx, nby;
    double xmin, xmax, ymin, ymax;
} tt_hist;

/* Apply a train to n rays: r_out may alias r_in. */
int tt_optics_apply(const tt_train *train, const double *r_in, int64_t ld_in, int64_t n, double *r_out,
                    int64_t ld_out, void *stream);

/* One detector of the fused detector pass.  mode 0: counts (uint32);
 * mode 1: interferogram, weight 1 + cos(phase - tilt_k * x_det) with x_det in
 * mm and tilt_k in rad/mm, summed as float32. */
typedef struct {
    tt_train train;
    tt_hist hist;
    int32_t mode;
    double tilt_k;
    uint32_t *counts; /* device, nby*nbx, mode 0 */
    float *wsum;      /* device, nby*nbx, mode 1 */
} tt_detector;
#define TT_MAX_DETECTORS 4
/* Runs up to TT_MAX_DETECTORS detectors over the same rays in one pass and
 * ADDS into their histograms.  phase may be NULL when no detector uses it. */
int tt_detect(const tt_detector *dets, int32_t ndet, const double *rf, int64_t ld, const double *phase,
              int64_t n, void *stream);

/* -------------------------------------------------------------- surrogate
 * The learned correction.  None of these has a counterpart in the reference
 * (it has no surrogate); they serve the cheap path estimate, the MLP that
 * corrects it, and its training. */

/* Straight-line z-block integrals of the node field for tt_path_features:
 * block b covers cells [b k, min((b + 1) k, nz - 1)); per node column,
 * G = sum over its cells of the cell's mid-plane value (the mean of the
 * cell's two node planes, the exact z-integral of the trilinear field) and
 * M = the exact first moment about the block centre,
 * sum_c (c + 1/2 - zmid) mid_c + (f[b1] - f[b0]) / 12.  GM4 is
 * [ny][nx][nb][2] float4 ({G, M} per column and block, a column's blocks
 * consecutive), nb = ceil((nz - 1) / k). */
int tt_field_blocks(const float *field4, const tt_grid *grid, int32_t k, float *GM4, void *stream);

/* Cheap path estimate of each ray's exit state and its 48 surrogate
 * features: one drift-kick-drift step per z-block of k cells (the block
 * integral G at the mid-block position, plus v . grad(M) for the ray's
 * transverse drift in the block; the exit position gets -s M_x / W for a
 * kick spread over the block; G, M from tt_field_blocks), against the eight
 * field evaluations per z-cell
 * of tt_trace_rk4.  A ray entering inside the grid takes single-cell steps
 * from field4 up to the next block boundary.  order (nullable) as in
 * tt_trace_rk4.  base: 5 x ldb float64 [x, theta, y, phi, phase]; feat: 48
 * float32 features per ray stored feature-major, feature q of ray r at
 * feat[q * ldf + r] (ldf >= n): 16 z-segment sums of the x kicks, 16 of the y kicks,
 * then U - U0, V - V0, W - W0, the exit displacement from the entry line (x,
 * y, cells), the phase, U0, V0, W0 - 1, the entry-cell fractions x, y, the x
 * and y column changes between steps, the summed second-order proxies
 * (dGx/dx Gx + dGx/dy Gy) / W^2 and (dGy/dx Gx + dGy/dy Gy) / W^2, and the
 * path length in cells.  With k = 1, GM4 may be NULL: the steps then take
 * a cell's G and M from its two node planes of field4 (the values
 * tt_field_blocks stores for k = 1), so no block fields are needed. */
int tt_path_features(const float *field4, const float *GM4, int32_t k, const tt_grid *grid, const double *s0,
                     int64_t ld, int64_t n, const int32_t *order, float *feat, int64_t ldf, double *base, int64_t ldb,
                     void *stream);

/* Apply the MLP output Y (n rows, ld ldy >= 7, fp32) to the cheap path
 * estimate: rf = base[0:4] + Y[:, {0, 2, 1, 3}] * {s_pos, s_ang, s_pos, s_ang}
 * and phase = base[4] + Y[:, 4] * s_ph, with scales = {s_pos, s_ang, s_ph,
 * e_pos, e_ang}.  A ray is flagged for the full tracer when
 * |Y[:, 5]| e_pos > thresholds[0] (metres) or |Y[:, 6]| e_ang > thresholds[1]
 * (radians) or the estimate is not finite; flag (optional) gets 0/1 per ray
 * and first + (index of each flagged ray) is appended to list (optional) with
 * *nflag (device int32, caller-zeroed) counting them, so batches of one ray
 * set can share one list.  Each wave of 64 rays takes consecutive list slots
 * for its flagged rays, in ray order. */
int tt_surrogate_apply(const double *base, int64_t ldb, const float *Y, int64_t ldy, int64_t n, const float *scales,
                       const double *thresholds, double *rf, int64_t ldr, double *phase, int32_t *flag,
                       int32_t *list, int32_t *nflag, int64_t first, void *stream);

/* Fused inference of the surrogate network nfeat -> 256 -> 256 -> 256 -> nout
 * (GELU, tanh form, on the hidden layers; bf16 MFMA, fp32 accumulation):
 * Y[r][o] for fp32 features feat[r * srow + f * scol] standardised as
 * (x - mu[f]) * inv_sd[f] and rounded to bf16.  nfeat <= 64, nout <= 16.
 * w_packed: bf16 fragments, layer by layer (tiles x chunks 16 x 2, 16 x 8,
 * 16 x 8, 1 x 8), fragment (nt, c) = 64 lanes x 8 with lane l element j =
 * W[16 nt + l%16][k], k = 32c + 8(l/16) + j for the first layer and
 * 32c + 16(j/4) + 4(l/16) + j%4 for the others (zero beyond the layer's
 * size); bias_packed: fp32 256 + 256 + 256 + 16.  Both 16-byte aligned. */
int tt_mlp_infer(const float *feat, int64_t srow, int64_t scol, int64_t n, int32_t nfeat, const float *mu,
                 const float *inv_sd,
                 const void *w_packed, const float *bias_packed, float *Y, int64_t ldy, int32_t nout, void *stream);

/* C[m][n] = act(sum_k A(m,k) B(k,n) + bias[n]) on bf16 MFMA with fp32
 * accumulation; A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn] (bf16).
 * gelu: 1 applies GELU (tanh form, 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))); Zpre (optional) keeps the pre-activation;
 * outputs fp32 Cf and/or bf16 Cb with row stride ldc.  ksplit > 0 splits K
 * into chunks whose raw products are ADDED atomically into Cf (no bias,
 * activation or bf16 output). */
typedef struct {
    const void *A;
    int64_t sam, sak;
    const void *B;
    int64_t sbk, sbn;
    int32_t M, N, K;
    const float *bias;
    int32_t gelu;
    float *Cf;
    void *Cb;
    int64_t ldc;
    float *Zpre;
    int32_t ksplit;
} tt_gemm_args;
int tt_gemm_bf16(const tt_gemm_args *args, void *stream);

/* dZ (bf16) = dH (fp32 or bf16) * GELU'(Zpre), elementwise over n */
int tt_gelu_backward(const float *dH_f32, const void *dH_bf16, const float *Zpre, void *dZ_bf16, int64_t n,
                     void *stream);
/* db[c] += sum over rows of dZ[r][c] (bf16 input, fp32 accumulate) */
int tt_bias_grad(const void *dZ_bf16, int64_t rows, int32_t cols, int64_t ld, float *db, void *stream);
/* Surrogate loss: weighted squared error on columns [0, n_reg) of Y against
 * T, plus pinball-loss error heads: column n_reg + h learns the tau-quantile
 * (0 < tau < 1) of the corrected error max |T - Y| over columns
 * [err_lo[h], err_hi[h]), i.e. a bound that tau of the rays stay under.
 * Writes the gradient / rows to dY (bf16, ld ldd, zero up to n_pad) and ADDS
 * the two loss sums to loss_out (2 doubles). */
int tt_surrogate_loss(const float *Y, int64_t ldy, const float *T, int64_t ldt, void *dY_bf16, int64_t ldd,
                      int64_t rows, int32_t n_reg, const float *w_reg, int32_t n_err, const int32_t *err_lo,
                      const int32_t *err_hi, double tau, int32_t n_pad, double *loss_out, void *stream);
/* Adam (bias-corrected, step >= 1) on fp32 master weights; refreshes the
 * bf16 copy w_bf16 when non-NULL. */
int tt_adam(float *w, const float *g, float *m, float *v, void *w_bf16, int64_t n, double lr, double b1, double b2,
            double eps, int64_t step, double weight_decay, void *stream);
/* out (bf16, [rows][ldo], zero-padded) = (x - mu) * inv_sd over cols, with
 * x element (r, c) at x[r * srow + c * scol] */
int tt_standardize_bf16(const float *x, int64_t srow, int64_t scol, int64_t rows, int32_t cols, const float *mu,
                        const float *inv_sd, void *out_bf16, int64_t ldo, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* TT_H */
