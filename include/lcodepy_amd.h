/*
 * lcodepy_amd -- C ABI of the MI355X (gfx950) quasistatic 3D xi step.
 *
 * Each entry point stands in for one factory closure of lcodePy's
 * plasma3d / beam3d packages, so the Python host layer can swap the
 * reference's Numba / CuPy kernels for these one call site at a time:
 *
 *   lc3d_move_wo_fields   plasma3d/move.py:331-349  (move_particles_wo_fields)
 *   lc3d_move_smart       plasma3d/move.py:351-381  (move_particles_smart)
 *   lc3d_deposit          plasma3d/weights.py:249-285 (deposit_plasma) for one
 *                         sort, with rhoj.py:11-20's background-ion column
 *   lc3d_compute_fields   plasma3d/fields.py:276-322 (compute_fields)
 *   lc3d_plasma_step      plasma3d/solver.py:25-78   (Plane2d3vPlasmaSolver.step_dxi)
 *   lc3d_deposit_beam     beam3d/weights.py:152-163  (deposit_beam)
 *   lc3d_layout_beam_layer beam3d/beam_calculator.py:82-93 (layout_beam_layer)
 *   lc3d_beam_substep     beam3d/beam_calculator.py:97-114 (start_moving_layer)
 *   lc3d_move_beam        beam3d/move.py:294-309     (move_beam_particles)
 *   lc3d_energy_flux      diagnostics/FXi/Strategy3D.py:6-25 (calculate_energy_fluxes)
 *   lc3d_probe_sample     diagnostics/FXi/Strategy3D.py:59-82 (the f(xi) probe values)
 *   lc3d_abs_max          diagnostics/FXi/Strategy3D.py:84-95 (process_chaotic*)
 *
 * Conventions
 *   - All pointers are device pointers (hipMalloc'd or torch-ROCm tensors),
 *     C-contiguous.  (N, N) grids are indexed [ix, iy] -> ix * N + iy;
 *     current grids are (2, N, N) with column = sort (0 electrons, 1 ions),
 *     as plasma3d/initialization.py:391 lays them out.
 *   - Precision is fixed per context:
 *       LC3D_FP64       every array float64 (the reference's precision);
 *       LC3D_FP32_BULK  x_init, y_init, x_offt, y_offt, pz and the fields of
 *                       layers k / k+1 (fields_prev, fields_full) float64;
 *                       px, py, q, m, currents and the half-step fields
 *                       float32.  pz stays float64 because gamma - pz, which
 *                       sets 1/(1 - v_z) and the position update, cancels
 *                       for forward-moving electrons.
 *     Beam arrays are float64 (ids and step counters int64, masks uint8) in
 *     both forms.
 *   - `stream` is a hipStream_t (NULL = default stream).  Work is
 *     stream-ordered; no call synchronises the host except
 *     lc3d_error_flags.
 *   - Return value: 0 on success, otherwise an LC3D_E* code; the message is
 *     available from lc3d_last_error(ctx).  The library never frees a
 *     caller's buffer and allocates only inside lc3d_create.
 *   - One context per host thread; calls on one context are not reentrant.
 */
#ifndef LCODEPY_AMD_H
#define LCODEPY_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LC3D_FP64 0
#define LC3D_FP32_BULK 1

/* Field-solve implementations (lc3d_params.field_solver). */
#define LC3D_SOLVER_MATRIX 0  /* DST-I/DCT-I as products with N x N transform matrices */
#define LC3D_SOLVER_FFT 1     /* DST-I/DCT-I as LDS FFTs; only in builds that define
                                 LC3D_FFT_SOLVER (the GPU library does not) */
#define LC3D_SOLVER_RFFT 2    /* DST-I/DCT-I as paired complex LDS FFTs of length 2(N-1) */

#define LC3D_OK 0
#define LC3D_EINVAL 1    /* bad argument (null pointer, size beyond the context) */
#define LC3D_EHIP 2      /* a HIP runtime call or kernel launch failed */
#define LC3D_ENOMEM 3    /* device allocation in lc3d_create failed */

/* Sticky device error bits (lc3d_error_flags). */
#define LC3D_ERR_PARTICLE_OFF_GRID 1u  /* plasma stencil left the grid: not deposited */
#define LC3D_ERR_NONFINITE 2u          /* non-finite plasma position: not deposited */
#define LC3D_ERR_BEAM_OFF_GRID 4u      /* beam stencil partly off the grid / substep cap hit */
#define LC3D_ERR_CELL_OVERFLOW 8u      /* > 16384 particles in one cell (summed in arrival order) */
#define LC3D_ERR_BEAM_RANGE 16u        /* beam charge contribution >= 2^40: not deposited */

typedef struct lc3d_ctx lc3d_ctx;

/* Geometry and numerics of one run; the values of the reference's config
 * keys after its 3D adjustment (config/config.py:185-234). */
typedef struct lc3d_params {
    int32_t grid_steps;          /* window-width-steps N: odd, N - 1 <= 1024 */
    double grid_step_size;       /* transverse-step h */
    double xi_step_size;         /* xi-step */
    double boundary;             /* h (N/2 - bound-padding-steps), move.py:321 */
    double subtraction_coeff;    /* field-solver-subtraction-coefficient */
    double lost_radius;          /* max(0.9 R, R - 1), R = h N / 2, beam3d/move.py:285 */
    int32_t precision;           /* LC3D_FP64 or LC3D_FP32_BULK */
    int64_t max_particles;       /* largest plasma sort the context will deposit */
    /* radix plan of the FFT: of length N-1 for LC3D_SOLVER_FFT, of length
     * 2(N-1) for LC3D_SOLVER_RFFT (radices 2 3 4 5 7 8 11 13 16);
     * 0 stages = library default */
    int32_t fft_nstages;
    int32_t fft_radix[12];
    /* host array of the N eigenvalues 4/h^2 sin^2(k pi / (2(N-1))) used by
     * the solver matrices (initialization.py:24, 43-44, 63); NULL = computed
     * here with the C library's sin */
    const double *eigenvalues;
    int32_t field_solver;        /* LC3D_SOLVER_MATRIX, LC3D_SOLVER_FFT or LC3D_SOLVER_RFFT */
    /* LC3D_SOLVER_RFFT launch shape: LDS of a block in complex values, two
     * buffers of half of it (float32 solve: 2048, 4096 or 8192; float64:
     * 1024, 2048 or 4096; 0 = the largest) and transforms per block, each
     * taking 2(N-1) values of a buffer (0 = as many as fit) */
    int32_t rfft_cap;
    int32_t rfft_tpb;
    /* declustering noise filter of the electrons (plasma3d/noise_filter.py:
     * 82-238; declustering-enabled): side of their (side, side) lattice, 0 =
     * off; the circular Gaussian smoothing as float64 taps for lattice
     * offsets tap0 .. tap0 + ntaps - 1 (host array, copied); the
     * declustering-force, the damping (read by the reference from the key
     * 'damping-declustering') and declustering-limit */
    int32_t declustering_side;
    int32_t declustering_ntaps;
    int32_t declustering_tap0;
    const double *declustering_taps;
    double declustering_force;
    double declustering_damping;
    double declustering_limit;
    /* LC3D_SOLVER_RFFT: the kernels in which each transform is done by one
     * wave in registers, four per block (2(N-1) = 512, 1024 or 2048 only), a
     * mask: 1 rows forward, 2 columns, 4 rows inverse; the radix plan and the
     * launch shape above serve the others (unused at 7) */
    int32_t rfft_wave;
} lc3d_params;

/* Allocates every workspace (cell bins, spectral scratch, transform
 * matrices or twiddles, eigenvalues) on the current HIP device. */
int lc3d_create(const lc3d_params *params, lc3d_ctx **out);
void lc3d_destroy(lc3d_ctx *ctx);
const char *lc3d_last_error(const lc3d_ctx *ctx);
int lc3d_precision(const lc3d_ctx *ctx);

/* Copies the sticky device error bits to *flags and clears them
 * (synchronises `stream`). */
int lc3d_error_flags(lc3d_ctx *ctx, uint32_t *flags, void *stream);

/* Measurement, with no counterpart in the reference: device time of the
 * three heaviest stages, from HIP events recorded on the launch stream around
 * them.  enable != 0 starts it (and clears the totals); 0 stops it.
 * lc3d_kernel_times waits for the recorded events and gives, for [smart
 * push, whole deposit (cell sort and gather, one launch each), solve
 * transform launch (a GEMM of the matrix solve, one of the three FFT kernels
 * of the RFFT solve)], the average seconds per launch and the launch counts
 * since timing was enabled. */
#define LC3D_KC_PUSH 0
#define LC3D_KC_DEPOSIT 1
#define LC3D_KC_SOLVE 2
int lc3d_kernel_timing(lc3d_ctx *ctx, int32_t enable);
int lc3d_kernel_times(lc3d_ctx *ctx, double seconds[3], int64_t launches[3]);

/* ---- field-solve autotuner -------------------------------------------------
 * No counterpart in the reference (its solve is fixed: scipy.fftpack on the
 * CPU, padded cuFFT on the GPU, plasma3d/fields.py:12-109). */
typedef struct lc3d_solve_choice {
    int32_t field_solver;        /* LC3D_SOLVER_MATRIX or LC3D_SOLVER_RFFT */
    int32_t fft_nstages;         /* RFFT radix plan (length 2(N-1)) */
    int32_t fft_radix[12];
    int32_t rfft_cap;            /* RFFT LDS size and transforms per block */
    int32_t rfft_tpb;
    double seconds;              /* device time of one lc3d_compute_fields (< 0: refused) */
    int32_t rfft_wave;           /* RFFT kernels by waves (lc3d_params.rfft_wave) */
} lc3d_solve_choice;
/* Times lc3d_compute_fields for the geometry and precision of *params on
 * synthetic inputs, `repeats` times after two warm-up calls, for the matrix
 * solve and the RFFT solve over its radix plans x LDS sizes x transforms per
 * block and, where the length allows, by waves; writes the fastest to *best and, when table != NULL, the first
 * table_cap candidates in the order tried (*n_tried = how many were timed).
 * Creates its own contexts and inputs and synchronises the device: a set-up
 * call, not for the xi loop.  Copy *best's fields into lc3d_params to use it. */
int lc3d_autotune_solve(const lc3d_params *params, int32_t repeats, lc3d_solve_choice *best,
                        lc3d_solve_choice *table, int32_t table_cap, int32_t *n_tried);
/* ---- plasma particles ------------------------------------------------------ */
typedef struct lc3d_fields6 {   /* the six fields a push reads */
    const void *Ex, *Ey, *Ez, *Bx, *By, *Bz;
} lc3d_fields6;

typedef struct lc3d_sort {
    int64_t n;                         /* macro-particles of this sort */
    const double *x_init, *y_init;     /* constant lattice positions */
    const void *m;                     /* mass, constant */
    const double *x_offt_prev, *y_offt_prev;   /* layer k */
    const void *px_prev, *py_prev, *q_prev;    /* bulk type */
    const double *pz_prev;
    double *x_offt, *y_offt;                   /* layer k+1 (written) */
    void *px, *py, *q;                         /* bulk type */
    double *pz;
    int32_t column;                    /* current column of this sort */
    /* chaotic displacements of the noise filter (float64, n each): layer k
     * read, layer k+1 written; NULL for a sort the filter does not touch */
    const double *dx_chaotic_prev, *dy_chaotic_prev;
    double *dx_chaotic, *dy_chaotic;
} lc3d_sort;

/* x_offt += px / (gamma - pz) * xi_step where m != 0 (in place). */
int lc3d_move_wo_fields(lc3d_ctx *ctx, int64_t n, const void *m, double *x_offt, double *y_offt,
                        const void *px, const void *py, const double *pz, void *stream);

/* Predictor-corrector push of one sort: reads the *_prev arrays and the
 * current estimate in x_offt / y_offt / q, writes x_offt, y_offt, px, py, pz
 * and q (absorbed particles get q = 0).  field_fp32 != 0 when the fields
 * are float32 (the half-step fields of the FP32-bulk form). */
int lc3d_move_smart(lc3d_ctx *ctx, const lc3d_sort *sort, const lc3d_fields6 *fields,
                    int32_t field_fp32, void *stream);

/* Deposits one sort's layer-k+1 state (x_offt, ..., q) into the (N, N)
 * grids ro, jx, jy, jz (overwritten).  With bg_ro != NULL also writes the
 * background-ion column: ro1 = bg_ro, jx1 = jy1 = jz1 = 0. */
int lc3d_deposit(lc3d_ctx *ctx, const lc3d_sort *sort, void *ro, void *jx, void *jy, void *jz,
                 const void *bg_ro, void *ro1, void *jx1, void *jy1, void *jz1, void *stream);

/* noise_filter(particles, particles_prev), plasma3d/noise_filter.py:102-238,
 * of one sort whose n is the context's declustering_side squared: kicks px,
 * py in place from x_offt, y_offt and writes dx_chaotic, dy_chaotic (the
 * sort's chaotic pointers must be set).  lc3d_plasma_step applies it to
 * sorts[0] after the last push when the context has the filter and sorts[0]
 * has the chaotic pointers (plasma3d/solver.py:69-73). */
int lc3d_noise_filter(lc3d_ctx *ctx, const lc3d_sort *sort, void *stream);
/* ---- fields ------------------------------------------------------------------- */
typedef struct lc3d_fields {   /* Ex Ey Ez Bx By Bz Phi, each (N, N) */
    void *Ex, *Ey, *Ez, *Bx, *By, *Bz, *Phi;
} lc3d_fields;

typedef struct lc3d_currents { /* ro jx jy jz, each (2, N, N) */
    void *ro, *jx, *jy, *jz;
} lc3d_currents;

/* compute_fields(fields, fields_prev, const, rho_beam_full, rho_beam_prev,
 * currents_prev, currents_full) -> (fields_full, fields_half).  `pred` is
 * the Helmholtz predictor (only Ex Ey Bx By are read); pred_fp32 != 0 when
 * it is float32. */
int lc3d_compute_fields(lc3d_ctx *ctx, const lc3d_fields *pred, int32_t pred_fp32,
                        const lc3d_fields *fields_prev, const double *rho_beam_full,
                        const double *rho_beam_prev, const lc3d_currents *currents_prev,
                        const lc3d_currents *currents_full, lc3d_fields *fields_full,
                        lc3d_fields *fields_half, void *stream);

/* One whole xi step of the plasma: free stream, 3 pushes, 3 deposits and 2
 * field solves in the reference's order.  Writes every sort's layer-k+1
 * arrays, fields_full, fields_half and currents_full (column 1 holds
 * bg_ro when there is one sort and bg_ro != NULL). */
int lc3d_plasma_step(lc3d_ctx *ctx, int32_t n_sorts, const lc3d_sort *sorts,
                     const lc3d_fields *fields_prev, const lc3d_currents *currents_prev,
                     const double *rho_beam_full, const double *rho_beam_prev, const void *bg_ro,
                     lc3d_fields *fields_full, lc3d_fields *fields_half,
                     const lc3d_currents *currents_full, void *stream);
/* The beam charge densities rho_beam_* are float64 (N, N) in both forms. */

/* ---- beam ----------------------------------------------------------------------- */
/* Adds q_norm * w * xi_loc into ro_running and q_norm * w * (1 - xi_loc)
 * into ro_next for the n particles of a layer (both accumulate).  Each
 * cell's sum is formed exactly (fixed point), so the result does not depend
 * on the order of the additions. */
int lc3d_deposit_beam(lc3d_ctx *ctx, int64_t n, int64_t plasma_layer_idx, const double *x,
                      const double *y, const double *xi, const double *q_norm,
                      double *ro_running, double *ro_next, void *stream);
/* The layer step of BeamCalculator.layout_beam_layer in one call: deposits
 * as lc3d_deposit_beam, then ro_running (the layer, with what the previous
 * layer's particles left in it) becomes its finished density divided by h^2,
 * and ro_next is overwritten with this layer's share of the next layer (its
 * previous contents are not read). */
int lc3d_layout_beam_layer(lc3d_ctx *ctx, int64_t n, int64_t plasma_layer_idx, const double *x,
                           const double *y, const double *xi, const double *q_norm,
                           double *ro_running, double *ro_next, void *stream);

/* dt = mult * time_step and remaining_steps = int(1 / mult) for n particles. */
int lc3d_beam_substep(lc3d_ctx *ctx, int64_t n, const double *q_m, const double *pz,
                      double substepping_energy, double time_step, double *dt,
                      int64_t *remaining_steps, void *stream);

/* Moves a layer through fields_k (previous plasma layer) and fields_k1 (the
 * one just computed); fields are float64.  lost/moved/fell are written. */
int lc3d_move_beam(lc3d_ctx *ctx, int64_t n, int64_t beam_layer_idx, const double *q_m,
                   const double *dt, const lc3d_fields6 *fields_k, const lc3d_fields6 *fields_k1,
                   int64_t *remaining_steps, int64_t *id, double *x, double *y, double *xi,
                   double *px, double *py, double *pz, uint8_t *lost, uint8_t *moved,
                   uint8_t *fell, void *stream);

/* The split of a moved layer into (lost, moved, fell), beam_calculator.py:
 * 124-139 (the three boolean-mask gathers of move_beam_layer): f holds nf
 * float64 rows and iv ni int64 rows of n particles (row r at f + r n),
 * flags three uint8 rows (lost, moved, fell; what lc3d_move_beam wrote).
 * Group g's particles go, in layer order, to the front of the rows of
 * out_f + g nf n and out_i + g ni n; their number to counts[g] (device
 * int64[3]). */
int lc3d_beam_split(lc3d_ctx *ctx, int64_t n, int32_t nf, int32_t ni, const double *f,
                    const int64_t *iv, const uint8_t *flags, double *out_f, int64_t *out_i,
                    int64_t *counts, void *stream);
/* Total energy flux of a plasma layer, diagnostics/FXi/Strategy3D.py:6-25
 * (calculate_energy_fluxes): the sum over the grid of (Ez^2 + Bz^2 +
 * (Ex - By)^2 + (Ey + Bx)^2) h^2 / 2 plus the sum over the n electrons of
 * sqrt(m^2 + px^2 + py^2 + pz^2) - m, with h the context's transverse step.
 * fields: the float64 (N, N) Ex..Bz; m, px, py of the bulk type, pz float64.
 * Writes one float64 to *psi (device memory), in a fixed summation order,
 * without synchronising the host. */
int lc3d_energy_flux(lc3d_ctx *ctx, const lc3d_fields6 *fields, int64_t n, const void *m,
                     const void *px, const void *py, const double *pz, double *psi,
                     void *stream);

/* Allocation-free form of lc3d_energy_flux.  It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops).  It needs no context: N, h and
 * the precision of m, px, py (LC3D_FP64 or LC3D_FP32_BULK) are arguments,
 * and the partial sums live in a static device array of the library
 * (2 x LC3D_NOALLOC_FLUX_BLOCKS float64 = 16 KiB per device, kept until the
 * process ends), so the call allocates nothing.  Limits: N <= 1025 and
 * n <= 2^31 - 1; beyond them it returns LC3D_EINVAL (message from
 * lc3d_last_error(NULL)).  The result equals lc3d_energy_flux's bit for bit.
 * Every call on a device shares that device's array: calls that overlap on
 * one device (other streams or threads) race, so order them on one stream. */
#define LC3D_NOALLOC_FLUX_BLOCKS 1024
#define LC3D_NOALLOC_FLUX_MAX_N 1025
#define LC3D_NOALLOC_FLUX_MAX_PARTICLES 2147483647LL
int lc3d_energy_flux_noalloc(int32_t N, double h, int32_t precision, const lc3d_fields6 *fields,
                             int64_t n, const void *m, const void *px, const void *py,
                             const double *pz, double *psi, void *stream);
/* ---- f(xi) probes ------------------------------------------------------------------ */
#define LC3D_PROBE_MAX_SOURCES 16
typedef struct lc3d_probe_sources {
    int32_t n;                                  /* sources used, <= LC3D_PROBE_MAX_SOURCES */
    const void *src[LC3D_PROBE_MAX_SOURCES];    /* (N, N) grids (a current column is one) */
    int32_t is_f32[LC3D_PROBE_MAX_SOURCES];     /* 1: float32 grid, 0: float64 */
} lc3d_probe_sources;

/* One xi step of diagnostics/FXi/Strategy3D.py:59-82 (process_field,
 * process_n, process_rho, process_Phi): every source at the n_points probe
 * cells (ix[p], iy[p]) (device int32 arrays) into out[s * n_points + p]
 * (device, float64), without synchronising; a cell off the grid reads NaN. */
int lc3d_probe_sample(lc3d_ctx *ctx, const lc3d_probe_sources *sources, int32_t n_points,
                      const int32_t *ix, const int32_t *iy, double *out, void *stream);

/* Strategy3D.py:84-95 (process_chaotic, process_chaotic_perp): max |a| over
 * rows [r0, r1) and columns [c0, c1) of an (nr, nc) float64 device array
 * into *out (device), without synchronising. */
int lc3d_abs_max(lc3d_ctx *ctx, const double *a, int64_t nr, int64_t nc, int64_t r0, int64_t r1,
                 int64_t c0, int64_t c1, double *out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* LCODEPY_AMD_H */
