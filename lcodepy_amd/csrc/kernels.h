// Host-side launch interfaces of the lcodepy_amd kernels (gfx950).
//
// TP is the bulk particle type (px, py, q, m): double in the FP64 form, float
// in the FP32-bulk form.  Positions (x_init, y_init, x_offt, y_offt) and pz
// are always double: gamma - pz, which sets 1/(1 - v_z) and the position
// update, cancels catastrophically for forward-moving electrons unless pz is
// kept in float64.  TF is the type of the field arrays a push reads, TG the
// type of the current / spectral grids.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace lc3d {

template <typename T> struct Fields6 {
    const T *Ex, *Ey, *Ez, *Bx, *By, *Bz;
};

struct PushGeom {
    int64_t n;        // particles in this sort
    int N;            // grid steps
    double h;         // transverse step
    double xi_step;
    double boundary;  // plasma reflection boundary
};

// Per-sort particle arrays of the previous (k) and next (k+1) layer.
template <typename TP> struct SortArrays {
    const double *x_init, *y_init;
    const TP *m;
    const double *xo_prev, *yo_prev;
    const TP *px_prev, *py_prev, *q_prev;
    const double *pz_prev;
    double *xo, *yo;
    TP *px, *py, *q;
    double *pz;
};

// Tiles of the cell-count scan (2048 cells each) covering cells 0..ncell.
constexpr int SCAN_TILE_CELLS = 2048;
inline int scan_tiles(int64_t ncell) { return (int)((ncell + SCAN_TILE_CELLS) / SCAN_TILE_CELLS); }

// Binning state shared by push and deposit (owned by the context).
struct BinState {
    int *count;       // N*N + 1, zero between deposits (the scan and the list
                      // sort clear them)
    int *start;       // N*N + 1: each cell's list start (written by k_order_cells)
    int *lstart;      // N*N + 1: each cell's list start within its scan tile
    int *base;        // per scan tile: the sum of the tiles before it (zero
                      // between deposits: the gather clears it)
    int nbase;        // scan tiles, scan_tiles(N*N)
    // key, rank, pid and rec are indexed by binning position: the thread
    // index j of the pass that binned the particle (the push visits particle
    // order[j], k_bin particle j).
    int *key;         // per position: cell index or -1
    int *rank;        // per position: its place among the particles counted into
                      // the same cell (or among the unbinned ones, counted in
                      // count[N*N]), in arrival order
    int *pid;         // per position: the particle binned there
    int *pos;         // per list slot: binning positions sorted by cell (each
                      // cell's list ascending), then the unbinned ones
    int *idx;         // per list slot: the particle ids of pos (the cell order
                      // the next push of this sort visits)
    unsigned *err;    // sticky error bits
    void *rec;        // per position: deposit records (deposit.hip)
    int64_t cap;      // particles the per-particle arrays hold
    int *big;         // dense cells (more than 16 particles), in queue order
    int *nbig;        // how many cells `big` holds
    int bigcap;       // capacity of `big`
    int *dslot;       // N*N: a cell's slot in `big`, or -1
    void *dense;      // per dense cell: its 5 x 5 x 4 deposit (deposit.hip)
};

// What the deposit needs of one binned plasma particle (weights.py:35-67):
// offsets xl, yl in [-1/2, 1/2) within its cell, dro = q / (1 - pz/gamma)
// and dro p / gamma; written at the particle's binning position.  Six values
// and no padding (24 bytes in float32, 48 in float64): every binning push
// writes one per particle and the gather reads them all.
template <typename T> struct alignas(2 * sizeof(T)) DepRec {
    T xl, yl, c0, c1, c2, c3;
};
static_assert(sizeof(DepRec<float>) == 24 && sizeof(DepRec<double>) == 48,
              "deposit records are six values, unpadded");

// Capacity of the dense-cell queue: cells with more than 16 of the at most
// `cap` particles.
inline int dense_cell_capacity(int64_t cap) { return (int)(cap / 16 + 2); }

inline size_t deposit_record_bytes(bool fp32_bulk)
{
    return fp32_bulk ? sizeof(DepRec<float>) : sizeof(DepRec<double>);
}

// Device time of the heaviest stages, measured with HIP events recorded on
// the launch stream around them (off unless enabled on the context): one
// smart push, one whole deposit (cell sort and gather, counted as one
// launch) and the solve's transform launches (the GEMMs of the matrix solve,
// the three FFT kernels of the LDS-FFT solve).  An interval may span several
// back-to-back launches of a class.
enum KernelClass { KC_PUSH = 0, KC_DEPOSIT = 1, KC_SOLVE = 2, KC_COUNT = 3 };

struct KernelTimer {
    static constexpr int CAP = 64;   // intervals recorded before they are read back
    bool on = false;
    hipEvent_t ev[2 * CAP] = {};
    int cls[CAP] = {}, nl[CAP] = {};
    int used = 0;
    double seconds[KC_COUNT] = {};
    int64_t launches[KC_COUNT] = {};

    hipError_t begin(int c, int nlaunch, hipStream_t s) {
        cls[used] = c;
        nl[used] = nlaunch;
        return hipEventRecord(ev[2 * used], s);
    }
    hipError_t end(hipStream_t s) {
        hipError_t e = hipEventRecord(ev[2 * used + 1], s);
        if (e != hipSuccess) return e;
        if (++used == CAP) return harvest();
        return hipSuccess;
    }
    // Waits for the recorded intervals and adds them up.
    hipError_t harvest() {
        for (int i = 0; i < used; ++i) {
            hipError_t e = hipEventSynchronize(ev[2 * i + 1]);
            if (e != hipSuccess) return e;
            float ms = 0;
            e = hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]);
            if (e != hipSuccess) return e;
            seconds[cls[i]] += 1e-3 * ms;
            launches[cls[i]] += nl[i];
        }
        used = 0;
        return hipSuccess;
    }
};

// plasma3d/move.py:151-165 (free-streaming estimate).
template <typename TP>
hipError_t launch_move_wo_fields(int64_t n, double xi_step, const TP *m, double *xo, double *yo,
                                 const TP *px, const TP *py, const double *pz, hipStream_t s);

// plasma3d/move.py:52-148.  first=true folds in the free-streaming estimate
// from the previous layer (move.py:331-349); bin=true also records each
// particle's deposit cell for the deposit that follows.  `order`, when not
// null, is a permutation of the n particles (the last deposit's cell order):
// thread j pushes particle order[j], so neighbouring threads read
// neighbouring field cells.  Each particle is pushed independently, so the
// result does not depend on the order.
template <typename TP, typename TF>
hipError_t launch_push(const PushGeom &g, const Fields6<TF> &f, const SortArrays<TP> &a,
                       bool first, bool bin, const BinState &bs, const int *order, hipStream_t s,
                       KernelTimer *timer = nullptr);
// The FP32-bulk push from float64 fields read as float32: the six fields are
// first rounded into f32 (6 fields32_stride(N) floats), then pushed as
// launch_push<float, float> does, both inside the push timing.
inline int64_t fields32_stride(int N) { return ((int64_t)N * N + 3) & ~(int64_t)3; }
hipError_t launch_push_f64_fields_as_f32(const PushGeom &g, const Fields6<double> &f,
                                         float *f32, const SortArrays<float> &a, bool first,
                                         bool bin, const BinState &bs, const int *order,
                                         hipStream_t s, KernelTimer *timer = nullptr);

// plasma3d/weights.py:35-67, 249-285 for one sort.  When `binned` is false
// the cells are computed here; otherwise the preceding push recorded them.
// bg_ro != nullptr also writes the background-ion column (rhoj.py:17-18)
// into ro1/jx1/jy1/jz1.
template <typename TP, typename TG>
hipError_t launch_deposit(const PushGeom &g, const SortArrays<TP> &a, bool use_full, bool binned,
                          const BinState &bs, TG *ro, TG *jx, TG *jy, TG *jz, const TG *bg_ro,
                          TG *ro1, TG *jx1, TG *jy1, TG *jz1, hipStream_t s,
                          KernelTimer *timer = nullptr);

// ---- field solve (plasma3d/fields.py:112-322) -------------------------------
struct FftPlan {
    int M;            // complex length N - 1
    int ns;           // stages
    int radix[12];
};

template <typename TG, typename TPRED> struct SolveIn {
    const TG *roF[2], *jxF[2], *jyF[2], *jzF[2];   // currents of layer k+1, per column
    const TG *roP[2], *jxP[2], *jyP[2], *jzP[2];   // currents of layer k
    const double *rbF, *rbP;                       // beam charge density k+1, k
    const TPRED *ExP, *EyP, *BxP, *ByP;            // Helmholtz predictor
};

template <typename TFULL, typename THALF> struct SolveOut {
    const TFULL *prev[7];   // Ex Ey Bx By Ez Bz Phi of layer k
    TFULL *full[7];
    THALF *half[7];
};

struct SolveGeom {
    int N;
    double h, xi_step, c;   // c: subtraction coefficient
    FftPlan plan;
    const double *lam;      // N eigenvalues 4/h^2 sin^2(k pi / 2(N-1))
    const void *tw;         // 2(N-1) roots exp(-i pi t / (N-1)), complex TG
    void *work;             // 7 N*N TG spectral scratch
    double *rowsum;         // N
};

template <typename TG, typename TPRED, typename TFULL, typename THALF>
hipError_t launch_compute_fields(const SolveGeom &g, const SolveIn<TG, TPRED> &in,
                                 const SolveOut<TFULL, THALF> &out, bool write_full,
                                 hipStream_t s);

// The same solve with each DST-I / DCT-I written as a product with its N x N
// transform matrix (transform.hip).
struct MatrixGeom {
    int N;
    double h, xi_step, c;   // c: subtraction coefficient
    const double *lam;      // N eigenvalues 4/h^2 sin^2(k pi / 2(N-1))
    const void *tdst;       // S[k][n] = 2 sin(pi k n / (N-1)), zero edge rows/columns (TG)
    const void *tdct;       // C[k][n] = w_n cos(pi k n / (N-1)), w = 1 at n = 0, N-1, else 2
    const void *tdct_t;     // C transposed
    void *work1, *work2;    // 7 N*N TG each
    double *rowsum;         // N
};

template <typename TG, typename TPRED, typename TFULL, typename THALF>
hipError_t launch_compute_fields_matrix(const MatrixGeom &g, const SolveIn<TG, TPRED> &in,
                                        const SolveOut<TFULL, THALF> &out, bool write_full,
                                        hipStream_t s, KernelTimer *timer = nullptr);

// The same solve with each DST-I / DCT-I computed by a complex FFT of length
// L = 2(N-1) in LDS, two real transforms per FFT (rfft.hip).
constexpr int RFFT_MAX_STAGES = 12;
struct RfftGeom {
    int N;
    double h, xi_step, c;   // c: subtraction coefficient
    uint64_t stages;        // radix of stage s in bits 5s..5s+4 (2 3 4 5 7 8 11 13 16),
                            // 0 after the last stage; the radices multiply to L
    int cap;                // LDS of a block in complex values, two buffers of cap / 2
                            // (an instantiated size)
    int tpb;                // transforms per block, tpb * L <= cap / 2
    int wave;               // kernels by waves, one transform per wave in registers
                            // (L = 512, 1024, 2048; four per block): bit 0 rows
                            // forward, 1 columns, 2 rows inverse; the others use the
                            // stages, cap and tpb above
    const double *lam;      // N eigenvalues 4/h^2 sin^2(k pi / 2(N-1))
    const void *tw;         // L roots exp(-2 pi i t / L), complex TG
    void *work;             // 7 N*N TG
    double *rowsum;         // N
};

// LDS capacities (complex values per block) the FFT kernels are built for.
bool rfft_cap_supported(bool fp32, int cap);
// Lengths of the FFT by waves.
bool rfft_wave_supported(int L);

template <typename TG, typename TPRED, typename TFULL, typename THALF>
hipError_t launch_compute_fields_rfft(const RfftGeom &g, const SolveIn<TG, TPRED> &in,
                                      const SolveOut<TFULL, THALF> &out, bool write_full,
                                      hipStream_t s, KernelTimer *timer = nullptr);

// ---- declustering noise filter (plasma3d/noise_filter.py) -------------------------
struct DeclustGeom {
    int side;               // lattice side n of the sort (n * n particles)
    int ntaps, tap0;        // circular Gaussian taps for offsets tap0 .. tap0 + ntaps - 1
    const double *taps;
    double xi_step, force, damping, limit;
    double *scratch;        // 4 n * n float64
};

// noise_filter(particles, particles_prev) of one sort: px, py kicked in place
// (bulk type TP), dx_chaotic, dy_chaotic written from x_offt, y_offt and the
// previous chaotic displacements.
template <typename TP>
hipError_t launch_noise_filter(const DeclustGeom &g, const double *x_offt, const double *y_offt,
                               TP *px, TP *py, const double *dx_prev, const double *dy_prev,
                               double *dx_chaotic, double *dy_chaotic, hipStream_t s);

// ---- diagnostics ---------------------------------------------------------------
// diagnostics/FXi/Strategy3D.py:6-25: total energy flux of a plasma layer into
// psi[0]; part holds 2 max_blocks float64 partial sums.
template <typename TP>
hipError_t launch_energy_flux(int N, double h, const Fields6<double> &f, int64_t n, const TP *m,
                              const TP *px, const TP *py, const double *pz, double *part,
                              int max_blocks, double *psi, hipStream_t s);

// diagnostics/FXi/Strategy3D.py:28-95: each of the ps.n (N, N) sources at
// the n_points probe cells (ix[p], iy[p]) into out[s * n_points + p], float64.
constexpr int PROBE_MAX_SOURCES = 16;
struct ProbeSources {
    int n;
    const void *src[PROBE_MAX_SOURCES];
    int is_f32[PROBE_MAX_SOURCES];
};
// The message of the last failed call that has no context (the
// allocation-free entry points), for lc3d_last_error(NULL); per host thread.
const char *noalloc_last_error();

hipError_t launch_probe_sample(const ProbeSources &ps, int N, int n_points, const int *ix,
                               const int *iy, double *out, hipStream_t s);

// max |a| over rows [r0, r1) and columns [c0, c1) of an nc-column row-major
// float64 array into out[0]; part holds max_blocks partial maxima.
hipError_t launch_absmax(const double *a, int64_t nc, int64_t r0, int64_t r1, int64_t c0,
                         int64_t c1, double *part, int max_blocks, double *out, hipStream_t s);

// ---- beam (beam3d) -----------------------------------------------------------
struct BeamGeom {
    int N;
    double h, xi_step, lost_radius;
};

// acc: 6 N*N words of fixed-point sums, zero between calls (the call leaves
// them zero).
// With layout, ro0 becomes (ro0 + sum0) / h^2 and ro1 = sum1 (its contents
// not read), every cell; otherwise ro0 += sum0 and ro1 += sum1.
hipError_t launch_beam_deposit(const BeamGeom &g, int64_t n, int64_t plasma_layer_idx,
                               const double *x, const double *y, const double *xi,
                               const double *q_norm, unsigned long long *acc, double *ro0,
                               double *ro1, unsigned *err, hipStream_t s, bool layout = false);

hipError_t launch_beam_substep(int64_t n, const double *q_m, const double *pz,
                               double substepping_energy, double time_step, double *dt,
                               int64_t *remaining_steps, hipStream_t s);

template <typename TF>
hipError_t launch_beam_move(const BeamGeom &g, int64_t n, int64_t beam_layer_idx,
                            const double *q_m, const double *dt, const Fields6<TF> &fk,
                            const Fields6<TF> &fk1, int64_t *remaining_steps, int64_t *id,
                            double *x, double *y, double *xi, double *px, double *py, double *pz,
                            uint8_t *lost, uint8_t *moved, uint8_t *fell, unsigned *err,
                            hipStream_t s);

// Order-preserving split of a moved layer's nf float64 and ni int64 rows (n
// particles each) by the three flag rows (lost, moved, fell): group g's
// particles to the front of out_f[g] (nf x n) and out_i[g] (ni x n), their
// number to counts[g].
hipError_t launch_beam_split(int64_t n, int nf, int ni, const double *f, const int64_t *iv,
                             const uint8_t *flags, double *out_f, int64_t *out_i, int64_t *counts,
                             hipStream_t s);

}  // namespace lc3d
