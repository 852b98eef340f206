/*
 * oracle/kernels.c -- CPU float64 restatement of lcodePy's 3D particle kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  Nothing in lcodepy_amd/ links or calls it.
 *
 * Arithmetic follows the reference's Numba kernels operation by operation so
 * that the end states agree with the reference's own golden fixtures:
 *   - built with -ffp-contract=off (Numba emits no fused multiply-adds);
 *   - `x ** k` for literal k is Numba's unrolled power-by-squaring:
 *     x**2 = x*x, x**3 = x*(x*x), x**4 = (x*x)*(x*x);
 *   - every sum keeps the reference's left-to-right association;
 *   - deposits add into each grid cell in particle order, as the reference's
 *     serial loops do (plasma3d/weights.py:43-67, beam3d/weights.py:26-41).
 *     The parallel form below gives each thread a band of grid rows and walks
 *     all particles in order, so every cell still sees the serial order and
 *     the result is bit-identical to the serial loop for any thread count.
 *
 * All arrays are C-contiguous float64 (int64 for ids / step counters, uint8
 * for the boolean masks); (N, N) grids are indexed [ix, iy] -> ix*N + iy.
 */
#include <math.h>
#include <stdint.h>
#include <stddef.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Quartic B-spline weight, taps -2..2.
 * plasma3d/weights.py:12-31 (identical copy in beam3d/weights.py:17-36). */
static inline double weight4(double x, int place)
{
    if (place == -2) {
        double t = 1.0 / 2.0 - x, t2 = t * t;
        return (t2 * t2) / 24.0;
    } else if (place == -1) {
        double x2 = x * x, x3 = x * x2, x4 = x2 * x2;
        return 19.0 / 96.0 - 11.0 / 24.0 * x + x2 / 4.0 + x3 / 6.0 - x4 / 6.0;
    } else if (place == 0) {
        double x2 = x * x, x4 = x2 * x2;
        return 115.0 / 192.0 - x2 * 5.0 / 8.0 + x4 / 4.0;
    } else if (place == 1) {
        double x2 = x * x, x3 = x * x2, x4 = x2 * x2;
        return 19.0 / 96.0 + 11.0 / 24.0 * x + x2 / 4.0 - x3 / 6.0 - x4 / 6.0;
    } else {
        double t = 1.0 / 2.0 + x, t2 = t * t;
        return (t2 * t2) / 24.0;
    }
}

double oracle_weight4(double x, int place) { return weight4(x, place); }

static int pick_threads(int nthreads)
{
#ifdef _OPENMP
    return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    (void)nthreads;
    return 1;
#endif
}

/* Free-streaming estimate, plasma3d/move.py:151-165. */
void oracle_move_wo_fields(int64_t n, double xi_step, const double *m,
                           double *x_offt, double *y_offt, const double *px,
                           const double *py, const double *pz, int nthreads)
{
    int nt = pick_threads(nthreads);
    (void)nt;
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int64_t k = 0; k < n; k++) {
        if (m[k] == 0) continue;
        double gamma_m = sqrt(m[k] * m[k] + pz[k] * pz[k] + px[k] * px[k] + py[k] * py[k]);
        x_offt[k] += px[k] / (gamma_m - pz[k]) * xi_step;
        y_offt[k] += py[k] / (gamma_m - pz[k]) * xi_step;
    }
}

/* Reflection at +-boundary, plasma3d/move.py:28-49.  Returns the reflected
 * position and updates x_offt/y_offt/px/py/pz in place. */
static inline void correct_boundary(double x_init, double y_init, double *x_offt,
                                    double *y_offt, double *px, double *py, double *pz,
                                    double boundary, double *x_out, double *y_out)
{
    double x = x_init + *x_offt;
    double y = y_init + *y_offt;
    if (x > +boundary) { x = +2 * boundary - x; *x_offt = x - x_init; *px = *py = *pz = 0; }
    if (x < -boundary) { x = -2 * boundary - x; *x_offt = x - x_init; *px = *py = *pz = 0; }
    if (y > +boundary) { y = +2 * boundary - y; *y_offt = y - y_init; *px = *py = *pz = 0; }
    if (y < -boundary) { y = -2 * boundary - y; *y_offt = y - y_init; *px = *py = *pz = 0; }
    *x_out = x;
    *y_out = y;
}

/* Predictor-corrector field push, plasma3d/move.py:52-148.
 * Fields are the six (N, N) arrays the particles move in; q is updated in
 * place (absorbed particles get q = 0, move.py:139-144). */
void oracle_move_smart(int64_t n, double xi_step, double boundary, double h, int N,
                       const double *Exf, const double *Eyf, const double *Ezf,
                       const double *Bxf, const double *Byf, const double *Bzf,
                       const double *x_init, const double *y_init,
                       const double *x_offt_prev, const double *y_offt_prev,
                       const double *px_prev, const double *py_prev, const double *pz_prev,
                       const double *ms, double *qs,
                       double *x_offt_full, double *y_offt_full,
                       double *px_full, double *py_full, double *pz_full, int nthreads)
{
    int nt = pick_threads(nthreads);
    (void)nt;
    const int half = N / 2;
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int64_t k = 0; k < n; k++) {
        if (qs[k] == 0) continue;
        double m = ms[k], q = qs[k];
        double opx = px_prev[k], opy = py_prev[k], opz = pz_prev[k];
        double px = opx, py = opy, pz = opz;
        double x_offt = x_offt_prev[k], y_offt = y_offt_prev[k];

        double x_offt_aver = (x_offt_prev[k] + x_offt_full[k]) / 2;
        double y_offt_aver = (y_offt_prev[k] + y_offt_full[k]) / 2;
        double x_half, y_half;
        correct_boundary(x_init[k], y_init[k], &x_offt_aver, &y_offt_aver,
                         &px, &py, &pz, boundary, &x_half, &y_half);

        double x_h = x_half / h + .5;
        double y_h = y_half / h + .5;
        double x_loc = x_h - floor(x_h) - .5;
        double y_loc = y_h - floor(y_h) - .5;
        int ix = (int)(floor(x_h) + half);
        int iy = (int)(floor(y_h) + half);

        double Ex = 0, Ey = 0, Ez = 0, Bx = 0, By = 0, Bz = 0;
        for (int kx = -2; kx <= 2; kx++) {
            double wx = weight4(x_loc, kx);
            for (int ky = -2; ky <= 2; ky++) {
                double w = wx * weight4(y_loc, ky);
                size_t idx = (size_t)(ix + kx) * N + (iy + ky);
                Ex += Exf[idx] * w; Bx += Bxf[idx] * w;
                Ey += Eyf[idx] * w; By += Byf[idx] * w;
                Ez += Ezf[idx] * w; Bz += Bzf[idx] * w;
            }
        }

        double gamma_m = sqrt(m * m + pz * pz + px * px + py * py);
        double vx = px / gamma_m, vy = py / gamma_m, vz = pz / gamma_m;
        double factor_1 = q * xi_step / (1 - pz / gamma_m);
        double dpx = factor_1 * (Ex + vy * Bz - vz * By);
        double dpy = factor_1 * (Ey - vx * Bz + vz * Bx);
        double dpz = factor_1 * (Ez + vx * By - vy * Bx);
        px = opx + dpx / 2; py = opy + dpy / 2; pz = opz + dpz / 2;

        gamma_m = sqrt(m * m + pz * pz + px * px + py * py);
        vx = px / gamma_m; vy = py / gamma_m; vz = pz / gamma_m;
        factor_1 = q * xi_step / (1 - pz / gamma_m);
        dpx = factor_1 * (Ex + vy * Bz - vz * By);
        dpy = factor_1 * (Ey - vx * Bz + vz * Bx);
        dpz = factor_1 * (Ez + vx * By - vy * Bx);
        px = opx + dpx / 2; py = opy + dpy / 2; pz = opz + dpz / 2;

        gamma_m = sqrt(m * m + pz * pz + px * px + py * py);
        x_offt += px / (gamma_m - pz) * xi_step;
        y_offt += py / (gamma_m - pz) * xi_step;

        px = opx + dpx; py = opy + dpy; pz = opz + dpz;

        double x_dummy, y_dummy;
        correct_boundary(x_init[k], y_init[k], &x_offt, &y_offt, &px, &py, &pz,
                         boundary, &x_dummy, &y_dummy);

        double x = x_init[k] + x_offt;
        double y = y_init[k] + y_offt;
        if (x >= +boundary || x <= -boundary || y >= +boundary || y <= -boundary) {
            x_offt = y_offt = px = py = pz = 0;
            qs[k] = 0;
        }
        x_offt_full[k] = x_offt; y_offt_full[k] = y_offt;
        px_full[k] = px; py_full[k] = py; pz_full[k] = pz;
    }
}

/* Plasma deposit of one sort into (N, N) ro/jx/jy/jz, plasma3d/weights.py:35-67
 * with x_h, y_h formed as in weights.py:271-274.  Accumulates into the given
 * arrays (the caller zeroes them, weights.py:255-258). */
void oracle_deposit_plasma(int64_t n, int N, double h,
                           const double *x_init, const double *y_init,
                           const double *x_offt, const double *y_offt,
                           const double *px, const double *py, const double *pz,
                           const double *q, const double *m,
                           double *ro, double *jx, double *jy, double *jz, int nthreads)
{
    int nt = pick_threads(nthreads);
    const int half = N / 2;
#pragma omp parallel num_threads(nt)
    {
        int t = 0, T = 1;
#ifdef _OPENMP
        t = omp_get_thread_num();
        T = omp_get_num_threads();
#endif
        int r0 = (int)((int64_t)N * t / T), r1 = (int)((int64_t)N * (t + 1) / T);
        for (int64_t k = 0; k < n; k++) {
            if (q[k] == 0) continue;
            double x_h = (x_init[k] + x_offt[k]) / h + 0.5;
            int ix = (int)(floor(x_h) + half);
            if (ix + 2 < r0 || ix - 2 >= r1) continue;
            double y_h = (y_init[k] + y_offt[k]) / h + 0.5;

            double gamma_m = sqrt(m[k] * m[k] + px[k] * px[k] + py[k] * py[k] + pz[k] * pz[k]);
            double dro = q[k] / (1 - pz[k] / gamma_m);
            double djx = px[k] * (dro / gamma_m);
            double djy = py[k] * (dro / gamma_m);
            double djz = pz[k] * (dro / gamma_m);

            double x_loc = x_h - floor(x_h) - .5;
            double y_loc = y_h - floor(y_h) - .5;
            int iy = (int)(floor(y_h) + half);
            for (int kx = -2; kx <= 2; kx++) {
                int gx = ix + kx;
                if (gx < r0 || gx >= r1) continue;
                double wx = weight4(x_loc, kx) / h;
                for (int ky = -2; ky <= 2; ky++) {
                    double w = wx * weight4(y_loc, ky) / h;
                    size_t idx = (size_t)gx * N + (iy + ky);
                    ro[idx] += dro * w;
                    jx[idx] += djx * w;
                    jy[idx] += djy * w;
                    jz[idx] += djz * w;
                }
            }
        }
    }
}

/* Beam deposit into the running layer (ro0, weight xi_loc) and the next one
 * (ro1, weight 1 - xi_loc), beam3d/weights.py:40-64, with x_h and xi_loc
 * formed as in beam3d/weights.py:152-160. */
void oracle_deposit_beam(int64_t n, int N, double h, double xi_step, int64_t plasma_layer_idx,
                         const double *x, const double *y, const double *xi,
                         const double *q_norm, double *ro0, double *ro1, int nthreads)
{
    int nt = pick_threads(nthreads);
    const int half = N / 2;
    const double xi_plasma_layer = -xi_step * (double)plasma_layer_idx;
#pragma omp parallel num_threads(nt)
    {
        int t = 0, T = 1;
#ifdef _OPENMP
        t = omp_get_thread_num();
        T = omp_get_num_threads();
#endif
        int r0 = (int)((int64_t)N * t / T), r1 = (int)((int64_t)N * (t + 1) / T);
        for (int64_t k = 0; k < n; k++) {
            double x_h = x[k] / h + 0.5;
            int ix = (int)(floor(x_h) + half);
            if (ix + 2 < r0 || ix - 2 >= r1) continue;
            double y_h = y[k] / h + 0.5;
            double xi_loc = (xi_plasma_layer - xi[k]) / xi_step;
            double x_loc = x_h - floor(x_h) - .5;
            double y_loc = y_h - floor(y_h) - .5;
            int iy = (int)(floor(y_h) + half);
            for (int kx = -2; kx <= 2; kx++) {
                int gx = ix + kx;
                if (gx < r0 || gx >= r1) continue;
                double wx = weight4(x_loc, kx);
                for (int ky = -2; ky <= 2; ky++) {
                    double w = wx * weight4(y_loc, ky);
                    double w0 = w * xi_loc;
                    double w1 = w * (1 - xi_loc);
                    size_t idx = (size_t)gx * N + (iy + ky);
                    ro0[idx] += q_norm[k] * w0;
                    ro1[idx] += q_norm[k] * w1;
                }
            }
        }
    }
}

/* Substep multiplier, beam3d/beam_calculator.py:27-34 (max_dt formed as the
 * numpy expression at :30). */
void oracle_beam_substep(int64_t n, const double *q_m, const double *pz,
                         double substepping_energy, double *dt_out)
{
    for (int64_t i = 0; i < n; i++) {
        double max_dt = sqrt(sqrt(1 / (q_m[i] * q_m[i]) + pz[i] * pz[i]) / substepping_energy);
        double dt = 1.0;
        while (dt > max_dt) dt /= 2.0;
        dt_out[i] = dt;
    }
}

static inline int sign_of(double x) { return x < 0 ? -1 : (x > 0 ? 1 : 0); }

/* Explicit-midpoint beam push over one layer, beam3d/move.py:31-127 (the CPU
 * kernel the golden fixtures come from).  fields_k are the fields of the
 * previous plasma layer, fields_k1 those just computed.  The trailing
 * "fell" re-check compares x, not xi, exactly as move.py:123 does. */
void oracle_move_beam(int64_t n, double xi_step, double lost_radius, int64_t beam_layer_idx,
                      double h, int N, const double *q_m, const double *dt,
                      const double *Ex_k, const double *Ey_k, const double *Ez_k,
                      const double *Bx_k, const double *By_k, const double *Bz_k,
                      const double *Ex_k1, const double *Ey_k1, const double *Ez_k1,
                      const double *Bx_k1, const double *By_k1, const double *Bz_k1,
                      int64_t *remaining_steps, int64_t *id, double *x, double *y,
                      double *xi, double *px, double *py, double *pz,
                      uint8_t *lost, uint8_t *moved, uint8_t *fell, int nthreads)
{
    int nt = pick_threads(nthreads);
    (void)nt;
    const int half = N / 2;
    const double xi_k = (double)beam_layer_idx * -xi_step;
    const double xi_k_1 = (double)(beam_layer_idx + 1) * -xi_step;
    const double r2max = lost_radius * lost_radius;
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int64_t k = 0; k < n; k++) {
        while (remaining_steps[k] > 0) {
            double iqm = 1 / q_m[k];
            double gamma_m = sqrt(iqm * iqm + px[k] * px[k] + py[k] * py[k] + pz[k] * pz[k]);
            double x_half = x[k] + dt[k] / 2 * (px[k] / gamma_m);
            double y_half = y[k] + dt[k] / 2 * (py[k] / gamma_m);
            double xi_half = xi[k] + dt[k] / 2 * (pz[k] / gamma_m - 1);

            if (xi_k_1 > xi_half) { fell[k] = 1; break; }
            if (x_half * x_half + y_half * y_half >= r2max) {
                x[k] = x_half; y[k] = y_half; xi[k] = xi_half;
                id[k] *= -1;
                lost[k] = 1;
                remaining_steps[k] = 0;
                break;
            }

            double x_h = x_half / h + .5;
            double y_h = y_half / h + .5;
            double x_loc = x_h - floor(x_h) - .5;
            double y_loc = y_h - floor(y_h) - .5;
            double xi_loc = (xi_half - xi_k) / xi_step;
            int ix = (int)(floor(x_h) + half);
            int iy = (int)(floor(y_h) + half);

            double Ex = 0, Ey = 0, Ez = 0, Bx = 0, By = 0, Bz = 0;
            for (int kx = -2; kx <= 2; kx++) {
                double wx = weight4(x_loc, kx);
                for (int ky = -2; ky <= 2; ky++) {
                    double w = wx * weight4(y_loc, ky);
                    double w0 = w * xi_loc;
                    double w1 = w * (1 - xi_loc);
                    size_t idx = (size_t)(ix + kx) * N + (iy + ky);
                    Ex += Ex_k[idx] * w0 + Ex_k1[idx] * w1;
                    Ey += Ey_k[idx] * w0 + Ey_k1[idx] * w1;
                    Ez += Ez_k[idx] * w0 + Ez_k1[idx] * w1;
                    Bx += Bx_k[idx] * w0 + Bx_k1[idx] * w1;
                    By += By_k[idx] * w0 + By_k1[idx] * w1;
                    Bz += Bz_k[idx] * w0 + Bz_k1[idx] * w1;
                }
            }

            double vx = px[k] / gamma_m, vy = py[k] / gamma_m, vz = pz[k] / gamma_m;
            int s = sign_of(q_m[k]);
            double px_half = px[k] + s * dt[k] / 2 * (Ex + vy * Bz - vz * By);
            double py_half = py[k] + s * dt[k] / 2 * (Ey + vz * Bx - vx * Bz);
            double pz_half = pz[k] + s * dt[k] / 2 * (Ez + vx * By - vy * Bx);

            gamma_m = sqrt(iqm * iqm + px_half * px_half + py_half * py_half + pz_half * pz_half);
            x[k] += dt[k] * (px_half / gamma_m);
            y[k] += dt[k] * (py_half / gamma_m);
            xi[k] += dt[k] * (pz_half / gamma_m - 1);

            px[k] = 2 * px_half - px[k];
            py[k] = 2 * py_half - py[k];
            pz[k] = 2 * pz_half - pz[k];

            if (x[k] * x[k] + y[k] * y[k] >= r2max) {
                id[k] *= -1;
                lost[k] = 1;
                remaining_steps[k] = 0;
                break;
            }
            remaining_steps[k] -= 1;
        }
        if (remaining_steps[k] == 0 && xi_k_1 > x[k]) fell[k] = 1;
        if (!fell[k] && !lost[k]) moved[k] = 1;
    }
}
