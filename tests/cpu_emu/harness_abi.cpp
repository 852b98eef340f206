// Drives every entry point of the C ABI (include/lcodepy_amd.h) on the CPU
// emulation, in both precision forms and with every field solver, with
// exactly sized host buffers: field solve, deposit, push, fused plasma step
// (1 and 2 sorts) and the beam kernels (including off-grid and non-finite
// beam particles); then checks that the FFT solves agree with the matrix one.
#include "hip/hip_runtime.h"
#include "../../include/lcodepy_amd.h"
#include <random>
#include <vector>
#include <cstring>
#include <cassert>
#include <cmath>
#include <algorithm>

static std::mt19937_64 rng(7);
static double nrm() { static std::normal_distribution<double> d; return d(rng); }
static double uni(double a, double b) { std::uniform_real_distribution<double> d(a, b); return d(rng); }

template <typename T> T *arr(size_t n, double scale = 0.0) {
    T *p; hipMalloc(&p, n * sizeof(T));
    for (size_t i = 0; i < n; ++i) p[i] = (T)(scale * nrm());
    return p;
}

struct Run {
    int N; double h = 0.1, dxi = 0.05; int prec; lc3d_ctx *ctx = nullptr; int64_t n = 0, side = 0;
    std::vector<double> grid;
    Run(int N_, int prec_, int solver) : N(N_), prec(prec_) {
        int steps = N - 30; side = steps / 2 * 2 * 2;   // fineness 2
        for (int i = 0; i < steps / 2 * 2; ++i) grid.push_back((0.5 + i) * h / 2);
        std::vector<double> g2; for (int i = (int)grid.size() - 1; i >= 0; --i) g2.push_back(-grid[i]);
        g2.insert(g2.end(), grid.begin(), grid.end()); grid = g2; side = grid.size(); n = side * side;
        lc3d_params p{}; p.grid_steps = N; p.grid_step_size = h; p.xi_step_size = dxi;
        p.boundary = h * (N / 2.0 - 10); p.subtraction_coeff = 1.0;
        double R = h * N / 2; p.lost_radius = std::max(0.9 * R, R - 1); p.precision = prec; p.max_particles = n;
        p.field_solver = solver;
        int rc = lc3d_create(&p, &ctx); if (rc) { printf("create failed %d\n", rc); exit(1); }
    }
};

template <typename TP> struct Sort {
    double *xi, *yi, *xo, *yo, *xo1, *yo1; TP *m, *px, *py, *q, *px1, *py1, *q1; double *pz, *pz1;
    lc3d_sort s{};
    Sort(Run &r, double spread, double pmax) {
        int64_t n = r.n;
        xi = arr<double>(n); yi = arr<double>(n);
        for (int64_t a = 0; a < r.side; ++a) for (int64_t b = 0; b < r.side; ++b) { xi[a * r.side + b] = r.grid[a]; yi[a * r.side + b] = r.grid[b]; }
        xo = arr<double>(n); yo = arr<double>(n); xo1 = arr<double>(n); yo1 = arr<double>(n);
        for (int64_t k = 0; k < n; ++k) { xo[k] = uni(-spread, spread) * r.h; yo[k] = uni(-spread, spread) * r.h; }
        m = arr<TP>(n); q = arr<TP>(n); px = arr<TP>(n, pmax); py = arr<TP>(n, pmax); pz = arr<double>(n, pmax);
        px1 = arr<TP>(n); py1 = arr<TP>(n); pz1 = arr<double>(n); q1 = arr<TP>(n);
        for (int64_t k = 0; k < n; ++k) { m[k] = (TP)(r.h * r.h / 4); q[k] = (TP)(uni(0, 1) < 0.01 ? 0.0 : -r.h * r.h / 4); }
        s.n = n; s.x_init = xi; s.y_init = yi; s.m = m; s.x_offt_prev = xo; s.y_offt_prev = yo;
        s.px_prev = px; s.py_prev = py; s.pz_prev = pz; s.q_prev = q;
        s.x_offt = xo1; s.y_offt = yo1; s.px = px1; s.py = py1; s.pz = pz1; s.q = q1; s.column = 0;
    }
};

template <typename TF> lc3d_fields fields(size_t nn, double scale) {
    lc3d_fields f; f.Ex = arr<TF>(nn, scale); f.Ey = arr<TF>(nn, scale); f.Ez = arr<TF>(nn, scale);
    f.Bx = arr<TF>(nn, scale); f.By = arr<TF>(nn, scale); f.Bz = arr<TF>(nn, scale); f.Phi = arr<TF>(nn, scale); return f;
}
template <typename TG> lc3d_currents currents(int N, double scale) {
    size_t nn = (size_t)N * N; lc3d_currents c;
    void **p[4] = {&c.ro, &c.jx, &c.jy, &c.jz};
    for (auto q : p) { TG *a = arr<TG>(2 * nn); for (int s = 0; s < 2; ++s) for (int i = 3; i < N - 3; ++i) for (int j = 3; j < N - 3; ++j) a[s * nn + i * N + j] = (TG)(scale * nrm()); *q = a; }
    return c;
}

#define CHECK(x) do { int rc_ = (x); if (rc_) { printf("FAIL %s -> %d\n", #x, rc_); exit(1); } } while (0)

template <typename TP, typename TG, typename TH> void run_all(int N, int prec, int solver) {
    printf("== N=%d prec=%d solver=%d\n", N, prec, solver); fflush(stdout);
    Run r(N, prec, solver); size_t nn = (size_t)N * N;
    lc3d_fields pred = fields<double>(nn, 0.01), prev = fields<double>(nn, 0.01), full = fields<double>(nn, 0), half = fields<TH>(nn, 0);
    lc3d_currents cp = currents<TG>(N, 0.1), cf = currents<TG>(N, 0.1);
    double *rbf = arr<double>(nn, 0.05), *rbp = arr<double>(nn, 0.05);
    CHECK(lc3d_compute_fields(r.ctx, &pred, 0, &prev, rbf, rbp, &cp, &cf, &full, &half, nullptr));
    printf("compute_fields ok\n"); fflush(stdout);
    Sort<TP> s(r, 0.3, 0.05);
    TG *ro = arr<TG>(nn), *jx = arr<TG>(nn), *jy = arr<TG>(nn), *jz = arr<TG>(nn);
    std::memcpy(s.xo1, s.xo, r.n * 8); std::memcpy(s.yo1, s.yo, r.n * 8);
    std::memcpy(s.px1, s.px, r.n * sizeof(TP)); std::memcpy(s.py1, s.py, r.n * sizeof(TP)); std::memcpy(s.pz1, s.pz, r.n * sizeof(double)); std::memcpy(s.q1, s.q, r.n * sizeof(TP));
    CHECK(lc3d_deposit(r.ctx, &s.s, ro, jx, jy, jz, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr));
    printf("deposit ok\n"); fflush(stdout);
    CHECK(lc3d_move_wo_fields(r.ctx, r.n, s.m, s.xo1, s.yo1, s.px1, s.py1, s.pz1, nullptr));
    lc3d_fields6 f6{pred.Ex, pred.Ey, pred.Ez, pred.Bx, pred.By, pred.Bz};
    CHECK(lc3d_move_smart(r.ctx, &s.s, &f6, 0, nullptr));
    printf("push ok\n"); fflush(stdout);
    lc3d_currents cf2 = currents<TG>(N, 0.0);
    TG *bg = arr<TG>(nn, 0.01);
    CHECK(lc3d_kernel_timing(r.ctx, 1));
    for (int step = 0; step < 2; ++step)
        CHECK(lc3d_plasma_step(r.ctx, 1, &s.s, &prev, &cp, rbf, rbp, bg, &full, &half, &cf2, nullptr));
    double sec[3]; int64_t nl[3];
    CHECK(lc3d_kernel_times(r.ctx, sec, nl));
    CHECK(lc3d_kernel_timing(r.ctx, 0));
    const int64_t want_gemm = solver == LC3D_SOLVER_MATRIX ? 16 : solver == LC3D_SOLVER_RFFT ? 12 : 0;
    if (nl[0] != 6 || nl[1] != 6 || nl[2] != want_gemm) { printf("FAIL kernel launch counts %lld %lld %lld\n", (long long)nl[0], (long long)nl[1], (long long)nl[2]); exit(1); }
    printf("plasma_step 1 sort ok (timed launches %lld %lld %lld)\n", (long long)nl[0], (long long)nl[1], (long long)nl[2]); fflush(stdout);
    Sort<TP> ions(r, 0.3, 0.01); ions.s.column = 1;
    lc3d_sort two[2] = {s.s, ions.s};
    CHECK(lc3d_plasma_step(r.ctx, 2, two, &prev, &cp, rbf, rbp, nullptr, &full, &half, &cf2, nullptr));
    printf("plasma_step 2 sorts ok\n"); fflush(stdout);
    // beam layer
    int64_t nb = 500; double *bx = arr<double>(nb, 0.5), *by = arr<double>(nb, 0.5), *bxi = arr<double>(nb), *bq = arr<double>(nb);
    double *bpx = arr<double>(nb, 0.01), *bpy = arr<double>(nb, 0.01), *bpz = arr<double>(nb), *bqm = arr<double>(nb), *bdt = arr<double>(nb);
    int64_t *bid = arr<int64_t>(nb), *bst = arr<int64_t>(nb); uint8_t *l = arr<uint8_t>(nb), *mv = arr<uint8_t>(nb), *fl = arr<uint8_t>(nb);
    for (int64_t k = 0; k < nb; ++k) { bxi[k] = -0.2 + uni(-0.05, 0); bq[k] = 0.01; bpz[k] = k % 7 == 0 ? 1.0 : 1000.0; bqm[k] = -1; bid[k] = k + 1; }
    bx[0] = 1e9; by[1] = NAN;
    double *r0 = arr<double>(nn), *r1 = arr<double>(nn);
    CHECK(lc3d_deposit_beam(r.ctx, nb, 4, bx, by, bxi, bq, r0, r1, nullptr));
    double *r2 = arr<double>(nn);
    CHECK(lc3d_layout_beam_layer(r.ctx, nb, 4, bx, by, bxi, bq, r1, r2, nullptr));
    CHECK(lc3d_layout_beam_layer(r.ctx, 0, 5, nullptr, nullptr, nullptr, nullptr, r2, r1, nullptr));
    CHECK(lc3d_beam_substep(r.ctx, nb, bqm, bpz, 2.0, 1.0, bdt, bst, nullptr));
    lc3d_fields6 fk{full.Ex, full.Ey, full.Ez, full.Bx, full.By, full.Bz};
    CHECK(lc3d_move_beam(r.ctx, nb, 3, bqm, bdt, &fk, &fk, bst, bid, bx, by, bxi, bpx, bpy, bpz, l, mv, fl, nullptr));
    uint32_t flags = 0; CHECK(lc3d_error_flags(r.ctx, &flags, nullptr));
    printf("beam ok, error flags %u\n", flags); fflush(stdout);
    // energy flux: the context entry and the allocation-free one (static scratch)
    double *psi = arr<double>(2);
    CHECK(lc3d_energy_flux(r.ctx, &fk, r.n, s.m, s.px, s.py, s.pz, psi, nullptr));
    CHECK(lc3d_energy_flux_noalloc(N, r.h, prec, &fk, r.n, s.m, s.px, s.py, s.pz, psi + 1, nullptr));
    CHECK(lc3d_energy_flux_noalloc(N, r.h, prec, &fk, r.n / 3, s.m, s.px, s.py, s.pz, psi + 1, nullptr));
    CHECK(lc3d_energy_flux(r.ctx, &fk, r.n / 3, s.m, s.px, s.py, s.pz, psi, nullptr));
    if (!(psi[0] == psi[1]) || !(psi[0] > 0)) { printf("FAIL energy flux noalloc %.17g vs %.17g\n", psi[1], psi[0]); exit(1); }
    if (lc3d_energy_flux_noalloc(1026, r.h, prec, &fk, 0, nullptr, nullptr, nullptr, nullptr, psi, nullptr) != LC3D_EINVAL) { printf("FAIL noalloc size check\n"); exit(1); }
    printf("energy flux ok (allocation-free form equal)\n"); fflush(stdout);
    lc3d_destroy(r.ctx);
}

// The same solve through both implementations; returns max |a - b| / max |b|
// over the seven full and half fields.
template <typename TG, typename TH> double cross_check(int N, int prec, int solver) {
    size_t nn = (size_t)N * N;
    Run a(N, prec, LC3D_SOLVER_MATRIX), b(N, prec, solver);
    lc3d_fields pred = fields<double>(nn, 0.01), prev = fields<double>(nn, 0.01);
    lc3d_currents cp = currents<TG>(N, 0.1), cf = currents<TG>(N, 0.1);
    double *rbf = arr<double>(nn, 0.05), *rbp = arr<double>(nn, 0.05);
    lc3d_fields fa = fields<double>(nn, 0), ha = fields<TH>(nn, 0), fb = fields<double>(nn, 0), hb = fields<TH>(nn, 0);
    CHECK(lc3d_compute_fields(a.ctx, &pred, 0, &prev, rbf, rbp, &cp, &cf, &fa, &ha, nullptr));
    CHECK(lc3d_compute_fields(b.ctx, &pred, 0, &prev, rbf, rbp, &cp, &cf, &fb, &hb, nullptr));
    double worst = 0;
    auto cmp = [&](const void *x, const void *y, bool half) {
        double d = 0, m = 0;
        for (size_t k = 0; k < nn; ++k) {
            double u = half ? (double)((const TH *)x)[k] : ((const double *)x)[k];
            double v = half ? (double)((const TH *)y)[k] : ((const double *)y)[k];
            d = std::max(d, std::fabs(u - v)); m = std::max(m, std::fabs(v));
        }
        worst = std::max(worst, m > 0 ? d / m : d);
    };
    const void *A[14] = {fa.Ex, fa.Ey, fa.Ez, fa.Bx, fa.By, fa.Bz, fa.Phi, ha.Ex, ha.Ey, ha.Ez, ha.Bx, ha.By, ha.Bz, ha.Phi};
    const void *B[14] = {fb.Ex, fb.Ey, fb.Ez, fb.Bx, fb.By, fb.Bz, fb.Phi, hb.Ex, hb.Ey, hb.Ez, hb.Bx, hb.By, hb.Bz, hb.Phi};
    for (int k = 0; k < 14; ++k) cmp(A[k], B[k], k >= 7);
    lc3d_destroy(a.ctx); lc3d_destroy(b.ctx);
    return worst;
}

int main(int argc, char **argv) {
    for (int i = 1; i < argc; ++i) {
        int N = atoi(argv[i]);
        for (int solver : {LC3D_SOLVER_MATRIX, LC3D_SOLVER_FFT, LC3D_SOLVER_RFFT}) {
            run_all<double, double, double>(N, LC3D_FP64, solver);
            run_all<float, float, float>(N, LC3D_FP32_BULK, solver);
        }
        for (int solver : {LC3D_SOLVER_FFT, LC3D_SOLVER_RFFT}) {
            const double e64 = cross_check<double, double>(N, LC3D_FP64, solver);
            const double e32 = cross_check<float, float>(N, LC3D_FP32_BULK, solver);
            printf("matrix vs solver %d N=%d: fp64 %.3e, fp32-bulk %.3e\n", solver, N, e64, e32);
            fflush(stdout);
            if (!(e64 < 1e-12) || !(e32 < 1e-4)) { printf("FAIL solver mismatch\n"); return 1; }
        }
    }
    {   // the solve autotuner over every candidate at the first size
        const int N = argc > 1 ? atoi(argv[1]) : 33;
        lc3d_params p{}; p.grid_steps = N; p.grid_step_size = 0.1; p.xi_step_size = 0.05;
        p.boundary = 0.1 * (N / 2.0 - 10); p.subtraction_coeff = 1.0; p.lost_radius = 1.0;
        for (int prec : {LC3D_FP64, LC3D_FP32_BULK}) {
            p.precision = prec;
            lc3d_solve_choice best{}, table[128]; int32_t n = 0;
            CHECK(lc3d_autotune_solve(&p, 1, &best, table, 128, &n));
            int refused = 0;
            for (int i = 0; i < n; ++i) refused += table[i].seconds < 0;
            printf("autotune N=%d prec=%d: %d candidates, %d refused, best solver %d\n", N, prec, n, refused, best.field_solver);
            if (n < 2 || refused) { printf("FAIL autotune\n"); return 1; }
        }
    }
    printf("ALL OK\n");
}
