// C ABI of lcodepy_amd (declared in include/lcodepy_amd.h): context set-up
// and the entry points that dispatch the gfx950 kernels by precision.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lcodepy_amd.h"
#include "kernels.h"

using namespace lc3d;

// Blocks of the energy-flux reduction (each leaves two float64 partials).
constexpr int DIAG_BLOCKS = 1024;
static_assert(DIAG_BLOCKS == LC3D_NOALLOC_FLUX_BLOCKS,
              "the allocation-free energy flux sums over the same blocks");

struct lc3d_ctx {
    lc3d_params prm{};
    int N = 0;
    int device = 0;
    int64_t cap = 0;
    FftPlan plan{};
    // binning
    int *count = nullptr, *start = nullptr, *lstart = nullptr, *base = nullptr, *key = nullptr;
    int *idx = nullptr, *rank = nullptr, *pid = nullptr, *pos = nullptr;
    int nbase = 0;
    unsigned *err = nullptr;
    void *rec = nullptr;
    int *big = nullptr, *nbig = nullptr, *dslot = nullptr;
    void *dense = nullptr;
    int bigcap = 0;
    // idx holds one list per current column (cap ints each): after a deposit
    // of the sort with this x_init and size into column k, its part k is a
    // permutation of the sort in cell order, which pushes of that sort walk.
    const void *order_x_init[2] = {nullptr, nullptr};
    int64_t order_n[2] = {-1, -1};
    // field solve
    int solver = LC3D_SOLVER_MATRIX;
    void *work = nullptr, *work2 = nullptr;
    double *rowsum = nullptr;
    double *diag = nullptr;   // energy-flux partial sums, 2 DIAG_BLOCKS
    unsigned long long *beam_acc = nullptr;   // beam deposit fixed-point sums, 6 N*N, zero
    float *fields32 = nullptr;   // FP32-bulk: float32 copies of six float64 fields
    void *tw = nullptr;                                      // FFT twiddles
    uint64_t rf_stages = 0;                                  // RFFT radix plan (packed)
    int rf_cap = 0, rf_tpb = 0, rf_wave = 0;                 // RFFT launch shape
    void *tdst = nullptr, *tdct = nullptr, *tdct_t = nullptr;   // transform matrices
    double *lam = nullptr;
    // declustering noise filter
    int dc_side = 0, dc_ntaps = 0, dc_tap0 = 0;
    double *dc_taps = nullptr, *dc_scratch = nullptr;
    KernelTimer timer;
    std::string last_error;
};

namespace {

int fail(lc3d_ctx *c, int code, const std::string &msg)
{
    if (c) c->last_error = msg;
    return code;
}

int hip_fail(lc3d_ctx *c, hipError_t e, const char *where)
{
    return fail(c, LC3D_EHIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(ctx, expr)                                   \
    do {                                                     \
        hipError_t e_ = (expr);                              \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #expr); \
    } while (0)

bool default_plan(int M, FftPlan &p)
{
    p.M = M;
    p.ns = 0;
    int m = M, twos = 0;
    while (m % 2 == 0) {
        m /= 2;
        ++twos;
    }
    while (twos >= 3) {
        p.radix[p.ns++] = 8;
        twos -= 3;
    }
    if (twos == 2) p.radix[p.ns++] = 4;
    if (twos == 1) p.radix[p.ns++] = 2;
    const int odd[] = {3, 5, 7, 11, 13};
    for (int r : odd) {
        while (m % r == 0) {
            if (p.ns >= 12) return false;
            p.radix[p.ns++] = r;
            m /= r;
        }
    }
    return m == 1 && p.ns <= 12;
}

bool valid_plan(const FftPlan &p)
{
    int prod = 1;
    for (int s = 0; s < p.ns; ++s) {
        const int r = p.radix[s];
        if (r != 2 && r != 3 && r != 4 && r != 5 && r != 7 && r != 8 && r != 11 && r != 13)
            return false;
        prod *= r;
    }
    return prod == p.M;
}

// Default radix plan of the length-L FFT of the RFFT solve: radix 16 while
// it divides, then 8, 4 or 2 for the rest of the power of two, then the odd
// primes.  Returns the packed plan (radix of stage s in bits 5s..5s+4), or 0
// when L has another prime factor or needs more stages than fit.
uint64_t rfft_default_plan(int L)
{
    std::vector<int> r;
    int m = L;
    while (m % 16 == 0) { r.push_back(16); m /= 16; }
    if (m % 8 == 0) { r.push_back(8); m /= 8; }
    if (m % 4 == 0) { r.push_back(4); m /= 4; }
    if (m % 2 == 0) { r.push_back(2); m /= 2; }
    for (int q : {3, 5, 7, 11, 13})
        while (m % q == 0) { r.push_back(q); m /= q; }
    if (m != 1 || r.size() > (size_t)RFFT_MAX_STAGES) return 0;
    uint64_t packed = 0;
    for (size_t s = 0; s < r.size(); ++s) packed |= (uint64_t)r[s] << (5 * s);
    return packed;
}

// Packs a caller's plan for the length-L FFT; 0 when it is not one.
uint64_t rfft_pack_plan(int L, int ns, const int32_t *radix)
{
    if (ns < 1 || ns > RFFT_MAX_STAGES) return 0;
    int64_t prod = 1;
    uint64_t packed = 0;
    for (int s = 0; s < ns; ++s) {
        const int r = radix[s];
        if (r != 2 && r != 3 && r != 4 && r != 5 && r != 7 && r != 8 && r != 11 && r != 13 &&
            r != 16)
            return 0;
        prod *= r;
        if (prod > L) return 0;
        packed |= (uint64_t)r << (5 * s);
    }
    return prod == L ? packed : 0;
}

bool fp32(const lc3d_ctx *c) { return c->prm.precision == LC3D_FP32_BULK; }

PushGeom push_geom(const lc3d_ctx *c, int64_t n)
{
    return PushGeom{n, c->N, c->prm.grid_step_size, c->prm.xi_step_size, c->prm.boundary};
}

// Binning state of a deposit into (or a push ordered by) current column col.
BinState bin_state(const lc3d_ctx *c, int col)
{
    const int64_t cap = std::max<int64_t>(1, c->cap);
    return BinState{c->count, c->start,  c->lstart, c->base,   c->nbase,
                    c->key,   c->rank,   c->pid,    c->pos,    c->idx + col * cap,
                    c->err,   c->rec,    cap,       c->big,    c->nbig,
                    c->bigcap, c->dslot, c->dense};
}

template <typename TP> SortArrays<TP> sort_arrays(const lc3d_sort *s)
{
    SortArrays<TP> a;
    a.x_init = s->x_init;
    a.y_init = s->y_init;
    a.m = static_cast<const TP *>(s->m);
    a.xo_prev = s->x_offt_prev;
    a.yo_prev = s->y_offt_prev;
    a.px_prev = static_cast<const TP *>(s->px_prev);
    a.py_prev = static_cast<const TP *>(s->py_prev);
    a.pz_prev = static_cast<const double *>(s->pz_prev);
    a.q_prev = static_cast<const TP *>(s->q_prev);
    a.xo = s->x_offt;
    a.yo = s->y_offt;
    a.px = static_cast<TP *>(s->px);
    a.py = static_cast<TP *>(s->py);
    a.pz = static_cast<double *>(s->pz);
    a.q = static_cast<TP *>(s->q);
    return a;
}

template <typename T> Fields6<T> f6(const lc3d_fields6 *f)
{
    return Fields6<T>{static_cast<const T *>(f->Ex), static_cast<const T *>(f->Ey),
                      static_cast<const T *>(f->Ez), static_cast<const T *>(f->Bx),
                      static_cast<const T *>(f->By), static_cast<const T *>(f->Bz)};
}

template <typename T> Fields6<T> f6(const lc3d_fields *f)
{
    return Fields6<T>{static_cast<const T *>(f->Ex), static_cast<const T *>(f->Ey),
                      static_cast<const T *>(f->Ez), static_cast<const T *>(f->Bx),
                      static_cast<const T *>(f->By), static_cast<const T *>(f->Bz)};
}

#ifdef LC3D_FFT_SOLVER
SolveGeom solve_geom(const lc3d_ctx *c)
{
    return SolveGeom{c->N, c->prm.grid_step_size, c->prm.xi_step_size, c->prm.subtraction_coeff,
                     c->plan, c->lam, c->tw, c->work, c->rowsum};
}
#endif

RfftGeom rfft_geom(const lc3d_ctx *c)
{
    return RfftGeom{c->N, c->prm.grid_step_size, c->prm.xi_step_size, c->prm.subtraction_coeff,
                    c->rf_stages, c->rf_cap, c->rf_tpb, c->rf_wave, c->lam, c->tw, c->work,
                    c->rowsum};
}

MatrixGeom matrix_geom(const lc3d_ctx *c)
{
    return MatrixGeom{c->N, c->prm.grid_step_size, c->prm.xi_step_size, c->prm.subtraction_coeff,
                      c->lam, c->tdst, c->tdct, c->tdct_t, c->work, c->work2, c->rowsum};
}

KernelTimer *timer_of(const lc3d_ctx *c)
{
    return c->timer.on ? const_cast<KernelTimer *>(&c->timer) : nullptr;
}

// Dispatches the solve to the context's implementation.
template <typename TG, typename TPRED, typename TFULL, typename THALF>
hipError_t run_solve(const lc3d_ctx *c, const SolveIn<TG, TPRED> &in,
                     const SolveOut<TFULL, THALF> &out, bool write_full, hipStream_t st)
{
#ifdef LC3D_FFT_SOLVER
    if (c->solver == LC3D_SOLVER_FFT)
        return launch_compute_fields<TG, TPRED, TFULL, THALF>(solve_geom(c), in, out, write_full, st);
#endif
    if (c->solver == LC3D_SOLVER_RFFT)
        return launch_compute_fields_rfft<TG, TPRED, TFULL, THALF>(rfft_geom(c), in, out,
                                                                   write_full, st, timer_of(c));
    return launch_compute_fields_matrix<TG, TPRED, TFULL, THALF>(
        matrix_geom(c), in, out, write_full, st, timer_of(c));
}

template <typename TG, typename TPRED>
SolveIn<TG, TPRED> solve_in(const lc3d_ctx *c, const lc3d_fields *pred, const double *rbF,
                            const double *rbP, const lc3d_currents *cp, const lc3d_currents *cf)
{
    const size_t nn = (size_t)c->N * c->N;
    SolveIn<TG, TPRED> in;
    auto col = [&](void *p, int k) { return static_cast<const TG *>(p) + k * nn; };
    for (int k = 0; k < 2; ++k) {
        in.roF[k] = col(cf->ro, k);
        in.jxF[k] = col(cf->jx, k);
        in.jyF[k] = col(cf->jy, k);
        in.jzF[k] = col(cf->jz, k);
        in.roP[k] = col(cp->ro, k);
        in.jxP[k] = col(cp->jx, k);
        in.jyP[k] = col(cp->jy, k);
        in.jzP[k] = col(cp->jz, k);
    }
    in.rbF = rbF;
    in.rbP = rbP;
    in.ExP = static_cast<const TPRED *>(pred->Ex);
    in.EyP = static_cast<const TPRED *>(pred->Ey);
    in.BxP = static_cast<const TPRED *>(pred->Bx);
    in.ByP = static_cast<const TPRED *>(pred->By);
    return in;
}

// Field order of the solve: Ex Ey Bx By Ez Bz Phi.
template <typename TFULL, typename THALF>
SolveOut<TFULL, THALF> solve_out(const lc3d_fields *prev, lc3d_fields *full, lc3d_fields *half)
{
    SolveOut<TFULL, THALF> o;
    void *const p[7] = {prev->Ex, prev->Ey, prev->Bx, prev->By, prev->Ez, prev->Bz, prev->Phi};
    void *const f[7] = {full->Ex, full->Ey, full->Bx, full->By, full->Ez, full->Bz, full->Phi};
    void *const h[7] = {half->Ex, half->Ey, half->Bx, half->By, half->Ez, half->Bz, half->Phi};
    for (int i = 0; i < 7; ++i) {
        o.prev[i] = static_cast<const TFULL *>(p[i]);
        o.full[i] = static_cast<TFULL *>(f[i]);
        o.half[i] = static_cast<THALF *>(h[i]);
    }
    return o;
}

bool sort_ok(const lc3d_ctx *c, const lc3d_sort *s)
{
    return s && s->n >= 0 && s->n <= c->cap && (s->n == 0 || (s->x_init && s->y_init && s->m &&
           s->x_offt_prev && s->y_offt_prev && s->px_prev && s->py_prev && s->pz_prev &&
           s->q_prev && s->x_offt && s->y_offt && s->px && s->py && s->pz && s->q)) &&
           (s->column == 0 || s->column == 1);
}

bool declust_ok(const lc3d_ctx *c, const lc3d_sort *s)
{
    return c->dc_side > 0 && s->n == (int64_t)c->dc_side * c->dc_side && s->dx_chaotic_prev &&
           s->dy_chaotic_prev && s->dx_chaotic && s->dy_chaotic;
}

int noise_filter_sort(lc3d_ctx *c, const lc3d_sort *s, hipStream_t st)
{
    const DeclustGeom g{c->dc_side, c->dc_ntaps, c->dc_tap0, c->dc_taps,
                        c->prm.xi_step_size, c->prm.declustering_force,
                        c->prm.declustering_damping, c->prm.declustering_limit, c->dc_scratch};
    hipError_t e;
    if (fp32(c))
        e = launch_noise_filter<float>(g, s->x_offt, s->y_offt, static_cast<float *>(s->px),
                                       static_cast<float *>(s->py), s->dx_chaotic_prev,
                                       s->dy_chaotic_prev, s->dx_chaotic, s->dy_chaotic, st);
    else
        e = launch_noise_filter<double>(g, s->x_offt, s->y_offt, static_cast<double *>(s->px),
                                        static_cast<double *>(s->py), s->dx_chaotic_prev,
                                        s->dy_chaotic_prev, s->dx_chaotic, s->dy_chaotic, st);
    if (e != hipSuccess) return hip_fail(c, e, "noise_filter");
    return LC3D_OK;
}

bool fields_ok(const lc3d_fields *f, bool phi)
{
    return f && f->Ex && f->Ey && f->Ez && f->Bx && f->By && f->Bz && (!phi || f->Phi);
}

bool currents_ok(const lc3d_currents *c) { return c && c->ro && c->jx && c->jy && c->jz; }

// Deposit of one sort into its column of `cur`.
int deposit_sort(lc3d_ctx *c, const lc3d_sort *s, bool binned, const lc3d_currents *cur,
                 const void *bg_ro, bool single_sort, hipStream_t st)
{
    const size_t nn = (size_t)c->N * c->N;
    const int col = s->column;
    const bool bg = bg_ro != nullptr && single_sort;
    hipError_t e;
    if (fp32(c)) {
        auto at = [&](void *p, int k) { return static_cast<float *>(p) + k * nn; };
        e = launch_deposit<float, float>(push_geom(c, s->n), sort_arrays<float>(s), true, binned,
                                         bin_state(c, col), at(cur->ro, col), at(cur->jx, col),
                                         at(cur->jy, col), at(cur->jz, col),
                                         bg ? static_cast<const float *>(bg_ro) : nullptr,
                                         at(cur->ro, 1), at(cur->jx, 1), at(cur->jy, 1),
                                         at(cur->jz, 1), st, timer_of(c));
    } else {
        auto at = [&](void *p, int k) { return static_cast<double *>(p) + k * nn; };
        e = launch_deposit<double, double>(push_geom(c, s->n), sort_arrays<double>(s), true, binned,
                                           bin_state(c, col), at(cur->ro, col), at(cur->jx, col),
                                           at(cur->jy, col), at(cur->jz, col),
                                           bg ? static_cast<const double *>(bg_ro) : nullptr,
                                           at(cur->ro, 1), at(cur->jx, 1), at(cur->jy, 1),
                                           at(cur->jz, 1), st, timer_of(c));
    }
    if (e != hipSuccess) return hip_fail(c, e, "deposit");
    c->order_x_init[col] = s->x_init;
    c->order_n[col] = s->n;
    return LC3D_OK;
}

// The cell order of the last deposit into this sort's column, when it
// belongs to this sort.
const int *push_order(const lc3d_ctx *c, const lc3d_sort *s)
{
    const int col = s->column;
    return (s->n > 0 && c->order_n[col] == s->n && c->order_x_init[col] == s->x_init)
               ? c->idx + col * std::max<int64_t>(1, c->cap)
               : nullptr;
}

// as_f32: an FP32-bulk context pushes float64 fields rounded to float32 (the
// step's first push, from the previous step's full fields).
int push_sort(lc3d_ctx *c, const lc3d_sort *s, const void *const *f, bool f32, bool first,
              bool bin, hipStream_t st, bool as_f32 = false)
{
    const lc3d_fields6 fl{f[0], f[1], f[2], f[3], f[4], f[5]};
    hipError_t e;
    if (fp32(c)) {
        if (!f32 && as_f32)
            e = launch_push_f64_fields_as_f32(push_geom(c, s->n), f6<double>(&fl), c->fields32,
                                              sort_arrays<float>(s), first, bin,
                                              bin_state(c, s->column), push_order(c, s), st,
                                              timer_of(c));
        else if (f32)
            e = launch_push<float, float>(push_geom(c, s->n), f6<float>(&fl), sort_arrays<float>(s),
                                          first, bin, bin_state(c, s->column), push_order(c, s), st,
                                          timer_of(c));
        else
            e = launch_push<float, double>(push_geom(c, s->n), f6<double>(&fl),
                                           sort_arrays<float>(s), first, bin, bin_state(c, s->column),
                                           push_order(c, s), st, timer_of(c));
    } else {
        if (f32) return fail(c, LC3D_EINVAL, "float32 fields need an FP32-bulk context");
        e = launch_push<double, double>(push_geom(c, s->n), f6<double>(&fl),
                                        sort_arrays<double>(s), first, bin, bin_state(c, s->column),
                                        push_order(c, s), st, timer_of(c));
    }
    if (e != hipSuccess) return hip_fail(c, e, "push");
    return LC3D_OK;
}

int solve(lc3d_ctx *c, const lc3d_fields *pred, bool pred_f32, const lc3d_fields *prev,
          const double *rbF, const double *rbP, const lc3d_currents *cp, const lc3d_currents *cf,
          lc3d_fields *full, lc3d_fields *half, bool write_full, hipStream_t st)
{
    hipError_t e;
    if (fp32(c)) {
        if (pred_f32)
            e = run_solve<float, float, double, float>(
                c, solve_in<float, float>(c, pred, rbF, rbP, cp, cf),
                solve_out<double, float>(prev, full, half), write_full, st);
        else
            e = run_solve<float, double, double, float>(
                c, solve_in<float, double>(c, pred, rbF, rbP, cp, cf),
                solve_out<double, float>(prev, full, half), write_full, st);
    } else {
        if (pred_f32) return fail(c, LC3D_EINVAL, "float32 predictor needs an FP32-bulk context");
        e = run_solve<double, double, double, double>(
            c, solve_in<double, double>(c, pred, rbF, rbP, cp, cf),
            solve_out<double, double>(prev, full, half), write_full, st);
    }
    if (e != hipSuccess) return hip_fail(c, e, "compute_fields");
    return LC3D_OK;
}

}  // namespace

extern "C" {

int lc3d_create(const lc3d_params *p, lc3d_ctx **out)
{
    if (!p || !out) return LC3D_EINVAL;
    *out = nullptr;
    const int N = p->grid_steps;
    if (N < 9 || N % 2 == 0 || N - 1 > 1024 || p->grid_step_size <= 0 || p->xi_step_size <= 0 ||
        p->max_particles < 0 || p->max_particles > 0x7fffffff ||
        (p->precision != LC3D_FP64 && p->precision != LC3D_FP32_BULK))
        return LC3D_EINVAL;
    lc3d_ctx *c = new lc3d_ctx();
    c->prm = *p;
    c->prm.eigenvalues = nullptr;
    c->N = N;
    c->cap = p->max_particles;
    const int M = N - 1;
    c->solver = p->field_solver;
#ifdef LC3D_FFT_SOLVER
    const bool fft = c->solver == LC3D_SOLVER_FFT;
#else
    const bool fft = false;
#endif
    const bool rfft = c->solver == LC3D_SOLVER_RFFT;
    if (c->solver != LC3D_SOLVER_MATRIX && !fft && !rfft) {
        delete c;
        return LC3D_EINVAL;   // unknown solver, or the FFT solver in a build without it
    }
    if (rfft) {
        const int L = 2 * M;
        const bool f32s = p->precision == LC3D_FP32_BULK;
        // Kernels by waves (a mask of the three); the others need the radix
        // plan and launch shape.
        c->rf_wave = p->rfft_wave;
        if (c->rf_wave < 0 || c->rf_wave > 7 || (c->rf_wave && !rfft_wave_supported(L))) {
            delete c;
            return LC3D_EINVAL;
        }
        if (c->rf_wave == 7) {
            c->rf_stages = 0;
            c->rf_cap = 0;
            c->rf_tpb = 0;
        } else {
            c->rf_stages = p->fft_nstages > 0 ? rfft_pack_plan(L, p->fft_nstages, p->fft_radix)
                                              : rfft_default_plan(L);
            c->rf_cap = p->rfft_cap > 0 ? p->rfft_cap : (f32s ? 8192 : 4096);
            c->rf_tpb = p->rfft_tpb > 0 ? p->rfft_tpb : c->rf_cap / (2 * L);
            if (c->rf_stages == 0 || !rfft_cap_supported(f32s, c->rf_cap) || c->rf_tpb < 1 ||
                2 * (int64_t)c->rf_tpb * L > c->rf_cap) {
                delete c;
                return LC3D_EINVAL;
            }
        }
    }
    if (fft) {
        if (p->fft_nstages > 0) {
            c->plan.M = M;
            c->plan.ns = p->fft_nstages;
            for (int s = 0; s < 12; ++s) c->plan.radix[s] = s < p->fft_nstages ? p->fft_radix[s] : 0;
            if (p->fft_nstages > 12 || !valid_plan(c->plan)) {
                delete c;
                return LC3D_EINVAL;
            }
        } else if (!default_plan(M, c->plan)) {
            delete c;
            return LC3D_EINVAL;
        }
    }
    if (p->declustering_side != 0) {
        const int side = p->declustering_side, nt = p->declustering_ntaps;
        if (side < 3 || side > 4096 || nt < 1 || nt > side || !p->declustering_taps ||
            p->declustering_limit <= 0) {
            delete c;
            return LC3D_EINVAL;
        }
        c->dc_side = side;
        c->dc_ntaps = nt;
        c->dc_tap0 = p->declustering_tap0;
    }
    c->prm.declustering_taps = nullptr;
    (void)hipGetDevice(&c->device);
    const size_t nn = (size_t)N * N;
    const size_t nblocks = (size_t)scan_tiles((int64_t)nn);
    c->nbase = (int)nblocks;
    const bool f32 = p->precision == LC3D_FP32_BULK;
    const size_t tg = f32 ? sizeof(float) : sizeof(double);
    const size_t cap = (size_t)std::max<int64_t>(1, c->cap);
    bool ok = hipMalloc(&c->count, (nn + 1) * sizeof(int)) == hipSuccess &&
              hipMalloc(&c->start, (nn + 1) * sizeof(int)) == hipSuccess &&
              hipMalloc(&c->lstart, (nn + 1) * sizeof(int)) == hipSuccess &&
              hipMalloc(&c->base, nblocks * sizeof(int)) == hipSuccess &&
              hipMalloc(&c->key, cap * sizeof(int)) == hipSuccess &&
              hipMalloc(&c->rank, cap * sizeof(int)) == hipSuccess &&
              hipMalloc(&c->idx, 2 * cap * sizeof(int)) == hipSuccess &&
              hipMalloc(&c->err, sizeof(unsigned)) == hipSuccess &&
              hipMalloc(&c->rec, cap * deposit_record_bytes(f32)) == hipSuccess &&
              hipMalloc(&c->pid, cap * sizeof(int)) == hipSuccess &&
              hipMalloc(&c->pos, cap * sizeof(int)) == hipSuccess &&
              hipMalloc(&c->big, dense_cell_capacity(cap) * sizeof(int)) == hipSuccess &&
              hipMalloc(&c->nbig, sizeof(int)) == hipSuccess &&
              hipMalloc(&c->dslot, nn * sizeof(int)) == hipSuccess &&
              hipMalloc(&c->dense, dense_cell_capacity(cap) * 100 * tg) == hipSuccess &&
              hipMalloc(&c->work, 7 * nn * tg) == hipSuccess &&
              hipMalloc(&c->rowsum, N * sizeof(double)) == hipSuccess &&
              hipMalloc(&c->diag, 2 * DIAG_BLOCKS * sizeof(double)) == hipSuccess &&
              hipMalloc(&c->beam_acc, 6 * nn * sizeof(unsigned long long)) == hipSuccess &&
              (!f32 || hipMalloc(&c->fields32, 6 * fields32_stride(N) * sizeof(float)) ==
                           hipSuccess) &&
              hipMemset(c->beam_acc, 0, 6 * nn * sizeof(unsigned long long)) == hipSuccess &&
              hipMalloc(&c->lam, N * sizeof(double)) == hipSuccess;
    if (ok && (fft || rfft)) ok = hipMalloc(&c->tw, 2 * (size_t)M * 2 * tg) == hipSuccess;
    if (ok && c->dc_side > 0)
        ok = hipMalloc(&c->dc_taps, c->dc_ntaps * sizeof(double)) == hipSuccess &&
             hipMalloc(&c->dc_scratch, 4 * (size_t)c->dc_side * c->dc_side * sizeof(double)) ==
                 hipSuccess;
    if (ok && !fft && !rfft)
        ok = hipMalloc(&c->work2, 7 * nn * tg) == hipSuccess &&
             hipMalloc(&c->tdst, nn * tg) == hipSuccess &&
             hipMalloc(&c->tdct, nn * tg) == hipSuccess &&
             hipMalloc(&c->tdct_t, nn * tg) == hipSuccess;
    if (!ok) {
        lc3d_destroy(c);
        return LC3D_ENOMEM;
    }
    const double pi = 3.14159265358979323846;
    std::vector<double> lam(N);
    for (int k = 0; k < N; ++k) {
        if (p->eigenvalues) {
            lam[k] = p->eigenvalues[k];
        } else {
            const double s = std::sin(k * pi / (2.0 * (N - 1)));
            lam[k] = 4 / (p->grid_step_size * p->grid_step_size) * (s * s);
        }
    }
    // Uploads a float64 host table in the solve's type.
    auto upload = [&](void *dst, const std::vector<double> &v) {
        if (f32) {
            std::vector<float> w(v.begin(), v.end());
            return hipMemcpy(dst, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice);
        }
        return hipMemcpy(dst, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice);
    };
    hipError_t e = hipSuccess;
    if (fft || rfft) {
        // Twiddles exp(-i pi t / M) = exp(-2 pi i t / L), L = 2M, t = 0..L-1.
        std::vector<double> twd(4 * (size_t)M);
        for (int t = 0; t < 2 * M; ++t) {
            twd[2 * t] = std::cos(pi * t / M);
            twd[2 * t + 1] = -std::sin(pi * t / M);
        }
        e = upload(c->tw, twd);
    } else {
        // S[k][n] = 2 sin(pi k n / M) (zero for k or n in {0, M}) and
        // C[k][n] = w_n cos(pi k n / M), w_0 = w_M = 1, else 2: the
        // unnormalised DST-I / DCT-I (fields.py:15-109) as N x N matrices.
        // k n is reduced mod 2M so every angle is at most 2 pi.
        std::vector<double> S(nn), C(nn), Ct(nn);
        for (int k = 0; k < N; ++k) {
            for (int n = 0; n < N; ++n) {
                const int r = (int)(((int64_t)k * n) % (2 * M));
                const double sn = (k == 0 || k == M || n == 0 || n == M) ? 0.0 : 2 * std::sin(pi * r / M);
                const double cs = (n == 0 || n == M ? 1.0 : 2.0) * std::cos(pi * r / M);
                S[(size_t)k * N + n] = sn;
                C[(size_t)k * N + n] = cs;
                Ct[(size_t)n * N + k] = cs;
            }
        }
        e = upload(c->tdst, S);
        if (e == hipSuccess) e = upload(c->tdct, C);
        if (e == hipSuccess) e = upload(c->tdct_t, Ct);
    }
    if (e == hipSuccess) e = hipMemcpy(c->lam, lam.data(), N * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess && c->dc_side > 0)
        e = hipMemcpy(c->dc_taps, p->declustering_taps, c->dc_ntaps * sizeof(double),
                      hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(c->count, 0, (nn + 1) * sizeof(int));
    if (e == hipSuccess) e = hipMemset(c->err, 0, sizeof(unsigned));
    if (e == hipSuccess) e = hipMemset(c->base, 0, c->nbase * sizeof(int));
    if (e == hipSuccess) e = hipMemset(c->nbig, 0, sizeof(int));
    c->bigcap = dense_cell_capacity(cap);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        lc3d_destroy(c);
        return LC3D_EHIP;
    }
    *out = c;
    return LC3D_OK;
}

void lc3d_destroy(lc3d_ctx *c)
{
    if (!c) return;
    if (c->timer.on) {
        (void)c->timer.harvest();
        for (hipEvent_t &e : c->timer.ev) (void)hipEventDestroy(e);
    }
    (void)hipFree(c->count);
    (void)hipFree(c->start);
    (void)hipFree(c->lstart);
    (void)hipFree(c->base);
    (void)hipFree(c->key);
    (void)hipFree(c->rank);
    (void)hipFree(c->idx);
    (void)hipFree(c->err);
    (void)hipFree(c->rec);
    (void)hipFree(c->pid);
    (void)hipFree(c->pos);
    (void)hipFree(c->big);
    (void)hipFree(c->nbig);
    (void)hipFree(c->dslot);
    (void)hipFree(c->dense);
    (void)hipFree(c->work);
    (void)hipFree(c->work2);
    (void)hipFree(c->rowsum);
    (void)hipFree(c->diag);
    (void)hipFree(c->beam_acc);
    (void)hipFree(c->fields32);
    (void)hipFree(c->tw);
    (void)hipFree(c->tdst);
    (void)hipFree(c->tdct);
    (void)hipFree(c->tdct_t);
    (void)hipFree(c->lam);
    (void)hipFree(c->dc_taps);
    (void)hipFree(c->dc_scratch);
    delete c;
}

const char *lc3d_last_error(const lc3d_ctx *c)
{
    return c ? c->last_error.c_str() : noalloc_last_error();
}

int lc3d_precision(const lc3d_ctx *c) { return c ? c->prm.precision : -1; }

int lc3d_kernel_timing(lc3d_ctx *c, int32_t enable)
{
    if (!c) return LC3D_EINVAL;
    KernelTimer &t = c->timer;
    if (t.on) {
        HIP_TRY(c, t.harvest());
        for (hipEvent_t &e : t.ev) (void)hipEventDestroy(e);
        t.on = false;
    }
    if (enable) {
        for (hipEvent_t &e : t.ev) HIP_TRY(c, hipEventCreate(&e));
        for (int k = 0; k < KC_COUNT; ++k) {
            t.seconds[k] = 0;
            t.launches[k] = 0;
        }
        t.used = 0;
        t.on = true;
    }
    return LC3D_OK;
}

int lc3d_kernel_times(lc3d_ctx *c, double *seconds, int64_t *launches)
{
    if (!c || !seconds || !launches) return LC3D_EINVAL;
    if (c->timer.on) HIP_TRY(c, c->timer.harvest());
    for (int k = 0; k < KC_COUNT; ++k) {
        launches[k] = c->timer.launches[k];
        seconds[k] = launches[k] ? c->timer.seconds[k] / (double)launches[k] : 0.0;
    }
    return LC3D_OK;
}

int lc3d_error_flags(lc3d_ctx *c, uint32_t *flags, void *stream)
{
    if (!c || !flags) return LC3D_EINVAL;
    hipStream_t st = static_cast<hipStream_t>(stream);
    unsigned v = 0;
    HIP_TRY(c, hipMemcpyAsync(&v, c->err, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIP_TRY(c, hipStreamSynchronize(st));
    HIP_TRY(c, hipMemsetAsync(c->err, 0, sizeof(unsigned), st));
    HIP_TRY(c, hipStreamSynchronize(st));
    *flags = v;
    return LC3D_OK;
}

int lc3d_move_wo_fields(lc3d_ctx *c, int64_t n, const void *m, double *x_offt, double *y_offt,
                        const void *px, const void *py, const double *pz, void *stream)
{
    if (!c || n < 0 || (n > 0 && (!m || !x_offt || !y_offt || !px || !py || !pz)))
        return fail(c, LC3D_EINVAL, "move_wo_fields: bad arguments");
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t e;
    if (fp32(c))
        e = launch_move_wo_fields<float>(n, c->prm.xi_step_size, static_cast<const float *>(m),
                                         x_offt, y_offt, static_cast<const float *>(px),
                                         static_cast<const float *>(py),
                                         pz, st);
    else
        e = launch_move_wo_fields<double>(n, c->prm.xi_step_size, static_cast<const double *>(m),
                                          x_offt, y_offt, static_cast<const double *>(px),
                                          static_cast<const double *>(py),
                                          pz, st);
    if (e != hipSuccess) return hip_fail(c, e, "move_wo_fields");
    return LC3D_OK;
}

int lc3d_move_smart(lc3d_ctx *c, const lc3d_sort *s, const lc3d_fields6 *f, int32_t field_fp32,
                    void *stream)
{
    if (!c || !sort_ok(c, s) || !f || !f->Ex || !f->Ey || !f->Ez || !f->Bx || !f->By || !f->Bz)
        return fail(c, LC3D_EINVAL, "move_smart: bad arguments");
    const void *fl[6] = {f->Ex, f->Ey, f->Ez, f->Bx, f->By, f->Bz};
    return push_sort(c, s, fl, field_fp32 != 0, false, false, static_cast<hipStream_t>(stream));
}

int lc3d_deposit(lc3d_ctx *c, const lc3d_sort *s, void *ro, void *jx, void *jy, void *jz,
                 const void *bg_ro, void *ro1, void *jx1, void *jy1, void *jz1, void *stream)
{
    if (!c || !sort_ok(c, s) || !ro || !jx || !jy || !jz ||
        (bg_ro && (!ro1 || !jx1 || !jy1 || !jz1)))
        return fail(c, LC3D_EINVAL, "deposit: bad arguments");
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t e;
    if (fp32(c))
        e = launch_deposit<float, float>(push_geom(c, s->n), sort_arrays<float>(s), true, false,
                                         bin_state(c, s->column), static_cast<float *>(ro),
                                         static_cast<float *>(jx), static_cast<float *>(jy),
                                         static_cast<float *>(jz), static_cast<const float *>(bg_ro),
                                         static_cast<float *>(ro1), static_cast<float *>(jx1),
                                         static_cast<float *>(jy1), static_cast<float *>(jz1), st, timer_of(c));
    else
        e = launch_deposit<double, double>(push_geom(c, s->n), sort_arrays<double>(s), true, false,
                                           bin_state(c, s->column), static_cast<double *>(ro),
                                           static_cast<double *>(jx), static_cast<double *>(jy),
                                           static_cast<double *>(jz),
                                           static_cast<const double *>(bg_ro),
                                           static_cast<double *>(ro1), static_cast<double *>(jx1),
                                           static_cast<double *>(jy1), static_cast<double *>(jz1),
                                           st, timer_of(c));
    if (e != hipSuccess) return hip_fail(c, e, "deposit");
    c->order_x_init[s->column] = s->x_init;
    c->order_n[s->column] = s->n;
    return LC3D_OK;
}

int lc3d_compute_fields(lc3d_ctx *c, const lc3d_fields *pred, int32_t pred_fp32,
                        const lc3d_fields *prev, const double *rbF, const double *rbP,
                        const lc3d_currents *cp, const lc3d_currents *cf, lc3d_fields *full,
                        lc3d_fields *half, void *stream)
{
    if (!c || !pred || !pred->Ex || !pred->Ey || !pred->Bx || !pred->By || !fields_ok(prev, true) ||
        !rbF || !rbP || !currents_ok(cp) || !currents_ok(cf) || !fields_ok(full, true) ||
        !fields_ok(half, true))
        return fail(c, LC3D_EINVAL, "compute_fields: bad arguments");
    return solve(c, pred, pred_fp32 != 0, prev, rbF, rbP, cp, cf, full, half, true,
                 static_cast<hipStream_t>(stream));
}

int lc3d_plasma_step(lc3d_ctx *c, int32_t n_sorts, const lc3d_sort *sorts,
                     const lc3d_fields *prev, const lc3d_currents *cp, const double *rbF,
                     const double *rbP, const void *bg_ro, lc3d_fields *full, lc3d_fields *half,
                     const lc3d_currents *cf, void *stream)
{
    if (!c || n_sorts < 1 || n_sorts > 2 || !sorts || !fields_ok(prev, true) ||
        !currents_ok(cp) || !rbF || !rbP || !fields_ok(full, true) || !fields_ok(half, true) ||
        !currents_ok(cf))
        return fail(c, LC3D_EINVAL, "plasma_step: bad arguments");
    for (int s = 0; s < n_sorts; ++s)
        if (!sort_ok(c, &sorts[s])) return fail(c, LC3D_EINVAL, "plasma_step: bad sort");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const bool f32 = fp32(c);
    const bool single = n_sorts == 1;
    const void *fprev[6] = {prev->Ex, prev->Ey, prev->Ez, prev->Bx, prev->By, prev->Bz};
    const void *fhalf[6] = {half->Ex, half->Ey, half->Ez, half->Bx, half->By, half->Bz};
    int r;
    // solver.py:29-42: free stream + push 1 in fields_prev, deposit 1.
    for (int s = 0; s < n_sorts; ++s) {
        if ((r = push_sort(c, &sorts[s], fprev, false, true, true, st, true))) return r;
        if ((r = deposit_sort(c, &sorts[s], true, cf, bg_ro, single, st))) return r;
    }
    // solver.py:44-47: solve 1 with fields_prev as predictor; only the half fields are kept.
    if ((r = solve(c, prev, false, prev, rbF, rbP, cp, cf, full, half, false, st))) return r;
    // solver.py:49-57: push 2 in fields_half, deposit 2.
    for (int s = 0; s < n_sorts; ++s) {
        if ((r = push_sort(c, &sorts[s], fhalf, f32, false, true, st))) return r;
        if ((r = deposit_sort(c, &sorts[s], true, cf, bg_ro, single, st))) return r;
    }
    // solver.py:59-62: solve 2 with fields_half as predictor.
    if ((r = solve(c, half, f32, prev, rbF, rbP, cp, cf, full, half, true, st))) return r;
    // solver.py:64-76: push 3 in the new fields_half, the electrons' noise
    // filter when on (it changes px, py, so that sort is binned by its
    // deposit instead of by the push), deposit 3.
    for (int s = 0; s < n_sorts; ++s) {
        const bool filter = s == 0 && declust_ok(c, &sorts[s]);
        if ((r = push_sort(c, &sorts[s], fhalf, f32, false, !filter, st))) return r;
        if (filter && (r = noise_filter_sort(c, &sorts[s], st))) return r;
        if ((r = deposit_sort(c, &sorts[s], !filter, cf, bg_ro, single, st))) return r;
    }
    return LC3D_OK;
}

int lc3d_noise_filter(lc3d_ctx *c, const lc3d_sort *s, void *stream)
{
    if (!c || !s || !sort_ok(c, s))
        return fail(c, LC3D_EINVAL, "noise_filter: bad arguments");
    if (!declust_ok(c, s))
        return fail(c, LC3D_EINVAL, "noise_filter: the context has no filter for a lattice of "
                                    "this size, or the sort lacks its chaotic arrays");
    return noise_filter_sort(c, s, static_cast<hipStream_t>(stream));
}

int lc3d_deposit_beam(lc3d_ctx *c, int64_t n, int64_t layer, const double *x, const double *y,
                      const double *xi, const double *q_norm, double *ro_running,
                      double *ro_next, void *stream)
{
    if (!c || n < 0 || !ro_running || !ro_next || (n > 0 && (!x || !y || !xi || !q_norm)))
        return fail(c, LC3D_EINVAL, "deposit_beam: bad arguments");
    const BeamGeom g{c->N, c->prm.grid_step_size, c->prm.xi_step_size, c->prm.lost_radius};
    hipError_t e = launch_beam_deposit(g, n, layer, x, y, xi, q_norm, c->beam_acc, ro_running,
                                       ro_next, c->err, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(c, e, "deposit_beam");
    return LC3D_OK;
}

int lc3d_layout_beam_layer(lc3d_ctx *c, int64_t n, int64_t layer, const double *x,
                           const double *y, const double *xi, const double *q_norm,
                           double *ro_running, double *ro_next, void *stream)
{
    if (!c || n < 0 || !ro_running || !ro_next || (n > 0 && (!x || !y || !xi || !q_norm)))
        return fail(c, LC3D_EINVAL, "layout_beam_layer: bad arguments");
    const BeamGeom g{c->N, c->prm.grid_step_size, c->prm.xi_step_size, c->prm.lost_radius};
    hipError_t e = launch_beam_deposit(g, n, layer, x, y, xi, q_norm, c->beam_acc, ro_running,
                                       ro_next, c->err, static_cast<hipStream_t>(stream), true);
    if (e != hipSuccess) return hip_fail(c, e, "layout_beam_layer");
    return LC3D_OK;
}

int lc3d_beam_substep(lc3d_ctx *c, int64_t n, const double *q_m, const double *pz, double e_sub,
                      double time_step, double *dt, int64_t *steps, void *stream)
{
    if (!c || n < 0 || (n > 0 && (!q_m || !pz || !dt || !steps)))
        return fail(c, LC3D_EINVAL, "beam_substep: bad arguments");
    hipError_t e = launch_beam_substep(n, q_m, pz, e_sub, time_step, dt, steps,
                                       static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(c, e, "beam_substep");
    return LC3D_OK;
}

int lc3d_move_beam(lc3d_ctx *c, int64_t n, int64_t beam_layer_idx, const double *q_m,
                   const double *dt, const lc3d_fields6 *fk, const lc3d_fields6 *fk1,
                   int64_t *steps, int64_t *id, double *x, double *y, double *xi, double *px,
                   double *py, double *pz, uint8_t *lost, uint8_t *moved, uint8_t *fell,
                   void *stream)
{
    if (!c || n < 0 || !fk || !fk1 ||
        (n > 0 && (!q_m || !dt || !steps || !id || !x || !y || !xi || !px || !py || !pz || !lost ||
                   !moved || !fell)))
        return fail(c, LC3D_EINVAL, "move_beam: bad arguments");
    const BeamGeom g{c->N, c->prm.grid_step_size, c->prm.xi_step_size, c->prm.lost_radius};
    hipError_t e = launch_beam_move<double>(g, n, beam_layer_idx, q_m, dt, f6<double>(fk),
                                            f6<double>(fk1), steps, id, x, y, xi, px, py, pz,
                                            lost, moved, fell, c->err,
                                            static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(c, e, "move_beam");
    return LC3D_OK;
}

int lc3d_beam_split(lc3d_ctx *c, int64_t n, int32_t nf, int32_t ni, const double *f,
                    const int64_t *iv, const uint8_t *flags, double *out_f, int64_t *out_i,
                    int64_t *counts, void *stream)
{
    if (!c || n < 0 || nf < 0 || ni < 0 || !counts ||
        (n > 0 && ((nf > 0 && (!f || !out_f)) || (ni > 0 && (!iv || !out_i)) || !flags)))
        return fail(c, LC3D_EINVAL, "beam_split: bad arguments");
    hipError_t e = launch_beam_split(n, nf, ni, f, iv, flags, out_f, out_i, counts,
                                     static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(c, e, "beam_split");
    return LC3D_OK;
}

int lc3d_energy_flux(lc3d_ctx *c, const lc3d_fields6 *f, int64_t n, const void *m,
                     const void *px, const void *py, const double *pz, double *psi, void *stream)
{
    if (!c || !f || !f->Ex || !f->Ey || !f->Ez || !f->Bx || !f->By || !f->Bz || !psi || n < 0 ||
        (n > 0 && (!m || !px || !py || !pz)))
        return fail(c, LC3D_EINVAL, "energy_flux: bad arguments");
    const double h = c->prm.grid_step_size;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t e;
    if (fp32(c))
        e = launch_energy_flux<float>(c->N, h, f6<double>(f), n, static_cast<const float *>(m),
                                      static_cast<const float *>(px),
                                      static_cast<const float *>(py), pz, c->diag, DIAG_BLOCKS,
                                      psi, st);
    else
        e = launch_energy_flux<double>(c->N, h, f6<double>(f), n, static_cast<const double *>(m),
                                       static_cast<const double *>(px),
                                       static_cast<const double *>(py), pz, c->diag, DIAG_BLOCKS,
                                       psi, st);
    if (e != hipSuccess) return hip_fail(c, e, "energy_flux");
    return LC3D_OK;
}

int lc3d_probe_sample(lc3d_ctx *c, const lc3d_probe_sources *src, int32_t n_points,
                      const int32_t *ix, const int32_t *iy, double *out, void *stream)
{
    if (!c || !src || src->n < 0 || src->n > LC3D_PROBE_MAX_SOURCES || n_points < 0 ||
        (n_points > 0 && src->n > 0 && (!ix || !iy || !out)))
        return fail(c, LC3D_EINVAL, "probe_sample: bad arguments");
    ProbeSources ps{};
    ps.n = src->n;
    for (int k = 0; k < src->n; ++k) {
        if (!src->src[k]) return fail(c, LC3D_EINVAL, "probe_sample: null source");
        ps.src[k] = src->src[k];
        ps.is_f32[k] = src->is_f32[k] != 0;
    }
    hipError_t e = launch_probe_sample(ps, c->N, n_points, ix, iy, out,
                                       static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(c, e, "probe_sample");
    return LC3D_OK;
}

int lc3d_abs_max(lc3d_ctx *c, const double *a, int64_t nr, int64_t nc, int64_t r0, int64_t r1,
                 int64_t c0, int64_t c1, double *out, void *stream)
{
    if (!c || !out || nr < 0 || nc < 0 || r0 < 0 || c0 < 0 || r1 > nr || c1 > nc || r0 > r1 ||
        c0 > c1 || (r1 > r0 && c1 > c0 && !a))
        return fail(c, LC3D_EINVAL, "abs_max: bad arguments");
    hipError_t e = launch_absmax(a, nc, r0, r1, c0, c1, c->diag, 2 * DIAG_BLOCKS, out,
                                 static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(c, e, "abs_max");
    return LC3D_OK;
}

}  // extern "C"
