// Autotuner of the field solve (lc3d_autotune_solve): times one xi step's
// field solve (lc3d_compute_fields) for each candidate implementation and
// launch shape on synthetic inputs of the caller's geometry and precision,
// and returns the fastest.  Candidates: the transform-matrix products
// (LC3D_SOLVER_MATRIX) and the paired LDS FFT (LC3D_SOLVER_RFFT) with every
// radix plan below x every LDS size x transforms per block (all that fit,
// half of them, one).  Uses only the public C ABI; allocates its contexts
// and inputs and synchronises, so it is a set-up step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/lcodepy_amd.h"
#include "kernels.h"

namespace {

// Radix plans of the length-L FFT: the power of two split greedily into
// 16s, into 8s, into 4s and into 2s (remainders 8, 4, 2 last), the odd
// primes (3 5 7 11 13) after them or before them.
std::vector<std::vector<int>> rfft_plans(int L)
{
    int m = L, twos = 0;
    while (m % 2 == 0) { m /= 2; ++twos; }
    std::vector<int> odd;
    for (int q : {3, 5, 7, 11, 13})
        while (m % q == 0) { odd.push_back(q); m /= q; }
    std::vector<std::vector<int>> out;
    if (m != 1) return out;
    for (int big : {16, 8, 4, 2}) {
        int lg = big == 16 ? 4 : big == 8 ? 3 : big == 4 ? 2 : 1;
        std::vector<int> p2;
        int t = twos;
        while (t >= lg) { p2.push_back(big); t -= lg; }
        if (t == 3) p2.push_back(8);
        else if (t == 2) p2.push_back(4);
        else if (t == 1) p2.push_back(2);
        for (int first_odd = 0; first_odd < 2; ++first_odd) {
            std::vector<int> p = first_odd ? odd : p2;
            const std::vector<int> &q = first_odd ? p2 : odd;
            p.insert(p.end(), q.begin(), q.end());
            if (p.empty() || p.size() > 12) continue;
            if (std::find(out.begin(), out.end(), p) == out.end()) out.push_back(p);
        }
    }
    return out;
}

struct Inputs {
    std::vector<void *> bufs;
    lc3d_fields pred{}, prev{}, full{}, half{};
    lc3d_currents cp{}, cf{};
    double *rbf = nullptr, *rbp = nullptr;
    bool ok = true;

    void *alloc(size_t bytes, unsigned char fill)
    {
        void *p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) { ok = false; return nullptr; }
        bufs.push_back(p);
        if (hipMemset(p, fill, bytes) != hipSuccess) ok = false;
        return p;
    }
    // Finite non-zero synthetic data: every byte 0x3c makes each float64 /
    // float32 element the same ordinary number; the timing does not depend
    // on the values.
    Inputs(int N, bool f32)
    {
        const size_t nn = (size_t)N * N, tg = f32 ? 4 : 8;
        void **fl[4][7] = {{&pred.Ex, &pred.Ey, &pred.Ez, &pred.Bx, &pred.By, &pred.Bz, &pred.Phi},
                           {&prev.Ex, &prev.Ey, &prev.Ez, &prev.Bx, &prev.By, &prev.Bz, &prev.Phi},
                           {&full.Ex, &full.Ey, &full.Ez, &full.Bx, &full.By, &full.Bz, &full.Phi},
                           {&half.Ex, &half.Ey, &half.Ez, &half.Bx, &half.By, &half.Bz, &half.Phi}};
        for (int a = 0; a < 4; ++a)
            for (int k = 0; k < 7; ++k)
                *fl[a][k] = alloc(nn * ((a == 3 && f32) ? 4 : 8), 0x3c);
        void **cu[2][4] = {{&cp.ro, &cp.jx, &cp.jy, &cp.jz}, {&cf.ro, &cf.jx, &cf.jy, &cf.jz}};
        for (auto &c : cu)
            for (void **q : c) *q = alloc(2 * nn * tg, 0x3c);
        rbf = static_cast<double *>(alloc(nn * 8, 0x3c));
        rbp = static_cast<double *>(alloc(nn * 8, 0x3c));
    }
    ~Inputs()
    {
        for (void *p : bufs) (void)hipFree(p);
    }
};

// Average seconds of one lc3d_compute_fields of a context made from p, or a
// negative value when the candidate is refused or fails.
double time_candidate(const lc3d_params &p, Inputs &in, int repeats, hipStream_t st)
{
    lc3d_ctx *ctx = nullptr;
    if (lc3d_create(&p, &ctx) != LC3D_OK) return -1;
    double sec = -1;
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess) {
        bool ok = true;
        for (int w = 0; w < 2 && ok; ++w)
            ok = lc3d_compute_fields(ctx, &in.pred, 0, &in.prev, in.rbf, in.rbp, &in.cp, &in.cf,
                                     &in.full, &in.half, st) == LC3D_OK;
        ok = ok && hipEventRecord(a, st) == hipSuccess;
        for (int r = 0; r < repeats && ok; ++r)
            ok = lc3d_compute_fields(ctx, &in.pred, 0, &in.prev, in.rbf, in.rbp, &in.cp, &in.cf,
                                     &in.full, &in.half, st) == LC3D_OK;
        ok = ok && hipEventRecord(b, st) == hipSuccess && hipEventSynchronize(b) == hipSuccess;
        float ms = 0;
        if (ok && hipEventElapsedTime(&ms, a, b) == hipSuccess) sec = 1e-3 * ms / repeats;
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    lc3d_destroy(ctx);
    return sec;
}

void fill_choice(const lc3d_params &p, double sec, lc3d_solve_choice *c)
{
    std::memset(c, 0, sizeof(*c));
    c->field_solver = p.field_solver;
    c->fft_nstages = p.fft_nstages;
    for (int s = 0; s < 12; ++s) c->fft_radix[s] = p.fft_radix[s];
    c->rfft_cap = p.rfft_cap;
    c->rfft_tpb = p.rfft_tpb;
    c->rfft_wave = p.rfft_wave;
    c->seconds = sec;
}

}  // namespace

extern "C" int lc3d_autotune_solve(const lc3d_params *params, int32_t repeats,
                                   lc3d_solve_choice *best, lc3d_solve_choice *table,
                                   int32_t table_cap, int32_t *n_tried)
{
    if (!params || !best || repeats < 1 || (table && table_cap < 0)) return LC3D_EINVAL;
    const int N = params->grid_steps;
    if (N < 9 || N % 2 == 0 || N - 1 > 1024) return LC3D_EINVAL;
    const bool f32 = params->precision == LC3D_FP32_BULK;
    Inputs in(N, f32);
    if (!in.ok) return LC3D_ENOMEM;
    std::vector<lc3d_params> cands;
    lc3d_params base = *params;
    base.max_particles = 1;   // the solve needs no particle workspaces
    base.fft_nstages = 0;
    base.rfft_cap = base.rfft_tpb = 0;
    base.rfft_wave = 0;
    base.field_solver = LC3D_SOLVER_MATRIX;
    cands.push_back(base);
    const int L = 2 * (N - 1);
    if (lc3d::rfft_wave_supported(L)) {
        lc3d_params q = base;
        q.field_solver = LC3D_SOLVER_RFFT;
        q.rfft_wave = 7;
        cands.push_back(q);
    }
    const int caps32[3] = {2048, 4096, 8192}, caps64[3] = {1024, 2048, 4096};
    for (const std::vector<int> &plan : rfft_plans(L)) {
        for (int ci = 0; ci < 3; ++ci) {
            const int cap = f32 ? caps32[ci] : caps64[ci];
            const int tmax = cap / (2 * L);
            if (tmax < 1) continue;
            std::vector<int> tpbs = {tmax};
            if (tmax / 2 >= 1 && tmax / 2 != tmax) tpbs.push_back(tmax / 2);
            if (tmax > 2) tpbs.push_back(1);
            for (int tpb : tpbs) {
                lc3d_params q = base;
                q.field_solver = LC3D_SOLVER_RFFT;
                q.fft_nstages = (int32_t)plan.size();
                for (size_t s = 0; s < plan.size(); ++s) q.fft_radix[s] = plan[s];
                q.rfft_cap = cap;
                q.rfft_tpb = tpb;
                cands.push_back(q);
            }
        }
    }
    hipStream_t st = nullptr;
    int tried = 0, best_i = -1;
    double best_sec = 0;
    for (size_t i = 0; i < cands.size(); ++i) {
        const double sec = time_candidate(cands[i], in, repeats, st);
        if (table && tried < table_cap) fill_choice(cands[i], sec, &table[tried]);
        ++tried;
        if (sec > 0 && (best_i < 0 || sec < best_sec)) {
            best_i = (int)i;
            best_sec = sec;
        }
    }
    // Then, from the fastest Stockham form, each kernel by waves in turn,
    // kept when it is faster.
    if (best_i >= 0 && lc3d::rfft_wave_supported(L) && cands[best_i].field_solver == LC3D_SOLVER_RFFT &&
        cands[best_i].rfft_wave == 0) {
        lc3d_params cur = cands[best_i];
        for (int k = 0; k < 3; ++k) {
            lc3d_params q = cur;
            q.rfft_wave |= 1 << k;
            const double sec = time_candidate(q, in, repeats, st);
            if (table && tried < table_cap) fill_choice(q, sec, &table[tried]);
            ++tried;
            cands.push_back(q);
            if (sec > 0 && sec < best_sec) {
                cur = q;
                best_i = (int)cands.size() - 1;
                best_sec = sec;
            }
        }
    }
    if (n_tried) *n_tried = tried;
    if (best_i < 0) return LC3D_EHIP;
    fill_choice(cands[best_i], best_sec, best);
    return LC3D_OK;
}
