// Grid execution of the CPU HIP emulation (see hip/hip_runtime.h).
//
// A launch hands its blocks to a pool of OS threads.  A worker runs one
// block at a time: every GPU thread of the block is a fiber with its own
// guarded stack; the worker resumes runnable fibers in thread order, and a
// fiber gives control back when it reaches a barrier or ends.  A barrier
// opens when every thread of its scope (block, or 64-thread wave) has
// arrived; if no fiber can run while some have not ended, the block is
// deadlocked and the run aborts with the block and barrier named.
#include "hip/hip_runtime.h"
#include <atomic>
#include <condition_variable>
#include <mutex>
#include <sys/mman.h>
#include <ucontext.h>
#include <unistd.h>

thread_local dim3 threadIdx, blockIdx;
dim3 gridDim, blockDim;
thread_local double emu_shfl_d[1024];
thread_local float emu_shfl_f[1024];
thread_local int emu_shfl_i[1024];
thread_local unsigned char *emu_lds = nullptr;

namespace {
constexpr size_t STACK = 256 * 1024;
enum { RUN, WAIT_BLOCK, WAIT_WAVE, WAIT_GROUP, DONE };

struct Fiber { ucontext_t ctx; int state; unsigned gen; char *stack; unsigned group = 0; };

struct BlockRun {
    std::vector<Fiber> f;
    ucontext_t sched;
    unsigned nt = 0, cur = 0;
    unsigned block_arrived = 0, block_gen = 0;
    unsigned wave_arrived[16] = {}, wave_gen[16] = {};
    // groups of `width` consecutive threads (width-limited shuffles), keyed
    // by the group's first thread
    unsigned group_arrived[1024] = {}, group_gen[1024] = {};
    const std::function<void()> *body = nullptr;
};
thread_local BlockRun R;

void fiber_main()
{
    (*R.body)();
    R.f[R.cur].state = DONE;   // uc_link returns to the scheduler
}

void ensure_stacks(unsigned nt)
{
    const size_t page = sysconf(_SC_PAGESIZE);
    while (R.f.size() < nt) {
        char *m = (char *)mmap(nullptr, STACK + page, PROT_READ | PROT_WRITE,
                               MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
        if (m == MAP_FAILED) { perror("emu: mmap"); abort(); }
        mprotect(m, page, PROT_NONE);   // guard page below the stack
        R.f.push_back(Fiber{{}, DONE, 0, m + page, 0});
    }
}

void run_block(const std::function<void()> &body, dim3 b, unsigned nt, size_t lds)
{
    ensure_stacks(nt);
    R.nt = nt;
    R.body = &body;
    R.block_arrived = R.block_gen = 0;
    for (int w = 0; w < 16; ++w) R.wave_arrived[w] = R.wave_gen[w] = 0;
    for (int g = 0; g < 1024; ++g) R.group_arrived[g] = R.group_gen[g] = 0;
    blockIdx = b;
    emu_lds = lds ? (unsigned char *)std::aligned_alloc(16, (lds + 15) / 16 * 16) : nullptr;
    for (unsigned t = 0; t < nt; ++t) {
        Fiber &f = R.f[t];
        getcontext(&f.ctx);
        f.ctx.uc_stack.ss_sp = f.stack;
        f.ctx.uc_stack.ss_size = STACK;
        f.ctx.uc_link = &R.sched;
        makecontext(&f.ctx, fiber_main, 0);
        f.state = RUN;
        f.gen = 0;
    }
    unsigned done = 0;
    while (done < nt) {
        bool progressed = false;
        for (unsigned t = 0; t < nt; ++t) {
            Fiber &f = R.f[t];
            bool ready = f.state == RUN || (f.state == WAIT_BLOCK && R.block_gen != f.gen) ||
                         (f.state == WAIT_WAVE && R.wave_gen[t / 64] != f.gen) ||
                         (f.state == WAIT_GROUP && R.group_gen[f.group] != f.gen);
            if (!ready) continue;
            f.state = RUN;
            R.cur = t;
            threadIdx = dim3(t, 0, 0);
            swapcontext(&R.sched, &f.ctx);
            if (f.state == DONE) ++done;
            progressed = true;
        }
        if (!progressed) {
            fprintf(stderr, "emu: deadlock in block (%u,%u,%u) of a %u-thread launch: %u of %u "
                    "threads ended, the rest wait at a barrier that not all reach\n",
                    b.x, b.y, b.z, nt, done, nt);
            abort();
        }
    }
    std::free(emu_lds);
    emu_lds = nullptr;
}

// Worker pool.
std::mutex mu;
std::condition_variable cv_go, cv_done;
unsigned gen = 0, busy = 0;
bool stop = false;
const std::function<void()> *cur_body;
dim3 cur_grid;
unsigned cur_nt;
size_t cur_lds;
std::atomic<unsigned long> next_block;
std::vector<std::thread> pool;

void worker()
{
    unsigned seen = 0;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_go.wait(lk, [&] { return stop || gen != seen; });
            if (stop) return;
            seen = gen;
        }
        const unsigned long nb = (unsigned long)cur_grid.x * cur_grid.y * cur_grid.z;
        for (unsigned long i; (i = next_block.fetch_add(1)) < nb;) {
            dim3 b(i % cur_grid.x, (i / cur_grid.x) % cur_grid.y, i / ((unsigned long)cur_grid.x * cur_grid.y));
            run_block(*cur_body, b, cur_nt, cur_lds);
        }
        std::lock_guard<std::mutex> lk(mu);
        if (--busy == 0) cv_done.notify_one();
    }
}

struct PoolGuard {
    ~PoolGuard()
    {
        { std::lock_guard<std::mutex> lk(mu); stop = true; }
        cv_go.notify_all();
        for (auto &t : pool) t.join();
    }
} guard;
}  // namespace

void emu_sync_block()
{
    Fiber &me = R.f[R.cur];
    if (++R.block_arrived == R.nt) {   // last to arrive opens the barrier
        R.block_arrived = 0;
        ++R.block_gen;
        return;
    }
    me.state = WAIT_BLOCK;
    me.gen = R.block_gen;
    swapcontext(&me.ctx, &R.sched);
}

void emu_sync_wave()
{
    Fiber &me = R.f[R.cur];
    const unsigned w = R.cur / 64, width = std::min(64u, R.nt - 64 * w);
    if (++R.wave_arrived[w] == width) {
        R.wave_arrived[w] = 0;
        ++R.wave_gen[w];
        return;
    }
    me.state = WAIT_WAVE;
    me.gen = R.wave_gen[w];
    swapcontext(&me.ctx, &R.sched);
}

void emu_sync_group(unsigned width)
{
    Fiber &me = R.f[R.cur];
    const unsigned first = R.cur / width * width;
    const unsigned n = std::min(width, R.nt - first);
    if (++R.group_arrived[first] == n) {
        R.group_arrived[first] = 0;
        ++R.group_gen[first];
        return;
    }
    me.state = WAIT_GROUP;
    me.group = first;
    me.gen = R.group_gen[first];
    swapcontext(&me.ctx, &R.sched);
}

void emu_launch(dim3 grid, dim3 block, size_t lds, const std::function<void()> &body)
{
    if (getenv("EMU_TRACE")) fprintf(stderr, "launch grid=%u,%u,%u block=%u lds=%zu\n", grid.x, grid.y, grid.z, block.x, lds);
    const unsigned nt = block.x * block.y * block.z;
    if (nt == 0 || nt > 1024 || block.y != 1 || block.z != 1) {
        fprintf(stderr, "emu: unsupported block shape %u,%u,%u\n", block.x, block.y, block.z);
        abort();
    }
    if ((unsigned long)grid.x * grid.y * grid.z == 0) return;
    std::unique_lock<std::mutex> lk(mu);
    if (pool.empty()) {
        unsigned n = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
        if (const char *e = getenv("EMU_THREADS")) n = std::max(1, atoi(e));
        for (unsigned i = 0; i < n; ++i) pool.emplace_back(worker);
    }
    gridDim = grid;
    blockDim = block;
    cur_body = &body;
    cur_grid = grid;
    cur_nt = nt;
    cur_lds = lds;
    next_block = 0;
    busy = pool.size();
    ++gen;
    cv_go.notify_all();
    cv_done.wait(lk, [&] { return busy == 0; });
}
