// CPU execution model of HIP kernels (test infrastructure; see
// hip/hip_runtime.h in this directory).
//
// Workgroups: a launch starts min(grid, EMU_RESIDENT) worker threads; each
// pulls workgroup indices in order (the resident set of a small GPU).
// Work-items: fibers with their own stacks, switched by a 6-register
// context switch; a per-workgroup scheduler resumes them round-robin.
// Rendezvous:
//   wave collectives  all live lanes of the wave must arrive at the same
//                     source line; the last one computes every lane's result
//   __syncthreads     all live work-items of the workgroup
//   s_sleep           the fiber yields; when only sleepers run, the OS
//                     thread yields too (cross-workgroup spin waits)
// A workgroup whose fibers can make no progress, or a launch that exceeds
// EMU_WATCHDOG_S seconds, prints each fiber's state and line and aborts.
#include <hip/hip_runtime.h>

#include <sys/mman.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

extern "C" void emu_ctx_switch(void **save_sp, void *load_sp);
asm(R"(
.text
.globl emu_ctx_switch
.hidden emu_ctx_switch
.type emu_ctx_switch,@function
emu_ctx_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
.size emu_ctx_switch,.-emu_ctx_switch
)");

uint64_t emu_realtime() {
  return (uint64_t)(std::chrono::duration_cast<std::chrono::nanoseconds>(
                        std::chrono::steady_clock::now().time_since_epoch()).count() / 10);
}

// dynamic LDS of the kernels (their `extern __shared__` arrays)
namespace tga {
__attribute__((aligned(16))) thread_local char smem[163840];
__attribute__((aligned(16))) thread_local uint32_t h[40960];
__attribute__((aligned(16))) thread_local uint32_t cnt[40960];
__attribute__((aligned(16))) thread_local uint32_t hist[40960];
}  // namespace tga

namespace emu {

thread_local dim3 t_block_idx, t_block_dim, t_grid_dim;

namespace {

constexpr size_t kStack = 256 * 1024;
constexpr size_t kLdsBytes = 163840;

enum State : int { RUNNABLE, WAIT_COLL, WAIT_BAR, SLEEP, DONE };

struct Fiber {
  void *sp = nullptr;
  char *stack = nullptr;
  dim3 tid;
  unsigned flat = 0, lane = 0, wave = 0;
  int state = RUNNABLE;
  int line = 0;
};

enum CollKind : int { C_BALLOT = 1, C_SHFL = 2, C_SHFL_UP = 3, C_FIRST = 4 };

struct WaveSync {
  int arrived = 0;
  uint64_t gen = 0;
  int kind = 0, line = 0;
  uint64_t val[64];
  int64_t arg[64];
  bool present[64];
  uint64_t res[64];
};

struct Block {
  std::vector<Fiber> f;
  std::vector<WaveSync> waves;
  unsigned n = 0, ndone = 0;
  unsigned bar_arrived = 0;
  uint64_t bar_gen = 0;
  void *sched_sp = nullptr;
  const std::function<void()> *body = nullptr;
  uint64_t block_id = 0;
};

thread_local Block *t_blk = nullptr;
thread_local Fiber *t_cur = nullptr;

std::mutex g_mu;
std::vector<Block *> g_running;  // workgroups in flight (for the watchdog dump)
void (*g_hang_hook)() = nullptr;

void dump_block(Block &b, const char *why) {
  std::fprintf(stderr, "[emu] workgroup %llu (%u work-items): %s\n", (unsigned long long)b.block_id, b.n, why);
  static const char *names[] = {"run", "wave-collective", "barrier", "spin", "done"};
  // one line per wave: state/line of each distinct (state, line) run
  for (size_t w = 0; w < b.waves.size(); ++w) {
    std::string s;
    int ps = -1, pl = -1, cnt = 0;
    auto flush = [&] {
      if (cnt) s += std::string(" ") + names[ps] + "@" + std::to_string(pl) + "x" + std::to_string(cnt);
    };
    for (unsigned l = 0; l < 64 && w * 64 + l < b.n; ++l) {
      const Fiber &f = b.f[w * 64 + l];
      if (f.state != ps || f.line != pl) { flush(); ps = f.state; pl = f.line; cnt = 0; }
      ++cnt;
    }
    flush();
    std::fprintf(stderr, "[emu]   wave %zu:%s\n", w, s.c_str());
  }
  std::fflush(stderr);
}

[[noreturn]] void die_block(Block &b, const char *why) {
  {
    std::lock_guard<std::mutex> lk(g_mu);
    dump_block(b, why);
  }
  std::abort();
}

void yield_to_sched() {
  Fiber *f = t_cur;
  emu_ctx_switch(&f->sp, t_blk->sched_sp);
}

void release_wave(WaveSync &w, Block &b, unsigned wave) {
  w.gen++;
  w.arrived = 0;
  for (unsigned l = 0; l < 64 && wave * 64 + l < b.n; ++l) {
    Fiber &f = b.f[wave * 64 + l];
    if (f.state == WAIT_COLL) f.state = RUNNABLE;
    w.present[l] = false;
  }
}

unsigned live_lanes(Block &b, unsigned wave) {
  unsigned c = 0;
  for (unsigned l = 0; l < 64 && wave * 64 + l < b.n; ++l) c += b.f[wave * 64 + l].state != DONE;
  return c;
}

void try_complete_wave(Block &b, unsigned wave) {
  WaveSync &w = b.waves[wave];
  if (w.arrived == 0 || (unsigned)w.arrived < live_lanes(b, wave)) return;
  const unsigned nl = std::min<unsigned>(64, b.n - wave * 64);
  if (w.kind == C_BALLOT) {
    uint64_t m = 0;
    for (unsigned l = 0; l < nl; ++l)
      if (w.present[l] && (w.val[l] & 1)) m |= 1ULL << l;
    for (unsigned l = 0; l < nl; ++l) w.res[l] = m;
  } else if (w.kind == C_FIRST) {
    unsigned f = 0;
    while (f < nl && !w.present[f]) ++f;
    for (unsigned l = 0; l < nl; ++l) w.res[l] = w.val[f];
  } else {
    for (unsigned l = 0; l < nl; ++l) {
      if (!w.present[l]) continue;
      const int64_t src = w.arg[l];
      if (src < 0 || src >= (int64_t)nl || !w.present[src]) w.res[l] = w.val[l];
      else w.res[l] = w.val[src];
    }
  }
  release_wave(w, b, wave);
}

void try_complete_barrier(Block &b) {
  if (b.bar_arrived == 0 || b.bar_arrived < b.n - b.ndone) return;
  b.bar_gen++;
  b.bar_arrived = 0;
  for (auto &f : b.f)
    if (f.state == WAIT_BAR) f.state = RUNNABLE;
}

uint64_t collective(int kind, int line, uint64_t v, int64_t arg) {
  Block &b = *t_blk;
  Fiber *f = t_cur;
  f->line = line;
  WaveSync &w = b.waves[f->wave];
  if (w.arrived == 0) {
    w.kind = kind;
    w.line = line;
  } else if (w.kind != kind || w.line != line) {
    char msg[160];
    std::snprintf(msg, sizeof msg, "wave %u: lanes at different collectives (line %d and line %d)", f->wave, w.line,
                  line);
    f->state = WAIT_COLL;
    die_block(b, msg);
  }
  w.val[f->lane] = v;
  w.arg[f->lane] = arg;
  w.present[f->lane] = true;
  w.arrived++;
  const uint64_t gen = w.gen;
  f->state = WAIT_COLL;
  try_complete_wave(b, f->wave);
  while (w.gen == gen) yield_to_sched();
  f->state = RUNNABLE;
  return w.res[f->lane];
}

void fiber_entry() {
  Fiber *f = t_cur;
  (*t_blk->body)();
  f->state = DONE;
  Block &b = *t_blk;
  b.ndone++;
  try_complete_wave(b, f->wave);
  try_complete_barrier(b);
  emu_ctx_switch(&f->sp, b.sched_sp);
  std::abort();  // never resumed
}

void run_block(Block &b, std::vector<char *> &stacks) {
  for (unsigned i = 0; i < b.n; ++i) {
    Fiber &f = b.f[i];
    f = Fiber{};
    f.flat = i;
    f.tid = dim3(i % t_block_dim.x, (i / t_block_dim.x) % t_block_dim.y, i / (t_block_dim.x * t_block_dim.y));
    f.lane = i % 64;
    f.wave = i / 64;
    f.stack = stacks[i];
    void **sp = (void **)(f.stack + kStack);
    *--sp = nullptr;                  // fake return address of fiber_entry
    *--sp = (void *)&fiber_entry;     // popped by the switch's `ret`
    for (int r = 0; r < 6; ++r) *--sp = nullptr;
    f.sp = sp;
  }
  for (auto &w : b.waves) { w = WaveSync{}; std::memset(w.present, 0, sizeof w.present); }
  b.ndone = 0;
  b.bar_arrived = 0;
  while (b.ndone < b.n) {
    unsigned ran = 0, spun = 0;
    for (unsigned i = 0; i < b.n; ++i) {
      Fiber &f = b.f[i];
      if (f.state != RUNNABLE && f.state != SLEEP) continue;
      const bool was_sleep = f.state == SLEEP;
      t_cur = &f;
      emu_ctx_switch(&b.sched_sp, f.sp);
      ++ran;
      if (was_sleep && f.state == SLEEP) ++spun;
    }
    if (ran == 0) die_block(b, "no work-item can proceed (barrier / collective mismatch)");
    if (ran == spun) std::this_thread::yield();
  }
}

struct Launch {
  dim3 grid, block;
  size_t lds;
  std::function<void()> body;
};

void run_launch(const Launch &L) {
  const uint64_t nblk = (uint64_t)L.grid.x * L.grid.y * L.grid.z;
  const unsigned nthr = L.block.x * L.block.y * L.block.z;
  if (nblk == 0 || nthr == 0) return;
  if (L.lds > kLdsBytes) {
    std::fprintf(stderr, "[emu] launch asks for %zu B of dynamic LDS (max %zu)\n", L.lds, kLdsBytes);
    std::abort();
  }
  if (nthr > 1024 || nthr % 64) {
    std::fprintf(stderr, "[emu] block of %u work-items (need a multiple of 64, <= 1024)\n", nthr);
    std::abort();
  }
  static const unsigned resident = [] {
    const char *e = std::getenv("EMU_RESIDENT");
    return e ? (unsigned)std::max(1, std::atoi(e)) : 8u;
  }();
  static const double watchdog = [] {
    const char *e = std::getenv("EMU_WATCHDOG_S");
    return e ? std::atof(e) : 600.0;
  }();
  const unsigned nw = (unsigned)std::min<uint64_t>(nblk, resident);
  std::atomic<uint64_t> next{0};
  std::atomic<unsigned> finished{0};
  std::vector<std::thread> th;
  for (unsigned w = 0; w < nw; ++w) {
    th.emplace_back([&, w] {
      (void)w;
      t_grid_dim = L.grid;
      t_block_dim = L.block;
      std::vector<char *> stacks(nthr);
      const size_t bytes = (size_t)nthr * kStack;
      char *mem = (char *)mmap(nullptr, bytes, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE,
                               -1, 0);
      if (mem == MAP_FAILED) { std::perror("[emu] mmap"); std::abort(); }
      for (unsigned i = 0; i < nthr; ++i) stacks[i] = mem + (size_t)i * kStack;
      Block b;
      b.n = nthr;
      b.f.resize(nthr);
      b.waves.resize((nthr + 63) / 64);
      b.body = &L.body;
      t_blk = &b;
      {
        std::lock_guard<std::mutex> lk(g_mu);
        g_running.push_back(&b);
      }
      for (;;) {
        const uint64_t id = next.fetch_add(1);
        if (id >= nblk) break;
        b.block_id = id;
        t_block_idx = dim3((uint32_t)(id % L.grid.x), (uint32_t)((id / L.grid.x) % L.grid.y),
                           (uint32_t)(id / ((uint64_t)L.grid.x * L.grid.y)));
        std::memset(tga::smem, 0xCD, std::min(L.lds + 64, kLdsBytes));  // LDS starts undefined
        run_block(b, stacks);
      }
      {
        std::lock_guard<std::mutex> lk(g_mu);
        g_running.erase(std::find(g_running.begin(), g_running.end(), &b));
      }
      t_blk = nullptr;
      munmap(mem, bytes);
      finished.fetch_add(1);
    });
  }
  const auto t0 = std::chrono::steady_clock::now();
  while (finished.load() < nw) {
    std::this_thread::sleep_for(std::chrono::milliseconds(2));
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (s > watchdog) {
      std::lock_guard<std::mutex> lk(g_mu);
      std::fprintf(stderr, "[emu] launch still running after %.0f s; workgroups in flight:\n", s);
      for (Block *b : g_running) dump_block(*b, "in flight");
      if (g_hang_hook) g_hang_hook();
      std::fflush(stderr);
      std::abort();
    }
  }
  for (auto &t : th) t.join();
}

}  // namespace

const dim3 &thread_idx() { return t_cur->tid; }
unsigned lane_id() { return t_cur->lane; }

uint64_t ballot(int line, bool pred) { return collective(C_BALLOT, line, pred ? 1 : 0, 0); }
uint32_t first_lane(int line, uint32_t v) { return (uint32_t)collective(C_FIRST, line, v, 0); }

int shfl(int line, int v, int src, int width) {
  const unsigned lane = t_cur->lane;
  const int s = (int)(lane & ~(unsigned)(width - 1)) + (src & (width - 1));
  return (int)(uint32_t)collective(C_SHFL, line, (uint32_t)v, s);
}

int shfl_up(int line, int v, unsigned delta, int width) {
  const int lane = (int)t_cur->lane;
  const int base = lane & ~(width - 1);
  const int s = lane - (int)delta;
  return (int)(uint32_t)collective(C_SHFL_UP, line, (uint32_t)v, s < base ? -1 : s);
}

void syncthreads(int line) {
  Block &b = *t_blk;
  Fiber *f = t_cur;
  f->line = line;
  b.bar_arrived++;
  const uint64_t gen = b.bar_gen;
  f->state = WAIT_BAR;
  try_complete_barrier(b);
  while (b.bar_gen == gen) yield_to_sched();
  f->state = RUNNABLE;
}

void spin_yield(int line) {
  Fiber *f = t_cur;
  f->line = line;
  f->state = SLEEP;
  yield_to_sched();
  f->state = RUNNABLE;
}

// ------------------------------------------------------------ streams/graphs
}  // namespace emu

struct emu_stream {
  bool capturing = false;
  std::vector<std::function<void()>> rec;
};
struct emu_graph {
  std::vector<std::function<void()>> ops;
};
struct emu_event {
  std::chrono::steady_clock::time_point t;
};

namespace emu {
namespace {
std::map<const void *, Invoker> &registry() {
  static std::map<const void *, Invoker> r;
  return r;
}
emu_stream g_null_stream;
emu_stream *S(hipStream_t s) { return s ? s : &g_null_stream; }
}  // namespace

void register_kernel(const void *f, Invoker inv) { registry()[f] = std::move(inv); }
void set_hang_hook(void (*f)()) { g_hang_hook = f; }

void submit(hipStream_t s, dim3 grid, dim3 block, size_t lds, std::function<void()> body) {
  auto run = [grid, block, lds, body]() { run_launch(Launch{grid, block, lds, body}); };
  if (S(s)->capturing) S(s)->rec.push_back(run);
  else run();
}

hipError_t launch_registered(const void *f, dim3 grid, dim3 block, void **args, size_t lds, hipStream_t s) {
  auto it = registry().find(f);
  if (it == registry().end()) {
    std::fprintf(stderr, "[emu] hipLaunchKernel of an unregistered kernel %p\n", f);
    return hipErrorInvalidValue;
  }
  submit(s, grid, block, lds, it->second(args));
  return hipSuccess;
}
}  // namespace emu

using emu::S;

hipError_t hipMalloc(void **p, size_t n) {
  const size_t b = std::max<size_t>(256, (n + 255) & ~(size_t)255);
  void *m = std::aligned_alloc(256, b);
  if (!m) return hipErrorOutOfMemory;
  std::memset(m, 0xA5, std::min<size_t>(b, 1ULL << 28));  // device memory starts undefined
  *p = m;
  return hipSuccess;
}
hipError_t hipFree(void *p) { std::free(p); return hipSuccess; }
hipError_t hipHostMalloc(void **p, size_t n, unsigned) { return hipMalloc(p, n); }
hipError_t hipHostFree(void *p) { std::free(p); return hipSuccess; }
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t n, hipMemcpyKind, hipStream_t s) {
  if (S(s)->capturing) S(s)->rec.push_back([dst, src, n] { std::memcpy(dst, src, n); });
  else if (n) std::memcpy(dst, src, n);
  return hipSuccess;
}
hipError_t hipMemcpy(void *dst, const void *src, size_t n, hipMemcpyKind k) { return hipMemcpyAsync(dst, src, n, k, nullptr); }
hipError_t hipMemGetInfo(size_t *fr, size_t *tot) {
  *fr = *tot = 16ULL << 30;
  return hipSuccess;
}
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int) {
  std::memset(p, 0, sizeof *p);
  std::snprintf(p->name, sizeof p->name, "cpu-emulation");
  std::snprintf(p->gcnArchName, sizeof p->gcnArchName, "emu");
  p->totalGlobalMem = 16ULL << 30;
  p->sharedMemPerBlock = 163840;
  const char *e = std::getenv("EMU_CUS");
  p->multiProcessorCount = e ? std::max(1, std::atoi(e)) : 8;
  p->l2CacheSize = 4 << 20;
  p->warpSize = 64;
  p->maxThreadsPerBlock = 1024;
  return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t a, int) {
  const char *e = std::getenv("EMU_CUS");
  *v = a == hipDeviceAttributeMultiprocessorCount ? (e ? std::max(1, std::atoi(e)) : 8) : 163840;
  return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipGetLastError() { return hipSuccess; }
const char *hipGetErrorString(hipError_t) { return "emulated HIP error"; }
hipError_t hipFuncSetAttribute(const void *, hipFuncAttribute, int) { return hipSuccess; }
hipError_t hipStreamCreate(hipStream_t *s) { *s = new emu_stream(); return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { return hipStreamCreate(s); }
hipError_t hipStreamDestroy(hipStream_t s) { delete s; return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  if (!s) return hipErrorInvalidValue;
  s->capturing = true;
  s->rec.clear();
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g) {
  if (!s || !s->capturing) return hipErrorInvalidValue;
  *g = new emu_graph{std::move(s->rec)};
  s->rec.clear();
  s->capturing = false;
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *, char *, size_t) {
  *e = new emu_graph{g->ops};
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s) {
  if (S(s)->capturing) {
    auto ops = e->ops;
    S(s)->rec.push_back([ops] { for (auto &o : ops) o(); });
  } else {
    for (auto &o : e->ops) o();
  }
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t *e) { *e = new emu_event(); return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
  if (S(s)->capturing) S(s)->rec.push_back([e] { e->t = std::chrono::steady_clock::now(); });
  else e->t = std::chrono::steady_clock::now();
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) {
  *ms = (float)std::chrono::duration<double, std::milli>(b->t - a->t).count();
  return hipSuccess;
}
