// C-ABI entry points (include/jaqsi_amd.h): plan objects, workspaces and the
// forward / adjoint drivers over the tile-pass kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "jaqsi_amd.h"
#include "jit.h"
#include "plan_internal.h"
#include "prep.h"
#include "pulse.h"
#include "program.h"
#include "sv_kernels.h"
#include "wide.h"
#include "evolve.h"
#include "noalloc.h"

namespace {
thread_local std::string g_err;

// Optional kernel timing: HIP events around each run of consecutive
// tile-pass launches of one kind (the forward passes of a call, its backward
// passes).  An event pair per run rather than per launch: event records
// between kernels stretch a 24-qubit gradient step by about 5%.
struct KTimer {
  struct Run {
    hipEvent_t e0, e1;
    int launches;
  };
  bool on = false;
  std::vector<Run> runs[3];
  Run open[3] = {};          // run in progress per kind (e0 recorded)
  bool is_open[3] = {false, false, false};
  std::vector<hipEvent_t> pool;  // events reused across reads
  hipEvent_t take() {
    hipEvent_t e = nullptr;
    if (!pool.empty()) {
      e = pool.back();
      pool.pop_back();
    } else {
      (void)hipEventCreate(&e);
    }
    return e;
  }
};
KTimer g_timer;

// A plan-specialised pass kernel (jit.cpp), or none.
struct JitFn {
  hipFunction_t f = nullptr;
  int nt = 0;            // threads per workgroup
  int64_t pool_off = 0;  // SvArgs::pool_off of the pass
  int resident = 0;      // workgroups of the device resident at once (persistent kernels)
  const jq::WideGeom* wide = nullptr;  // a wide-block pass: run the wide-block kernels
};

hipError_t launch_pass(const JitFn& jf, int variant, int dtype, int T, const jq::SvArgs& a, int ntiles, int batch,
                       size_t smem, hipStream_t st) {
  if (jf.wide != nullptr) return jq::launch_wide(variant & 1, dtype, a, *jf.wide, batch, st);
  if (jf.f == nullptr) return jq::launch_sv(variant, dtype, T, a, ntiles, batch, smem, st);
  jq::SvArgs arg = a;
  arg.pool_off = jf.pool_off;
  void* params[] = {(void*)&arg};
  if (jf.resident > 0) {
    // persistent kernels: one wave of resident workgroups over the flat
    // (circuit, tile) items of all rows
    arg.rows = batch;
    const int64_t items = (int64_t)batch * ntiles;
    const int gx = (int)std::min<int64_t>(items, jf.resident);
    return hipModuleLaunchKernel(jf.f, gx, 1, 1, jf.nt, 1, 1, (unsigned)smem, st, params, nullptr);
  }
  return hipModuleLaunchKernel(jf.f, ntiles, batch, 1, jf.nt, 1, 1, (unsigned)smem, st, params, nullptr);
}

// which: 0 forward pass, 1 backward pass; `full` selects the kernel variant
// that also carries 4x4 register blocks and the dense 3-wire path.  With a
// specialised kernel `jf` the pass runs that kernel instead.
// `first` / `last`: this launch opens / closes a run of launches of kind
// `which` on stream `st` (timing only).
hipError_t timed_launch(int which, bool full, int dtype, int T, const jq::SvArgs& a, int ntiles, int batch,
                        size_t smem, hipStream_t st, bool wide, JitFn jf, bool first, bool last) {
  const int variant = which + (full ? 2 : 0) + (which == 1 && wide ? 4 : 0);
  if (!g_timer.on) return launch_pass(jf, variant, dtype, T, a, ntiles, batch, smem, st);
  KTimer::Run& r = g_timer.open[which];
  if (first) {
    if (g_timer.is_open[which]) g_timer.pool.push_back(r.e0);  // a run left open by a failed launch
    r.e0 = g_timer.take();
    r.launches = 0;
    g_timer.is_open[which] = true;
    hipEventRecord(r.e0, st);
  }
  hipError_t e = launch_pass(jf, variant, dtype, T, a, ntiles, batch, smem, st);
  if (g_timer.is_open[which]) ++r.launches;
  if (last && g_timer.is_open[which]) {
    r.e1 = g_timer.take();
    hipEventRecord(r.e1, st);
    g_timer.runs[which].push_back(r);
    g_timer.is_open[which] = false;
  }
  return e;
}

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                            \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess) return fail(JQ_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  // Grow to `want` bytes.  A captured HIP graph keeps the workspace
  // addresses it saw, so a workspace never grows while `st` is capturing:
  // the call fails instead (one eager call at the largest size first).
  hipError_t ensure(size_t want, hipStream_t st) {
    if (want <= bytes) return hipSuccess;
    if (st != nullptr) {
      hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
      if (hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)
        return hipErrorStreamCaptureUnsupported;
    }
    if (p) hipFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) bytes = want;
    return e;
  }
  void release() {
    if (p) hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

template <typename T>
hipError_t upload(DevBuf& b, const std::vector<T>& v) {
  size_t n = std::max<size_t>(1, v.size()) * sizeof(T);
  hipError_t e = b.ensure(n, nullptr);
  if (e != hipSuccess) return e;
  if (!v.empty()) return hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
  return hipSuccess;
}
}  // namespace

// Plan-specialised kernels: compiled on the host (possibly on a worker
// thread: hipRTC only, no HIP runtime call), loaded as a module on first use
// by the launching thread.
struct JitState {
  std::thread worker;
  std::atomic<int> phase{JQ_JIT_NONE};
  int mode = 0;  // 0 interpreter, 1 specialised kernels once compiled
  jq::JitPlan plan;                       // distinct pass bodies and the pass -> kernel map
  std::vector<std::pair<int, std::string>> kref;  // per kernel id: (cache code object, kernel name)
  std::string log;
  std::vector<hipFunction_t> fn;          // per kernel id
  std::vector<int> resident;              // per kernel id: resident workgroups on the device
  bool persist[2] = {false, true};        // persistent tile loops (forward, backward)
  bool dma = false;                       // forward streams tiles into LDS (two tiles of LDS)
  bool declined = false;                  // passes too long to specialise: the interpreter runs the plan
  bool required = false;                  // no prebuilt kernel serves the plan (14-bit tiles)
  int nt[2] = {0, 0};
};

// Buffers a call works in, and the token of the checkpoints the last
// jq_sv_forward_ckpt on this stream left: valid while no other call has run
// on the plan on this stream since (every call that touches the workspaces
// advances `generation`).
struct Workspace {
  DevBuf pool, partials, bars, ev_part, lam, ckpt, obs, fstate, groups;
  int64_t generation = 0;
  int64_t ckpt_token = -1;
  int ckpt_batch = 0;
};

struct jq_plan {
  jq::HostProgram hp;
  int dtype = 0;
  int n_params = 0;
  int64_t n_matrix_elems = 0;
  int checkpoints = 1;
  DevBuf prog, prog_b, block_table, member_list, member_table, member_values, slot_map;
  int n_slots = 0;
  std::vector<int32_t> slot_words;
  std::vector<char> pass_full;  // pass holds a 4x4 register block or a dense 3-wire block
  bool wide_bwd = false;        // backward program built for 16 amplitudes x 512 threads (complex64)
  int variant_flags = 0;        // JQ_TUNE_* kernel variants
  std::vector<int32_t> words_bj;  // backward program of the plan-specialised kernels
  DevBuf prog_bj;
  int regions_b = 2, regions_bj = 2;  // LDS state regions of the interpreter's / specialised backward
  jq::TailFold fold;                  // trailing blocks measured through the observables (fold_tail)
  std::vector<int32_t> fold_words;    // jq_plan_tables(6)
  bool uploaded = false;
  int class_mask = 0;  // block classes of the prep / chain kernels (jq::block_class)
  int max_members2 = 0;  // most members of a 2x2 (class 2) block: the chain's threads per block
  std::vector<int32_t> wg_blocks;  // blocks of the workgroup prep / chain classes
  DevBuf wg_list;
  std::vector<char> pass_wide;               // pass q is one wide block (program.h PF_WIDE)
  std::vector<jq::WideGeom> wide_geom;       // its launch geometry
  bool any_wide = false;
  JitState jit;
  // Workspaces, one set per stream the plan runs on, so calls on different
  // streams never share buffers (a plan is re-entrant per stream; calls on
  // one stream are ordered by the stream itself).  Created on a stream's
  // first call, freed with the plan.
  std::mutex ws_mu;
  std::map<hipStream_t, std::unique_ptr<Workspace>> ws_map;
};

Workspace& WS(jq_plan* p, hipStream_t st) {
  std::lock_guard<std::mutex> lock(p->ws_mu);
  std::unique_ptr<Workspace>& w = p->ws_map[st];
  if (!w) w.reset(new Workspace());
  return *w;
}

namespace {

// LDS bytes of a pass workgroup: `regions` state regions (forward 1;
// backward the program's lds_regions), the observable table, 64 expectation
// reals and `extra_reals` cotangent slot reals
size_t smem_for(jq_plan* p, int extra_reals, int regions = 1);
int max_slots(jq_plan* p);

// Process-wide cache of compiled pass kernels.  A kernel serves every plan
// whose pass body (and register tile, dtype, loop variant) is the same, so
// plans of one ansatz at different depths compile their shared passes once.
// Compiles are serialised under the lock (the run-time compiler serialises
// concurrent compiles anyway); code objects and modules live until the
// process ends.
struct JitCache {
  std::mutex mu;
  std::unordered_map<std::string, std::pair<int, std::string>> kernels;  // key -> (code object, name)
  std::vector<std::vector<char>> code;
  std::map<std::pair<int, int>, hipModule_t> mods;  // (device, code object) -> module
  int next = 0;                                     // kernel name counter
};
JitCache& jit_cache() {
  static JitCache c;
  return c;
}

// Compile the plan's specialised kernels that the cache lacks (host only;
// any thread).
// Longest pass (in gates) that is specialised: straight-line code grows
// with the gate count and so does the compile time; longer passes (deep
// circuits on few wires) stay with the interpreter.
constexpr int kJitMaxPassGates = 256;
// Blocks per pass of a 14-bit-tile plan, so that its straight-line code
// stays within reach of the instruction cache (JQ_T14_PASS_BLOCKS overrides).
constexpr int kT14PassBlocks = 160;

void jit_build(jq_plan* p) {
  {
    const int* W = p->hp.words.data();
    int longest = 0;
    for (int q = 0; q < p->hp.n_passes; ++q) {
      const int* pass = W + W[jq::HDR_WORDS + q];
      int ng = 0;
      for (int s = 0; s < pass[0]; ++s) ng += W[pass[4 + p->hp.T + s]];
      longest = std::max(longest, ng);
    }
    if (longest > kJitMaxPassGates) {
      p->jit.log = "a pass of " + std::to_string(longest) + " gates exceeds the specialisation limit (" +
                   std::to_string(kJitMaxPassGates) + "); the interpreter runs this plan";
      p->jit.declined = true;
      p->jit.phase.store(JQ_JIT_FAILED, std::memory_order_release);
      return;
    }
  }
  jq::JitSpec spec;
  spec.words_f = &p->hp.words;
  spec.words_b = &p->words_bj;
  spec.dtype = p->dtype;
  // persistent backward (with its next-tile prefetch registers) only for
  // 512-thread or larger workgroups
  spec.persist_bwd = !(p->variant_flags & JQ_TUNE_NO_PERSIST_BWD) && !p->jit.required &&
                     (p->words_bj[2] - p->words_bj[3]) >= 9;
  spec.dma = !(p->variant_flags & JQ_TUNE_NO_DMA_FWD);
  p->jit.dma = spec.dma && p->dtype == 0;
  p->jit.persist[0] = false;
  p->jit.persist[1] = spec.persist_bwd;
  jq::jit_plan(spec, p->jit.plan);
  const int nk = (int)p->jit.plan.kernels.size();
  JitCache& C = jit_cache();
  std::vector<std::string> keys(nk);
  std::vector<int> todo;
  std::vector<std::string> names;
  p->jit.kref.assign(nk, {-1, std::string()});
  {
    std::lock_guard<std::mutex> lock(C.mu);
    for (int id = 0; id < nk; ++id) {
      keys[id] = jq::jit_key(spec, p->jit.plan, id);
      auto it = C.kernels.find(keys[id]);
      if (it != C.kernels.end()) {
        p->jit.kref[id] = it->second;
      } else {
        todo.push_back(id);
        names.push_back(jq::jit_kernel_name(p->jit.plan.kernels[id].dir, C.next++));
      }
    }
  }
  bool ok = true;
  std::string log;
  if (!todo.empty()) {
    // compiled outside the cache lock (other plans may look kernels up
    // meanwhile); a kernel two plans compile at once is compiled twice, and
    // each keeps its own copy
    const std::string src = jq::jit_source(spec, p->jit.plan, todo, names);
    if (const char* dump = std::getenv("JQ_JIT_DUMP")) {  // debugging aid: keep the generated source
      if (FILE* fh = std::fopen((std::string(dump) + "." + names[0] + ".hip").c_str(), "w")) {
        std::fwrite(src.data(), 1, src.size(), fh);
        std::fclose(fh);
      }
    }
    std::vector<char> code;
    ok = jq::jit_compile(src, code, log);
    if (ok) {
      std::lock_guard<std::mutex> lock(C.mu);
      const int ci = (int)C.code.size();
      C.code.push_back(std::move(code));
      for (size_t j = 0; j < todo.size(); ++j) {
        C.kernels.emplace(keys[todo[j]], std::make_pair(ci, names[j]));
        p->jit.kref[todo[j]] = {ci, names[j]};
      }
    }
  }
  p->jit.log = log;
  p->jit.phase.store(ok ? JQ_JIT_COMPILED : JQ_JIT_FAILED, std::memory_order_release);
}

void jit_start(jq_plan* p) {
  int expect = JQ_JIT_NONE;
  if (!p->jit.phase.compare_exchange_strong(expect, JQ_JIT_COMPILING)) return;
  p->jit.worker = std::thread(jit_build, p);
}

// Load the compiled module (launching thread).  False on failure.
bool jit_load(jq_plan* p) {
  if (p->jit.phase.load(std::memory_order_acquire) != JQ_JIT_COMPILED) return false;
  if (p->jit.worker.joinable()) p->jit.worker.join();
  p->jit.fn.assign(p->jit.plan.kernels.size(), nullptr);
  hipError_t e = hipSuccess;
  {
    JitCache& C = jit_cache();
    std::lock_guard<std::mutex> lock(C.mu);
    int dev = 0;
    e = hipGetDevice(&dev);
    for (size_t id = 0; id < p->jit.kref.size() && e == hipSuccess; ++id) {
      const int ci = p->jit.kref[id].first;
      auto it = C.mods.find({dev, ci});
      if (it == C.mods.end()) {
        hipModule_t m = nullptr;
        e = hipModuleLoadData(&m, C.code[ci].data());
        if (e != hipSuccess) break;
        it = C.mods.emplace(std::make_pair(dev, ci), m).first;
      }
      e = hipModuleGetFunction(&p->jit.fn[id], it->second, p->jit.kref[id].second.c_str());
    }
  }
  if (e != hipSuccess) {
    p->jit.log = std::string("module load: ") + hipGetErrorString(e);
    p->jit.phase.store(JQ_JIT_FAILED);
    return false;
  }
  p->jit.nt[0] = 1 << (p->hp.words[2] - p->hp.words[3]);
  p->jit.nt[1] = 1 << (p->words_bj[2] - p->words_bj[3]);
  int dev = 0, cus = 1;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  p->jit.resident.assign(p->jit.fn.size(), cus);
  for (size_t id = 0; id < p->jit.fn.size(); ++id) {
    const int dir = p->jit.plan.kernels[id].dir;
    const size_t sm = dir == 0 ? smem_for(p, 0) : smem_for(p, 2 * max_slots(p), p->regions_bj);
    int nb = 0;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nb, p->jit.fn[id], p->jit.nt[dir], sm) == hipSuccess &&
        nb > 0)
      p->jit.resident[id] = nb * cus;
  }
  // cap on the resident workgroups of a persistent launch (tests use it to
  // make every workgroup walk several tiles on small vectors)
  if (const char* e = std::getenv("JQ_JIT_MAX_BLOCKS")) {
    const int cap = std::atoi(e);
    if (cap > 0)
      for (int& r : p->jit.resident) r = std::min(r, cap);
  }
  p->jit.phase.store(JQ_JIT_LOADED);
  return true;
}

// Side streams of ragged-batch calls, per device (created once, default
// priority, non-blocking), with the events that fork and join them.
struct SideStreams {
  static constexpr int kStreams = 8;
  hipStream_t s[kStreams];
  hipEvent_t fork;
  hipEvent_t join[kStreams];
  std::mutex mu;  // held while a call records on the shared events
};

hipError_t side_streams(SideStreams** out) {
  static std::mutex mu;
  static std::map<int, SideStreams*> per_dev;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lock(mu);
  auto it = per_dev.find(dev);
  if (it == per_dev.end()) {
    SideStreams* ss = new SideStreams();
    for (int k = 0; k < SideStreams::kStreams && e == hipSuccess; ++k)
      e = hipStreamCreateWithFlags(&ss->s[k], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ss->fork, hipEventDisableTiming);
    for (int k = 0; k < SideStreams::kStreams && e == hipSuccess; ++k)
      e = hipEventCreateWithFlags(&ss->join[k], hipEventDisableTiming);
    if (e != hipSuccess) return e;
    it = per_dev.emplace(dev, ss).first;
  }
  *out = it->second;
  return hipSuccess;
}

// Whether this call's passes run the specialised kernels.
bool jit_ready(jq_plan* p) {
  if (p->jit.mode == JQ_JIT_OFF) return false;
  if (p->jit.required) {  // no prebuilt kernel can stand in: wait for the compile
    int expect = JQ_JIT_NONE;
    if (p->jit.phase.compare_exchange_strong(expect, JQ_JIT_COMPILING)) jit_build(p);
    if (p->jit.worker.joinable()) p->jit.worker.join();
  }
  const int ph = p->jit.phase.load(std::memory_order_acquire);
  if (ph == JQ_JIT_LOADED) return true;
  if (ph == JQ_JIT_COMPILED) return jit_load(p);
  return false;
}

JitFn jit_fn(jq_plan* p, bool on, int dir, int q) {
  JitFn f;
  if (p->pass_wide[q]) {
    f.wide = &p->wide_geom[q];
    return f;
  }
  if (on) {
    f.f = p->jit.fn[p->jit.plan.kernel_of[dir][q]];
    f.nt = p->jit.nt[dir];
    f.pool_off = p->jit.plan.pool_off[dir][q];
    f.resident = p->jit.persist[dir] ? p->jit.resident[p->jit.plan.kernel_of[dir][q]] : 0;
  }
  return f;
}

}  // namespace

extern "C" {

const char* jq_last_error(void) { return g_err.c_str(); }

const char* jq_version(void) { return "jaqsi_amd 0.1 (gfx950)"; }

static int elem_bytes(int dtype) { return dtype == 0 ? 8 : 16; }
// state regions a backward workgroup keeps in LDS (the kernels' lds_regions)
static size_t lds_regions(int dtype, int T, int RB) {
  return 2 * ((size_t)elem_bytes(dtype) << T) <= 128 * 1024 && T - RB >= 9 ? 2 : 1;
}

int jq_plan_create(const jq_plan_desc* desc, jq_plan** out) {
  if (!desc || !out) return fail(JQ_ERR_ARG, "null argument");
  if (desc->dtype != 0 && desc->dtype != 1) return fail(JQ_ERR_ARG, "dtype must be 0 (complex64) or 1 (complex128)");
  if (desc->n_ops < 0 || (desc->n_ops > 0 && !desc->ops)) return fail(JQ_ERR_ARG, "bad op list");
  int T = desc->tile_bits;
  // amplitudes per thread: forward 16 (2^4), backward 8 (2^3, psi and lambda)
  // for both dtypes; the workgroup covers one 64 KiB tile (sv_f32 / sv_f64).
  // complex64 plans with 4x4 / dense 3-wire blocks are rebuilt below with a
  // 16-amplitude backward.
  const int RB = 4;
  const int RBb = 3;
  if (T == 0) T = desc->dtype == 0 ? 13 : 12;
  const bool okT = desc->dtype == 0 ? (T == 13 || T == 14) : (T == 12);
  if (!okT) return fail(JQ_ERR_ARG, "tile_bits must be 13 or 14 (complex64) or 12 (complex128)");
  // A 14-bit tile (complex64) is 2^14 amplitudes, 16 per thread of a
  // 1024-thread workgroup: a 14-qubit circuit is one tile, so each pass
  // reads and writes its vector once however many of its wires the gates
  // touch.  Only plan-specialised kernels run such plans (non-persistent
  // backward: 1024 threads leave 128 VGPRs, no room for a prefetch set).
  const bool t14 = T == 14;
  if (t14 && desc->n_wires < 14) return fail(JQ_ERR_ARG, "tile_bits 14 needs at least 14 wires");
  // LDS per backward workgroup: the state region(s) (psi and lambda side by
  // side, or one region they take in turn for 14-bit tiles), the observable
  // table, 64 expectation reals and two buffers of the pass's cotangent
  // slots (the persistent backward alternates them from tile to tile).
  const size_t lds_state = (size_t)elem_bytes(desc->dtype) << T;
  // (the slot budget holds for every backward program of the plan: 13-bit
  // tiles keep the interpreter's two regions)
  const size_t regions = t14 ? 1 : 2;
  const int slot_budget = ((int)((160 * 1024 - regions * lds_state - sizeof(jq::ObsTable)) / 8) - 64) / 2;
  // a pass's specialised code grows with its gates: 14-bit passes, which
  // could otherwise take a whole circuit, close at a block budget
  int max_blocks = 0;
  if (t14) {
    max_blocks = kT14PassBlocks;
    if (const char* e = std::getenv("JQ_T14_PASS_BLOCKS")) max_blocks = std::max(1, std::atoi(e));
    max_blocks = std::min(max_blocks, kJitMaxPassGates);
  }
  if (desc->fold_tail != 0 && desc->fold_tail != 1) return fail(JQ_ERR_ARG, "fold_tail must be 0 or 1");
  jq_plan* p = new jq_plan();
  // the passes run the ops before the folded tail
  std::vector<jq_op> kept_ops;
  jq_plan_desc dd = *desc;
  if (desc->fold_tail) {
    jq::fold_tail(*desc, kept_ops, p->fold);
    if (p->fold.n_folded > 0) {
      dd.ops = kept_ops.data();
      dd.n_ops = (int32_t)kept_ops.size();
    }
  }
  {
    std::vector<int32_t>& fw = p->fold_words;
    fw.push_back(p->fold.n_folded);
    fw.push_back((int32_t)(uint32_t)p->fold.flip);
    fw.push_back((int32_t)(uint32_t)(p->fold.flip >> 32));
    for (uint64_t r : p->fold.rows) {
      fw.push_back((int32_t)(uint32_t)r);
      fw.push_back((int32_t)(uint32_t)(r >> 32));
    }
  }
  desc = &dd;
  try {
    jq::build_program(*desc, T, RB, RBb, p->hp, slot_budget, max_blocks);
  } catch (const jq::PlanError& e) {
    delete p;
    return fail(e.code, e.what());
  } catch (const std::exception& e) {
    delete p;
    return fail(JQ_ERR_ARG, e.what());
  }
  if (desc->variant_flags & ~(JQ_TUNE_NO_PERSIST_BWD | JQ_TUNE_NO_DMA_FWD | JQ_TUNE_NARROW_BWD | JQ_TUNE_BWD16)) {
    delete p;
    return fail(JQ_ERR_ARG, "unknown variant_flags bits");
  }
  p->variant_flags = desc->variant_flags;
  p->dtype = desc->dtype;
  p->n_params = desc->n_params;
  p->n_matrix_elems = desc->n_matrix_elems;
  p->checkpoints = desc->checkpoints;
  // (a wide pass reads its input state from the checkpoints in the backward)
  // per-pass slot capacity in LDS (the planner keeps every pass within it)
  const int* W = p->hp.words.data();
  p->pass_wide.assign(p->hp.n_passes, 0);
  p->wide_geom.assign(p->hp.n_passes, jq::WideGeom());
  for (int q = 0; q < p->hp.n_passes; ++q) {
    const int* pass = W + W[jq::HDR_WORDS + q];
    if (pass[1] & jq::PF_WIDE) {  // one substep, one gate record
      const int* sub = W + pass[4 + T];
      p->pass_wide[q] = 1;
      p->any_wide = true;
      jq::wide_geom(sub + 1 + T, desc->n_wires, p->hp.ntiles, pass[3], desc->dtype, p->wide_geom[q]);
      continue;
    }
    if (regions * lds_state + sizeof(jq::ObsTable) + (64 + 2 * (size_t)pass[2]) * 8 > 160 * 1024) {
      delete p;
      return fail(JQ_ERR_UNSUPPORTED, "too many gradient slots in one pass for LDS");
    }
  }
  if (p->any_wide) p->checkpoints = 1;
  p->pass_full.assign(p->hp.n_passes, 0);
  bool any_full = false;
  for (int q = 0; q < p->hp.n_passes; ++q) {
    const int* pass = W + W[jq::HDR_WORDS + q];
    const int* soff = pass + 4 + T;
    for (int s = 0; s < pass[0]; ++s) {
      const int* sub = W + soff[s];
      const int* g = sub + 1 + T;
      for (int k = 0; k < sub[0]; ++k, g += jq::GW)
        if (g[jq::G_KIND] == jq::AK_MC2 || g[jq::G_KIND] == jq::AK_DENSEK) p->pass_full[q] = 1;
    }
    any_full = any_full || p->pass_full[q];
  }
  // Default kernel choice: plan-specialised kernels for plans of
  // JQ_JIT_MIN_WIRES or more wires (JQ_JIT=0 / 1 overrides).
  bool jit_on = desc->n_wires >= JQ_JIT_MIN_WIRES;
  if (const char* env = std::getenv("JQ_JIT")) {
    if (env[0] == '0') jit_on = false;
    if (env[0] == '1') jit_on = true;
  }
  // complex64: a second backward program for 16 amplitudes x 512 threads
  // (same passes: the tile choice does not depend on the register tile).
  // The plan-specialised backward always runs it (its straight-line code
  // keeps 16 amplitudes of psi and lambda in registers without spilling, and
  // fewer LDS re-layouts beat the lower occupancy); the interpreter runs it
  // for plans with 4x4 / dense 3-wire blocks, which would spill under the
  // 1024-thread register cap.
  p->words_bj = p->hp.words_b;
  if (desc->dtype == 0) {
    // The specialised backward's register tile: 32 amplitudes per thread
    // (13-bit tiles: 256 threads, two workgroups per CU, each exchanging psi
    // and lambda in turn through one LDS region; 14-bit tiles: 512 threads),
    // or with JQ_TUNE_BWD16 16 amplitudes (13-bit tiles: 512 threads, one
    // persistent workgroup per CU with next-tile prefetch; 14-bit tiles: 1024
    // threads).  (The prebuilt kernels keep their own 16-amplitude program
    // for 13-bit plans with 4x4 / dense 3-wire blocks.)
    const bool bwd16 = (p->variant_flags & JQ_TUNE_BWD16) != 0;
    const int rbw = bwd16 ? 4 : 5;
    jq::HostProgram wide, wide5;
    try {
      jq::build_program(*desc, T, RB, 4, wide, slot_budget, max_blocks);
      if (rbw == 5) jq::build_program(*desc, T, RB, 5, wide5, slot_budget, max_blocks);
    } catch (const std::exception& e) {
      delete p;
      return fail(JQ_ERR_ARG, e.what());
    }
    if (wide.n_passes != p->hp.n_passes || (rbw == 5 && wide5.n_passes != p->hp.n_passes)) {
      delete p;
      return fail(JQ_ERR_ARG, "internal: a backward program changed the pass structure");
    }
    // (JQ_TUNE_NARROW_BWD keeps the 8-amplitude program for the specialised
    // backward too, unless 4x4 / dense 3-wire blocks or 14-bit tiles rule it out)
    if ((p->variant_flags & JQ_TUNE_NARROW_BWD) && !any_full && !t14)
      p->words_bj = p->hp.words_b;
    else
      p->words_bj = rbw == 5 ? wide5.words_b : wide.words_b;
    if (any_full || t14) {
      // (14-bit plans never run the prebuilt kernels: their program is the
      // specialised one, which is what the CPU program interpreter then checks)
      p->hp.words_b = t14 ? p->words_bj : wide.words_b;
      p->wide_bwd = true;
    }
  }
  p->regions_b = (int)lds_regions(desc->dtype, T, p->hp.words_b[3]);
  p->regions_bj = (int)lds_regions(desc->dtype, T, p->words_bj[3]);
  if (t14) {
    bool dense3 = false;
    for (size_t r = 0; r + jq::BW <= p->hp.block_table.size(); r += jq::BW)
      dense3 = dense3 || p->hp.block_table[r + jq::B_KIND] == jq::AK_DENSEK;
    if (p->any_wide || dense3) {
      delete p;
      return fail(JQ_ERR_UNSUPPORTED, "tile_bits 14: wide-pass and dense 3-wire blocks need 13-bit tiles");
    }
    // compiled at its first use or jq_plan_jit call, not on a worker thread
    // now: groups of a ragged batch then compile in turn and reuse the
    // kernels of the passes they share
    p->jit.required = true;
    jit_on = false;
    p->jit.mode = JQ_JIT_ASYNC;
  }
  std::vector<int32_t> sm;
  std::vector<jq::SlotMap> smv = p->hp.slot_map;
  std::sort(smv.begin(), smv.end(), [](const jq::SlotMap& a, const jq::SlotMap& b) { return a.bar < b.bar; });
  for (auto& s : smv) {
    const int* pass = W + W[jq::HDR_WORDS + s.pass];
    sm.insert(sm.end(), {s.base, pass[2], s.local, s.bar, s.size});
  }
  p->n_slots = (int)smv.size();
  p->slot_words = sm;
  for (size_t r = 0; r + jq::BW <= p->hp.block_table.size(); r += jq::BW) {
    const int cls = jq::block_class(p->hp.block_table[r + jq::B_DIAGF] != 0, p->hp.block_table[r + jq::B_D]);
    if (cls == 2)
      p->max_members2 = std::max(p->max_members2, p->hp.block_table[r + jq::B_MEND] - p->hp.block_table[r + jq::B_MBEG]);
    if (cls >= jq::WG_FIRST_CLASS)
      p->wg_blocks.push_back((int32_t)(r / jq::BW));
    else
      p->class_mask |= 1 << cls;
  }
  int ndev = 0;
  if (jit_on && hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0) {  // nothing to run them on otherwise
    p->jit.mode = JQ_JIT_ASYNC;
    jit_start(p);
  }
  *out = p;
  return JQ_OK;
}

int jq_plan_jit(jq_plan* p, int32_t mode, int32_t* phase) {
  if (!p) return fail(JQ_ERR_ARG, "null plan");
  if (mode < JQ_JIT_QUERY || mode > JQ_JIT_SYNC) return fail(JQ_ERR_ARG, "unknown specialisation mode");
  int rc = JQ_OK;
  if (mode == JQ_JIT_OFF) {
    p->jit.mode = JQ_JIT_OFF;
  } else if (mode == JQ_JIT_ASYNC) {
    p->jit.mode = JQ_JIT_ASYNC;
    jit_start(p);
  } else if (mode == JQ_JIT_SYNC) {
    p->jit.mode = JQ_JIT_ASYNC;
    int expect = JQ_JIT_NONE;
    if (p->jit.phase.compare_exchange_strong(expect, JQ_JIT_COMPILING)) jit_build(p);
    if (p->jit.worker.joinable()) p->jit.worker.join();
    const int ph = p->jit.phase.load(std::memory_order_acquire);
    if (ph == JQ_JIT_COMPILED) jit_load(p);
    if (p->jit.phase.load() != JQ_JIT_LOADED && !p->jit.declined)
      rc = fail(JQ_ERR_DEVICE, "plan-specialised kernels: " + p->jit.log);
  }
  if (phase) *phase = p->jit.phase.load(std::memory_order_acquire);
  return rc;
}

int jq_plan_destroy(jq_plan* p) {
  if (!p) return JQ_OK;
  if (p->jit.worker.joinable()) p->jit.worker.join();
  for (DevBuf* b : {&p->prog, &p->prog_b, &p->prog_bj, &p->block_table, &p->member_list, &p->member_table, &p->member_values, &p->slot_map,
                    &p->wg_list})
    b->release();
  for (auto& kv : p->ws_map) {
    Workspace& w = *kv.second;
    for (DevBuf* b : {&w.pool, &w.partials, &w.bars, &w.ev_part, &w.lam, &w.ckpt, &w.obs, &w.fstate, &w.groups})
      b->release();
  }
  delete p;
  return JQ_OK;
}

int jq_plan_info(const jq_plan* p, int32_t* n_passes, int32_t* n_substeps, int32_t* n_blocks, int32_t* n_grad_reals,
                 int32_t* tile_bits) {
  if (!p) return fail(JQ_ERR_ARG, "null plan");
  if (n_passes) *n_passes = p->hp.n_passes;
  if (n_substeps) *n_substeps = p->hp.n_substeps;
  if (n_blocks) *n_blocks = p->hp.n_dev_blocks;
  if (n_grad_reals) *n_grad_reals = p->hp.bars * 2;
  if (tile_bits) *tile_bits = p->hp.T;
  return JQ_OK;
}

int jq_plan_program(const jq_plan* p, int32_t* out, int64_t capacity, int64_t* size) {
  if (!p) return fail(JQ_ERR_ARG, "null plan");
  if (size) *size = (int64_t)p->hp.words.size();
  if (out) {
    if (capacity < (int64_t)p->hp.words.size()) return fail(JQ_ERR_ARG, "capacity too small");
    std::memcpy(out, p->hp.words.data(), p->hp.words.size() * sizeof(int32_t));
  }
  return JQ_OK;
}

int jq_plan_tables(const jq_plan* p, int32_t which, int32_t* out, int64_t capacity, int64_t* size) {
  if (!p) return fail(JQ_ERR_ARG, "null plan");
  const std::vector<int32_t>* v = nullptr;
  switch (which) {
    case 0: v = &p->hp.block_table; break;
    case 1: v = &p->hp.member_list; break;
    case 2: v = &p->hp.member_table; break;
    case 3: v = &p->slot_words; break;
    case 4: v = &p->hp.words_b; break;
    case 5: v = &p->words_bj; break;
    case 6: v = &p->fold_words; break;
    default: return fail(JQ_ERR_ARG, "unknown table");
  }
  if (size) *size = (int64_t)v->size();
  if (out) {
    if (capacity < (int64_t)v->size()) return fail(JQ_ERR_ARG, "capacity too small");
    std::memcpy(out, v->data(), v->size() * sizeof(int32_t));
  }
  return JQ_OK;
}

}  // extern "C"

namespace {

int ensure_uploaded(jq_plan* p) {
  if (p->uploaded) return JQ_OK;
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = upload(p->prog, p->hp.words);
  if (e == hipSuccess) e = upload(p->prog_b, p->hp.words_b);
  if (e == hipSuccess) e = upload(p->prog_bj, p->words_bj);
  if (e == hipSuccess) e = upload(p->block_table, p->hp.block_table);
  if (e == hipSuccess) e = upload(p->member_list, p->hp.member_list);
  if (e == hipSuccess) e = upload(p->member_table, p->hp.member_table);
  if (e == hipSuccess) e = upload(p->member_values, p->hp.member_values);
  if (e == hipSuccess) e = upload(p->slot_map, p->slot_words);
  if (e == hipSuccess) e = upload(p->wg_list, p->wg_blocks);
  if (e != hipSuccess) return fail(JQ_ERR_DEVICE, std::string("plan upload: ") + hipGetErrorString(e));
  p->uploaded = true;
  return JQ_OK;
}

int fill_obs(jq_plan* p, jq::SvArgs& a, int n_wires, int obs_kind, int n_obs, const int32_t* obs_wires,
             const double* obs_d, hipStream_t st, void* table = nullptr) {
  a.obs_kind = obs_kind;
  a.n_obs = 0;
  a.obs = nullptr;
  if (obs_kind == JQ_OBS_NONE) return JQ_OK;
  if (obs_kind != JQ_OBS_SV_DIAG && obs_kind != JQ_OBS_DM_DIAG && obs_kind != JQ_OBS_SV_PARITY)
    return fail(JQ_ERR_ARG, "unknown observable kind");
  if (n_obs < 0 || n_obs > 64) return fail(JQ_ERR_ARG, "between 0 and 64 observables are supported");
  if (n_obs > 0 && (!obs_wires || !obs_d)) return fail(JQ_ERR_ARG, "null observable arrays");
  const bool folded = p->fold.n_folded > 0;
  if (folded && obs_kind == JQ_OBS_DM_DIAG)
    return fail(JQ_ERR_UNSUPPORTED, "a plan with a folded tail measures statevector observables only");
  const int span = obs_kind == JQ_OBS_DM_DIAG ? n_wires / 2 : n_wires;
  jq::ObsTable t;
  std::memset(&t, 0, sizeof(t));
  for (int o = 0; o < n_obs; ++o) {
    double d0 = obs_d[2 * o], d1 = obs_d[2 * o + 1];
    if (obs_kind == JQ_OBS_SV_PARITY || folded) {
      uint64_t wm;
      if (obs_kind == JQ_OBS_SV_PARITY) {
        wm = (uint64_t)(uint32_t)obs_wires[2 * o] | ((uint64_t)(uint32_t)obs_wires[2 * o + 1] << 32);
        if (span < 64 && (wm >> span) != 0) return fail(JQ_ERR_ARG, "observable wire out of range");
      } else {
        if (obs_wires[o] < 0 || obs_wires[o] >= span) return fail(JQ_ERR_ARG, "observable wire out of range");
        wm = 1ull << obs_wires[o];
      }
      if (folded) {
        // the parity of the final bits in wm is the parity of the pre-tail
        // bits in the XOR of their rows, flipped by their flip bits
        uint64_t m2 = 0;
        int fl = 0;
        for (int w = 0; w < span; ++w)
          if ((wm >> w) & 1ull) {
            m2 ^= p->fold.rows[w];
            fl ^= (int)((p->fold.flip >> w) & 1ull);
          }
        wm = m2;
        if (fl) std::swap(d0, d1);
      }
      uint64_t gm = 0;
      for (int w = 0; w < span; ++w)
        if ((wm >> w) & 1ull) gm |= 1ull << (n_wires - 1 - w);
      t.mask[o] = gm;
    } else {
      if (obs_wires[o] < 0 || obs_wires[o] >= span) return fail(JQ_ERR_ARG, "observable wire out of range");
      t.gbit[o] = n_wires - 1 - obs_wires[o];
    }
    t.d[2 * o] = d0;
    t.d[2 * o + 1] = d1;
  }
  if (folded) a.obs_kind = obs_kind = JQ_OBS_SV_PARITY;
  if (!table) {
    HIP_TRY(WS(p, st).obs.ensure(sizeof(jq::ObsTable), st));
    table = WS(p, st).obs.p;
  }
  HIP_TRY(jq::launch_obs_upload(t, table, st));
  a.obs = table;
  a.n_obs = n_obs;
  return JQ_OK;
}

int run_prep(jq_plan* p, int batch, const double* angles, const void* matrices, hipStream_t st,
             void* pool = nullptr) {
  const size_t eb = elem_bytes(p->dtype);
  if (!pool) {
    HIP_TRY(WS(p, st).pool.ensure(std::max<size_t>(1, (size_t)p->hp.pool * batch) * eb, st));
    pool = WS(p, st).pool.p;
  }
  jq::PrepArgs pa;
  pa.n_blocks = p->hp.n_dev_blocks;
  pa.batch = batch;
  pa.block_table = (const int*)p->block_table.p;
  pa.member_list = (const int*)p->member_list.p;
  pa.member_table = (const int*)p->member_table.p;
  pa.member_values = (const double*)p->member_values.p;
  pa.angles = angles;
  pa.n_params = p->n_params;
  pa.matrices = matrices;
  pa.n_matrix_elems = p->n_matrix_elems;
  pa.pool = pool;
  pa.pool_stride = p->hp.pool;
  pa.class_mask = p->class_mask;
  pa.wg_blocks = (const int*)p->wg_list.p;
  pa.n_wg = (int)p->wg_blocks.size();
  if (p->n_params > 0 && !angles) return fail(JQ_ERR_ARG, "angles required");
  if (p->n_matrix_elems > 0 && !matrices) return fail(JQ_ERR_ARG, "explicit matrices required");
  HIP_TRY(jq::launch_prep(p->dtype, pa, st));
  return JQ_OK;
}

jq::SvArgs base_args(jq_plan* p, hipStream_t st) {
  jq::SvArgs a;
  std::memset(&a, 0, sizeof(a));
  a.prog = (const int*)p->prog.p;
  a.n_wires = p->hp.n_wires;
  a.vec_len = 1ull << p->hp.n_wires;
  a.mats = WS(p, st).pool.p;
  a.pool_stride = p->hp.pool;
  return a;
}

size_t smem_for(jq_plan* p, int extra_reals, int regions) {
  // state region(s) | observable table (64 ints + 128 doubles) | 64 expval + slot reals
  return (size_t)regions * ((size_t)elem_bytes(p->dtype) << p->hp.T) + sizeof(jq::ObsTable) + (size_t)(64 + extra_reals) * 8;
}

int max_slots(jq_plan* p) {
  int m = 0;
  const int* W = p->hp.words.data();
  for (int q = 0; q < p->hp.n_passes; ++q)
    if (!p->pass_wide[q]) m = std::max(m, W[W[jq::HDR_WORDS + q] + 2]);
  return m;
}


// per-tile cotangent partials -> block cotangents -> angle / matrix gradients
// quat_t: the partials come from the plan-specialised backward, whose
// rotation blocks flush their cotangent in su(2) form (ChainArgs::quat_t).
int finish_grads(jq_plan* p, int batch, const double* angles, const void* matrices, double* grad_part,
                 int64_t part_stride, double* grad_angles, void* grad_matrices, bool quat_t, hipStream_t st) {
  const int ntiles = p->hp.ntiles;
  const int nreals = p->hp.bars * 2;
  if (grad_angles && p->n_params > 0)
    HIP_TRY(jq::launch_zero(grad_angles, (size_t)batch * p->n_params * 8, st));
  if (grad_matrices && p->n_matrix_elems > 0)
    HIP_TRY(jq::launch_zero(grad_matrices, (size_t)batch * p->n_matrix_elems * 16, st));
  if (nreals > 0 && (grad_angles || grad_matrices)) {
    HIP_TRY(WS(p, st).bars.ensure((size_t)batch * nreals * 8, st));
    jq::ReduceArgs r;
    r.slot_map = (const int*)p->slot_map.p;
    r.n_slots = p->n_slots;
    r.partials = grad_part;
    r.part_stride = part_stride;
    r.ntiles = ntiles;
    r.bars = (double*)WS(p, st).bars.p;
    r.n_reals = nreals;
    r.batch = batch;
    HIP_TRY(jq::launch_reduce_grad(r, st));
    jq::ChainArgs c;
    c.n_blocks = p->hp.n_dev_blocks;
    c.batch = batch;
    c.block_table = (const int*)p->block_table.p;
    c.member_list = (const int*)p->member_list.p;
    c.member_table = (const int*)p->member_table.p;
    c.member_values = (const double*)p->member_values.p;
    c.angles = angles;
    c.n_params = p->n_params;
    c.matrices = matrices;
    c.n_matrix_elems = p->n_matrix_elems;
    c.bars = (const double*)WS(p, st).bars.p;
    c.bars_stride = p->hp.bars;
    c.grad_angles = grad_angles;
    c.grad_matrices = grad_matrices;
    c.class_mask = p->class_mask;
    c.max_members2 = p->max_members2;
    c.quat_t = quat_t ? 1 : 0;
    c.wg_blocks = (const int*)p->wg_list.p;
    c.n_wg = (int)p->wg_blocks.size();
    if (grad_angles || grad_matrices) HIP_TRY(jq::launch_chain(c, st));
  }
  return JQ_OK;
}

// Output of pass q: the checkpoint slice when checkpoints are kept, else the
// final vector (passes then run in place).
void* pass_out(jq_plan* p, bool ck, void* state, size_t vbytes, int q, hipStream_t st) {
  if (q == p->hp.n_passes - 1 || !ck) return state;
  return (char*)WS(p, st).ckpt.p + vbytes * (size_t)q;
}

// Forward passes of one call (prep, passes, expectation values).
int run_forward(jq_plan* p, int batch, const double* angles, const void* matrices, const void* psi0,
                int psi0_batched, void* state, int obs_kind, int n_obs, const int32_t* obs_wires, const double* obs_d,
                double* expvals, bool ck, hipStream_t st) {
  int rc = ensure_uploaded(p);
  if (rc) return rc;
  ++WS(p, st).generation;  // the workspaces (pool, checkpoints) change from here on
  rc = run_prep(p, batch, angles, matrices, st);
  if (rc) return rc;
  jq::SvArgs a = base_args(p, st);
  rc = fill_obs(p, a, p->hp.n_wires, expvals ? obs_kind : JQ_OBS_NONE, n_obs, obs_wires, obs_d, st);
  if (rc) return rc;
  const int ntiles = p->hp.ntiles;
  const int P = p->hp.n_passes;
  if (a.obs_kind != JQ_OBS_NONE) {
    HIP_TRY(WS(p, st).ev_part.ensure(std::max<size_t>(1, (size_t)batch * ntiles * a.n_obs) * 8, st));
    a.ev_part = (double*)WS(p, st).ev_part.p;
  }
  const size_t vbytes = (size_t)a.vec_len * elem_bytes(p->dtype) * batch;
  ck = ck && P > 1;
  if (ck) HIP_TRY(WS(p, st).ckpt.ensure(vbytes * (size_t)(P - 1), st));
  const bool jit = jit_ready(p);
  if (!jit && p->jit.required)
    return fail(JQ_ERR_UNSUPPORTED, "tile_bits 14 plans run only on plan-specialised kernels: " +
                                        (p->jit.log.empty() ? std::string("specialisation is off") : p->jit.log));
  for (int q = 0; q < P; ++q) {
    a.pass = q;
    a.src = q == 0 ? psi0 : pass_out(p, ck, state, vbytes, q - 1, st);
    a.src_stride = q == 0 ? (psi0_batched ? (int64_t)a.vec_len : 0) : (int64_t)a.vec_len;
    a.dst = pass_out(p, ck, state, vbytes, q, st);
    a.dst_stride = (int64_t)a.vec_len;
    HIP_TRY(timed_launch(0, p->pass_full[q], p->dtype, p->hp.T, a, ntiles, batch, smem_for(p, 0), st, false,
                         jit_fn(p, jit, 0, q), q == 0, q == P - 1));
  }
  if (a.obs_kind != JQ_OBS_NONE && expvals)
    HIP_TRY(jq::launch_reduce_ev(a.ev_part, expvals, ntiles, a.n_obs, batch, st));
  return JQ_OK;
}

// Backward sweep over the passes of the preceding run_forward (same ck).
int run_backward(jq_plan* p, int batch, const double* angles, const void* matrices, const void* psi0,
                 int psi0_batched, void* state, int obs_kind, int n_obs, const int32_t* obs_wires,
                 const double* obs_d, const double* ct, void* lam_seed, double* grad_angles, void* grad_matrices,
                 void* grad_psi0, bool ck, hipStream_t st) {
  jq::SvArgs a = base_args(p, st);
  if (lam_seed && p->fold.n_folded > 0)
    return fail(JQ_ERR_UNSUPPORTED, "a plan with a folded tail takes observables, not a final-vector cotangent");
  if (lam_seed) {
    a.lam_given = 1;
  } else {
    int rc = fill_obs(p, a, p->hp.n_wires, obs_kind, n_obs, obs_wires, obs_d, st);
    if (rc) return rc;
    a.ct = ct;
  }
  const int ntiles = p->hp.ntiles;
  const int P = p->hp.n_passes;
  const size_t vbytes = (size_t)a.vec_len * elem_bytes(p->dtype) * batch;
  ck = ck && P > 1;
  HIP_TRY(WS(p, st).ev_part.ensure(std::max<size_t>(1, (size_t)batch * ntiles * std::max(1, a.n_obs)) * 8, st));
  HIP_TRY(WS(p, st).partials.ensure(std::max<size_t>(1, (size_t)batch * p->hp.partial_reals) * 8, st));
  a.grad_part = (double*)WS(p, st).partials.p;
  a.part_stride = p->hp.partial_reals;
  a.ev_part = (double*)WS(p, st).ev_part.p;
  // the running adjoint vector ends as the initial-state cotangent
  void* lam = grad_psi0;
  if (lam_seed && grad_psi0) {
    HIP_TRY(hipMemcpyAsync(grad_psi0, lam_seed, vbytes, hipMemcpyDeviceToDevice, st));
  } else if (lam_seed) {
    lam = lam_seed;
  } else if (!lam) {
    HIP_TRY(WS(p, st).lam.ensure(vbytes, st));
    lam = WS(p, st).lam.p;
  }
  a.lam = lam;
  a.lam_stride = (int64_t)a.vec_len;
  const bool jit = jit_ready(p);
  if (!jit && p->jit.required)
    return fail(JQ_ERR_UNSUPPORTED, "tile_bits 14 plans run only on plan-specialised kernels: " +
                                        (p->jit.log.empty() ? std::string("specialisation is off") : p->jit.log));
  const size_t sm = smem_for(p, 2 * max_slots(p), jit ? p->regions_bj : p->regions_b);
  a.prog = (const int*)(jit ? p->prog_bj.p : p->prog_b.p);
  for (int q = P - 1; q >= 0; --q) {
    a.pass = q;
    if (p->pass_wide[q]) {  // reads the pass's input state
      a.src = q == 0 ? psi0 : pass_out(p, ck, state, vbytes, q - 1, st);
      a.src_stride = q == 0 ? (psi0_batched ? (int64_t)a.vec_len : 0) : (int64_t)a.vec_len;
      a.dst = nullptr;
    } else {
      a.src = pass_out(p, ck, state, vbytes, q, st);
      a.src_stride = (int64_t)a.vec_len;
      a.dst = ck ? nullptr : state;  // without checkpoints psi is un-applied in place
    }
    a.dst_stride = (int64_t)a.vec_len;
    HIP_TRY(timed_launch(1, p->pass_full[q], p->dtype, p->hp.T, a, ntiles, batch, sm, st, p->wide_bwd,
                         jit_fn(p, jit, 1, q), q == P - 1, q == 0));
  }
  return finish_grads(p, batch, angles, matrices, a.grad_part, a.part_stride, grad_angles, grad_matrices, jit, st);
}
}  // namespace

extern "C" {

int jq_sv_forward(jq_plan* p, int32_t batch, const double* angles, const void* matrices, const void* psi0,
                  int32_t psi0_batched, void* state, int32_t obs_kind, int32_t n_obs, const int32_t* obs_wires,
                  const double* obs_d, double* expvals, void* stream) {
  if (!p || !state || batch < 1) return fail(JQ_ERR_ARG, "bad arguments to jq_sv_forward");
  if (batch > 65535) return fail(JQ_ERR_ARG, "batch above 65535 per call");
  return run_forward(p, batch, angles, matrices, psi0, psi0_batched, state, obs_kind, n_obs, obs_wires, obs_d, expvals,
                     false, (hipStream_t)stream);
}

int jq_sv_forward_noalloc(jq_plan* p, int32_t batch, const double* angles, const void* matrices, const void* psi0,
                          int32_t psi0_batched, void* state, int32_t obs_kind, int32_t n_obs,
                          const int32_t* obs_wires, const double* obs_d, double* expvals, void* stream) {
  if (!p || !state || batch < 1) return fail(JQ_ERR_ARG, "bad arguments to jq_sv_forward_noalloc");
  if (batch > 65535) return fail(JQ_ERR_ARG, "batch above 65535 per call");
  if (!p->uploaded)
    return fail(JQ_ERR_ARG, "jq_sv_forward_noalloc: the plan's program is not resident on the device yet "
                            "(run any other entry point on the plan once first)");
  hipStream_t st = (hipStream_t)stream;
  // scratch layout: observable table | block matrices | expectation partials
  const size_t eb = elem_bytes(p->dtype);
  const int ntiles = p->hp.ntiles;
  const size_t off_pool = 4096;  // >= sizeof(jq::ObsTable), 256-aligned
  const size_t pool_bytes = ((std::max<size_t>(1, (size_t)p->hp.pool * batch) * eb + 255) / 256) * 256;
  const size_t off_ev = off_pool + pool_bytes;
  const size_t ev_bytes = expvals ? (size_t)batch * ntiles * std::max(1, n_obs) * 8 : 0;
  static_assert(sizeof(jq::ObsTable) <= 4096, "observable table exceeds its scratch slot");
  if (off_ev + ev_bytes > (size_t)JQ_NOALLOC_SCRATCH_BYTES)
    return fail(JQ_ERR_ARG, "jq_sv_forward_noalloc: the call needs " + std::to_string(off_ev + ev_bytes) +
                                " bytes of scratch, above JQ_NOALLOC_SCRATCH_BYTES (" +
                                std::to_string((size_t)JQ_NOALLOC_SCRATCH_BYTES) + "); use jq_sv_forward");
  void* scr = nullptr;
  HIP_TRY(jq::noalloc_scratch(&scr));
  char* base = (char*)scr;
  int rc = run_prep(p, batch, angles, matrices, st, base + off_pool);
  if (rc) return rc;
  ++WS(p, st).generation;  // the plan's own checkpoints no longer match its pool
  jq::SvArgs a = base_args(p, st);
  a.mats = base + off_pool;
  rc = fill_obs(p, a, p->hp.n_wires, expvals ? obs_kind : JQ_OBS_NONE, n_obs, obs_wires, obs_d, st, base);
  if (rc) return rc;
  if (a.obs_kind != JQ_OBS_NONE) a.ev_part = (double*)(base + off_ev);
  // plan-specialised kernels only when already loaded (loading allocates)
  const bool jit = p->jit.mode != JQ_JIT_OFF && p->jit.phase.load(std::memory_order_acquire) == JQ_JIT_LOADED;
  if (!jit && p->jit.required)
    return fail(JQ_ERR_UNSUPPORTED, "tile_bits 14 plans run only on plan-specialised kernels: " +
                                        (p->jit.log.empty() ? std::string("specialisation is off") : p->jit.log));
  const int P = p->hp.n_passes;
  for (int q = 0; q < P; ++q) {
    a.pass = q;
    a.src = q == 0 ? psi0 : state;
    a.src_stride = q == 0 ? (psi0_batched ? (int64_t)a.vec_len : 0) : (int64_t)a.vec_len;
    a.dst = state;
    a.dst_stride = (int64_t)a.vec_len;
    const int variant = 0 + (p->pass_full[q] ? 2 : 0);
    HIP_TRY(launch_pass(jit_fn(p, jit, 0, q), variant, p->dtype, p->hp.T, a, ntiles, batch, smem_for(p, 0), st));
  }
  if (a.obs_kind != JQ_OBS_NONE && expvals)
    HIP_TRY(jq::launch_reduce_ev(a.ev_part, expvals, ntiles, a.n_obs, batch, st));
  return JQ_OK;
}

int jq_sv_adjoint(jq_plan* p, int32_t batch, const double* angles, const void* matrices, const void* psi0,
                  int32_t psi0_batched, void* state, int32_t obs_kind, int32_t n_obs, const int32_t* obs_wires,
                  const double* obs_d, const double* ct, double* expvals, double* grad_angles, void* grad_matrices,
                  void* grad_psi0, void* stream) {
  if (!p || !state || batch < 1 || !ct) return fail(JQ_ERR_ARG, "bad arguments to jq_sv_adjoint");
  if (batch > 65535) return fail(JQ_ERR_ARG, "batch above 65535 per call");
  if (obs_kind == JQ_OBS_NONE) return fail(JQ_ERR_ARG, "the adjoint needs observables");
  hipStream_t st = (hipStream_t)stream;
  const bool ck = p->checkpoints != 0;
  int rc = run_forward(p, batch, angles, matrices, psi0, psi0_batched, state, obs_kind, n_obs, obs_wires, obs_d,
                       expvals, ck, st);
  if (rc) return rc;
  return run_backward(p, batch, angles, matrices, psi0, psi0_batched, state, obs_kind, n_obs, obs_wires, obs_d, ct,
                      nullptr, grad_angles, grad_matrices, grad_psi0, ck, st);
}

int jq_sv_forward_ckpt(jq_plan* p, int32_t batch, const double* angles, const void* matrices, const void* psi0,
                       int32_t psi0_batched, void* state, int32_t obs_kind, int32_t n_obs, const int32_t* obs_wires,
                       const double* obs_d, double* expvals, int64_t* token, void* stream) {
  if (!p || !state || batch < 1 || !token) return fail(JQ_ERR_ARG, "bad arguments to jq_sv_forward_ckpt");
  if (batch > 65535) return fail(JQ_ERR_ARG, "batch above 65535 per call");
  *token = -1;
  hipStream_t st = (hipStream_t)stream;
  int rc = run_forward(p, batch, angles, matrices, psi0, psi0_batched, state, obs_kind, n_obs, obs_wires, obs_d,
                       expvals, true, st);
  if (rc) return rc;
  Workspace& w = WS(p, st);
  w.ckpt_token = w.generation;
  w.ckpt_batch = batch;
  *token = w.ckpt_token;
  return JQ_OK;
}

int jq_sv_backward(jq_plan* p, int64_t token, int32_t batch, const double* angles, const void* matrices,
                   const void* psi0, int32_t psi0_batched, void* state, int32_t obs_kind, int32_t n_obs,
                   const int32_t* obs_wires, const double* obs_d, const double* ct, void* lam_seed,
                   double* grad_angles, void* grad_matrices, void* grad_psi0, void* stream) {
  if (!p || !state || batch < 1) return fail(JQ_ERR_ARG, "bad arguments to jq_sv_backward");
  if (batch > 65535) return fail(JQ_ERR_ARG, "batch above 65535 per call");
  if (!lam_seed && (obs_kind == JQ_OBS_NONE || !ct))
    return fail(JQ_ERR_ARG, "jq_sv_backward needs observables and their cotangents, or lam_seed");
  hipStream_t st = (hipStream_t)stream;
  Workspace& w = WS(p, st);
  const bool valid = token >= 0 && token == w.ckpt_token && token == w.generation && batch == w.ckpt_batch;
  if (!valid) {  // another call has used the plan on this stream since: run the forward again
    int rc = run_forward(p, batch, angles, matrices, psi0, psi0_batched, state, obs_kind, n_obs, obs_wires, obs_d,
                         nullptr, true, st);
    if (rc) return rc;
    w.ckpt_token = w.generation;
    w.ckpt_batch = batch;
  }
  return run_backward(p, batch, angles, matrices, psi0, psi0_batched, state, obs_kind, n_obs, obs_wires, obs_d, ct,
                      lam_seed, grad_angles, grad_matrices, grad_psi0, true, st);
}

int jq_sv_adjoint_groups(int32_t n_groups, const jq_group* groups, int32_t obs_kind, int32_t n_obs,
                         const int32_t* obs_wires, const double* obs_d, void* stream) {
  if (n_groups < 1 || !groups) return fail(JQ_ERR_ARG, "no groups");
  if (obs_kind == JQ_OBS_NONE) return fail(JQ_ERR_ARG, "the adjoint needs observables");
  hipStream_t st = (hipStream_t)stream;
  jq_plan* p0 = groups[0].plan;
  if (!p0) return fail(JQ_ERR_ARG, "null plan");
  std::vector<int> order(n_groups);
  int64_t total = 0;
  for (int g = 0; g < n_groups; ++g) {
    const jq_group& G = groups[g];
    if (!G.plan || G.batch < 1 || !G.ct || !G.state) return fail(JQ_ERR_ARG, "bad group");
    if (G.plan->hp.n_wires != p0->hp.n_wires || G.plan->dtype != p0->dtype)
      return fail(JQ_ERR_ARG, "groups must share width and dtype");
    for (int h = 0; h < g; ++h)
      if (groups[h].plan == G.plan)
        return fail(JQ_ERR_ARG, "two groups share a plan: merge circuits of one structure into one group");
    order[g] = g;
    total += G.batch;
  }
  if (total > 65535) return fail(JQ_ERR_ARG, "more than 65535 circuits per call");
  // Every plan has its specialised kernels: run the groups one after the
  // other through the single-plan adjoint (a few launches per pass and group
  // instead of one per pass step, but each pass runs its straight-line
  // kernel).  Otherwise all groups share the interpreter's launches below.
  bool all_jit = true;
  bool any_wide = false;
  for (int g = 0; g < n_groups; ++g) {
    all_jit = all_jit && jit_ready(groups[g].plan);
    any_wide = any_wide || groups[g].plan->any_wide;
  }
  for (int g = 0; g < n_groups; ++g)
    if (groups[g].plan->jit.required && !all_jit)
      return fail(JQ_ERR_UNSUPPORTED, "tile_bits 14 groups need every group's plan-specialised kernels");
  // (groups that run their own specialised adjoints may differ in tile size;
  // the shared launches below need one program geometry)
  if (!(all_jit || any_wide))
    for (int g = 0; g < n_groups; ++g) {
      if (groups[g].plan->hp.T != p0->hp.T || groups[g].plan->wide_bwd != p0->wide_bwd)
        return fail(JQ_ERR_ARG, "groups on the shared launches must share tile size and backward register tile");
      // (one observable table serves every group there)
      if (!(groups[g].plan->fold == p0->fold))
        return fail(JQ_ERR_UNSUPPORTED, "groups on the shared launches must fold the same tail");
    }
  // (wide-block passes run their own kernels: such groups take this path too)
  if (all_jit || any_wide) {
    // Groups share nothing (each plan owns its workspaces), so they run on
    // side streams forked from and joined back to the caller's stream with
    // events: small groups' launches overlap instead of leaving the device
    // partly idle.  Event fork / join keeps the call capturable.
    SideStreams* ss = nullptr;
    hipError_t e = side_streams(&ss);
    if (e != hipSuccess) return fail(JQ_ERR_DEVICE, std::string("side streams: ") + hipGetErrorString(e));
    const int ns = SideStreams::kStreams;
    std::lock_guard<std::mutex> lock(ss->mu);
    HIP_TRY(hipEventRecord(ss->fork, st));
    for (int k = 0; k < ns; ++k) HIP_TRY(hipStreamWaitEvent(ss->s[k], ss->fork, 0));
    // largest groups first, one per stream in turn
    std::vector<int> by_work(n_groups);
    for (int g = 0; g < n_groups; ++g) by_work[g] = g;
    std::stable_sort(by_work.begin(), by_work.end(), [&](int x, int y) {
      return (int64_t)groups[x].batch * groups[x].plan->hp.n_passes > (int64_t)groups[y].batch * groups[y].plan->hp.n_passes;
    });
    // A plan's workspaces serve one launch sequence at a time: groups that
    // share a plan share its stream.
    std::map<const jq_plan*, int> stream_of;
    int rc = JQ_OK;
    for (int j = 0; j < n_groups && rc == JQ_OK; ++j) {
      const jq_group& G = groups[by_work[j]];
      const int k = stream_of.emplace(G.plan, (int)stream_of.size() % ns).first->second;
      rc = jq_sv_adjoint(G.plan, G.batch, G.angles, G.matrices, G.psi0, G.psi0_batched, G.state, obs_kind, n_obs,
                         obs_wires, obs_d, G.ct, G.expvals, G.grad_angles, nullptr, nullptr, ss->s[k]);
    }
    for (int k = 0; k < ns; ++k) {
      HIP_TRY(hipEventRecord(ss->join[k], ss->s[k]));
      HIP_TRY(hipStreamWaitEvent(st, ss->join[k], 0));
    }
    return rc;
  }
  // groups with more passes first: the circuits still active at a step are a prefix
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return groups[x].plan->hp.n_passes > groups[y].plan->hp.n_passes; });
  jq::SvArgs a = base_args(p0, st);
  int rc = fill_obs(p0, a, p0->hp.n_wires, obs_kind, n_obs, obs_wires, obs_d, st);
  if (rc) return rc;
  const size_t eb = elem_bytes(p0->dtype);
  const int ntiles = p0->hp.ntiles;
  std::vector<jq::SvGroup> table(n_groups);
  std::vector<int64_t> row_begin(n_groups + 1, 0);
  int maxP = 0;
  size_t smem_f = smem_for(p0, 0), smem_b = 0;
  for (int k = 0; k < n_groups; ++k) {
    const jq_group& G = groups[order[k]];
    jq_plan* p = G.plan;
    rc = ensure_uploaded(p);
    if (rc) return rc;
    ++WS(p, st).generation;
    rc = run_prep(p, G.batch, G.angles, G.matrices, st);
    if (rc) return rc;
    const int P = p->hp.n_passes;
    const size_t vbytes = (size_t)a.vec_len * eb * G.batch;
    HIP_TRY(WS(p, st).ckpt.ensure(std::max<size_t>(1, vbytes * (size_t)std::max(0, P - 1)), st));
    HIP_TRY(WS(p, st).lam.ensure(vbytes, st));
    HIP_TRY(WS(p, st).ev_part.ensure(std::max<size_t>(1, (size_t)G.batch * ntiles * std::max(1, n_obs)) * 8, st));
    HIP_TRY(WS(p, st).partials.ensure(std::max<size_t>(1, (size_t)G.batch * p->hp.partial_reals) * 8, st));
    jq::SvGroup& t = table[k];
    std::memset(&t, 0, sizeof(t));
    t.prog_f = (const int*)p->prog.p;
    t.prog_b = (const int*)p->prog_b.p;
    t.mats = WS(p, st).pool.p;
    t.pool_stride = p->hp.pool;
    t.psi0 = G.psi0;
    t.psi0_stride = G.psi0_batched ? (int64_t)a.vec_len : 0;
    t.ckpt = WS(p, st).ckpt.p;
    t.state = G.state;
    t.lam = WS(p, st).lam.p;
    t.ev_part = (double*)WS(p, st).ev_part.p;
    t.grad_part = (double*)WS(p, st).partials.p;
    t.part_stride = p->hp.partial_reals;
    t.ct = G.ct;
    t.n_passes = P;
    t.circ_begin = (int)row_begin[k];
    t.batch = G.batch;
    row_begin[k + 1] = row_begin[k] + G.batch;
    maxP = std::max(maxP, P);
    smem_b = std::max(smem_b, smem_for(p, 2 * max_slots(p), p->regions_b));
  }
  HIP_TRY(WS(p0, st).groups.ensure(sizeof(jq::SvGroup) * n_groups, st));
  HIP_TRY(jq::launch_bytes_upload(table.data(), sizeof(jq::SvGroup) * n_groups, WS(p0, st).groups.p, st));
  a.groups = (const jq::SvGroup*)WS(p0, st).groups.p;
  a.n_groups = n_groups;
  // rows active at step s: groups k with n_passes > s (a prefix of the order)
  auto active_rows = [&](int s, bool& full, int dir) -> int {
    int rows = 0;
    full = false;
    for (int k = 0; k < n_groups; ++k) {
      jq_plan* p = groups[order[k]].plan;
      const int P = p->hp.n_passes;
      if (P <= s) break;
      rows = (int)row_begin[k + 1];
      const int q = dir == 0 ? s : P - 1 - s;
      full = full || p->pass_full[q];
    }
    return rows;
  };
  for (int s = 0; s < maxP; ++s) {
    bool full;
    const int rows = active_rows(s, full, 0);
    a.step = s;
    a.dir = 0;
    HIP_TRY(timed_launch(0, full, p0->dtype, p0->hp.T, a, ntiles, rows, smem_f, st, false, JitFn(), s == 0,
                         s == maxP - 1));
  }
  for (int k = 0; k < n_groups; ++k) {
    const jq_group& G = groups[order[k]];
    jq_plan* p = G.plan;
    if (G.expvals) HIP_TRY(jq::launch_reduce_ev((double*)WS(p, st).ev_part.p, G.expvals, ntiles, n_obs, G.batch, st));
  }
  for (int s = 0; s < maxP; ++s) {
    bool full;
    const int rows = active_rows(s, full, 1);
    a.step = s;
    a.dir = 1;
    HIP_TRY(timed_launch(1, full, p0->dtype, p0->hp.T, a, ntiles, rows, smem_b, st, p0->wide_bwd, JitFn(), s == 0,
                         s == maxP - 1));
  }
  for (int k = 0; k < n_groups; ++k) {
    const jq_group& G = groups[order[k]];
    jq_plan* p = G.plan;
    rc = finish_grads(p, G.batch, G.angles, G.matrices, (double*)WS(p, st).partials.p, p->hp.partial_reals, G.grad_angles,
                      nullptr, false, st);
    if (rc) return rc;
  }
  return JQ_OK;
}

namespace {
int evolve_args(const jq_evolve_desc* d, const double* coef, const double* hmats, jq::EvolveArgs& a) {
  if (!d || !coef || !hmats) return fail(JQ_ERR_ARG, "null argument to jq_evolve");
  if (d->dim != 2 && d->dim != 4)
    return fail(JQ_ERR_UNSUPPORTED, "propagators of one- and two-qubit Hamiltonians (dim 2 or 4) are supported");
  if (d->n_gates < 0 || d->n_terms < 1 || d->n_factors < 1) return fail(JQ_ERR_ARG, "bad jq_evolve_desc");
  if (d->n_gates > 2147483647 / 64) return fail(JQ_ERR_ARG, "too many gates in one call");
  a.n_gates = d->n_gates;
  a.dim = d->dim;
  a.n_terms = d->n_terms;
  a.n_factors = d->n_factors;
  a.hmats_batched = d->hmats_batched != 0;
  a.coef = coef;
  a.hmats = hmats;
  a.unitaries = nullptr;
  a.cot = nullptr;
  a.grad_coef = nullptr;
  return JQ_OK;
}
}  // namespace

int jq_evolve(const jq_evolve_desc* d, const double* coef, const double* hmats, double* unitaries, void* stream) {
  jq::EvolveArgs a;
  int rc = evolve_args(d, coef, hmats, a);
  if (rc) return rc;
  if (!unitaries) return fail(JQ_ERR_ARG, "null unitaries");
  a.unitaries = unitaries;
  HIP_TRY(jq::launch_evolve(a, 0, (hipStream_t)stream));
  return JQ_OK;
}

int jq_evolve_vjp(const jq_evolve_desc* d, const double* coef, const double* hmats, const double* cot,
                  double* grad_coef, void* stream) {
  jq::EvolveArgs a;
  int rc = evolve_args(d, coef, hmats, a);
  if (rc) return rc;
  if (!cot || !grad_coef) return fail(JQ_ERR_ARG, "null cotangent or gradient");
  a.cot = cot;
  a.grad_coef = grad_coef;
  HIP_TRY(jq::launch_evolve(a, 1, (hipStream_t)stream));
  return JQ_OK;
}

int jq_sv_probs(int32_t n_wires, int32_t batch, int32_t dtype, const void* state, void* probs, void* stream) {
  if (!state || !probs || n_wires < 1 || n_wires > 40 || batch < 1) return fail(JQ_ERR_ARG, "bad arguments to jq_sv_probs");
  HIP_TRY(jq::launch_probs(dtype, state, probs, ((int64_t)1 << n_wires) * batch, (hipStream_t)stream));
  return JQ_OK;
}

int jq_sv_outer(int32_t n_wires, int32_t batch, int32_t dtype, const void* state, void* rho, void* stream) {
  if (!state || !rho || n_wires < 1 || n_wires > 20 || batch < 1) return fail(JQ_ERR_ARG, "bad arguments to jq_sv_outer");
  HIP_TRY(jq::launch_outer(dtype, state, rho, (int64_t)1 << n_wires, batch, (hipStream_t)stream));
  return JQ_OK;
}

int jq_dm_diag(int32_t n_qubits, int32_t batch, int32_t dtype, const void* rho, void* probs, void* stream) {
  if (!rho || !probs || n_qubits < 1 || n_qubits > 20 || batch < 1) return fail(JQ_ERR_ARG, "bad arguments to jq_dm_diag");
  HIP_TRY(jq::launch_dm_diag(dtype, rho, probs, (int64_t)1 << n_qubits, batch, (hipStream_t)stream));
  return JQ_OK;
}

int jq_expval_matrix(int32_t n_qubits, int32_t batch, int32_t dtype, int32_t density, const void* state, int32_t k,
                     const int32_t* wires, const double* matrix, double* out, void* stream) {
  if (!state || !out || !matrix || !wires || k < 1 || k > 5 || k > n_qubits || batch < 1)
    return fail(JQ_ERR_ARG, "bad arguments to jq_expval_matrix");
  jq::ExpvalArgs a;
  std::memset(&a, 0, sizeof(a));
  a.n_qubits = n_qubits;
  a.k = k;
  a.density = density;
  a.state = state;
  a.stride = density ? ((int64_t)1 << (2 * n_qubits)) : ((int64_t)1 << n_qubits);
  for (int j = 0; j < k; ++j) {
    if (wires[j] < 0 || wires[j] >= n_qubits) return fail(JQ_ERR_ARG, "observable wire out of range");
    a.gbits[j] = n_qubits - 1 - wires[j];
    a.wmask |= 1ll << a.gbits[j];
  }
  a.mat = matrix;
  a.out = out;
  hipStream_t st = (hipStream_t)stream;
  HIP_TRY(jq::launch_zero(out, (size_t)batch * 8, st));
  HIP_TRY(jq::launch_expval_matrix(dtype, a, batch, st));
  return JQ_OK;
}

int jq_timing_enable(int32_t on) {
  g_timer.on = on != 0;
  return JQ_OK;
}

int jq_timing_read(int32_t which, double* total_ms, int32_t* count) {
  if (which < 0 || which > 2) return fail(JQ_ERR_ARG, "which must be 0 (forward), 1 (backward) or 2 (fused)");
  double tot = 0;
  int n = 0;
  for (auto& r : g_timer.runs[which]) {
    HIP_TRY(hipEventSynchronize(r.e1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, r.e0, r.e1));
    tot += ms;
    n += r.launches;
    g_timer.pool.push_back(r.e0);
    g_timer.pool.push_back(r.e1);
  }
  g_timer.runs[which].clear();
  if (total_ms) *total_ms = tot;
  if (count) *count = n;
  return JQ_OK;
}

int jq_pulse_propagate(const jq_pulse_desc* desc, const double* params, const double* target, double* unitaries,
                       double* infidelity, double* grad, void* stream) {
  if (!desc || !params || !target || !infidelity) return fail(JQ_ERR_ARG, "null argument to jq_pulse_propagate");
  if (desc->envelope != JQ_ENV_DRAG_GAUSSIAN) return fail(JQ_ERR_UNSUPPORTED, "unknown pulse envelope");
  if (desc->n_sets < 1 || desc->n_sets > (1 << 30) || desc->n_steps < 1 || !(desc->t_total > 0))
    return fail(JQ_ERR_ARG, "bad pulse batch sizes");
  jq::PulseArgs a;
  std::memset(&a, 0, sizeof(a));
  a.n_sets = desc->n_sets;
  a.n_steps = desc->n_steps;
  a.T = desc->t_total;
  a.J = desc->coupling_j;
  a.params = params;
  std::memcpy(a.V, target, sizeof(a.V));
  a.U = unitaries;
  a.infid = infidelity;
  a.grad = grad;
  HIP_TRY(jq::launch_pulse(a, (hipStream_t)stream));
  return JQ_OK;
}

}  // extern "C"
