// Plan-specialised tile-pass kernels: source generation and hipRTC compile.
//
// The generated body of a pass repeats, gate by gate, what the interpreter
// (InterpBody in sv_impl.h) does with the same program words: the layout
// check before each substep, the control tests, the gate arithmetic and the
// gradient flushes, in the same order and through the same device
// functions.  Only the record decoding and the dispatch disappear, so the
// two paths produce the same amplitudes and gradients.
#include "jit.h"

#include <hip/hiprtc.h>

#include <cstdio>
#include <sstream>

#include "program.h"

namespace jq {

namespace {

#include "jit_headers.inc"  // kJitHeaderNames / kJitHeaderTexts (build.py embeds the kernel headers)

// Types of <stdint.h> for the run-time compiler, which declares them in its
// own namespace.
const char* kPrelude =
    "typedef __hip_internal::int8_t int8_t;\n"
    "typedef __hip_internal::uint8_t uint8_t;\n"
    "typedef __hip_internal::int16_t int16_t;\n"
    "typedef __hip_internal::uint16_t uint16_t;\n"
    "typedef __hip_internal::int32_t int32_t;\n"
    "typedef __hip_internal::uint32_t uint32_t;\n"
    "typedef __hip_internal::int64_t int64_t;\n"
    "typedef __hip_internal::uint64_t uint64_t;\n"
    "typedef unsigned long uintptr_t;\n"
    "#include \"sv_impl.h\"\n"
    "namespace jq {\n";

struct ProgView {
  const int32_t* W;
  int T, RB, n_passes;
  const int32_t* pass(int q) const { return W + W[HDR_WORDS + q]; }
  const int32_t* sub(int q, int s) const { return W + pass(q)[4 + T + s]; }
};

std::string hex32(uint32_t v) {
  char b[24];
  std::snprintf(b, sizeof(b), "0x%xu", v);
  return b;
}
std::string hex64(uint64_t v) {
  char b[32];
  std::snprintf(b, sizeof(b), "0x%llxull", (unsigned long long)v);
  return b;
}

// layout words of a substep record: RB register sources, then the thread sources
std::string layout_args(const ProgView& pv, const int32_t* sub) {
  std::string o;
  for (int j = 0; j < pv.T; ++j) o += ", " + std::to_string(sub[1 + j]);
  return o;
}

bool same_rpos(const ProgView& pv, const int32_t* sub, const std::vector<int>& cur) {
  for (int b = 0; b < pv.RB; ++b)
    if (sub[1 + b] != cur[b]) return false;
  return true;
}

std::vector<int> canonical_rpos(const ProgView& pv) {
  std::vector<int> r(pv.RB);
  r[0] = 0;
  for (int b = 1; b < pv.RB; ++b) r[b] = pv.T - pv.RB + b;
  return r;
}

struct GateLit {
  int kind, ra, rb, hi, lo, k, mat, inv, slot, w0, w1;
  bool quat;
  uint32_t creg, cthr;
  uint64_t cout;
  GateLit(const int32_t* g, int rebase) {
    kind = g[G_KIND];
    ra = g[G_RA];
    rb = g[G_RB];
    hi = ra > rb ? ra : rb;
    lo = ra > rb ? rb : ra;
    k = g[G_K];
    mat = g[G_MAT] - rebase;
    inv = g[G_INV] >= 0 ? g[G_INV] - rebase : -1;
    slot = g[G_SLOT];
    quat = kind == AK_MC1 && g[G_QUAT] != 0;
    w0 = g[G_DW0];
    w1 = g[G_DW0 + 1];
    creg = (uint32_t)g[G_CREG];
    cthr = (uint32_t)g[G_CTHR];
    cout = ((uint64_t)(uint32_t)g[G_COUT_HI] << 32) | (uint32_t)g[G_COUT_LO];
  }
  bool ctl() const { return cthr != 0 || cout != 0; }
  std::string ok() const { return ctl() ? "ctl_ok(L, base, " + hex32(cthr) + ", " + hex64(cout) + ")" : "true"; }
};

void emit_fwd_gate(std::ostringstream& o, const GateLit& g) {
  const std::string mp = "pool + " + std::to_string(g.mat);
  const std::string nr = "(1 << RB)";
  std::string call;
  switch (g.kind) {
    case AK_MC1:
      call = "jf_mc1<R, RB, " + std::to_string(g.ra) + ", " + hex32(g.creg) + ">(re, im, " + mp + ");";
      break;
    case AK_MC2:
      call = "jf_mc2<R, RB, " + std::to_string(g.hi) + ", " + std::to_string(g.lo) + ", " + hex32(g.creg) + ", " +
             (g.ra < g.rb ? "true" : "false") + ">(re, im, " + mp + ");";
      break;
    case AK_MCX:
      call = "ren_x<R, " + nr + ", " + std::to_string(g.ra) + ", " + hex32(g.creg) + ">(re, im);";
      break;
    case AK_MCSWAP:
      call = "ren_swap<R, " + nr + ", " + std::to_string(g.hi) + ", " + std::to_string(g.lo) + ", " + hex32(g.creg) +
             ">(re, im);";
      break;
    case AK_DIAG:  // controls are part of the diagonal
      o << "    jf_diag<R, RB, " << g.k << ", " << g.w0 << ", " << g.w1 << ">(re, im, " << mp << ", L, base);\n";
      return;
    case AK_DENSEK:  // no controls; contains barriers, so every thread enters
      o << "    if constexpr (RB >= 3) dense3<R, RB>(re, im, " << mp << ", 0, L, lds);\n";
      return;
    default:
      return;
  }
  if (g.ctl())
    o << "    if (" << g.ok() << ") " << call << "\n";
  else
    o << "    " << call << "\n";
}

void emit_bwd_gate(std::ostringstream& o, const GateLit& g) {
  const std::string mp = "pool + " + std::to_string(g.mat);
  const std::string ip = "pool + " + std::to_string(g.inv >= 0 ? g.inv : g.mat);
  const std::string inv = g.inv >= 0 ? "true" : "false";
  const std::string nr = "(1 << RB)";
  const std::string st = std::to_string(g.slot);
  switch (g.kind) {
    case AK_MC1:
      o << "    jb_mc1<R, RB, " << g.ra << ", " << hex32(g.creg) << ", " << inv << ", " << (g.quat ? "true" : "false")
        << ">(pr, pi, lr, li, " << mp << ", " << ip << ", " << g.ok() << ", slots, " << st << ");\n";
      return;
    case AK_MC2:
      o << "    jb_mc2<R, RB, " << g.hi << ", " << g.lo << ", " << hex32(g.creg) << ", " << inv << ", "
        << (g.ra < g.rb ? "true" : "false") << ">(pr, pi, lr, li, " << mp << ", " << ip << ", " << g.ok()
        << ", slots, " << st << ");\n";
      return;
    case AK_MCX: {
      const std::string t = "<R, " + nr + ", " + std::to_string(g.ra) + ", " + hex32(g.creg) + ">";
      o << "    if (" << g.ok() << ") { ren_x" << t << "(pr, pi); ren_x" << t << "(lr, li); }\n";
      return;
    }
    case AK_MCSWAP: {
      const std::string t =
          "<R, " + nr + ", " + std::to_string(g.hi) + ", " + std::to_string(g.lo) + ", " + hex32(g.creg) + ">";
      o << "    if (" << g.ok() << ") { ren_swap" << t << "(pr, pi); ren_swap" << t << "(lr, li); }\n";
      return;
    }
    case AK_DIAG:
      o << "    jb_diag<R, RB, " << g.k << ", " << g.w0 << ", " << g.w1 << ", " << inv << ">(pr, pi, lr, li, " << mp
        << ", " << ip << ", L, base, slots, " << st << ");\n";
      return;
    case AK_DENSEK:
      o << "    if constexpr (RB >= 3) jb_dense3<R, RB, " << inv << ">(pr, pi, lr, li, " << mp << ", " << ip
        << ", L, lds);\n";
      return;
    default:
      return;
  }
}

// Smallest matrix offset referenced by pass q (its gates' offsets are
// written relative to it).
int pass_rebase(const ProgView& pv, int q) {
  int m = -1;
  const int nsub = pv.pass(q)[0];
  for (int s = 0; s < nsub; ++s) {
    const int32_t* sub = pv.sub(q, s);
    const int32_t* g = sub + 1 + pv.T;
    for (int k = 0; k < sub[0]; ++k, g += GW) {
      const int kind = g[G_KIND];
      if (kind == AK_NOP || kind == AK_MCX || kind == AK_MCSWAP) continue;
      for (int v : {g[G_MAT], g[G_INV]})
        if (v >= 0 && (m < 0 || v < m)) m = v;
    }
  }
  return m < 0 ? 0 : m;
}

// Layout words (`first` / `last`, with the tile bit of each register and
// thread bit) at the ends of a pass body, empty for the canonical layout.
struct Ends {
  std::string first, last;
  const int32_t* first_sub = nullptr;
  const int32_t* last_sub = nullptr;
};

// Forward pass body.  `e.first` receives the layout of the first substep
// when the pass starts with a re-layout; `o` receives the statements after
// that re-layout.
void emit_forward(std::ostringstream& o, Ends& e, const ProgView& pv, int q, int rebase) {
  std::vector<int> cur = canonical_rpos(pv);
  const int nsub = pv.pass(q)[0];
  e = Ends();
  for (int s = 0; s < nsub; ++s) {
    const int32_t* sub = pv.sub(q, s);
    if (!same_rpos(pv, sub, cur)) {
      if (s == 0) {
        e.first = layout_args(pv, sub);
        e.first_sub = sub;
      } else {
        o << "    jit_relayout<R, RB, LNT" << layout_args(pv, sub) << ">(re, im, lds, L);\n";
      }
      e.last = layout_args(pv, sub);
      e.last_sub = sub;
      for (int b = 0; b < pv.RB; ++b) cur[b] = sub[1 + b];
    }
    const int32_t* g = sub + 1 + pv.T;
    for (int k = 0; k < sub[0]; ++k, g += GW) emit_fwd_gate(o, GateLit(g, rebase));
  }
}

// How a tile can move between HBM and registers under the layout of `sub`:
// 1 when register bit 0 and thread bits 0, 1 hold tile bits 0, 1, 2 (16-byte
// accesses, 4 lanes per 64 contiguous bytes), 2 when thread bits 0, 1, 2 do
// (8-byte accesses, 8 lanes per 64 bytes), else 0 (exchange through LDS into
// the canonical layout).  Tile bits 0..2 must be global bits 0..2 in every
// pass (`lowrun`).
int access_mode(const ProgView& pv, const int32_t* sub, bool lowrun) {
  if (sub == nullptr || !lowrun || pv.T - pv.RB < 3) return 0;
  const int32_t* w = sub + 1;
  const int rb = pv.RB;
  if (w[0] == 0 && w[rb] == 1 && w[rb + 1] == 2) return 1;
  if (w[rb] == 0 && w[rb + 1] == 1 && w[rb + 2] == 2) return 2;
  return 0;
}

// layout words of the canonical layout (register bits 0, T-RB+1 .. T-1;
// thread bit j <-> tile bit 1 + j)
std::string canonical_args(const ProgView& pv) {
  std::string o;
  for (int r : canonical_rpos(pv)) o += ", " + std::to_string(r);
  for (int j = 0; j < pv.T - pv.RB; ++j) o += ", " + std::to_string(1 + j);
  return o;
}

// Backward pass body (substeps and gates in reverse); `e` and `o` as for
// the forward.
void emit_backward(std::ostringstream& o, Ends& e, const ProgView& pv, int q, int rebase) {
  std::vector<int> cur = canonical_rpos(pv);
  const int nsub = pv.pass(q)[0];
  e = Ends();
  for (int s = nsub - 1; s >= 0; --s) {
    const int32_t* sub = pv.sub(q, s);
    if (!same_rpos(pv, sub, cur)) {
      if (s == nsub - 1) {
        e.first = layout_args(pv, sub);
        e.first_sub = sub;
      } else {
        o << "    jit_relayout2<R, RB, LNT" << layout_args(pv, sub) << ">(pr, pi, lr, li, lds, L);\n";
      }
      e.last = layout_args(pv, sub);
      e.last_sub = sub;
      for (int b = 0; b < pv.RB; ++b) cur[b] = sub[1 + b];
    }
    const int32_t* g0 = sub + 1 + pv.T;
    for (int k = sub[0] - 1; k >= 0; --k) emit_bwd_gate(o, GateLit(g0 + k * GW, rebase));
  }
}

ProgView view(const std::vector<int32_t>& w) {
  ProgView pv;
  pv.W = w.data();
  pv.n_passes = w[1];
  pv.T = w[2];
  pv.RB = w[3];
  return pv;
}

}  // namespace

std::string jit_kernel_name(int dir, int id) {
  return std::string(dir == 0 ? "jq_sv_forward_s" : "jq_sv_backward_s") + std::to_string(id);
}

void jit_plan(const JitSpec& spec, JitPlan& out) {
  out = JitPlan();
  // Direct HBM access under substep layouts needs full tiles, 32-bit byte
  // offsets (2^27 amplitudes at most) and tile bits 0..2 = global bits 0..2
  // in every pass of both programs.
  const ProgView pf = view(*spec.words_f);
  const bool full = pf.W[4] >= pf.T;
  bool lowrun = full && pf.W[4] <= 27 && pf.T >= 3;
  for (int dir = 0; dir < 2 && lowrun; ++dir) {
    const ProgView pv = view(dir == 0 ? *spec.words_f : *spec.words_b);
    for (int q = 0; q < pv.n_passes && lowrun; ++q) {
      const int32_t* tb = pv.pass(q) + 4;
      lowrun = tb[0] == 0 && tb[1] == 1 && tb[2] == 2;
    }
  }
  for (int dir = 0; dir < 2; ++dir) {
    const ProgView pv = view(dir == 0 ? *spec.words_f : *spec.words_b);
    for (int q = 0; q < pv.n_passes; ++q) {
      if (pv.pass(q)[1] & PF_WIDE) {  // wide-block passes run their own kernels
        out.kernel_of[dir].push_back(-1);
        out.pool_off[dir].push_back(0);
        continue;
      }
      const int rebase = pass_rebase(pv, q);
      std::ostringstream o;
      Ends e;
      if (dir == 0)
        emit_forward(o, e, pv, q, rebase);
      else
        emit_backward(o, e, pv, q, rebase);
      const std::string body0 = o.str();
      const std::string relayout = dir == 0 ? "    jit_relayout<R, RB, LNT" + e.first + ">(re, im, lds, L);\n"
                                            : "    jit_relayout2<R, RB, LNT" + e.first + ">(pr, pi, lr, li, lds, L);\n";
      const std::string body = e.first.empty() ? body0 : relayout + body0;
      int id = -1;
      for (size_t k = 0; k < out.kernels.size(); ++k)
        if (out.kernels[k].dir == dir && out.kernels[k].body == body) id = (int)k;
      if (id < 0) {
        id = (int)out.kernels.size();
        JitPlan::Kernel k{dir, body, e.first, body0, e.last};
        k.full = full;
        // (the forward reads directly under its first layout when it can,
        // else through LDS: sv_forward_dma)
        k.load = access_mode(pv, e.first_sub, lowrun);
        k.store = access_mode(pv, e.last_sub, lowrun);
        out.kernels.push_back(k);
      }
      out.kernel_of[dir].push_back(id);
      out.pool_off[dir].push_back(rebase);
    }
  }
}

std::string jit_key(const JitSpec& spec, const JitPlan& plan, int id) {
  const JitPlan::Kernel& k = plan.kernels[id];
  const ProgView pv = view(k.dir == 0 ? *spec.words_f : *spec.words_b);
  return std::to_string(k.dir) + " " + std::to_string(spec.dtype) + " " + std::to_string(pv.T) + " " +
         std::to_string(pv.RB) + " " + std::to_string(k.dir == 1 && spec.persist_bwd ? 1 : 0) + " " +
         std::to_string(spec.dma ? 1 : 0) + " " + std::to_string(k.full ? 1 : 0) + " " + std::to_string(k.load) +
         " " + std::to_string(k.store) + "\n" + k.body;
}

// `body` with every exchange `name<` rewritten to `name_split<LEAD, `: no
// leading barrier for the first (a barrier precedes the body), one for the
// others; a dense 3-wire gate after an exchange gets a barrier of its own.
std::string split_exchanges(const std::string& b, const std::string& name) {
  std::string out;
  bool first = true;
  size_t pos = 0;
  const std::string from = name + "<";
  while (pos < b.size()) {
    const size_t nl = b.find('\n', pos);
    std::string line = b.substr(pos, nl == std::string::npos ? std::string::npos : nl - pos + 1);
    const size_t at = line.find(from);
    if (at != std::string::npos) {
      line.replace(at, from.size(), name + "_split<" + (first ? "false" : "true") + ", ");
      first = false;
    } else if (line.find("dense3") != std::string::npos && !first) {
      out += "    __syncthreads();\n";
    }
    out += line;
    if (nl == std::string::npos) break;
    pos = nl + 1;
  }
  return out;
}

// `body` with its last occurrence of `from` replaced by `to`; unchanged when
// absent or when a dense 3-wire gate (which exchanges through the same LDS
// region) follows it.
std::string replace_last(std::string body, const std::string& from, const std::string& to) {
  const size_t at = body.rfind(from);
  if (at != std::string::npos && body.find("dense3", at) == std::string::npos) body.replace(at, from.size(), to);
  return body;
}

std::string jit_source(const JitSpec& spec, const JitPlan& plan, const std::vector<int>& ids,
                       const std::vector<std::string>& names) {
  std::ostringstream o;
  o << kPrelude;
  const char* rt = spec.dtype == 0 ? "float" : "double";
  const ProgView f = view(*spec.words_f), b = view(*spec.words_b);
  const int occ = spec.dtype == 0 ? 4 : 2;  // the interpreter's forward launch bounds
  for (size_t j = 0; j < ids.size(); ++j) {
    const int id = (int)j;  // struct names are local to this source
    const JitPlan::Kernel& k = plan.kernels[ids[j]];
    const std::string& kname = names[j];
    const ProgView& pv = k.dir == 0 ? f : b;
    const int nt = 1 << (pv.T - pv.RB);
    if (k.dir == 0) {
      const bool dma = spec.dma && spec.dtype == 0;
      const char* sig =
          "(cint*, int, R* re, R* im, Layout& L, uint64_t base, C2<R>* lds, cmat<R>* pool) {\n"
          "    constexpr int LNT = __builtin_ctz(NT);\n";
      // one tile per workgroup: every exchange after the first opens with the
      // barrier the previous one would have closed with (a store through the
      // canonical layout opens with one too)
      const std::string fbody = dma ? split_exchanges(k.body, "jit_relayout") : k.body;
      const std::string fbody0 = dma ? split_exchanges(k.body0, "jit_relayout") : k.body0;
      o << "struct JF" << id << " {\n"
        << "  static constexpr bool kFull = " << (k.full ? "true" : "false") << ";\n"
        << "  static constexpr int kStore = " << (dma ? k.store : 0) << ";\n"
        << "  static constexpr int kLoad = " << (dma ? k.load : 0) << ";\n"
        << "  static constexpr bool kSplit = " << (dma ? "true" : "false") << ";\n"
        << "  template <int RB, int LNT>\n"
        << "  static __device__ __forceinline__ void last_layout(cint* tb, Layout& L) {\n"
        << "    layout_lit_g<RB, LNT" << (k.last.empty() ? canonical_args(pv) : k.last) << ">(tb, L);\n  }\n"
        << "  template <typename R, int RB, int NT, bool FULL>\n"
        << "  static __device__ __forceinline__ void fwd" << sig << fbody
        << "    (void)lds; (void)base; (void)pool;\n  }\n"
        << "  template <typename R, int RB, int NT, bool FULL>\n"
        << "  static __device__ __forceinline__ void fwd0" << sig << fbody0
        << "    (void)lds; (void)base; (void)pool;\n  }\n"
        << "  template <int RB, int LNT>\n"
        << "  static __device__ __forceinline__ void first_layout(cint* tb, Layout& L) {\n"
        << "    layout_lit<RB, LNT" << (k.first.empty() ? canonical_args(pv) : k.first) << ">(tb, L);\n  }\n"
        << "  template <int RB, int LNT>\n"
        << "  static __device__ __forceinline__ void first_layout_g(cint* tb, Layout& L) {\n"
        << "    layout_lit_g<RB, LNT" << (k.first.empty() ? canonical_args(pv) : k.first) << ">(tb, L);\n  }\n};\n";
      if (dma)
        o << "extern \"C\" __global__ __launch_bounds__(" << nt << ", " << occ << ") void " << kname
          << "(SvArgs a) {\n  sv_forward_dma<" << rt << ", " << pv.RB << ", " << nt << ", JF" << id
          << ">(a);\n}\n\n";
      else
        o << "extern \"C\" __global__ __launch_bounds__(" << nt << ", " << occ << ") void " << kname
          << "(SvArgs a) {\n  sv_forward_impl<" << rt << ", " << pv.RB << ", " << nt << ", false, JF" << id
          << ">(a);\n}\n\n";
    } else {
      const char* sig =
          "(cint*, int, R* pr, R* pi, R* lr, R* li, Layout& L, uint64_t base, C2<R>* lds, cmat<R>* pool, "
          "double* slots) {\n    constexpr int LNT = __builtin_ctz(NT);\n";
      // (the persistent loop and the one-tile-per-workgroup kernel of
      // narrower workgroups both read and write under the body's end layouts;
      // the 1024-thread one-tile kernel goes through the canonical layout)
      const bool ends = spec.persist_bwd || nt < 1024;
      const int load = ends ? k.load : 0, store = ends ? k.store : 0;
      // persistent loop: every exchange after the first of a body opens with
      // the barrier that the previous one would have closed with
      const bool split = spec.persist_bwd;
      const std::string bbody = split ? split_exchanges(k.body, "jit_relayout2")
                                      : (store != 0 ? replace_last(k.body, "jit_relayout2<", "jit_relayout2_nt<") : k.body);
      const std::string bbody0 = split ? split_exchanges(k.body0, "jit_relayout2")
                                       : (store != 0 ? replace_last(k.body0, "jit_relayout2<", "jit_relayout2_nt<") : k.body0);
      o << "struct JB" << id << " {\n"
        << "  static constexpr bool kFull = " << (k.full ? "true" : "false") << ";\n"
        << "  static constexpr int kLoad = " << load << ";\n"
        << "  static constexpr int kStore = " << store << ";\n"
        << "  static constexpr bool kSplit = " << (split ? "true" : "false") << ";\n"
        << "  template <int RB, int LNT>\n"
        << "  static __device__ __forceinline__ void first_layout(cint* tb, Layout& L) {\n"
        << "    layout_lit_g<RB, LNT" << (k.first.empty() ? canonical_args(pv) : k.first) << ">(tb, L);\n  }\n"
        << "  template <int RB, int LNT>\n"
        << "  static __device__ __forceinline__ void last_layout(cint* tb, Layout& L) {\n"
        << "    layout_lit_g<RB, LNT" << (k.last.empty() ? canonical_args(pv) : k.last) << ">(tb, L);\n  }\n"
        << "  template <typename R, int RB, int NT, bool FULL>\n"
        << "  static __device__ __forceinline__ void bwd" << sig << bbody
        << "    (void)lds; (void)base; (void)pool; (void)slots;\n  }\n";
      if (load != 0)
        o << "  template <typename R, int RB, int NT, bool FULL>\n"
          << "  static __device__ __forceinline__ void bwd0" << sig << bbody0
          << "    (void)lds; (void)base; (void)pool; (void)slots;\n  }\n";
      o << "};\n"
        << "extern \"C\" __global__ __launch_bounds__(" << nt << (nt < 512 ? ", 2" : "") << ") void " << kname
        << "(SvArgs a) {\n  "
        << (spec.persist_bwd ? "sv_backward_loop<" : ends ? "sv_backward_direct<" : "sv_backward_impl<") << rt << ", "
        << pv.RB << ", " << nt << (spec.persist_bwd || ends ? "" : ", false") << ", JB" << id << ">(a);\n}\n\n";
    }
  }
  o << "}  // namespace jq\n";
  return o.str();
}
bool jit_compile(const std::string& src, std::vector<char>& code, std::string& log) {
  hiprtcProgram prog;
  const int nh = (int)(sizeof(kJitHeaderNames) / sizeof(kJitHeaderNames[0]));
  hiprtcResult r = hiprtcCreateProgram(&prog, src.c_str(), "jaqsi_amd_plan.hip", nh, kJitHeaderTexts, kJitHeaderNames);
  if (r != HIPRTC_SUCCESS) {
    log = std::string("hiprtcCreateProgram: ") + hiprtcGetErrorString(r);
    return false;
  }
  // packed FP32 instructions stay off, as for the prebuilt kernels
  const char* opts[] = {"--offload-arch=gfx950", "-O3",   "-std=c++17",       "-Xclang",
                        "-target-feature",       "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"};
  r = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
  size_t ls = 0;
  hiprtcGetProgramLogSize(prog, &ls);
  log.assign(ls, '\0');
  if (ls) hiprtcGetProgramLog(prog, &log[0]);
  bool ok = r == HIPRTC_SUCCESS;
  if (ok) {
    size_t cs = 0;
    ok = hiprtcGetCodeSize(prog, &cs) == HIPRTC_SUCCESS && cs > 0;
    if (ok) {
      code.resize(cs);
      ok = hiprtcGetCode(prog, code.data()) == HIPRTC_SUCCESS;
    }
  } else {
    log = std::string("hiprtcCompileProgram: ") + hiprtcGetErrorString(r) + "\n" + log;
  }
  hiprtcDestroyProgram(&prog);
  return ok;
}

}  // namespace jq
