// runtime.hip -- host side of the MI355X hot path behind the C-ABI in
// include/necstar_hip.h.  It mirrors the reference's QuantumState facade
// (crates/necstar/src/state/mod.rs) and compiler
// (crates/necstar/src/state/compiler/mod.rs) with the per-term arithmetic
// on the GPU: CH-form tables stay device resident; the host only keeps exact
// coefficient scalars, gate lists, and the RNG draws of sampling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "internal.hpp"
#include "necstar_hip.h"
#include "rng.hpp"
#include "random_clifford.hpp"

using cplx = std::complex<double>;

namespace nx {

// ---------------------------------------------------------------------------
// errors
thread_local std::string g_err;

struct NxError {
  int code;
};

[[noreturn]] static void fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  throw NxError{code};
}

static void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(NX_E_DEVICE, "HIP error in %s: %s", what, hipGetErrorString(e));
}

// reference Display strings (crates/necstar/src/error/mod.rs,
// crates/stabilizer-ch-form-rust/src/error/mod.rs)
[[noreturn]] static void err_oob(int q, int n) {
  fail(NX_E_VALUE, "Qubit index %d is out of bounds for %d qubits.", q, n);
}
[[noreturn]] static void err_dup_gate(int q) {
  fail(NX_E_VALUE, "Two-qubit gate requires two different qubit indices, but got the same: %d.", q);
}
static const char* gate_name(int op) {
  static const char* names[] = {"H", "X", "Y", "Z", "S", "Sdg", "SqrtX", "SqrtXdg", "CX", "CZ", "Swap",
                                "T", "Tdg", "CCX"};
  return (op >= 0 && op <= 13) ? names[op] : "?";
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free norm (nx_state_norm_noalloc): the one
// device buffer that entry point uses, one copy per GPU, held until the
// process ends.  Host code never passes this name to a kernel or a copy: it
// is a host-side stand-in, not a device address; each device's address is
// looked up once (hipGetSymbolAddress) when its context is set up.
static constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

// device allocations requested from the library's allocator (pool hits
// included) and page-locked host allocations, for tests of the
// allocation-free entry point
static std::atomic<int64_t> g_alloc_calls{0};

// ---------------------------------------------------------------------------
// device context: stream, FP64 tables, caching allocator, scratch address
struct DeviceCtx {
  int dev = 0;
  hipStream_t stream = nullptr;
  double* tabs = nullptr;
  unsigned char* scratch = nullptr;  // this device's copy of `scratch`
  // ordering against a caller's stream (asynchronous entry points)
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  std::mutex mu;
  std::map<size_t, std::vector<void*>> free_lists;
};

static std::mutex g_ctx_mu;
static std::map<int, DeviceCtx*> g_ctx;
static thread_local int t_dev = 0;

static DeviceCtx* ctx() {
  std::lock_guard<std::mutex> lk(g_ctx_mu);
  auto it = g_ctx.find(t_dev);
  if (it != g_ctx.end()) {
    hip_check(hipSetDevice(t_dev), "hipSetDevice");
    return it->second;
  }
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= t_dev)
    fail(NX_E_DEVICE, "no HIP device %d available (the necstar HIP path needs an MI355X / gfx950)", t_dev);
  hip_check(hipSetDevice(t_dev), "hipSetDevice");
  hipDeviceProp_t prop;
  hip_check(hipGetDeviceProperties(&prop, t_dev), "hipGetDeviceProperties");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    fail(NX_E_DEVICE, "device %d is %s; this library is built for gfx950 only", t_dev, prop.gcnArchName);
  DeviceCtx* c = new DeviceCtx();
  c->dev = t_dev;
  hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
  hip_check(hipEventCreateWithFlags(&c->ev_in, hipEventDisableTiming), "hipEventCreate");
  hip_check(hipEventCreateWithFlags(&c->ev_out, hipEventDisableTiming), "hipEventCreate");
  // PhaseFactor::to_complex (phase_factor.rs:19-22) and powf(2, -r/2)
  // (scalar.rs:30-38): computed with the same libm calls on the host
  std::vector<double> t(TAB_SIZE);
  for (int k = 0; k < 8; ++k) {
    const double angle = (double)k * M_PI_4;
    t[TAB_COS + k] = std::cos(angle);
    t[TAB_SIN + k] = std::sin(angle);
  }
  for (int r = -TAB_RMAX; r <= TAB_RMAX; ++r) t[TAB_POW + r] = std::pow(2.0, -((double)r) / 2.0);
  hip_check(hipMalloc(&c->tabs, sizeof(double) * TAB_SIZE), "hipMalloc");
  hip_check(hipMemcpy(c->tabs, t.data(), sizeof(double) * TAB_SIZE, hipMemcpyHostToDevice), "hipMemcpy");
  void* sp = nullptr;
  hip_check(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  c->scratch = (unsigned char*)sp;
  g_ctx[t_dev] = c;
  return c;
}

static size_t size_class(size_t b) {
  size_t c = 256;
  while (c < b) c <<= 1;
  return c;
}

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  DeviceCtx* c = nullptr;
  DevBuf() = default;
  explicit DevBuf(size_t bytes) { alloc(bytes); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), cap(o.cap), c(o.c) { o.p = nullptr; o.cap = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p; cap = o.cap; c = o.c;
      o.p = nullptr; o.cap = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t bytes) {
    release();
    if (bytes == 0) return;
    g_alloc_calls.fetch_add(1, std::memory_order_relaxed);
    c = ctx();
    cap = size_class(bytes);
    {
      std::lock_guard<std::mutex> lk(c->mu);
      auto& fl = c->free_lists[cap];
      if (!fl.empty()) {
        p = fl.back();
        fl.pop_back();
        return;
      }
    }
    if (hipMalloc(&p, cap) != hipSuccess) {
      p = nullptr;
      cap = 0;
      fail(NX_E_DEVICE, "out of device memory (%zu bytes)", bytes);
    }
  }
  void release() {
    if (p && c) {
      std::lock_guard<std::mutex> lk(c->mu);
      c->free_lists[cap].push_back(p);
    }
    p = nullptr;
    cap = 0;
  }
  template <class T>
  T* as() const { return (T*)p; }
};

// page-locked host memory (staging for small per-pass uploads)
struct PinnedBuf {
  void* p = nullptr;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
  void alloc(size_t bytes) {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    g_alloc_calls.fetch_add(1, std::memory_order_relaxed);
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
      p = nullptr;
      fail(NX_E_DEVICE, "out of pinned host memory (%zu bytes)", bytes);
    }
  }
  template <class T>
  T* as(size_t off = 0) const {
    return (T*)((char*)p + off);
  }
};

// The stream the calling thread's work goes to: the device context's own
// (non-blocking) stream, or the caller's stream set with nx_set_stream.
static thread_local bool t_user_stream = false;
static thread_local hipStream_t t_stream = nullptr;
static hipStream_t stream() {
  DeviceCtx* c = ctx();
  return t_user_stream ? t_stream : c->stream;
}

// Orders the library's work after the caller's earlier work on `us` (and
// the caller's later work after the library's) when they are different
// streams; nothing to do when `us` is the library's stream for this thread.
struct StreamOrder {
  DeviceCtx* c;
  hipStream_t us, lib;
  StreamOrder(void* user) : c(ctx()), us((hipStream_t)user), lib(stream()) {
    if (us == lib) return;
    hip_check(hipEventRecord(c->ev_in, us), "hipEventRecord");
    hip_check(hipStreamWaitEvent(lib, c->ev_in, 0), "hipStreamWaitEvent");
  }
  void done() {
    if (us == lib) return;
    hip_check(hipEventRecord(c->ev_out, lib), "hipEventRecord");
    hip_check(hipStreamWaitEvent(us, c->ev_out, 0), "hipStreamWaitEvent");
  }
};
static void sync() { hip_check(hipStreamSynchronize(stream()), "hipStreamSynchronize"); }
static void h2d(void* d, const void* h, size_t b) {
  if (b) hip_check(hipMemcpyAsync(d, h, b, hipMemcpyHostToDevice, stream()), "hipMemcpyAsync H2D");
}
static void d2h(void* h, const void* d, size_t b) {
  if (b) hip_check(hipMemcpyAsync(h, d, b, hipMemcpyDeviceToHost, stream()), "hipMemcpyAsync D2H");
}
static void d2d(void* d, const void* s, size_t b) {
  if (b) hip_check(hipMemcpyAsync(d, s, b, hipMemcpyDeviceToDevice, stream()), "hipMemcpyAsync D2D");
}

// kernel timing (HIP events on the library stream)
// Per-kernel timing with HIP events on the library stream.  Scopes record an
// event pair and do not wait; the pairs are read (and recycled) when the
// totals are asked for, so profiled runs keep the stream's own pacing.
struct Prof {
  bool on = false;
  std::string only;  // ",name,name," : only these scopes are timed (empty: all)
  struct Rec {
    std::string name;
    hipEvent_t e0, e1;
  };
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
  std::vector<Rec> pending;
  std::map<std::string, std::pair<double, int64_t>> acc;
  // caller holds g_prof_mu
  void collect() {
    for (Rec& r : pending) {
      hip_check(hipEventSynchronize(r.e1), "hipEventSynchronize");
      float ms = 0.f;
      hip_check(hipEventElapsedTime(&ms, r.e0, r.e1), "hipEventElapsedTime");
      auto& a = acc[r.name];
      a.first += ms;
      a.second += 1;
      pool.push_back({r.e0, r.e1});
    }
    pending.clear();
  }
};
static std::mutex g_prof_mu;
static Prof g_prof;

// ext: the timed launch records the pair itself (start_ev() and end_ev() go
// to a launch that takes timing events), so the stream carries no event
// records of its own
struct ProfScope {
  const char* name;
  bool on;
  bool ext;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  explicit ProfScope(const char* nm, bool ext_ = false) : name(nm), on(false), ext(ext_) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if (!g_prof.on) return;
    if (!g_prof.only.empty() && g_prof.only.find("," + std::string(nm) + ",") == std::string::npos) return;
    on = true;
    if (g_prof.pool.empty()) {
      hipEvent_t a, b;
      hip_check(hipEventCreate(&a), "hipEventCreate");
      hip_check(hipEventCreate(&b), "hipEventCreate");
      g_prof.pool.push_back({a, b});
    }
    e0 = g_prof.pool.back().first;
    e1 = g_prof.pool.back().second;
    g_prof.pool.pop_back();
    if (!ext) hip_check(hipEventRecord(e0, stream()), "hipEventRecord");
  }
  hipEvent_t start_ev() const { return on && ext ? e0 : nullptr; }
  hipEvent_t end_ev() const { return on && ext ? e1 : nullptr; }
  // after the launch (ext: only once it was issued with start_ev / end_ev)
  void end() {
    if (!on) return;
    on = false;
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if (!ext) hip_check(hipEventRecord(e1, stream()), "hipEventRecord");
    g_prof.pending.push_back({name, e0, e1});
    if (g_prof.pending.size() >= 4096) g_prof.collect();
  }
  ~ProfScope() {
    if (!on) return;
    if (ext) {  // the launch did not happen: the events were never recorded
      std::lock_guard<std::mutex> lk(g_prof_mu);
      g_prof.pool.push_back({e0, e1});
      return;
    }
    try { end(); } catch (...) {}
  }
};

static int words_for(int n) { return n <= 64 ? 1 : (n + 63) / 64; }
static const int MAX_QUBITS = 256;

// ---------------------------------------------------------------------------
// term tables
struct Forms {
  DevBuf G, F, M, g, v, s, ph;
  int64_t chi = 0;
  int n = 0, W = 1;
  void alloc(int64_t chi_, int n_) {
    chi = chi_;
    n = n_;
    W = words_for(n_);
    const size_t R = 64 * (size_t)W;
    G.alloc(chi * R * W * 8);
    F.alloc(chi * R * W * 8);
    M.alloc(chi * R * W * 8);
    g.alloc(chi * R);
    v.alloc(chi * W * 8);
    s.alloc(chi * W * 8);
    ph.alloc(chi);
  }
  FormsDev desc() const {
    FormsDev d;
    d.G = G.as<uint64_t>(); d.F = F.as<uint64_t>(); d.M = M.as<uint64_t>();
    d.g = g.as<uint8_t>(); d.v = v.as<uint64_t>(); d.s = s.as<uint64_t>(); d.ph = ph.as<uint8_t>();
    d.chi = chi; d.n = n; d.W = W;
    return d;
  }
};

struct QForms {
  DevBuf Q, B, C, CT, a, alo, ahi, rhs, nc, k, p0, meta;
  // the amplitude kernel's per-term tables, built on first use after the
  // forms were (re)computed
  DevBuf T;
  bool t_ok = false;
  int64_t chi = 0;
  int n = 0, W = 1;
  // the pair kernels index rows with 32-bit word offsets
  static void check_size(int64_t chi_, int W_) {
    if ((uint64_t)chi_ * 64u * (uint64_t)W_ * (uint64_t)W_ >= (1ull << 32))
      fail(NX_E_VALUE, "stabilizer rank %lld at %d words per row exceeds the device table limit",
           (long long)chi_, W_);
  }
  void alloc(int64_t chi_, int n_) {
    chi = chi_;
    n = n_;
    W = words_for(n_);
    check_size(chi, W);
    const size_t R = 64 * (size_t)W;
    Q.alloc(chi * R * W * 8);
    B.alloc(chi * R * W * 8);
    C.alloc(chi * R * W * 8);
    CT.alloc(chi * R * W * 8);
    a.alloc(chi * R);
    alo.alloc(chi * W * 8);
    ahi.alloc(chi * W * 8);
    rhs.alloc(chi * W * 8);
    nc.alloc(chi * 4);
    k.alloc(chi * 4);
    p0.alloc(chi * 4);
    meta.alloc(chi * 16);
  }
  // only the fields the pair kernels read (projected forms)
  void alloc_pair(int64_t chi_, int n_) {
    chi = chi_;
    n = n_;
    W = words_for(n_);
    check_size(chi, W);
    const size_t R = 64 * (size_t)W;
    B.alloc(chi * R * W * 8);
    C.alloc(chi * R * W * 8);
    a.alloc(chi * R);
    rhs.alloc(chi * W * 8);
    nc.alloc(chi * 4);
    k.alloc(chi * 4);
    p0.alloc(chi * 4);
  }
  QFormDev desc() const {
    QFormDev d;
    d.Q = Q.as<uint64_t>(); d.B = B.as<uint64_t>(); d.C = C.as<uint64_t>(); d.CT = CT.as<uint64_t>();
    d.a = a.as<uint8_t>();
    d.alo = alo.as<uint64_t>(); d.ahi = ahi.as<uint64_t>(); d.rhs = rhs.as<uint64_t>();
    d.nc = nc.as<int32_t>(); d.k = k.as<int32_t>(); d.p0 = p0.as<int32_t>();
    d.meta = meta.as<int4>();
    d.chi = chi; d.n = n; d.W = W;
    return d;
  }
};

static Forms init_zero(int64_t chi, int n) {
  Forms f;
  f.alloc(chi, n);
  hip_check(launch_init_zero(f.desc(), stream()), "k_init_zero");
  return f;
}

// Runs a program on every term.  n_out: qubit count after the program.
static Forms run_program(Forms& in, const std::vector<ProgOp>& prog, int n_out, std::vector<int4>* status) {
  Forms out;
  const bool inplace = words_for(n_out) == in.W;
  if (!inplace) out.alloc(in.chi, n_out);
  DevBuf dprog(prog.size() * sizeof(ProgOp));
  DevBuf dstat(in.chi * sizeof(int4));
  h2d(dprog.p, prog.data(), prog.size() * sizeof(ProgOp));
  FormsDev od = inplace ? in.desc() : out.desc();
  {
    ProfScope ps("k_program");
    hip_check(launch_program(in.desc(), od, dprog.as<ProgOp>(), (int)prog.size(), dstat.as<int4>(), stream()),
              "k_program");
    ps.end();
  }
  std::vector<int4> st(in.chi);
  d2h(st.data(), dstat.p, in.chi * sizeof(int4));
  sync();
  for (int64_t i = 0; i < in.chi; ++i) {
    if (st[i].w == DEV_ERR_SAME_VEC_MINUS) fail(NX_E_INTERNAL, "Inconsistent state: superposition with -1 coefficient.");
    if (st[i].w == DEV_ERR_G_ROW) fail(NX_E_INTERNAL, "Cannot zero G row due to G matrix state.");
    if (st[i].w == DEV_ERR_G_COL) fail(NX_E_INTERNAL, "Cannot zero G column due to v vector state.");
    if (st[i].w != 0 && st[i].w != DEV_ERR_CANNOT_DISCARD) fail(NX_E_INTERNAL, "device program error %d", st[i].w);
  }
  if (status) *status = std::move(st);
  if (inplace) {
    in.n = n_out;
    return std::move(in);
  }
  out.n = n_out;
  return out;
}

static Forms gather(const Forms& in, const std::vector<int64_t>& idx) {
  Forms out;
  out.alloc((int64_t)idx.size(), in.n);
  if (idx.empty()) return out;
  DevBuf didx(idx.size() * 8);
  h2d(didx.p, idx.data(), idx.size() * 8);
  hip_check(launch_gather(in.desc(), out.desc(), didx.as<int64_t>(), stream()), "k_gather");
  sync();
  return out;
}

static Forms clone_forms(const Forms& in) {
  Forms out;
  out.alloc(in.chi, in.n);
  const size_t R = 64 * (size_t)in.W;
  d2d(out.G.p, in.G.p, in.chi * R * in.W * 8);
  d2d(out.F.p, in.F.p, in.chi * R * in.W * 8);
  d2d(out.M.p, in.M.p, in.chi * R * in.W * 8);
  d2d(out.g.p, in.g.p, in.chi * R);
  d2d(out.v.p, in.v.p, in.chi * in.W * 8);
  d2d(out.s.p, in.s.p, in.chi * in.W * 8);
  d2d(out.ph.p, in.ph.p, in.chi);
  return out;
}

static Forms concat(const std::vector<const Forms*>& parts) {
  Forms out;
  int64_t chi = 0;
  for (auto* p : parts) chi += p->chi;
  out.alloc(chi, parts[0]->n);
  const size_t R = 64 * (size_t)out.W, W = out.W;
  int64_t off = 0;
  for (auto* p : parts) {
    d2d(out.G.as<uint8_t>() + off * R * W * 8, p->G.p, p->chi * R * W * 8);
    d2d(out.F.as<uint8_t>() + off * R * W * 8, p->F.p, p->chi * R * W * 8);
    d2d(out.M.as<uint8_t>() + off * R * W * 8, p->M.p, p->chi * R * W * 8);
    d2d(out.g.as<uint8_t>() + off * R, p->g.p, p->chi * R);
    d2d(out.v.as<uint8_t>() + off * W * 8, p->v.p, p->chi * W * 8);
    d2d(out.s.as<uint8_t>() + off * W * 8, p->s.p, p->chi * W * 8);
    d2d(out.ph.as<uint8_t>() + off, p->ph.p, p->chi);
    off += p->chi;
  }
  return out;
}

static Forms kron(const Forms& A, const Forms& B) {
  Forms out;
  out.alloc(A.chi * B.chi, A.n + B.n);
  hip_check(launch_kron(A.desc(), B.desc(), out.desc(), stream()), "k_kron");
  return out;
}

// quadratic forms of f into q (its buffers are reused when the shape
// matches); with_tables (one word per row): the amplitude tables as well, in
// the same launch
static void fill_qform(const Forms& f, QForms& q, bool with_tables = false) {
  if (q.chi != f.chi || q.n != f.n || !q.meta.p) q.alloc(f.chi, f.n);
  q.t_ok = false;
  void* tab = nullptr;
  if (with_tables && q.W == 1) {
    const size_t bytes = (size_t)q.chi * amp_table_bytes(1);
    if (!q.T.p || q.T.cap < bytes) q.T.alloc(bytes);
    tab = q.T.p;
  }
  ProfScope ps("k_qform");
  hip_check(launch_qform(f.desc(), q.desc(), stream(), tab), "k_qform");
  ps.end();
  q.t_ok = tab != nullptr;
}

static QForms make_qform(const Forms& f) {
  QForms q;
  q.alloc(f.chi, f.n);
  ProfScope ps("k_qform");
  hip_check(launch_qform(f.desc(), q.desc(), stream()), "k_qform");
  ps.end();
  return q;
}

// ---------------------------------------------------------------------------
// exact coefficients (crates/necstar/src/state/types/scalar.rs)
struct Scal {
  uint8_t z = 0, k = 0;
  int64_t r = 0;
};
static Scal smul(Scal a, Scal b) {
  Scal o;
  o.z = a.z || b.z;
  o.k = o.z ? 0 : (uint8_t)((a.k + b.k) % 8);
  o.r = o.z ? 0 : a.r + b.r;
  return o;
}
// Scalar::to_complex (scalar.rs:30-38)
static cplx scal_value(Scal s) {
  if (s.z) return cplx(0.0, 0.0);
  const double nrm = std::pow(2.0, -((double)s.r) / 2.0);
  const double angle = (double)s.k * M_PI_4;
  return cplx(std::cos(angle) * nrm, std::sin(angle) * nrm);
}

// num-complex products (no FMA, same operand order)
static cplx cmul(cplx a, cplx b) {
  return cplx(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
}

}  // namespace nx

// ---------------------------------------------------------------------------
// the state object
struct nx_state {
  int n = 0;
  nx::Forms forms;
  bool is_complex = false;
  std::vector<nx::Scal> sc;
  std::vector<cplx> cc;
  cplx g{1.0, 0.0};
  // caches (device)
  mutable bool q_ok = false;
  mutable nx::QForms q;
  mutable bool c_ok = false;
  mutable nx::DevBuf d_c, d_cs;

  int64_t chi() const { return forms.chi; }
  void invalidate() { q_ok = false; c_ok = false; }
  cplx coeff_value(int64_t i) const { return is_complex ? cc[i] : nx::scal_value(sc[i]); }
};

namespace nx {

static const QForms& qform_of(const nx_state* s) {
  if (!s->q_ok) {
    fill_qform(s->forms, s->q);
    s->q_ok = true;
  }
  return s->q;
}

// the amplitude tables of the state's forms: with the forms when they are
// computed here at one word per row, else k_amp_tables once per forms
static const void* amp_tables_of(const nx_state* s) {
  if (!s->q_ok) {
    fill_qform(s->forms, s->q, true);
    s->q_ok = true;
  }
  QForms& q = s->q;
  if (!q.t_ok) {
    const size_t bytes = (size_t)q.chi * amp_table_bytes(q.W);
    if (!q.T.p || q.T.cap < bytes) q.T.alloc(bytes);
    ProfScope ps("k_amp_tables");
    hip_check(launch_amp_tables(q.desc(), q.T.p, stream()), "k_amp_tables");
    ps.end();
    q.t_ok = true;
  }
  return q.T.p;
}

static CoefDev coef_of(const nx_state* s) {
  const int64_t chi = s->chi();
  if (!s->c_ok) {
    std::vector<double> c(2 * chi);
    for (int64_t i = 0; i < chi; ++i) {
      cplx v = s->coeff_value(i);
      c[2 * i] = v.real();
      c[2 * i + 1] = v.imag();
    }
    s->d_c.alloc(chi * 16);
    h2d(s->d_c.p, c.data(), chi * 16);
    if (!s->is_complex) {
      std::vector<int32_t> cs(chi);
      for (int64_t i = 0; i < chi; ++i) {
        if (s->sc[i].r < -TAB_RMAX / 2 || s->sc[i].r > TAB_RMAX / 2)
          fail(NX_E_INTERNAL, "coefficient exponent %lld outside the supported range", (long long)s->sc[i].r);
        cs[i] = s->sc[i].z ? -1 : (int32_t)((s->sc[i].k & 7) | ((s->sc[i].r + CS_ROFF) << 3));
      }
      s->d_cs.alloc(chi * 4);
      h2d(s->d_cs.p, cs.data(), chi * 4);
      sync();
    }
    sync();
    s->c_ok = true;
  }
  CoefDev d;
  d.scalar = s->is_complex ? 0 : 1;
  d.cs = s->d_cs.as<int32_t>();
  d.c = s->d_c.as<double2>();
  return d;
}

static const int PAIR_J = 8;


// work units (PAIR_J pairs of one fixed index each) of a pair plan
static int64_t pair_units(int64_t nfixed, int64_t nother, bool triangle) {
  return triangle ? tri_units_upto(nother, PAIR_J) : nfixed * ((nother + PAIR_J - 1) / PAIR_J);
}

// at most three rounds of resident waves (8 per SIMD on 256 CUs): with
// 8-pair units a state's pairs spread evenly (64 qubits, 729 terms: 0.197 ms,
// against 0.215 ms for 16-pair units in 16,384 waves)
static int pair_blocks(int64_t units) { return (int)std::max<int64_t>(1, std::min<int64_t>((units + 3) / 4, 6144)); }

// the sum A + sqrt(2) B of a reduced partial {A_re, A_im, B_re, B_im}
static cplx ab_value(const double* r) {
  const double s2 = 1.4142135623730951;
  return cplx(r[0] + s2 * r[2], r[1] + s2 * r[3]);
}

// sum over pairs; triangle: o in [f, nother) (nfixed == nother) else
// [0, nother)
static cplx run_pairs(const QForms& X, const QForms& Y, int64_t nfixed, int64_t nother, bool triangle, int mode,
                      const CoefDev& cx, const CoefDev& cy, uint32_t* d_exact = nullptr, int64_t ld = 0) {
  if (nfixed == 0 || nother == 0) return cplx(0.0, 0.0);
  const int nblocks = pair_blocks(pair_units(nfixed, nother, triangle));
  DevBuf partial((size_t)nblocks * 32);
  // the reduction writes the sum straight into page-locked host memory (one
  // buffer per calling thread): no copy after it
  static thread_local PinnedBuf t_res;
  if (!t_res.p) t_res.alloc(64);
  ProfScope ps("k_pairs", true);
  hip_check(launch_pairs(X.desc(), Y.desc(), nfixed, nother, triangle ? 1 : 0, PAIR_J, mode, cx, cy, ctx()->tabs,
                         partial.as<double4>(), nblocks, d_exact, ld, stream(), ps.start_ev(), ps.end_ev()),
            "k_pairs");
  ps.end();
  {
    ProfScope pr("k_reduce");
    hip_check(launch_reduce(partial.as<double4>(), nblocks, t_res.as<double4>(), stream()), "k_reduce");
    pr.end();
  }
  sync();
  return ab_value(t_res.as<double>());
}

// norm (norm.rs:32-35) into device memory d_out[0], enqueued on the library
// stream (no host wait)
static void norm_enqueue(const nx_state* s, double* d_out) {
  const double g2 = s->g.real() * s->g.real() + s->g.imag() * s->g.imag();
  if (s->chi() == 0) {
    hip_check(hipMemsetAsync(d_out, 0, 8, stream()), "hipMemsetAsync");
    return;
  }
  const QForms& q = qform_of(s);
  CoefDev c = coef_of(s);
  const int nblocks = pair_blocks(pair_units(s->chi(), s->chi(), true));
  DevBuf partial((size_t)nblocks * 32);
  ProfScope ps("k_pairs", true);
  hip_check(launch_pairs(q.desc(), q.desc(), s->chi(), s->chi(), 1, PAIR_J, /*fixed bra, herm*/ 2, c, c, ctx()->tabs,
                         partial.as<double4>(), nblocks, nullptr, 0, stream(), ps.start_ev(), ps.end_ev()),
            "k_pairs");
  ps.end();
  ProfScope pr("k_reduce");
  hip_check(launch_reduce(partial.as<double4>(), nblocks, nullptr, stream(), d_out, g2), "k_reduce");
  pr.end();
  // `partial` returns to the pool; its next user is ordered after these
  // launches on the same stream
}

// norm.rs:8-30
static double norm_squared(const nx_state* s) {
  if (s->chi() == 0) return 0.0 * std::norm(s->g);
  const QForms& q = qform_of(s);
  CoefDev c = coef_of(s);
  cplx sum = run_pairs(q, q, s->chi(), s->chi(), true, /*fixed bra, herm*/ 2, c, c);
  return sum.real() * (s->g.real() * s->g.real() + s->g.imag() * s->g.imag());
}

// The same sum with every device buffer carved out of this device's copy of
// `scratch` (no allocation of any kind): the quadratic forms of all terms,
// the coefficient copies, the per-wave partial sums and the reduced result.  The kernels, launch shape
// and summation order are those of norm_squared, so the two agree bit for
// bit.  Calls on one GPU share its scratch: calls that overlap on one GPU
// (from several threads) race.
static double norm_squared_noalloc(const nx_state* s) {
  DeviceCtx* c = ctx();
  const hipStream_t st = stream();
  const int64_t chi = s->chi();
  const double g2 = s->g.real() * s->g.real() + s->g.imag() * s->g.imag();
  if (chi == 0) return 0.0 * g2;
  const int W = s->forms.W, n = s->n;
  const size_t R = 64 * (size_t)W;
  const int nblocks = pair_blocks(pair_units(chi, chi, true));
  // layout (256-byte aligned pieces)
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t r = o;
    o += (bytes + 255) & ~(size_t)255;
    return r;
  };
  const size_t rows = (size_t)chi * R * W * 8;
  const size_t oQ = take(rows), oB = take(rows), oC = take(rows), oCT = take(rows), oa = take((size_t)chi * R);
  const size_t oalo = take((size_t)chi * W * 8), oahi = take((size_t)chi * W * 8), orhs = take((size_t)chi * W * 8);
  const size_t onc = take((size_t)chi * 4), ok = take((size_t)chi * 4), op0 = take((size_t)chi * 4);
  const size_t ometa = take((size_t)chi * 16);
  const size_t ocoef = take((size_t)chi * (s->is_complex ? 16 : 4));
  const size_t opart = take((size_t)nblocks * 32), ores = take(32);
  if (o > NOALLOC_SCRATCH_BYTES)
    fail(NX_E_VALUE,
         "nx_state_norm_noalloc: %lld terms on %d qubits need %zu bytes of scratch, more than its %zu bytes",
         (long long)chi, n, o, NOALLOC_SCRATCH_BYTES);
  unsigned char* b = c->scratch;
  QFormDev q;
  q.Q = (uint64_t*)(b + oQ); q.B = (uint64_t*)(b + oB); q.C = (uint64_t*)(b + oC); q.CT = (uint64_t*)(b + oCT);
  q.a = b + oa;
  q.alo = (uint64_t*)(b + oalo); q.ahi = (uint64_t*)(b + oahi); q.rhs = (uint64_t*)(b + orhs);
  q.nc = (int32_t*)(b + onc); q.k = (int32_t*)(b + ok); q.p0 = (int32_t*)(b + op0);
  q.meta = (int4*)(b + ometa);
  q.chi = chi; q.n = n; q.W = W;
  CoefDev cd;
  cd.scalar = s->is_complex ? 0 : 1;
  cd.cs = (const int32_t*)(b + ocoef);
  cd.c = (const double2*)(b + ocoef);
  if (s->is_complex) {
    // std::complex<double> is laid out as (re, im), as double2
    hip_check(hipMemcpyAsync(b + ocoef, s->cc.data(), (size_t)chi * 16, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  } else {
    // packed Scalars (CoefDev), staged through a fixed host block
    int32_t blk[2048];
    for (int64_t i0 = 0; i0 < chi; i0 += 2048) {
      const int64_t m = std::min<int64_t>(2048, chi - i0);
      for (int64_t i = 0; i < m; ++i) {
        const Scal& v = s->sc[i0 + i];
        if (v.r < -TAB_RMAX / 2 || v.r > TAB_RMAX / 2)
          fail(NX_E_INTERNAL, "coefficient exponent %lld outside the supported range", (long long)v.r);
        blk[i] = v.z ? -1 : (int32_t)((v.k & 7) | ((v.r + CS_ROFF) << 3));
      }
      hip_check(hipMemcpyAsync(b + ocoef + (size_t)i0 * 4, blk, (size_t)m * 4, hipMemcpyHostToDevice, st),
                "hipMemcpyAsync");
      hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");  // blk is reused
    }
  }
  double4* part = (double4*)(b + opart);
  double4* res = (double4*)(b + ores);
  hip_check(launch_qform(s->forms.desc(), q, st), "k_qform");
  hip_check(launch_pairs(q, q, chi, chi, 1, PAIR_J, /*fixed bra, herm*/ 2, cd, cd, c->tabs, part, nblocks, nullptr, 0,
                         st),
            "k_pairs");
  hip_check(launch_reduce(part, nblocks, res, st), "k_reduce");
  double sum[4] = {0.0, 0.0, 0.0, 0.0};
  hip_check(hipMemcpyAsync(sum, res, 32, hipMemcpyDeviceToHost, st), "hipMemcpyAsync D2H");
  hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
  return ab_value(sum).real() * g2;
}

// ---------------------------------------------------------------------------
// gate validation in the reference's error order (left_multiplication/*.rs)
static void validate_gate(int op, int q0, int q1, int n) {
  switch (op) {
    case OP_CX:
    case OP_CZ:
    case OP_SWAP:
      if (q0 < 0 || q0 >= n) err_oob(q0, n);
      if (q1 < 0 || q1 >= n) err_oob(q1, n);
      if (q0 == q1) err_dup_gate(q0);
      return;
    default:
      if (q0 < 0 || q0 >= n) err_oob(q0, n);
  }
}

// ---------------------------------------------------------------------------
// |T>^{(x)t} decomposition (magic_states/cat_state.rs, t_state.rs)
struct Decomp {
  Forms f;
  std::vector<Scal> c;
};

static Forms single(int m, const std::vector<ProgOp>& prog) {
  Forms f = init_zero(1, m);
  if (prog.empty()) return f;
  return run_program(f, prog, m, nullptr);
}

static std::vector<ProgOp> zmio_prog(int m) {  // zero_minus_i_one_state
  std::vector<ProgOp> p{{OP_H, 0, 0, 0}};
  for (int i = 1; i < m; ++i) p.push_back({OP_CX, 0, i, 0});
  p.push_back({OP_SDG, 0, 0, 0});
  return p;
}
static std::vector<ProgOp> even_parity_prog(int m, bool flipped) {
  std::vector<ProgOp> p;
  for (int i = 0; i < m - 1; ++i) p.push_back({OP_H, i, 0, 0});
  for (int i = 0; i < m - 1; ++i) p.push_back({OP_CX, i, m - 1, 0});
  if (flipped)
    for (int i = 0; i < m; ++i)
      for (int j = i + 1; j < m; ++j) p.push_back({OP_CZ, i, j, 0});
  return p;
}

static void check_alive(const std::vector<int4>& st) {
  for (auto& x : st)
    if (x.x || x.w) fail(NX_E_INTERNAL, "magic-state construction hit an impossible projection");
}

static Decomp cat_state(int m) {
  Decomp d;
  switch (m) {
    case 1:
      d.f = init_zero(1, 1);
      d.c = {Scal{0, 0, 0}};
      return d;
    case 2:
      d.f = single(2, {{OP_H, 0, 0, 0}, {OP_CX, 0, 1, 0}, {OP_S, 1, 0, 0}});
      d.c = {Scal{0, 0, 0}};
      return d;
    case 4: {
      Forms a = single(4, zmio_prog(4)), b = single(4, even_parity_prog(4, false));
      d.f = concat({&a, &b});
      d.c = {Scal{0, 7, 1}, Scal{0, 2, 0}};
      return d;
    }
    case 6: {
      Forms a = single(6, zmio_prog(6)), b = single(6, even_parity_prog(6, false)),
            c = single(6, even_parity_prog(6, true));
      d.f = concat({&a, &b, &c});
      d.c = {Scal{0, 0, 2}, Scal{0, 3, 1}, Scal{0, 5, 1}};
      return d;
    }
    case 3:
    case 5: {  // reduce_cat_state
      Decomp big = cat_state(m + 1);
      std::vector<int4> st;
      d.f = run_program(big.f, {{OP_PROJ, m, 0, 0}, {OP_DISC, m, 0, 0}}, m, &st);
      check_alive(st);
      d.c = big.c;
      return d;
    }
    default: {
      Decomp a = cat_state(m - 4), b = cat_state(6);
      Forms k = kron(a.f, b.f);
      for (auto& ca : a.c)
        for (auto& cb : b.c) d.c.push_back(smul(ca, cb));
      const int q0 = m - 5, q1 = m - 4;
      std::vector<int4> st;
      d.f = run_program(k,
                        {{OP_SDG, q0, 0, 0}, {OP_CX, q0, q1, 0}, {OP_H, q0, 0, 0}, {OP_PROJ, q0, 0, 0},
                         {OP_PROJ, q1, 0, 0}, {OP_DISC, q1, 0, 0}, {OP_DISC, q0, 0, 0}},
                        m, &st);
      check_alive(st);
      return d;
    }
  }
}

static Decomp t_tensor_state(int t) {
  Decomp cat = cat_state(t);
  Forms xs = clone_forms(cat.f);
  xs = run_program(xs, {{OP_X, 0, 0, 0}, {OP_S, 0, 0, 0}}, t, nullptr);
  Decomp d;
  d.f = concat({&cat.f, &xs});
  for (auto& c : cat.c) d.c.push_back(smul(c, Scal{0, 0, 1}));            // * ONE_OVER_SQRT_2
  for (auto& c : cat.c) d.c.push_back(smul(smul(c, Scal{0, 0, 1}), Scal{0, 7, 0}));  // * e^{-i pi/4}
  return d;
}

// ---------------------------------------------------------------------------
// compile (compiler/mod.rs:44-140)
static nx_state* compile(int n, int64_t ngates, const int32_t* gates) {
  std::vector<ProgOp> ops;
  int t = 0;
  for (int64_t i = 0; i < ngates; ++i) {
    const int op = gates[4 * i], q0 = gates[4 * i + 1], q1 = gates[4 * i + 2];
    if (op >= OP_H && op <= OP_SWAP) {
      ops.push_back({op, q0, q1, 0});
    } else if (op == 11 || op == 12) {
      const int anc = n + t;
      ops.push_back({OP_CX, q0, anc, 0});
      if (op == 12) ops.push_back({OP_SDG, q0, 0, 0});
      ++t;
    } else {
      fail(NX_E_VALUE, "Gate %s is not supported.", gate_name(op));
    }
  }
  if (n <= 0) fail(NX_E_VALUE, "Number of qubits must be greater than zero, got %d.", n);
  const int ntot = n + t;
  if (ntot > MAX_QUBITS) fail(NX_E_VALUE, "circuits above %d qubits (including one ancilla per T gate) are not supported", MAX_QUBITS);
  for (auto& o : ops) validate_gate(o.code, o.a, o.b, ntot);
  nx_state* s = new nx_state();
  try {
    s->n = n;
    if (t == 0) {
      s->forms = single(n, ops);
      s->sc = {Scal{0, 0, 0}};
      return s;
    }
    Decomp ts = t_tensor_state(t);
    Forms zero = init_zero(1, n);
    Forms full = kron(zero, ts.f);
    std::vector<ProgOp> prog = ops;
    for (int q = ntot - 1; q >= n; --q) prog.push_back({OP_PROJ, q, 0, 0});
    for (int q = ntot - 1; q >= n; --q) prog.push_back({OP_DISC, q, 0, 0});
    std::vector<int4> st;
    Forms out = run_program(full, prog, n, &st);
    std::vector<int64_t> keep;
    for (int64_t i = 0; i < out.chi; ++i)
      if (!st[i].x) {
        if (st[i].w) fail(NX_E_INTERNAL, "Cannot discard qubit after successful projection");
        keep.push_back(i);
        Scal c = ts.c[i];
        if (!c.z) c.r -= st[i].y;  // amplify(#deterministic)
        s->sc.push_back(c);
      }
    s->forms = gather(out, keep);
    s->forms.n = n;
    return s;
  } catch (...) {
    delete s;
    throw;
  }
}

// ---------------------------------------------------------------------------
static void apply_program_all(nx_state* s, const std::vector<ProgOp>& prog) {
  if (prog.empty() || s->chi() == 0) return;
  s->forms = run_program(s->forms, prog, s->n, nullptr);
  s->invalidate();
}

// projection.rs:27-57
static void project_unnormalized(nx_state* s, int q, int outcome) {
  if (q < 0 || q >= s->n) err_oob(q, s->n);
  if (s->chi() == 0) return;
  std::vector<int4> st;
  Forms f = run_program(s->forms, {{OP_PROJ, q, outcome ? 1 : 0, 0}}, s->n, &st);
  std::vector<int64_t> keep;
  std::vector<Scal> sc;
  std::vector<cplx> cc;
  for (int64_t i = 0; i < f.chi; ++i) {
    if (st[i].x) continue;
    keep.push_back(i);
    const bool rnd = st[i].z != 0;
    if (s->is_complex) {
      cplx c = s->cc[i];
      if (rnd) {  // Amplify for Complex64 (coefficient.rs:28-34)
        const double scale = std::pow(2.0, -1.0 / 2.0);
        c = cplx(c.real() * scale, c.imag() * scale);
      }
      cc.push_back(c);
    } else {
      Scal c = s->sc[i];
      if (rnd && !c.z) c.r += 1;  // amplify(-1)
      sc.push_back(c);
    }
  }
  if ((int64_t)keep.size() == f.chi) s->forms = std::move(f);
  else s->forms = gather(f, keep);
  s->sc = std::move(sc);
  s->cc = std::move(cc);
  s->invalidate();
}

// stabilizer_decomposed_state/discard.rs:5-12
static void discard(nx_state* s, int q) {
  if (s->chi() > 0) {
    if (q < 0 || q >= s->n) err_oob(q, s->n);
    std::vector<int4> st;
    Forms f = clone_forms(s->forms);
    f = run_program(f, {{OP_DISC, q, 0, 0}}, s->n - 1, &st);
    for (auto& x : st)
      if (x.w == DEV_ERR_CANNOT_DISCARD) fail(NX_E_VALUE, "Cannot discard qubit %d because it is not in a proper state", q);
    s->forms = std::move(f);
  } else {
    s->forms.n = s->n - 1;
  }
  s->n -= 1;
  s->invalidate();
}

static nx_state* clone_state(const nx_state* s) {
  nx_state* c = new nx_state();
  c->n = s->n;
  c->forms = clone_forms(s->forms);
  c->forms.n = s->forms.n;
  c->is_complex = s->is_complex;
  c->sc = s->sc;
  c->cc = s->cc;
  c->g = s->g;
  return c;
}

struct StateGuard {
  nx_state* p;
  explicit StateGuard(nx_state* s) : p(s) {}
  ~StateGuard() { delete p; }
  StateGuard(const StateGuard&) = delete;
  nx_state* release() { nx_state* r = p; p = nullptr; return r; }
};

// stabilizer_decomposed_state/mod.rs:44-67
static void validate_qargs(const nx_state* s, int nq, const int32_t* qargs) {
  if (nq == 0) fail(NX_E_VALUE, "Qubit index list is empty.");
  for (int i = 0; i < nq; ++i)
    if (qargs[i] < 0 || qargs[i] >= s->n) err_oob(qargs[i], s->n);
  std::vector<int> v(qargs, qargs + nq);
  std::sort(v.begin(), v.end());
  for (size_t i = 0; i + 1 < v.size(); ++i)
    if (v[i] == v[i + 1]) fail(NX_E_VALUE, "Duplicate qubit index found: %d.", v[i]);
}

static void seed_or_entropy(const uint8_t* seed32, uint8_t out[32]) {
  if (seed32) {
    std::memcpy(out, seed32, 32);
    return;
  }
  std::random_device rd;
  for (int i = 0; i < 32; i += 4) {
    uint32_t x = rd();
    std::memcpy(out + i, &x, 4);
  }
}

// norm_squared() of the state projected onto qubit q = 0 and = 1, summed as
// the reference sums it (Scalar coefficients; defined with the sampling
// windows below)
static bool ref_outcome_norms(const nx_state* s, int q, double* p0, double* p1);

// measurement.rs:34-103
static int measure_single(nx_state* s, int q, ChaCha12Rng& rng) {
  StateGuard zero(clone_state(s)), one(clone_state(s));
  project_unnormalized(zero.p, q, 0);
  project_unnormalized(one.p, q, 1);
  double p0 = 0.0, p1 = 0.0;
  if (!ref_outcome_norms(s, q, &p0, &p1)) {
    p0 = norm_squared(zero.p);
    p1 = norm_squared(one.p);
  }
  const double total = p0 + p1;
  if (std::fabs(total) < 1e-12) fail(NX_E_VALUE, "Not implemented: Measurement on a zero-norm state is not possible.");
  const double pz = p0 / total;
  const double sample = rng.gen_f64();
  const int res = sample < pz ? 0 : 1;
  const double nsq = res ? p1 : p0;
  nx_state* chosen = res ? one.release() : zero.release();
  s->forms = std::move(chosen->forms);
  s->sc = std::move(chosen->sc);
  s->cc = std::move(chosen->cc);
  s->g = chosen->g;
  delete chosen;
  s->invalidate();
  const double nrm = std::sqrt(nsq);
  if (std::fabs(nrm) < 1e-12)
    fail(NX_E_VALUE, "Not implemented: Measurement resulted in a zero-norm state, which is not physical.");
  s->g = cmul(s->g, cplx(1.0 / nrm, 0.0));
  return res;
}

// Pauli program (gate_application.rs:186-210) with the length checks of
// exp_value.rs:9-27
static std::vector<ProgOp> pauli_prog(const nx_state* s, int dense, int32_t len, const int32_t* ops,
                                      const int32_t* qubits) {
  std::vector<ProgOp> p;
  if (dense) {
    if (len != s->n) fail(NX_E_VALUE, "Invalid Pauli string length: expected %d, found %d.", s->n, len);
    for (int i = 0; i < len; ++i)
      if (ops[i] >= 1 && ops[i] <= 3) p.push_back({ops[i] == 1 ? OP_X : (ops[i] == 2 ? OP_Y : OP_Z), i, 0, 0});
  } else {
    int mx = 0;
    for (int i = 0; i < len; ++i) mx = std::max(mx, (int)qubits[i]);
    if (mx >= s->n) fail(NX_E_VALUE, "Invalid Pauli string length: expected %d, found %d.", s->n, mx + 1);
    for (int i = 0; i < len; ++i)
      if (ops[i] >= 1 && ops[i] <= 3) p.push_back({ops[i] == 1 ? OP_X : (ops[i] == 2 ? OP_Y : OP_Z), qubits[i], 0, 0});
  }
  return p;
}

static Forms run_pairs_prep(Forms& f, const std::vector<ProgOp>& prog, int n) {
  return run_program(f, prog, n, nullptr);
}

// exp_value.rs:7-61: sum over i <= j of c_j* c_i <phi_j|P phi_i>
static double exp_value(const nx_state* s, const std::vector<ProgOp>& prog) {
  if (s->chi() == 0) return 0.0;
  const QForms& q = qform_of(s);
  Forms pf = clone_forms(s->forms);
  if (!prog.empty()) pf = run_program(pf, prog, s->n, nullptr);
  QForms qp = make_qform(pf);
  CoefDev c = coef_of(s);
  // fixed index = ket i (the Pauli image), bras j in [i, chi)
  cplx sum = run_pairs(q, qp, s->chi(), s->chi(), true, /*fixed ket | herm*/ 3, c, c);
  return sum.real() * (s->g.real() * s->g.real() + s->g.imag() * s->g.imag());
}

// ---------------------------------------------------------------------------
// Pauli variants and variant pair plans.  One k_pairs_var launch per word
// width evaluates, for many jobs at once (a job = one term set and up to
// 64*VW Pauli strings), every pair (bra o >= ket f) with all of the job's
// Paulis applied to the ket; the identity string gives the norm.

// sparse Pauli op lists, q << 2 | op (1 X, 2 Y, 3 Z), in application order
struct PauliSet {
  std::vector<int32_t> off{0}, ops;
  int size() const { return (int)off.size() - 1; }
  void add(const std::vector<ProgOp>& prog) {
    for (const ProgOp& o : prog) ops.push_back(o.a << 2 | (o.code == OP_X ? 1 : (o.code == OP_Y ? 2 : 3)));
    off.push_back((int32_t)ops.size());
  }
};

struct VarSet {
  DevBuf hi, rhs, ph;
  int nvar = 0, VW = 1;
  VarDev desc() const { return VarDev{hi.as<uint64_t>(), rhs.as<uint64_t>(), ph.as<uint64_t>()}; }
};

// variants first .. first + count - 1 of ps, for every term of f
static VarSet pauli_vars(const Forms& f, const PauliSet& ps, int first, int count, int VW) {
  if (count < 0 || count > 64 * VW) fail(NX_E_INTERNAL, "bad variant count %d", count);
  VarSet vs;
  vs.nvar = count;
  vs.VW = VW;
  const size_t R = 64 * (size_t)f.W, chi = (size_t)f.chi;
  vs.hi.alloc(chi * R * VW * 8);
  vs.rhs.alloc(chi * R * VW * 8);
  vs.ph.alloc(chi * 3 * VW * 8);
  std::vector<int32_t> off(count + 1), ops;
  for (int k = 0; k < count; ++k) {
    off[k] = (int32_t)ops.size();
    ops.insert(ops.end(), ps.ops.begin() + ps.off[first + k], ps.ops.begin() + ps.off[first + k + 1]);
  }
  off[count] = (int32_t)ops.size();
  if (ops.empty()) ops.push_back(0);
  DevBuf doff(off.size() * 4), dops(ops.size() * 4);
  h2d(doff.p, off.data(), off.size() * 4);
  h2d(dops.p, ops.data(), ops.size() * 4);
  hip_check(launch_pauli_var(f.desc(), doff.as<int32_t>(), dops.as<int32_t>(), count, VW, vs.hi.as<uint64_t>(),
                             vs.rhs.as<uint64_t>(), vs.ph.as<uint64_t>(), stream()),
            "k_pauli_var");
  return vs;
}

struct VarJob {
  const QForms* X = nullptr;
  const QForms* Y = nullptr;
  const VarSet* V = nullptr;  // null: unit-constraint job (nvar variants, no ket tables)
  int nvar = 0;
  int split_ab = 0;           // write the A and B parts of the sums separately
  CoefDev cx{}, cy{};
  int64_t out_base = 0;
  int64_t base_out = -1;  // >= 0: also the base sum (norm^2 / |g|^2) there
};

struct VarPlan {
  int VW = 1, J = 16, units = 0;
  int64_t total_waves = 0, nout = 0;
  int njobs = 0;
  struct Launch {
    int W;
    int64_t wave_base, nwaves;
  };
  std::vector<Launch> launches;
  int64_t nseg = 0;
  DevBuf jobs, wave_job, row_off, partial, out, seg_job, seg_part;
  // a plan built into a PlanStage: its tables live there (these pointers,
  // not the buffers above)
  const PairJobDev* jobs_p = nullptr;
  const int32_t *wave_job_p = nullptr, *seg_job_p = nullptr;
  double *partial_p = nullptr, *seg_part_p = nullptr;
  std::vector<PairJobDev> hjobs;  // host copy of the job table
  // jobs_dev: the job table to read (default: this plan's own)
  // out_override: where the reduction writes instead of out (page-locked
  // host memory for results the host reads right after)
  // rec: the jobs carry pair records (PairJobDev::rec; unit plans, VW = 1)
  void enqueue(hipStream_t st, const PairJobDev* jobs_dev = nullptr, double* out_override = nullptr,
               bool rec = false) const {
    const PairJobDev* jd = jobs_dev ? jobs_dev : (jobs_p ? jobs_p : jobs.as<PairJobDev>());
    const int32_t* wj = wave_job_p ? wave_job_p : wave_job.as<int32_t>();
    double* part = partial_p ? partial_p : partial.as<double>();
    for (const Launch& L : launches)
      hip_check(launch_pairs_var(L.W, VW, rec ? 2 : units, jd, wj, L.wave_base, L.nwaves, J, ctx()->tabs, part, st),
                "k_pairs_var");
    hip_check(launch_reduce_var(jd, njobs, VW, part, seg_job_p ? seg_job_p : seg_job.as<int32_t>(), nseg,
                                seg_part_p ? seg_part_p : seg_part.as<double>(),
                                out_override ? out_override : out.as<double>(), st),
              "k_reduce_var");
  }
};

// Page-locked staging of a plan's tables and the device block they are
// uploaded to in one copy, plus the plan's partial-sum buffers; kept across
// the plans of a sampling batch (grown between rounds, never while a round's
// launches may still read them)
struct PlanStage {
  PinnedBuf h;
  DevBuf d, partial, seg_part;
  size_t cap = 0, used = 0, pcap = 0, scap = 0;
  PairJobDev* jobs_h = nullptr;  // the staged job table (host side)
};

// Builds the device job table.  Jobs are grouped by word width (one launch
// each); within a launch the waves are shared out in proportion to each
// job's work (units x n^2), at least one wave per job and at most one per
// unit.  row_frac < 1 keeps only that fraction of each job's ket rows (a
// shortened plan of the same shape, for timing launch shapes).  With ps the
// tables go into ps (uploaded by the caller: ps->used bytes of ps->h to
// ps->d) and the plan has no result buffer (enqueue with out_override).
static VarPlan make_plan(const std::vector<VarJob>& in, int VW, int J, int64_t target_waves, double row_frac = 1.0,
                         PlanStage* ps = nullptr) {
  VarPlan pl;
  pl.VW = VW;
  pl.J = J;
  std::vector<int> order(in.size());
  for (size_t i = 0; i < in.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return in[a].X->W < in[b].X->W; });
  std::vector<int64_t> units(in.size()), nrows(in.size());
  double work_total = 0.0;
  int64_t rows_total = 0;
  for (size_t i = 0; i < in.size(); ++i) {
    const int64_t chi = in[i].X->chi;
    nrows[i] = row_frac >= 1.0 ? chi : std::max<int64_t>(1, (int64_t)std::ceil(row_frac * (double)chi));
    units[i] = 0;
    for (int64_t f = 0; f < nrows[i]; ++f) units[i] += (chi - f + J - 1) / J;
    const double n = (double)in[i].X->n;
    work_total += (double)units[i] * n * n;
    rows_total += nrows[i] + 1;
  }
  std::vector<int64_t> roff;
  roff.reserve(rows_total);
  std::vector<int64_t> roff_base(in.size());
  std::vector<PairJobDev> hj(in.size());
  std::vector<int32_t> wj;
  int64_t wave = 0;
  int curW = -1;
  for (int idx : order) {
    const VarJob& j = in[idx];
    const int64_t chi = j.X->chi;
    if (j.X->W != curW) {
      curW = j.X->W;
      pl.launches.push_back({curW, wave, 0});
    }
    roff_base[idx] = (int64_t)roff.size();
    int64_t acc = 0;
    for (int64_t f = 0; f < nrows[idx]; ++f) {
      roff.push_back(acc);
      acc += (chi - f + J - 1) / J;
    }
    roff.push_back(acc);
    const double n = (double)j.X->n;
    int64_t nw = work_total > 0 ? (int64_t)std::llround((double)units[idx] * n * n / work_total * target_waves) : 1;
    nw = std::max<int64_t>(1, std::min<int64_t>(nw, std::max<int64_t>(units[idx], 1)));
    PairJobDev& d = hj[idx];
    d.X = j.X->desc();
    d.Y = j.Y->desc();
    d.V = j.V ? j.V->desc() : VarDev{nullptr, nullptr, nullptr};
    d.cx = j.cx;
    d.cy = j.cy;
    d.row_off = nullptr;  // set below
    d.nfixed = nrows[idx];
    d.nother = chi;
    d.triangle = 1;
    d.nvar = j.V ? j.V->nvar : j.nvar;
    d.nunit = 0;
    d.uq = nullptr;
    d.ub = nullptr;
    d.termvar = 0;
    d.split_ab = j.split_ab;
    d.rec = nullptr;
    for (int w = 0; w < 4; ++w) d.removed[w] = 0ull;
    d.wave0 = wave;
    d.nwaves = nw;
    d.out_base = j.out_base;
    d.base_out = j.base_out;
    for (int64_t w = 0; w < nw; ++w) wj.push_back(idx);
    wave += nw;
    pl.launches.back().nwaves += nw;
    pl.nout = std::max<int64_t>(pl.nout, std::max<int64_t>(j.out_base + (j.split_ab ? 2 : 1) * d.nvar,
                                                           j.base_out + 1));
  }
  pl.total_waves = wave;
  pl.njobs = (int)in.size();
  std::vector<int32_t> sj;
  for (size_t i = 0; i < in.size(); ++i) {
    hj[i].seg0 = (int64_t)sj.size();
    hj[i].nseg = (int32_t)((hj[i].nwaves + RED_SEG - 1) / RED_SEG);
    for (int32_t k = 0; k < hj[i].nseg; ++k) sj.push_back((int32_t)i);
  }
  pl.nseg = (int64_t)sj.size();
  pl.units = (!in.empty() && in[0].V == nullptr) ? 1 : 0;
  if (ps) {
    const auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t need = al(sj.size() * 4) + al(roff.size() * 8) + al(hj.size() * sizeof(PairJobDev)) +
                        al(wj.size() * 4);
    if (ps->cap < need) {
      ps->cap = 2 * need;
      ps->h.alloc(ps->cap);
      ps->d.alloc(ps->cap);
    }
    size_t o = 0;
    const auto put = [&](const void* src, size_t bytes) {
      std::memcpy(ps->h.as<char>(o), src, bytes);
      char* dp = ps->d.as<char>() + o;
      o += al(bytes);
      return dp;
    };
    pl.seg_job_p = (const int32_t*)put(sj.data(), sj.size() * 4);
    const int64_t* roff_d = (const int64_t*)put(roff.data(), roff.size() * 8);
    for (size_t i = 0; i < in.size(); ++i) hj[i].row_off = roff_d + roff_base[i];
    ps->jobs_h = ps->h.as<PairJobDev>(o);
    pl.jobs_p = (const PairJobDev*)put(hj.data(), hj.size() * sizeof(PairJobDev));
    pl.wave_job_p = (const int32_t*)put(wj.data(), wj.size() * 4);
    ps->used = o;
    const size_t pbytes = (size_t)std::max<int64_t>(wave, 1) * 2 * (64 * VW + 1) * 8;
    const size_t sbytes = (size_t)std::max<int64_t>(pl.nseg, 1) * 2 * (64 * VW + 1) * 8;
    if (ps->pcap < pbytes) {
      ps->pcap = 2 * pbytes;
      ps->partial.alloc(ps->pcap);
    }
    if (ps->scap < sbytes) {
      ps->scap = 2 * sbytes;
      ps->seg_part.alloc(ps->scap);
    }
    pl.partial_p = ps->partial.as<double>();
    pl.seg_part_p = ps->seg_part.as<double>();
    pl.hjobs = hj;
    return pl;
  }
  pl.seg_job.alloc(std::max<size_t>(sj.size(), 1) * 4);
  h2d(pl.seg_job.p, sj.data(), sj.size() * 4);
  pl.seg_part.alloc(std::max<int64_t>(pl.nseg, 1) * 2 * (64 * VW + 1) * 8);
  pl.row_off.alloc(std::max<size_t>(roff.size(), 1) * 8);
  h2d(pl.row_off.p, roff.data(), roff.size() * 8);
  for (size_t i = 0; i < in.size(); ++i) hj[i].row_off = pl.row_off.as<int64_t>() + roff_base[i];
  pl.jobs.alloc(std::max<size_t>(hj.size(), 1) * sizeof(PairJobDev));
  h2d(pl.jobs.p, hj.data(), hj.size() * sizeof(PairJobDev));
  pl.hjobs = hj;
  pl.wave_job.alloc(std::max<size_t>(wj.size(), 1) * 4);
  h2d(pl.wave_job.p, wj.data(), wj.size() * 4);
  pl.partial.alloc(std::max<int64_t>(wave, 1) * 2 * (64 * VW + 1) * 8);
  pl.out.alloc(std::max<int64_t>(pl.nout, 1) * 8);
  return pl;
}

// Launch-shape autotuner: pair-chunk length J and total wave count, timed
// with HIP events on the library stream; results are cached per (word
// width, variant words, log2 of the pair count).
struct TuneKey {
  int W, VW, lg;
  bool operator<(const TuneKey& o) const {
    return std::tie(W, VW, lg) < std::tie(o.W, o.VW, o.lg);
  }
};
static std::mutex g_tune_mu;
static std::map<TuneKey, std::pair<int, int64_t>> g_tune;

static const int64_t DEFAULT_WAVES = 16384;

static TuneKey tune_key(const std::vector<VarJob>& jobs, int VW) {
  int W = 0;
  double pairs = 0.0;
  for (const VarJob& j : jobs) {
    W = std::max(W, j.X->W);
    pairs += 0.5 * (double)j.X->chi * (double)(j.X->chi + 1);
  }
  return TuneKey{W, VW, (int)std::floor(std::log2(std::max(pairs, 1.0)))};
}

static std::pair<int, int64_t> tuned_shape(const std::vector<VarJob>& jobs, int VW) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(tune_key(jobs, VW));
  if (it != g_tune.end()) return it->second;
  return {16, DEFAULT_WAVES};
}

// Each candidate is timed on a shortened plan (about 2^22 pairs: the same
// jobs with a fraction of their ket rows); the two fastest are timed again
// on a plan four times as long, since a short plan has few units per wave,
// which favours short chunks.
static std::pair<int, int64_t> autotune(const std::vector<VarJob>& jobs, int VW) {
  static const int Js[] = {8, 16, 32};
  static const int64_t Ws[] = {8192, 16384, 32768};
  double pairs = 0.0;
  for (const VarJob& j : jobs) pairs += 0.5 * (double)j.X->chi * (double)(j.X->chi + 1);
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "hipEventCreate");
  hip_check(hipEventCreate(&e1), "hipEventCreate");
  auto time_shape = [&](int J, int64_t nw, double frac) {
    VarPlan pl = make_plan(jobs, VW, J, nw, frac);
    pl.enqueue(stream());  // warm-up
    hip_check(hipEventRecord(e0, stream()), "hipEventRecord");
    pl.enqueue(stream());
    hip_check(hipEventRecord(e1, stream()), "hipEventRecord");
    hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    return ms;
  };
  const double frac = std::min(1.0, 4194304.0 / std::max(pairs, 1.0));
  std::vector<std::pair<float, std::pair<int, int64_t>>> t;
  for (int J : Js)
    for (int64_t nw : Ws) t.push_back({time_shape(J, nw, frac), {J, nw}});
  std::stable_sort(t.begin(), t.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
  std::pair<int, int64_t> arg = t[0].second;
  if (frac < 1.0) {
    const double frac4 = std::min(1.0, 4.0 * frac);
    const float a = time_shape(t[0].second.first, t[0].second.second, frac4);
    const float b = time_shape(t[1].second.first, t[1].second.second, frac4);
    if (b < a) arg = t[1].second;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[tune_key(jobs, VW)] = arg;
  return arg;
}

// Dense Pauli strings -> PauliSet, with the length checks of exp_value.rs:9-27
static void add_dense_paulis(const nx_state* s, int32_t npauli, const int32_t* ops, PauliSet& ps) {
  for (int32_t k = 0; k < npauli; ++k) ps.add(pauli_prog(s, 1, s->n, ops + (size_t)k * s->n, nullptr));
}

// Unnormalised <P_k> for every string of ps (exp_value.rs:7-61), all
// strings per pair in one pass.
static std::vector<double> exp_values_var(const nx_state* s, const PauliSet& ps) {
  const int np = ps.size();
  std::vector<double> res(np, 0.0);
  if (s->chi() == 0 || np == 0) return res;
  const int VW = np <= 64 ? 1 : 2;
  const int per = 64 * VW;
  const QForms& q = qform_of(s);
  CoefDev c = coef_of(s);
  std::vector<VarSet> vs;
  vs.reserve((np + per - 1) / per);
  for (int first = 0; first < np; first += per) vs.push_back(pauli_vars(s->forms, ps, first, std::min(per, np - first), VW));
  std::vector<VarJob> jobs;
  for (size_t ci = 0; ci < vs.size(); ++ci) {
    VarJob j;
    j.X = &q;
    j.Y = &q;
    j.V = &vs[ci];
    j.cx = c;
    j.cy = c;
    j.out_base = (int64_t)ci * per;
    jobs.push_back(j);
  }
  const auto shape = tuned_shape(jobs, VW);
  VarPlan pl = make_plan(jobs, VW, shape.first, shape.second);
  {
    ProfScope ps_("k_pairs_var");
    pl.enqueue(stream());
    ps_.end();
  }
  std::vector<double> o(pl.nout);
  d2h(o.data(), pl.out.p, pl.nout * 8);
  sync();
  const double g2 = s->g.real() * s->g.real() + s->g.imag() * s->g.imag();
  for (int k = 0; k < np; ++k) res[k] = o[k] * g2;
  return res;
}

// Coefficient backward of L = sum_k w_k exp_value(P_k) (exp_value.rs:7-61)
// for the strings of ps, device weights d_wts and coefficient values c
// (device complex128 [chi]): d_loss[0] = L and d_grad = 2 |g|^2 G c, G =
// sum_k w_k G_k, from one pass over the pairs per 64 VW strings (k_grad_var:
// tiles of pairs, no chi x chi matrix), enqueued on the library stream.
static void grad_enqueue(const nx_state* s, const PauliSet& ps, const double* d_wts, const double2* c, double* d_loss,
                         double2* d_grad) {
  const int np = ps.size();
  const int64_t chi = s->chi();
  const int VW = np <= 64 ? 1 : 2;
  const int per = 64 * VW;
  const QForms& q = qform_of(s);
  const double g2 = s->g.real() * s->g.real() + s->g.imag() * s->g.imag();
  const int64_t nt = grad_tiles(chi);
  const size_t tile_vals = (size_t)nt * grad_tile_size();
  // three rounds of resident waves (tiles of 16 x 16 pairs even out)
  const int nblocks = (int)std::max<int64_t>(1, std::min<int64_t>((nt + 3) / 4, 6144));
  const int64_t nwaves = (int64_t)nblocks * 4;
  DevBuf trow(tile_vals * 16), tcol(tile_vals * 16), part((size_t)nwaves * per * 8), vals((size_t)np * 8);
  for (int first = 0; first < np; first += per) {
    const int cnt = std::min(per, np - first);
    VarSet vs = pauli_vars(s->forms, ps, first, cnt, VW);
    ProfScope pg("k_grad_var");
    hip_check(launch_grad_var(q.desc(), vs.desc(), VW, c, cnt, d_wts + first, ctx()->tabs, nblocks, trow.as<double2>(),
                              tcol.as<double2>(), part.as<double>(), stream()),
              "k_grad_var");
    pg.end();
    hip_check(launch_grad_reduce(trow.as<double2>(), tcol.as<double2>(), chi, 2.0 * g2, first > 0, d_grad, stream()),
              "k_grad_reduce");
    hip_check(launch_grad_vals(part.as<double>(), nwaves, per, cnt, g2, vals.as<double>() + first, stream()),
              "k_grad_vals");
    // vs returns to the pool; its next user is ordered after these launches
  }
  hip_check(launch_grad_loss(vals.as<double>(), d_wts, np, d_loss, stream()), "k_grad_loss");
}

// One state's Z-basis projection norms, 2^d outcomes per pass.  The plan
// (job table, unit offsets, wave map) is built once; a pass only uploads the
// unit constraints: the fixed outcomes of the already processed qubits and
// the d look-ahead qubits whose 2^d outcomes are the variants (variant k
// takes outcome (k >> i) & 1 on look-ahead qubit i).
struct ProjPlan {
  const nx_state* s = nullptr;
  VarPlan plan;
  QForms projv;  // the forms projected onto the fixed and look-ahead outcomes (base + tables)
  DevBuf vhi, vrhs, vph;
  std::vector<PairJobDev> hj;
  int VW = 1, maxunit = 0;
  double g2 = 1.0;
  // one upload per pass: [job | fixed qubits | fixed outcomes | look-ahead
  // qubits | look-ahead outcome masks], staged in page-locked memory: this
  // plan's own buffers, or slices of a sampling batch's (one upload for the
  // windows of all its states)
  DevBuf stage_d_own;
  PinnedBuf stage_h_own, res_h;
  char* sh = nullptr;  // host staging
  char* sd = nullptr;  // device staging
  size_t o_fq = 0, o_fb = 0, o_uq = 0, o_ub = 0, stage_bytes = 0;
  int nfix_staged = 0;  // fixed outcomes of the staged pass
  static int vw_for(int max_look) { return (1 << max_look) <= 64 ? 1 : 2; }
  static size_t stage_size(int max_units, int max_look, size_t* ofq = nullptr, size_t* ofb = nullptr,
                           size_t* ouq = nullptr, size_t* oub = nullptr) {
    const size_t mu = (size_t)std::max(max_units, 1);
    const int vw = vw_for(max_look);
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t a = up(sizeof(PairJobDev)), b = a + up(mu * 4), c = b + up(mu * 4), d = c + up(mu * 4);
    if (ofq) *ofq = a;
    if (ofb) *ofb = b;
    if (ouq) *ouq = c;
    if (oub) *oub = d;
    return d + up(mu * (vw + 1) * 8);
  }
  // ext_h / ext_d: stage_size() bytes of page-locked host and device memory
  // owned by a sampling batch, whose own launch plans run the passes (null:
  // this plan allocates its staging and builds its own one-job launch plan)
  void init(const nx_state* st, int max_units, int max_look, char* ext_h = nullptr, char* ext_d = nullptr) {
    s = st;
    VW = vw_for(max_look);
    maxunit = max_units;
    if (ext_h && ext_d) {
      hj.assign(1, PairJobDev{});
    } else {
      VarJob j;
      j.X = &qform_of(st);
      j.Y = j.X;
      j.cx = coef_of(st);
      j.cy = j.cx;
      j.nvar = 1 << max_look;
      j.split_ab = 1;
      // chunks of 8 pairs and three rounds of resident waves: one window pass
      // of a single state is short, and fine work units keep its waves even
      // (1,024 shots of a 64-qubit, 729-term state on all qubits: 1.26 s with
      // 16-pair chunks in 16k waves, 1.19 s so)
      plan = make_plan({j}, VW, 8, std::max<int64_t>(3 * resident_waves(j.X->W, VW), 1024));
      hj = plan.hjobs;
    }
    stage_bytes = stage_size(max_units, max_look, &o_fq, &o_fb, &o_uq, &o_ub);
    if (ext_h && ext_d) {
      sh = ext_h;
      sd = ext_d;
    } else {
      stage_d_own.alloc(stage_bytes);
      stage_h_own.alloc(stage_bytes);
      sh = stage_h_own.as<char>();
      sd = stage_d_own.as<char>();
      res_h.alloc((size_t)2 * 64 * VW * 8);
    }
    projv.alloc_pair(st->chi(), st->n);
    const size_t rows = (size_t)st->chi() * 64 * projv.W;
    vhi.alloc(std::max<size_t>(rows * VW * 8, 8));
    vrhs.alloc(std::max<size_t>(rows * VW * 8, 8));
    vph.alloc(std::max<size_t>((size_t)st->chi() * 3 * VW * 8, 8));
    sync();
    g2 = st->g.real() * st->g.real() + st->g.imag() * st->g.imag();
  }
  // waves of k_pairs_var<W, VW, units> resident on this device (cached)
  static int64_t resident_waves(int W, int VW) {
    static std::mutex res_mu;
    static std::map<std::tuple<int, int, int>, int64_t> resident;  // (device, W, VW) -> waves
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    const auto key = std::make_tuple(dev, W, VW);
    std::lock_guard<std::mutex> lk(res_mu);
    auto it = resident.find(key);
    if (it == resident.end()) it = resident.emplace(key, pairs_var_resident_waves(W, VW, 1)).first;
    return it->second;
  }
  VarDev tables() const { return VarDev{vhi.as<uint64_t>(), vrhs.as<uint64_t>(), vph.as<uint64_t>()}; }
  // Host side of a pass: the fixed outcomes (fq, fb) and the look-ahead
  // qubits into the staging block, and the pass's job (hj[0], also staged:
  // its pointers refer to the device staging block).  Returns 2^d.
  int stage(const std::vector<int>& fq, const std::vector<int>& fb, const std::vector<int>& look, uint64_t* rec) {
    const int d = (int)look.size(), nv = 1 << d;
    if ((int)(fq.size() + look.size()) > maxunit) fail(NX_E_INTERNAL, "too many projected qubits");
    if (nv > 64 * VW) fail(NX_E_INTERNAL, "too many look-ahead qubits");
    PairJobDev& job = hj[0];
    for (int w = 0; w < 4; ++w) job.removed[w] = 0ull;
    int32_t* hfq = (int32_t*)(sh + o_fq);
    int32_t* hfb = (int32_t*)(sh + o_fb);
    for (size_t i = 0; i < fq.size(); ++i) {
      hfq[i] = fq[i];
      hfb[i] = fb[i];
      job.removed[fq[i] >> 6] |= 1ull << (fq[i] & 63);
    }
    nfix_staged = (int)fq.size();
    int32_t* uq = (int32_t*)(sh + o_uq);
    uint64_t* ub = (uint64_t*)(sh + o_ub);
    for (int i = 0; i < d; ++i) {
      uq[i] = look[i];
      uint64_t* e = ub + (size_t)i * (VW + 1);
      e[0] = 0ull;
      for (int v = 0; v < VW; ++v) {
        uint64_t m = 0;
        for (int k = 0; k < 64; ++k)
          if (64 * v + k < nv && (((64 * v + k) >> i) & 1)) m |= 1ull << k;
        e[1 + v] = m;
      }
    }
    job.nunit = (int32_t)d;
    job.uq = (const int32_t*)(sd + o_uq);
    job.ub = (const uint64_t*)(sd + o_ub);
    job.nvar = nv;
    // the look-ahead outcomes are substituted once per term as well
    // (k_qproject_var): the pairs start from the projected forms and tables
    job.X = job.Y = projv.desc();
    job.V = tables();
    job.termvar = 1;
    job.rec = rec;
    for (int i = 0; i < d; ++i) job.removed[look[i] >> 6] |= 1ull << (look[i] & 63);
    std::memcpy(sh, &job, sizeof(PairJobDev));
    return nv;
  }
  // the projection job of the staged pass (for launch_qproject_var_jobs)
  QProjJobDev qproj_job() const {
    QProjJobDev q;
    q.in = qform_of(s).desc();
    q.out = projv.desc();
    q.vout = tables();
    q.fq = (const int32_t*)(sd + o_fq);
    q.fb = (const int32_t*)(sd + o_fb);
    q.nfix = nfix_staged;
    q.uq = hj[0].uq;
    q.ub = hj[0].ub;
    q.nunit = hj[0].nunit;
    q.block0 = 0;
    return q;
  }
  // norm_squared / |g|^2 of Pi psi for the 2^d outcomes of qubits look[0..d),
  // as exact parts: value = A[k] + sqrt(2) B[k] (A, B = the returned halves)
  // rec != null: the pass also writes every pair's exact per-outcome values
  // (PairJobDev::rec)
  std::vector<double> run(const std::vector<int>& fq, const std::vector<int>& fb, const std::vector<int>& look,
                          uint64_t* rec = nullptr) {
    if (rec && VW != 1) fail(NX_E_INTERNAL, "pair records need one word of variants");
    const int nv = stage(fq, fb, look, rec);
    h2d(sd, sh, stage_bytes);
    {
      ProfScope ps("k_qproject_var");
      const QProjJobDev q = qproj_job();
      hip_check(launch_qproject_var(q.in, q.out, q.vout, VW, q.fq, q.fb, q.nfix, q.uq, q.ub, q.nunit, stream()),
                "k_qproject_var");
      ps.end();
    }
    {
      ProfScope ps("k_pairs_var");
      plan.enqueue(stream(), (const PairJobDev*)sd, res_h.as<double>(), rec != nullptr);
      ps.end();
    }
    sync();
    return std::vector<double>(res_h.as<double>(), res_h.as<double>() + 2 * nv);
  }
  // the reference's norm_squared: (sum) * |g|^2, the sum being A + sqrt(2) B
  double value(double A, double B) const { return (A + 1.4142135623730951 * B) * g2; }
};

// Norms of every node of a DFS window: the window rooted at depth `root`
// holds, for relative depth j = 1 .. d and outcome index k < 2^j (bit i =
// outcome of processed qubit root + i), the norm of that node as the sum of
// its look-ahead leaves.
struct Window {
  int root = 0, d = 0;
  bool ref = false;                          // norms in nr (reference order), not la/lb
  std::vector<std::vector<double>> la, lb;  // exact parts A, B of node [j][k]
  std::vector<std::vector<double>> nr;      // norm_squared of node [j][k] as the reference sums it
};

// Buffers of the reference-order fold (fold.hip), kept across the windows of
// one sample call: pair records, per-(term, node) projection counts, one
// chunk of per-(pair, node) terms, the node sums and the error flag.
struct RefFold {
  DevBuf rec, D, t, cnt, sums;
  PinnedBuf host;
  int64_t chunk = 0;
};

// The largest D such that the Z-basis marginal of the qubits order[0..D)
// is provably uniform: no pair value <phi_o| Z^t |phi_f> with t != 0
// supported on them is nonzero, so <psi| Z^t |psi> = 0 for all such t and
// ||Pi_b psi||^2 = 2^-D ||psi||^2 for every outcome b on them.  A sufficient
// condition (cancellation between pairs is not used); at most 128.  With
// `fixed` qubits, q holds the forms projected onto their outcomes (the
// conditional marginal given those outcomes).
static int uniform_depth(const QForms& q, int n, const std::vector<int>& order, const std::vector<int>& fixed) {
  const int64_t chi = q.chi;
  const int ncoord = std::min<int>((int)order.size(), 128);
  if (chi == 0 || ncoord == 0) return 0;
  const int J = 16;
  std::vector<int64_t> off(chi + 1);
  off[0] = 0;
  for (int64_t f = 0; f < chi; ++f) off[f + 1] = off[f] + (chi - f + J - 1) / J;
  std::vector<int32_t> coord(n, -1);
  for (int k = 0; k < ncoord; ++k) coord[order[k]] = k;
  uint64_t removed[4] = {0, 0, 0, 0};
  for (int qq : fixed) removed[qq >> 6] |= 1ull << (qq & 63);
  const int32_t init = ncoord + 1;
  DevBuf doff(off.size() * 8), dcoord(coord.size() * 4), dmin(4), drem(32);
  h2d(doff.p, off.data(), off.size() * 8);
  h2d(dcoord.p, coord.data(), coord.size() * 4);
  h2d(dmin.p, &init, 4);
  h2d(drem.p, removed, 32);
  const int64_t units = off[chi];
  int nblocks = (int)std::min<int64_t>((units + 3) / 4, 4096);
  if (nblocks < 1) nblocks = 1;
  ProfScope ps("k_pairs_sym");
  hip_check(launch_pairs_sym(q.desc(), doff.as<int64_t>(), J, ncoord > 64 ? 2 : 1, dcoord.as<int32_t>(),
                             fixed.empty() ? nullptr : drem.as<uint64_t>(), dmin.as<int>(), nblocks, stream()),
            "k_pairs_sym");
  ps.end();
  int32_t d = 0;
  d2h(&d, dmin.p, 4);
  sync();
  return d - 1;
}

static int uniform_prefix(const nx_state* s, const std::vector<int>& order) {
  return uniform_depth(qform_of(s), s->n, order, {});
}

// Leaf norms of a sampling tail (Scalar coefficients): for the 2^d
// completions of the path `base` on qubits qs, |psi(x)|^2 / |g|^2 = A + sqrt2 B
// from exact amplitudes (components in the basis 1, zeta, zeta^2, zeta^3):
// every term restricted to the d tail variables (tails.hip).  Device buffers
// and the page-locked staging / result block are kept per state size.
struct TailBufs {
  int64_t nv = 0, chi = 0;
  DevBuf recs, al, ab, meta;
  PinnedBuf host;
};

// job fields of one tail (pointers into the caller's buffers); d <= 16
static TailJobDev tail_job(const nx_state* s, int d) {
  if (d < 1 || d > 16) fail(NX_E_INTERNAL, "sampling tail of %d qubits", d);
  if (words_for(s->n) > 2) fail(NX_E_INTERNAL, "sampling tails need at most 128 qubits");
  TailJobDev t{};
  t.q = qform_of(s).desc();
  t.cs = coef_of(s).cs;
  t.d = d;
  t.nv = (int64_t)1 << d;
  t.nchunks = tail_chunks(s->chi(), t.nv);
  return t;
}

static void tail_leaves(const nx_state* s, TailBufs& tb, const std::vector<uint64_t>& base,
                        const std::vector<int32_t>& qs, std::vector<double>& A, std::vector<double>& B) {
  const int d = (int)qs.size();
  TailJobDev t = tail_job(s, d);
  const int64_t nv = t.nv, chi = s->chi();
  if (tb.nv < nv || tb.chi < chi) {
    tb.nv = std::max(tb.nv, nv);
    tb.chi = std::max(tb.chi, chi);
    tb.recs.alloc((size_t)tb.chi * TAIL_REC * 4);
    tb.al.alloc((size_t)64 * tb.nv * 4 * 8);
    tb.ab.alloc((size_t)tb.nv * 2 * 8);
    tb.meta.alloc(512 + sizeof(TailJobDev));
    tb.host.alloc(std::max<size_t>((size_t)tb.nv * 2 * 8, 512 + sizeof(TailJobDev)));
  }
  // [job | base words | qubit list] through one staged upload
  const size_t ob = 256 * ((sizeof(TailJobDev) + 255) / 256), oq = ob + 32;
  std::memset(tb.host.as<char>(ob), 0, 32);
  std::memcpy(tb.host.as<char>(ob), base.data(), base.size() * 8);
  std::memcpy(tb.host.as<char>(oq), qs.data(), (size_t)d * 4);
  t.base = (const uint64_t*)(tb.meta.as<char>() + ob);
  t.qs = (const int32_t*)(tb.meta.as<char>() + oq);
  t.recs = tb.recs.as<uint32_t>();
  t.al = tb.al.as<double>();
  t.ab = tb.ab.as<double>();
  t.term0 = t.block0 = t.str0 = 0;
  std::memcpy(tb.host.p, &t, sizeof t);
  h2d(tb.meta.p, tb.host.p, oq + (size_t)d * 4);
  {
    ProfScope ps("k_tail");
    const int64_t blocks = ((nv + 1023) / 1024) * t.nchunks;
    hip_check(launch_tails(t.q.W, tb.meta.as<TailJobDev>(), 1, chi, blocks, nv, stream()), "launch_tails");
    ps.end();
  }
  d2h(tb.host.p, tb.ab.p, (size_t)nv * 16);
  sync();
  A.assign(tb.host.as<double>(), tb.host.as<double>() + nv);
  B.assign(tb.host.as<double>() + nv, tb.host.as<double>() + 2 * nv);
}

struct SampleBatch;

struct SampleCtx {
  SampleBatch* batch = nullptr;  // non-null: windows come from a sampling batch's dispatcher
  int uniform = 0;  // the first `uniform` qubits of q have a uniform marginal
  int tail_m = 0;   // windows of at most this many levels come from exact amplitudes (0: never)
  const nx_state* s = nullptr;
  TailBufs tb;
  std::vector<int> q, bit;
  std::vector<uint64_t> bits, counts;
  std::vector<int> path;  // outcomes of the processed qubits so far
  ChaCha12Rng* rng;
  ProjPlan pp;
  int look = 6;
  int64_t passes = 0, tails = 0;
  bool ref = false;  // node norms in the reference's FP64 summation order
  RefFold rf;
  // NX_SAMPLE_PROGRESS set: one line on stderr per window (long runs)
  bool progress = false;
  std::chrono::steady_clock::time_point t0;
};

static void sample_progress(const SampleCtx& cx, const char* what, int level) {
  if (!cx.progress) return;
  const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - cx.t0).count();
  std::fprintf(stderr, "[necstar sample] %s at level %d: %lld windows, %lld tails, %.2f s\n", what, level,
               (long long)cx.passes, (long long)cx.tails, s);
  std::fflush(stderr);
}

// The window rooted at depth `root` of the path cx.path: its look-ahead
// qubits and the fixed outcomes above it
static void window_args(const SampleCtx& cx, int root, Window& w, std::vector<int>& fq, std::vector<int>& fb,
                        std::vector<int>& lq) {
  w.root = root;
  w.d = std::min<int>(cx.look, (int)cx.q.size() - root);
  fq.assign(cx.q.begin(), cx.q.begin() + root);
  fb.assign(cx.path.begin(), cx.path.begin() + root);
  lq.assign(cx.q.begin() + root, cx.q.begin() + root + w.d);
}

// leaves A, B (2^d each) into the window, then its inner nodes as exact
// sums of their leaves (powers of two in A and B)
static void window_fill(Window& w, const double* A, const double* B) {
  const size_t nv = (size_t)1 << w.d;
  w.la.resize(w.d + 1);
  w.lb.resize(w.d + 1);
  w.la[w.d].assign(A, A + nv);
  w.lb[w.d].assign(B, B + nv);
  for (int j = w.d; j >= 2; --j) {
    const size_t m = (size_t)1 << (j - 1);
    w.la[j - 1].assign(m, 0.0);
    w.lb[j - 1].assign(m, 0.0);
    for (size_t k = 0; k < m; ++k) {
      w.la[j - 1][k] = w.la[j][k] + w.la[j][k | m];
      w.lb[j - 1][k] = w.lb[j][k] + w.lb[j][k | m];
    }
  }
}

static Window open_window(SampleCtx& cx, int root) {
  Window w;
  std::vector<int> fq, fb, lq;
  window_args(cx, root, w, fq, fb, lq);
  const std::vector<double> ab = cx.pp.run(fq, fb, lq);
  const size_t nv = (size_t)1 << w.d;
  window_fill(w, ab.data(), ab.data() + nv);
  ++cx.passes;
  return w;
}

// Node norms in the reference's FP64 summation order: for the window of
// look-ahead qubits lq below the fixed outcomes (fq, fb), norm_squared() of
// the state projected on each node's path (norm.rs:8-30 over the terms that
// projection.rs:35-51 keeps), summed in its order, times |g|^2 as norm.rs:29
// does; node u = 2^j - 2 + k (relative depth j, outcome index k).  The pass
// writes the exact per-leaf values of every pair; k_fold_diag /
// k_fold_terms / k_fold_sum turn them into each node's sequential sum.
static int64_t ref_fold_max_pairs() { return (int64_t)1 << 28; }

static std::vector<double> ref_node_norms(const nx_state* s, ProjPlan& pp, RefFold& rf, const std::vector<int>& fq,
                                          const std::vector<int>& fb, const std::vector<int>& lq) {
  const int d = (int)lq.size();
  const int64_t chi = s->chi();
  const int64_t np = chi * (chi + 1) / 2;
  const int nn = (2 << d) - 2;
  if (np > ref_fold_max_pairs()) fail(NX_E_INTERNAL, "too many term pairs for the reference-order fold");
  if (!rf.rec.p) {
    rf.rec.alloc((size_t)np * FOLD_REC * 8);
    rf.D.alloc((size_t)chi * FOLD_NODES);
    rf.chunk = (std::min<int64_t>(np, (int64_t)1 << 20) + 255) / 256 * 256;
    rf.t.alloc((size_t)rf.chunk * FOLD_NODES * 8);
    rf.cnt.alloc((size_t)(rf.chunk / 256) * FOLD_NODES * 4);
    rf.sums.alloc(FOLD_NODES * 8 + 256);
    rf.host.alloc(FOLD_NODES * 8 + 256);
  }
  double* sums = rf.sums.as<double>();
  int* err = (int*)(rf.sums.as<char>() + FOLD_NODES * 8);
  pp.run(fq, fb, lq, rf.rec.as<uint64_t>());
  {
    ProfScope ps("k_fold");
    hip_check(hipMemsetAsync(rf.sums.p, 0, FOLD_NODES * 8 + 4, stream()), "hipMemsetAsync");
    hip_check(launch_fold_diag(rf.rec.as<uint64_t>(), chi, d, rf.D.as<uint8_t>(), err, stream()), "k_fold_diag");
    for (int64_t p0 = 0; p0 < np; p0 += rf.chunk) {
      const int64_t m = std::min<int64_t>(rf.chunk, np - p0);
      hip_check(launch_fold_terms(rf.rec.as<uint64_t>(), chi, p0, m, rf.chunk, d, rf.D.as<uint8_t>(),
                                  coef_of(s).cs, ctx()->tabs, rf.t.as<double>(), rf.cnt.as<int32_t>(), err, stream()),
                "k_fold_terms");
      hip_check(launch_fold_sum(rf.t.as<double>(), rf.cnt.as<int32_t>(), m, rf.chunk, nn, sums, stream()),
                "k_fold_sum");
    }
    ps.end();
  }
  d2h(rf.host.p, rf.sums.p, FOLD_NODES * 8 + 4);
  sync();
  if (*(const int*)(rf.host.as<char>() + FOLD_NODES * 8))
    fail(NX_E_INTERNAL, "a projected pair value is not of the form zeta^k 2^(-r/2)");
  std::vector<double> out(nn);
  for (int u = 0; u < nn; ++u) out[u] = rf.host.as<double>()[u] * pp.g2;
  return out;
}

static bool ref_outcome_norms(const nx_state* s, int q, double* p0, double* p1) {
  const int64_t chi = s->chi();
  if (s->is_complex || chi == 0 || chi * (chi + 1) / 2 > ref_fold_max_pairs()) return false;
  ProjPlan pp;
  pp.init(s, 1, 1);
  RefFold rf;
  const std::vector<double> nv = ref_node_norms(s, pp, rf, {}, {}, {q});
  *p0 = nv[0];
  *p1 = nv[1];
  return true;
}

static Window open_window_ref(SampleCtx& cx, int root) {
  Window w;
  w.root = root;
  w.ref = true;
  w.d = std::min<int>(cx.look, (int)cx.q.size() - root);
  std::vector<int> fq(cx.q.begin(), cx.q.begin() + root), fb(cx.path.begin(), cx.path.begin() + root);
  std::vector<int> lq(cx.q.begin() + root, cx.q.begin() + root + w.d);
  const std::vector<double> nv = ref_node_norms(cx.s, cx.pp, cx.rf, fq, fb, lq);
  ++cx.passes;
  w.nr.resize(w.d + 1);
  for (int j = 1; j <= w.d; ++j) w.nr[j].assign(nv.begin() + ((1 << j) - 2), nv.begin() + ((2 << j) - 2));
  return w;
}

// The last levels of a path (every qubit sampled, Scalar coefficients): the
// leaf norms are |psi(x)|^2 of the 2^d completions x of the path, from exact
// amplitudes: with psi = sum_j a_j zeta^j, |psi|^2 = sum_j a_j^2 +
// sqrt2 (a0 a1 + a1 a2 + a2 a3 - a0 a3), A and B formed with exact products
// and rounded once -- the values the pair sums give whenever those are exact.
static Window open_tail(SampleCtx& cx, int root) {
  Window w;
  w.root = root;
  w.d = (int)cx.q.size() - root;
  const int W = words_for(cx.s->n);
  std::vector<uint64_t> base(W, 0ull);
  for (int i = 0; i < root; ++i)
    if (cx.path[i]) base[cx.q[i] >> 6] |= 1ull << (cx.q[i] & 63);
  std::vector<int32_t> qs(cx.q.begin() + root, cx.q.end());
  w.la.resize(w.d + 1);
  w.lb.resize(w.d + 1);
  tail_leaves(cx.s, cx.tb, base, qs, w.la[w.d], w.lb[w.d]);
  ++cx.tails;
  for (int j = w.d; j >= 2; --j) {
    const size_t m = (size_t)1 << (j - 1);
    w.la[j - 1].assign(m, 0.0);
    w.lb[j - 1].assign(m, 0.0);
    for (size_t k = 0; k < m; ++k) {
      w.la[j - 1][k] = w.la[j][k] + w.la[j][k | m];
      w.lb[j - 1][k] = w.lb[j][k] + w.lb[j][k | m];
    }
  }
  return w;
}

// ---------------------------------------------------------------------------
// Sampling many states at once (exact mode).  Each state's DFS runs on a host
// thread of its own, in its own RNG order, and stops at every window it
// needs; once every running DFS waits, the dispatcher (the calling thread,
// the only one that touches the GPU) evaluates all waiting windows -- the
// projections of all states in one launch per row width
// (k_qproject_var_jobs), all window passes in one launch plan (one
// k_pairs_var launch per row width) -- and resumes them.  Each state draws
// exactly what sample_exact draws for it alone.
struct SampleBatch {
  struct Req {
    SampleCtx* cx;
    int root;
    bool tail;
    bool done = false;
    Window* out;
  };
  std::mutex mu;
  std::condition_variable cv_disp, cv_work;
  int running = 0;  // DFS threads not finished
  int waiting = 0;  // of which waiting for a window
  bool abort = false;
  std::vector<Req*> pending;
  // staging of all states' window passes (one upload per round) and the
  // job tables after it
  PinnedBuf stage_h, res_h, tab_h;
  DevBuf stage_d;
  size_t stage_bytes = 0, jobs_off = 0, tab_cap = 0;
  // the round's pair plan tables and the tails' work buffers, kept across
  // rounds (grown when a round needs more)
  PlanStage plan_stage;
  DevBuf trec, tal, tab;
  size_t trec_cap = 0, tal_cap = 0, tabd_cap = 0;
  int64_t rounds = 0, windows = 0, tails = 0;
  double round_s = 0.0, wait_s = 0.0;  // in batch_round / waiting for the DFS threads
  double sync_s = 0.0;                 // of round_s: waiting for the round's launches
};

struct BatchAbort {};

static Window batch_window(SampleCtx& cx, int root, bool tail) {
  SampleBatch& b = *cx.batch;
  Window w;
  SampleBatch::Req r{&cx, root, tail, false, &w};
  std::unique_lock<std::mutex> lk(b.mu);
  if (b.abort) throw BatchAbort{};
  b.pending.push_back(&r);
  ++b.waiting;
  b.cv_disp.notify_one();
  b.cv_work.wait(lk, [&] { return r.done; });
  if (b.abort) throw BatchAbort{};
  return w;
}

// one round: every waiting state's window or exact tail, all enqueued
// before one wait
static void batch_round(SampleBatch& b, std::vector<SampleBatch::Req*>& reqs) {
  std::vector<SampleBatch::Req*> win, tl;
  for (SampleBatch::Req* r : reqs) (r->tail ? tl : win).push_back(r);
  size_t jo = b.jobs_off;  // staging cursor after the states' own slices
  auto take = [&](size_t bytes) {
    const size_t o = jo;
    jo += (bytes + 255) & ~(size_t)255;
    if (jo > b.stage_bytes) fail(NX_E_INTERNAL, "sampling batch staging too small");
    return o;
  };
  // windows: host staging of every pass, projection jobs grouped by (W, VW)
  std::vector<int> nvs;
  std::vector<VarJob> vj(win.size());
  std::map<std::pair<int, int>, std::vector<QProjJobDev>> groups;
  for (size_t i = 0; i < win.size(); ++i) {
    SampleBatch::Req* r = win[i];
    Window& w = *r->out;
    std::vector<int> fq, fb, lq;
    window_args(*r->cx, r->root, w, fq, fb, lq);
    ProjPlan& pp = r->cx->pp;
    nvs.push_back(pp.stage(fq, fb, lq, nullptr));
    groups[{pp.projv.W, pp.VW}].push_back(pp.qproj_job());
    vj[i].X = vj[i].Y = &pp.projv;
    vj[i].cx = vj[i].cy = coef_of(pp.s);
    vj[i].nvar = nvs[i];
    vj[i].split_ab = 1;
    vj[i].out_base = (int64_t)i * 2 * 64 * pp.VW;
  }
  std::vector<std::tuple<int, int, size_t, int64_t, int>> qlaunch;  // W, VW, table offset, blocks, jobs
  for (auto& g : groups) {
    int64_t blocks = 0;
    for (QProjJobDev& q : g.second) {
      q.block0 = blocks;
      blocks += (q.in.chi + 3) / 4;
    }
    const size_t o = take(g.second.size() * sizeof(QProjJobDev));
    std::memcpy(b.stage_h.as<char>(o), g.second.data(), g.second.size() * sizeof(QProjJobDev));
    qlaunch.emplace_back(g.first.first, g.first.second, o, blocks, (int)g.second.size());
  }
  // tails: the 2^d completions of each path, grouped by row words
  struct TailAt {
    int64_t nv;
    size_t ab;  // offset (doubles) of its leaf norms in the round's results
  };
  std::vector<TailAt> tat(tl.size());
  std::map<int, std::vector<TailJobDev>> tgroups;
  size_t nrec = 0, nal = 0, nab = 0;
  for (size_t i = 0; i < tl.size(); ++i) {
    const SampleCtx& cx = *tl[i]->cx;
    const int d = (int)cx.q.size() - tl[i]->root;
    TailJobDev t = tail_job(cx.s, d);
    const size_t ob = take(32 + 4 * (size_t)d);
    uint64_t* base = b.stage_h.as<uint64_t>(ob);
    for (int w = 0; w < 4; ++w) base[w] = 0ull;
    for (int k = 0; k < tl[i]->root; ++k)
      if (cx.path[k]) base[cx.q[k] >> 6] |= 1ull << (cx.q[k] & 63);
    std::memcpy(b.stage_h.as<char>(ob + 32), cx.q.data() + tl[i]->root, 4 * (size_t)d);
    t.base = (const uint64_t*)(b.stage_d.as<char>() + ob);
    t.qs = (const int32_t*)(b.stage_d.as<char>() + ob + 32);
    // offsets into the round's buffers for now, pointers once allocated
    t.recs = (uint32_t*)(uintptr_t)nrec;
    t.al = (double*)(uintptr_t)nal;
    t.ab = (double*)(uintptr_t)nab;
    tat[i] = {t.nv, nab};
    nrec += (size_t)cx.s->chi() * TAIL_REC;
    nal += (size_t)t.nchunks * t.nv * 4;
    nab += 2 * (size_t)t.nv;
    tgroups[t.q.W].push_back(t);
  }
  DevBuf &trec = b.trec, &tal = b.tal, &tab = b.tab;
  std::vector<std::tuple<int, size_t, int, int64_t, int64_t, int64_t>> tlaunch;  // W, table, jobs, terms, blocks, strings
  if (!tl.empty()) {
    if (b.trec_cap < nrec) {
      b.trec_cap = 2 * nrec;
      trec.alloc(b.trec_cap * 4);
    }
    if (b.tal_cap < nal) {
      b.tal_cap = 2 * nal;
      tal.alloc(b.tal_cap * 8);
    }
    if (b.tabd_cap < nab) {
      b.tabd_cap = 2 * nab;
      tab.alloc(b.tabd_cap * 8);
    }
    if (b.tab_cap < nab) {
      b.tab_h.alloc(nab * 8);
      b.tab_cap = nab;
    }
    for (auto& g : tgroups) {
      int64_t term0 = 0, str0 = 0, blocks = 0;
      for (TailJobDev& t : g.second) {
        t.recs = trec.as<uint32_t>() + (uintptr_t)t.recs;
        t.al = tal.as<double>() + (uintptr_t)t.al;
        t.ab = tab.as<double>() + (uintptr_t)t.ab;
        t.term0 = term0;
        t.str0 = str0;
        t.block0 = blocks;
        term0 += t.q.chi;
        str0 += t.nv;
        blocks += ((t.nv + 1023) / 1024) * t.nchunks;
      }
      const size_t o = take(g.second.size() * sizeof(TailJobDev));
      std::memcpy(b.stage_h.as<char>(o), g.second.data(), g.second.size() * sizeof(TailJobDev));
      tlaunch.emplace_back(g.first, o, (int)g.second.size(), term0, blocks, str0);
    }
  }
  // the windows' pair plan, staged (its tables uploaded with one copy)
  VarPlan plan;
  if (!win.empty()) {
    int VW = 1;
    int64_t target = 0;
    for (size_t i = 0; i < win.size(); ++i) {
      VW = std::max(VW, win[i]->cx->pp.VW);
      target = std::max<int64_t>(target, 3 * ProjPlan::resident_waves(win[i]->cx->pp.projv.W, win[i]->cx->pp.VW));
    }
    plan = make_plan(vj, VW, 8, std::max<int64_t>(target, 1024), 1.0, &b.plan_stage);
    // the plan's jobs carry each window's unit constraints, as a lone pass does
    for (size_t i = 0; i < win.size(); ++i) {
      const PairJobDev& src = win[i]->cx->pp.hj[0];
      PairJobDev& d = b.plan_stage.jobs_h[i];
      d.X = src.X;
      d.Y = src.Y;
      d.V = src.V;
      d.nvar = src.nvar;
      d.nunit = src.nunit;
      d.uq = src.uq;
      d.ub = src.ub;
      d.termvar = src.termvar;
      d.rec = nullptr;
      for (int w = 0; w < 4; ++w) d.removed[w] = src.removed[w];
    }
  }
  // one upload of the staged block (and of the plan's tables), then the launches
  h2d(b.stage_d.p, b.stage_h.p, jo);
  if (!win.empty()) h2d(b.plan_stage.d.p, b.plan_stage.h.p, b.plan_stage.used);
  {
    ProfScope ps("k_qproject_var");
    for (auto& L : qlaunch)
      hip_check(launch_qproject_var_jobs(std::get<0>(L), std::get<1>(L),
                                         (const QProjJobDev*)(b.stage_d.as<char>() + std::get<2>(L)), std::get<4>(L),
                                         std::get<3>(L), stream()),
                "k_qproject_var_jobs");
    ps.end();
  }
  if (!tl.empty()) {
    ProfScope ps("k_tail");
    for (auto& L : tlaunch)
      hip_check(launch_tails(std::get<0>(L), (const TailJobDev*)(b.stage_d.as<char>() + std::get<1>(L)),
                             std::get<2>(L), std::get<3>(L), std::get<4>(L), std::get<5>(L), stream()),
                "launch_tails");
    ps.end();
    d2h(b.tab_h.p, tab.p, nab * 8);
  }
  if (!win.empty()) {
    ProfScope ps("k_pairs_var");
    plan.enqueue(stream(), nullptr, b.res_h.as<double>(), false);
    ps.end();
  }
  const auto ts = std::chrono::steady_clock::now();
  sync();
  b.sync_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - ts).count();
  for (size_t i = 0; i < win.size(); ++i) {
    const double* r = b.res_h.as<double>() + vj[i].out_base;
    window_fill(*win[i]->out, r, r + nvs[i]);
    ++win[i]->cx->passes;
  }
  for (size_t i = 0; i < tl.size(); ++i) {
    Window& w = *tl[i]->out;
    w.root = tl[i]->root;
    w.d = (int)tl[i]->cx->q.size() - w.root;
    const double* A = b.tab_h.as<double>() + tat[i].ab;
    window_fill(w, A, A + tat[i].nv);
    ++tl[i]->cx->tails;
  }
  b.windows += (int64_t)win.size();
  b.tails += (int64_t)tl.size();
}

// dispatcher: until every DFS has finished, wait for all running ones to
// stop at a window, then evaluate the round
static void batch_dispatch(SampleBatch& b) {
  for (;;) {
    std::vector<SampleBatch::Req*> reqs;
    const auto tw = std::chrono::steady_clock::now();
    {
      std::unique_lock<std::mutex> lk(b.mu);
      b.cv_disp.wait(lk, [&] { return b.waiting == b.running; });
      if (b.running == 0) return;
      reqs.swap(b.pending);
    }
    const auto tr = std::chrono::steady_clock::now();
    b.wait_s += std::chrono::duration<double>(tr - tw).count();
    bool ok = true;
    std::string err;
    int code = NX_OK;
    try {
      batch_round(b, reqs);
      ++b.rounds;
      b.round_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - tr).count();
    } catch (const NxError& e) {
      ok = false;
      code = e.code;
      err = g_err;
    } catch (const std::exception& e) {
      ok = false;
      code = NX_E_INTERNAL;
      err = e.what();
    }
    {
      std::lock_guard<std::mutex> lk(b.mu);
      if (!ok) b.abort = true;
      for (SampleBatch::Req* r : reqs) r->done = true;
      b.waiting -= (int)reqs.size();
    }
    b.cv_work.notify_all();
    if (!ok) {
      // every DFS thread was waiting in this round: they wake, see `abort`,
      // unwind (BatchAbort) and finish; a later request throws at once
      std::unique_lock<std::mutex> lk(b.mu);
      b.cv_disp.wait(lk, [&] { return b.running == 0; });
      g_err = err;
      throw NxError{code};
    }
  }
}

// sampling.rs:80-189.  The norm of a child is ||Pi psi||^2 for the projector
// onto the outcomes chosen so far (the reference projects, then discards
// the measured qubit, which leaves the norm unchanged); each window pass
// yields the norms of all nodes up to `look` levels below its root, so the
// draws below happen in the reference's DFS order without further passes.
static void recursive_sample(SampleCtx& cx, size_t idx, uint64_t shots, uint64_t lo, uint64_t hi, const Window* w,
                             uint64_t rel) {
  if (shots == 0) return;
  if (idx == cx.q.size()) {
    cx.bits.push_back(lo);
    cx.bits.push_back(hi);
    cx.counts.push_back(shots);
    return;
  }
  Window own;
  double pz = 0.5;  // both children hold half of a uniform marginal's mass
  int j = 0;        // children's relative depth in the window
  if ((int)idx >= cx.uniform) {
    if (!w || (int)idx + 1 > w->root + w->d) {
      const bool tail = !cx.ref && (int)cx.q.size() - (int)idx <= cx.tail_m;
      own = cx.batch ? batch_window(cx, (int)idx, tail)
                     : (cx.ref ? open_window_ref(cx, (int)idx) : (tail ? open_tail(cx, (int)idx)
                                                                       : open_window(cx, (int)idx)));
      sample_progress(cx, "window", (int)idx);
      w = &own;
      rel = 0;
    }
    j = (int)idx - w->root + 1;
    const uint64_t r1 = rel | ((uint64_t)1 << (j - 1));
    const double nz = w->ref ? w->nr[j][rel] : cx.pp.value(w->la[j][rel], w->lb[j][rel]);
    const double no = w->ref ? w->nr[j][r1] : cx.pp.value(w->la[j][r1], w->lb[j][r1]);
    pz = nz / (nz + no);
    if (!std::isnan(pz)) pz = pz < 0.0 ? 0.0 : (pz > 1.0 ? 1.0 : pz);
    if (!(pz >= 0.0)) fail(NX_E_VALUE, "p < 0 or is NaN in binomial distribution");
  }
  const uint64_t nzeros = binomial(*cx.rng, shots, pz);
  const uint64_t nones = shots - nzeros;
  const int b = cx.bit[idx];
  if (nzeros > 0) {
    cx.path.push_back(0);
    recursive_sample(cx, idx + 1, nzeros, lo, hi, w, rel);
    cx.path.pop_back();
  }
  if (nones > 0) {
    uint64_t nlo = lo, nhi = hi;
    if (b < 64) nlo |= 1ull << b;
    else nhi |= 1ull << (b - 64);
    cx.path.push_back(1);
    recursive_sample(cx, idx + 1, nones, nlo, nhi, w, j ? rel | ((uint64_t)1 << (j - 1)) : 0);
    cx.path.pop_back();
  }
}

}  // namespace nx

// ---------------------------------------------------------------------------
// a batch of states with their Pauli lists, evaluated by one launch plan
struct nx_batch {
  std::vector<nx_state*> states;  // private clones
  std::vector<nx::VarSet> vars;
  std::vector<int64_t> state_out;
  std::vector<int32_t> state_np;
  nx::VarPlan plan;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int64_t nout = 0, waves = 0;
  int J = 16, VW = 1;
  ~nx_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    plan = nx::VarPlan();
    vars.clear();
    for (nx_state* s : states) delete s;
  }
};

// ---------------------------------------------------------------------------
// C-ABI
using namespace nx;

#define NX_TRY(...)                                   \
  try {                                               \
    __VA_ARGS__;                                      \
    return NX_OK;                                     \
  } catch (const NxError& e) {                        \
    return e.code;                                    \
  } catch (const std::bad_alloc&) {                   \
    g_err = "host out of memory";                     \
    return NX_E_INTERNAL;                             \
  } catch (const std::exception& e) {                 \
    g_err = e.what();                                 \
    return NX_E_INTERNAL;                             \
  }

extern "C" {

const char* nx_last_error(void) { return g_err.c_str(); }

int nx_init(int device) { NX_TRY(t_dev = device; (void)ctx()) }

int nx_set_stream(void* stream_, int use) {
  NX_TRY({
    (void)ctx();
    hip_check(hipStreamSynchronize(stream()), "hipStreamSynchronize");  // the previous stream's work first
    t_user_stream = use != 0;
    t_stream = (hipStream_t)stream_;
  })
}

int nx_state_compile(int num_qubits, int64_t ngates, const int32_t* gates, nx_state** out) {
  NX_TRY(*out = compile(num_qubits, ngates, gates))
}

int nx_state_clone(const nx_state* s, nx_state** out) { NX_TRY(*out = clone_state(s)) }

void nx_state_free(nx_state* s) { delete s; }

int nx_state_num_qubits(const nx_state* s, int32_t* out) { NX_TRY(*out = s->n) }

int nx_state_stabilizer_rank(const nx_state* s, int64_t* out) { NX_TRY(*out = s->chi()) }

int nx_state_apply_gates(nx_state* s, int64_t ngates, const int32_t* gates) {
  NX_TRY({
    std::vector<ProgOp> prog;
    for (int64_t i = 0; i < ngates; ++i) {
      const int op = gates[4 * i], q0 = gates[4 * i + 1], q1 = gates[4 * i + 2];
      bool bad = false;
      std::string msg;
      int code = NX_OK;
      if (op == 11 || op == 12 || op == 13) {
        bad = true;
        char buf[160];
        snprintf(buf, sizeof buf, "Direct application of non-Clifford gate %s is not supported.", gate_name(op));
        msg = buf;
        code = NX_E_VALUE;
      } else if (op < 0 || op > 10) {
        bad = true;
        msg = "unknown gate code";
        code = NX_E_VALUE;
      } else if (s->chi() > 0) {
        try {
          validate_gate(op, q0, q1, s->n);
        } catch (const NxError& e) {
          bad = true;
          msg = g_err;
          code = e.code;
        }
      }
      if (bad) {
        apply_program_all(s, prog);  // gates before the failing one are applied
        g_err = msg;
        throw NxError{code};
      }
      prog.push_back({op, q0, q1, 0});
    }
    apply_program_all(s, prog);
  })
}

int nx_state_norm_squared(const nx_state* s, double* out) { NX_TRY(*out = norm_squared(s)) }

int nx_state_norm(const nx_state* s, double* out) { NX_TRY(*out = std::sqrt(norm_squared(s))) }

int nx_state_norm_noalloc(const nx_state* s, double* out) { NX_TRY(*out = std::sqrt(norm_squared_noalloc(s))) }

int nx_state_norm_async(const nx_state* s, double* d_out, void* stream_) {
  NX_TRY({
    StreamOrder order(stream_);  // NULL: the null stream
    norm_enqueue(s, d_out);
    order.done();
  })
}

int nx_state_exp_value(const nx_state* s, int dense, int32_t len, const int32_t* ops, const int32_t* qubits,
                       double* out) {
  NX_TRY(*out = exp_value(s, pauli_prog(s, dense, len, ops, qubits)))
}

int nx_state_inner_product(const nx_state* a, const nx_state* b, double* re, double* im) {
  NX_TRY({
    cplx res(0.0, 0.0);
    if (a->chi() > 0 && b->chi() > 0) {
      if (a->n != b->n)
        fail(NX_E_VALUE, "The qubit counts of the two states must match for calculating inner product, got %d and %d.",
             a->n, b->n);
      const QForms& qa = qform_of(a);
      const QForms& qb = qform_of(b);
      CoefDev ca = coef_of(a), cb = coef_of(b);
      res = run_pairs(qa, qb, a->chi(), b->chi(), false, /*fixed bra, complex*/ 0, ca, cb);
    }
    res = cmul(cmul(res, std::conj(a->g)), b->g);
    *re = res.real();
    *im = res.imag();
  })
}

// enqueues the amplitude pass on the library stream (no host wait)
static void amplitudes_enqueue(const nx_state* s, int64_t nbits, const uint64_t* d_bits, double* d_out,
                               uint8_t* d_per_term, int enumerate) {
  if (nbits <= 0) return;  // no strings: no launch (a grid of zero blocks is not a valid launch)
  const void* tab = amp_tables_of(s);
  const QForms& q = qform_of(s);
  CoefDev c = coef_of(s);
  // the term loop is split into chunks when the strings alone give too
  // few blocks to fill the GPU
  const int nchunks = amp_chunks(q.W, nbits, s->chi());
  DevBuf partial(nchunks > 1 ? (size_t)nchunks * nbits * 16 : 0);
  ProfScope ps("k_amplitude");
  hip_check(launch_amplitude(q.desc(), tab, c.c, d_bits, nbits, nchunks, ctx()->tabs, (double2*)d_out,
                             partial.as<double2>(), d_per_term, enumerate, s->g.real(), s->g.imag(), stream()),
            "k_amplitude");
  ps.end();
  ProfScope pr("k_amp_reduce");
  hip_check(launch_amp_reduce(partial.as<double2>(), nbits, (int)nchunks, (double2*)d_out, s->g.real(), s->g.imag(),
                              stream()),
            "k_amp_reduce");
  pr.end();
  // `partial` returns to the pool now; its next user is ordered after these
  // launches on the same stream
}

static void amplitudes_impl(const nx_state* s, int64_t nbits, const uint64_t* d_bits, double* d_out,
                            uint8_t* d_per_term, int enumerate) {
  amplitudes_enqueue(s, nbits, d_bits, d_out, d_per_term, enumerate);
  sync();
}

int nx_state_amplitudes(const nx_state* s, int64_t nbits, const uint64_t* bits, int32_t words, double* out,
                        uint8_t* per_term, int32_t* per_term_r) {
  NX_TRY({
    const int W = words_for(s->n);
    if (words < W) fail(NX_E_VALUE, "bitstrings need %d words for %d qubits", W, s->n);
    if (nbits < 0) fail(NX_E_VALUE, "negative bitstring count %lld", (long long)nbits);
    std::vector<uint64_t> packed((size_t)nbits * W);
    for (int64_t b = 0; b < nbits; ++b)
      for (int w = 0; w < W; ++w) {
        uint64_t x = bits[(size_t)b * words + w];
        const int hi = s->n - 64 * w;
        if (hi < 64) x &= hi <= 0 ? 0ull : ((1ull << hi) - 1ull);
        packed[(size_t)b * W + w] = x;
      }
    DevBuf dbits(packed.size() * 8), dout((size_t)nbits * 16);
    DevBuf dpt(per_term ? (size_t)s->chi() * nbits : 0);
    h2d(dbits.p, packed.data(), packed.size() * 8);
    if (s->chi() == 0) {
      std::memset(out, 0, (size_t)nbits * 16);
    } else {
      amplitudes_impl(s, nbits, dbits.as<uint64_t>(), dout.as<double>(), per_term ? dpt.as<uint8_t>() : nullptr, 0);
      d2h(out, dout.p, (size_t)nbits * 16);
      if (per_term) d2h(per_term, dpt.p, (size_t)s->chi() * nbits);
      if (per_term_r) {
        std::vector<int32_t> k(s->chi());
        d2h(k.data(), qform_of(s).k.p, s->chi() * 4);
        sync();
        std::memcpy(per_term_r, k.data(), s->chi() * 4);
      }
    }
    sync();
  })
}

int nx_state_amplitudes_device(const nx_state* s, int64_t nbits, const uint64_t* d_bits, int32_t words,
                               double* d_out) {
  NX_TRY({
    if (words != words_for(s->n)) fail(NX_E_VALUE, "device bitstrings must use %d words", words_for(s->n));
    if (nbits < 0) fail(NX_E_VALUE, "negative bitstring count %lld", (long long)nbits);
    if (nbits == 0) return NX_OK;
    if (s->chi() == 0) {
      hip_check(hipMemsetAsync(d_out, 0, (size_t)nbits * 16, stream()), "hipMemsetAsync");
    } else {
      amplitudes_impl(s, nbits, d_bits, d_out, nullptr, 0);
    }
    sync();
  })
}

int nx_state_amplitudes_async(const nx_state* s, int64_t nbits, const uint64_t* d_bits, int32_t words, double* d_out,
                              void* stream_) {
  NX_TRY({
    if (words != words_for(s->n)) fail(NX_E_VALUE, "device bitstrings must use %d words", words_for(s->n));
    if (nbits < 0) fail(NX_E_VALUE, "negative bitstring count %lld", (long long)nbits);
    if (nbits == 0) return NX_OK;
    // the caller's stream (NULL: the null stream): its earlier work (the
    // inputs) comes first, its later work sees the results
    StreamOrder order(stream_);
    if (s->chi() == 0) {
      hip_check(hipMemsetAsync(d_out, 0, (size_t)nbits * 16, stream()), "hipMemsetAsync");
    } else {
      amplitudes_enqueue(s, nbits, d_bits, d_out, nullptr, 0);
    }
    order.done();
  })
}

int nx_state_statevector(const nx_state* s, double* out) {
  NX_TRY({
    if (s->n > 28) fail(NX_E_VALUE, "Calculating the statevector for a state with %d qubits is not feasible.", s->n);
    const int64_t dim = (int64_t)1 << s->n;
    if (s->chi() == 0) {
      std::memset(out, 0, (size_t)dim * 16);
    } else {
      DevBuf dout((size_t)dim * 16);
      amplitudes_impl(s, dim, nullptr, dout.as<double>(), nullptr, 1);
      d2h(out, dout.p, (size_t)dim * 16);
      sync();
    }
  })
}

int nx_state_project(nx_state* s, int32_t qubit, int32_t outcome, int32_t normalized) {
  NX_TRY({
    project_unnormalized(s, qubit, outcome);
    if (normalized) {  // projection.rs:10-21
      const double nrm = std::sqrt(norm_squared(s));
      if (std::fabs(nrm) < 1e-12 || std::isnan(nrm))
        fail(NX_E_VALUE, "Impossible projection on qubit %d: cannot project determined state |%d> onto |%d>.", qubit,
             outcome ? 0 : 1, outcome ? 1 : 0);
      s->g = cplx(s->g.real() / nrm, s->g.imag() / nrm);
    }
  })
}

int nx_state_discard(nx_state* s, int32_t qubit) { NX_TRY(discard(s, qubit)) }

int nx_state_measure(nx_state* s, int32_t nq, const int32_t* qargs, const uint8_t* seed32, uint8_t* outcomes) {
  NX_TRY({
    validate_qargs(s, nq, qargs);
    uint8_t seed[32];
    seed_or_entropy(seed32, seed);
    ChaCha12Rng rng(seed);
    for (int i = 0; i < nq; ++i) outcomes[i] = (uint8_t)measure_single(s, qargs[i], rng);
  })
}

// sampling.rs:13-189.  ref: every node norm as the reference's FP64 sum
// (Scalar coefficients; seeded outcomes equal the reference's); otherwise
// exact node norms, certified uniform levels and exact tails.
static int sample_impl(const nx_state* s, int32_t nq, const int32_t* qargs, uint64_t shots, const uint8_t* seed32,
                       int64_t max_out, uint64_t* out_bits, uint64_t* out_counts, int64_t* nout, bool ref) {
  NX_TRY({
    validate_qargs(s, nq, qargs);
    uint8_t seed[32];
    seed_or_entropy(seed32, seed);
    if (nq > 128) fail(NX_E_VALUE, "Sampling more than 128 qubits is not supported.");
    ChaCha12Rng rng(seed);
    SampleCtx cx;
    cx.rng = &rng;
    std::vector<std::pair<int, int>> sorted;
    for (int i = 0; i < nq; ++i) sorted.push_back({qargs[i], i});
    std::stable_sort(sorted.begin(), sorted.end(), [](auto& a, auto& b) { return a.first > b.first; });
    for (auto& p : sorted) {
      cx.q.push_back(p.first);
      cx.bit.push_back(p.second);
    }
    if (shots > 0 && nq > 0 && s->chi() == 0)  // every norm is 0: p = 0/0
      fail(NX_E_VALUE, "p < 0 or is NaN in binomial distribution");
    if (shots > 0 && nq > 0) {
      cx.s = s;
      cx.ref = ref && !s->is_complex;
      cx.progress = std::getenv("NX_SAMPLE_PROGRESS") != nullptr;
      cx.t0 = std::chrono::steady_clock::now();
      if (cx.ref) {
        const int64_t chi = s->chi();
        if (chi * (chi + 1) / 2 > ref_fold_max_pairs())
          fail(NX_E_VALUE, "reference-order sampling holds one record per term pair: at most 23,169 terms, got %lld",
               (long long)chi);
      } else {
        // at least the last level keeps its window: it also reports a zero state
        cx.uniform = std::min(uniform_prefix(s, cx.q), nq - 1);
        sample_progress(cx, "certified uniform levels", cx.uniform);
        // exact tails need every qubit sampled (leaf = one amplitude) and
        // Scalar coefficients; 2^14 completions cost chi 2^14 amplitude
        // evaluations against ceil(14/6) windows of chi^2 / 2 pairs
        cx.tail_m = (!s->is_complex && nq == s->n) ? 14 : 0;
      }
      cx.pp.init(s, nq, cx.look);
      recursive_sample(cx, 0, shots, 0, 0, nullptr, 0);
    } else if (shots > 0) {
      cx.bits.push_back(0);
      cx.bits.push_back(0);
      cx.counts.push_back(shots);
    }
    const int64_t k = (int64_t)cx.counts.size();
    // the caller learns the capacity it needs even when this call fails
    *nout = k;
    if (k > max_out) fail(NX_E_VALUE, "sample output buffer too small (%lld outcomes)", (long long)k);
    std::memcpy(out_bits, cx.bits.data(), k * 16);
    std::memcpy(out_counts, cx.counts.data(), k * 8);
    *nout = k;
  })
}

int nx_state_sample(const nx_state* s, int32_t nq, const int32_t* qargs, uint64_t shots, const uint8_t* seed32,
                    int64_t max_out, uint64_t* out_bits, uint64_t* out_counts, int64_t* nout) {
  return sample_impl(s, nq, qargs, shots, seed32, max_out, out_bits, out_counts, nout, true);
}

int nx_state_sample_exact(const nx_state* s, int32_t nq, const int32_t* qargs, uint64_t shots, const uint8_t* seed32,
                          int64_t max_out, uint64_t* out_bits, uint64_t* out_counts, int64_t* nout) {
  return sample_impl(s, nq, qargs, shots, seed32, max_out, out_bits, out_counts, nout, false);
}

int nx_batch_sample_exact(int32_t nstates, const nx_state* const* states, const int32_t* nq, const int32_t* qargs,
                          const uint64_t* shots, const uint8_t* seeds, int64_t max_out, uint64_t* out_bits,
                          uint64_t* out_counts, int64_t* nout, int64_t* stats) {
  NX_TRY({
    if (nstates < 0) fail(NX_E_VALUE, "negative state count");
    std::vector<std::unique_ptr<SampleCtx>> cxs;
    std::vector<std::unique_ptr<ChaCha12Rng>> rngs;
    std::vector<uint64_t> nshots(nstates);
    SampleBatch b;
    size_t qoff = 0, soff = 0;
    std::vector<size_t> stage_at(nstates, 0);
    for (int32_t i = 0; i < nstates; ++i) {
      const nx_state* s = states[i];
      if (!s) fail(NX_E_VALUE, "state %d is null", i);
      if (nq[i] < 0) fail(NX_E_VALUE, "state %d: negative qubit count", i);
      validate_qargs(s, nq[i], qargs + qoff);
      if (nq[i] > 128) fail(NX_E_VALUE, "Sampling more than 128 qubits is not supported.");
      uint8_t seed[32];
      seed_or_entropy(seeds ? seeds + (size_t)32 * i : nullptr, seed);
      rngs.emplace_back(new ChaCha12Rng(seed));
      cxs.emplace_back(new SampleCtx());
      SampleCtx& cx = *cxs.back();
      cx.rng = rngs.back().get();
      std::vector<std::pair<int, int>> sorted;
      for (int k = 0; k < nq[i]; ++k) sorted.push_back({qargs[qoff + k], k});
      std::stable_sort(sorted.begin(), sorted.end(), [](auto& a, auto& c) { return a.first > c.first; });
      for (auto& pq : sorted) {
        cx.q.push_back(pq.first);
        cx.bit.push_back(pq.second);
      }
      qoff += (size_t)nq[i];
      nshots[i] = shots[i];
      if (shots[i] > 0 && nq[i] > 0 && s->chi() == 0)  // every norm is 0: p = 0/0
        fail(NX_E_VALUE, "p < 0 or is NaN in binomial distribution");
      if (shots[i] > 0 && nq[i] > 0) {
        cx.s = s;
        stage_at[i] = soff;
        soff += ProjPlan::stage_size(nq[i], cx.look);
      }
    }
    // one staging block for every state's window passes, the projection job
    // tables after it
    b.jobs_off = soff;
    b.stage_bytes = soff + (size_t)std::max(nstates, 1) * (sizeof(QProjJobDev) + sizeof(TailJobDev) + 1024) +
                    32 * 256;
    b.stage_h.alloc(b.stage_bytes);
    b.stage_d.alloc(b.stage_bytes);
    b.res_h.alloc((size_t)std::max(nstates, 1) * 2 * 64 * 2 * 8);
    std::vector<int32_t> live;
    for (int32_t i = 0; i < nstates; ++i) {
      SampleCtx& cx = *cxs[i];
      if (!cx.s) {
        if (nshots[i] > 0) {  // no qubits: one empty outcome
          cx.bits.push_back(0);
          cx.bits.push_back(0);
          cx.counts.push_back(nshots[i]);
        }
        continue;
      }
      const nx_state* s = cx.s;
      const int n = (int)cx.q.size();
      cx.batch = &b;
      cx.uniform = std::min(uniform_prefix(s, cx.q), n - 1);
      cx.tail_m = (!s->is_complex && n == s->n) ? 14 : 0;
      cx.pp.init(s, n, cx.look, b.stage_h.as<char>(stage_at[i]), b.stage_d.as<char>() + stage_at[i]);
      live.push_back(i);
    }
    // the DFS threads; this thread dispatches their windows
    std::vector<int> codes(nstates, NX_OK);
    std::vector<std::string> msgs(nstates);
    b.running = (int)live.size();
    std::vector<std::thread> th;
    for (int32_t i : live) {
      th.emplace_back([&, i] {
        SampleCtx& cx = *cxs[i];
        try {
          recursive_sample(cx, 0, nshots[i], 0, 0, nullptr, 0);
        } catch (const BatchAbort&) {
        } catch (const NxError& e) {
          codes[i] = e.code;
          msgs[i] = g_err;
        } catch (const std::exception& e) {
          codes[i] = NX_E_INTERNAL;
          msgs[i] = e.what();
        }
        {
          std::lock_guard<std::mutex> lk(b.mu);
          --b.running;
        }
        b.cv_disp.notify_one();
      });
    }
    int dcode = NX_OK;
    std::string dmsg;
    try {
      batch_dispatch(b);
    } catch (const NxError& e) {
      dcode = e.code;
      dmsg = g_err;
    }
    for (std::thread& t : th) t.join();
    if (dcode != NX_OK) fail(dcode, "%s", dmsg.c_str());
    for (int32_t i = 0; i < nstates; ++i)
      if (codes[i] != NX_OK) fail(codes[i], "state %d: %s", i, msgs[i].c_str());
    int64_t total = 0;
    for (int32_t i = 0; i < nstates; ++i) {
      nout[i] = (int64_t)cxs[i]->counts.size();
      total += nout[i];
    }
    if (stats) {
      stats[0] = b.rounds;
      stats[1] = b.windows;
      stats[2] = b.tails;
      stats[3] = (int64_t)(b.round_s * 1e6);
      stats[4] = (int64_t)(b.wait_s * 1e6);
      stats[5] = (int64_t)(b.sync_s * 1e6);
    }
    // the caller learns the capacity it needs (sum of nout) even when this fails
    if (total > max_out) fail(NX_E_VALUE, "sample output buffer too small (%lld outcomes)", (long long)total);
    int64_t o = 0;
    for (int32_t i = 0; i < nstates; ++i) {
      const int64_t k = nout[i];
      if (k) {
        std::memcpy(out_bits + 2 * o, cxs[i]->bits.data(), (size_t)k * 16);
        std::memcpy(out_counts + o, cxs[i]->counts.data(), (size_t)k * 8);
      }
      o += k;
    }
  })
}

int nx_state_completion_norms(const nx_state* s, const uint64_t* base, int32_t d, const int32_t* qs, double* A,
                              double* B) {
  NX_TRY({
    if (s->is_complex) fail(NX_E_VALUE, "completion norms need Scalar coefficients");
    if (s->n > 128) fail(NX_E_VALUE, "completion norms are limited to 128 qubits");
    if (d < 1 || d > 16) fail(NX_E_VALUE, "between 1 and 16 completion qubits are supported, got %d", d);
    std::vector<int32_t> q(qs, qs + d);
    validate_qargs(s, d, qs);
    const int W = words_for(s->n);
    std::vector<uint64_t> b(base, base + W);
    for (int i = 0; i < d; ++i)
      if ((b[q[i] >> 6] >> (q[i] & 63)) & 1ull) fail(NX_E_VALUE, "completion qubit %d is set in the base", q[i]);
    if (s->chi() == 0) {
      std::fill(A, A + ((int64_t)1 << d), 0.0);
      std::fill(B, B + ((int64_t)1 << d), 0.0);
      return NX_OK;
    }
    TailBufs tb;
    std::vector<double> va, vb;
    tail_leaves(s, tb, b, q, va, vb);
    std::copy(va.begin(), va.end(), A);
    std::copy(vb.begin(), vb.end(), B);
  })
}

int nx_state_uniform_prefix(const nx_state* s, int32_t nq, const int32_t* qargs, int32_t* out) {
  NX_TRY({
    validate_qargs(s, nq, qargs);
    *out = uniform_prefix(s, std::vector<int>(qargs, qargs + nq));
  })
}

int nx_state_uniform_prefix_given(const nx_state* s, int32_t nfixed, const int32_t* fixed_q, const int32_t* fixed_b,
                                  int32_t nq, const int32_t* qargs, int32_t* out) {
  NX_TRY({
    if (nfixed < 0) fail(NX_E_VALUE, "negative qubit count");
    std::vector<int32_t> all(fixed_q, fixed_q + nfixed);
    all.insert(all.end(), qargs, qargs + nq);
    validate_qargs(s, (int32_t)all.size(), all.data());
    const std::vector<int> order(qargs, qargs + nq);
    if (nfixed == 0) {
      *out = uniform_prefix(s, order);
    } else if (s->chi() == 0) {
      *out = 0;
    } else {
      QForms proj;
      proj.alloc_pair(s->chi(), s->n);
      DevBuf dq((size_t)nfixed * 4), db((size_t)nfixed * 4);
      std::vector<int32_t> b(fixed_b, fixed_b + nfixed);
      for (int32_t& x : b) x = x ? 1 : 0;
      h2d(dq.p, fixed_q, (size_t)nfixed * 4);
      h2d(db.p, b.data(), (size_t)nfixed * 4);
      hip_check(launch_qproject(qform_of(s).desc(), proj.desc(), dq.as<int32_t>(), db.as<int32_t>(), nfixed, stream()),
                "k_qproject");
      *out = uniform_depth(proj, s->n, order, std::vector<int>(fixed_q, fixed_q + nfixed));
    }
  })
}

int nx_state_projection_norms(const nx_state* s, int32_t nfixed, const int32_t* fixed_q, const int32_t* fixed_b,
                              int32_t nlook, const int32_t* look_q, double* out) {
  NX_TRY({
    if (nlook < 0 || nlook > 7) fail(NX_E_VALUE, "between 0 and 7 look-ahead qubits are supported");
    if (nfixed < 0) fail(NX_E_VALUE, "negative qubit count");
    std::vector<int32_t> all(fixed_q, fixed_q + nfixed);
    all.insert(all.end(), look_q, look_q + nlook);
    if (!all.empty()) validate_qargs(s, (int32_t)all.size(), all.data());
    const int nv = 1 << nlook;
    if (s->chi() == 0) {
      for (int k = 0; k < nv; ++k) out[k] = 0.0;
      return NX_OK;
    }
    ProjPlan pp;
    pp.init(s, (int)all.size(), nlook);
    std::vector<int> fq(fixed_q, fixed_q + nfixed), fb(nfixed), lq(look_q, look_q + nlook);
    for (int i = 0; i < nfixed; ++i) fb[i] = fixed_b[i] ? 1 : 0;
    const std::vector<double> r = pp.run(fq, fb, lq);
    for (int k = 0; k < nv; ++k) out[k] = pp.value(r[k], r[nv + k]);
  })
}

int nx_state_export_term(const nx_state* s, int64_t term, uint8_t* G, uint8_t* F, uint8_t* M, uint8_t* gamma,
                         uint8_t* v, uint8_t* sv, uint8_t* phase) {
  NX_TRY({
    if (term < 0 || term >= s->chi()) fail(NX_E_VALUE, "term index out of range");
    const int n = s->n, W = s->forms.W, R = 64 * W;
    std::vector<uint64_t> g(R * W), f(R * W), m(R * W), vv(W), ss(W);
    std::vector<uint8_t> gm(R);
    uint8_t ph;
    const size_t mo = (size_t)term * R * W * 8;
    d2h(g.data(), s->forms.G.as<uint8_t>() + mo, R * W * 8);
    d2h(f.data(), s->forms.F.as<uint8_t>() + mo, R * W * 8);
    d2h(m.data(), s->forms.M.as<uint8_t>() + mo, R * W * 8);
    d2h(gm.data(), s->forms.g.as<uint8_t>() + (size_t)term * R, R);
    d2h(vv.data(), s->forms.v.as<uint8_t>() + (size_t)term * W * 8, W * 8);
    d2h(ss.data(), s->forms.s.as<uint8_t>() + (size_t)term * W * 8, W * 8);
    d2h(&ph, s->forms.ph.as<uint8_t>() + term, 1);
    sync();
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        const size_t wi = (size_t)i * W + (j >> 6);
        G[(size_t)i * n + j] = (uint8_t)((g[wi] >> (j & 63)) & 1);
        F[(size_t)i * n + j] = (uint8_t)((f[wi] >> (j & 63)) & 1);
        M[(size_t)i * n + j] = (uint8_t)((m[wi] >> (j & 63)) & 1);
      }
    for (int i = 0; i < n; ++i) {
      gamma[i] = (uint8_t)(2 * gm[i]);
      v[i] = (uint8_t)((vv[i >> 6] >> (i & 63)) & 1);
      sv[i] = (uint8_t)((ss[i >> 6] >> (i & 63)) & 1);
    }
    *phase = ph;
  })
}

int nx_state_get_scalar_coefficients(const nx_state* s, uint8_t* zero, uint8_t* k, int64_t* r) {
  NX_TRY({
    if (s->is_complex) fail(NX_E_VALUE, "state holds Complex64 coefficients");
    for (int64_t i = 0; i < s->chi(); ++i) {
      zero[i] = s->sc[i].z;
      k[i] = s->sc[i].k;
      r[i] = s->sc[i].r;
    }
  })
}

int nx_state_get_coefficients(const nx_state* s, double* re_im) {
  NX_TRY(for (int64_t i = 0; i < s->chi(); ++i) {
    cplx c = s->coeff_value(i);
    re_im[2 * i] = c.real();
    re_im[2 * i + 1] = c.imag();
  })
}

int nx_state_set_coefficients(nx_state* s, const double* re_im) {
  NX_TRY({
    s->cc.resize(s->chi());
    for (int64_t i = 0; i < s->chi(); ++i) s->cc[i] = cplx(re_im[2 * i], re_im[2 * i + 1]);
    s->is_complex = true;
    s->sc.clear();
    s->c_ok = false;
  })
}

int nx_state_global_factor(const nx_state* s, double* re, double* im) {
  NX_TRY(*re = s->g.real(); *im = s->g.imag())
}

int nx_state_pair_table(const nx_state* s, uint32_t* out) {
  NX_TRY({
    const int64_t chi = s->chi();
    if (chi == 0) return NX_OK;
    const QForms& q = qform_of(s);
    CoefDev c = coef_of(s);
    DevBuf ex((size_t)chi * chi * 4);
    run_pairs(q, q, chi, chi, false, 0, c, c, ex.as<uint32_t>(), chi);
    d2h(out, ex.p, (size_t)chi * chi * 4);
    sync();
  })
}

int nx_state_cross_pair_table(const nx_state* a, const nx_state* b, uint32_t* out) {
  NX_TRY({
    if (a->n != b->n)
      fail(NX_E_VALUE, "The qubit counts of the two states must match for calculating inner product, got %d and %d.",
           a->n, b->n);
    if (a->chi() == 0 || b->chi() == 0) return NX_OK;
    const QForms& qa = qform_of(a);
    const QForms& qb = qform_of(b);
    CoefDev ca = coef_of(a), cb = coef_of(b);
    DevBuf ex((size_t)a->chi() * b->chi() * 4);
    run_pairs(qa, qb, a->chi(), b->chi(), false, 0, ca, cb, ex.as<uint32_t>(), b->chi());
    d2h(out, ex.p, (size_t)a->chi() * b->chi() * 4);
    sync();
  })
}

int nx_state_exp_values(const nx_state* s, int32_t npauli, const int32_t* ops, double* out) {
  NX_TRY({
    if (npauli < 0) fail(NX_E_VALUE, "negative Pauli count %d", npauli);
    if (npauli == 0) return NX_OK;
    PauliSet ps;
    add_dense_paulis(s, npauli, ops, ps);
    const std::vector<double> r = exp_values_var(s, ps);
    for (int32_t k = 0; k < npauli; ++k) out[k] = r[k];
  })
}

int nx_state_grad_exp_values(const nx_state* s, int32_t npauli, const int32_t* ops, const double* weights,
                             double* loss, double* grad) {
  NX_TRY({
    const int64_t chi = s->chi();
    PauliSet ps;
    add_dense_paulis(s, npauli, ops, ps);
    for (int64_t j = 0; j < 2 * chi; ++j) grad[j] = 0.0;
    *loss = 0.0;
    if (chi == 0 || npauli == 0) return NX_OK;
    DevBuf dw((size_t)npauli * 8), dl(8), dg((size_t)chi * 16);
    h2d(dw.p, weights, (size_t)npauli * 8);
    grad_enqueue(s, ps, dw.as<double>(), coef_of(s).c, dl.as<double>(), dg.as<double2>());
    d2h(grad, dg.p, (size_t)chi * 16);
    d2h(loss, dl.p, 8);
    sync();
  })
}

int nx_state_grad_exp_values_device(const nx_state* s, int32_t npauli, const int32_t* ops, const double* weights,
                                    const double* d_coef, double* d_loss, double* d_grad, void* stream_) {
  NX_TRY({
    const int64_t chi = s->chi();
    PauliSet ps;
    add_dense_paulis(s, npauli, ops, ps);
    StreamOrder order(stream_);
    if (chi == 0 || npauli == 0) {
      hip_check(hipMemsetAsync(d_loss, 0, 8, stream()), "hipMemsetAsync");
      if (chi) hip_check(hipMemsetAsync(d_grad, 0, (size_t)chi * 16, stream()), "hipMemsetAsync");
    } else {
      DevBuf dw((size_t)npauli * 8);
      h2d(dw.p, weights, (size_t)npauli * 8);
      const double2* c = d_coef ? (const double2*)d_coef : coef_of(s).c;
      grad_enqueue(s, ps, dw.as<double>(), c, d_loss, (double2*)d_grad);
    }
    order.done();
  })
}

// ---------------------------------------------------------------------------
// batches

int nx_batch_create(int32_t nstates, const nx_state* const* states, const int32_t* npauli, const int32_t* ops,
                    int32_t flags, nx_batch** out) {
  NX_TRY({
    *out = nullptr;
    if (nstates < 0) fail(NX_E_VALUE, "negative state count");
    std::unique_ptr<nx_batch> b(new nx_batch());
    // Pauli lists first: every length error is reported before any device work
    std::vector<PauliSet> sets(nstates);
    size_t pos = 0;
    int maxvar = 1;
    for (int32_t i = 0; i < nstates; ++i) {
      if (npauli[i] < 0) fail(NX_E_VALUE, "negative Pauli count for state %d", i);
      add_dense_paulis(states[i], npauli[i], ops + pos, sets[i]);
      pos += (size_t)npauli[i] * states[i]->n;
      maxvar = std::max(maxvar, npauli[i]);
    }
    const int VW = maxvar <= 64 ? 1 : 2;
    const int per = 64 * VW;
    b->state_out.resize(nstates);
    b->state_np.assign(npauli, npauli + nstates);
    int64_t nout = 0;
    for (int32_t i = 0; i < nstates; ++i) {
      b->state_out[i] = nout;
      nout += 1 + npauli[i];
    }
    b->nout = nout;
    b->states.reserve(nstates);
    for (int32_t i = 0; i < nstates; ++i) b->states.push_back(clone_state(states[i]));
    std::vector<VarJob> jobs;
    size_t njobs = 0;
    for (int32_t i = 0; i < nstates; ++i)
      if (b->states[i]->chi() > 0) njobs += std::max(1, (sets[i].size() + per - 1) / per);
    b->vars.reserve(njobs);
    for (int32_t i = 0; i < nstates; ++i) {
      const nx_state* st = b->states[i];
      if (st->chi() == 0) continue;
      const QForms& q = qform_of(st);
      CoefDev c = coef_of(st);
      // the first job of a state also yields its norm (the base, i.e. the
      // identity string, comes with every variant pass)
      for (int first = 0; first == 0 || first < sets[i].size(); first += per) {
        b->vars.push_back(pauli_vars(st->forms, sets[i], first, std::min(per, sets[i].size() - first), VW));
        VarJob j;
        j.X = &q;
        j.Y = &q;
        j.V = &b->vars.back();
        j.cx = c;
        j.cy = c;
        j.out_base = b->state_out[i] + 1 + first;
        j.base_out = first == 0 ? b->state_out[i] : -1;
        jobs.push_back(j);
      }
    }
    b->VW = VW;
    if (!jobs.empty()) {
      std::pair<int, int64_t> shape = (flags & NX_BATCH_AUTOTUNE) ? autotune(jobs, VW) : tuned_shape(jobs, VW);
      b->J = shape.first;
      b->waves = shape.second;
      b->plan = make_plan(jobs, VW, b->J, b->waves);
      if (b->plan.nout < nout) b->plan.out.alloc(nout * 8);
      b->plan.nout = nout;
      sync();
      if (flags & NX_BATCH_GRAPH) {
        // captured on the context's own stream (a caller's stream set with
        // nx_set_stream may be the null stream, which cannot be captured)
        hipStream_t st = ctx()->stream;
        hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
          b->plan.enqueue(st);
        } catch (...) {
          hipGraph_t g = nullptr;
          (void)hipStreamEndCapture(st, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        hip_check(hipStreamEndCapture(st, &b->graph), "hipStreamEndCapture");
        hip_check(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
      }
    }
    *out = b.release();
  })
}

int nx_batch_num_outputs(const nx_batch* b, int64_t* out) { NX_TRY(*out = b->nout) }

int nx_batch_shape(const nx_batch* b, int32_t* J, int64_t* waves, int32_t* variant_words) {
  NX_TRY(*J = b->J; *waves = b->plan.total_waves; *variant_words = b->VW)
}

int nx_batch_run(nx_batch* b, double* out) {
  NX_TRY({
    std::vector<double> o(b->nout, 0.0);
    if (b->plan.njobs > 0) {
      ProfScope ps("batch_pairs");
      if (b->exec) hip_check(hipGraphLaunch(b->exec, stream()), "hipGraphLaunch");
      else b->plan.enqueue(stream());
      ps.end();
      d2h(o.data(), b->plan.out.p, (size_t)b->nout * 8);
      sync();
    }
    for (size_t i = 0; i < b->states.size(); ++i) {
      const nx_state* st = b->states[i];
      const double g2 = st->g.real() * st->g.real() + st->g.imag() * st->g.imag();
      const int64_t base = b->state_out[i];
      if (st->chi() == 0) {
        for (int k = 0; k <= b->state_np[i]; ++k) out[base + k] = 0.0;
        continue;
      }
      out[base] = std::sqrt(o[base] * g2);  // norm.rs:32-35
      for (int k = 1; k <= b->state_np[i]; ++k) out[base + k] = o[base + k] * g2;
    }
  })
}

void nx_batch_free(nx_batch* b) { delete b; }

int nx_profile_enable(int on) {
  NX_TRY({
    (void)ctx();
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof.collect();
    g_prof.on = on != 0;
    g_prof.acc.clear();
  })
}

int nx_profile_only(const char* kernels) {
  NX_TRY({
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof.only = (kernels && *kernels) ? "," + std::string(kernels) + "," : std::string();
  })
}

int nx_profile_get(const char* kernel, double* total_ms, int64_t* count) {
  NX_TRY({
    if (kernel && std::strcmp(kernel, "allocations") == 0) {
      *total_ms = 0.0;
      *count = g_alloc_calls.load(std::memory_order_relaxed);
      return NX_OK;
    }
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof.collect();
    auto it = g_prof.acc.find(kernel ? kernel : "");
    *total_ms = it == g_prof.acc.end() ? 0.0 : it->second.first;
    *count = it == g_prof.acc.end() ? 0 : it->second.second;
  })
}

int nx_random_clifford(int32_t n, const uint8_t* seed32, int64_t max_gates, int32_t* gates, int64_t* ngates) {
  NX_TRY({
    if (n < 0) fail(NX_E_VALUE, "negative qubit count");
    if (n > 1024) fail(NX_E_VALUE, "random Clifford circuits are limited to 1024 qubits");
    uint8_t seed[32];
    seed_or_entropy(seed32, seed);
    ChaCha12Rng rng(seed);
    const std::vector<int32_t> g = random_clifford_gates(n, rng);
    const int64_t k = (int64_t)g.size() / 4;
    if (k > max_gates) fail(NX_E_VALUE, "gate buffer too small (%lld gates)", (long long)k);
    if (k) std::memcpy(gates, g.data(), g.size() * sizeof(int32_t));
    *ngates = k;
  })
}

int nx_state_drop_cache(nx_state* s) {
  NX_TRY({
    // the quadratic forms are recomputed by the next call, into the same
    // device buffers (the coefficient copies are inputs: they stay resident
    // until the coefficients change)
    s->q_ok = false;
  })
}

}  // extern "C"
