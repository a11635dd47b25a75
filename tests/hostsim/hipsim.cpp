// Host stand-in for the HIP runtime calls and kernel launchers that the
// library's host code (fft_amd/csrc/plan.cpp) makes, for CPU-only tests of
// the planner under AddressSanitizer and UndefinedBehaviorSanitizer. Test
// infrastructure: never linked into the product.
//
// "Device" memory is host memory, streams, events and graphs are tokens, and
// the kernel launchers do no arithmetic: they check each launch against what
// the kernels assume (workgroup shape, LDS footprint, points per thread, the
// radix schedule) and count it. Failures can be injected: the n-th device
// allocation, pinned allocation or kernel launch fails.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "gofft.h"
#include "kernels.h"

namespace {
std::mutex g_mu;
long g_allocs = 0, g_fail_alloc = 0;    // fail allocation number g_fail_alloc (1-based; 0: never)
long g_nt = 0;  // launches that take a non-temporal form (hipsim_nt_launches)
long g_launches = 0, g_fail_launch = 0; // likewise for kernel launches
long g_hostallocs = 0, g_fail_host = 0;
long g_live = 0;                        // live device allocations
std::map<std::string, long> g_counts;   // launches per launcher
std::string g_bad;                      // first launch that broke a kernel assumption
std::map<uintptr_t, size_t> g_regions;  // "device" memory: allocations, the scratch symbol, registered buffers

// Whether [p, p + bytes) lies in one region.
bool in_region(const void* p, size_t bytes) {
  std::lock_guard<std::mutex> lk(g_mu);
  const uintptr_t x = (uintptr_t)p;
  auto it = g_regions.upper_bound(x);
  if (it == g_regions.begin()) return false;
  --it;
  return x >= it->first && x + bytes <= it->first + it->second;
}
int g_token = 0;

void* token() { return (void*)(uintptr_t)(0x1000 + 16 * (++g_token)); }

hipError_t launch(const char* what) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_launches++;
  g_counts[what]++;
  if (g_fail_launch && g_launches == g_fail_launch) return hipErrorLaunchFailure;
  return hipSuccess;
}

void bad(const std::string& msg) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_bad.empty()) g_bad = msg;
}

bool pow2(int x) { return x > 0 && (x & (x - 1)) == 0; }

// The shape every LDS-staged kernel assumes: thread tid owns column
// tid >> log2tpc (< C) and points lt + u * tpc, u < maxp; the column pitch
// holds R points and their pads; the stages multiply to R.
void check_sub(const char* k, const gofft::SubFFT& f, int threads, int maxp, long lds_cap, size_t elt) {
  char b[512];
  if (threads < 1 || threads > 1024 || f.C < 1 || !pow2(f.C) || !pow2(f.tpc) || (1 << f.log2C) != f.C ||
      (1 << f.log2tpc) != f.tpc || f.C * f.tpc != threads) {
    snprintf(b, sizeof b, "%s: workgroup shape C=%d tpc=%d threads=%d", k, f.C, f.tpc, threads);
    return bad(b);
  }
  const long rp = (long)f.R + (f.R >> f.psh) + 1;
  if (f.Rp != rp) {
    snprintf(b, sizeof b, "%s: pitch %d for R=%d psh=%d", k, f.Rp, f.R, f.psh);
    return bad(b);
  }
  if ((long)f.C * f.Rp * (long)elt > lds_cap) {
    snprintf(b, sizeof b, "%s: LDS %ld > %ld", k, (long)f.C * f.Rp * (long)elt, lds_cap);
    return bad(b);
  }
  if ((long)f.R > (long)maxp * f.tpc) {
    snprintf(b, sizeof b, "%s: R=%d > %d points x %d threads per column", k, f.R, maxp, f.tpc);
    return bad(b);
  }
  long prod = 1, ns = 1;
  for (int i = 0; i < f.nstages; i++) {
    const int r = f.st[i].radix > 100 ? f.st[i].radix - 100 : f.st[i].radix;
    if (r < 2 || f.st[i].ns != ns || f.st[i].nb * r != f.R) {
      snprintf(b, sizeof b, "%s: stage %d radix %d ns %d nb %d (R=%d)", k, i, r, f.st[i].ns, f.st[i].nb, f.R);
      return bad(b);
    }
    // kernels.hip lds_fft: the radices each form compiles, and every
    // butterfly j < nb reached by j = lt + m * tpc, m < max(maxp / radix, 1)
    if (f.st[i].radix < 100) {
      static const int k16[] = {2, 3, 4, 5, 6, 7, 8, 10, 11, 12, 13, 16};
      static const int k24[] = {2, 3, 4, 5, 6, 7, 8, 10, 11, 12, 13, 16};
      bool known = false;
      for (int x : k16) known |= maxp == 16 && x == r;
      for (int x : k24) known |= maxp == 24 && x == r;
      known |= maxp == 8 && (r == 2 || r == 4 || r == 8);  // lds_fft_pow2
      const int cap = maxp / r > 0 ? maxp / r : 1;
      if (!known || (f.st[i].nb + f.tpc - 1) / f.tpc > cap) {
        snprintf(b, sizeof b, "%s: stage %d radix %d (%d butterflies, %d threads per column) not run by the %d-point form",
                 k, i, r, f.st[i].nb, f.tpc, maxp);
        return bad(b);
      }
    }
    prod *= r;
    ns *= r;
  }
  if (f.R > 1 && prod != f.R) {
    snprintf(b, sizeof b, "%s: stages multiply to %ld, not R=%d", k, prod, f.R);
    return bad(b);
  }
}
}  // namespace

extern "C" {
// Test controls.
void hipsim_fail_alloc(long n) { g_fail_alloc = n; }
void hipsim_fail_host_alloc(long n) { g_fail_host = n; }
void hipsim_fail_launch(long n) { g_fail_launch = n; }
void hipsim_reset_counts() {
  g_allocs = g_launches = g_hostallocs = 0;
  g_counts.clear();
}
long hipsim_allocs() { return g_allocs; }
long hipsim_launches() { return g_launches; }
long hipsim_nt_launches() { return g_nt; }
long hipsim_live() { return g_live; }
void hipsim_region(void* p, size_t n) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_regions[(uintptr_t)p] = n;
}
void hipsim_unregion(void* p) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_regions.erase((uintptr_t)p);
}
const char* hipsim_bad() { return g_bad.c_str(); }
}

// --- the runtime ------------------------------------------------------------
hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int) {
  *v = a == hipDeviceAttributeMaxSharedMemoryPerBlock ? 160 * 1024 : a == hipDeviceAttributeMultiprocessorCount ? 256 : 0;
  return hipSuccess;
}
hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int) {
  memset(p, 0, sizeof *p);
  strcpy(p->gcnArchName, "gfx950:sim");
  return hipSuccess;
}
hipError_t hipGetSymbolAddress(void** out, const void* sym) {
  *out = const_cast<void*>(sym);
  hipsim_region(*out, (size_t)GOFFT_NOALLOC_SCRATCH_BYTES);  // the library's only device symbol
  return hipSuccess;
}
hipError_t hipMalloc(void** p, size_t n) {
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_allocs++;
    if (g_fail_alloc && g_allocs == g_fail_alloc) {
      *p = nullptr;
      return hipErrorOutOfMemory;
    }
    g_live++;
  }
  *p = aligned_alloc(256, (n + 255) / 256 * 256);
  if (*p) hipsim_region(*p, n);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p) {
  if (p) {
    hipsim_unregion(p);
    std::lock_guard<std::mutex> lk(g_mu);
    g_live--;
  }
  free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t n, unsigned int) {
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_hostallocs++;
    if (g_fail_host && g_hostallocs == g_fail_host) {
      *p = nullptr;
      return hipErrorOutOfMemory;
    }
  }
  *p = aligned_alloc(256, (n + 255) / 256 * 256);
  return hipSuccess;
}
hipError_t hipHostFree(void* p) {
  free(p);
  return hipSuccess;
}
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
  memmove(d, s, n);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) {
  memmove(d, s, n);
  return hipSuccess;
}
hipError_t hipMemGetInfo(size_t* fr, size_t* tot) {
  *fr = (size_t)16 << 30;
  *tot = (size_t)288 << 30;
  return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) {
  *s = (hipStream_t)token();
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned int) { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t* e) {
  *e = (hipEvent_t)token();
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
  *ms = 1.0f;
  return hipSuccess;
}
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipGetLastError() { return hipSuccess; }
const char* hipGetErrorString(hipError_t e) {
  return e == hipErrorOutOfMemory ? "hipErrorOutOfMemory" : e == hipErrorLaunchFailure ? "hipErrorLaunchFailure" : "hipError";
}
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipSuccess; }
hipError_t hipStreamIsCapturing(hipStream_t, hipStreamCaptureStatus* st) {
  *st = hipStreamCaptureStatusNone;
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
  *g = (hipGraph_t)token();
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t, hipGraphNode_t*, char*, size_t) {
  *e = (hipGraphExec_t)token();
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }

// --- the launchers (fft_amd/csrc/kernels.h) -------------------------------
namespace gofft {
static const long kLds = 160 * 1024;

// Non-temporal line accesses (PassArgs::nt, RealFwdArgs::nt, RealInvArgs::nt)
// only on data the launch reads or writes in whole 128-byte runs through the
// plain load and store code: no offset table, no hook, a 128-byte-aligned
// base, and not the real forms' bins.
void count_nt(int nt) {
  if (nt) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_nt++;
  }
}
void check_nt_pass(const PassArgs& a) {
  count_nt(a.nt);
  if ((a.nt & kNtLoad) && (a.in_off || a.in_hook || a.real_first || (uintptr_t)a.in % 128))
    bad("k_pass: non-temporal loads of an offset-table, hooked, real-pack or unaligned input");
  if ((a.nt & kNtStore) && (a.out_off || a.out_hook || a.real_last || (uintptr_t)a.out % 128))
    bad("k_pass: non-temporal stores to an offset-table, hooked, real-untangle or unaligned output");
}

hipError_t launch_pass(int prec, int, const PassArgs& a, int threads, hipStream_t) {
  const size_t e = prec == 0 ? 16 : 8;
  check_nt_pass(a);
  check_sub("k_pass", a.f, threads, 16, kLds - kPassStaticLds, e);
  if (a.mid_spec) check_sub("k_pass (second sub-FFT)", a.f2, threads, 16, kLds - kPassStaticLds, e);
  if (!a.real_last && !a.real_first && a.ncols <= 0) bad("k_pass: no columns");
  if ((a.in_hook == 2 || a.out_hook == 2) && !a.x0buf) bad("k_pass: Rader hooks without x0 slots");
  return launch("pass");
}
hipError_t launch_real_fwd(int prec, int maxp, const RealFwdArgs& a, int threads, hipStream_t) {
  const size_t re = prec == 0 ? 8 : 4;
  count_nt(a.nt);
  if (a.lds_tab) {
    // table form: complex64 only, 8 points per thread, radices 2/4/8, the
    // tables beside the columns, at most 512 threads, a positive grid
    if (prec != 1 || maxp != 8) bad("k_real_fwd_tab: complex64 8-point form only");
    for (int i = 0; i < a.f.nstages; i++) {
      const int r = a.f.st[i].radix;
      if (r != 2 && r != 4 && r != 8) bad("k_real_fwd_tab: radix other than 2, 4, 8");
    }
    if (threads > 512) bad("k_real_fwd_tab: more than 512 threads");
    if (a.wgs <= 0) bad("k_real_fwd_tab: empty grid");
    if (a.stw_n < 1) bad("k_real_fwd_tab: no stage twiddle entries");
    const long tab = (long)(a.stw_n + 2 * a.f.R + 1) * 8;
    check_sub("k_real_fwd_tab", a.f, threads, 8, kLds - tab, 8);
    return launch("real_fwd");
  }
  if ((a.nt & kNtLoad) && (!a.vec_in || (uintptr_t)a.in % 128 || (a.in_stride * re) % 128 ||
                           a.in_stride < 2 * (long long)a.f.R))
    bad("k_real_fwd: non-temporal loads of unaligned or overlapping rows");
  if (a.nt & kNtStore) bad("k_real_fwd: non-temporal stores are not used");
  check_sub("k_real_fwd", a.f, threads, 16, kLds, prec == 0 ? 16 : 8);
  return launch("real_fwd");
}
hipError_t launch_real_inv(int prec, int maxp, const RealInvArgs& a, int threads, hipStream_t) {
  const size_t re = prec == 0 ? 8 : 4;
  count_nt(a.nt);
  if ((a.nt & kNtStore) && ((uintptr_t)a.out % 128 || (a.out_stride * re) % 128 || a.out_stride < 2 * (long long)a.f.R))
    bad("k_real_inv: non-temporal stores to unaligned or overlapping rows");
  if (a.lds_tab) {
    if (prec != 1 || maxp != 8) bad("k_real_inv_tab: complex64 8-point form only");
    for (int i = 0; i < a.f.nstages; i++) {
      const int r = a.f.st[i].radix;
      if (r != 2 && r != 4 && r != 8) bad("k_real_inv_tab: radix other than 2, 4, 8");
    }
    if (threads > 512) bad("k_real_inv_tab: more than 512 threads");
    if (a.wgs <= 0) bad("k_real_inv_tab: empty grid");
    if (a.stw_n < 1) bad("k_real_inv_tab: no stage twiddle entries");
    const long tab = (long)(a.stw_n + a.f.R / 2 + 1) * 8;
    check_sub("k_real_inv_tab", a.f, threads, 8, kLds - tab, 8);
    return launch("real_inv");
  }
  if ((a.nt & kNtStore) && ((uintptr_t)a.out % 128 || (a.out_stride * re) % 128 || a.out_stride < 2 * (long long)a.f.R))
    bad("k_real_inv: non-temporal stores to unaligned or overlapping rows");
  if (a.nt & kNtLoad) bad("k_real_inv: non-temporal loads are not used");
  check_sub("k_real_inv", a.f, threads, 16, kLds, prec == 0 ? 16 : 8);
  return launch("real_inv");
}
hipError_t launch_conv(int prec, int maxp, const ConvArgs& a, int threads, hipStream_t) {
  check_sub("k_conv", a.f, threads, maxp, kLds, prec == 0 ? 16 : 8);
  if (maxp != 16 && !(maxp == 24 && prec == 1)) bad("k_conv: unsupported points per thread");
  if (a.nlines <= 0) bad("k_conv: no lines");
  if (a.kind != 2 && (!a.perm || !a.oinv)) bad("k_conv: Rader without its gather or output-order table");
  return launch("conv");
}
// Mirror of kernels.hip pass_form (the form k_pass would run).
int pass_form(int prec, const PassArgs& a) {
  if (a.real_last || a.real_first) return a.real_last ? 11 : 12;
  const bool ext = a.in_hook || a.out_hook || a.mid_spec;
  int mode = a.mid256 && prec != 0 ? 2 : ext ? 1 : 0;
  const SubFFT& f = a.f;
  const int A = f.st[0].radix;
  if (mode <= 1 && f.nstages == 2 && f.tpc == 16 && f.st[1].radix == 16 && f.R == 16 * A && f.st[0].cmul == 0 &&
      f.st[1].cmul == 0 && !a.mid_spec) {
    const bool hooks = a.in_hook || a.out_hook;
    const bool rader = a.in_hook == 2 || a.out_hook == 2;
    if (hooks && prec != 0 && !rader) {
      if (A == 10) mode = 6;
      if (A == 12) mode = 7;
      if (A == 16) mode = 5;
    } else if (!hooks || prec == 0) {
      if (A == 10) mode = 3;
      if (A == 12) mode = 4;
    }
  }
  return mode;
}
unsigned queue_pass_forms(int prec) { return prec == 0 ? 0x1bu : 0xffu; }
bool queue_conv_form(int prec, int maxp) { return maxp == 16 || (prec != 0 && maxp == 24); }
size_t queue_pass_tables(int prec) { return (size_t)kMaxCols * (7 * 8 + (prec == 0 ? 16 : 8)); }
size_t queue_static_lds() { return 64; }
// LDS and waves per CU (2048 threads); complex128's queue kernel holds two
// 256-thread or one 512-thread workgroup per CU by its registers.
int queue_blocks_per_cu(int prec, int threads, size_t lds) {
  const long by_lds = kLds / (long)(lds + queue_static_lds());
  long by_regs = 2048 / threads;
  if (prec == 0) by_regs = threads == 256 ? 2 : 1;
  return (int)(by_lds < by_regs ? by_lds : by_regs);
}
// The setup kernel's work, done on the host ("device" memory is host memory).
hipError_t launch_queue_setup(const QueueSetupArgs& a, int ndesc, hipStream_t) {
  if (ndesc < 1 || !a.tmpl || !a.descs || !a.d || !a.ctr || a.nctr < 4) bad("k_queue_setup: arguments");
  for (int i = 0; i < ndesc; i++) {
    const QueueDesc& d = a.d[i];
    if (d.off % 8 || d.bytes % 8 || d.nreloc < 0 || d.nreloc > 4) bad("k_queue_setup: descriptor layout");
    memcpy(a.descs + d.off, a.tmpl + d.off, (size_t)d.bytes);
    for (int k = 0; k < d.nreloc; k++) {
      if (d.sel[k] < 1 || d.sel[k] > 3 || d.field[k] < 0 || d.field[k] + 8 > d.bytes) bad("k_queue_setup: relocation");
      const char* v = a.base[d.sel[k]] + d.add[k];
      memcpy(a.descs + d.off + d.field[k], &v, sizeof v);
    }
  }
  for (int i = 0; i < a.nctr; i++) a.ctr[i] = 0;
  return launch("queue_setup");
}
// A pointer the queue resolved: set, and no stand-in left (those sit at or
// above 1 << 60).
static bool resolved(const void* p) { return p && (uintptr_t)p < ((uintptr_t)1 << 56); }
hipError_t launch_queue(int prec, const QueueArgs& a, int threads, size_t lds, long long grid, hipStream_t) {
  char b[256];
  const size_t e = prec == 0 ? 16 : 8;
  if (a.nstages < 1 || a.nstages > kMaxQueueStages || grid < 1 || !a.ticket || !a.error || !a.done ||
      a.need_extra != 0 || a.spin_limit <= 0)
    bad("k_queue: arguments");
  if (queue_blocks_per_cu(prec, threads, lds) < 1) bad("k_queue: workgroup does not fit a CU");
  if (!in_region(a.ticket, 8) || !in_region(a.error, 4) || !in_region(a.done, 4 * (size_t)a.nstages))
    bad("k_queue: counters outside device memory");
  long long start = 0;
  for (int i = 0; i < a.nstages; i++) {
    const QueueStage& st = a.st[i];
    snprintf(b, sizeof b, "k_queue stage %d", i);
    if (st.start != start || st.nitems < 1 || st.dep >= i || st.dep < -1 || (st.dep >= 0 && !a.st[st.dep].signal)) {
      bad(std::string(b) + ": table (start, items or dependency)");
      return launch("queue");
    }
    start += st.nitems;
    if (st.kind == 0) {
      const PassArgs& p = *reinterpret_cast<const PassArgs*>(a.descs + st.desc);
      // the columns end before the per-column tables, which end within the LDS
      check_sub(b, p.f, threads, 16, (long)a.sh_off, e);
      if (p.mid_spec) check_sub(b, p.f2, threads, 16, (long)a.sh_off, e);
      if (a.sh_off % 256 || a.sh_off + (long long)queue_pass_tables(prec) > (long long)lds) bad(std::string(b) + ": tables");
      if (pass_form(prec, p) != st.form || !((queue_pass_forms(prec) >> st.form) & 1u)) bad(std::string(b) + ": form");
      if ((p.ncols + p.f.C - 1) / p.f.C != st.nitems) bad(std::string(b) + ": items");
      if (!resolved(p.in) || !resolved(p.out)) bad(std::string(b) + ": data pointers");
      if ((p.in_hook == 2 || p.out_hook == 2) && !resolved(p.x0buf)) bad(std::string(b) + ": x0 slots");
      if (p.dc && !resolved(p.dc)) bad(std::string(b) + ": DC slots");
      // every pointer inside device memory; lines without an offset table
      // (workspace buffers) whole: ncols columns of R points
      const size_t span = (size_t)p.ncols * (size_t)p.f.R * e;
      if (!in_region(p.in, p.in_off || p.in_hook ? 1 : span)) bad(std::string(b) + ": input outside device memory");
      if (!in_region(p.out, p.out_off || p.out_hook ? 1 : span)) bad(std::string(b) + ": output outside device memory");
      if (p.x0buf && !in_region(p.x0buf, e)) bad(std::string(b) + ": x0 slots outside device memory");
      if (p.dc && !in_region(p.dc, e)) bad(std::string(b) + ": DC slots outside device memory");
      if (p.xcd_group != 1) bad(std::string(b) + ": XCD grouping in the queue");
    } else {
      const ConvArgs& c = *reinterpret_cast<const ConvArgs*>(a.descs + st.desc);
      check_sub(b, c.f, threads, st.form, (long)lds, e);
      if (!queue_conv_form(prec, st.form)) bad(std::string(b) + ": points per thread");
      if ((c.nlines + c.f.C - 1) / c.f.C != st.nitems) bad(std::string(b) + ": items");
      if (!resolved(c.in) || !resolved(c.out)) bad(std::string(b) + ": data pointers");
      if (!in_region(c.in, 1) || !in_region(c.out, 1)) bad(std::string(b) + ": data outside device memory");
    }
  }
  if (start != a.nitems) bad("k_queue: item count");
  return launch("queue");
}
int real_kernel_vgprs(int prec, int dir) { return prec == 0 ? 126 : dir == 1 ? 78 : 88; }
int real_tab_vgprs(int) { return 89; }
long long residual_parts(long long per) {
  long long p = (per + 2047) / 2048;
  return p < 1 ? 1 : p > 1024 ? 1024 : p;
}
#define SIM_EW(NAME, ...) \
  hipError_t NAME(__VA_ARGS__) { return launch(#NAME); }
SIM_EW(launch_cmul_table, int, const void*, void*, const void*, long long, long long, long long, long long, int, int, double,
       const long long*, const long long*, hipStream_t)
SIM_EW(launch_rader_gather, int, const void*, void*, const int*, long long, long long, long long, long long, long long, int,
       const long long*, hipStream_t)
SIM_EW(launch_rader_scatter, int, const void*, const void*, void*, const int*, long long, long long, long long, long long,
       long long, long long, int, int, double, const long long*, const long long*, hipStream_t)
SIM_EW(launch_chirp_in, int, const void*, void*, const void*, long long, long long, long long, long long, long long, int,
       const long long*, hipStream_t)
SIM_EW(launch_untangle, int, const void*, void*, const void*, long long, long long, long long, long long, int, double, double,
       double, hipStream_t)
SIM_EW(launch_herm_pack, int, const void*, void*, const void*, long long, long long, long long, long long, long long, double,
       hipStream_t)
SIM_EW(launch_weight_bins, int, const void*, void*, long long, long long, long long, long long, long long, double, double,
       hipStream_t)
SIM_EW(launch_promote, int, const void*, void*, long long, long long, long long, long long, hipStream_t)
SIM_EW(launch_copy_lines, int, const void*, void*, long long, long long, long long, long long, hipStream_t)
SIM_EW(launch_herm_full, int, const void*, void*, long long, long long, long long, long long, long long, double, hipStream_t)
SIM_EW(launch_take_real, int, const void*, void*, long long, long long, long long, long long, double, hipStream_t)
SIM_EW(launch_power, int, const void*, void*, long long, long long, long long, long long, long long, double, hipStream_t)
SIM_EW(launch_frames, int, const void*, const void*, void*, long long, long long, long long, hipStream_t)
SIM_EW(launch_axis_copy, int, const void*, void*, long long, long long, long long, int, hipStream_t)
SIM_EW(launch_fill, int, void*, long long, double, hipStream_t)
SIM_EW(launch_scale, int, const void*, void*, long long, double, hipStream_t)
SIM_EW(launch_filter_mul, int, void*, const void*, long long, long long, int, hipStream_t)
SIM_EW(launch_residual_loss, int, void*, const void*, long long, long long, void*, void*, hipStream_t)
// the radix-16 stage probe (kernels.hip k_probe_dft16_*): one launch over the
// given lines; gofft_probe_dft16 checks its arguments before it gets here
hipError_t launch_probe_dft16(int method, const void* in, void* out, long long transforms, int reps,
                              const float* wmat, hipStream_t) {
  if ((method != 0 && method != 1) || !in || !out || transforms <= 0 || reps < 0 || !wmat)
    bad("launch_probe_dft16: arguments");
  return launch("launch_probe_dft16");
}
}  // namespace gofft
