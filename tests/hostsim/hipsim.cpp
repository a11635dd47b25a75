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

#include "kernels.h"

namespace {
std::mutex g_mu;
long g_allocs = 0, g_fail_alloc = 0;    // fail allocation number g_fail_alloc (1-based; 0: never)
long g_launches = 0, g_fail_launch = 0; // likewise for kernel launches
long g_hostallocs = 0, g_fail_host = 0;
long g_live = 0;                        // live device allocations
std::map<std::string, long> g_counts;   // launches per launcher
std::string g_bad;                      // first launch that broke a kernel assumption
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
long hipsim_live() { return g_live; }
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
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p) {
  if (p) {
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

hipError_t launch_pass(int prec, int, const PassArgs& a, int threads, hipStream_t) {
  const size_t e = prec == 0 ? 16 : 8;
  check_sub("k_pass", a.f, threads, 16, kLds - kPassStaticLds, e);
  if (a.mid_spec) check_sub("k_pass (second sub-FFT)", a.f2, threads, 16, kLds - kPassStaticLds, e);
  if (!a.real_last && !a.real_first && a.ncols <= 0) bad("k_pass: no columns");
  if ((a.in_hook == 2 || a.out_hook == 2) && !a.x0buf) bad("k_pass: Rader hooks without x0 slots");
  return launch("pass");
}
hipError_t launch_real_fwd(int prec, int, const RealFwdArgs& a, int threads, hipStream_t) {
  check_sub("k_real_fwd", a.f, threads, 16, kLds, prec == 0 ? 16 : 8);
  return launch("real_fwd");
}
hipError_t launch_real_inv(int prec, int, const RealInvArgs& a, int threads, hipStream_t) {
  check_sub("k_real_inv", a.f, threads, 16, kLds, prec == 0 ? 16 : 8);
  return launch("real_inv");
}
hipError_t launch_conv(int prec, int maxp, const ConvArgs& a, int threads, hipStream_t) {
  check_sub("k_conv", a.f, threads, maxp, kLds, prec == 0 ? 16 : 8);
  if (maxp != 16 && !(maxp == 24 && prec == 1)) bad("k_conv: unsupported points per thread");
  if (a.nlines <= 0) bad("k_conv: no lines");
  return launch("conv");
}
int real_kernel_vgprs(int prec) { return prec == 0 ? 126 : 84; }
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
}  // namespace gofft
