// Host side of libgofft_hip.so: plan construction (routing, radix schedules,
// twiddle/chirp/kernel tables), per-stream workspace, execution, and the C ABI
// declared in include/gofft.h.
//
// Routing follows the reference's NewPlan (plan.go:42-73): n <= 1 trivial,
// 13-smooth lengths by Stockham FFT kernels, primes >= 700 by Rader's
// algorithm, every other length by Bluestein's chirp-z. All tables are generated
// here in long double from exact integer phases and rounded once to the working
// precision.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstddef>
#include <cstdlib>
#include <cstdio>
#include <functional>
#include <limits>
#include <tuple>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "gofft.h"
#include "kernels.h"

using namespace gofft;
typedef long long i64;
typedef std::complex<long double> cld;

namespace {

// ---------------------------------------------------------------------------
// Errors.

thread_local std::string t_err;

int fail(int code, const std::string& msg) {
  t_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(GOFFT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
  } while (0)

#define RC_TRY(expr)            \
  do {                          \
    int rc_ = (expr);           \
    if (rc_ != GOFFT_OK) return rc_; \
  } while (0)

// ---------------------------------------------------------------------------
// Device limits.

// Library version (gofft_version) and the short tag that the autotuner's
// cache file records with each entry.
const char* const kVersionText = "gofft-hip 0.2 (gfx950)";
const char* const kVersionTag = "0.2";

// Scratch of the no-allocation entry point (gofft_exec_c2c_noalloc): a
// file-scope device array, one copy per device, held until the process
// ends. Host code never uses its name as an address: current_device() reads
// each device's copy's address once, with that device current.
constexpr size_t kScratchBytes = (size_t)GOFFT_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct DevInfo {
  int id = -1;
  unsigned char* scratch = nullptr;  // this device's copy of `scratch`
  int lds_bytes = 65536;  // dynamic LDS per workgroup beside k_pass's static tables
  int lds_full = 65536;   // dynamic LDS per workgroup of the kernels without static LDS
  int cus = 0;
  std::string arch;       // e.g. "gfx950:sramecc+:xnack-"
};
std::mutex g_dev_mu;
std::map<int, DevInfo> g_devs;

int current_device(DevInfo* out) {
  int id = 0;
  HIP_TRY(hipGetDevice(&id));
  std::lock_guard<std::mutex> lk(g_dev_mu);
  auto it = g_devs.find(id);
  if (it == g_devs.end()) {
    DevInfo d;
    d.id = id;
    int lds = 0, cus = 0;
    HIP_TRY(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, id));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, id));
    // k_pass keeps per-column tables in static LDS; k_conv, k_real_fwd and
    // k_real_inv have none.
    d.lds_full = lds;
    d.lds_bytes = lds - kPassStaticLds;
    if (const char* e = std::getenv("GOFFT_LDS_BYTES")) {
      int v = std::atoi(e);
      if (v > 0 && v < d.lds_bytes) d.lds_bytes = v;
      if (v > 0 && v < d.lds_full) d.lds_full = v;
    }
    d.cus = cus;
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, id));
    d.arch = prop.gcnArchName;
    for (char& ch : d.arch)
      if (ch == ' ') ch = '_';
    if (d.arch.empty()) d.arch = "unknown";
    HIP_TRY(hipGetSymbolAddress((void**)&d.scratch, HIP_SYMBOL(scratch)));
    it = g_devs.emplace(id, d).first;
  }
  *out = it->second;
  return GOFFT_OK;
}

// ---------------------------------------------------------------------------
// Integer helpers (restated from the reference's plan.go / rader.go; the
// product keeps its own copy so it never depends on the test oracle).

bool factors_are_small(i64 n) {  // plan.go:134-141
  for (int p : {2, 3, 5, 7, 11, 13})
    while (n % p == 0) n /= p;
  return n == 1;
}

bool is_prime(i64 n) {  // rader.go:232-245
  if (n < 2) return false;
  if (n % 2 == 0) return n == 2;
  for (i64 d = 3; d * d <= n; d += 2)
    if (n % d == 0) return false;
  return true;
}

i64 mod_pow(i64 b, i64 e, i64 m) {  // rader.go:217-228
  i64 r = 1;
  b %= m;
  while (e > 0) {
    if (e & 1) r = (i64)((__int128)r * b % m);
    e >>= 1;
    b = (i64)((__int128)b * b % m);
  }
  return r;
}

i64 primitive_root(i64 p) {  // rader.go:178-214
  i64 phi = p - 1, m = phi;
  std::vector<i64> fs;
  for (i64 d = 2; d * d <= m; d++)
    if (m % d == 0) {
      fs.push_back(d);
      while (m % d == 0) m /= d;
    }
  if (m > 1) fs.push_back(m);
  if (fs.empty()) return 1;
  for (i64 g = 2;; g++) {
    bool gen = true;
    for (i64 f : fs)
      if (mod_pow(g, phi / f, p) == 1) {
        gen = false;
        break;
      }
    if (gen) return g;
  }
}

const i64 kRaderThreshold = 700;  // plan.go:83

int route(i64 n) {  // plan.go:42-73
  if (n <= 1) return 0;
  if ((n & (n - 1)) == 0) return 1;
  if (factors_are_small(n)) return 2;
  if (n >= kRaderThreshold && is_prime(n)) return 3;
  return 4;
}

// ---------------------------------------------------------------------------
// Twiddles: W_M^p = exp(-2 pi i p / M), exact integer phase reduced mod M,
// evaluated in long double.

cld root(i64 p, i64 M) {
  p %= M;
  if (p < 0) p += M;
  if (p == 0) return cld(1, 0);
  const long double twopi = 6.283185307179586476925286766559005768L;
  long double a = twopi * (long double)p / (long double)M;
  return cld(cosl(a), -sinl(a));
}

// Host long-double FFT (forward, unnormalised) used only at plan time to
// transform convolution kernels. Recursive mixed radix with a direct DFT for
// prime factors.
void host_fft(std::vector<cld>& x) {
  const i64 n = (i64)x.size();
  if (n <= 1) return;
  i64 p = 0;
  for (i64 d : {4, 2, 3, 5, 7, 11, 13})
    if (n % d == 0) {
      p = d;
      break;
    }
  if (p == 0) {
    for (i64 d = 17; d * d <= n; d += 2)
      if (n % d == 0) {
        p = d;
        break;
      }
    if (p == 0) p = n;
  }
  const i64 m = n / p;
  if (m == 1) {
    std::vector<cld> y(n);
    for (i64 k = 0; k < n; k++) {
      cld s = 0;
      for (i64 j = 0; j < n; j++) s += x[j] * root((j * k) % n, n);
      y[k] = s;
    }
    x.swap(y);
    return;
  }
  std::vector<std::vector<cld>> sub(p, std::vector<cld>(m));
  for (i64 r = 0; r < p; r++)
    for (i64 j = 0; j < m; j++) sub[r][j] = x[j * p + r];
  for (i64 r = 0; r < p; r++) host_fft(sub[r]);
  std::vector<cld> wp(p);
  for (i64 q = 0; q < p; q++) wp[q] = root(q, p);
  for (i64 k1 = 0; k1 < m; k1++) {
    std::vector<cld> t(p);
    for (i64 r = 0; r < p; r++) t[r] = sub[r][k1] * root(r * k1, n);
    for (i64 k2 = 0; k2 < p; k2++) {
      cld s = 0;
      for (i64 r = 0; r < p; r++) s += t[r] * wp[(r * k2) % p];
      x[k1 + m * k2] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// Device buffers.

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() {}
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t b) {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (b == 0) return GOFFT_OK;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      return fail(GOFFT_ERR_OOM, "hipMalloc of " + std::to_string(b) + " bytes failed");
    }
    bytes = b;
    return GOFFT_OK;
  }
  int upload(const void* h, size_t b) {
    RC_TRY(alloc(b));
    if (b) HIP_TRY(hipMemcpy(p, h, b, hipMemcpyHostToDevice));
    return GOFFT_OK;
  }
};

size_t elt(int prec) { return prec == 0 ? 16 : 8; }  // complex element bytes
size_t relt(int prec) { return prec == 0 ? 8 : 4; }  // real element bytes

// Upload complex long-double values in the working precision.
int upload_complex(DevBuf& b, const std::vector<cld>& v, int prec) {
  if (prec == 0) {
    std::vector<double> h(2 * v.size());
    for (size_t i = 0; i < v.size(); i++) {
      h[2 * i] = (double)v[i].real();
      h[2 * i + 1] = (double)v[i].imag();
    }
    return b.upload(h.data(), h.size() * sizeof(double));
  }
  std::vector<float> h(2 * v.size());
  for (size_t i = 0; i < v.size(); i++) {
    h[2 * i] = (float)v[i].real();
    h[2 * i + 1] = (float)v[i].imag();
  }
  return b.upload(h.data(), h.size() * sizeof(float));
}

int upload_real(DevBuf& b, const std::vector<long double>& v, int prec) {
  if (prec == 0) {
    std::vector<double> h(v.begin(), v.end());
    return b.upload(h.data(), h.size() * sizeof(double));
  }
  std::vector<float> h(v.size());
  for (size_t i = 0; i < v.size(); i++) h[i] = (float)v[i];
  return b.upload(h.data(), h.size() * sizeof(float));
}

// ---------------------------------------------------------------------------
// Sub-FFT schedules.

int ilog2c(int x) {
  int l = 0;
  while ((1 << l) < x) l++;
  return l;
}

std::vector<int> radix_schedule(int R, int maxrad = 16) {
  std::vector<int> r;
  int x = R;
  while (maxrad >= 16 && x % 16 == 0) {
    r.push_back(16);
    x /= 16;
  }
  while (x % 8 == 0) {
    r.push_back(8);
    x /= 8;
  }
  if (x % 4 == 0) {
    r.push_back(4);
    x /= 4;
  }
  if (x % 2 == 0) {
    r.push_back(2);
    x /= 2;
  }
  for (int p : {3, 5, 7, 11, 13})
    while (x % p == 0) {
      r.push_back(p);
      x /= p;
    }
  if (x != 1) r.clear();
  return r;
}

constexpr int kMaxP = 16;   // points per thread per stage (kernel template)
// Radix-16 stages first, then the remaining factors in as few stages as the
// composite prime-factor radices 6, 10 and 12 allow (kernels.hip dft_pfa);
// empty when R has a prime factor above 13.
std::vector<int> composite_schedule(int R) {
  std::vector<int> r;
  int x = R;
  while (x % 16 == 0) {
    r.push_back(16);
    x /= 16;
  }
  int e = 0;
  while (x % 2 == 0) {
    e++;
    x /= 2;
  }
  std::vector<int> odd;
  for (int p : {3, 5, 7, 11, 13})
    while (x % p == 0) {
      odd.push_back(p);
      x /= p;
    }
  if (x != 1) return {};
  // Exhaustive search over the few remaining factors: e twos (e < 4) and
  // the odd primes; radices 2, 4, 8, the odd primes, 6 = 2*3, 10 = 2*5,
  // 12 = 4*3.
  std::vector<int> best;
  std::function<void(int, std::vector<int>, std::vector<int>)> go = [&](int e2, std::vector<int> rest,
                                                                        std::vector<int> cur) {
    if (!best.empty() && cur.size() >= best.size()) return;
    if (e2 == 0 && rest.empty()) {
      best = cur;
      return;
    }
    if (!rest.empty()) {
      const int p = rest.back();
      rest.pop_back();
      auto with = [&](int rad, int use2) {
        std::vector<int> c2 = cur;
        c2.push_back(rad);
        go(e2 - use2, rest, c2);
      };
      with(p, 0);
      if (p == 3 && e2 >= 1) with(6, 1);
      if (p == 5 && e2 >= 1) with(10, 1);
      if (p == 3 && e2 >= 2) with(12, 2);
      return;
    }
    for (int k = std::min(e2, 3); k >= 1; k--) {
      std::vector<int> c2 = cur;
      c2.push_back(1 << k);
      go(e2 - k, rest, c2);
    }
  };
  go(e, odd, {});
  std::sort(best.begin(), best.end(), [](int a, int b) { return a > b; });
  r.insert(r.end(), best.begin(), best.end());
  return r;
}

constexpr int kWideP = 24;  // the complex64 fused convolution's wide form
constexpr int kNoPad = 16;  // pad shift of an unpadded column (every point < 2^16)

// Schedule of the wide form (one column, tpc threads, at most 24 points per
// thread per stage, so a radix-r stage's R/r butterflies need
// ceil(R/r/tpc) <= 24/r): the fewest stages over radices 2-8, 10, 12, 11
// and 13, ties broken by an estimate of the work per point (the butterfly's
// and twiddle products' flops, plus 16 instructions of addressing per
// butterfly); odd radices first (kernels.hip lpm: their writes land at odd
// strides, which spread over the LDS banks), then the others from the
// largest. Empty when R has a prime factor above 13. Measured on Rader
// 10007 x 583 complex64 (20160 points): 7,5,3,3,8,8 0.143-0.147 ms;
// 7,12,12,10,2 0.141-0.143; 7,5,12,12,4 0.140-0.142; 7,12,10,6,4 (the
// choice) 0.135-0.138.
std::vector<int> wide_schedule(int R, int tpc) {
  static const int kRad[] = {12, 10, 8, 7, 6, 5, 4, 3, 2, 11, 13};
  auto cost = [](int r) {
    double flops;
    switch (r) {
      case 2: flops = 5.0; break;
      case 3: flops = 9.3; break;
      case 4: flops = 8.5; break;
      case 5: flops = 15.0; break;
      case 6: flops = 12.3; break;
      case 7: flops = 20.0; break;
      case 8: flops = 12.0; break;
      case 10: flops = 17.4; break;
      case 12: flops = 14.8; break;
      default: flops = 30.0; break;
    }
    return flops + 16.0 / r;
  };
  auto ok = [&](int r) {
    const int nb = R / r;
    return r <= kWideP && (nb + tpc - 1) / tpc <= kWideP / r;
  };
  std::vector<int> best, cur;
  double best_cost = 0;
  std::function<void(int, double)> go = [&](int x, double c) {
    if (!best.empty() && cur.size() > best.size()) return;
    if (x == 1) {
      if (best.empty() || cur.size() < best.size() || c < best_cost) {
        best = cur;
        best_cost = c;
      }
      return;
    }
    // radices in non-increasing order of kRad's list position, so each
    // multiset is visited once
    const int last = cur.empty() ? 0 : (int)(std::find(std::begin(kRad), std::end(kRad), cur.back()) - std::begin(kRad));
    for (int i = last; i < (int)(sizeof(kRad) / sizeof(kRad[0])); i++) {
      const int r = kRad[i];
      if (x % r || !ok(r)) continue;
      cur.push_back(r);
      go(x / r, c + cost(r));
      cur.pop_back();
    }
  };
  go(R, 0.0);
  std::stable_sort(best.begin(), best.end(), [](int a, int b) {
    const bool oa = a % 2, ob = b % 2;
    return oa != ob ? oa : a > b;
  });
  return best;
}

// LDS column pitch for R points (kernels.hip: one pad slot per 16 points plus
// one, odd so adjacent columns fall in different banks).
i64 pitch(i64 R, int sh = 4) { return R + (R >> sh) + 1; }

// Preferred LDS footprint (points) of a batched single-pass workgroup: about
// 64 KB, so that two or more workgroups share a CU's 160 KB.
i64 soft_pts(int prec) { return prec == 0 ? 4352 : 8704; }

struct SubPlan {
  SubFFT f{};
  int threads = 0;
  int maxp = kMaxP;     // points per thread per stage (kernel template)
  std::vector<cld> tw;  // stage twiddles
};

std::string describe_sub(i64 R, const SubPlan& sp) {
  std::string s = std::to_string(R) + "[C" + std::to_string(sp.f.C) + " T" + std::to_string(sp.threads) +
                  (sp.maxp != kMaxP ? " P" + std::to_string(sp.maxp) : "") +
                  (sp.f.psh == kNoPad ? " unpadded" : sp.f.psh != 4 ? " pad" + std::to_string(1 << sp.f.psh) : "") + ":";
  for (int i = 0; i < sp.f.nstages; i++) {
    const int r = sp.f.st[i].radix;
    s += (i ? "," : "") + (r > 100 ? "r2^" + std::to_string(ilog2c(r - 100)) : std::to_string(r));
  }
  return s + "]";
}

int ilog2(int x) {
  int l = 0;
  while ((1 << l) < x) l++;
  return l;
}

// Builds the LDS schedule for C columns of R points; false if impossible.
bool make_sub(int R, int C, SubPlan& sp, int min_threads = 0, int sh = 4, int maxp = kMaxP, int maxrad = 16) {
  if (R < 1 || C < 1 || C > kMaxCols || (C & (C - 1))) return false;
  std::vector<int> rads;
  if (R > 1) {
    rads = radix_schedule(R, maxrad);
    if (rads.empty()) return false;
  }
  if ((int)rads.size() > kMaxStages) return false;
  // The wide form (one whole-LDS convolution) runs its small odd radices
  // first: their writes land at odd strides, which spread over the LDS banks.
  // Measured: Rader 10007 x 585 complex64 0.207 -> 0.203 ms, mixed-length
  // batch 0.772 -> 0.767 ms.
  if (maxp == kWideP) std::reverse(rads.begin(), rads.end());
  // every thread holds at most maxp points per stage (and per load round of
  // the pass kernels, which have one round)
  auto feasible = [&](const std::vector<int>& rs, int T) {
    const int tpc = T / C;
    if (maxp == kMaxP && (i64)T * maxp < (i64)C * R) return false;
    for (int rad : rs) {
      const int nb = R / rad;
      const int per = (nb + tpc - 1) / tpc;
      const int cap = maxp / rad > 0 ? maxp / rad : 1;
      if (rad > maxp || per > cap) return false;
    }
    return true;
  };
  int T = 64;
  while (T < C || T < min_threads) T <<= 1;
  while ((i64)T * maxp < (i64)C * R && T < 1024) T <<= 1;
  bool ok = false;
  for (; T <= 1024; T <<= 1)
    if ((ok = feasible(rads, T))) break;
  if (!ok) return false;
  // Fewer stages with composite radices when they fit the same threads.
  if (maxp == kMaxP && maxrad >= 16 && R > 1) {
    std::vector<int> comp = composite_schedule(R);
    if (!comp.empty() && comp.size() < rads.size() && feasible(comp, T)) rads = comp;
  }
  if (maxp == kWideP && R > 1) {
    std::vector<int> comp = wide_schedule(R, T / C);
    if (!comp.empty() && comp.size() < rads.size() && feasible(comp, T)) rads = comp;
  }
  // A x 16 points with 16 threads per column: radix A first, the order the
  // pass kernel's two-stage register form runs (kernels.hip reg2).
  if (maxp == kMaxP && T / C == 16 && rads.size() == 2 && rads[0] == 16 && rads[1] != 16 && R == 16 * rads[1] &&
      feasible({rads[1], 16}, T))
    rads = {rads[1], 16};
  sp = SubPlan();
  sp.threads = T;
  sp.maxp = maxp;
  SubFFT& f = sp.f;
  f.R = R;
  f.Rp = (int)pitch(R, sh);
  f.psh = sh;
  f.C = C;
  f.log2C = ilog2(C);
  f.tpc = T / C;
  f.log2tpc = ilog2(f.tpc);
  f.nstages = (int)rads.size();
  f.Rd = fastdiv_make((uint32_t)R);
  int ns = 1, off = 0;
  for (size_t i = 0; i < rads.size(); i++) {
    int rad = rads[i];
    Stage& s = f.st[i];
    s.radix = rad;
    s.ns = ns;
    s.nb = R / rad;
    s.tw_off = off;
    s.kmul = rad - 1;
    s.cmul = 0;
    s.nsd = fastdiv_make((uint32_t)ns);
    if (ns > 1) {
      for (int k = 0; k < ns; k++)
        for (int j = 1; j < rad; j++) sp.tw.push_back(root((i64)j * k, (i64)ns * rad));
      off += ns * (rad - 1);
    }
    ns *= rad;
  }
  return true;
}

// Schedule for one pass of the reference-twiddle transform: the pass's
// radix-2 layers grouped four at a time into register stages (radix codes
// 100 + 2^t, kernels.hip lds_stage_rec). A stage starting at local size ns
// reads layer tables of the one table holding every layer L = 2, 4, ..., n
// (layer L at offset L/2 - 1), indexed with the column digit c and the
// pass's global Ns.
bool make_sub_r2(int R, int C, i64 Ns, SubPlan& sp) {
  if (R < 2 || (R & (R - 1)) || C < 1 || C > kMaxCols || (C & (C - 1))) return false;
  const int q = ilog2(R);
  std::vector<int> ts;
  for (int left = q; left > 0;) {
    int t = std::min(left, 4);
    ts.push_back(t);
    left -= t;
  }
  if ((int)ts.size() > kMaxStages) return false;
  int T = 64;
  while (T < C) T <<= 1;
  while ((i64)T * kMaxP < (i64)C * R && T < 1024) T <<= 1;
  if ((i64)T * kMaxP < (i64)C * R) return false;
  const int tpc = T / C;
  for (int t : ts) {
    int nb = R >> t;
    int cap = kMaxP >> t;
    if ((nb + tpc - 1) / tpc > cap) return false;
  }
  sp = SubPlan();
  sp.threads = T;
  SubFFT& f = sp.f;
  f.R = R;
  f.Rp = (int)pitch(R);
  f.psh = 4;
  f.C = C;
  f.log2C = ilog2(C);
  f.tpc = tpc;
  f.log2tpc = ilog2(f.tpc);
  f.nstages = (int)ts.size();
  f.Rd = fastdiv_make((uint32_t)R);
  // Twiddles of layer u of a stage (local size ns, layer length
  // L = Ns*ns*2^(u+1)): entry i*Ns + cdig of the reference's table for L,
  // read at tw_off + Ns*ns*(2^u - 1) + cdig*ns*2^u + i (kernels.hip
  // lds_stage_rec). With Ns = 1 that is the shared table of every layer
  // (layer L at L/2 - 1); a pass with Ns > 1 reads its own copy laid out so
  // (rec_pass_table), adjacent butterflies reading adjacent entries.
  int ns = 1;
  i64 base = 0;
  for (size_t i = 0; i < ts.size(); i++) {
    Stage& g = f.st[i];
    g.radix = 100 + (1 << ts[i]);
    g.ns = ns;
    g.nb = R >> ts[i];
    g.tw_off = (int)(Ns == 1 ? ns - 1 : base);
    g.kmul = (int)Ns;
    g.cmul = 1;
    g.nsd = fastdiv_make((uint32_t)ns);
    base += Ns * ns * (((i64)1 << ts[i]) - 1);
    ns <<= ts[i];
  }
  return true;
}


// Go's math.Sin / math.Cos (Go standard library math/sin.go, the Cephes
// sin.c algorithm of S. L. Moshier) for |x| < 2^29. The reference's Bluestein
// twiddle recurrence (internal/kernels/kernels.go:44-55) amplifies any
// one-ulp difference in its starting pair, so reproducing the reference's
// rounding needs Go's values, not the C library's.
#pragma clang fp contract(off)
void go_sincos(double x, double* s_out, double* c_out) {
  static const double S[6] = {1.58962301576546568060e-10, -2.50507477628578072866e-8,
                              2.75573136213857245213e-6,  -1.98412698295895385996e-4,
                              8.33333333332211858878e-3,  -1.66666666666666307295e-1};
  static const double Cc[6] = {-1.13585365213876817300e-11, 2.08757008419747316778e-9,
                               -2.75573141792967388112e-7,  2.48015872888517045348e-5,
                               -1.38888888888730564116e-3,  4.16666666666665929218e-2};
  const double PI4A = 7.85398125648498535156e-1, PI4B = 3.77489470793079817668e-8,
               PI4C = 2.69515142907905952645e-15;
  const double fourOverPi = 1.27323954473516268615107010698011489627567716592365;
  bool ss = false, cs = false;
  if (x < 0) {
    x = -x;
    ss = true;
  }
  uint64_t j = (uint64_t)(x * fourOverPi);
  double y = (double)j;
  if (j & 1) {
    j++;
    y++;
  }
  j &= 7;
  if (j > 3) {
    j -= 4;
    ss = !ss;
    cs = !cs;
  }
  if (j > 1) cs = !cs;
  double z = ((x - y * PI4A) - y * PI4B) - y * PI4C;
  double zz = z * z;
  double c = 1.0 - 0.5 * zz + zz * zz * ((((((Cc[0] * zz) + Cc[1]) * zz + Cc[2]) * zz + Cc[3]) * zz + Cc[4]) * zz + Cc[5]);
  double sn = z + z * zz * ((((((S[0] * zz) + S[1]) * zz + S[2]) * zz + S[3]) * zz + S[4]) * zz + S[5]);
  if (j == 1 || j == 2) std::swap(sn, c);
  *s_out = ss ? -sn : sn;
  *c_out = cs ? -c : c;
}

struct dcx {
  double re, im;
};
static inline dcx dmul(dcx a, dcx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }

// The per-layer twiddles of the reference's Radix2 (kernels.go:38-59):
// w_L[k] = w_L[k-1] * wlen, wlen = (cos, sin)(-2*pi/L), for L = 2 .. n; the
// table for L sits at offset L/2 - 1.
std::vector<dcx> recurrence_table(i64 n) {
  std::vector<dcx> t;
  t.reserve(n);
  const double kPi = 3.14159265358979323846264338327950288;
  for (i64 L = 2; L <= n; L <<= 1) {
    double ang = -1.0 * 2 * kPi / (double)L;
    double sn, c;
    go_sincos(ang, &sn, &c);
    dcx wlen{c, sn}, w{1, 0};
    for (i64 k = 0; k < L / 2; k++) {
      t.push_back(w);
      w = dmul(w, wlen);
    }
  }
  return t;
}

// Reference Bluestein set-up (bluestein.go:28-63) in float64 with separately
// rounded products: chirps from Go's sin/cos and kernel spectra by
// BitReverse + Radix2 with recurrence twiddles.
void reference_bluestein_tables(i64 n, i64 m, const std::vector<i64>& jj, std::vector<dcx>& wF,
                                std::vector<dcx>& wI, std::vector<dcx>& bF, std::vector<dcx>& bI) {
  const double kPi = 3.14159265358979323846264338327950288;
  wF.resize(n);
  wI.resize(n);
  for (i64 j = 0; j < n; j++) {
    double ang = kPi * (double)jj[j] / (double)n;
    double s1, c1, s2, c2;
    go_sincos(-ang, &s1, &c1);
    go_sincos(ang, &s2, &c2);
    wF[j] = {c1, s1};
    wI[j] = {c2, s2};
  }
  std::vector<dcx> tab = recurrence_table(m);
  auto spectrum = [&](const std::vector<dcx>& w, std::vector<dcx>& b) {
    b.assign(m, dcx{0, 0});
    b[0] = {w[0].re, -w[0].im};
    for (i64 j = 1; j < n; j++) {
      dcx c{w[j].re, -w[j].im};
      b[j] = c;
      b[m - j] = c;
    }
    for (i64 i = 1, j = 0; i < m; i++) {  // kernels.go:20-32
      i64 bit = m >> 1;
      for (; j & bit; bit >>= 1) j ^= bit;
      j ^= bit;
      if (i < j) std::swap(b[i], b[j]);
    }
    for (i64 L = 2; L <= m; L <<= 1) {  // kernels.go:38-59
      const dcx* wl = tab.data() + (L / 2 - 1);
      i64 half = L >> 1;
      for (i64 st = 0; st < m; st += L)
        for (i64 k = 0; k < half; k++) {
          dcx u = b[st + k];
          dcx v = dmul(b[st + k + half], wl[k]);
          b[st + k] = {u.re + v.re, u.im + v.im};
          b[st + k + half] = {u.re - v.re, u.im - v.im};
        }
    }
  };
  spectrum(wF, bF);
  spectrum(wI, bI);
}
#pragma clang fp contract(on)

// Chirp phase indices j^2 mod 2n, j < n (bluestein.go:35-41), exact in
// 128-bit integers: the chirp is w_j = exp(-pi i (j^2 mod 2n) / n).
std::vector<i64> chirp_index(i64 n) {
  std::vector<i64> jj(n);
  for (i64 j = 0; j < n; j++) jj[j] = (i64)((__int128)j * j % (2 * n));
  return jj;
}

int upload_dcx(DevBuf& b, const std::vector<dcx>& v, int prec, double scale = 1.0) {
  std::vector<cld> t(v.size());
  for (size_t i = 0; i < v.size(); i++) t[i] = cld(v[i].re * scale, v[i].im * scale);
  return upload_complex(b, t, prec);
}

// Largest power-of-two column count <= cap with C*R points within the LDS
// budget `pts` and a feasible schedule.
bool pick_cols(int R, int cap, i64 pts, SubPlan& sp, int* Cout) {
  int C = 1;
  while (C * 2 <= cap && (i64)C * 2 * pitch(R) <= pts) C *= 2;
  for (; C >= 1; C /= 2) {
    if ((i64)C * pitch(R) > pts) continue;
    if (make_sub(R, C, sp)) {
      *Cout = C;
      return true;
    }
  }
  return false;
}

// Largest power-of-two column count (at most kMaxCols) that still gives
// every CU about two workgroups over `lines` lines: packing many lines into
// few workgroups would leave most of the chip idle on a small batch.
int spread_cols(i64 lines, int cus) {
  const i64 want = 2 * (i64)std::max(cus, 1);
  int c = 1;
  while (c < kMaxCols && lines / (2 * c) >= want) c *= 2;
  return c;
}

// Schedule of a fused single-kernel convolution of length m (k_conv): C
// columns within the preferred LDS budget when they fit (fewer when `lines`
// lines would otherwise fill too few workgroups, spread_cols), else one
// column with the whole LDS. The complex64 kernel also has a 24-points-per-thread form
// (radix <= 8 stages), and may pad every 32 or 64 points instead of 16, which
// takes a single column up to 20480 points.
bool fused_conv_plan(i64 m, int prec, i64 pts, SubPlan& sp, int colcap = 16) {
  int C = 0;
  if (pitch(m) <= pts &&
      pick_cols((int)m, colcap, std::min<i64>(pts, std::max<i64>(pitch(m), soft_pts(prec))), sp, &C))
    return true;
  if (prec != GOFFT_C64 || m > 24 * 1024) return false;
  for (int sh : {4, 5, 6})
    if (pitch(m, sh) <= pts)
      for (int maxrad : {16, 8})
        if (make_sub((int)m, 1, sp, 0, sh, kMaxP, maxrad)) return true;
  // the wide form, unpadded (kernels.hip lpm): psh = kNoPad makes r >> psh
  // zero for every point, so the pitch is m + 1
  if (pitch(m, kNoPad) <= pts)
    for (int maxrad : {16, 8})
      if (make_sub((int)m, 1, sp, 0, kNoPad, kWideP, maxrad)) return true;
  return false;
}

// ---------------------------------------------------------------------------
// Engines: a complex transform of length n over `batch` lines.

struct LineIO {
  const i64* off = nullptr;  // device offsets (elements) per line, or null
  i64 stride = 0;            // elements between lines when off is null
};

struct Ctx {
  hipStream_t s;
  char* ws;  // workspace for this engine (ws_bytes() bytes)
};

// A real inverse transform's pack, fused into the first pass of its
// half-length multi-pass transform (kernels.hip pack_step).
struct RealFirst {
  const void* utw = nullptr;  // W_n^k, k = 0..m
  const double2* uhi = nullptr;  // or hi[k >> ushift] * lo[k mod 2^ushift] (FP64)
  const double2* ulo = nullptr;
  int ushift = 0;
  i64 bins = 0;               // bins per line present in the input
  double w_mid = 1.0;         // scale of bins 1..m-1
};

// A real forward transform's untangle, fused into the last pass of its
// half-length multi-pass transform (kernels.hip untangle_store).
struct RealLast {
  const void* utw = nullptr;  // W_n^k, k = 0..m
  const double2* uhi = nullptr;  // or hi[k >> ushift] * lo[k mod 2^ushift] (FP64)
  const double2* ulo = nullptr;
  int ushift = 0;
  i64 ostride = 0;            // output elements between lines
  int mode = 0;               // 0 bins, 1 one-sided PSD, 2 weighted bins
  double psd = 1.0, w_edge = 1.0, w_mid = 1.0;
};

// Launches of one exec per kernel family (GOFFT_KERNEL_*) and the data bytes
// they move: each launch's input and output lines once, tables excluded.
struct Work {
  i64 launches[GOFFT_KERNEL_COUNT] = {};
  i64 bytes[GOFFT_KERNEL_COUNT] = {};
  void add(int k, i64 l, i64 b) {
    launches[k] += l;
    bytes[k] += b;
  }
};

class Engine {
 public:
  int prec = 0;
  i64 n = 0, batch = 0;
  // Set by the autotuner: the convolution length (0: the cost model's) and
  // the columns per workgroup of multi-pass passes (0: the default).
  i64 force_m = 0;
  int force_cols = 0;
  bool tuned = false;
  virtual ~Engine() {}
  virtual size_t ws_bytes() const = 0;
  virtual int launches() const = 0;
  // Adds the launches and data bytes of a run over `lines` lines.
  virtual void work(Work& w, i64 lines) const = 0;
  // One line of text naming the engine and its launches.
  virtual std::string describe() const = 0;
  // Whether a workgroup of this engine takes more than half a CU's LDS, so
  // that it can start only on a CU the other launches have left.
  virtual bool whole_cu() const { return false; }
  // Whether run() enqueues only pass and fused-convolution launches (so that
  // a mixed-length plan's work queue can take them over).
  virtual bool queueable() const { return false; }
  // out = scale * conj?( FFT( conj?(in) ) ) per line, for the first `nl`
  // lines (all `batch` lines when nl < 0).
  virtual int run(const void* in, LineIO li, void* out, LineIO lo, int conj_in, int conj_out,
                  double scale, const Ctx& c, i64 nl = -1) const = 0;
};

// Lines of a multi-kernel transform are processed in chunks whose
// intermediate buffers stay within this many bytes each (GOFFT_CHUNK_BYTES).
// Measured on MI355X, small chunks lose more to partly filled launches than
// they gain from cache residency, so the default only bounds workspace size.
i64 chunk_bytes() {
  static i64 v = [] {
    const char* e = std::getenv("GOFFT_CHUNK_BYTES");
    i64 x = e ? std::atoll(e) : 0;
    return x > 0 ? x : (i64)2 << 30;
  }();
  return v;
}

// Where the pass kernels stream line data through non-temporal loads and
// stores (PassArgs::nt): bits 1/2 a convolution chain's workspace reads and
// writes, 4/8 a multi-pass transform's workspace reads and writes, 16/32 a
// plain transform's reads and writes of aligned caller lines.
#ifndef GOFFT_REAL_TAB
#define GOFFT_REAL_TAB 1
#endif
constexpr bool kRealTab = GOFFT_REAL_TAB != 0;

#ifndef GOFFT_NT_POLICY
#define GOFFT_NT_POLICY 63
#endif
constexpr int kNtPolicy = GOFFT_NT_POLICY;
int nt_bits(bool load_ok, int load_class, bool store_ok, int store_class) {
  return (load_ok && (kNtPolicy & load_class) ? kNtLoad : 0) | (store_ok && (kNtPolicy & store_class) ? kNtStore : 0);
}

i64 chunk_lines(i64 per_line_bytes, i64 batch) {
  i64 cb = std::max<i64>(1, chunk_bytes() / std::max<i64>(1, per_line_bytes));
  return std::max<i64>(1, std::min<i64>(cb, batch));
}

// Advance a line address by b0 lines: offset tables index from b0, strided
// lines move the base pointer.
const void* line_ptr(const void* p, const LineIO& li, i64 b0, size_t eb) {
  return li.off ? p : (const void*)((const char*)p + (size_t)(b0 * li.stride) * eb);
}
void* line_ptr(void* p, const LineIO& li, i64 b0, size_t eb) {
  return li.off ? p : (void*)((char*)p + (size_t)(b0 * li.stride) * eb);
}
LineIO line_sub(const LineIO& li, i64 b0) { return LineIO{li.off ? li.off + b0 : nullptr, li.stride}; }

std::unique_ptr<Engine> make_engine(i64 n, i64 batch, int prec, unsigned flags, int* rc);

// Launch recording: while t_record is set on a thread, the engines' pass and
// fused-convolution launches are appended to it instead of being enqueued.
// A mixed-length plan records its groups' runs this way, at plan creation
// with stand-in data pointers, to build its work-queue stages.
struct RecordedLaunch {
  int kind = 0;  // 0 pass, 1 fused convolution
  PassArgs pa{};
  ConvArgs ca{};
  int threads = 0;
  int maxp = 0;
};
thread_local std::vector<RecordedLaunch>* t_record = nullptr;

hipError_t issue_pass(int prec, int maxp, const PassArgs& a, int threads, hipStream_t s) {
  if (t_record) {
    RecordedLaunch r;
    r.kind = 0;
    r.pa = a;
    r.threads = threads;
    r.maxp = maxp;
    t_record->push_back(r);
    return hipSuccess;
  }
  return launch_pass(prec, maxp, a, threads, s);
}

hipError_t issue_conv(int prec, int maxp, const ConvArgs& a, int threads, hipStream_t s) {
  if (t_record) {
    RecordedLaunch r;
    r.kind = 1;
    r.ca = a;
    r.threads = threads;
    r.maxp = maxp;
    t_record->push_back(r);
    return hipSuccess;
  }
  return launch_conv(prec, maxp, a, threads, s);
}

// Trivial engine (n <= 1): out = scale * in (conjugations cancel for n = 1).
class TrivialEngine : public Engine {
 public:
  DevBuf one;
  int init() {
    std::vector<cld> v(1, cld(1, 0));
    return upload_complex(one, v, prec);
  }
  size_t ws_bytes() const override { return 0; }
  int launches() const override { return n == 1 ? 1 : 0; }
  void work(Work& w, i64 lines) const override {
    if (n == 1 && lines > 0) w.add(GOFFT_KERNEL_AUX, 1, 2 * lines * (i64)elt(prec));
  }
  std::string describe() const override { return "trivial n=" + std::to_string(n); }
  int run(const void* in, LineIO li, void* out, LineIO lo, int ci, int co, double scale,
          const Ctx& c, i64 nl) const override {
    const i64 lines = nl < 0 ? batch : nl;
    if (n != 1 || lines == 0) return GOFFT_OK;
    HIP_TRY(launch_cmul_table(prec, in, out, one.p, 1, lines, li.stride, lo.stride, ci, co, scale, li.off,
                              lo.off, c.s));
    return GOFFT_OK;
  }
};

// The twiddle copy of a recurrence pass with Ns > 1 (make_sub_r2): for each
// stage and layer, entry cdig*ns*2^u + i is the reference's w_L[i*Ns + cdig]
// (`full` holds layer L at offset L/2 - 1).
void rec_pass_table(const std::vector<dcx>& full, const SubPlan& sp, i64 Ns, std::vector<dcx>& out) {
  for (int si = 0; si < sp.f.nstages; si++) {
    const Stage& g = sp.f.st[si];
    const int t = ilog2(g.radix - 100);
    for (int u = 0; u < t; u++) {
      const i64 S = (i64)g.ns << u;  // entries per column digit
      const i64 L = Ns * S * 2;
      for (i64 cd = 0; cd < Ns; cd++)
        for (i64 i = 0; i < S; i++) out.push_back(full[(size_t)(L / 2 - 1 + i * Ns + cd)]);
    }
  }
}

// One pass of a global Stockham decomposition (or the single pass).
struct Pass {
  SubPlan sp;
  i64 Ns = 1, R = 1;
  ColMap im{}, om{};
  i64 tw_M = 0;
  int tw_shift = 0;
  int tw_digit = 0;
  size_t stw_off = 0;  // element offset into the stage-twiddle buffer
  size_t ptw_hi = 0, ptw_lo = 0;  // double2 offsets into the pass-twiddle buffer
  int xcd_group = 1;  // > 1: consecutive column groups share cache lines, placed on one XCD
};

// Factor n into pass lengths each <= rmax (13-smooth n), balancing products.
bool factor_passes(i64 n, i64 rmax, int P, std::vector<i64>& out) {
  std::vector<i64> primes;
  i64 x = n;
  for (int p : {13, 11, 7, 5, 3, 2})
    while (x % p == 0) {
      primes.push_back(p);
      x /= p;
    }
  if (x != 1) return false;
  std::sort(primes.begin(), primes.end(), [](i64 a, i64 b) { return a > b; });
  std::vector<i64> prod(P, 1);
  for (i64 p : primes) {
    int best = -1;
    for (int i = 0; i < P; i++)
      if (prod[i] * p <= rmax && (best < 0 || prod[i] < prod[best])) best = i;
    if (best < 0) return false;
    prod[best] *= p;
  }
  std::sort(prod.begin(), prod.end(), [](i64 a, i64 b) { return a > b; });
  while (!prod.empty() && prod.back() == 1) prod.pop_back();
  out = prod;
  return !out.empty();
}

// Shortest column run (bytes) a two-pass 1-D plan may use. Measured on
// MI355X against the three-pass plans (ms, complex64 / complex128):
//   2^21 x 4  (1024 x 2048, 128- and 64-byte runs)  0.114 -> 0.092, 0.203 -> 0.185
//   2^23      (2048 x 4096, 64 and 32 bytes)          0.097 -> 0.100, 0.202 -> 0.240
//   2^24      (4096 x 4096, 32 bytes)                 0.195 -> 0.233, 0.364 -> 0.578
//   RFFT 2^24 float32 (pairs of 32 bytes)             0.133 -> 0.136; PSD 2^26 (16 bytes) 0.519 -> 0.842
// Passes of 4096 or more points take a whole CU's LDS for a few columns,
// one workgroup per CU, and lose more to latency than the third pass costs.
constexpr i64 kTwoPassMinRun = 64;
static i64 two_pass_min_run() { return kTwoPassMinRun; }

// Stockham engine over lines that may be strided: a line is n elements at
// element stride `st`; `inner` lines share an outer block (adjacent in memory),
// For a plain 1-D batch st = 1 and inner = 1.
class StockhamEngine : public Engine {
 public:
  i64 st = 1;        // element stride along the transform axis
  i64 inner = 1;     // lines interleaved at distance 1 (the axis stride)
  std::vector<Pass> passes;
  DevBuf stw, ptw;
  // complex64 pass-twiddle form (kernels.hip PassTw): 0 FP64-formed, 1 FP32,
  // 2 hi + lo float pairs; ptwx holds the tables of forms 1 and 2
  int tw_mode = 0;
  DevBuf ptwx;
  // The half-length transform of a real plan: the pass whose columns must
  // pair with their mirrors (0 none, 1 the last: the forward untangle, 2 the
  // first: the inverse pack), so that twice its columns fit LDS.
  int pair_req = 0;
  bool two_pass = false;  // planned by two_pass_lengths
  int P() const { return (int)passes.size(); }
  bool queueable() const override { return st == 1 && inner == 1 && tw_mode == 0; }

  // Power-of-two transform with the reference's recurrence twiddles
  // (kernels.Radix2): radix-2 layers only, grouped into passes of up to 2^8
  // layers; no pass twiddles (each layer reads its own table).
  // log2 pass lengths of a recurrence transform: one pass when it fits the
  // preferred LDS budget, else balanced passes of at most 2^8 points.
  static std::vector<int> recurrence_qs(i64 n, int prec, i64 pts) {
    const int q = ilog2((int)std::min<i64>(n, 1 << 30));
    std::vector<int> qs;
    if (pitch(n) <= std::min<i64>(pts, soft_pts(prec))) {
      qs.push_back(q);
    } else {
      int P = (q + 7) / 8;
      for (int i = 0; i < P; i++) qs.push_back(q / P + (i < q % P ? 1 : 0));
    }
    return qs;
  }

  int init_recurrence(std::vector<int> qs = {}) {
    DevInfo dev;
    RC_TRY(current_device(&dev));
    if (n < 2 || (n & (n - 1)) || st != 1 || inner != 1)
      return fail(GOFFT_ERR_UNSUPPORTED, "recurrence transform needs a contiguous power of two");
    const i64 pts = dev.lds_bytes / (i64)elt(prec);
    if (qs.empty()) qs = recurrence_qs(n, prec, pts);
    passes.clear();
    i64 Ns = 1;
    for (int qp : qs) {
      Pass p;
      p.R = (i64)1 << qp;
      p.Ns = Ns;
      int C = qs.size() == 1 ? 1 : 8;
      if (qs.size() == 1) {
        while (C * 2 <= kMaxCols && C * 2 <= batch && (i64)C * 2 * pitch(p.R) <= std::min<i64>(pts, soft_pts(prec)))
          C *= 2;
      }
      if (!make_sub_r2((int)p.R, C, Ns, p.sp))
        return fail(GOFFT_ERR_UNSUPPORTED, "no radix-2 schedule for pass length " + std::to_string(p.R));
      passes.push_back(p);
      Ns *= p.R;
    }
    for (auto& p : passes) {
      const i64 R = p.R, Nsp = p.Ns;
      p.im.D0 = 1;
      p.im.D1 = Nsp;
      p.im.D2 = n / (R * Nsp);
      p.im.a0 = 1;
      p.im.a1 = 1;
      p.im.a2 = Nsp;
      p.im.a3 = n;
      p.im.g = n / R;
      p.im.rowmajor = p.im.g == 1;
      p.om = p.im;
      p.om.a2 = Nsp * R;
      p.om.g = Nsp;
      p.om.rowmajor = p.om.g == 1;
      colmap_magic(p.im, batch * (n / R));
      colmap_magic(p.om, batch * (n / R));
      p.stw_off = 0;
      p.tw_M = 0;
      p.tw_digit = 1;
    }
    std::vector<dcx> tab = recurrence_table(n);
    const size_t nfull = tab.size();
    std::vector<dcx> extra;
    for (auto& p : passes) {
      if (p.Ns == 1) continue;
      p.stw_off = nfull + extra.size();
      rec_pass_table(tab, p.sp, p.Ns, extra);
    }
    tab.insert(tab.end(), extra.begin(), extra.end());
    return upload_dcx(stw, tab, prec);
  }

  int init(unsigned flags) {
    if (prec == GOFFT_C64) tw_mode = (flags & GOFFT_FLAG_TWIDDLES_SPLIT) ? 2 : (flags & GOFFT_FLAG_TWIDDLES_FP32) ? 1 : 0;
    DevInfo dev;
    RC_TRY(current_device(&dev));
    const i64 pts = dev.lds_bytes / (i64)elt(prec);
    const i64 lines = batch * inner;  // columns of a single pass
    // Single pass: all n points of a line in one workgroup. Strided lines need
    // several adjacent lines per workgroup for coalesced access.
    const int cwant = st == 1 ? 1 : (int)std::min<i64>(prec == 0 ? 8 : 16, st);
    const int ccap = (int)std::min<i64>(std::min<i64>(kMaxCols, std::max<i64>(cwant, lines)),
                                        std::max(cwant, spread_cols(lines, dev.cus)));
    const i64 budget = std::min<i64>(pts, std::max<i64>(pitch(n) * cwant, soft_pts(prec)));
    SubPlan sp;
    int C = 0;
    if (pitch(n) * cwant <= pts && pick_cols((int)n, ccap, budget, sp, &C) && C >= cwant) {
      Pass p;
      p.sp = sp;
      p.R = n;
      p.Ns = 1;
      passes.push_back(p);
    } else if (int Cn = narrow_cols(cwant, pts, sp)) {
      // Strided lines too long for cwant columns: one pass of Cn columns,
      // the cwant / Cn workgroups that split each 128-byte run on one XCD.
      Pass p;
      p.sp = sp;
      p.R = n;
      p.Ns = 1;
      p.xcd_group = cwant / Cn;
      passes.push_back(p);
    } else {
      std::vector<i64> lens;
      RC_TRY(multi_lengths(lens));
      std::vector<i64> l2;
      std::vector<int> c2;
      if (lens.size() >= 3 && two_pass_lengths(pts, l2, c2)) {
        two_pass = true;
        return init_lengths(l2, c2);
      }
      return init_lengths(lens);
    }
    return finish();
  }

  // Two passes instead of three or more for a contiguous 1-D transform: n =
  // R1 x R2 with each pass taking fewer columns than the usual 128-byte runs
  // (and, for a real plan's paired pass, half of what fits), at least
  // 16-byte runs; consecutive column groups then share cache lines and run
  // on one XCD (xcd_group). Of the factor pairs, the one whose narrowest
  // pass has the widest runs, then the more balanced.
  bool two_pass_lengths(i64 pts, std::vector<i64>& lens, std::vector<int>& cols) const {
    if (st != 1 || inner != 1 || force_cols > 0 || two_pass_min_run() <= 0) return false;
    const int cwant = prec == 0 ? 8 : 16;
    i64 best_run = 0, best_bal = 0;
    for (i64 R1 = 2; R1 <= std::min<i64>(16384, n / 2); R1++) {
      if (n % R1) continue;
      const i64 R2 = n / R1;
      if (R2 > 16384) continue;
      std::vector<i64> ls = {R1, R2};
      std::vector<int> cs(2, 0);
      i64 run = 1ll << 40;
      bool ok = true;
      for (int i = 0; i < 2 && ok; i++) {
        const bool paired = (pair_req == 1 && i == 1) || (pair_req == 2 && i == 0);
        int C = cwant;
        while (C >= 1 && (i64)(paired ? 2 * C : C) * pitch(ls[i]) > pts) C /= 2;
        SubPlan sp, sp2;
        // a paired pass needs a schedule of 2C columns with the same stages
        // (StockhamEngine::paired)
        auto fits = [&](int c) {
          if (!make_sub((int)ls[i], c, sp)) return false;
          if (!paired) return true;
          if (!make_sub((int)ls[i], 2 * c, sp2) || sp2.f.nstages != sp.f.nstages || sp2.tw != sp.tw) return false;
          for (int k = 0; k < sp.f.nstages; k++)
            if (sp2.f.st[k].radix != sp.f.st[k].radix) return false;
          return true;
        };
        while (C >= 1 && !fits(C)) C /= 2;
        if (C < 1) ok = false;
        cs[i] = C;
        run = std::min<i64>(run, (i64)C * (i64)elt(prec));
      }
      if (!ok || run < two_pass_min_run()) continue;
      const i64 bal = -std::abs(ilog2((int)R1) - ilog2((int)R2));
      if (run > best_run || (run == best_run && bal > best_bal)) {
        best_run = run;
        best_bal = bal;
        lens = ls;
        cols = cs;
      }
    }
    return best_run > 0;
  }

  // Columns for one pass over strided lines that do not fit cwant columns:
  // a power of two below cwant whose runs are still at least 32 bytes (so at
  // most four workgroups split a 128-byte run); 0 when none fits or the
  // lines are contiguous. Measured on MI355X: FFT2 4096^2 complex64 (4
  // columns) 0.219 -> 0.150 ms and 2048^2 0.067 -> 0.044 ms; 16-byte runs
  // (RFFT2 8192^2's columns, 2 of them) were 9% slower than two passes.
  int narrow_cols(int cwant, i64 pts, SubPlan& sp) const {
    if (st == 1 || force_cols > 0) return 0;
    for (int c = cwant / 2; c >= 2 && c * (i64)elt(prec) >= 32; c /= 2)
      if (pitch(n) * c <= pts && make_sub((int)n, c, sp)) return c;
    return 0;
  }

  // Pass lengths of a multi-pass decomposition: each <= rmax so that cmp
  // columns of it fit in LDS, as few passes as possible, balanced.
  int multi_lengths(std::vector<i64>& lens) const {
    DevInfo dev;
    RC_TRY(current_device(&dev));
    const i64 pts = dev.lds_bytes / (i64)elt(prec);
    int cmp = prec == 0 ? 8 : 16;
    i64 rmax = std::min<i64>(pts / cmp * 16 / 17, 1024);
    bool ok = false;
    for (int Pn = 2; Pn <= 6 && !ok; Pn++) ok = factor_passes(n, rmax, Pn, lens);
    if (!ok) return fail(GOFFT_ERR_UNSUPPORTED, "no pass decomposition for n=" + std::to_string(n));
    return GOFFT_OK;
  }

  // Multi-pass with the given pass lengths (product n), in order; `cols`,
  // if given, caps each pass's columns (narrow passes, placed by XCD).
  int init_lengths(const std::vector<i64>& lens, const std::vector<int>& cols = {}) {
    DevInfo dev;
    RC_TRY(current_device(&dev));
    const i64 pts = dev.lds_bytes / (i64)elt(prec);
    int cmp = force_cols > 0 ? force_cols : prec == 0 ? 8 : 16;
    // complex128 runs 16 columns (256-byte runs) per pass when 16 columns of
    // every pass fit 48 KB of LDS (three workgroups per CU). Measured: 16384
    // x 1024 0.278 -> 0.267 ms; Rader chains 5003 x 1024 0.346 -> 0.314 ms,
    // 10007 x 585 0.343 -> 0.332 ms (passes of 128 and 160). Passes of 256
    // to 512 points keep 8: 262144 x 64 was 10% and 40009 x 128 24% slower
    // with 16.
    if (prec == GOFFT_C128 && force_cols == 0) {
      bool fits = true;
      for (i64 R : lens) fits = fits && pitch(R) * 16 * (i64)elt(prec) <= 48 * 1024;
      if (fits) cmp = 16;
    }
    passes.clear();
    i64 Ns = 1;
    for (size_t li = 0; li < lens.size(); li++) {
      const i64 R = lens[li];
      Pass p;
      int Cp = 0;
      bool ok = pick_cols((int)R, cols.empty() ? cmp : cols[li], pts, p.sp, &Cp);
      if (!ok)
        return fail(GOFFT_ERR_UNSUPPORTED, "no LDS schedule for pass length " + std::to_string(R));
      if (!cols.empty() && (i64)Cp * (i64)elt(prec) < 128) p.xcd_group = 2;
      p.R = R;
      p.Ns = Ns;
      passes.push_back(p);
      Ns *= R;
    }
    return finish();
  }

  // Column maps and tables for the chosen passes.
  int finish() {
    std::vector<cld> stw_h;
    std::vector<cld> ptw_h;
    for (size_t i = 0; i < passes.size(); i++) {
      Pass& p = passes[i];
      const i64 R = p.R, Ns = p.Ns;
      const i64 D2 = n / (R * Ns);
      // digits: d0 = inner, d1 = j % Ns, d2 = j / Ns, d3 = outer line (batch)
      p.im.D0 = inner;
      p.im.D1 = Ns;
      p.im.D2 = D2;
      p.im.a0 = 1;
      p.im.a1 = st;
      p.im.a2 = Ns * st;
      p.im.a3 = n * st;  // replaced by the caller's line stride for 1-D
      p.im.g = (n / R) * st;
      p.im.rowmajor = p.im.g == 1;
      p.om = p.im;
      p.om.a2 = Ns * R * st;
      p.om.g = Ns * st;
      p.om.rowmajor = p.om.g == 1;
      colmap_magic(p.im, batch * inner * (n / R));
      colmap_magic(p.om, batch * inner * (n / R));
      p.stw_off = stw_h.size();
      stw_h.insert(stw_h.end(), p.sp.tw.begin(), p.sp.tw.end());
      if (Ns > 1) {
        const i64 M = Ns * R;
        int sh = (ilog2((int)std::min<i64>(M, 1 << 30)) + 1) / 2;
        i64 S = 1ll << sh;
        i64 nh = (M + S - 1) / S;
        p.tw_M = M;
        p.tw_shift = sh;
        p.tw_digit = 1;
        if (p.im.rowmajor) return fail(GOFFT_ERR_UNSUPPORTED, "twiddled pass must walk across columns");
        p.ptw_hi = ptw_h.size();
        for (i64 h = 0; h < nh; h++) ptw_h.push_back(root(h * S, M));
        p.ptw_lo = ptw_h.size();
        for (i64 l = 0; l < S; l++) ptw_h.push_back(root(l, M));
      }
    }
    if (stw_h.empty()) stw_h.push_back(cld(1, 0));
    RC_TRY(upload_complex(stw, stw_h, prec));
    if (!ptw_h.empty()) RC_TRY(upload_complex(ptw, ptw_h, 0));
    if (tw_mode == 1 && !ptw_h.empty()) RC_TRY(upload_complex(ptwx, ptw_h, GOFFT_C64));
    if (tw_mode == 2 && !ptw_h.empty()) {
      // each entry as (hi.re, hi.im, lo.re, lo.im): hi = fl(W), lo = fl(W - hi)
      std::vector<float> h(4 * ptw_h.size());
      for (size_t i = 0; i < ptw_h.size(); i++) {
        const long double re = ptw_h[i].real(), im = ptw_h[i].imag();
        const float hr = (float)re, hi = (float)im;
        h[4 * i] = hr;
        h[4 * i + 1] = hi;
        h[4 * i + 2] = (float)(re - (long double)hr);
        h[4 * i + 3] = (float)(im - (long double)hi);
      }
      RC_TRY(ptwx.upload(h.data(), h.size() * sizeof(float)));
    }
    return GOFFT_OK;
  }

  // Outer units (d3 digit) per chunk of a multi-pass run.
  i64 chunk() const { return chunk_lines((i64)n * inner * (i64)elt(prec), batch); }

  size_t ws_bytes() const override {
    // ping-pong buffers for one chunk of outer units
    if (P() <= 1) return 0;
    size_t one = (size_t)n * (size_t)inner * (size_t)chunk() * elt(prec);
    return P() == 2 ? one : 2 * one;
  }
  int launches() const override {
    if (P() <= 1) return P();
    return (int)(P() * ((batch + chunk() - 1) / chunk()));
  }
  void work(Work& w, i64 lines) const override {
    if (lines <= 0 || P() == 0) return;
    const i64 chunks = P() <= 1 ? 1 : (lines + chunk() - 1) / chunk();
    w.add(GOFFT_KERNEL_PASS, P() * chunks, (i64)P() * 2 * n * inner * lines * (i64)elt(prec));
  }
  std::string passes_text() const {
    std::string s;
    for (int i = 0; i < P(); i++) s += (i ? " x " : "") + describe_sub(passes[i].R, passes[i].sp);
    return s;
  }
  std::string describe() const override {
    std::string s = std::string(tuned ? "tuned " : "") + "stockham n=" + std::to_string(n) +
                    " lines=" + std::to_string(batch);
    if (st != 1) s += " stride=" + std::to_string(st);
    if (P() > 1) s += " chunk=" + std::to_string(chunk());
    return s + " passes=" + passes_text();
  }

  int run(const void* in, LineIO li, void* out, LineIO lo, int ci, int co, double scale,
          const Ctx& c, i64 nl) const override {
    const i64 lines = nl < 0 ? batch : nl;
    if (lines == 0) return GOFFT_OK;
    if (P() <= 1) return run_chunk(in, li, out, lo, ci, co, scale, c, lines);
    // Strided (N-D) lines advance by whole outer blocks of n*st elements.
    const i64 cb = chunk();
    const size_t eb = elt(prec);
    for (i64 b0 = 0; b0 < lines; b0 += cb) {
      const i64 nb = std::min<i64>(cb, lines - b0);
      LineIO lis = inner == 1 ? li : LineIO{nullptr, n * st};
      LineIO los = inner == 1 ? lo : LineIO{nullptr, n * st};
      RC_TRY(run_chunk(line_ptr(in, lis, b0, eb), line_sub(li, b0), line_ptr(out, los, b0, eb),
                       line_sub(lo, b0), ci, co, scale, c, nb));
    }
    return GOFFT_OK;
  }

  // Whether lines at p (element stride or offset table io) may start off the
  // 128-byte grid.
  bool unaligned_lines(const void* p, LineIO io) const {
    return io.off != nullptr || ((size_t)io.stride * elt(prec)) % 128 != 0 || (uintptr_t)p % 128 != 0;
  }

  int run_chunk(const void* in, LineIO li, void* out, LineIO lo, int ci, int co, double scale,
                const Ctx& c, i64 lines, const RealLast* rl = nullptr, const RealFirst* rf = nullptr) const {
    const size_t one = (size_t)n * (size_t)inner * (size_t)chunk() * elt(prec);
    char* bufs[2] = {c.ws, c.ws + one};
    for (int i = 0; i < P(); i++) {
      const Pass& p = passes[i];
      PassArgs a{};
      a.in = i == 0 ? in : bufs[(i - 1) % 2];
      a.out = i == P() - 1 ? out : bufs[i % 2];
      a.stw = (const char*)stw.p + p.stw_off * elt(prec);
      a.ptw_hi = p.tw_M ? (const double2*)ptw.p + p.ptw_hi : nullptr;
      a.ptw_lo = p.tw_M ? (const double2*)ptw.p + p.ptw_lo : nullptr;
      if (p.tw_M && tw_mode) {
        // same layout, entries of 8 (FP32) or 16 (hi + lo) bytes
        const size_t xb = tw_mode == 1 ? 8 : 16;
        a.tw_mode = tw_mode;
        a.ptwx_hi = (const char*)ptwx.p + p.ptw_hi * xb;
        a.ptwx_lo = (const char*)ptwx.p + p.ptw_lo * xb;
      }
      a.ncols = lines * inner * (n / p.R);
      a.tw_M = p.tw_M;
      a.tw_shift = p.tw_shift;
      a.tw_digit = p.tw_digit;
      a.conj_in = i == 0 ? ci : 0;
      a.conj_out = i == P() - 1 ? co : 0;
      a.scale = i == P() - 1 ? scale : 1.0;
      a.im = p.im;
      a.om = p.om;
      if (i == 0) {
        a.in_off = li.off;
        if (inner == 1) a.im.a3 = li.stride;
      }
      if (i == P() - 1) {
        a.out_off = lo.off;
        if (inner == 1) a.om.a3 = lo.stride;
      }
      a.f = p.sp.f;
      a.xcd_group = p.xcd_group;
      // Caller lines at any alignment (a mixed-length plan's packed lines, or
      // a line stride or base off the 128-byte grid): the first pass's
      // column runs straddle cache lines that the neighbouring column group
      // also reads, and the last pass's runs lines that it also writes.
      // Consecutive groups then go to one XCD and meet in its L2; partial
      // lines written from two XCDs cost the most. Measured on the 4096-
      // transform mixed-length batch: FP32 0.850 -> 0.786 ms, FP64 2.173 ->
      // 2.094 ms. The same holds in any pass whose columns per line are not
      // a multiple of C, where a group's runs straddle two lines: 40000 x
      // 256 complex64 0.106 -> 0.100 ms, 20000 x 256 complex128 0.114 ->
      // 0.111 ms.
      if (P() > 1 && inner == 1 &&
          ((i == 0 && unaligned_lines(in, li)) || (i == P() - 1 && unaligned_lines(out, lo)) ||
           (n / p.R) % p.sp.f.C != 0))
        a.xcd_group = std::max(a.xcd_group, 2);
      // Strided axis passes (N-D, one or more passes) over rows that are not
      // a whole number of C-column groups or of 128-byte lines, such as an
      // RFFT2 spectrum's cols/2+1 columns: column groups straddle rows the
      // same way. Measured: RFFT2 8192^2 0.484 -> 0.445 ms and IRFFT2 0.582
      // -> 0.547 ms; RFFT2 of 64 x 1024^2 0.386 -> 0.350 ms and IRFFT2 0.476
      // -> 0.435 ms; FFTN 100^3 0.045 -> 0.042 ms.
      if (inner > 1 && (inner % p.sp.f.C != 0 || ((size_t)inner * elt(prec)) % 128 != 0))
        a.xcd_group = std::max(a.xcd_group, 2);
      int threads = p.sp.threads;
      if (rf && i == 0) {
        a.f = pair_first_sp.f;
        threads = pair_first_sp.threads;
        const i64 H = pair_first_sp.f.C / 2, Mc = n / p.R;
        a.real_first = 1;
        a.in_hook = 3;  // bins read up to rf->bins, zero beyond
        a.nvalid = rf->bins;
        a.ru_conj = a.conj_in;  // conjugation after the pack
        a.conj_in = 0;
        a.ru_m = n;
        a.ru_lines = lines;
        a.ru_wpl = (Mc / 2 + 1 + H - 1) / H;
        a.ru_utw = rf->utw;
        a.ru_uhi = rf->uhi;
        a.ru_ulo = rf->ulo;
        a.ru_ushift = rf->ushift;
        a.ru_w_mid = rf->w_mid;
        a.in_off = nullptr;
      }
      if (rl && i == P() - 1) {
        a.f = pair_sp.f;
        threads = pair_sp.threads;
        const i64 H = pair_sp.f.C / 2, Ns = n / p.R;
        a.real_last = 1;
        a.ru_m = n;
        a.ru_lines = lines;
        a.ru_wpl = (Ns / 2 + 1 + H - 1) / H;
        a.ru_utw = rl->utw;
        a.ru_uhi = rl->uhi;
        a.ru_ulo = rl->ulo;
        a.ru_ushift = rl->ushift;
        a.ru_ostride = rl->ostride;
        a.ru_mode = rl->mode;
        a.ru_psd = rl->psd;
        a.ru_w_edge = rl->w_edge;
        a.ru_w_mid = rl->w_mid;
        a.out_off = nullptr;
        // The mirrored half of a workgroup (columns Ns-j) writes runs of H
        // bins that start one bin past the 128-byte grid, so consecutive
        // workgroups share cache lines: keep them on one XCD. Measured: RFFT
        // 2^24 float32 0.136 -> 0.130 ms, float64 0.281 -> 0.270 ms (the
        // inverse's fused pack gains nothing from it).
        a.xcd_group = 2;
      }
      // Non-temporal line accesses where this launch's column groups own
      // whole 128-byte runs: 1-D lines (a single pass: whole lines per
      // workgroup; several: column groups that neither straddle lines nor
      // split runs), on workspace buffers or aligned caller lines.
      // (Strided N-D axes stay off: FFTN 256^3 complex64 0.150 -> 0.168 ms
      // with their aligned passes streamed.)
      if (inner == 1 && p.xcd_group <= 1 &&
          (P() == 1 || ((n / p.R) % p.sp.f.C == 0 && (size_t)p.sp.f.C * elt(prec) >= 128))) {
        const bool in_ws = i > 0, out_ws = i < P() - 1;
        // workspace lines: n points each, back to back
        const LineIO wio{nullptr, n};
        // The fused real forms stream only their workspace side: the
        // untangle's and the pack's bins go through their own code.
        // Measured: RFFT / IRFFT 2^24 float64 0.254 / 0.262 -> 0.232 / 0.240
        // ms (float32 unchanged).
        const bool ld = in_ws ? !a.real_first && !unaligned_lines(a.in, wio)
                              : !a.real_first && !a.real_last && !unaligned_lines(in, li);
        const bool st = out_ws ? !a.real_last && !unaligned_lines(a.out, wio)
                               : !a.real_first && !a.real_last && !unaligned_lines(out, lo);
        a.nt = nt_bits(ld, in_ws ? 4 : 16, st, out_ws ? 8 : 32);
      }
      HIP_TRY(issue_pass(prec, kMaxP, a, threads, c.s));
    }
    return GOFFT_OK;
  }

  // Whether run_real can fuse a real forward untangle into the last pass:
  // multi-pass, 1-D lines, and a schedule of the last pass with twice its
  // columns (each workgroup takes C columns and their C mirrors, so both
  // halves still read and write whole 128-byte runs) that fits LDS and keeps
  // the same stage twiddles.
  SubPlan pair_sp, pair_first_sp;
  // A schedule of pass sp with twice its columns and the same stages.
  bool paired(const SubPlan& sp, SubPlan& out) const {
    DevInfo dev;
    if (current_device(&dev) != GOFFT_OK) return false;
    const int C2 = 2 * sp.f.C;
    if (C2 > kMaxCols || (i64)C2 * pitch(sp.f.R) * (i64)elt(prec) > dev.lds_bytes) return false;
    if (!make_sub(sp.f.R, C2, out) || out.f.nstages != sp.f.nstages || out.tw != sp.tw) return false;
    for (int i = 0; i < sp.f.nstages; i++)
      if (out.f.st[i].radix != sp.f.st[i].radix) return false;
    return true;
  }
  bool prepare_untangle() { return P() >= 2 && inner == 1 && paired(passes.back().sp, pair_sp); }
  bool prepare_pack() { return P() >= 2 && inner == 1 && paired(passes.front().sp, pair_first_sp); }

  // The m-point transform of `lines` real lines read as complex pairs (istride
  // complex elements between lines), with the untangle to bins (or PSD) done
  // by the last pass: out holds ostride output elements per line (complex
  // bins, or reals for the PSD).
  int run_real(const void* in, i64 istride, void* out, const RealLast& rl, const Ctx& c, i64 lines) const {
    const i64 cb = chunk();
    const size_t eb = elt(prec), ob = rl.mode == 1 ? eb / 2 : eb;
    for (i64 b0 = 0; b0 < lines; b0 += cb) {
      const i64 nb = std::min<i64>(cb, lines - b0);
      const char* ib = (const char*)in + (size_t)(b0 * istride) * eb;
      char* o = (char*)out + (size_t)(b0 * rl.ostride) * ob;
      RC_TRY(run_chunk(ib, LineIO{nullptr, istride}, o, LineIO{nullptr, 0}, 0, 0, 1.0, c, nb, &rl));
    }
    return GOFFT_OK;
  }

  // The m-point inverse (conj(FFT(conj)) scaled) of `lines` bin lines
  // (istride complex elements between lines, rf.bins present), packed by the
  // first pass; out gets ostride complex elements (real pairs) per line.
  int run_real_inv(const void* in, i64 istride, void* out, i64 ostride, double scale, const RealFirst& rf,
                   const Ctx& c, i64 lines) const {
    const i64 cb = chunk();
    const size_t eb = elt(prec);
    for (i64 b0 = 0; b0 < lines; b0 += cb) {
      const i64 nb = std::min<i64>(cb, lines - b0);
      const char* ib = (const char*)in + (size_t)(b0 * istride) * eb;
      char* o = (char*)out + (size_t)(b0 * ostride) * eb;
      RC_TRY(run_chunk(ib, LineIO{nullptr, istride}, o, LineIO{nullptr, ostride}, 1, 1, scale, c, nb, nullptr, &rf));
    }
    return GOFFT_OK;
  }
};

// Cyclic convolution of length m over lines in 2P-1 kernels. The forward
// transform runs passes with radices (R1..RP) and the inverse runs them
// reversed (RP..R1): the forward's last pass writes column j's outputs to
// j + r*Ns with Ns = m/RP, which are exactly the inputs of the inverse's first
// pass column j. One kernel therefore runs both around the spectrum multiply.
// The Bluestein chirp or the Rader gather rides on the first load, and the
// chirp or the Rader scatter on the last store.
struct ConvChain {
  int prec = 0;
  i64 m = 0, cb = 0;
  StockhamEngine fwd, inv;

  struct Hooks {
    int in_hook = 0, out_hook = 0;
    i64 nvalid = 0, conv_off = 0;
    const void* w = nullptr;
    const int* perm = nullptr;
    const int* iperm = nullptr;
    const void* spec = nullptr;
  };

  int init(i64 m_, i64 cb_, int prec_, bool recurrence, int cols = 0) {
    m = m_;
    cb = cb_;
    prec = prec_;
    for (StockhamEngine* e : {&fwd, &inv}) {
      e->prec = prec;
      e->n = m;
      e->batch = cb;
      e->force_cols = cols;
    }
    if (recurrence) {
      DevInfo dev;
      RC_TRY(current_device(&dev));
      std::vector<int> qs = StockhamEngine::recurrence_qs(m, prec, dev.lds_bytes / (i64)elt(prec));
      if (qs.size() < 2) {
        int q = qs[0];
        qs = {q - q / 2, q / 2};
      }
      RC_TRY(fwd.init_recurrence(qs));
      std::reverse(qs.begin(), qs.end());
      RC_TRY(inv.init_recurrence(qs));
    } else {
      std::vector<i64> lens;
      RC_TRY(fwd.multi_lengths(lens));
      // The middle launch runs the forward's last pass length twice (forward
      // and inverse): give it the length with the fewest LDS stages.
      auto stages = [](i64 R) { return radix_schedule((int)R).size(); };
      auto mid = std::min_element(lens.begin(), lens.end(), [&](i64 x, i64 y) { return stages(x) < stages(y); });
      std::iter_swap(mid, lens.end() - 1);
      RC_TRY(fwd.init_lengths(lens));
      std::reverse(lens.begin(), lens.end());
      RC_TRY(inv.init_lengths(lens));
    }
    const Pass& a = fwd.passes.back();
    const Pass& b = inv.passes.front();
    if (fwd.P() < 2 || a.R != b.R || a.sp.f.C != b.sp.f.C || a.sp.threads != b.sp.threads)
      return fail(GOFFT_ERR_UNSUPPORTED, "convolution passes do not pair for m=" + std::to_string(m));
    return GOFFT_OK;
  }

  // two ping-pong buffers of cb lines, then per line the Rader DC term A[0]
  // and the saved x0
  size_t ws_bytes() const { return 2 * (size_t)m * cb * elt(prec) + 2 * (size_t)cb * elt(prec); }
  int launches() const { return 2 * fwd.P() - 1; }
  std::string describe() const {
    return "chain m=" + std::to_string(m) + " chunk=" + std::to_string(cb) + " fwd=" + fwd.passes_text() +
           " inv=" + inv.passes_text();
  }
  // One run over `lines` lines in chunks of cb: the first pass reads nin
  // points of each line, the last writes nout.
  void work(Work& w, i64 lines, i64 nin, i64 nout) const {
    if (lines <= 0) return;
    const i64 K = launches(), chunks = (lines + cb - 1) / cb;
    w.add(GOFFT_KERNEL_PASS, K * chunks, lines * (i64)elt(prec) * (nin + 2 * m * (K - 1) + nout));
  }

  // The mid kernel reads the spectrum at forward output position j + r*Ns
  // for column j and point r; storing it as spec[j*R + r] makes a column's
  // reads contiguous.
  int upload_spec(const std::vector<cld>& b, DevBuf& out) const {
    const i64 R = fwd.passes.back().R, Ns = m / R;
    std::vector<cld> t(m);
    for (i64 j = 0; j < Ns; j++)
      for (i64 r = 0; r < R; r++) t[j * R + r] = b[j + r * Ns];
    return upload_complex(out, t, prec);
  }

  void fill(PassArgs& a, const StockhamEngine& e, const Pass& p, i64 nl) const {
    a.stw = (const char*)e.stw.p + p.stw_off * elt(prec);
    a.ptw_hi = p.tw_M ? (const double2*)e.ptw.p + p.ptw_hi : nullptr;
    a.ptw_lo = p.tw_M ? (const double2*)e.ptw.p + p.ptw_lo : nullptr;
    a.ncols = nl * (m / p.R);
    a.tw_M = p.tw_M;
    a.tw_shift = p.tw_shift;
    a.tw_digit = p.tw_digit;
    a.scale = 1.0;
    a.im = p.im;
    a.om = p.om;
    a.f = p.sp.f;
  }

  // One chunk of nl <= cb lines.
  int run(const Hooks& h, const void* in, LineIO li, void* out, LineIO lo, int ci, int co, double scale,
          char* ws, hipStream_t s, i64 nl) const {
    char* bufs[2] = {ws, ws + (size_t)m * cb * elt(prec)};
    void* dc = ws + 2 * (size_t)m * cb * elt(prec);
    void* x0buf = (char*)dc + (size_t)cb * elt(prec);
    const int P = fwd.P();
    const int K = 2 * P - 1;
    for (int k = 0; k < K; k++) {
      PassArgs a{};
      int threads = 0;
      if (k < P - 1) {
        fill(a, fwd, fwd.passes[k], nl);
        threads = fwd.passes[k].sp.threads;
      } else if (k == P - 1) {
        const Pass& pf = fwd.passes[P - 1];
        const Pass& pi = inv.passes[0];
        fill(a, fwd, pf, nl);
        a.mm = pf.om;  // column j, point r -> spectrum entry j*R + r (see upload_spec)
        a.mm.a0 = 0;
        a.mm.a1 = pf.R;
        a.mm.a2 = 0;
        a.mm.g = 1;
        a.om = pi.om;
        a.mid_spec = h.spec;
        a.f2 = pi.sp.f;
        a.stw2 = (const char*)inv.stw.p + pi.stw_off * elt(prec);
        auto is16x16 = [](const SubPlan& sp) {
          return sp.f.R == 256 && sp.f.nstages == 2 && sp.f.st[0].radix == 16 && sp.f.st[1].radix == 16 &&
                 sp.f.tpc == 16 && sp.f.psh <= 4;
        };
        // (complex64 only: the complex128 form would spill registers)
        a.mid256 = prec == GOFFT_C64 && K == 3 && is16x16(pf.sp) && is16x16(pi.sp) &&
                   pf.sp.f.C == pi.sp.f.C && a.om.rowmajor;
        if (h.out_hook == 2) a.dc = dc;
        threads = pf.sp.threads;
      } else {
        const Pass& pi = inv.passes[k - (P - 1)];
        fill(a, inv, pi, nl);
        threads = pi.sp.threads;
      }
      a.in = k == 0 ? in : bufs[(k - 1) % 2];
      a.out = k == K - 1 ? out : bufs[k % 2];
      if (k == 0) {
        a.im.a3 = li.stride;
        a.in_off = li.off;
        a.in_hook = h.in_hook;
        a.nvalid = h.nvalid;
        a.hook_w = h.w;
        a.hook_perm = h.perm;
        a.conj_in = ci;
        if (h.in_hook == 2) a.x0buf = x0buf;
      }
      if (k == K - 1) {
        a.om.a3 = lo.stride;
        a.out_off = lo.off;
        a.out_hook = h.out_hook;
        a.nvalid = h.nvalid;
        a.conv_off = h.conv_off;
        a.hook_w = h.w;
        a.hook_perm = h.iperm;
        a.conj_out = co;
        a.scale = scale;
        if (h.out_hook == 2) {
          a.x0buf = x0buf;
          a.x0_conj = ci;
          a.dc = dc;
        }
      }
      // The first launch reads, and the last writes, the caller's lines:
      // n points each, packed back to back at any alignment, and gathered or
      // scattered by the generator permutation for Rader. A column group's
      // runs straddle cache lines that its neighbours also touch, so
      // consecutive groups go to one XCD and share its L2. Measured: Rader
      // 10007 x 585 complex128 0.395 -> 0.343 ms, Bluestein 20014 x 585
      // complex64 0.277 -> 0.271 ms; the middle launches (aligned workspace
      // rows) gain nothing from it.
      if (k == 0 || k == K - 1) a.xcd_group = 2;
      // workspace rows (m points each) stream through non-temporal accesses
      // when they lie on the 128-byte grid; the caller's lines do not (see
      // xcd_group)
      auto grid = [&](const void* q) { return (uintptr_t)q % 128 == 0 && ((size_t)m * elt(prec)) % 128 == 0; };
      a.nt = nt_bits(k > 0 && grid(a.in), 1, k < K - 1 && grid(a.out), 2);
      HIP_TRY(issue_pass(prec, kMaxP, a, threads, s));
    }
    return GOFFT_OK;
  }
};

// Convolution length for Rader's linear path and for Bluestein outside the
// reference-twiddle mode: a 7-smooth length in [lo, ceil(1.4 lo)] (the window
// of plan.go:241-259). A length that one LDS
// pass holds wins outright; otherwise the cost is the HBM passes of the
// Stockham decomposition, each weighted by its LDS stage count (measured on
// MI355X: radix-16 passes run about 40% faster per byte than passes of four
// small radices).
std::vector<i64> conv_candidates(i64 lo, int prec, int k);

i64 choose_conv_len(i64 lo, int prec) {
  std::vector<i64> c = conv_candidates(lo, prec, 1);
  return c.empty() ? -1 : c[0];
}

// The k best convolution lengths in [lo, ceil(1.4 lo)] by the cost model:
// lengths that one workgroup holds first, then by cost.
std::vector<i64> conv_candidates(i64 lo, int prec, int k) {
  DevInfo dev;
  current_device(&dev);
  const i64 pts = dev.lds_bytes / (i64)elt(prec);
  const i64 pts_full = dev.lds_full / (i64)elt(prec);
  const int cmp = prec == 0 ? 8 : 16;
  const i64 rmax = std::min<i64>(pts / cmp * 16 / 17, 1024);
  i64 hi = lo + (lo * 4 + 9) / 10;
  std::vector<std::tuple<int, double, i64>> all;  // (-fit, cost, m)
  for (i64 m = lo; m <= hi; m++) {
    i64 x = m;
    for (int p : {2, 3, 5, 7})
      while (x % p == 0) x /= p;
    if (x != 1) continue;
    SubPlan tmp;
    int fit = fused_conv_plan(m, prec, pts_full, tmp) ? 1 : 0;
    double cost;
    if (fit) {
      cost = (double)m * (1.0 + 0.25 * (double)radix_schedule((int)m).size());
    } else {
      std::vector<i64> lens;
      bool ok = false;
      for (int P = 2; P <= 6 && !ok; P++) ok = factor_passes(m, rmax, P, lens);
      if (!ok) continue;
      cost = 0;
      for (i64 R : lens) cost += (double)m * (1.0 + 0.25 * (double)radix_schedule((int)R).size());
    }
    all.emplace_back(-fit, cost, m);
  }
  std::sort(all.begin(), all.end());
  std::vector<i64> out;
  for (size_t i = 0; i < all.size() && (int)out.size() < k; i++) out.push_back(std::get<2>(all[i]));
  return out;
}

// Rader's algorithm for prime n (rader.go:39-172): gather by the generator
// permutation, cyclic convolution with the precomputed kernel spectrum,
// scatter by the inverse permutation.
class RaderEngine : public Engine {
 public:
  i64 q = 0, cl = 0, g = 0;  // g: the primitive root whose powers order the gather
  bool cyclic = true;
  DevBuf perm, iperm, oinv, dlog, spec;
  // fused path
  bool fused = false;
  bool big_lds = false;  // its workgroup takes more than half a CU's LDS
  SubPlan sp;
  DevBuf stw;
  // multi-kernel path
  ConvChain chain;

  int init(unsigned flags) {
    (void)flags;
    q = n - 1;
    g = primitive_root(n);
    if (factors_are_small(q)) {
      cyclic = true;
      cl = q;
    } else {
      cyclic = false;
      cl = force_m > 0 ? force_m : choose_conv_len(2 * q, prec);
      if (cl < 2 * q)
        return fail(GOFFT_ERR_UNSUPPORTED, "no convolution length >= " + std::to_string(2 * q) + " for Rader n=" +
                                               std::to_string(n));
    }
    std::vector<int> ph(q), iph(q);
    i64 x = 1;
    for (i64 k = 0; k < q; k++) {
      ph[k] = (int)x;
      x = x * g % n;
    }
    for (i64 k = 0; k < q; k++) iph[k] = ph[(q - k) % q];
    RC_TRY(perm.upload(ph.data(), ph.size() * sizeof(int)));
    RC_TRY(iperm.upload(iph.data(), iph.size() * sizeof(int)));
    // kernel ker[m] = W_n^{perm[(q - m) mod q]} replicated over 2q for the
    // linear case (rader.go:101-125); spectrum scaled by 1/cl.
    std::vector<cld> b(cl, cld(0, 0));
    i64 lim = cyclic ? q : 2 * q;
    for (i64 m = 0; m < lim; m++) b[m] = root(ph[(q - (m % q)) % q], n);
    host_fft(b);
    for (auto& v : b) v /= (long double)cl;
    RC_TRY(upload_complex(spec, b, prec));  // natural order (fused kernel)
    DevInfo dev;
    RC_TRY(current_device(&dev));
    if (fused_conv_plan(cl, prec, dev.lds_full / (i64)elt(prec), sp, spread_cols(batch, dev.cus))) {
      fused = true;
      big_lds = 2 * (i64)sp.f.C * sp.f.Rp * (i64)elt(prec) > dev.lds_full;
      RC_TRY(upload_complex(stw, sp.tw.empty() ? std::vector<cld>(1, cld(1, 0)) : sp.tw, prec));
      // the fused kernel writes X[p], p = 1..q, in order: oinv[p - 1] is the
      // k with iperm[k] = p
      std::vector<int> oi(q);
      for (i64 k = 0; k < q; k++) oi[iph[k] - 1] = (int)k;
      RC_TRY(oinv.upload(oi.data(), oi.size() * sizeof(int)));
      // and reads x[i], i = 1..q, in order: dlog[i - 1] is the k with
      // perm[k] = i (the discrete logarithm of i to the base g)
      std::vector<int> dl(q);
      for (i64 k = 0; k < q; k++) dl[ph[k] - 1] = (int)k;
      RC_TRY(dlog.upload(dl.data(), dl.size() * sizeof(int)));
    } else {
      cb = chunk_lines(2 * cl * (i64)elt(prec), batch);
      RC_TRY(chain.init(cl, cb, prec, false, force_cols));
      RC_TRY(chain.upload_spec(b, spec));
    }
    return GOFFT_OK;
  }
  i64 cb = 1;  // lines per chunk (multi-kernel path)

  size_t ws_bytes() const override { return fused ? 0 : chain.ws_bytes(); }
  bool whole_cu() const override { return fused && big_lds; }
  bool queueable() const override { return true; }
  int launches() const override {
    return fused ? 1 : (int)(chain.launches() * ((batch + cb - 1) / cb));
  }
  void work(Work& w, i64 lines) const override {
    if (lines <= 0) return;
    if (fused) w.add(GOFFT_KERNEL_CONV, 1, 2 * n * lines * (i64)elt(prec));
    else chain.work(w, lines, q, n);
  }
  std::string describe() const override {
    std::string s = std::string(tuned ? "tuned " : "") + "rader n=" + std::to_string(n) + " lines=" + std::to_string(batch) +
                    (cyclic ? " cyclic" : " linear") + " cl=" + std::to_string(cl);
    return s + (fused ? " fused " + describe_sub(cl, sp) : " " + chain.describe());
  }

  int run(const void* in, LineIO li, void* out, LineIO lo, int ci, int co, double scale,
          const Ctx& c, i64 nl) const override {
    const i64 lines = nl < 0 ? batch : nl;
    if (lines == 0) return GOFFT_OK;
    if (fused) {
      ConvArgs a{};
      a.stw = stw.p;
      a.spec = spec.p;
      a.perm = (const int*)perm.p;
      a.dlog = (const int*)dlog.p;
      a.oinv = (const int*)oinv.p;
      a.in_stride = li.stride;
      a.out_stride = lo.stride;
      a.n = (int)n;
      a.kind = cyclic ? 0 : 1;
      a.conj_in = ci;
      a.conj_out = co;
      a.scale = scale;
      a.f = sp.f;
      a.in = in;
      a.out = out;
      a.in_off = li.off;
      a.out_off = lo.off;
      a.nlines = lines;
      HIP_TRY(issue_conv(prec, sp.maxp, a, sp.threads, c.s));
      return GOFFT_OK;
    }
    ConvChain::Hooks h;
    h.in_hook = 2;
    h.out_hook = 2;
    h.nvalid = q;
    h.conv_off = cyclic ? 0 : q;
    h.perm = (const int*)perm.p;
    h.iperm = (const int*)iperm.p;
    h.spec = spec.p;
    const size_t eb = elt(prec);
    for (i64 b0 = 0; b0 < lines; b0 += cb) {
      const i64 nb = std::min<i64>(cb, lines - b0);
      RC_TRY(chain.run(h, line_ptr(in, li, b0, eb), line_sub(li, b0), line_ptr(out, lo, b0, eb),
                       line_sub(lo, b0), ci, co, scale, c.ws, c.s, nb));
    }
    return GOFFT_OK;
  }
};

// Bluestein's chirp-z (bluestein.go:17-95): x*w zero-padded to a power of two
// m >= 2n-1, convolution with the conj-chirp kernel spectrum, times w.
class BluesteinEngine : public Engine {
 public:
  i64 m = 0;
  // Reference-twiddle mode (plain complex128): the convolution transforms
  // use the reference's recurrence twiddles and its forward/inverse kernel
  // spectra, so results round the way the reference's do.
  bool faithful = false;
  DevBuf chirp, spec, chirpI, specI;
  std::vector<i64> jj;  // chirp phase indices j^2 mod 2n (the tables were built from these)
  bool fused = false;
  bool big_lds = false;  // its workgroup takes more than half a CU's LDS
  SubPlan sp;
  DevBuf stw;
  ConvChain chain;

  int init(unsigned flags) {
    faithful = prec == 0 && (flags & GOFFT_FLAG_REFERENCE_TWIDDLES) && !(flags & GOFFT_FLAG_ACCURATE_TWIDDLES);
    // The reference pads to a power of two (bluestein.go:29-32); that length
    // is kept in reference mode, otherwise any smooth m >= 2n-1 serves.
    if (faithful) {
      m = 1;
      while (m < 2 * n - 1) m <<= 1;
    } else {
      m = force_m > 0 ? force_m : choose_conv_len(2 * n - 1, prec);
      if (m < 2 * n - 1)
        return fail(GOFFT_ERR_UNSUPPORTED, "no convolution length >= " + std::to_string(2 * n - 1) +
                                               " for Bluestein n=" + std::to_string(n));
    }
    std::vector<cld> hF, hI;  // kernel spectra (1/m folded in), natural order
    jj = chirp_index(n);
    if (faithful) {
      std::vector<dcx> wF, wI, bF, bI;
      reference_bluestein_tables(n, m, jj, wF, wI, bF, bI);
      const double invm = 1.0 / (double)m;  // exact: m is a power of two
      RC_TRY(upload_dcx(chirp, wF, prec));
      RC_TRY(upload_dcx(chirpI, wI, prec));
      hF.resize(m);
      hI.resize(m);
      for (i64 i = 0; i < m; i++) {
        hF[i] = cld(bF[i].re * invm, bF[i].im * invm);
        hI[i] = cld(bI[i].re * invm, bI[i].im * invm);
      }
      RC_TRY(upload_complex(spec, hF, prec));
      RC_TRY(upload_complex(specI, hI, prec));
    } else {
      std::vector<cld> w(n), b(m, cld(0, 0));
      for (i64 j = 0; j < n; j++) w[j] = root(jj[j], 2 * n);  // exp(-pi i jj / n)
      b[0] = std::conj(w[0]);
      for (i64 j = 1; j < n; j++) {
        b[j] = std::conj(w[j]);
        b[m - j] = std::conj(w[j]);
      }
      host_fft(b);
      for (auto& v : b) v /= (long double)m;
      RC_TRY(upload_complex(chirp, w, prec));
      RC_TRY(upload_complex(spec, b, prec));
      hF = std::move(b);
    }
    DevInfo dev;
    RC_TRY(current_device(&dev));
    const i64 pts = dev.lds_full / (i64)elt(prec);  // k_conv: no static LDS
    int C = 0;
    if (faithful) {
      if (pitch(m) <= std::min<i64>(pts, soft_pts(prec))) {
        C = 1;
        const int cap = std::min(16, spread_cols(batch, dev.cus));
        while (C * 2 <= cap && (i64)C * 2 * pitch(m) <= std::min<i64>(pts, soft_pts(prec))) C *= 2;
        while (C >= 1 && !make_sub_r2((int)m, C, 1, sp)) C /= 2;
        if (C >= 1) {
          fused = true;
          RC_TRY(upload_dcx(stw, recurrence_table(m), prec));
          return GOFFT_OK;
        }
      }
      cb = chunk_lines(2 * m * (i64)elt(prec), batch);
      RC_TRY(chain.init(m, cb, prec, true));
      RC_TRY(chain.upload_spec(hF, spec));
      return chain.upload_spec(hI, specI);
    }
    if (fused_conv_plan(m, prec, pts, sp, spread_cols(batch, dev.cus))) {
      fused = true;
      big_lds = 2 * (i64)sp.f.C * sp.f.Rp * (i64)elt(prec) > dev.lds_full;
      RC_TRY(upload_complex(stw, sp.tw.empty() ? std::vector<cld>(1, cld(1, 0)) : sp.tw, prec));
    } else {
      cb = chunk_lines(2 * m * (i64)elt(prec), batch);
      RC_TRY(chain.init(m, cb, prec, false, force_cols));
      RC_TRY(chain.upload_spec(hF, spec));
    }
    return GOFFT_OK;
  }
  i64 cb = 1;  // lines per chunk (multi-kernel path)

  size_t ws_bytes() const override { return fused ? 0 : chain.ws_bytes(); }
  bool whole_cu() const override { return fused && big_lds; }
  bool queueable() const override { return true; }
  int launches() const override {
    return fused ? 1 : (int)(chain.launches() * ((batch + cb - 1) / cb));
  }
  void work(Work& w, i64 lines) const override {
    if (lines <= 0) return;
    if (fused) w.add(GOFFT_KERNEL_CONV, 1, 2 * n * lines * (i64)elt(prec));
    else chain.work(w, lines, n, n);
  }
  std::string describe() const override {
    std::string s = std::string(tuned ? "tuned " : "") + "bluestein n=" + std::to_string(n) + " lines=" + std::to_string(batch) +
                    (faithful ? " reference-twiddles" : "") + " m=" + std::to_string(m);
    return s + (fused ? " fused " + describe_sub(m, sp) : " " + chain.describe());
  }

  int run(const void* in, LineIO li, void* out, LineIO lo, int ci, int co, double scale,
          const Ctx& c, i64 nl) const override {
    const i64 lines = nl < 0 ? batch : nl;
    if (lines == 0) return GOFFT_OK;
    // An inverse (ci and co) in reference mode runs the reference's inverse
    // (bluestein.go:67-95 with the inverse chirp and spectrum) instead of
    // conjugating a forward transform.
    const bool ref_inverse = faithful && ci && co;
    const void* w = ref_inverse ? chirpI.p : chirp.p;
    const void* b = ref_inverse ? specI.p : spec.p;
    if (ref_inverse) ci = co = 0;
    if (fused) {
      ConvArgs a{};
      a.stw = stw.p;
      a.spec = b;
      a.chirp = w;
      a.in_stride = li.stride;
      a.out_stride = lo.stride;
      a.n = (int)n;
      a.kind = 2;
      a.conj_in = ci;
      a.conj_out = co;
      a.scale = scale;
      a.f = sp.f;
      a.in = in;
      a.out = out;
      a.in_off = li.off;
      a.out_off = lo.off;
      a.nlines = lines;
      HIP_TRY(issue_conv(prec, sp.maxp, a, sp.threads, c.s));
      return GOFFT_OK;
    }
    ConvChain::Hooks h;
    h.in_hook = 1;
    h.out_hook = 1;
    h.nvalid = n;
    h.w = w;
    h.spec = b;
    const size_t eb = elt(prec);
    for (i64 b0 = 0; b0 < lines; b0 += cb) {
      const i64 nb = std::min<i64>(cb, lines - b0);
      RC_TRY(chain.run(h, line_ptr(in, li, b0, eb), line_sub(li, b0), line_ptr(out, lo, b0, eb),
                       line_sub(lo, b0), ci, co, scale, c.ws, c.s, nb));
    }
    return GOFFT_OK;
  }
};

std::unique_ptr<Engine> make_engine_fixed(i64 n, i64 batch, int prec, unsigned flags, int* rc, i64 force_m,
                                          int force_cols) {
  *rc = GOFFT_OK;
  int r = route(n);
  if (r == 0) {
    auto e = std::make_unique<TrivialEngine>();
    e->prec = prec;
    e->n = n;
    e->batch = batch;
    *rc = e->init();
    return *rc == GOFFT_OK ? std::move(e) : nullptr;
  }
  if (r == 1 || r == 2) {
    auto e = std::make_unique<StockhamEngine>();
    e->prec = prec;
    e->n = n;
    e->batch = batch;
    e->force_m = force_m;
    e->force_cols = force_cols;
    *rc = e->init(flags);
    return *rc == GOFFT_OK ? std::move(e) : nullptr;
  }
  if (r == 3) {
    auto e = std::make_unique<RaderEngine>();
    e->prec = prec;
    e->n = n;
    e->batch = batch;
    e->force_m = force_m;
    e->force_cols = force_cols;
    *rc = e->init(flags);
    return *rc == GOFFT_OK ? std::move(e) : nullptr;
  }
  auto e = std::make_unique<BluesteinEngine>();
  e->prec = prec;
  e->n = n;
  e->batch = batch;
  e->force_m = force_m;
  e->force_cols = force_cols;
  *rc = e->init(flags);
  return *rc == GOFFT_OK ? std::move(e) : nullptr;
}

// ---------------------------------------------------------------------------
// Autotuner (GOFFT_FLAG_AUTOTUNE, or GOFFT_AUTOTUNE=1): where the planner has
// a choice that the device decides best, the candidates are built and timed
// on the current device, and the fastest kept. Choices: the convolution
// length of Rader's linear case and of Bluestein outside the reference-
// twiddle mode (the cost model's three best, the reference's bestConvLen
// question, plan.go:192-259), and the columns per workgroup of multi-pass
// transforms and convolution chains. Results are cached per (route, length,
// lines, precision) for the process, and in the file GOFFT_TUNE_CACHE names
// (one "route n lines prec m cols" line each) across processes.

struct TuneChoice {
  i64 m = 0;
  int cols = 0;
};
// (device architecture, route, length, lines, precision)
typedef std::tuple<std::string, int, i64, i64, int> TuneKey;
std::mutex g_tune_mu;
std::map<TuneKey, TuneChoice> g_tune;
bool g_tune_loaded = false;
// Cache file lines carry this tag, the device architecture and the library
// version; lines with another tag or version are ignored.
const char* const kTuneTag = "gofft-tune-v2";

// Columns per workgroup that the autotuner tries for multi-pass passes.
std::vector<int> tune_cols(int prec) {
  return prec == GOFFT_C64 ? std::vector<int>{16, 8, 32} : std::vector<int>{8, 4, 16};
}

// Whether a (cached) choice is one the autotuner could have made: columns
// from the candidate set, and for a convolution a 7-smooth length in the
// planner's window [lo, lo + ceil(0.4 lo)] (none for other routes).
bool tune_choice_valid(const TuneChoice& c, bool conv, i64 lo, int prec) {
  const std::vector<int> cols = tune_cols(prec);
  if (std::find(cols.begin(), cols.end(), c.cols) == cols.end()) return false;
  if (!conv) return c.m == 0;
  if (c.m < lo || c.m > lo + (lo * 4 + 9) / 10) return false;
  i64 x = c.m;
  for (int p : {2, 3, 5, 7})
    while (x % p == 0) x /= p;
  return x == 1;
}

bool autotune_on(unsigned flags) {
  if (flags & GOFFT_FLAG_AUTOTUNE) return true;
  const char* e = std::getenv("GOFFT_AUTOTUNE");
  return e && std::atoi(e) != 0;
}

void tune_load_locked() {
  if (g_tune_loaded) return;
  g_tune_loaded = true;
  const char* path = std::getenv("GOFFT_TUNE_CACHE");
  if (!path) return;
  FILE* f = std::fopen(path, "r");
  if (!f) return;
  char line[512];
  while (std::fgets(line, sizeof line, f)) {
    char tag[64], arch[64], ver[64];
    long long r, n, b, pr, m, c;
    if (std::sscanf(line, "%63s %63s %63s %lld %lld %lld %lld %lld %lld", tag, ver, arch, &r, &n, &b, &pr, &m,
                    &c) != 9)
      continue;
    if (std::strcmp(tag, kTuneTag) != 0 || std::strcmp(ver, kVersionTag) != 0) continue;
    // entries are checked against the planner's constraints when used
    g_tune[TuneKey(arch, (int)r, (i64)n, (i64)b, (int)pr)] = TuneChoice{(i64)m, (int)c};
  }
  std::fclose(f);
}

void tune_store_locked(const TuneKey& key, const TuneChoice& c) {
  g_tune[key] = c;
  const char* path = std::getenv("GOFFT_TUNE_CACHE");
  if (!path) return;
  if (FILE* f = std::fopen(path, "a")) {
    std::fprintf(f, "%s %s %s %d %lld %lld %d %lld %d\n", kTuneTag, kVersionTag, std::get<0>(key).c_str(),
                 std::get<1>(key), (long long)std::get<2>(key), (long long)std::get<3>(key), std::get<4>(key),
                 (long long)c.m, c.cols);
    std::fclose(f);
  }
}

// Average milliseconds of one run of e over `lines` lines of zeros.
int time_engine(const Engine& e, i64 lines, float* ms) {
  DevBuf in, out, ws;  // freed after the stream below is drained
  hipStream_t s = nullptr;
  HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct Cleanup {
    hipStream_t s;
    hipEvent_t a = nullptr, b = nullptr;
    ~Cleanup() {
      if (a) (void)hipEventDestroy(a);
      if (b) (void)hipEventDestroy(b);
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  } cl{s};
  const size_t bytes = (size_t)e.n * (size_t)lines * elt(e.prec);
  RC_TRY(in.alloc(bytes));
  RC_TRY(out.alloc(bytes));
  if (e.ws_bytes()) RC_TRY(ws.alloc(e.ws_bytes()));
  HIP_TRY(launch_fill(e.prec, in.p, 2 * (long long)e.n * lines, 0.0, s));
  Ctx c{s, (char*)ws.p};
  for (int i = 0; i < 2; i++) RC_TRY(e.run(in.p, LineIO{nullptr, e.n}, out.p, LineIO{nullptr, e.n}, 0, 0, 1.0, c, lines));
  HIP_TRY(hipEventCreate(&cl.a));
  HIP_TRY(hipEventCreate(&cl.b));
  const int reps = 5;
  HIP_TRY(hipEventRecord(cl.a, s));
  for (int i = 0; i < reps; i++)
    RC_TRY(e.run(in.p, LineIO{nullptr, e.n}, out.p, LineIO{nullptr, e.n}, 0, 0, 1.0, c, lines));
  HIP_TRY(hipEventRecord(cl.b, s));
  HIP_TRY(hipEventSynchronize(cl.b));
  float t = 0;
  HIP_TRY(hipEventElapsedTime(&t, cl.a, cl.b));
  *ms = t / reps;
  return GOFFT_OK;
}

std::unique_ptr<Engine> make_engine(i64 n, i64 batch, int prec, unsigned flags, int* rc) {
  const int r = route(n);
  const bool faithful_blue =
      r == 4 && prec == GOFFT_C128 && (flags & GOFFT_FLAG_REFERENCE_TWIDDLES) && !(flags & GOFFT_FLAG_ACCURATE_TWIDDLES);
  const bool conv = (r == 3 && !factors_are_small(n - 1)) || (r == 4 && !faithful_blue);
  if (!autotune_on(flags) || batch <= 0 || r == 0 || faithful_blue || (!conv && r != 1 && r != 2))
    return make_engine_fixed(n, batch, prec, flags, rc, 0, 0);
  DevInfo dev;
  if ((*rc = current_device(&dev)) != GOFFT_OK) return nullptr;
  const TuneKey key(dev.arch, r, n, batch, prec);
  // Candidates: (conv length, columns); columns only matter for multi-pass
  // passes and chains, so a single-pass candidate is tried once.
  std::vector<TuneChoice> cand;
  const std::vector<int> cols = tune_cols(prec);
  const i64 lo = r == 3 ? 2 * (n - 1) : 2 * n - 1;
  if (conv) {
    for (i64 m : conv_candidates(lo, prec, 3))
      for (int c : cols) cand.push_back(TuneChoice{m, c});
  } else {
    for (int c : cols) cand.push_back(TuneChoice{0, c});
  }
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_load_locked();
    auto it = g_tune.find(key);
    // A cached choice is used only if the planner could have made it: a
    // convolution length of the cost model's window, columns from the
    // candidate set. Anything else (a stale or edited file) is dropped.
    if (it != g_tune.end() && !tune_choice_valid(it->second, conv, lo, prec)) {
      g_tune.erase(it);
      it = g_tune.end();
    }
    if (it != g_tune.end()) {
      auto e = make_engine_fixed(n, batch, prec, flags, rc, it->second.m, it->second.cols);
      if (e) {
        e->tuned = true;
        return e;
      }
      g_tune.erase(key);  // not plannable here: plan afresh below
    }
  }
  // Time at most 512 MB of lines, or all of them.
  const i64 lines = std::max<i64>(1, std::min<i64>(batch, ((i64)512 << 20) / std::max<i64>(1, n * (i64)elt(prec))));
  std::unique_ptr<Engine> best;
  TuneChoice bestc;
  float bestms = 0;
  std::vector<std::string> seen;
  for (const TuneChoice& c : cand) {
    int rc2 = GOFFT_OK;
    auto e = make_engine_fixed(n, batch, prec, flags, &rc2, c.m, c.cols);
    if (!e) continue;
    const std::string d = e->describe();
    if (std::find(seen.begin(), seen.end(), d) != seen.end()) continue;  // same plan as a tried one
    seen.push_back(d);
    float ms = 0;
    if (time_engine(*e, lines, &ms) != GOFFT_OK) continue;
    if (!best || ms < bestms) {
      best = std::move(e);
      bestc = c;
      bestms = ms;
    }
  }
  if (!best) return make_engine_fixed(n, batch, prec, flags, rc, 0, 0);
  best->tuned = true;
  *rc = GOFFT_OK;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  tune_store_locked(key, bestc);
  return best;
}

// ---------------------------------------------------------------------------
// Real-input engine (realplan.go): even n packs pairs into an n/2-point complex
// transform plus an untangle; odd n promotes to a full complex transform.

// Columns per workgroup of the single-pass real kernels (k_real_fwd,
// k_real_inv). The choice of pick_cols (about 64 KB of LDS per workgroup)
// stands unless it leaves a CU under 60% of the waves the kernels' registers
// allow; then the column count with the most resident waves wins (fewest
// columns on ties). A SIMD holds 512 / (registers per thread, in granules
// of 8) waves, at most 8, and a CU four SIMDs; a CU also holds only as many
// workgroups as fit its LDS. Measured: complex128 frames of 512 points (126
// registers) ran 11-19% faster at one column (16 waves per CU) than at four
// (8 waves); complex64 (84 registers, 16 of 20 waves at eight columns) ran
// 5-13% slower at one column, so it keeps pick_cols' choice.
bool real_cols(int R, int cap, i64 pts, i64 budget, int prec, const DevInfo& dev, SubPlan& sp, int* Cout) {
  if (!pick_cols(R, cap, budget, sp, Cout)) return false;
  const int vg = real_kernel_vgprs(prec);
  if (vg <= 0) return true;
  const int wave_cap = 4 * std::min(8, 512 / ((vg + 7) / 8 * 8));
  auto waves = [&](const SubPlan& t, int C) {
    const i64 waves_wg = (t.threads + 63) / 64;
    const i64 wgs = std::max<i64>(1, (i64)dev.lds_full / ((i64)C * pitch(R) * (i64)elt(prec)));
    return (int)std::min<i64>(wave_cap, wgs * waves_wg);
  };
  if (waves(sp, *Cout) * 5 >= wave_cap * 3) return true;
  int bestW = -1;
  for (int C = 1; C <= cap && (i64)C * pitch(R) <= pts; C *= 2) {
    SubPlan t;
    if (!make_sub(R, C, t)) continue;
    const int w = waves(t, C);
    if (w > bestW) {
      bestW = w;
      sp = t;
      *Cout = C;
    }
  }
  return true;
}

class RealEngine {
 public:
  int prec = 0;
  i64 n = 0, batch = 0, m = 0;
  bool even = false, single = false;
  SubPlan sp;
  DevBuf stw, utw;
  // The fused untangle and pack's twiddles W_n^k (k = 0..m) as two FP64
  // tables: hi[h] = W_n^(h * 2^ushift), lo[l] = W_n^l (l < 2^ushift).
  DevBuf utw2;
  int ushift = 0;
  i64 unh = 0;  // entries of hi
  std::unique_ptr<Engine> half;  // m-point (even, multi) or n-point (odd) complex engine

  int init(unsigned flags) {
    even = n >= 2 && n % 2 == 0;
    m = n / 2;
    if (even) {
      std::vector<cld> u(m + 1);
      for (i64 k = 0; k <= m; k++) u[k] = root(k, n);
      RC_TRY(upload_complex(utw, u, prec));
      DevInfo dev;
      RC_TRY(current_device(&dev));
      const i64 pts = dev.lds_bytes / (i64)elt(prec);
      // complex64 takes about 17 KB of LDS per workgroup: two waves for
      // 256-, 512- and 1024-point halves (8, 4 and 2 lines of pitch 273,
      // 545 and 1089 points). Measured on the 2^26 Spectrogram (131071
      // frames of 1024), 4 frames per workgroup 0.304-0.305 ms against 2
      // frames 0.316-0.317 ms, 8 frames 0.312-0.324 ms and 16 frames
      // 0.318-0.319 ms (with the pair stores: 4 frames 0.290 ms, 2 frames
      // 0.291-0.300 ms). complex128 keeps the larger budget (real_cols then
      // picks by resident waves).
      const i64 soft = prec == GOFFT_C64 ? 2200 : soft_pts(prec);
      const i64 budget = std::min<i64>(pts, std::max<i64>(pitch(m), soft));
      int C = 0;
      const int cap = (int)std::min<i64>(kMaxCols, std::max<i64>(1, batch));
      if (pitch(m) <= pts && factors_are_small(m) &&
          real_cols((int)m, cap, pts, budget, prec, dev, sp, &C)) {
        single = true;
        RC_TRY(upload_complex(stw, sp.tw.empty() ? std::vector<cld>(1, cld(1, 0)) : sp.tw, prec));
        if (prec == GOFFT_C64 && kRealTab) RC_TRY(plan_tab(dev, pts));
        return GOFFT_OK;
      }
      int rc = 0;
      half = make_real_half(m, 1, flags, &rc);
      if (!half) return rc;
      StockhamEngine* se = dynamic_cast<StockhamEngine*>(half.get());
      fused_last = se && se->prepare_untangle();
      // Two-pass halves pair different passes in the two directions, so
      // the inverse gets its own (pass order and columns chosen for the
      // pack in its first pass).
      if (se && se->two_pass) {
        half_inv = make_real_half(m, 2, flags, &rc);
        if (!half_inv) return rc;
      }
      StockhamEngine* si = dynamic_cast<StockhamEngine*>(inv_half());
      fused_first = si && si->prepare_pack();
      if (fused_last || fused_first) {
        // Read from two tables of about sqrt(m) entries instead of the m + 1
        // of utw: for PSD 2^26 that is 268 MB (complex64) of HBM reads less
        // per call.
        ushift = (ilog2((int)std::min<i64>(m + 1, 1 << 30)) + 1) / 2;
        const i64 S = 1ll << ushift;
        unh = m / S + 1;
        std::vector<cld> t;
        for (i64 h = 0; h < unh; h++) t.push_back(root(h * S, n));
        for (i64 l = 0; l < S; l++) t.push_back(root(l, n));
        RC_TRY(upload_complex(utw2, t, GOFFT_C128));
        // with both directions fused nothing reads the (m+1)-entry table
        if (fused_last && fused_first) RC_TRY(utw.alloc(0));
      }
      return GOFFT_OK;
    }
    int rc = 0;
    half = make_engine(n, batch, prec, flags, &rc);
    return half ? GOFFT_OK : rc;
  }
  // Table forms of the complex64 single-pass real kernels (kernels.hip
  // k_real_fwd_tab, k_real_inv_tab): 8 points per thread, radices 2, 4 and
  // 8, the twiddles (and the forward's window) in LDS, tab_wgs[dir]
  // workgroups looping over the line groups with the next group's loads in
  // flight. dir 0 forward, 1 inverse; tab_wgs 0 keeps the plain kernel.
  SubPlan sp_tab;
  DevBuf stw_tab;
  i64 tab_wgs[2] = {0, 0};

  int plan_tab(const DevInfo& dev, i64 pts) {
    if (m < 64 || (m & (m - 1)) || batch < 2 * (i64)dev.cus) return GOFFT_OK;
    for (int dir = 0; dir < 2; dir++) {
      const int vg = real_tab_vgprs(dir);
      if (vg <= 0) return GOFFT_OK;
      const int wave_cap = 4 * std::min(8, 512 / ((vg + 7) / 8 * 8));
      // LDS points of the tables beside the lines: forward stage + untangle
      // twiddles + window, inverse stage + pack twiddles
      auto tab_pts = [&](const SubPlan& t) {
        return (i64)std::max<size_t>(t.tw.size(), 1) + (dir == 0 ? 2 * m + 1 : m / 2 + 1);
      };
      auto waves = [&](const SubPlan& t) {
        const i64 wg_pts = (i64)t.f.C * pitch(m) + tab_pts(t);
        if (wg_pts > pts) return 0;
        const int waves_wg = (t.threads + 63) / 64;
        const i64 by_lds = dev.lds_full / (wg_pts * (i64)elt(prec));
        return (int)std::min<i64>(by_lds, wave_cap / waves_wg) * waves_wg;
      };
      if (dir == 0) {
        // The column count, chosen for the forward: the most waves resident
        // among the counts whose line groups fill the resident grid at least
        // once (smaller batches keep the plain kernel); ties go to more
        // lines per wave.
        int best = 0;
        for (int C = 1; C <= kMaxCols && C <= batch; C *= 2) {
          SubPlan t;
          if (!make_sub((int)m, C, t, 0, 4, 8, 8) || t.threads > 512) continue;
          const int w = waves(t);
          const i64 grid = (i64)std::max(dev.cus, 1) * (w / ((t.threads + 63) / 64));
          if (!w || (batch + C - 1) / C < grid) continue;
          if (w >= best) {
            best = w;
            sp_tab = t;
          }
        }
        if (!best) return GOFFT_OK;
      }
      const int w = waves(sp_tab);
      if (!w) continue;
      // The tables take LDS from the lines (about 6 KB per 1024 points of
      // half length in the forward): keep the plain kernel when it keeps
      // more waves resident.
      const int vr = real_kernel_vgprs(prec, dir);
      if (vr > 0) {
        const int cap_r = 4 * std::min(8, 512 / ((vr + 7) / 8 * 8));
        const int wr = (sp.threads + 63) / 64;
        const i64 by = dev.lds_full / ((i64)sp.f.C * pitch(m) * (i64)elt(prec));
        if (w < (int)std::min<i64>(by, cap_r / wr) * wr) continue;
      }
      tab_wgs[dir] = (i64)std::max(dev.cus, 1) * (w / ((sp_tab.threads + 63) / 64));
    }
    if (tab_wgs[0] || tab_wgs[1])
      RC_TRY(upload_complex(stw_tab, sp_tab.tw.empty() ? std::vector<cld>(1, cld(1, 0)) : sp_tab.tw, prec));
    return GOFFT_OK;
  }

  bool fused_last = false;   // forward untangle done by the half transform's last pass
  bool fused_first = false;  // inverse pack done by the half transform's first pass
  std::unique_ptr<Engine> half_inv;  // the inverse's own half transform (two-pass halves)
  Engine* inv_half() const { return half_inv ? half_inv.get() : half.get(); }

  // The m-point half transform, planned with `pair` (StockhamEngine::pair_req)
  // when it is a Stockham transform the autotuner does not plan.
  std::unique_ptr<Engine> make_real_half(i64 len, int pair, unsigned flags, int* rc) const {
    const int r = route(len);
    if ((r == 1 || r == 2) && !autotune_on(flags)) {
      auto e = std::make_unique<StockhamEngine>();
      e->prec = prec;
      e->n = len;
      e->batch = batch;
      e->pair_req = pair;
      *rc = e->init(flags);
      return *rc == GOFFT_OK ? std::move(e) : nullptr;
    }
    return make_engine(len, batch, prec, flags, rc);
  }

  size_t ws_bytes() const {
    if (single) return 0;
    size_t hw = half ? half->ws_bytes() : 0;
    if (half_inv) hw = std::max(hw, half_inv->ws_bytes());
    // (the inverse still packs into m points, so the workspace keeps them)
    if (even) return (size_t)m * batch * elt(prec) + hw;
    return 2 * (size_t)n * batch * elt(prec) + hw;
  }
  // Launches of a forward run (the inverse always adds its pack launch).
  int launches() const {
    if (single) return 1;
    return half->launches() + (even ? (fused_last ? 0 : 1) : 2);
  }
  std::string describe() const {
    std::string s = "real n=" + std::to_string(n) + " lines=" + std::to_string(batch);
    if (single)
      return s + " fused " + describe_sub(m, sp) +
             (tab_wgs[0] ? " forward table form " + describe_sub(m, sp_tab) + " x" + std::to_string(tab_wgs[0]) : "") +
             (tab_wgs[1] ? " inverse table form " + describe_sub(m, sp_tab) + " x" + std::to_string(tab_wgs[1]) : "");
    if (even && fused_last) {
      const StockhamEngine* se = static_cast<const StockhamEngine*>(half.get());
      s += " packed: " + half->describe() + "; untangle in the last pass, paired " +
           describe_sub(se->passes.back().R, se->pair_sp);
      if (half_inv) {
        const StockhamEngine* si = static_cast<const StockhamEngine*>(half_inv.get());
        s += "; inverse: " + half_inv->describe();
        if (fused_first) s += "; pack in the first pass, paired " + describe_sub(si->passes.front().R, si->pair_first_sp);
      }
      return s;
    }
    return s + (even ? " packed: " : " promoted: ") + half->describe();
  }
  // One forward (mode 0 bins, 1 power) or inverse (bins read per line) run.
  void work(Work& w, int inverse, int mode, i64 bins) const {
    if (batch == 0 || n <= 0) return;
    const i64 e = (i64)elt(prec), nb = n / 2 + 1;
    const i64 rb = n * e / 2;  // bytes of one real line
    const i64 sb = mode ? nb * e / 2 : nb * e;
    const i64 ib = std::min<i64>(std::max<i64>(bins, 0), nb) * e;
    if (single) {
      w.add(GOFFT_KERNEL_REAL, 1, batch * (inverse ? ib + rb : rb + sb));
      return;
    }
    (inverse ? inv_half() : half.get())->work(w, batch);
    if (even) {
      if (fused_last && !inverse) {
        w.add(GOFFT_KERNEL_PASS, 0, batch * (sb - m * e));  // the last pass writes the bins
        return;
      }
      if (fused_first && inverse) {
        w.add(GOFFT_KERNEL_PASS, 0, batch * (ib - m * e));  // the first pass reads the bins
        return;
      }
      w.add(GOFFT_KERNEL_AUX, 1, batch * (inverse ? ib + m * e : m * e + sb));
      return;
    }
    w.add(GOFFT_KERNEL_AUX, 1, batch * (inverse ? ib + n * e : rb + n * e));
    w.add(GOFFT_KERNEL_AUX, 1, batch * (inverse ? n * e + rb : nb * e + sb));
  }

  // mode 0: complex bins (stride ostride per line), mode 1: PSD with psd_scale,
  // mode 2: complex bins times w_edge (DC, Nyquist) or w_mid (the others).
  // (the first nl lines when nl >= 0, every line otherwise)
  int forward(const void* in, i64 istride, void* out, i64 ostride, int mode, double psd_scale,
              const void* window, const Ctx& c, double w_edge = 1.0, double w_mid = 1.0, i64 nl = -1) const {
    const i64 lines = nl < 0 ? batch : nl;
    if (lines == 0 || n == 0) return GOFFT_OK;
    if (single) {
      RealFwdArgs a{};
      a.in = in;
      a.out = out;
      a.stw = stw.p;
      a.utw = utw.p;
      a.window = window;
      a.nlines = lines;
      a.in_stride = istride;
      a.out_stride = ostride;
      a.mode = mode;
      a.psd_scale = psd_scale;
      a.w_edge = w_edge;
      a.w_mid = w_mid;
      a.vec_in = istride % 2 == 0 && (uintptr_t)in % (2 * relt(prec)) == 0;
      // whole rows of n samples on the 128-byte grid, not shared with the
      // next line (frames that do not overlap): non-temporal loads
      if (a.vec_in && istride >= n && ((size_t)istride * relt(prec)) % 128 == 0 && (uintptr_t)in % 128 == 0 &&
          ((size_t)n * relt(prec)) % 128 == 0)
        a.nt = nt_bits(true, 16, false, 0);
      a.nbd = fastdiv_make((uint32_t)(sp.f.R + 1));
      a.f = sp.f;
      if (tab_wgs[0]) {
        a.lds_tab = 1;
        a.stw = stw_tab.p;
        a.stw_n = (int)std::max<size_t>(sp_tab.tw.size(), 1);
        a.wgs = tab_wgs[0];
        a.f = sp_tab.f;
        HIP_TRY(launch_real_fwd(prec, 8, a, sp_tab.threads, c.s));
        return GOFFT_OK;
      }
      HIP_TRY(launch_real_fwd(prec, kMaxP, a, sp.threads, c.s));
      return GOFFT_OK;
    }
    if (window) return fail(GOFFT_ERR_UNSUPPORTED, "window needs the single-pass real kernel");
    if (even) {
      if (istride % 2) return fail(GOFFT_ERR_UNSUPPORTED, "odd real line stride");
      char* Z = c.ws;
      Ctx ic{c.s, Z + (size_t)m * batch * elt(prec)};
      if (fused_last) {
        RealLast rl;
        rl.utw = utw.p;
        rl.uhi = (const double2*)utw2.p;
        rl.ulo = rl.uhi ? rl.uhi + unh : nullptr;
        rl.ushift = ushift;
        rl.ostride = ostride;
        rl.mode = mode;
        rl.psd = psd_scale;
        rl.w_edge = w_edge;
        rl.w_mid = w_mid;
        return static_cast<const StockhamEngine*>(half.get())->run_real(in, istride / 2, out, rl, ic, lines);
      }
      RC_TRY(half->run(in, LineIO{nullptr, istride / 2}, Z, LineIO{nullptr, m}, 0, 0, 1.0, ic, lines));
      HIP_TRY(launch_untangle(prec, Z, out, utw.p, m, lines, m, ostride, mode, psd_scale, w_edge, w_mid, c.s));
      return GOFFT_OK;
    }
    char* A = c.ws;
    char* B = c.ws + (size_t)n * batch * elt(prec);
    Ctx ic{c.s, B + (size_t)n * batch * elt(prec)};
    HIP_TRY(launch_promote(prec, in, A, n, lines, istride, n, c.s));
    RC_TRY(half->run(A, LineIO{nullptr, n}, B, LineIO{nullptr, n}, 0, 0, 1.0, ic, lines));
    if (mode == 0)
      HIP_TRY(launch_copy_lines(prec, B, out, n / 2 + 1, lines, n, ostride, c.s));
    else if (mode == 2)
      HIP_TRY(launch_weight_bins(prec, B, out, n / 2 + 1, n, lines, n, ostride, w_edge, w_mid, c.s));
    else
      HIP_TRY(launch_power(prec, B, out, n / 2 + 1, n, lines, n, ostride, psd_scale, c.s));
    return GOFFT_OK;
  }

  // IRFFT: `bins` bins per input line (stride istride), n reals per output
  // line; the output times gain, the bins other than DC and Nyquist times w_mid.
  int inverse(const void* in, i64 istride, i64 bins, void* out, i64 ostride, const Ctx& c, double gain = 1.0,
              double w_mid = 1.0, i64 nl = -1) const {
    const i64 lines = nl < 0 ? batch : nl;
    if (lines == 0 || n <= 0) return GOFFT_OK;
    if (single) {
      RealInvArgs a{};
      a.in = in;
      a.out = out;
      a.stw = stw.p;
      a.utw = utw.p;
      a.nlines = lines;
      a.in_stride = istride;
      a.in_bins = bins;
      a.out_stride = ostride;
      a.scale = gain / (double)m;
      a.w_mid = w_mid;
      a.halfd = fastdiv_make((uint32_t)(sp.f.R / 2 + 1));
      a.f = sp.f;
      // whole rows of n samples on the 128-byte grid, one workgroup each:
      // written through non-temporal stores
      if (ostride >= n && ((size_t)ostride * relt(prec)) % 128 == 0 && (uintptr_t)out % 128 == 0 &&
          ((size_t)n * relt(prec)) % 128 == 0)
        a.nt = nt_bits(false, 0, true, 32);
      if (tab_wgs[1]) {
        a.lds_tab = 1;
        a.stw = stw_tab.p;
        a.stw_n = (int)std::max<size_t>(sp_tab.tw.size(), 1);
        a.wgs = tab_wgs[1];
        a.f = sp_tab.f;
        HIP_TRY(launch_real_inv(prec, 8, a, sp_tab.threads, c.s));
        return GOFFT_OK;
      }
      HIP_TRY(launch_real_inv(prec, kMaxP, a, sp.threads, c.s));
      return GOFFT_OK;
    }
    if (even) {
      if (ostride % 2) return fail(GOFFT_ERR_UNSUPPORTED, "odd real line stride");
      char* Z = c.ws;
      Ctx ic{c.s, Z + (size_t)m * batch * elt(prec)};
      if (fused_first) {
        RealFirst rf;
        rf.utw = utw.p;
        rf.uhi = (const double2*)utw2.p;
        rf.ulo = rf.uhi ? rf.uhi + unh : nullptr;
        rf.ushift = ushift;
        rf.bins = std::max<i64>(0, std::min<i64>(bins, m + 1));
        rf.w_mid = w_mid;
        return static_cast<const StockhamEngine*>(inv_half())
            ->run_real_inv(in, istride, out, ostride / 2, gain / (double)m, rf, ic, lines);
      }
      HIP_TRY(launch_herm_pack(prec, in, Z, utw.p, m, lines, istride, bins, m, w_mid, c.s));
      RC_TRY(inv_half()->run(Z, LineIO{nullptr, m}, out, LineIO{nullptr, ostride / 2}, 1, 1, gain / (double)m, ic,
                             lines));
      return GOFFT_OK;
    }
    char* A = c.ws;
    char* B = c.ws + (size_t)n * batch * elt(prec);
    Ctx ic{c.s, B + (size_t)n * batch * elt(prec)};
    HIP_TRY(launch_herm_full(prec, in, A, n, lines, istride, bins, n, w_mid, c.s));
    RC_TRY(half->run(A, LineIO{nullptr, n}, B, LineIO{nullptr, n}, 1, 1, gain / (double)n, ic, lines));
    HIP_TRY(launch_take_real(prec, B, out, n, lines, n, ostride, 1.0, c.s));
    return GOFFT_OK;
  }
};

struct NDAxis {
  i64 n = 1, st = 1, outer = 1;
  std::unique_ptr<Engine> eng;  // strided Stockham, or contiguous engine
  bool gather = false;          // non-smooth strided axis: copy lines out and back
};

// ---------------------------------------------------------------------------
// Per-stream workspace.

struct WorkspacePool {
  std::mutex mu;
  std::map<hipStream_t, std::unique_ptr<DevBuf>> bufs;
  int get(hipStream_t s, size_t bytes, char** out) {
    std::lock_guard<std::mutex> lk(mu);
    auto& b = bufs[s];
    if (!b) b = std::make_unique<DevBuf>();
    if (b->bytes < bytes) RC_TRY(b->alloc(bytes));
    *out = (char*)b->p;
    return GOFFT_OK;
  }
};

// Pinned host memory (hipHostMalloc), for the host-staged calls.
struct PinnedBuf {
  void* p = nullptr;
  size_t bytes = 0;
  PinnedBuf() {}
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
  int alloc(size_t b) {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
    if (b == 0) return GOFFT_OK;
    hipError_t e = hipHostMalloc(&p, b, hipHostMallocDefault);
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      return fail(GOFFT_ERR_OOM, "hipHostMalloc of " + std::to_string(b) + " bytes failed");
    }
    bytes = b;
    return GOFFT_OK;
  }
};

// One host-staged call in flight: its own stream, device copies of the
// call's input and output, and two pinned staging buffers through which
// chunks of the caller's (pageable) memory are copied while the previous
// chunk's DMA runs. Sessions are made on first use and reused; a plan keeps
// as many as calls have overlapped, so concurrent host calls on one plan run
// on separate streams and wait for their own stream only.
struct HostSession {
  hipStream_t s = nullptr;
  DevBuf din, dout;
  PinnedBuf stage[2];
  hipEvent_t done[2] = {nullptr, nullptr};  // the DMA of each staging buffer's last chunk
  ~HostSession() {
    if (s) (void)hipStreamSynchronize(s);
    for (hipEvent_t e : done)
      if (e) (void)hipEventDestroy(e);
    if (s) (void)hipStreamDestroy(s);
  }
};

struct HostPool {
  std::mutex mu;
  std::vector<std::unique_ptr<HostSession>> all;
  std::vector<HostSession*> idle;
  int acquire(HostSession** out) {
    {
      std::lock_guard<std::mutex> lk(mu);
      if (!idle.empty()) {
        *out = idle.back();
        idle.pop_back();
        return GOFFT_OK;
      }
    }
    auto h = std::make_unique<HostSession>();
    HIP_TRY(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
    for (hipEvent_t& e : h->done) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    std::lock_guard<std::mutex> lk(mu);
    *out = h.get();
    all.push_back(std::move(h));
    return GOFFT_OK;
  }
  void release(HostSession* h) {
    std::lock_guard<std::mutex> lk(mu);
    idle.push_back(h);
  }
};

}  // namespace

// ---------------------------------------------------------------------------
// Public plan object.

enum PlanKind { K_C2C = 0, K_MANY, K_R2C, K_ND_C2C, K_ND_R2C, K_PSD, K_SPEC };

struct gofft_plan {
  PlanKind kind = K_C2C;
  int prec = 0;
  unsigned flags = 0;
  i64 n = 0, batch = 0;
  // C2C
  std::unique_ptr<Engine> eng;
  // MANY
  struct Group {
    i64 len = 0;
    std::unique_ptr<Engine> eng;
    DevBuf off;
  };
  std::vector<std::unique_ptr<Group>> groups;
  i64 total_elems = 0;
  // R2C / PSD / SPEC
  std::unique_ptr<RealEngine> real;
  double d = 1.0;
  // ND
  std::vector<i64> shape;
  std::vector<NDAxis> axes;
  i64 nd_total = 1;
  std::unique_ptr<RealEngine> nd_row;   // R2C rows
  i64 rows = 0, cols = 0, rcols = 0;
  // SPEC
  i64 seg = 0, step = 0, frames = 0, sig_len = 0;
  DevBuf window;
  bool spec_gather = false;
  size_t ws_need = 0;
  size_t ws_engines = 0;  // ND_R2C: the part of ws_need the row and column engines use
  mutable WorkspacePool pool;
  mutable WorkspacePool loss_pool;  // gofft_filter2_loss_grad: per-stream chunk buffers
  mutable HostPool host;  // sessions of the host-staged calls
  // Mixed-length plans: streams (default priority) on which length groups run
  // beside the caller's stream, joined back by events. lane_of[g] is group
  // g's lane, 0 being the caller's stream.
  std::vector<hipStream_t> lanes;
  hipEvent_t fork_ev = nullptr;
  std::vector<hipEvent_t> join_ev;
  std::vector<int> lane_of;
  std::vector<size_t> order;  // enqueue order of the groups
  mutable std::mutex lane_mu;  // one fork/join enqueue at a time
  // Work queue (GOFFT_FLAG_QUEUE): the groups it takes run as the stages of
  // one persistent kernel (kernels.hip k_queue) on the caller's stream; the
  // others run on the side lanes beside it. The workspace starts with the
  // queue's counters and descriptor block, then each queued group's own
  // workspace.
  struct QueueVariant {
    DevBuf tmpl, table;  // descriptor template block, its QueueDesc table
    int ndesc = 0;
    QueueArgs args{};    // stage table; pointers set per exec
  };
  struct Queue {
    bool on = false;
    bool check = false;  // GOFFT_QUEUE_CHECK: each exec waits and reads the error flag
    int threads = 0;
    size_t lds = 0;
    long long grid = 0;
    size_t off_desc = 0, off_groups = 0, ctr_words = 0;
    std::vector<char> in_queue;  // per group
    QueueVariant var[4];         // index: conjugation + 2 * (1/n scaling)
    std::string text;
  } q;

  int ws(hipStream_t s, char** out) const {
    if (ws_need == 0) {
      *out = nullptr;
      return GOFFT_OK;
    }
    return pool.get(s, ws_need, out);
  }

  ~gofft_plan() {
    host.all.clear();  // drains each session's stream
    for (hipStream_t l : lanes) (void)hipStreamDestroy(l);
    for (hipEvent_t e : join_ev) (void)hipEventDestroy(e);
    if (fork_ev) (void)hipEventDestroy(fork_ev);
  }
};

namespace {

// Lanes of a mixed-length plan (GOFFT_LANES streams, default 5, at most one
// per group): each group, largest data first, goes to the lane with the least
// data so far. Lane workspaces are allocated here, so that an exec enqueues
// no allocation.
int assign_lanes(gofft_plan& p) {
  int want = 5;
  if (const char* e = std::getenv("GOFFT_LANES")) want = std::max(1, std::min(16, std::atoi(e)));
  // With a work queue on the caller's stream, the groups it does not take
  // share the side lanes.
  size_t lane_groups = p.groups.size();
  if (p.q.on) {
    lane_groups = 0;
    for (size_t g = 0; g < p.groups.size(); g++) lane_groups += p.q.in_queue[g] ? 0 : 1;
    if (lane_groups > 0) lane_groups++;
  }
  const int n = (int)std::min<size_t>((size_t)want, lane_groups);
  p.lane_of.assign(p.groups.size(), 0);
  // Enqueue order: groups whose workgroups need more than half a CU's LDS
  // first, while every CU is still empty; the others after, largest first.
  // Measured on the 4096-transform mixed-length batch (Rader 10007's fused
  // convolution first): FP32 0.785 -> 0.773 ms.
  p.order.clear();
  for (int pass = 0; pass < 2; pass++)
    for (size_t g = 0; g < p.groups.size(); g++)
      if (p.groups[g]->eng->whole_cu() == (pass == 0)) p.order.push_back(g);
  if (n <= 1) return GOFFT_OK;
  std::vector<double> load(n, 0.0);
  if (p.q.on) load[0] = std::numeric_limits<double>::infinity();  // the queue's stream
  for (size_t g = 0; g < p.groups.size(); g++) {
    if (p.q.on && p.q.in_queue[g]) continue;
    const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    p.lane_of[g] = k;
    load[k] += (double)p.groups[g]->len * (double)p.groups[g]->eng->batch;
  }
  p.lanes.resize(n - 1, nullptr);
  p.join_ev.resize(n - 1, nullptr);
  for (int i = 0; i < n - 1; i++) {
    HIP_TRY(hipStreamCreateWithFlags(&p.lanes[i], hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&p.join_ev[i], hipEventDisableTiming));
    char* w = nullptr;
    RC_TRY(p.ws(p.lanes[i], &w));
  }
  HIP_TRY(hipEventCreateWithFlags(&p.fork_ev, hipEventDisableTiming));
  return GOFFT_OK;
}

// ---------------------------------------------------------------------------
// Work queue of a mixed-length plan (GOFFT_FLAG_QUEUE, or GOFFT_QUEUE=1).
// Each queueable group's run is recorded once per variant (conjugation and
// 1/n scaling) with stand-in data pointers; every recorded launch becomes a
// stage of the queue, re-planned to the queue's workgroup size (columns or
// lines per workgroup = threads / threads per column), and its pointer
// fields that hold a stand-in become relocations that the setup kernel
// resolves per exec. A group's stages run in order (each waits for the
// previous one); groups do not wait for each other.

// Stand-in bases: 1 source, 2 destination, 3 the queued groups' workspace.
constexpr uintptr_t kFakeSpan = (uintptr_t)1 << 56;
const void* fake_base(int sel) { return (const void*)((uintptr_t)sel << 60); }
int fake_sel(const void* v, long long* add) {
  const uintptr_t x = (uintptr_t)v;
  for (int sel = 1; sel <= 3; sel++) {
    const uintptr_t b = (uintptr_t)sel << 60;
    if (x >= b && x < b + kFakeSpan) {
      *add = (long long)(x - b);
      return sel;
    }
  }
  return 0;
}

struct QStage {
  int group = 0, level = 0;
  int kind = 0, form = 0;
  i64 nitems = 0;
  size_t lds = 0;
  double work = 0;                 // data bytes of the stage (ordering only)
  std::vector<char> desc[4];       // descriptor bytes per variant (relocated fields zero)
  QueueDesc rel[4];                // relocations per variant (off and bytes set later)
};

i64 env_int(const char* name, i64 dflt) {
  const char* e = std::getenv(name);
  return e && *e ? std::atoll(e) : dflt;
}

// Re-plans one recorded launch for a workgroup of `threads` threads and
// records its descriptor; false if the queue kernel cannot run it.
bool queue_stage(int prec, int threads, const RecordedLaunch& r, QStage& st, int var) {
  const size_t e2 = elt(prec);
  auto relocate = [&](char* base, size_t bytes, std::initializer_list<size_t> fields, QueueDesc& qd) -> bool {
    qd.nreloc = 0;
    for (size_t f : fields) {
      void* v = nullptr;
      std::memcpy(&v, base + f, sizeof(v));
      long long add = 0;
      const int sel = fake_sel(v, &add);
      if (!sel) continue;
      if (qd.nreloc >= 4) return false;
      qd.field[qd.nreloc] = (int)f;
      qd.sel[qd.nreloc] = sel;
      qd.add[qd.nreloc] = add;
      qd.nreloc++;
      std::memset(base + f, 0, sizeof(v));
    }
    // no other 8-byte word may hold a stand-in pointer
    for (size_t o = 0; o + 8 <= bytes; o += 8) {
      void* v = nullptr;
      std::memcpy(&v, base + o, sizeof(v));
      long long add = 0;
      if (fake_sel(v, &add)) return false;
    }
    return true;
  };
  QueueDesc qd{};
  if (r.kind == 0) {
    PassArgs a = r.pa;
    const int form = pass_form(prec, a);
    if (form >= 11 || !((queue_pass_forms(prec) >> form) & 1u)) return false;
    if (r.threads != a.f.C * a.f.tpc || a.tw_mode != 0) return false;
    const int C = threads / a.f.tpc;
    if (C < 1 || C > kMaxCols || C * a.f.tpc != threads) return false;
    a.f.C = C;
    a.f.log2C = ilog2(C);
    int rp = a.f.Rp;
    if (a.mid_spec) {
      if (a.f2.tpc != a.f.tpc) return false;
      a.f2.C = C;
      a.f2.log2C = ilog2(C);
      rp = std::max(rp, a.f2.Rp);
    }
    a.xcd_group = 1;  // the queue hands out column groups in ticket order
    st.kind = 0;
    st.form = form;
    st.nitems = (a.ncols + C - 1) / C;
    st.lds = (size_t)C * (size_t)rp * e2;
    st.work = 2.0 * (double)a.ncols * (double)a.f.R * (double)e2;
    st.desc[var].resize(sizeof(PassArgs));
    std::memcpy(st.desc[var].data(), &a, sizeof(PassArgs));
    if (!relocate(st.desc[var].data(), sizeof(PassArgs),
                  {offsetof(PassArgs, in), offsetof(PassArgs, out), offsetof(PassArgs, x0buf), offsetof(PassArgs, dc)},
                  qd))
      return false;
  } else {
    ConvArgs a = r.ca;
    if (!queue_conv_form(prec, r.maxp) || r.threads != a.f.C * a.f.tpc) return false;
    const int C = threads / a.f.tpc;
    if (C < 1 || C > kMaxCols || C * a.f.tpc != threads) return false;
    a.f.C = C;
    a.f.log2C = ilog2(C);
    st.kind = 1;
    st.form = r.maxp;
    st.nitems = (a.nlines + C - 1) / C;
    st.lds = (size_t)C * (size_t)a.f.Rp * e2;
    st.work = 2.0 * (double)a.nlines * (double)a.f.R * (double)e2;
    st.desc[var].resize(sizeof(ConvArgs));
    std::memcpy(st.desc[var].data(), &a, sizeof(ConvArgs));
    if (!relocate(st.desc[var].data(), sizeof(ConvArgs), {offsetof(ConvArgs, in), offsetof(ConvArgs, out)}, qd))
      return false;
  }
  st.rel[var] = qd;
  return true;
}

// Builds the plan's work queue; leaves it off (the lanes run every group)
// when no group fits it. Sets p.q and may raise p.ws_need.
int build_queue(gofft_plan& p) {
  auto& Q = p.q;  // (fresh: called once, at plan creation)
  DevInfo dev;
  RC_TRY(current_device(&dev));
  const int prec = p.prec;
  // Workgroup size and workgroups per CU: complex64 1024 threads, one per CU
  // (every group of the mixed-length batches fits, the whole-CU fused
  // convolutions included); complex128 256 threads, two per CU (its kernel
  // holds 216 registers).
  const int threads = (int)env_int("GOFFT_QUEUE_THREADS", prec == 0 ? 256 : 1024);
  const int per_cu = (int)std::max<i64>(1, env_int("GOFFT_QUEUE_WGS", prec == 0 ? 2 : 1));
  // LDS of a workgroup: a pass stage's columns plus its per-column tables,
  // or a fused convolution's lines
  const size_t tables = queue_pass_tables(prec);
  const i64 budget = (i64)dev.lds_full / per_cu - (i64)queue_static_lds();
  if (budget <= 0) return GOFFT_OK;
  const size_t G = p.groups.size();
  Q.in_queue.assign(G, 0);
  std::vector<std::vector<QStage>> gst(G);
  size_t area = 0;  // queued groups' workspace
  for (size_t g = 0; g < G; g++) {
    const Engine& e = *p.groups[g]->eng;
    if (!e.queueable() || e.batch == 0) continue;
    const size_t off = area;
    bool ok = true;
    for (int var = 0; var < 4 && ok; var++) {
      const int c = var & 1;
      const bool scaled = var >= 2;
      std::vector<RecordedLaunch> rec;
      t_record = &rec;
      const i64* o = (const i64*)p.groups[g]->off.p;
      const int rc = e.run(fake_base(1), LineIO{o, 0}, (void*)fake_base(2), LineIO{o, 0}, c, c,
                           scaled ? 1.0 / (double)p.groups[g]->len : 1.0, Ctx{nullptr, (char*)fake_base(3) + off});
      t_record = nullptr;
      if (rc != GOFFT_OK || (int)rec.size() != e.launches() || rec.empty()) {
        ok = false;
        break;
      }
      if (var == 0) gst[g].resize(rec.size());
      if (gst[g].size() != rec.size()) ok = false;
      for (size_t k = 0; k < rec.size() && ok; k++) {
        QStage& st = gst[g][k];
        const int kind = st.kind, form = st.form;
        const i64 ni = st.nitems;
        ok = queue_stage(prec, threads, rec[k], st, var);
        if (ok && var > 0) ok = st.kind == kind && st.form == form && st.nitems == ni;
        if (ok) ok = (i64)(st.kind == 0 ? ((st.lds + 255) & ~(size_t)255) + tables : st.lds) <= budget;
        st.group = (int)g;
        st.level = (int)k;
      }
    }
    if (!ok) {
      gst[g].clear();
      continue;
    }
    Q.in_queue[g] = 1;
    area += (e.ws_bytes() + 255) & ~(size_t)255;
  }
  size_t nq = 0;
  for (size_t g = 0; g < G; g++) nq += gst[g].size();
  i64 nitems_all = 0;
  for (size_t g = 0; g < G; g++)
    for (const QStage& st : gst[g]) nitems_all += st.nitems;
  if (nq == 0 || nq > (size_t)kMaxQueueStages || nitems_all >= ((i64)1 << 30)) {
    Q.in_queue.assign(G, 0);
    return GOFFT_OK;
  }
  // Stage order. Multi-stage groups (chains, multi-pass transforms) take
  // their first stages at the head of the queue, longest chain first; the
  // single-stage groups fill the gaps between later levels (split by data
  // bytes), heaviest items first, so that a dependent stage is reached only
  // after other work has given its dependency time to finish.
  std::vector<const QStage*> order;
  int maxl = 0;
  for (size_t g = 0; g < G; g++) maxl = std::max(maxl, (int)gst[g].size());
  std::vector<size_t> multi, single;
  for (size_t g = 0; g < G; g++) {
    if (gst[g].size() > 1) multi.push_back(g);
    if (gst[g].size() == 1) single.push_back(g);
  }
  std::stable_sort(multi.begin(), multi.end(), [&](size_t a, size_t b) { return gst[a].size() > gst[b].size(); });
  auto item_work = [&](size_t g) { return gst[g][0].work / (double)std::max<i64>(1, gst[g][0].nitems); };
  std::stable_sort(single.begin(), single.end(), [&](size_t a, size_t b) { return item_work(a) > item_work(b); });
  double fill_total = 0;
  for (size_t g : single) fill_total += gst[g][0].work;
  const int gaps = std::max(1, maxl - 1);
  size_t si = 0;
  double placed = 0;
  for (int L = 0; L < std::max(1, maxl); L++) {
    for (size_t g : multi)
      if ((int)gst[g].size() > L) order.push_back(&gst[g][L]);
    if (L < gaps) {
      const double upto = fill_total * (double)(L + 1) / (double)gaps;
      while (si < single.size() && (L == gaps - 1 || placed < upto)) {
        placed += gst[single[si]][0].work;
        order.push_back(&gst[single[si++]][0]);
      }
    }
  }
  while (si < single.size()) order.push_back(&gst[single[si++]][0]);
  // Stage table, dependencies, LDS and the descriptor blocks.
  std::map<const QStage*, int> pos;
  for (size_t i = 0; i < order.size(); i++) pos[order[i]] = (int)i;
  size_t pass_lds = 0, conv_lds = 0;
  bool any_pass = false;
  i64 items = 0;
  for (const QStage* st : order) {
    if (st->kind == 0) {
      pass_lds = std::max(pass_lds, st->lds);
      any_pass = true;
    } else {
      conv_lds = std::max(conv_lds, st->lds);
    }
    items += st->nitems;
  }
  const size_t sh_off = (pass_lds + 255) & ~(size_t)255;
  Q.threads = threads;
  Q.lds = (std::max(any_pass ? sh_off + tables : 0, conv_lds) + 15) & ~(size_t)15;
  const int bpc = queue_blocks_per_cu(prec, threads, Q.lds);
  if (bpc < 1) {
    Q.in_queue.assign(G, 0);
    return GOFFT_OK;
  }
  Q.grid = std::min<i64>(items, (i64)bpc * dev.cus);
  Q.ctr_words = 4 + order.size();  // ticket (2 words), error, pad, done[stage]
  Q.off_desc = 256 * ((Q.ctr_words * 4 + 255) / 256);
  size_t desc_bytes = 0;
  for (int var = 0; var < 4; var++) {
    auto& V = Q.var[var];
    std::vector<char> block;
    std::vector<QueueDesc> table;
    QueueArgs& A = V.args;
    A = QueueArgs{};
    A.nstages = (int)order.size();
    A.nitems = items;
    A.spin_limit = env_int("GOFFT_QUEUE_SPIN_TICKS", 200000000);  // 2 s of the 100 MHz clock
    A.need_extra = (int)env_int("GOFFT_QUEUE_TEST_STALL", 0);
    A.sh_off = (long long)sh_off;
    i64 start = 0;
    for (size_t i = 0; i < order.size(); i++) {
      const QStage& st = *order[i];
      const size_t off = (block.size() + 63) & ~(size_t)63;
      block.resize(off + st.desc[var].size());
      std::memcpy(block.data() + off, st.desc[var].data(), st.desc[var].size());
      // (a workspace relocation's offset already holds its group's offset:
      // the group ran on the stand-in base plus that offset)
      QueueDesc qd = st.rel[var];
      qd.off = (long long)off;
      qd.bytes = (long long)st.desc[var].size();
      table.push_back(qd);
      QueueStage& qs = A.st[i];
      qs.start = start;
      qs.nitems = st.nitems;
      qs.desc = (long long)off;
      qs.kind = st.kind;
      qs.form = st.form;
      qs.dep = st.level > 0 ? pos[&gst[st.group][st.level - 1]] : -1;
      qs.signal = st.level + 1 < (int)gst[st.group].size() ? 1 : 0;
      start += st.nitems;
    }
    desc_bytes = std::max(desc_bytes, block.size());
    RC_TRY(V.tmpl.upload(block.data(), block.size()));
    RC_TRY(V.table.upload(table.data(), table.size() * sizeof(QueueDesc)));
    V.ndesc = (int)table.size();
  }
  Q.off_groups = Q.off_desc + 256 * ((desc_bytes + 255) / 256);
  p.ws_need = std::max(p.ws_need, Q.off_groups + area);
  Q.check = env_int("GOFFT_QUEUE_CHECK", 0) != 0;
  Q.on = true;
  int ng = 0;
  for (size_t g = 0; g < G; g++) ng += Q.in_queue[g];
  Q.text = "queue threads=" + std::to_string(threads) + " lds=" + std::to_string(Q.lds) +
           " grid=" + std::to_string(Q.grid) + " (" + std::to_string(bpc) + "/CU) stages=" +
           std::to_string(order.size()) + " items=" + std::to_string(items) + " groups=" + std::to_string(ng) + "/" +
           std::to_string(G) + "\n  order:";
  // stage list: length.launch (items, LDS KiB) and "<k" for a wait on stage k
  for (size_t i = 0; i < order.size(); i++) {
    const QStage& st = *order[i];
    const QueueStage& qs = Q.var[0].args.st[i];
    Q.text += " " + std::to_string(p.groups[st.group]->len) + "." + std::to_string(st.level) + "(" +
              std::to_string(st.nitems) + "," + std::to_string((st.lds + 1023) / 1024) + "K" +
              (qs.dep >= 0 ? ",<" + std::to_string(qs.dep) : std::string()) + ")";
  }
  return GOFFT_OK;
}

// Workspace of one exec: the plan's per-stream pool (a null WsSet), or fixed
// buffers of at least the plan's workspace bytes each: `main` for the
// caller's stream and lanes[i] for a mixed-length plan's side lane i (a
// graph's own workspace, or the no-allocation entry point's scratch).
struct WsSet {
  char* main = nullptr;
  std::vector<char*> lanes;
};

int exec_ws(const gofft_plan* p, hipStream_t s, const WsSet* w, char** out) {
  if (w) {
    *out = w->main;
    return GOFFT_OK;
  }
  return p->ws(s, out);
}

// Runs every group of a mixed-length plan: on the caller's stream alone, or
// forked over the plan's lanes and joined back.
// The queued groups of a mixed-length plan on stream s: the setup kernel
// resolves the variant's descriptors into the workspace and zeroes the
// counters, then the queue kernel runs every stage.
int run_queue(const gofft_plan* p, const void* src, void* dst, int c, bool scaled, hipStream_t s,
              const WsSet* w) {
  const auto& Q = p->q;
  char* ws = nullptr;
  RC_TRY(exec_ws(p, s, w, &ws));
  const auto& V = Q.var[(c ? 1 : 0) + (scaled ? 2 : 0)];
  QueueSetupArgs sa{};
  sa.tmpl = (const char*)V.tmpl.p;
  sa.descs = ws + Q.off_desc;
  sa.d = (const QueueDesc*)V.table.p;
  sa.base[1] = (const char*)src;
  sa.base[2] = (const char*)dst;
  sa.base[3] = ws + Q.off_groups;
  sa.ctr = (unsigned*)ws;
  sa.nctr = (int)Q.ctr_words;
  HIP_TRY(launch_queue_setup(sa, V.ndesc, s));
  QueueArgs a = V.args;
  a.descs = ws + Q.off_desc;
  a.ticket = (unsigned long long*)ws;
  a.error = (unsigned*)(ws + 8);
  a.done = (unsigned*)(ws + 16);
  HIP_TRY(launch_queue(p->prec, a, Q.threads, Q.lds, Q.grid, s));
  if (Q.check) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIP_TRY(hipStreamIsCapturing(s, &cs));
    if (cs == hipStreamCaptureStatusNone) {
      unsigned err = 0;
      HIP_TRY(hipStreamSynchronize(s));
      HIP_TRY(hipMemcpy(&err, a.error, sizeof(err), hipMemcpyDeviceToHost));
      if (err) return fail(GOFFT_ERR_HIP, "work queue: a wait for an earlier stage exceeded its time limit");
    }
  }
  return GOFFT_OK;
}

int run_groups(const gofft_plan* p, const void* src, void* dst, int ci, int co, bool inverse_scale,
               hipStream_t s, const WsSet* w) {
  auto one = [&](const gofft_plan::Group& g, hipStream_t st, int lane) -> int {
    char* ws = nullptr;
    if (w)
      ws = lane == 0 ? w->main : w->lanes[lane - 1];
    else
      RC_TRY(p->ws(st, &ws));
    Ctx c{st, ws};
    const i64* o = (const i64*)g.off.p;
    return g.eng->run(src, LineIO{o, 0}, dst, LineIO{o, 0}, ci, co, inverse_scale ? 1.0 / (double)g.len : 1.0, c);
  };
  const bool queue = p->q.on;
  auto queued = [&](size_t g) { return queue && p->q.in_queue[g]; };
  if (p->lanes.empty()) {
    for (size_t g : p->order)
      if (!queued(g)) RC_TRY(one(*p->groups[g], s, 0));
    if (queue) RC_TRY(run_queue(p, src, dst, ci, inverse_scale, s, w));
    return GOFFT_OK;
  }
  std::lock_guard<std::mutex> lk(p->lane_mu);
  HIP_TRY(hipEventRecord(p->fork_ev, s));
  for (hipStream_t l : p->lanes) HIP_TRY(hipStreamWaitEvent(l, p->fork_ev, 0));
  if (queue) RC_TRY(run_queue(p, src, dst, ci, inverse_scale, s, w));
  for (size_t g : p->order) {
    if (queued(g)) continue;
    const int k = p->lane_of[g];
    RC_TRY(one(*p->groups[g], k == 0 ? s : p->lanes[k - 1], k));
  }
  for (size_t i = 0; i < p->lanes.size(); i++) {
    HIP_TRY(hipEventRecord(p->join_ev[i], p->lanes[i]));
    HIP_TRY(hipStreamWaitEvent(s, p->join_ev[i], 0));
  }
  return GOFFT_OK;
}

int new_plan(gofft_plan** out, std::unique_ptr<gofft_plan>& p) {
  *out = p.release();
  return GOFFT_OK;
}

int check_dtype(int dtype) {
  if (dtype != GOFFT_C128 && dtype != GOFFT_C64)
    return fail(GOFFT_ERR_INVALID_ARG, "dtype must be GOFFT_C128 or GOFFT_C64");
  return GOFFT_OK;
}

// The plain complex128 path reproduces the reference's Bluestein rounding
// unless accurate twiddles are requested.
unsigned default_flags(int dtype, unsigned flags) {
  if (dtype == GOFFT_C128 && !(flags & GOFFT_FLAG_ACCURATE_TWIDDLES)) flags |= GOFFT_FLAG_REFERENCE_TWIDDLES;
  if (const char* e = std::getenv("GOFFT_QUEUE"))
    if (std::atoi(e) != 0) flags |= GOFFT_FLAG_QUEUE;
  return flags;
}

// One axis of an N-D transform on a row-major array: `outer` blocks of n
// elements at stride st, each block holding st interleaved lines.
int build_axis(int prec, unsigned flags, i64 n, i64 st, i64 outer, NDAxis& a, size_t* ws_need) {
  a.n = n;
  a.st = st;
  a.outer = outer;
  if (n == 1) return GOFFT_OK;
  int rc = GOFFT_OK;
  if (st == 1) {
    a.eng = make_engine(n, outer, prec, flags, &rc);
  } else if (factors_are_small(n)) {
    auto e = std::make_unique<StockhamEngine>();
    e->prec = prec;
    e->n = n;
    e->batch = outer;
    e->st = st;
    e->inner = st;
    rc = e->init(flags);
    if (rc == GOFFT_OK) a.eng = std::move(e);
  } else {
    a.gather = true;
    a.eng = make_engine(n, outer * st, prec, flags, &rc);
  }
  if (!a.eng) return rc;
  size_t need = a.eng->ws_bytes();
  if (a.gather) need += 2 * (size_t)n * st * outer * elt(prec);
  *ws_need = std::max(*ws_need, need);
  return GOFFT_OK;
}

int build_nd_axes(int prec, unsigned flags, const std::vector<i64>& shape, i64 batch,
                  std::vector<NDAxis>& axes, size_t* ws_need) {
  const int rank = (int)shape.size();
  i64 total = 1;
  for (i64 s : shape) total *= s;
  for (int ax = 0; ax < rank; ax++) {
    i64 st = 1;
    for (int k = ax + 1; k < rank; k++) st *= shape[k];
    NDAxis a;
    RC_TRY(build_axis(prec, flags, shape[ax], st, total / (shape[ax] * st) * batch, a, ws_need));
    axes.push_back(std::move(a));
  }
  return GOFFT_OK;
}

// Runs the axes of an N-D complex transform: src -> dst on the first
// non-trivial axis, then in place on dst. `total` is the element count of the
// whole (batched) array. An inverse conjugates on the first load and the last
// store only: conj(F_all(conj x)) composes per axis.
int run_nd_axes(const std::vector<NDAxis>& axes, int prec, i64 total, const void* src, void* dst,
                int inverse, double final_scale, hipStream_t s, char* ws) {
  const void* cur = src;
  int last = -1;
  for (int i = 0; i < (int)axes.size(); i++)
    if (axes[i].n > 1) last = i;
  if (last < 0) {
    if (total > 0 && src != dst) HIP_TRY(launch_copy_lines(prec, src, dst, total, 1, total, total, s));
    if (total > 0 && final_scale != 1.0) HIP_TRY(launch_scale(prec, dst, dst, 2 * total, final_scale, s));
    return GOFFT_OK;
  }
  int first = 1;
  for (int i = 0; i < (int)axes.size(); i++) {
    const NDAxis& a = axes[i];
    if (a.n == 1) continue;
    int ci = inverse && first ? 1 : 0;
    int co = inverse && i == last ? 1 : 0;
    double sc = i == last ? final_scale : 1.0;
    if (a.st == 1) {
      Ctx c{s, ws};
      RC_TRY(a.eng->run(cur, LineIO{nullptr, a.n}, dst, LineIO{nullptr, a.n}, ci, co, sc, c));
    } else if (!a.gather) {
      Ctx c{s, ws};
      RC_TRY(a.eng->run(cur, LineIO{nullptr, 0}, dst, LineIO{nullptr, 0}, ci, co, sc, c));
    } else {
      const i64 elems = a.n * a.st * a.outer;
      char* A = ws;
      char* B = ws + (size_t)elems * elt(prec);
      Ctx c{s, B + (size_t)elems * elt(prec)};
      i64 lines = elems / a.n;
      HIP_TRY(launch_axis_copy(prec, cur, A, a.n, a.st, lines, 0, s));
      RC_TRY(a.eng->run(A, LineIO{nullptr, a.n}, B, LineIO{nullptr, a.n}, ci, co, sc, c));
      HIP_TRY(launch_axis_copy(prec, B, dst, a.n, a.st, lines, 1, s));
    }
    cur = dst;
    first = 0;
  }
  return GOFFT_OK;
}

// One axis of an N-D plan over its first `nouter` outer blocks (an N-D
// transform's axis a.outer blocks are the arrays times the other axes'
// blocks; for the column axis of an RFFT2 plan they are the arrays), in place
// or src -> dst.
int run_axis_blocks(const NDAxis& a, int prec, const void* src, void* dst, int ci, int co, double sc, i64 nouter,
                    hipStream_t s, char* ws) {
  if (a.n == 1) {
    const i64 elems = a.st * nouter;
    if (src != dst) HIP_TRY(launch_copy_lines(prec, src, dst, elems, 1, elems, elems, s));
    if (sc != 1.0) HIP_TRY(launch_scale(prec, dst, dst, 2 * elems, sc, s));
    return GOFFT_OK;
  }
  if (a.st == 1) return a.eng->run(src, LineIO{nullptr, a.n}, dst, LineIO{nullptr, a.n}, ci, co, sc, Ctx{s, ws}, nouter);
  if (!a.gather) return a.eng->run(src, LineIO{nullptr, 0}, dst, LineIO{nullptr, 0}, ci, co, sc, Ctx{s, ws}, nouter);
  const i64 elems = a.n * a.st * nouter;
  char* A = ws;
  char* B = ws + (size_t)a.n * a.st * a.outer * elt(prec);
  char* rest = B + (size_t)a.n * a.st * a.outer * elt(prec);
  const i64 lines = elems / a.n;
  HIP_TRY(launch_axis_copy(prec, src, A, a.n, a.st, lines, 0, s));
  RC_TRY(a.eng->run(A, LineIO{nullptr, a.n}, B, LineIO{nullptr, a.n}, ci, co, sc, Ctx{s, rest}, lines));
  HIP_TRY(launch_axis_copy(prec, B, dst, a.n, a.st, lines, 1, s));
  return GOFFT_OK;
}

}  // namespace

// ===========================================================================
// C ABI.

extern "C" {

int gofft_init(void) {
  DevInfo d;
  return current_device(&d);
}

const char* gofft_last_error(void) { return t_err.c_str(); }

const char* gofft_version(void) { return kVersionText; }

int gofft_route(int64_t n) { return route(n); }

int gofft_plan_c2c(gofft_plan** plan, int64_t n, int64_t batch, int dtype, unsigned flags) {
  if (!plan) return fail(GOFFT_ERR_INVALID_ARG, "null plan pointer");
  *plan = nullptr;
  RC_TRY(check_dtype(dtype));
  if (n < 0 || batch < 0) return fail(GOFFT_ERR_INVALID_ARG, "negative length or batch");
  RC_TRY(gofft_init());
  auto p = std::make_unique<gofft_plan>();
  p->kind = K_C2C;
  p->prec = dtype;
  p->flags = default_flags(dtype, flags);
  p->n = n;
  p->batch = batch;
  int rc = GOFFT_OK;
  p->eng = make_engine(n, batch, dtype, p->flags, &rc);
  if (!p->eng) return rc;
  p->ws_need = p->eng->ws_bytes();
  return new_plan(plan, p);
}

int gofft_plan_c2c_many(gofft_plan** plan, const int64_t* lengths, int64_t count, int dtype,
                        unsigned flags) {
  if (!plan) return fail(GOFFT_ERR_INVALID_ARG, "null plan pointer");
  *plan = nullptr;
  RC_TRY(check_dtype(dtype));
  if (count < 0 || (count > 0 && !lengths)) return fail(GOFFT_ERR_INVALID_ARG, "bad length list");
  RC_TRY(gofft_init());
  auto p = std::make_unique<gofft_plan>();
  p->kind = K_MANY;
  p->prec = dtype;
  p->flags = default_flags(dtype, flags);
  p->batch = count;
  std::map<i64, std::vector<i64>> offs;
  i64 off = 0;
  for (i64 i = 0; i < count; i++) {
    if (lengths[i] < 0) return fail(GOFFT_ERR_INVALID_ARG, "negative length");
    offs[lengths[i]].push_back(off);
    off += lengths[i];
  }
  p->total_elems = off;
  // Largest lengths first so the long transforms start early on the stream.
  std::vector<i64> lens;
  for (auto& kv : offs) lens.push_back(kv.first);
  std::sort(lens.begin(), lens.end(), [](i64 a, i64 b) { return a > b; });
  for (i64 len : lens) {
    if (len == 0) continue;
    auto g = std::make_unique<gofft_plan::Group>();
    g->len = len;
    const auto& o = offs[len];
    int rc = GOFFT_OK;
    g->eng = make_engine(len, (i64)o.size(), dtype, p->flags, &rc);
    if (!g->eng) return rc;
    RC_TRY(g->off.upload(o.data(), o.size() * sizeof(i64)));
    p->ws_need = std::max(p->ws_need, g->eng->ws_bytes());
    p->groups.push_back(std::move(g));
  }
  if (p->flags & GOFFT_FLAG_QUEUE) RC_TRY(build_queue(*p));
  RC_TRY(assign_lanes(*p));
  return new_plan(plan, p);
}

int gofft_plan_r2c(gofft_plan** plan, int64_t n, int64_t batch, int dtype, unsigned flags) {
  if (!plan) return fail(GOFFT_ERR_INVALID_ARG, "null plan pointer");
  *plan = nullptr;
  RC_TRY(check_dtype(dtype));
  if (n < 0 || batch < 0) return fail(GOFFT_ERR_INVALID_ARG, "negative length or batch");
  RC_TRY(gofft_init());
  auto p = std::make_unique<gofft_plan>();
  p->kind = K_R2C;
  p->prec = dtype;
  p->flags = default_flags(dtype, flags);
  p->n = n;
  p->batch = batch;
  p->real = std::make_unique<RealEngine>();
  p->real->prec = dtype;
  p->real->n = n;
  p->real->batch = batch;
  if (n > 0) RC_TRY(p->real->init(p->flags));
  p->ws_need = n > 0 ? p->real->ws_bytes() : 0;
  return new_plan(plan, p);
}

int gofft_plan_psd(gofft_plan** plan, int64_t n, int64_t batch, double d, int dtype, unsigned flags) {
  int rc = gofft_plan_r2c(plan, n, batch, dtype, flags);
  if (rc != GOFFT_OK) return rc;
  (*plan)->kind = K_PSD;
  (*plan)->d = d;
  return GOFFT_OK;
}

int gofft_plan_nd(gofft_plan** plan, const int64_t* shape, int rank, int kind, int64_t batch,
                  int dtype, unsigned flags) {
  if (!plan) return fail(GOFFT_ERR_INVALID_ARG, "null plan pointer");
  *plan = nullptr;
  RC_TRY(check_dtype(dtype));
  if (rank < 0 || (rank > 0 && !shape) || batch < 0) return fail(GOFFT_ERR_INVALID_ARG, "bad shape");
  for (int i = 0; i < rank; i++)
    if (shape[i] <= 0) return fail(GOFFT_ERR_SHAPE, "fft: shape lengths must be positive");
  RC_TRY(gofft_init());
  auto p = std::make_unique<gofft_plan>();
  p->prec = dtype;
  p->flags = default_flags(dtype, flags);
  p->batch = batch;
  p->shape.assign(shape, shape + rank);
  p->nd_total = 1;
  for (int i = 0; i < rank; i++) p->nd_total *= shape[i];
  if (kind == GOFFT_ND_C2C) {
    p->kind = K_ND_C2C;
    p->n = p->nd_total;
    RC_TRY(build_nd_axes(dtype, p->flags, p->shape, batch, p->axes, &p->ws_need));
    return new_plan(plan, p);
  }
  if (kind != GOFFT_ND_R2C || rank != 2)
    return fail(GOFFT_ERR_UNSUPPORTED, "R2C plans are two-dimensional (RFFT2/IRFFT2)");
  p->kind = K_ND_R2C;
  p->rows = shape[0];
  p->cols = shape[1];
  p->rcols = p->cols / 2 + 1;
  p->n = p->rows * p->cols;
  p->nd_row = std::make_unique<RealEngine>();
  p->nd_row->prec = dtype;
  p->nd_row->n = p->cols;
  p->nd_row->batch = p->rows * batch;
  RC_TRY(p->nd_row->init(p->flags));
  // Column transform over the rows x rcols complex half-spectrum (the only
  // complex axis of RFFT2, rfft2.go:55-71).
  NDAxis col;
  size_t need = p->nd_row->ws_bytes();
  RC_TRY(build_axis(dtype, p->flags, p->rows, p->rcols, batch, col, &need));
  p->axes.push_back(std::move(col));
  // IRFFT2 works on a padded copy of the spectrum.
  p->ws_engines = need;
  p->ws_need = need + (size_t)p->rows * p->rcols * batch * elt(dtype);
  return new_plan(plan, p);
}

int gofft_plan_spectrogram(gofft_plan** plan, int64_t length, int64_t segment, int64_t overlap,
                           const double* window, int64_t window_len, double d, int dtype,
                           unsigned flags) {
  if (!plan) return fail(GOFFT_ERR_INVALID_ARG, "null plan pointer");
  *plan = nullptr;
  RC_TRY(check_dtype(dtype));
  // Argument checks in the reference's order and wording (freq.go:108-116).
  if (segment <= 0) return fail(GOFFT_ERR_INVALID_ARG, "fft: segment length must be positive");
  if (window_len != segment || !window)
    return fail(GOFFT_ERR_INVALID_ARG, "fft: window length must equal segment length");
  if (overlap < 0 || overlap >= segment)
    return fail(GOFFT_ERR_INVALID_ARG, "fft: overlap must satisfy 0 <= overlap < segment");
  if (length < 0) return fail(GOFFT_ERR_INVALID_ARG, "negative signal length");
  RC_TRY(gofft_init());
  auto p = std::make_unique<gofft_plan>();
  p->kind = K_SPEC;
  p->prec = dtype;
  p->flags = default_flags(dtype, flags);
  p->seg = segment;
  p->step = segment - overlap;
  p->sig_len = length;
  p->d = d;
  p->frames = 0;
  for (i64 start = 0; start + segment <= length; start += p->step) p->frames++;
  p->n = segment;
  p->batch = p->frames;
  std::vector<long double> w(window, window + segment);
  RC_TRY(upload_real(p->window, w, dtype));
  p->real = std::make_unique<RealEngine>();
  p->real->prec = dtype;
  p->real->n = segment;
  p->real->batch = p->frames;
  if (p->frames > 0) RC_TRY(p->real->init(p->flags));
  // The fused kernel reads frames in place; otherwise frames are gathered.
  p->spec_gather = !p->real->single;
  p->ws_need = p->frames > 0 ? p->real->ws_bytes() : 0;
  if (p->spec_gather && p->frames > 0) p->ws_need += (size_t)p->frames * segment * relt(dtype);
  return new_plan(plan, p);
}

int gofft_plan_destroy(gofft_plan* plan) {
  delete plan;
  return GOFFT_OK;
}

int64_t gofft_plan_len(const gofft_plan* p) { return p ? p->n : -1; }
int64_t gofft_plan_frames(const gofft_plan* p) { return p ? p->frames : -1; }
int64_t gofft_plan_workspace_bytes(const gofft_plan* p) { return p ? (int64_t)p->ws_need : -1; }

// Bins per line of a real plan: n/2 + 1, none for an empty line.
static i64 real_bins(const gofft_plan* p) { return p->n > 0 ? p->n / 2 + 1 : 0; }

int64_t gofft_plan_in_elems(const gofft_plan* p, int direction) {
  if (!p) return -1;
  switch (p->kind) {
    case K_C2C: return p->n * p->batch;
    case K_MANY: return p->total_elems;
    case K_R2C: return direction == 0 ? p->n * p->batch : real_bins(p) * p->batch;
    case K_PSD: return p->n * p->batch;
    case K_ND_C2C: return p->nd_total * p->batch;
    case K_ND_R2C: return direction == 0 ? p->rows * p->cols * p->batch : p->rows * p->rcols * p->batch;
    case K_SPEC: return p->sig_len;
  }
  return -1;
}

int64_t gofft_plan_out_elems(const gofft_plan* p, int direction) {
  if (!p) return -1;
  switch (p->kind) {
    case K_C2C: return p->n * p->batch;
    case K_MANY: return p->total_elems;
    case K_R2C: return direction == 0 ? real_bins(p) * p->batch : p->n * p->batch;
    case K_PSD: return real_bins(p) * p->batch;
    case K_ND_C2C: return p->nd_total * p->batch;
    case K_ND_R2C: return direction == 0 ? p->rows * p->rcols * p->batch : p->rows * p->cols * p->batch;
    case K_SPEC: return p->frames * (p->seg / 2 + 1);
  }
  return -1;
}

int gofft_plan_num_launches(const gofft_plan* p, int direction) {
  if (!p) return -1;
  (void)direction;
  switch (p->kind) {
    case K_C2C: return p->eng->launches();
    case K_MANY: {
      int k = p->q.on ? 2 : 0;  // the queue's setup and queue kernels
      for (size_t g = 0; g < p->groups.size(); g++)
        if (!p->q.on || !p->q.in_queue[g]) k += p->groups[g]->eng->launches();
      return k;
    }
    case K_R2C:
    case K_PSD: return p->n > 0 ? p->real->launches() : 0;
    case K_ND_C2C: {
      int k = 0;
      for (auto& a : p->axes) k += a.eng ? a.eng->launches() + (a.gather ? 2 : 0) : 0;
      return k;
    }
    case K_ND_R2C: {
      int k = p->nd_row->launches();
      for (auto& a : p->axes) k += a.eng ? a.eng->launches() + (a.gather ? 2 : 0) : 0;
      return k;
    }
    case K_SPEC: return p->frames > 0 ? p->real->launches() + (p->spec_gather ? 1 : 0) : 0;
  }
  return -1;
}

int64_t gofft_plan_describe(const gofft_plan* p, char* buf, int64_t len) {
  if (!p) return fail(GOFFT_ERR_INVALID_ARG, "null plan");
  std::string s;
  auto line = [&](const std::string& t) { s += t + "\n"; };
  switch (p->kind) {
    case K_C2C: line(p->eng->describe()); break;
    case K_MANY:
      if (p->q.on) line(p->q.text);
      for (size_t g = 0; g < p->groups.size(); g++)
        line(std::string(p->q.on && p->q.in_queue[g] ? "queued: " : "") + p->groups[g]->eng->describe());
      break;
    case K_R2C:
    case K_PSD:
    case K_SPEC:
      if (p->real) line(p->real->describe());
      break;
    case K_ND_C2C:
    case K_ND_R2C:
      if (p->nd_row) line("rows: " + p->nd_row->describe());
      for (const auto& a : p->axes)
        if (a.eng) line(std::string(a.gather ? "axis (gathered): " : "axis: ") + a.eng->describe());
      break;
  }
  if (buf && len > 0) {
    const size_t k = std::min<size_t>((size_t)len - 1, s.size());
    std::memcpy(buf, s.data(), k);
    buf[k] = 0;
  }
  return (int64_t)s.size() + 1;
}

}  // extern "C"

namespace {

// The plan's engine for transform length n (a length group, an axis, the
// half-length transform of a real plan), or null.
const Engine* find_engine(const gofft_plan* p, i64 n) {
  auto pick = [&](const Engine* e) { return e && e->n == n ? e : nullptr; };
  const Engine* e = nullptr;
  switch (p->kind) {
    case K_C2C: e = pick(p->eng.get()); break;
    case K_MANY:
      for (const auto& g : p->groups)
        if (!e) e = pick(g->eng.get());
      break;
    case K_R2C:
    case K_PSD:
    case K_SPEC:
      if (p->real) e = pick(p->real->half.get());
      break;
    case K_ND_C2C:
    case K_ND_R2C:
      if (p->nd_row) e = pick(p->nd_row->half.get());
      for (const auto& a : p->axes)
        if (!e) e = pick(a.eng.get());
      break;
  }
  return e;
}

}  // namespace

extern "C" {

int64_t gofft_plan_index_table(const gofft_plan* p, int64_t n, int table, int64_t* out, int64_t cap) {
  if (!p) return fail(GOFFT_ERR_INVALID_ARG, "null plan");
  const Engine* e = find_engine(p, n);
  if (!e) return fail(GOFFT_ERR_INVALID_ARG, "the plan has no transform of length " + std::to_string(n));
  std::vector<i64> v;
  const RaderEngine* ra = dynamic_cast<const RaderEngine*>(e);
  const BluesteinEngine* bl = dynamic_cast<const BluesteinEngine*>(e);
  switch (table) {
    case GOFFT_TABLE_RADER_GENERATOR:
    case GOFFT_TABLE_RADER_CONV:
      if (!ra) return fail(GOFFT_ERR_INVALID_ARG, "length " + std::to_string(n) + " is not routed to Rader");
      if (table == GOFFT_TABLE_RADER_GENERATOR) v = {ra->g};
      else v = {ra->cl, ra->cyclic ? 1 : 0};
      break;
    case GOFFT_TABLE_RADER_PERM:
    case GOFFT_TABLE_RADER_IPERM: {
      if (!ra) return fail(GOFFT_ERR_INVALID_ARG, "length " + std::to_string(n) + " is not routed to Rader");
      // read back the device tables the kernels index with
      const DevBuf& b = table == GOFFT_TABLE_RADER_PERM ? ra->perm : ra->iperm;
      std::vector<int> h((size_t)ra->q);
      if (ra->q > 0) HIP_TRY(hipMemcpy(h.data(), b.p, h.size() * sizeof(int), hipMemcpyDeviceToHost));
      v.assign(h.begin(), h.end());
      break;
    }
    case GOFFT_TABLE_CHIRP_INDEX:
    case GOFFT_TABLE_BLUESTEIN_CONV:
      if (!bl) return fail(GOFFT_ERR_INVALID_ARG, "length " + std::to_string(n) + " is not routed to Bluestein");
      if (table == GOFFT_TABLE_CHIRP_INDEX) v = bl->jj;
      else v = {bl->m};
      break;
    default:
      return fail(GOFFT_ERR_INVALID_ARG, "unknown table");
  }
  if (out && cap > 0) std::memcpy(out, v.data(), sizeof(i64) * (size_t)std::min<i64>(cap, (i64)v.size()));
  return (int64_t)v.size();
}

int gofft_plan_kernel_work(const gofft_plan* p, int direction, int kernel, int64_t* launches,
                           int64_t* bytes) {
  if (!p || !launches || !bytes) return fail(GOFFT_ERR_INVALID_ARG, "null argument");
  if (kernel < 0 || kernel >= GOFFT_KERNEL_COUNT) return fail(GOFFT_ERR_INVALID_ARG, "unknown kernel family");
  if (direction != GOFFT_FORWARD && direction != GOFFT_INVERSE)
    return fail(GOFFT_ERR_INVALID_ARG, "direction must be GOFFT_FORWARD or GOFFT_INVERSE");
  const int inv = direction == GOFFT_INVERSE;
  const i64 e = (i64)elt(p->prec);
  Work w;
  auto axes_work = [&](const std::vector<NDAxis>& axes, i64 total) {
    bool any = false;
    for (const auto& a : axes) {
      if (a.n == 1 || !a.eng) continue;
      any = true;
      if (a.gather) {
        const i64 elems = a.n * a.st * a.outer;
        w.add(GOFFT_KERNEL_AUX, 2, 4 * elems * e);
        a.eng->work(w, a.outer * a.st);
      } else {
        a.eng->work(w, a.outer);
      }
    }
    if (!any && total > 0) w.add(GOFFT_KERNEL_AUX, 1, 2 * total * e);
  };
  switch (p->kind) {
    case K_C2C:
      if (p->n > 0) p->eng->work(w, p->batch);
      break;
    case K_MANY:
      for (const auto& g : p->groups) g->eng->work(w, g->eng->batch);
      break;
    case K_R2C: p->real->work(w, inv, 0, p->n / 2 + 1); break;
    case K_PSD: p->real->work(w, 0, 1, 0); break;
    case K_SPEC:
      if (p->frames > 0) {
        if (p->spec_gather) w.add(GOFFT_KERNEL_AUX, 1, 2 * p->frames * p->seg * e / 2);
        p->real->work(w, 0, 1, 0);
      }
      break;
    case K_ND_C2C: axes_work(p->axes, p->nd_total * p->batch); break;
    case K_ND_R2C: {
      const i64 htotal = p->rows * p->rcols * p->batch;
      if (p->batch == 0) break;
      if (!inv) {
        p->nd_row->work(w, 0, 0, p->rcols);
        axes_work(p->axes, 0);
      } else {
        w.add(GOFFT_KERNEL_AUX, 1, 2 * htotal * e);
        axes_work(p->axes, htotal);
        p->nd_row->work(w, 1, 0, p->rcols);
      }
      break;
    }
  }
  *launches = w.launches[kernel];
  *bytes = w.bytes[kernel];
  return GOFFT_OK;
}

// ---------------------------------------------------------------------------
// Execution. Every exec and VJP entry point, device or host, goes through
// exec_op; graph capture records one exec_op.

}  // extern "C"

namespace {

enum ExecOp {
  OP_C2C = 0,
  OP_R2C,
  OP_C2R,
  OP_ND_R2C,
  OP_PSD,
  OP_SPECTROGRAM,
  OP_VJP_C2C,
  OP_VJP_R2C,
  OP_VJP_C2R,
  OP_VJP_ND_R2C,
  OP_COUNT
};

bool is_complex_plan(const gofft_plan* p) { return p->kind == K_C2C || p->kind == K_MANY || p->kind == K_ND_C2C; }

// The plan kind an op needs, and the error text when the plan is another.
int check_kind(const gofft_plan* p, int op) {
  switch (op) {
    case OP_C2C:
    case OP_VJP_C2C:
      if (!is_complex_plan(p)) return fail(GOFFT_ERR_INVALID_ARG, "plan is not a complex-to-complex plan");
      return GOFFT_OK;
    case OP_R2C:
    case OP_C2R:
    case OP_VJP_R2C:
    case OP_VJP_C2R:
      if (p->kind != K_R2C) return fail(GOFFT_ERR_INVALID_ARG, "plan is not an r2c plan");
      return GOFFT_OK;
    case OP_ND_R2C:
    case OP_VJP_ND_R2C:
      if (p->kind != K_ND_R2C) return fail(GOFFT_ERR_INVALID_ARG, "plan is not an nd r2c plan");
      return GOFFT_OK;
    case OP_PSD:
      if (p->kind != K_PSD) return fail(GOFFT_ERR_INVALID_ARG, "plan is not a psd plan");
      return GOFFT_OK;
    case OP_SPECTROGRAM:
      if (p->kind != K_SPEC) return fail(GOFFT_ERR_INVALID_ARG, "plan is not a spectrogram plan");
      return GOFFT_OK;
  }
  return fail(GOFFT_ERR_INVALID_ARG, "unknown operation");
}

// Whether the op's direction argument is meaningful (and so checked).
bool op_has_direction(int op) { return op == OP_C2C || op == OP_VJP_C2C || op == OP_ND_R2C || op == OP_VJP_ND_R2C; }

// Whether the op takes a bin count (the inverse real transforms).
bool op_takes_bins(int op, int direction) {
  return op == OP_C2R || (op == OP_ND_R2C && direction == GOFFT_INVERSE);
}

// Bytes of the op's input and output arrays.
void io_bytes(const gofft_plan* p, int op, int direction, i64 bins, size_t* ib, size_t* ob) {
  const size_t e = elt(p->prec), r = relt(p->prec);
  const i64 nb = p->n > 0 ? p->n / 2 + 1 : 0;
  const i64 B = p->batch;
  size_t in = 0, out = 0;
  switch (op) {
    case OP_C2C:
    case OP_VJP_C2C:
      in = out = (size_t)gofft_plan_in_elems(p, GOFFT_FORWARD) * e;
      break;
    case OP_R2C:
      in = (size_t)(p->n * B) * r;
      out = (size_t)(nb * B) * e;
      break;
    case OP_C2R:
      in = (size_t)(std::max<i64>(bins, 0) * B) * e;
      out = (size_t)(p->n * B) * r;
      break;
    case OP_VJP_R2C:
      in = (size_t)(nb * B) * e;
      out = (size_t)(p->n * B) * r;
      break;
    case OP_VJP_C2R:
      in = (size_t)(p->n * B) * r;
      out = (size_t)(nb * B) * e;
      break;
    case OP_PSD:
      in = (size_t)(p->n * B) * r;
      out = (size_t)(nb * B) * r;
      break;
    case OP_SPECTROGRAM:
      in = (size_t)p->sig_len * r;
      out = (size_t)(p->frames * (p->seg / 2 + 1)) * r;
      break;
    case OP_ND_R2C:
      if (direction == GOFFT_FORWARD) {
        in = (size_t)(p->rows * p->cols * B) * r;
        out = (size_t)(p->rows * p->rcols * B) * e;
      } else {
        in = (size_t)(std::max<i64>(bins, 0) * B) * e;
        out = (size_t)(p->rows * p->cols * B) * r;
      }
      break;
    case OP_VJP_ND_R2C:
      if (direction == GOFFT_FORWARD) {
        in = (size_t)(p->rows * p->rcols * B) * e;
        out = (size_t)(p->rows * p->cols * B) * r;
      } else {
        in = (size_t)(p->rows * p->cols * B) * r;
        out = (size_t)(p->rows * p->rcols * B) * e;
      }
      break;
  }
  *ib = in;
  *ob = out;
}

// Argument checks shared by every exec path, before any work is enqueued: the
// plan kind, the direction, the bin count, data pointers for the elements
// the call reads and writes, and (real transforms) no overlap between input
// and output. The kernels of a batched real transform write line l's output
// over where line l+1's input would lie, and the reference's real calls
// cannot alias either (their input and output slices have different element
// types).
int check_op(const gofft_plan* p, int op, const void* src, i64 bins, const void* dst, int direction) {
  if (!p) return fail(GOFFT_ERR_INVALID_ARG, "null plan");
  RC_TRY(check_kind(p, op));
  if (op_has_direction(op) && direction != GOFFT_FORWARD && direction != GOFFT_INVERSE)
    return fail(GOFFT_ERR_INVALID_ARG, "direction must be GOFFT_FORWARD or GOFFT_INVERSE");
  if (op_takes_bins(op, direction) && bins < 0) return fail(GOFFT_ERR_INVALID_ARG, "negative bin count");
  size_t ib = 0, ob = 0;
  io_bytes(p, op, direction, bins, &ib, &ob);
  if ((ib > 0 && !src) || (ob > 0 && !dst)) return fail(GOFFT_ERR_INVALID_ARG, "null data pointer");
  if (op != OP_C2C && op != OP_VJP_C2C && ib > 0 && ob > 0) {
    const uintptr_t a = (uintptr_t)src, b = (uintptr_t)dst;
    if (a < b + ob && b < a + ib)
      return fail(GOFFT_ERR_INVALID_ARG, "input and output of a real transform must not overlap");
  }
  return GOFFT_OK;
}

// Enqueues one checked op on stream s.
int run_op(const gofft_plan* p, int op, const void* src, i64 bins, void* dst, int direction, hipStream_t s,
           const WsSet* w) {
  char* ws = nullptr;
  if (p->kind != K_MANY) RC_TRY(exec_ws(p, s, w, &ws));
  const int inv = direction == GOFFT_INVERSE;
  const i64 B = p->batch;
  switch (op) {
    case OP_C2C:
    case OP_VJP_C2C: {
      // VJP: y = F x  =>  gx = F^H gy = conj(F conj(gy)) = n * IFFT(gy);
      // y = IFFT x = F^H x / n  =>  gx = F gy / n.
      const int c = op == OP_C2C ? inv : !inv;
      const bool scaled = inv;
      if (p->kind == K_C2C) {
        if (p->n == 0 || B == 0) return GOFFT_OK;
        return p->eng->run(src, LineIO{nullptr, p->n}, dst, LineIO{nullptr, p->n}, c, c,
                           scaled ? 1.0 / (double)p->n : 1.0, Ctx{s, ws});
      }
      if (p->kind == K_MANY) return run_groups(p, src, dst, c, c, scaled, s, w);
      if (p->nd_total * B == 0) return GOFFT_OK;
      return run_nd_axes(p->axes, p->prec, p->nd_total * B, src, dst, c, scaled ? 1.0 / (double)p->nd_total : 1.0,
                         s, ws);
    }
    case OP_R2C:
      if (p->n == 0) return GOFFT_OK;
      return p->real->forward(src, p->n, dst, p->n / 2 + 1, 0, 1.0, nullptr, Ctx{s, ws});
    case OP_C2R: {
      if (p->n <= 0) return GOFFT_OK;
      const i64 keep = std::min<i64>(bins, p->n / 2 + 1);
      return p->real->inverse(src, bins, keep, dst, p->n, Ctx{s, ws});
    }
    case OP_PSD:
      if (p->n == 0) return GOFFT_OK;
      return p->real->forward(src, p->n, dst, p->n / 2 + 1, 1, p->d / (double)p->n, nullptr, Ctx{s, ws});
    case OP_SPECTROGRAM: {
      if (p->frames == 0) return GOFFT_OK;
      const i64 nbins = p->seg / 2 + 1;
      const double sc = p->d / (double)p->seg;
      if (!p->spec_gather) return p->real->forward(src, p->step, dst, nbins, 1, sc, p->window.p, Ctx{s, ws});
      char* F = ws;
      char* rest = ws + (size_t)p->frames * p->seg * relt(p->prec);
      HIP_TRY(launch_frames(p->prec, src, p->window.p, F, p->seg, p->step, p->frames, s));
      return p->real->forward(F, p->seg, dst, nbins, 1, sc, nullptr, Ctx{s, rest});
    }
    case OP_VJP_R2C: {
      // y = RFFT(x): dL/dx_j = Re sum_k gy_k W^{-jk} over the n/2+1 bins,
      // which is n * IRFFT(w * gy) with w = 1 at DC and Nyquist and 1/2
      // elsewhere (the IRFFT counts every interior bin twice through its
      // Hermitian mirror).
      if (p->n <= 0) return GOFFT_OK;
      const i64 nb = p->n / 2 + 1;
      return p->real->inverse(src, nb, nb, dst, p->n, Ctx{s, ws}, (double)p->n, 0.5);
    }
    case OP_VJP_C2R: {
      // x = IRFFT(X) reads Re X_0, Re X_{n/2} once and every other bin
      // twice: gX = RFFT(gx) / n with those bins doubled (Im of DC and
      // Nyquist is 0).
      if (p->n <= 0) return GOFFT_OK;
      const double in = 1.0 / (double)p->n;
      return p->real->forward(src, p->n, dst, p->n / 2 + 1, 2, 0.0, nullptr, Ctx{s, ws}, in, 2.0 * in);
    }
    case OP_ND_R2C: {
      if (B == 0) return GOFFT_OK;
      const i64 half = p->rows * p->rcols;
      const i64 htotal = half * B;
      if (!inv) {
        // Rows: real transform of each row into dst (rows x rcols), then the
        // column FFT in place (rfft2.go:22-74).
        RC_TRY(p->nd_row->forward(src, p->cols, dst, p->rcols, 0, 1.0, nullptr, Ctx{s, ws}));
        return run_nd_axes(p->axes, p->prec, htotal, dst, dst, 0, 1.0, s, ws);
      }
      // IRFFT2 (rfft2.go:87-138): copy zero-padded to rows x rcols per
      // array, column inverse, row IRFFT. `bins` complex values are supplied
      // per array.
      char* H = ws;
      char* rest = ws + (size_t)htotal * elt(p->prec);
      const i64 keep = std::min<i64>(bins, half);
      if (keep < half) HIP_TRY(launch_fill(p->prec, H, 2 * htotal, 0.0, s));
      if (keep > 0) HIP_TRY(launch_copy_lines(p->prec, src, H, keep, B, bins, half, s));
      RC_TRY(run_nd_axes(p->axes, p->prec, htotal, H, H, 1, 1.0 / (double)p->rows, s, rest));
      return p->nd_row->inverse(H, p->rcols, p->rcols, dst, p->cols, Ctx{s, rest});
    }
    case OP_VJP_ND_R2C: {
      // direction 0, y = RFFT2(x): gx = rows*cols * IRFFT2(w * gy) with the
      // r2c weights w over the bin axis (they commute with the column
      // transforms). direction 1, x = IRFFT2(X): gX = w' * RFFT2(gx) /
      // (rows*cols) with the c2r weights w'.
      if (B == 0) return GOFFT_OK;
      const i64 htotal = p->rows * p->rcols * B;
      if (!inv) {
        char* H = ws;
        char* rest = ws + (size_t)htotal * elt(p->prec);
        RC_TRY(run_nd_axes(p->axes, p->prec, htotal, src, H, 1, 1.0, s, rest));
        return p->nd_row->inverse(H, p->rcols, p->rcols, dst, p->cols, Ctx{s, rest}, (double)p->cols, 0.5);
      }
      const double in = 1.0 / ((double)p->rows * (double)p->cols);
      RC_TRY(p->nd_row->forward(src, p->cols, dst, p->rcols, 2, 0.0, nullptr, Ctx{s, ws}, in, 2.0 * in));
      return run_nd_axes(p->axes, p->prec, htotal, dst, dst, 0, 1.0, s, ws);
    }
  }
  return fail(GOFFT_ERR_INVALID_ARG, "unknown operation");
}

int exec_op(const gofft_plan* p, int op, const void* src, i64 bins, void* dst, int direction, void* stream,
            const WsSet* w = nullptr) {
  RC_TRY(check_op(p, op, src, bins, dst, direction));
  return run_op(p, op, src, bins, dst, direction, (hipStream_t)stream, w);
}

// Host-staged op: the caller's input goes through pinned staging to a device
// copy, the op runs on the session's own stream, and the output comes back
// the same way; the call waits for that stream only. Chunks of the caller's
// memory are copied into one staging buffer while the other's DMA runs.
const size_t kStageChunk = (size_t)4 << 20;

int host_op(const gofft_plan* p, int op, const void* src, i64 bins, void* dst, int direction) {
  RC_TRY(check_op(p, op, src, bins, dst, direction));
  size_t ib = 0, ob = 0;
  io_bytes(p, op, direction, bins, &ib, &ob);
  HostSession* h = nullptr;
  RC_TRY(p->host.acquire(&h));
  struct Release {
    const gofft_plan* p;
    HostSession* h;
    ~Release() {
      (void)hipStreamSynchronize(h->s);
      p->host.release(h);
    }
  } rel{p, h};
  if (h->din.bytes < ib) RC_TRY(h->din.alloc(ib));
  if (h->dout.bytes < ob) RC_TRY(h->dout.alloc(ob));
  const size_t chunk = std::min(kStageChunk, std::max(ib, ob));
  for (PinnedBuf& b : h->stage)
    if (chunk > 0 && b.bytes < chunk) RC_TRY(b.alloc(chunk));
  // in: fill staging buffer k % 2 once its previous DMA has finished
  int k = 0;
  for (size_t off = 0; off < ib; off += chunk, k++) {
    const size_t b = std::min(chunk, ib - off);
    const int j = k & 1;
    if (k >= 2) HIP_TRY(hipEventSynchronize(h->done[j]));
    std::memcpy(h->stage[j].p, (const char*)src + off, b);
    HIP_TRY(hipMemcpyAsync((char*)h->din.p + off, h->stage[j].p, b, hipMemcpyHostToDevice, h->s));
    HIP_TRY(hipEventRecord(h->done[j], h->s));
  }
  RC_TRY(run_op(p, op, ib ? h->din.p : nullptr, bins, ob ? h->dout.p : nullptr, direction, h->s, nullptr));
  // out: the DMA of chunk k + 1 runs while chunk k is copied out
  const size_t nout = chunk ? (ob + chunk - 1) / chunk : 0;
  auto issue = [&](size_t c) -> int {
    const size_t off = c * chunk, b = std::min(chunk, ob - off);
    HIP_TRY(hipMemcpyAsync(h->stage[c & 1].p, (const char*)h->dout.p + off, b, hipMemcpyDeviceToHost, h->s));
    HIP_TRY(hipEventRecord(h->done[c & 1], h->s));
    return GOFFT_OK;
  };
  if (nout > 0) RC_TRY(issue(0));
  for (size_t c = 0; c < nout; c++) {
    if (c + 1 < nout) RC_TRY(issue(c + 1));
    HIP_TRY(hipEventSynchronize(h->done[c & 1]));
    const size_t off = c * chunk, b = std::min(chunk, ob - off);
    std::memcpy((char*)dst + off, h->stage[c & 1].p, b);
  }
  HIP_TRY(hipStreamSynchronize(h->s));
  return GOFFT_OK;
}

}  // namespace

extern "C" {

int gofft_exec_c2c(const gofft_plan* p, const void* src, void* dst, int direction, void* stream) {
  return exec_op(p, OP_C2C, src, 0, dst, direction, stream);
}
int gofft_exec_r2c(const gofft_plan* p, const void* src, void* dst, void* stream) {
  return exec_op(p, OP_R2C, src, 0, dst, GOFFT_FORWARD, stream);
}
int gofft_exec_c2r(const gofft_plan* p, const void* src, int64_t bins, void* dst, void* stream) {
  return exec_op(p, OP_C2R, src, bins, dst, GOFFT_INVERSE, stream);
}
int gofft_exec_psd(const gofft_plan* p, const void* src, void* dst, void* stream) {
  return exec_op(p, OP_PSD, src, 0, dst, GOFFT_FORWARD, stream);
}
int gofft_exec_spectrogram(const gofft_plan* p, const void* src, void* dst, void* stream) {
  return exec_op(p, OP_SPECTROGRAM, src, 0, dst, GOFFT_FORWARD, stream);
}
int gofft_exec_nd_r2c(const gofft_plan* p, const void* src, int64_t bins, void* dst, int direction,
                      void* stream) {
  return exec_op(p, OP_ND_R2C, src, bins, dst, direction, stream);
}
int gofft_vjp_c2c(const gofft_plan* p, const void* gy, void* gx, int direction, void* stream) {
  return exec_op(p, OP_VJP_C2C, gy, 0, gx, direction, stream);
}
int gofft_vjp_r2c(const gofft_plan* p, const void* gy, void* gx, void* stream) {
  return exec_op(p, OP_VJP_R2C, gy, 0, gx, GOFFT_FORWARD, stream);
}
int gofft_vjp_c2r(const gofft_plan* p, const void* gx, void* gX, void* stream) {
  return exec_op(p, OP_VJP_C2R, gx, 0, gX, GOFFT_FORWARD, stream);
}
int gofft_vjp_nd_r2c(const gofft_plan* p, const void* g_out, void* g_in, int direction, void* stream) {
  return exec_op(p, OP_VJP_ND_R2C, g_out, 0, g_in, direction, stream);
}

// The complex transform with the plan's workspace in the device's static
// scratch instead of the per-stream pool: no allocation of any kind. A
// mixed-length plan's side lanes each get their own slice.
int gofft_exec_c2c_noalloc(const gofft_plan* p, const void* src, void* dst, int direction, void* stream) {
  RC_TRY(check_op(p, OP_C2C, src, 0, dst, direction));
  DevInfo dev;
  RC_TRY(current_device(&dev));
  const size_t slice = (p->ws_need + 255) / 256 * 256;
  const size_t sets = p->ws_need ? 1 + p->lanes.size() : 0;
  if (sets * slice > kScratchBytes)
    return fail(GOFFT_ERR_UNSUPPORTED, "plan needs " + std::to_string(sets * slice) +
                                           " bytes of workspace; gofft_exec_c2c_noalloc has " +
                                           std::to_string(kScratchBytes) + " bytes of scratch per device");
  WsSet w;
  w.main = (char*)dev.scratch;
  for (size_t i = 0; i < p->lanes.size(); i++) w.lanes.push_back((char*)dev.scratch + (i + 1) * slice);
  return run_op(p, OP_C2C, src, 0, dst, direction, (hipStream_t)stream, &w);
}

// The spectral filter loss and its gradient over an ND_R2C plan's arrays, in
// chunks of `chunk` arrays whose spectrum and residual are formed, used and
// dropped in turn (the forward recomputed per chunk inside the backward
// rather than stored for the whole batch).
int gofft_filter2_loss_grad(const gofft_plan* p, const void* x, const void* H, const void* y, void* loss, void* gx,
                            int64_t chunk, void* stream) {
  if (!p || p->kind != K_ND_R2C) return fail(GOFFT_ERR_INVALID_ARG, "plan is not an nd r2c plan");
  const i64 B = p->batch;
  if (B == 0) return GOFFT_OK;
  if (!x || !H || !y || !loss || !gx) return fail(GOFFT_ERR_INVALID_ARG, "null data pointer");
  hipStream_t s = (hipStream_t)stream;
  const size_t e = elt(p->prec), re = relt(p->prec);
  const i64 half = p->rows * p->rcols, per = p->rows * p->cols;
  const i64 parts = residual_parts(per);
  // spectrum, residual and the loss's partial sums of one array
  const size_t per_arr = (size_t)half * e + (size_t)per * re + (size_t)parts * re;
  i64 ch = chunk;
  if (ch <= 0) {
    // the largest chunk that fits 80% of the free device memory, at most all
    size_t fr = 0, tot = 0;
    HIP_TRY(hipMemGetInfo(&fr, &tot));
    const double room = 0.8 * (double)fr - (double)p->ws_engines;
    ch = room > 0 ? (i64)(room / (double)per_arr) : 1;
  }
  ch = std::max<i64>(1, std::min<i64>(ch, B));
  // the engines' workspace, then the chunk buffers (aligned to 256 bytes)
  const size_t eng = (p->ws_engines + 255) / 256 * 256;
  const size_t sb = ((size_t)ch * half * e + 255) / 256 * 256, rb = ((size_t)ch * per * re + 255) / 256 * 256;
  char* buf = nullptr;
  RC_TRY(p->loss_pool.get(s, eng + sb + rb + (size_t)ch * parts * re, &buf));
  char* ws = buf;
  char* S = buf + eng;
  char* r = S + sb;
  char* partial = r + rb;
  const NDAxis& col = p->axes[0];
  const double inv = 1.0 / ((double)p->rows * (double)p->cols);
  for (i64 b0 = 0; b0 < B; b0 += ch) {
    const i64 nb = std::min<i64>(ch, B - b0);
    const char* xc = (const char*)x + (size_t)(b0 * per) * re;
    const char* yc = (const char*)y + (size_t)(b0 * per) * re;
    char* gc = (char*)gx + (size_t)(b0 * per) * re;
    char* lc = (char*)loss + (size_t)b0 * re;
    // S = RFFT2(x) (rfft2.go:22-74)
    RC_TRY(p->nd_row->forward(xc, p->cols, S, p->rcols, 0, 1.0, nullptr, Ctx{s, ws}, 1.0, 1.0, nb * p->rows));
    RC_TRY(run_axis_blocks(col, p->prec, S, S, 0, 0, 1.0, nb, s, ws));
    // r = IRFFT2(S H) (rfft2.go:87-138)
    HIP_TRY(launch_filter_mul(p->prec, S, H, half, nb, 0, s));
    RC_TRY(run_axis_blocks(col, p->prec, S, S, 1, 1, 1.0 / (double)p->rows, nb, s, ws));
    RC_TRY(p->nd_row->inverse(S, p->rcols, p->rcols, r, p->cols, Ctx{s, ws}, 1.0, 1.0, nb * p->rows));
    // loss = ||r - y||^2 per array, r <- 2 (r - y)
    HIP_TRY(launch_residual_loss(p->prec, r, yc, per, nb, partial, lc, s));
    // gS = VJP of IRFFT2 at r, times conj(H)
    RC_TRY(p->nd_row->forward(r, p->cols, S, p->rcols, 2, 0.0, nullptr, Ctx{s, ws}, inv, 2.0 * inv, nb * p->rows));
    RC_TRY(run_axis_blocks(col, p->prec, S, S, 0, 0, 1.0, nb, s, ws));
    HIP_TRY(launch_filter_mul(p->prec, S, H, half, nb, 1, s));
    // gx = VJP of RFFT2 at gS
    RC_TRY(run_axis_blocks(col, p->prec, S, S, 1, 1, 1.0, nb, s, ws));
    RC_TRY(p->nd_row->inverse(S, p->rcols, p->rcols, gc, p->cols, Ctx{s, ws}, (double)p->cols, 0.5, nb * p->rows));
  }
  return GOFFT_OK;
}

// ---------------------------------------------------------------------------
// Host convenience calls (the reference's slice API).

int gofft_exec_c2c_host(const gofft_plan* p, const void* src, void* dst, int direction) {
  return host_op(p, OP_C2C, src, 0, dst, direction);
}
int gofft_exec_r2c_host(const gofft_plan* p, const void* src, void* dst) {
  return host_op(p, OP_R2C, src, 0, dst, GOFFT_FORWARD);
}
int gofft_exec_c2r_host(const gofft_plan* p, const void* src, int64_t bins, void* dst) {
  return host_op(p, OP_C2R, src, bins, dst, GOFFT_INVERSE);
}
int gofft_exec_nd_r2c_host(const gofft_plan* p, const void* src, int64_t bins, void* dst, int direction) {
  return host_op(p, OP_ND_R2C, src, bins, dst, direction);
}
int gofft_exec_psd_host(const gofft_plan* p, const void* src, void* dst) {
  return host_op(p, OP_PSD, src, 0, dst, GOFFT_FORWARD);
}
int gofft_exec_spectrogram_host(const gofft_plan* p, const void* src, void* dst) {
  return host_op(p, OP_SPECTROGRAM, src, 0, dst, GOFFT_FORWARD);
}
int gofft_vjp_c2c_host(const gofft_plan* p, const void* gy, void* gx, int direction) {
  return host_op(p, OP_VJP_C2C, gy, 0, gx, direction);
}
int gofft_vjp_r2c_host(const gofft_plan* p, const void* gy, void* gx) {
  return host_op(p, OP_VJP_R2C, gy, 0, gx, GOFFT_FORWARD);
}
int gofft_vjp_c2r_host(const gofft_plan* p, const void* gx, void* gX) {
  return host_op(p, OP_VJP_C2R, gx, 0, gX, GOFFT_FORWARD);
}
int gofft_vjp_nd_r2c_host(const gofft_plan* p, const void* g_out, void* g_in, int direction) {
  return host_op(p, OP_VJP_ND_R2C, g_out, 0, g_in, direction);
}

// ---------------------------------------------------------------------------
// Graphs: the launches (and, for mixed-length plans, the stream fork and
// join) of one exec on fixed buffers, replayed as one hipGraph. Each graph
// owns its workspace (one set per lane of a mixed-length plan), allocated at
// capture, so its replays do not touch the plan's per-stream workspaces.

struct gofft_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<std::unique_ptr<DevBuf>> ws;
  ~gofft_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

int gofft_graph_capture(gofft_graph** out, const gofft_plan* p, int op, const void* src, int64_t bins, void* dst,
                        int direction, void* stream) {
  if (!out) return fail(GOFFT_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  if (op < 0 || op >= OP_COUNT) return fail(GOFFT_ERR_INVALID_ARG, "unknown operation");
  RC_TRY(check_op(p, op, src, bins, dst, direction));
  hipStream_t s = (hipStream_t)stream;
  if (!s) return fail(GOFFT_ERR_INVALID_ARG, "capture needs a stream other than the null stream");
  auto g = std::make_unique<gofft_graph>();
  // the graph's own workspace, allocated before capture: no allocation is captured
  WsSet w;
  const size_t nws = p->ws_need ? 1 + p->lanes.size() : 0;
  for (size_t i = 0; i < nws; i++) {
    g->ws.push_back(std::make_unique<DevBuf>());
    RC_TRY(g->ws.back()->alloc(p->ws_need));
  }
  w.main = nws ? (char*)g->ws[0]->p : nullptr;
  for (size_t i = 0; i < p->lanes.size(); i++) w.lanes.push_back(nws ? (char*)g->ws[1 + i]->p : nullptr);
  HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const int rc = run_op(p, op, src, bins, dst, direction, s, &w);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(s, &graph);
  if (rc != GOFFT_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  HIP_TRY(ec);
  g->graph = graph;
  HIP_TRY(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
  *out = g.release();
  return GOFFT_OK;
}

int gofft_graph_c2c(gofft_graph** out, const gofft_plan* p, const void* src, void* dst, int direction,
                    void* stream) {
  return gofft_graph_capture(out, p, OP_C2C, src, 0, dst, direction, stream);
}

int gofft_graph_launch(const gofft_graph* g, void* stream) {
  if (!g) return fail(GOFFT_ERR_INVALID_ARG, "null graph");
  HIP_TRY(hipGraphLaunch(g->exec, (hipStream_t)stream));
  return GOFFT_OK;
}

int gofft_graph_destroy(gofft_graph* g) {
  delete g;
  return GOFFT_OK;
}

int gofft_sync(void) {
  HIP_TRY(hipDeviceSynchronize());
  return GOFFT_OK;
}

int gofft_probe_dft16(int method, const void* in, void* out, int64_t transforms, int reps, void* stream) {
  if (method != 0 && method != 1) return fail(GOFFT_ERR_INVALID_ARG, "method must be 0 (VALU) or 1 (matrix cores)");
  if (transforms < 0 || reps < 0) return fail(GOFFT_ERR_INVALID_ARG, "negative count");
  if (transforms == 0) return GOFFT_OK;
  if (!in || !out) return fail(GOFFT_ERR_INVALID_ARG, "null data pointer");
  if (transforms > ((int64_t)1 << 34)) return fail(GOFFT_ERR_INVALID_ARG, "too many transforms");
  // The A operands of the matrix-core form (kernels.hip k_probe_dft16_mfma):
  // lane l, slice q: W[l % 32][row(q, l / 32)], row(q, h) = 8*(q/4) + 4h + q%4,
  // W the unitary 16-point forward DFT as a real 32 x 32 matrix.
  static std::mutex mu;
  static std::map<int, std::unique_ptr<DevBuf>> tabs;
  DevInfo dev;
  RC_TRY(current_device(&dev));
  const float* w = nullptr;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto& b = tabs[dev.id];
    if (!b) {
      std::vector<float> h(64 * 16);
      for (int l = 0; l < 64; l++)
        for (int q = 0; q < 16; q++) {
          const int i = l % 32, c = 8 * (q / 4) + 4 * (l / 32) + q % 4;
          const int k = i % 16, nn = c % 16;
          const cld wr = root((i64)k * nn, 16);  // exp(-2 pi i k n / 16)
          const long double cs = wr.real(), sn = -wr.imag();
          long double v;
          if (i < 16) v = c < 16 ? cs : sn;  // Re Y = cos Re x + sin Im x
          else v = c < 16 ? -sn : cs;        // Im Y = -sin Re x + cos Im x
          h[l * 16 + q] = (float)(v * 0.25L);
        }
      auto nb = std::make_unique<DevBuf>();
      RC_TRY(nb->upload(h.data(), h.size() * sizeof(float)));
      b = std::move(nb);
    }
    w = (const float*)b->p;
  }
  HIP_TRY(launch_probe_dft16(method, in, out, transforms, reps, w, (hipStream_t)stream));
  return GOFFT_OK;
}

}  // extern "C"
