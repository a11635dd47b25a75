// CPU-only test driver of the library's host code (plan.cpp) against the
// host stand-in runtime (hipsim.cpp), built with AddressSanitizer and
// UndefinedBehaviorSanitizer. It builds and runs plans of every kind and
// many shapes (every launch checked against the kernels' assumptions),
// drives the out-of-memory and launch-failure paths by injected failures and
// checks their status codes and messages, feeds the autotuner a corrupt
// cache file, and checks the argument errors. Prints "ok NAME" or
// "FAIL NAME: why" per scenario; exits non-zero on any failure.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gofft.h"

extern "C" {
void hipsim_fail_alloc(long n);
void hipsim_fail_host_alloc(long n);
void hipsim_fail_launch(long n);
void hipsim_reset_counts();
long hipsim_allocs();
long hipsim_launches();
long hipsim_nt_launches();
long hipsim_live();
const char* hipsim_bad();
void hipsim_region(void* p, size_t n);
void hipsim_unregion(void* p);
}

static int g_failures = 0;

static void report(const std::string& name, const std::string& why) {
  if (why.empty()) {
    std::printf("ok %s\n", name.c_str());
  } else {
    std::printf("FAIL %s: %s\n", name.c_str(), why.c_str());
    g_failures++;
  }
}

static std::string err() { return gofft_last_error(); }

// A "device" buffer of the simulated runtime (host memory).
struct Buf {
  void* p = nullptr;
  explicit Buf(size_t bytes) {
    p = std::calloc(1, bytes + 64);
    hipsim_region(p, bytes + 64);
  }
  ~Buf() {
    hipsim_unregion(p);
    std::free(p);
  }
};

static size_t elt(int prec) { return prec == GOFFT_C128 ? 16 : 8; }

// Runs every op a plan supports on zeroed buffers (and as a graph, and the
// host-staged form); returns "" or what failed.
static std::string exercise(gofft_plan* p, int kind, int prec) {
  const size_t e = elt(prec), r = e / 2;
  const int64_t in0 = gofft_plan_in_elems(p, 0), out0 = gofft_plan_out_elems(p, 0);
  const int64_t in1 = gofft_plan_in_elems(p, 1), out1 = gofft_plan_out_elems(p, 1);
  const size_t big = (size_t)std::max<int64_t>(std::max(in0, out0), std::max(in1, out1)) * e + 64;
  Buf a(big), b(big), c(big);
  void* s = (void*)0x10;  // any stream token
  auto rc = [&](int v, const char* what) -> std::string {
    return v == GOFFT_OK ? "" : std::string(what) + ": " + err();
  };
  std::string w;
  gofft_graph* g = nullptr;
  switch (kind) {
    case 0:  // complex
      for (int d = 0; d < 2 && w.empty(); d++) {
        w = rc(gofft_exec_c2c(p, a.p, b.p, d, s), "exec_c2c");
        if (w.empty()) w = rc(gofft_exec_c2c(p, a.p, a.p, d, s), "exec_c2c in place");
        if (w.empty()) w = rc(gofft_vjp_c2c(p, a.p, b.p, d, s), "vjp_c2c");
        if (w.empty()) w = rc(gofft_exec_c2c_host(p, a.p, b.p, d), "exec_c2c_host");
        if (w.empty()) w = rc(gofft_graph_capture(&g, p, GOFFT_OP_C2C, a.p, 0, b.p, d, s), "graph");
        if (g) {
          if (w.empty()) w = rc(gofft_graph_launch(g, s), "graph launch");
          gofft_graph_destroy(g);
          g = nullptr;
        }
        if (w.empty() && gofft_plan_workspace_bytes(p) <= (1 << 24)) w = rc(gofft_exec_c2c_noalloc(p, a.p, b.p, d, s), "noalloc");
      }
      break;
    case 1:  // real
      w = rc(gofft_exec_r2c(p, a.p, b.p, s), "r2c");
      if (w.empty()) w = rc(gofft_exec_c2r(p, a.p, in1 ? 3 : 0, b.p, s), "c2r short");
      if (w.empty()) w = rc(gofft_vjp_r2c(p, a.p, b.p, s), "vjp_r2c");
      if (w.empty()) w = rc(gofft_vjp_c2r(p, a.p, b.p, s), "vjp_c2r");
      if (w.empty()) w = rc(gofft_exec_r2c_host(p, a.p, b.p), "r2c_host");
      if (w.empty()) w = rc(gofft_graph_capture(&g, p, GOFFT_OP_C2R, a.p, 2, b.p, 1, s), "graph c2r");
      if (g) gofft_graph_destroy(g);
      break;
    case 2:  // psd
      w = rc(gofft_exec_psd(p, a.p, b.p, s), "psd");
      if (w.empty()) w = rc(gofft_exec_psd_host(p, a.p, b.p), "psd_host");
      break;
    case 3:  // spectrogram
      w = rc(gofft_exec_spectrogram(p, a.p, b.p, s), "spectrogram");
      if (w.empty()) w = rc(gofft_exec_spectrogram_host(p, a.p, b.p), "spectrogram_host");
      break;
    case 4:  // nd r2c
      w = rc(gofft_exec_nd_r2c(p, a.p, 0, b.p, 0, s), "rfft2");
      if (w.empty()) w = rc(gofft_exec_nd_r2c(p, a.p, in1 / 2, b.p, 1, s), "irfft2 short");
      if (w.empty()) w = rc(gofft_vjp_nd_r2c(p, a.p, b.p, 0, s), "vjp rfft2");
      if (w.empty()) w = rc(gofft_vjp_nd_r2c(p, a.p, b.p, 1, s), "vjp irfft2");
      if (w.empty()) w = rc(gofft_exec_nd_r2c_host(p, a.p, 0, b.p, 0), "rfft2_host");
      if (w.empty()) w = rc(gofft_filter2_loss_grad(p, a.p, b.p, a.p, c.p, b.p, 2, s), "filter loss");
      break;
  }
  (void)r;
  if (w.empty() && *hipsim_bad()) w = std::string("launch check: ") + hipsim_bad();
  return w;
}

static void sweep() {
  const std::vector<int64_t> lens = {1, 2, 3, 8, 16, 17, 60, 64, 100, 128, 641, 769, 1000, 1008, 1009, 1024, 1296, 2017,
                                     4096, 5003, 6720, 8192, 9973, 10006, 10007, 15015, 16384, 20011, 20014, 20160,
                                     40009, 65536, 1 << 18, 600011, 600014, 1 << 20};
  for (int prec : {GOFFT_C128, GOFFT_C64})
    for (int64_t n : lens)
      for (int64_t batch : {1, 3, 64}) {
        if (n * batch > (1 << 22)) continue;
        const std::string name = "c2c n=" + std::to_string(n) + " batch=" + std::to_string(batch) + " prec=" +
                                 std::to_string(prec);
        gofft_plan* p = nullptr;
        if (gofft_plan_c2c(&p, n, batch, prec, 0) != GOFFT_OK) {
          report(name, "plan: " + err());
          continue;
        }
        std::string w = exercise(p, 0, prec);
        gofft_plan_destroy(p);
        report(name, w);
      }
  {
    std::vector<int64_t> mix = {1024, 4096, 65536, 6720, 1009, 10007, 20014, 1009, 4096, 600011, 17, 1, 0};
    for (int prec : {GOFFT_C128, GOFFT_C64}) {
      gofft_plan* p = nullptr;
      std::string w = gofft_plan_c2c_many(&p, mix.data(), (int64_t)mix.size(), prec, 0) == GOFFT_OK ? exercise(p, 0, prec)
                                                                                                   : "plan: " + err();
      if (p) gofft_plan_destroy(p);
      report("many prec=" + std::to_string(prec), w);
    }
  }
  for (int prec : {GOFFT_C128, GOFFT_C64})
    for (int64_t n : {1, 2, 7, 8, 30, 1024, 2018, 4096, 20014, 1 << 17, 1 << 20})
      for (int64_t batch : {1, 5}) {
        gofft_plan* p = nullptr;
        const std::string name = "r2c n=" + std::to_string(n) + " batch=" + std::to_string(batch);
        std::string w = gofft_plan_r2c(&p, n, batch, prec, 0) == GOFFT_OK ? exercise(p, 1, prec) : "plan: " + err();
        if (p) gofft_plan_destroy(p);
        report(name, w);
        p = nullptr;
        w = gofft_plan_psd(&p, n, batch, 0.5, prec, 0) == GOFFT_OK ? exercise(p, 2, prec) : "plan: " + err();
        if (p) gofft_plan_destroy(p);
        report("psd n=" + std::to_string(n) + " batch=" + std::to_string(batch), w);
      }
  for (int prec : {GOFFT_C128, GOFFT_C64})
    for (int64_t seg : {1, 8, 64, 1024, 4097}) {
      std::vector<double> win(seg, 1.0);
      gofft_plan* p = nullptr;
      std::string w = gofft_plan_spectrogram(&p, 1 << 16, seg, seg / 2, win.data(), seg, 1.0, prec, 0) == GOFFT_OK
                          ? exercise(p, 3, prec)
                          : "plan: " + err();
      if (p) gofft_plan_destroy(p);
      report("spectrogram seg=" + std::to_string(seg), w);
    }
  const std::vector<std::vector<int64_t>> shapes = {{}, {1}, {5}, {4, 6}, {16, 1009}, {64, 64, 20}, {256, 10007},
                                                    {3, 1, 2}, {8, 600011}, {100, 100, 100}, {1024, 1024}};
  for (int prec : {GOFFT_C128, GOFFT_C64})
    for (const auto& sh : shapes) {
      int64_t tot = 1;
      for (int64_t v : sh) tot *= v;
      if (tot > (1 << 23)) continue;
      std::string name = "nd c2c rank " + std::to_string(sh.size()) + " total " + std::to_string(tot);
      gofft_plan* p = nullptr;
      std::string w = gofft_plan_nd(&p, sh.data(), (int)sh.size(), GOFFT_ND_C2C, 2, prec, 0) == GOFFT_OK
                          ? exercise(p, 0, prec)
                          : "plan: " + err();
      if (p) gofft_plan_destroy(p);
      report(name, w);
    }
  for (int prec : {GOFFT_C128, GOFFT_C64})
    for (auto sh : std::vector<std::vector<int64_t>>{{1, 1}, {6, 1}, {1, 6}, {32, 64}, {384, 1000}, {8192, 64}, {1024, 1024}}) {
      gofft_plan* p = nullptr;
      std::string w = gofft_plan_nd(&p, sh.data(), 2, GOFFT_ND_R2C, 3, prec, 0) == GOFFT_OK ? exercise(p, 4, prec)
                                                                                          : "plan: " + err();
      if (p) gofft_plan_destroy(p);
      report("nd r2c " + std::to_string(sh[0]) + "x" + std::to_string(sh[1]), w);
    }
}

// Every allocation a plan's creation (and first exec) makes fails in turn:
// GOFFT_ERR_OOM with the runtime's message, no plan, nothing leaked.
static void oom_paths() {
  struct Case {
    const char* name;
    int64_t n, batch;
    int prec;
  };
  for (Case c : {Case{"rader chain", 10007, 4, GOFFT_C128}, Case{"bluestein chain", 20014, 3, GOFFT_C64},
                 Case{"two-pass", 65536, 2, GOFFT_C64}, Case{"fused rader", 1009, 8, GOFFT_C64}}) {
    hipsim_reset_counts();
    gofft_plan* p = nullptr;
    gofft_plan_c2c(&p, c.n, c.batch, c.prec, 0);
    const long allocs = hipsim_allocs();
    gofft_plan_destroy(p);
    std::string w;
    for (long k = 1; k <= allocs && w.empty(); k++) {
      hipsim_reset_counts();
      hipsim_fail_alloc(k);
      gofft_plan* q = (gofft_plan*)0x1;
      const int rc = gofft_plan_c2c(&q, c.n, c.batch, c.prec, 0);
      hipsim_fail_alloc(0);
      if (rc != GOFFT_ERR_OOM) w = "allocation " + std::to_string(k) + ": status " + std::to_string(rc);
      else if (err().find("hipMalloc of ") != 0) w = "allocation " + std::to_string(k) + ": message '" + err() + "'";
      else if (q != nullptr) w = "a plan came back";
      if (q && q != (gofft_plan*)0x1) gofft_plan_destroy(q);
    }
    const long live = hipsim_live();
    if (w.empty() && live != 0) w = std::to_string(live) + " device allocations still live";
    report(std::string("oom during creation: ") + c.name + " (" + std::to_string(allocs) + " allocations)", w);
  }
  // the first exec on a new stream allocates its workspace
  gofft_plan* p = nullptr;
  gofft_plan_c2c(&p, 65536, 2, GOFFT_C64, 0);
  Buf a(65536 * 2 * 8), b(65536 * 2 * 8);
  hipsim_reset_counts();
  hipsim_fail_alloc(1);
  int rc = gofft_exec_c2c(p, a.p, b.p, GOFFT_FORWARD, (void*)0x20);
  hipsim_fail_alloc(0);
  std::string w = rc == GOFFT_ERR_OOM && err().find("hipMalloc of ") == 0 ? "" : "exec: status " + std::to_string(rc) + " '" + err() + "'";
  if (w.empty() && gofft_exec_c2c(p, a.p, b.p, GOFFT_FORWARD, (void*)0x20) != GOFFT_OK) w = "retry failed: " + err();
  // pinned staging of a host call
  hipsim_fail_host_alloc(1);
  rc = gofft_exec_c2c_host(p, a.p, b.p, GOFFT_FORWARD);
  hipsim_fail_host_alloc(0);
  if (w.empty() && !(rc == GOFFT_ERR_OOM && err().find("hipHostMalloc of ") == 0))
    w = "host staging: status " + std::to_string(rc) + " '" + err() + "'";
  if (w.empty() && gofft_exec_c2c_host(p, a.p, b.p, GOFFT_FORWARD) != GOFFT_OK) w = "host retry: " + err();
  gofft_plan_destroy(p);
  report("oom during exec (workspace, pinned staging)", w);
}

// Each launch of an exec fails in turn: GOFFT_ERR_HIP naming the launcher
// and the runtime's error.
static void launch_failures() {
  for (int64_t n : {10007, 20014, 65536, 4096}) {
    gofft_plan* p = nullptr;
    gofft_plan_c2c(&p, n, 3, GOFFT_C64, 0);
    Buf a(n * 3 * 8), b(n * 3 * 8);
    const int launches = gofft_plan_num_launches(p, 0);
    std::string w;
    for (int k = 1; k <= launches && w.empty(); k++) {
      hipsim_reset_counts();
      hipsim_fail_launch(k);
      const int rc = gofft_exec_c2c(p, a.p, b.p, GOFFT_FORWARD, (void*)0x30);
      hipsim_fail_launch(0);
      if (rc != GOFFT_ERR_HIP) w = "launch " + std::to_string(k) + ": status " + std::to_string(rc);
      else if ((err().find("launch_") == std::string::npos && err().find("issue_") == std::string::npos) ||
               err().find("hipErrorLaunchFailure") == std::string::npos)
        w = "message '" + err() + "'";
    }
    gofft_plan_destroy(p);
    report("launch failure n=" + std::to_string(n) + " (" + std::to_string(launches) + " launches)", w);
  }
  // a work-queue plan: its setup and queue launches
  for (int prec : {GOFFT_C128, GOFFT_C64}) {
    std::vector<int64_t> mix = {1009, 20014, 65536, 4096, 10007};
    gofft_plan* p = nullptr;
    gofft_plan_c2c_many(&p, mix.data(), (int64_t)mix.size(), prec, GOFFT_FLAG_QUEUE);
    Buf a(200000 * 16), b(200000 * 16);
    const int launches = gofft_plan_num_launches(p, 0);
    std::string w;
    for (int k = 1; k <= launches && w.empty(); k++) {
      hipsim_reset_counts();
      hipsim_fail_launch(k);
      const int rc = gofft_exec_c2c(p, a.p, b.p, GOFFT_FORWARD, (void*)0x30);
      hipsim_fail_launch(0);
      if (rc != GOFFT_ERR_HIP) w = "launch " + std::to_string(k) + ": status " + std::to_string(rc);
      else if (err().find("hipErrorLaunchFailure") == std::string::npos) w = "message '" + err() + "'";
    }
    gofft_plan_destroy(p);
    report("launch failure queue prec=" + std::to_string(prec) + " (" + std::to_string(launches) + " launches)", w);
  }
}

// The autotuner's cache file with entries the planner could not have made.
static void corrupt_tune_cache() {
  const char* path = "_build/tune_cache.txt";
  FILE* f = std::fopen(path, "w");
  std::fprintf(f, "gofft-tune-v2 0.2 gfx950:sim 3 10007 4 1 1000 16\n");    // cl < 2q
  std::fprintf(f, "gofft-tune-v2 0.2 gfx950:sim 3 10007 5 1 40000 16\n");   // outside the window
  std::fprintf(f, "gofft-tune-v2 0.2 gfx950:sim 3 10007 6 1 20160 5\n");    // not a column candidate
  std::fprintf(f, "gofft-tune-v2 0.1 gfx950:sim 4 20014 4 1 40960 16\n");   // another version
  std::fprintf(f, "gofft-tune-v2 0.2 gfx950:sim 4 20014 5 1 1 16\n");       // m < 2n - 1
  std::fprintf(f, "garbage\n3 10007 7 1 10 16\n");
  std::fclose(f);
  setenv("GOFFT_TUNE_CACHE", path, 1);
  std::string w;
  for (int64_t batch : {4, 5, 6, 7}) {
    gofft_plan* p = nullptr;
    if (gofft_plan_c2c(&p, 10007, batch, GOFFT_C64, GOFFT_FLAG_AUTOTUNE) != GOFFT_OK) w = "rader: " + err();
    if (p) gofft_plan_destroy(p);
    p = nullptr;
    if (w.empty() && gofft_plan_c2c(&p, 20014, batch, GOFFT_C64, GOFFT_FLAG_AUTOTUNE) != GOFFT_OK) w = "bluestein: " + err();
    if (p) gofft_plan_destroy(p);
  }
  unsetenv("GOFFT_TUNE_CACHE");
  if (w.empty() && *hipsim_bad()) w = std::string("launch check: ") + hipsim_bad();
  report("corrupt autotuner cache", w);
}

static void argument_errors() {
  std::string w;
  auto expect = [&](int rc, int want, const char* text, const char* what) {
    if (!w.empty()) return;
    if (rc != want || err().find(text) == std::string::npos)
      w = std::string(what) + ": status " + std::to_string(rc) + " '" + err() + "'";
  };
  gofft_plan* p = nullptr;
  expect(gofft_plan_c2c(nullptr, 8, 1, GOFFT_C64, 0), GOFFT_ERR_INVALID_ARG, "null plan pointer", "null out");
  expect(gofft_plan_c2c(&p, -1, 1, GOFFT_C64, 0), GOFFT_ERR_INVALID_ARG, "negative", "negative n");
  expect(gofft_plan_c2c(&p, 8, 1, 7, 0), GOFFT_ERR_INVALID_ARG, "dtype", "dtype");
  const int64_t bad_shape[2] = {4, 0};
  expect(gofft_plan_nd(&p, bad_shape, 2, GOFFT_ND_C2C, 1, GOFFT_C128, 0), GOFFT_ERR_SHAPE,
         "fft: shape lengths must be positive", "shape");
  std::vector<double> win(4, 1.0);
  expect(gofft_plan_spectrogram(&p, 100, 0, 0, win.data(), 4, 1.0, GOFFT_C128, 0), GOFFT_ERR_INVALID_ARG,
         "fft: segment length must be positive", "segment");
  expect(gofft_plan_spectrogram(&p, 100, 4, 0, win.data(), 3, 1.0, GOFFT_C128, 0), GOFFT_ERR_INVALID_ARG,
         "fft: window length must equal segment length", "window");
  expect(gofft_plan_spectrogram(&p, 100, 4, 4, win.data(), 4, 1.0, GOFFT_C128, 0), GOFFT_ERR_INVALID_ARG,
         "fft: overlap must satisfy 0 <= overlap < segment", "overlap");
  gofft_plan* c = nullptr;
  gofft_plan_c2c(&c, 1024, 2, GOFFT_C64, 0);
  Buf a(1 << 16), b(1 << 16);
  expect(gofft_exec_c2c(c, nullptr, b.p, 0, nullptr), GOFFT_ERR_INVALID_ARG, "null data pointer", "null src");
  expect(gofft_exec_c2c(c, a.p, b.p, 2, nullptr), GOFFT_ERR_INVALID_ARG, "direction", "direction");
  expect(gofft_exec_r2c(c, a.p, b.p, nullptr), GOFFT_ERR_INVALID_ARG, "not an r2c plan", "kind");
  expect(gofft_exec_c2c_noalloc(c, a.p, nullptr, 0, nullptr), GOFFT_ERR_INVALID_ARG, "null data pointer", "noalloc");
  gofft_plan* r = nullptr;
  gofft_plan_r2c(&r, 1024, 4, GOFFT_C64, 0);
  expect(gofft_exec_r2c(r, a.p, a.p, nullptr), GOFFT_ERR_INVALID_ARG, "must not overlap", "alias");
  expect(gofft_exec_c2r(r, a.p, -1, b.p, nullptr), GOFFT_ERR_INVALID_ARG, "negative bin count", "bins");
  expect(gofft_plan_index_table(r, 1024, GOFFT_TABLE_RADER_PERM, nullptr, 0), GOFFT_ERR_INVALID_ARG, "", "table");
  gofft_plan* big = nullptr;
  gofft_plan_c2c(&big, 65536, 512, GOFFT_C64, 0);
  Buf x((size_t)65536 * 512 * 8);
  expect(gofft_exec_c2c_noalloc(big, x.p, x.p, 0, nullptr), GOFFT_ERR_UNSUPPORTED, "bytes of scratch", "scratch");
  gofft_plan_destroy(big);
  gofft_plan_destroy(r);
  gofft_plan_destroy(c);
  report("argument errors", w);
}

// The integer tables the kernels index with, read back through the ABI:
// g^k mod n and the inverse order, and j^2 mod 2n, in exact arithmetic.
static void index_tables() {
  std::string w;
  for (int64_t n : {1009, 10007, 600011}) {
    gofft_plan* p = nullptr;
    gofft_plan_c2c(&p, n, 1, GOFFT_C64, 0);
    int64_t g = 0;
    gofft_plan_index_table(p, n, GOFFT_TABLE_RADER_GENERATOR, &g, 1);
    std::vector<int64_t> perm(n - 1), iperm(n - 1);
    gofft_plan_index_table(p, n, GOFFT_TABLE_RADER_PERM, perm.data(), n - 1);
    gofft_plan_index_table(p, n, GOFFT_TABLE_RADER_IPERM, iperm.data(), n - 1);
    int64_t x = 1;
    for (int64_t k = 0; k < n - 1 && w.empty(); k++) {
      if (perm[k] != x) w = "perm";
      if (iperm[k] != perm[(n - 1 - k) % (n - 1)]) w = "iperm";
      x = x * g % n;
      if (k < n - 2 && x == 1) w = "g is not a generator";
    }
    gofft_plan_destroy(p);
  }
  for (int64_t n : {641, 20014}) {
    gofft_plan* p = nullptr;
    gofft_plan_c2c(&p, n, 1, GOFFT_C128, 0);
    std::vector<int64_t> jj(n);
    if (gofft_plan_index_table(p, n, GOFFT_TABLE_CHIRP_INDEX, jj.data(), n) != n) w = "chirp table length";
    for (int64_t j = 0; j < n && w.empty(); j++)
      if (jj[j] != (int64_t)((__int128)j * j % (2 * n))) w = "chirp index";
    gofft_plan_destroy(p);
  }
  report("index tables", w);
}

// Mixed-length plans with the work queue: every op (the queue launchers of
// the stand-in resolve each exec's descriptors and check every stage), the
// queue's launch count and plan text, groups the queue cannot take (complex64
// twiddle-form flags) staying on the lanes, and every allocation of the
// plan's creation failing in turn without a leak.
static void queue_plans() {
  std::vector<int64_t> mix;
  for (int r = 0; r < 40; r++)
    for (int64_t n : {1024, 4096, 65536, 6720, 1009, 10007, 20014}) mix.push_back(n);
  mix.push_back(600011);
  mix.push_back(17);
  mix.push_back(1);
  mix.push_back(0);
  for (int prec : {GOFFT_C128, GOFFT_C64})
    for (unsigned extra : {0u, (unsigned)GOFFT_FLAG_TWIDDLES_FP32}) {
      if (extra && prec == GOFFT_C128) continue;
      const std::string name = "queue prec=" + std::to_string(prec) + (extra ? " fp32 twiddles" : "");
      gofft_plan* p = nullptr;
      if (gofft_plan_c2c_many(&p, mix.data(), (int64_t)mix.size(), prec, GOFFT_FLAG_QUEUE | extra) != GOFFT_OK) {
        report(name, "plan: " + err());
        continue;
      }
      std::vector<char> text(1 << 16);
      gofft_plan_describe(p, text.data(), (int64_t)text.size());
      const std::string t(text.data());
      std::string w;
      if (t.find("queue threads=") == std::string::npos || t.find("queued: ") == std::string::npos)
        w = "no queue in the plan: " + t;
      // the queue takes the Stockham, Rader and Bluestein groups; with FP32
      // twiddles the multi-pass groups (65536, 600011) stay on the lanes
      if (w.empty() && extra && t.find("queued: stockham n=65536") != std::string::npos) w = "65536 queued with FP32 twiddles";
      if (w.empty() && !extra && t.find("queued: stockham n=65536") == std::string::npos) w = "65536 not queued";
      if (w.empty() && t.find("trivial n=1") != std::string::npos && t.find("queued: trivial") != std::string::npos)
        w = "trivial group queued";
      if (w.empty()) {
        hipsim_reset_counts();
        void* s = (void*)0x10;
        Buf a(sizeof(double) * 2 * 8000000), b(sizeof(double) * 2 * 8000000);
        if (gofft_exec_c2c(p, a.p, b.p, GOFFT_FORWARD, s) != GOFFT_OK) w = "exec: " + err();
        else if (hipsim_launches() != gofft_plan_num_launches(p, GOFFT_FORWARD))
          w = "launches " + std::to_string(hipsim_launches()) + " != " + std::to_string(gofft_plan_num_launches(p, 0));
      }
      if (w.empty()) w = exercise(p, 0, prec);
      gofft_plan_destroy(p);
      report(name, w);
    }
  for (int prec : {GOFFT_C128, GOFFT_C64}) {
    std::vector<int64_t> small = {1009, 20014, 65536, 4096, 1009, 10007};
    hipsim_reset_counts();
    gofft_plan* p = nullptr;
    gofft_plan_c2c_many(&p, small.data(), (int64_t)small.size(), prec, GOFFT_FLAG_QUEUE);
    if (p) gofft_plan_destroy(p);
    const long total = hipsim_allocs();
    std::string w;
    for (long k = 1; k <= total && w.empty(); k++) {
      hipsim_reset_counts();
      const long live0 = hipsim_live();
      hipsim_fail_alloc(k);
      p = nullptr;
      const int rc = gofft_plan_c2c_many(&p, small.data(), (int64_t)small.size(), prec, GOFFT_FLAG_QUEUE);
      hipsim_fail_alloc(0);
      if (rc != GOFFT_ERR_OOM || p) w = "allocation " + std::to_string(k) + ": status " + std::to_string(rc);
      else if (hipsim_live() != live0) w = "allocation " + std::to_string(k) + ": leaked";
      if (p) gofft_plan_destroy(p);
    }
    report("queue oom prec=" + std::to_string(prec) + " (" + std::to_string(total) + " allocations)", w);
  }
}

// The radix-16 stage probe: argument errors, and one launch per method.
// (It keeps its 4 KB matrix table per device until the process ends.)
static void probe_calls() {
  Buf in(4096 * 16 * 8), out(4096 * 16 * 8);
  std::string w;
  if (gofft_probe_dft16(2, in.p, out.p, 4096, 1, nullptr) != GOFFT_ERR_INVALID_ARG) w = "method 2 accepted";
  if (w.empty() && gofft_probe_dft16(0, nullptr, out.p, 4096, 1, nullptr) != GOFFT_ERR_INVALID_ARG)
    w = "null input accepted";
  if (w.empty() && gofft_probe_dft16(1, in.p, out.p, -1, 1, nullptr) != GOFFT_ERR_INVALID_ARG)
    w = "negative count accepted";
  if (w.empty() && gofft_probe_dft16(1, in.p, out.p, 0, 1, nullptr) != GOFFT_OK) w = "empty call: " + err();
  for (int m = 0; m < 2 && w.empty(); m++) {
    hipsim_reset_counts();
    if (gofft_probe_dft16(m, in.p, out.p, 4096, 3, nullptr) != GOFFT_OK) w = "method " + std::to_string(m) + ": " + err();
    else if (hipsim_launches() != 1) w = "launches " + std::to_string(hipsim_launches());
  }
  if (w.empty() && *hipsim_bad()) w = hipsim_bad();
  report("probe dft16", w);
}

int main() {
  gofft_init();
  sweep();
  oom_paths();
  launch_failures();
  corrupt_tune_cache();
  argument_errors();
  index_tables();
  queue_plans();
  probe_calls();  // last: its table stays allocated for the process (the leak checks count from zero)
  // the sweep's aligned workspaces and buffers took the non-temporal forms
  // (each such launch was checked above)
  report("non-temporal launches seen", hipsim_nt_launches() > 0 ? "" : "none");
  std::printf("%s: %d failed\n", g_failures ? "FAILED" : "PASSED", g_failures);
  return g_failures ? 1 : 0;
}
