// TEST INFRASTRUCTURE ONLY — the parity oracle and the timed CPU baseline.
//
// A plain C++ FP64 restatement of go-fft's CPU transform path (the reference is
// pure Go, and no Go toolchain exists here, so the reference itself cannot be
// built or run). Nothing in the product library links, calls or executes this
// file: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// load liboracle.so, and only as the checker or the CPU baseline.
//
// Every function cites the reference file:line it restates. It is compiled with
// -ffp-contract=off so that every product and sum rounds separately, the way
// amd64 Go (GOAMD64=v1) evaluates complex128 arithmetic.
//
// Pinning (see tests/test_oracle_*.py): integer helpers against the reference's
// known-answer tests; transforms against a long-double naive DFT, numpy
// (pocketfft) FP64 and the reference's closed-form tests; the recurrence
// twiddles of kernels.Radix2 against correctly-rounded cos/sin.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace oracle {

typedef long long i64;

struct cx {
  double re, im;
};
static inline cx mk(double r, double i) { cx z; z.re = r; z.im = i; return z; }
static inline cx add(cx a, cx b) { return mk(a.re + b.re, a.im + b.im); }
static inline cx sub(cx a, cx b) { return mk(a.re - b.re, a.im - b.im); }
// Go complex128 product: (ar*br - ai*bi) + (ar*bi + ai*br)i, each term rounded.
static inline cx mul(cx a, cx b) {
  return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline cx conj(cx a) { return mk(a.re, -a.im); }
static const double kPi = 3.14159265358979323846264338327950288;

// ---------------------------------------------------------------------------
// Go math.Sin / math.Cos (Go standard library, math/sin.go; the algorithm is
// Cephes sin.c by S. L. Moshier). Used only where the reference's output depends
// on the exact bits of one cos/sin pair: the per-stage `wlen` of kernels.Radix2
// (internal/kernels/kernels.go:45-46), which the twiddle recurrence amplifies.
// Valid for |x| < 2^29 (the reference only passes |x| <= pi).
static const double kSin[6] = {
    1.58962301576546568060e-10, -2.50507477628578072866e-8,
    2.75573136213857245213e-6,  -1.98412698295895385996e-4,
    8.33333333332211858878e-3,  -1.66666666666666307295e-1};
static const double kCos[6] = {
    -1.13585365213876817300e-11, 2.08757008419747316778e-9,
    -2.75573141792967388112e-7,  2.48015872888517045348e-5,
    -1.38888888888730564116e-3,  4.16666666666665929218e-2};
static const double PI4A = 7.85398125648498535156e-1;
static const double PI4B = 3.77489470793079817668e-8;
static const double PI4C = 2.69515142907905952645e-15;

static void go_sincos(double x, double* s_out, double* c_out) {
  bool sin_sign = false, cos_sign = false;
  if (x < 0) { x = -x; sin_sign = true; }
  // Go folds the constant 4/Pi exactly, then rounds it once.
  const double fourOverPi = 1.27323954473516268615107010698011489627567716592365;
  uint64_t j = (uint64_t)(x * fourOverPi);
  double y = (double)j;
  if (j & 1) { j++; y++; }
  j &= 7;
  if (j > 3) { j -= 4; sin_sign = !sin_sign; cos_sign = !cos_sign; }
  if (j > 1) cos_sign = !cos_sign;
  double z = ((x - y * PI4A) - y * PI4B) - y * PI4C;
  double zz = z * z;
  double c = 1.0 - 0.5 * zz +
             zz * zz * ((((((kCos[0] * zz) + kCos[1]) * zz + kCos[2]) * zz + kCos[3]) * zz + kCos[4]) * zz + kCos[5]);
  double s = z + z * zz * ((((((kSin[0] * zz) + kSin[1]) * zz + kSin[2]) * zz + kSin[3]) * zz + kSin[4]) * zz + kSin[5]);
  if (j == 1 || j == 2) { double t = s; s = c; c = t; }
  if (cos_sign) c = -c;
  if (sin_sign) s = -s;
  *s_out = s;
  *c_out = c;
}

// ---------------------------------------------------------------------------
// Integer helpers — plan.go:134-279, rader.go:178-245.

// factorsAreSmall, plan.go:134-141.
static bool factorsAreSmall(i64 n) {
  static const int primes[6] = {2, 3, 5, 7, 11, 13};
  for (int p : primes)
    while (n % p == 0) n /= p;
  return n == 1;
}

// nextSmoothConv, plan.go:154-173.
static i64 nextSmoothConv(i64 lo) {
  if (lo < 1) lo = 1;
  for (i64 m = lo;; m++) {
    i64 x = m;
    while (x % 2 == 0) x /= 2;
    while (x % 3 == 0) x /= 3;
    while (x % 5 == 0) x /= 5;
    if (x == 1) return m;
  }
}

// convCost, plan.go:192-220.
static bool convCost(i64 m, double* out) {
  const double w2 = 2.0, w3 = 2.0, w5 = 2.8, w7 = 2.6;
  double cost = 0.0;
  i64 x = m;
  while (x % 4 == 0) { cost += w2; x /= 4; }
  while (x % 2 == 0) { cost += w2; x /= 2; }
  while (x % 3 == 0) { cost += w3; x /= 3; }
  while (x % 5 == 0) { cost += w5; x /= 5; }
  while (x % 7 == 0) { cost += w7; x /= 7; }
  if (x != 1) { *out = 0; return false; }
  *out = (double)m * cost;
  return true;
}

// bestConvLen, plan.go:241-259.
static i64 bestConvLen(i64 lo) {
  if (lo < 1) lo = 1;
  i64 hi = lo + (lo * 4 + 9) / 10;
  i64 best = lo;
  double bestCost = 0;
  convCost(lo, &bestCost);
  bool haveBest = false;
  for (i64 m = lo; m <= hi; m++) {
    double c;
    if (!convCost(m, &c)) continue;
    if (!haveBest || c < bestCost) { best = m; bestCost = c; haveBest = true; }
  }
  return best;
}

// factorize, plan.go:265-279.
static std::vector<int> factorize(i64 n) {
  std::vector<int> f;
  while (n % 4 == 0) { f.push_back(4); n /= 4; }
  static const int primes[6] = {2, 3, 5, 7, 11, 13};
  for (int p : primes)
    while (n % p == 0) { f.push_back(p); n /= p; }
  return f;
}

// modPow, rader.go:217-228.
static i64 modPow(i64 base, i64 exp, i64 m) {
  i64 result = 1;
  base %= m;
  while (exp > 0) {
    if (exp & 1) result = result * base % m;
    exp >>= 1;
    base = base * base % m;
  }
  return result;
}

// primeFactorsDistinct, rader.go:200-214.
static std::vector<i64> primeFactorsDistinct(i64 m) {
  std::vector<i64> f;
  for (i64 d = 2; d * d <= m; d++) {
    if (m % d == 0) {
      f.push_back(d);
      while (m % d == 0) m /= d;
    }
  }
  if (m > 1) f.push_back(m);
  return f;
}

// primitiveRoot, rader.go:178-197.
static i64 primitiveRoot(i64 p) {
  i64 phi = p - 1;
  std::vector<i64> fs = primeFactorsDistinct(phi);
  if (fs.empty()) return 1;
  for (i64 g = 2;; g++) {
    bool gen = true;
    for (i64 f : fs)
      if (modPow(g, phi / f, p) == 1) { gen = false; break; }
    if (gen) return g;
  }
}

// isPrime, rader.go:232-245.
static bool isPrime(i64 n) {
  if (n < 2) return false;
  if (n % 2 == 0) return n == 2;
  for (i64 d = 3; d * d <= n; d += 2)
    if (n % d == 0) return false;
  return true;
}

// twiddleTable, plan.go:320-327: t[k] = exp(-2*pi*i*k/n), the angle formed as
// (-2*pi)*k/n in float64.
static std::vector<cx> twiddleTable(i64 n) {
  std::vector<cx> t(n);
  const double m2pi = -2.0 * kPi;
  for (i64 k = 0; k < n; k++) {
    double ang = m2pi * (double)k / (double)n;
    t[k] = mk(std::cos(ang), std::sin(ang));
  }
  return t;
}

// ---------------------------------------------------------------------------
// internal/kernels: scalar stage oracles and the Bluestein helpers.

// Radix2StageScalar, internal/kernels/butterfly.go:53-67.
static void radix2Stage(cx* a, i64 n, i64 span, const cx* tw) {
  i64 step = 2 * span;
  for (i64 base = 0; base < n; base += step)
    for (i64 k = 0; k < span; k++) {
      i64 i = base + k, j = i + span;
      cx t = mul(a[j], tw[k]);
      cx u = a[i];
      a[i] = add(u, t);
      a[j] = sub(u, t);
    }
}

// Radix4StageScalar, internal/kernels/butterfly.go:80-108 (cmul1 at :123-127 is
// the separately rounded product `mul`).
static void radix4Stage(cx* a, i64 n, i64 span, const cx* w1, const cx* w2,
                        const cx* w3, bool inverse) {
  i64 step = 4 * span;
  for (i64 base = 0; base < n; base += step) {
    i64 i0 = base, i1 = i0 + span, i2 = i1 + span, i3 = i2 + span;
    for (i64 k = 0; k < span; k++) {
      cx m1 = mul(a[i1 + k], w1[k]);
      cx m2 = mul(a[i2 + k], w2[k]);
      cx m3 = mul(a[i3 + k], w3[k]);
      cx a0 = a[i0 + k];
      cx t0 = add(a0, m2), t1 = sub(a0, m2), t2 = add(m1, m3), d = sub(m1, m3);
      cx t3 = inverse ? mk(-d.im, d.re) : mk(d.im, -d.re);
      a[i0 + k] = add(t0, t2);
      a[i1 + k] = add(t1, t3);
      a[i2 + k] = sub(t0, t2);
      a[i3 + k] = sub(t1, t3);
    }
  }
}

// BitReverse, internal/kernels/kernels.go:20-32.
static void bitReverse(cx* x, i64 n) {
  for (i64 i = 1, j = 0; i < n; i++) {
    i64 bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(x[i], x[j]);
  }
}

// Radix2, internal/kernels/kernels.go:38-59: per-stage twiddles by the
// recurrence w *= wlen, wlen = (cos, sin)((sign*2)*pi/length).
static void radix2Rec(cx* x, i64 n, bool inverse) {
  double sign = inverse ? 1.0 : -1.0;
  for (i64 length = 2; length <= n; length <<= 1) {
    double ang = sign * 2 * kPi / (double)length;
    double s, c;
    go_sincos(ang, &s, &c);
    cx wlen = mk(c, s);
    i64 half = length >> 1;
    for (i64 start = 0; start < n; start += length) {
      cx w = mk(1, 0);
      for (i64 k = 0; k < half; k++) {
        cx u = x[start + k];
        cx v = mul(x[start + k + half], w);
        x[start + k] = add(u, v);
        x[start + k + half] = sub(u, v);
        w = mul(w, wlen);
      }
    }
  }
}

// CMulScalar, internal/kernels/cmul.go:26-36.
static void cmulInPlace(cx* a, const cx* b, i64 n) {
  for (i64 i = 0; i < n; i++) a[i] = mul(a[i], b[i]);
}

// ---------------------------------------------------------------------------
// Iterative power-of-two engine — iterative.go.

struct StagePlan {
  int radix;
  i64 span;
  std::vector<cx> tw, twC;
};

struct ITPlan {
  i64 n;
  std::vector<i64> revPos;
  std::vector<StagePlan> stages;
};

// bitReversalForRadices, iterative.go:87-103.
static std::vector<i64> bitReversalForRadices(i64 n, const std::vector<int>& radices) {
  std::vector<i64> rev(n);
  for (i64 i = 0; i < n; i++) {
    i64 x = i, r = 0;
    for (int rad : radices) {
      r = r * rad + x % rad;
      x /= rad;
    }
    rev[i] = r;
  }
  return rev;
}

// newITPlan + buildStages, iterative.go:59-80, 108-150.
static ITPlan* newITPlan(i64 n) {
  ITPlan* p = new ITPlan();
  p->n = n;
  int log2 = 0;
  while (((i64)1 << log2) < n) log2++;
  std::vector<int> radices;
  if (log2 % 2 == 1) radices.push_back(2);
  for (int i = 0; i < log2 / 2; i++) radices.push_back(4);
  p->revPos = bitReversalForRadices(n, radices);
  std::vector<cx> root = twiddleTable(n);
  i64 span = 1;
  for (int rad : radices) {
    i64 next = span * rad;
    StagePlan st;
    st.radix = rad;
    st.span = span;
    if (rad == 2) {
      st.tw.resize(span);
      st.twC.resize(span);
      for (i64 k = 0; k < span; k++) {
        cx w = root[(k * (n / next)) % n];
        st.tw[k] = w;
        st.twC[k] = conj(w);
      }
    } else {
      st.tw.resize(span * 3);
      st.twC.resize(span * 3);
      i64 step = n / next;
      for (i64 k = 0; k < span; k++) {
        cx w1 = root[(k * step) % n], w2 = root[(2 * k * step) % n], w3 = root[(3 * k * step) % n];
        st.tw[k] = w1; st.tw[span + k] = w2; st.tw[2 * span + k] = w3;
        st.twC[k] = conj(w1); st.twC[span + k] = conj(w2); st.twC[2 * span + k] = conj(w3);
      }
    }
    p->stages.push_back(std::move(st));
    span = next;
  }
  return p;
}

// itPlan.run, iterative.go:200-214.
static void itRun(const ITPlan* p, cx* a, bool inverse) {
  for (const StagePlan& st : p->stages) {
    const cx* tw = inverse ? st.twC.data() : st.tw.data();
    if (st.radix == 2)
      radix2Stage(a, p->n, st.span, tw);
    else
      radix4Stage(a, p->n, st.span, tw, tw + st.span, tw + 2 * st.span, inverse);
  }
}

// itPlan.transform, iterative.go:154-166 (src copied when it aliases dst).
static void itTransform(const ITPlan* p, cx* dst, const cx* src, bool inverse) {
  std::vector<cx> copy;
  if (dst == src) {
    copy.assign(src, src + p->n);
    src = copy.data();
  }
  for (i64 i = 0; i < p->n; i++) dst[i] = src[p->revPos[i]];
  itRun(p, dst, inverse);
}

// ---------------------------------------------------------------------------
// Mixed-radix Cooley–Tukey engine — mixedradix.go.

struct CTPlan {
  i64 n;
  std::vector<int> factors;
  std::vector<cx> tw, twConj;
};

static CTPlan* newCTPlan(i64 n) {  // mixedradix.go:28-40
  CTPlan* p = new CTPlan();
  p->n = n;
  p->tw = twiddleTable(n);
  p->twConj.resize(n);
  for (i64 i = 0; i < n; i++) p->twConj[i] = conj(p->tw[i]);
  p->factors = factorize(n);
  return p;
}

static inline cx rotNeg90(cx z, bool inverse) {  // mixedradix.go:165-170
  return inverse ? mk(-z.im, z.re) : mk(z.im, -z.re);
}
static inline cx rotI(cx z) { return mk(-z.im, z.re); }  // mixedradix.go:301-303

static void ctRadix2(cx* out, i64 m, i64 step, const cx* tw) {  // :125-133
  for (i64 k = 0; k < m; k++) {
    cx w = tw[k * step];
    cx a = out[k];
    cx b = mul(out[k + m], w);
    out[k] = add(a, b);
    out[k + m] = sub(a, b);
  }
}

static void ctRadix4(cx* out, i64 m, i64 step, const cx* tw, bool inverse) {  // :138-160
  for (i64 k = 0; k < m; k++) {
    i64 ks = k * step;
    cx w1 = tw[ks], w2 = tw[2 * ks], w3 = tw[3 * ks];
    cx a = out[k];
    cx b = mul(out[k + m], w1);
    cx c = mul(out[k + 2 * m], w2);
    cx d = mul(out[k + 3 * m], w3);
    cx t0 = add(a, c), t1 = sub(a, c), t2 = add(b, d), t3 = sub(b, d);
    cx t3r = rotNeg90(t3, inverse);
    out[k] = add(t0, t2);
    out[k + m] = add(t1, t3r);
    out[k + 2 * m] = sub(t0, t2);
    out[k + 3 * m] = sub(t1, t3r);
  }
}

static const double kSin120 = 0.8660254037844386467637231707529361834714026269051903140279;

static void ctRadix3(cx* out, i64 m, i64 step, const cx* tw, bool inverse) {  // :178-200
  double s = inverse ? kSin120 : -kSin120;
  for (i64 k = 0; k < m; k++) {
    i64 ks = k * step;
    cx w1 = tw[ks], w2 = tw[2 * ks];
    cx a = out[k];
    cx b = mul(out[k + m], w1);
    cx c = mul(out[k + 2 * m], w2);
    cx t = add(b, c);
    out[k] = add(a, t);
    cx u = sub(a, mk(t.re * 0.5, t.im * 0.5));
    cx d = sub(b, c);
    cx v = mk(-s * d.im, s * d.re);
    out[k + m] = add(u, v);
    out[k + 2 * m] = sub(u, v);
  }
}

static void ctRadix5(cx* out, i64 m, i64 step, const cx* tw, bool inverse) {  // :205-242
  const double c1 = 0.30901699437494742410229341718281905886015458990288;
  const double s1 = 0.95105651629515357211643933337938214340569863412575;
  const double c2 = -0.80901699437494742410229341718281905886015458990289;
  const double s2 = 0.58778525229247312916870595463907276859765243764314;
  double sg = inverse ? 1.0 : -1.0;
  for (i64 k = 0; k < m; k++) {
    i64 ks = k * step;
    cx a = out[k];
    cx b = mul(out[k + m], tw[ks]);
    cx c = mul(out[k + 2 * m], tw[2 * ks]);
    cx d = mul(out[k + 3 * m], tw[3 * ks]);
    cx e = mul(out[k + 4 * m], tw[4 * ks]);
    cx t1 = add(b, e), t2 = sub(b, e), t3 = add(c, d), t4 = sub(c, d);
    out[k] = add(add(a, t1), t3);
    cx r1 = add(a, mk(c1 * t1.re + c2 * t3.re, c1 * t1.im + c2 * t3.im));
    cx r2 = add(a, mk(c2 * t1.re + c1 * t3.re, c2 * t1.im + c1 * t3.im));
    cx i1 = rotI(mk(sg * (s1 * t2.re + s2 * t4.re), sg * (s1 * t2.im + s2 * t4.im)));
    cx i2 = rotI(mk(sg * (s2 * t2.re - s1 * t4.re), sg * (s2 * t2.im - s1 * t4.im)));
    out[k + m] = add(r1, i1);
    out[k + 4 * m] = sub(r1, i1);
    out[k + 2 * m] = add(r2, i2);
    out[k + 3 * m] = sub(r2, i2);
  }
}

static void ctRadix7(cx* out, i64 m, i64 step, const cx* tw, bool inverse) {  // :251-298
  const double c1 = 0.6234898018587335305250048840042398106322747308237;
  const double s1 = 0.7818314824680298087084445266740577502323345187493;
  const double c2 = -0.2225209339563144042889025644967948758319743752712;
  const double s2 = 0.9749279121818236070181316829939312172327858006199;
  const double c3 = -0.9009688679024191262361023195074450511659191621318;
  const double s3 = 0.4338837391175581204757683328483587546099907277859;
  double sg = inverse ? 1.0 : -1.0;
  for (i64 k = 0; k < m; k++) {
    i64 ks = k * step;
    cx a = out[k];
    cx b = mul(out[k + m], tw[ks]);
    cx c = mul(out[k + 2 * m], tw[2 * ks]);
    cx d = mul(out[k + 3 * m], tw[3 * ks]);
    cx e = mul(out[k + 4 * m], tw[4 * ks]);
    cx f = mul(out[k + 5 * m], tw[5 * ks]);
    cx g = mul(out[k + 6 * m], tw[6 * ks]);
    cx s16 = add(b, g), d16 = sub(b, g), s25 = add(c, f), d25 = sub(c, f);
    cx s34 = add(d, e), d34 = sub(d, e);
    out[k] = add(add(add(a, s16), s25), s34);
    cx r1 = add(a, mk(c1 * s16.re + c2 * s25.re + c3 * s34.re, c1 * s16.im + c2 * s25.im + c3 * s34.im));
    cx r2 = add(a, mk(c2 * s16.re + c3 * s25.re + c1 * s34.re, c2 * s16.im + c3 * s25.im + c1 * s34.im));
    cx r3 = add(a, mk(c3 * s16.re + c1 * s25.re + c2 * s34.re, c3 * s16.im + c1 * s25.im + c2 * s34.im));
    cx i1 = rotI(mk(sg * (s1 * d16.re + s2 * d25.re + s3 * d34.re), sg * (s1 * d16.im + s2 * d25.im + s3 * d34.im)));
    cx i2 = rotI(mk(sg * (s2 * d16.re - s3 * d25.re - s1 * d34.re), sg * (s2 * d16.im - s3 * d25.im - s1 * d34.im)));
    cx i3 = rotI(mk(sg * (s3 * d16.re - s1 * d25.re + s2 * d34.re), sg * (s3 * d16.im - s1 * d25.im + s2 * d34.im)));
    out[k + m] = add(r1, i1);
    out[k + 6 * m] = sub(r1, i1);
    out[k + 2 * m] = add(r2, i2);
    out[k + 5 * m] = sub(r2, i2);
    out[k + 3 * m] = add(r3, i3);
    out[k + 4 * m] = sub(r3, i3);
  }
}

static void ctRadixP(i64 n, cx* out, i64 r, i64 m, i64 step, const cx* tw) {  // :310-335
  cx buf[13];
  i64 rstep = n / r;
  for (i64 k = 0; k < m; k++) {
    for (i64 i = 0; i < r; i++) buf[i] = mul(out[k + i * m], tw[i * k * step % n]);
    for (i64 q = 0; q < r; q++) {
      cx sum = mk(0, 0);
      for (i64 i = 0; i < r; i++) sum = add(sum, mul(buf[i], tw[i * q * rstep % n]));
      out[k + q * m] = sum;
    }
  }
}

// ctPlan.rec + butterfly, mixedradix.go:69-119.
static void ctRec(const CTPlan* p, cx* out, const cx* src, i64 length, i64 stride,
                  const int* factors, const cx* tw, bool inverse) {
  i64 r = factors[0];
  i64 m = length / r;
  if (m == 1) {
    for (i64 i = 0; i < r; i++) out[i] = src[i * stride];
  } else {
    for (i64 i = 0; i < r; i++)
      ctRec(p, out + i * m, src + i * stride, m, r * stride, factors + 1, tw, inverse);
  }
  i64 step = p->n / length;
  switch (r) {
    case 2: ctRadix2(out, m, step, tw); break;
    case 3: ctRadix3(out, m, step, tw, inverse); break;
    case 4: ctRadix4(out, m, step, tw, inverse); break;
    case 5: ctRadix5(out, m, step, tw, inverse); break;
    case 7: ctRadix7(out, m, step, tw, inverse); break;
    default: ctRadixP(p->n, out, r, m, step, tw); break;
  }
}

// ctPlan.transform, mixedradix.go:55-62.
static void ctTransform(const CTPlan* p, cx* dst, const cx* src, bool inverse) {
  std::vector<cx> scratch(src, src + p->n);
  ctRec(p, dst, scratch.data(), p->n, 1, p->factors.data(),
        inverse ? p->twConj.data() : p->tw.data(), inverse);
}

// ---------------------------------------------------------------------------
// Plan router and cache — plan.go.

struct RaderPlan;
struct BluesteinPlan;

struct Plan {
  i64 n = 0;
  ITPlan* it = nullptr;
  CTPlan* ct = nullptr;
  RaderPlan* rader = nullptr;
  BluesteinPlan* bluestein = nullptr;
};

static Plan* newPlan(i64 n);
static void planFFT(const Plan* p, cx* dst, const cx* src);
static void planIFFT(const Plan* p, cx* dst, const cx* src);

// planCache / cachedPlan, plan.go:287-317: built outside the lock (Rader
// re-enters the cache for its convolution sub-plan), first store wins.
static std::mutex g_planMu;
static std::map<i64, Plan*> g_planCache;

static Plan* cachedPlan(i64 n) {
  {
    std::lock_guard<std::mutex> lk(g_planMu);
    auto it = g_planCache.find(n);
    if (it != g_planCache.end()) return it->second;
  }
  Plan* p = newPlan(n);
  std::lock_guard<std::mutex> lk(g_planMu);
  auto it = g_planCache.find(n);
  if (it != g_planCache.end()) return it->second;  // the losing build is leaked, as a GC would drop it
  g_planCache[n] = p;
  return p;
}

// ---------------------------------------------------------------------------
// Rader's prime engine — rader.go.

struct RaderPlan {
  i64 n, cl;
  bool cyclic;
  std::vector<i64> perm, iperm;
  std::vector<cx> bF, bI;
};

// buildKernel, rader.go:101-125.
static std::vector<cx> raderBuildKernel(const RaderPlan* p, i64 n, bool inverse) {
  i64 q = n - 1;
  std::vector<cx> b(p->cl, mk(0, 0));
  double sign = inverse ? 1.0 : -1.0;
  i64 lim = p->cyclic ? q : 2 * q;
  for (i64 m = 0; m < lim; m++) {
    i64 s = m % q;
    i64 e = p->perm[(q - s) % q];
    double ang = sign * 2 * kPi * (double)e / (double)n;
    b[m] = mk(std::cos(ang), std::sin(ang));
  }
  std::vector<cx> out(p->cl);
  planFFT(cachedPlan(p->cl), out.data(), b.data());
  return out;
}

// newRaderPlan, rader.go:51-91.
static RaderPlan* newRaderPlan(i64 n) {
  i64 g = primitiveRoot(n);
  i64 q = n - 1;
  RaderPlan* p = new RaderPlan();
  p->n = n;
  if (factorsAreSmall(q)) {
    p->cyclic = true;
    p->cl = q;
  } else {
    p->cyclic = false;
    p->cl = bestConvLen(2 * q);
  }
  p->perm.resize(q);
  p->iperm.resize(q);
  i64 x = 1;
  for (i64 k = 0; k < q; k++) {
    p->perm[k] = x;
    x = x * g % n;
  }
  for (i64 qi = 0; qi < q; qi++) p->iperm[qi] = p->perm[(q - qi) % q];
  p->bF = raderBuildKernel(p, n, false);
  p->bI = raderBuildKernel(p, n, true);
  return p;
}

// raderPlan.transform, rader.go:129-172.
static void raderTransform(const RaderPlan* p, cx* dst, const cx* src, bool inverse) {
  i64 n = p->n, cl = p->cl, q = n - 1;
  const std::vector<cx>& bSpec = inverse ? p->bI : p->bF;
  cx x0 = src[0];
  cx sum = mk(0, 0);
  for (i64 i = 0; i < n; i++) sum = add(sum, src[i]);
  std::vector<cx> a(cl, mk(0, 0));
  for (i64 k = 0; k < q; k++) a[k] = src[p->perm[k]];
  const Plan* plan = cachedPlan(cl);
  std::vector<cx> A(cl), conv(cl);
  planFFT(plan, A.data(), a.data());
  for (i64 i = 0; i < cl; i++) A[i] = mul(A[i], bSpec[i]);
  planIFFT(plan, conv.data(), A.data());
  dst[0] = sum;
  i64 off = p->cyclic ? 0 : q;
  for (i64 qi = 0; qi < q; qi++) dst[p->iperm[qi]] = add(x0, conv[qi + off]);
}

// ---------------------------------------------------------------------------
// Bluestein chirp-z — bluestein.go.

struct BluesteinPlan {
  i64 n, m;
  std::vector<cx> wF, wI, bF, bI;
};

// buildKernelSpectrum, bluestein.go:52-63.
static std::vector<cx> bluesteinKernelSpectrum(const std::vector<cx>& w, i64 n, i64 m) {
  std::vector<cx> b(m, mk(0, 0));
  b[0] = conj(w[0]);
  for (i64 j = 1; j < n; j++) {
    cx c = conj(w[j]);
    b[j] = c;
    b[m - j] = c;
  }
  bitReverse(b.data(), m);
  radix2Rec(b.data(), m, false);
  return b;
}

// newBluesteinPlan, bluestein.go:28-48.
static BluesteinPlan* newBluesteinPlan(i64 n) {
  i64 m = 1;
  while (m < 2 * n - 1) m <<= 1;
  BluesteinPlan* p = new BluesteinPlan();
  p->n = n;
  p->m = m;
  p->wF.resize(n);
  p->wI.resize(n);
  for (i64 j = 0; j < n; j++) {
    i64 jj = (j * j) % (2 * n);
    double ang = kPi * (double)jj / (double)n;
    p->wF[j] = mk(std::cos(-ang), std::sin(-ang));
    p->wI[j] = mk(std::cos(ang), std::sin(ang));
  }
  p->bF = bluesteinKernelSpectrum(p->wF, n, m);
  p->bI = bluesteinKernelSpectrum(p->wI, n, m);
  return p;
}

// bluesteinPlan.transform, bluestein.go:67-95.
static void bluesteinTransform(const BluesteinPlan* p, cx* dst, const cx* src, bool inverse) {
  i64 n = p->n, m = p->m;
  const std::vector<cx>& w = inverse ? p->wI : p->wF;
  const std::vector<cx>& bSpec = inverse ? p->bI : p->bF;
  std::vector<cx> a(m, mk(0, 0));
  for (i64 j = 0; j < n; j++) a[j] = mul(src[j], w[j]);
  bitReverse(a.data(), m);
  radix2Rec(a.data(), m, false);
  cmulInPlace(a.data(), bSpec.data(), m);
  bitReverse(a.data(), m);
  radix2Rec(a.data(), m, true);
  cx invM = mk(1 / (double)m, 0);
  for (i64 i = 0; i < m; i++) a[i] = mul(a[i], invM);
  for (i64 k = 0; k < n; k++) dst[k] = mul(a[k], w[k]);
}

// ---------------------------------------------------------------------------

static const i64 kRaderThreshold = 700;  // plan.go:83

// NewPlan, plan.go:42-73.
static Plan* newPlan(i64 n) {
  Plan* p = new Plan();
  p->n = n;
  if (n <= 1) return p;
  if ((n & (n - 1)) == 0) {
    p->it = newITPlan(n);
    return p;
  }
  if (factorsAreSmall(n)) {
    p->ct = newCTPlan(n);
    return p;
  }
  if (n >= kRaderThreshold && isPrime(n)) {
    p->rader = newRaderPlan(n);
    return p;
  }
  p->bluestein = newBluesteinPlan(n);
  return p;
}

// Plan.execute, plan.go:113-129.
static void planExecute(const Plan* p, cx* dst, const cx* src, bool inverse) {
  if (p->n <= 1) {
    if (p->n == 1) dst[0] = src[0];
    return;
  }
  if (p->it) itTransform(p->it, dst, src, inverse);
  else if (p->ct) ctTransform(p->ct, dst, src, inverse);
  else if (p->rader) raderTransform(p->rader, dst, src, inverse);
  else bluesteinTransform(p->bluestein, dst, src, inverse);
}

// Plan.FFT / Plan.IFFT, plan.go:91-109 (the inverse multiplies by complex(1/n, 0)).
static void planFFT(const Plan* p, cx* dst, const cx* src) { planExecute(p, dst, src, false); }
static void planIFFT(const Plan* p, cx* dst, const cx* src) {
  planExecute(p, dst, src, true);
  if (p->n == 0) return;
  cx inv = mk(1 / (double)p->n, 0);
  for (i64 i = 0; i < p->n; i++) dst[i] = mul(dst[i], inv);
}

// ---------------------------------------------------------------------------
// Real-input plans — realplan.go, rfft.go.

struct RealPlan {
  i64 n = 0;
  Plan* half = nullptr;
  std::vector<cx> tw;
  Plan* full = nullptr;
};

// NewRealPlan, realplan.go:37-58.
static RealPlan* newRealPlan(i64 n) {
  RealPlan* p = new RealPlan();
  p->n = n;
  if (n <= 1 || n % 2 != 0) {
    if (n >= 1) p->full = cachedPlan(n);
    return p;
  }
  i64 m = n / 2;
  p->half = cachedPlan(m);
  p->tw.resize(m + 1);
  const double m2pi = -2.0 * kPi;
  for (i64 k = 0; k <= m; k++) {
    double ang = m2pi * (double)k / (double)n;
    p->tw[k] = mk(std::cos(ang), std::sin(ang));
  }
  return p;
}

static std::mutex g_realMu;
static std::map<i64, RealPlan*> g_realCache;

// cachedRealPlan, rfft.go:47-64 (built under its own lock).
static RealPlan* cachedRealPlan(i64 n) {
  std::lock_guard<std::mutex> lk(g_realMu);
  auto it = g_realCache.find(n);
  if (it != g_realCache.end()) return it->second;
  RealPlan* p = newRealPlan(n);
  g_realCache[n] = p;
  return p;
}

// rfftUntangle, realplan.go:138-169.
static void rfftUntangle(cx* dst, const cx* Z, const cx* tw, i64 m) {
  double z0r = Z[0].re, z0i = Z[0].im;
  dst[0] = mk(z0r + z0i, 0);
  dst[m] = mk(z0r - z0i, 0);
  i64 half = (m - 1) / 2;
  for (i64 k = 1; k <= half; k++) {
    cx zk = Z[k], zmk = Z[m - k];
    double xer = (zk.re + zmk.re) * 0.5;
    double xei = (zk.im - zmk.im) * 0.5;
    double xor_ = (zk.im + zmk.im) * 0.5;
    double xoi = -(zk.re - zmk.re) * 0.5;
    double wr = tw[k].re, wi = tw[k].im;
    double tr = wr * xor_ - wi * xoi;
    double ti = wr * xoi + wi * xor_;
    dst[k] = mk(xer + tr, xei + ti);
    dst[m - k] = mk(xer - tr, -(xei - ti));
  }
  if ((m & 1) == 0) {
    i64 k = m / 2;
    dst[k] = mk(Z[k].re, -Z[k].im);
  }
}

// RealPlan.RFFT, realplan.go:77-116. dst holds n/2+1 bins.
static void realRFFT(const RealPlan* p, cx* dst, const double* src) {
  i64 n = p->n;
  if (n == 0) return;
  if (n == 1) { dst[0] = mk(src[0], 0); return; }
  if (!p->half) {
    std::vector<cx> c(n), full(n);
    for (i64 i = 0; i < n; i++) c[i] = mk(src[i], 0);
    planFFT(p->full, full.data(), c.data());
    for (i64 i = 0; i < n / 2 + 1; i++) dst[i] = full[i];
    return;
  }
  i64 m = n / 2;
  std::vector<cx> z(m), Z(m);
  if (p->half->it) {
    // transformRealPacked, iterative.go:191-197: the pair packing fused into the
    // digit-reversal gather.
    const ITPlan* it = p->half->it;
    for (i64 i = 0; i < m; i++) {
      i64 j = it->revPos[i];
      Z[i] = mk(src[2 * j], src[2 * j + 1]);
    }
    itRun(it, Z.data(), false);
  } else {
    for (i64 j = 0; j < m; j++) z[j] = mk(src[2 * j], src[2 * j + 1]);
    planFFT(p->half, Z.data(), z.data());
  }
  rfftUntangle(dst, Z.data(), p->tw.data(), m);
}

// irfftFull, realplan.go:197-213.
static void irfftFull(const RealPlan* p, double* dst, const cx* src, i64 srclen) {
  i64 n = p->n;
  std::vector<cx> full(n, mk(0, 0)), inv(n);
  i64 half = n / 2 + 1;
  for (i64 k = 0; k < half && k < srclen; k++) {
    full[k] = src[k];
    if (k > 0 && k < n - k) full[n - k] = conj(src[k]);
  }
  planIFFT(cachedPlan(n), inv.data(), full.data());
  for (i64 i = 0; i < n; i++) dst[i] = inv[i].re;
}

// irfftPacked, realplan.go:221-296.
static void irfftPacked(const RealPlan* p, double* dst, const cx* src, i64 srclen) {
  i64 n = p->n, m = n / 2;
  auto bin = [&](i64 k) { return k < srclen ? src[k] : mk(0, 0); };
  std::vector<cx> z(m), Z(m);
  double x0 = bin(0).re, xm = bin(m).re;
  Z[0] = mk((x0 + xm) * 0.5, (x0 - xm) * 0.5);
  for (i64 k = 1; k <= m - k; k++) {
    cx xk = bin(k), xmk = bin(m - k);
    cx cmk = conj(xmk);
    double xer = (xk.re + cmk.re) * 0.5;
    double xei = (xk.im + cmk.im) * 0.5;
    double dr = (xk.re - cmk.re) * 0.5;
    double di = (xk.im - cmk.im) * 0.5;
    double wr = p->tw[k].re, wi = p->tw[k].im;
    double xor_ = wr * dr + wi * di;
    double xoi = wr * di - wi * dr;
    Z[k] = mk(xer - xoi, xei + xor_);
    if (k != m - k) Z[m - k] = mk(xer + xoi, -(xei - xor_));
  }
  if (p->half->it) {
    // transformScratch, iterative.go:177-183, then normalise on unpack.
    const ITPlan* it = p->half->it;
    for (i64 i = 0; i < m; i++) z[i] = Z[it->revPos[i]];
    itRun(it, z.data(), true);
    double inv = 1 / (double)m;
    for (i64 j = 0; j < m; j++) {
      dst[2 * j] = z[j].re * inv;
      dst[2 * j + 1] = z[j].im * inv;
    }
    return;
  }
  planIFFT(p->half, z.data(), Z.data());
  for (i64 j = 0; j < m; j++) {
    dst[2 * j] = z[j].re;
    dst[2 * j + 1] = z[j].im;
  }
}

// RealPlan.IRFFT, realplan.go:181-192.
static void realIRFFT(const RealPlan* p, double* dst, const cx* src, i64 srclen) {
  if (p->n <= 0) return;
  if (!p->half) irfftFull(p, dst, src, srclen);
  else irfftPacked(p, dst, src, srclen);
}

// ---------------------------------------------------------------------------
// Line-parallel helpers — parallel.go.

static int g_parWorkers = (int)std::max(1u, std::thread::hardware_concurrency());
static const i64 kParThreshold = 1 << 14;  // parallel.go:25

template <typename F>
static void parChunks(i64 n, F body) {  // parallel.go:32-55
  i64 w = g_parWorkers;
  if (w > n) w = n;
  if (w <= 1) { body(0, n); return; }
  i64 chunk = (n + w - 1) / w;
  std::vector<std::thread> ts;
  for (i64 lo = 0; lo < n; lo += chunk) {
    i64 hi = std::min(lo + chunk, n);
    ts.emplace_back([&body, lo, hi]() { body(lo, hi); });
  }
  for (auto& t : ts) t.join();
}

static bool parallelizeLines(i64 lineCount, i64 lineLen) {  // parallel.go:59-61
  return g_parWorkers > 1 && lineCount > 1 && lineCount * lineLen >= kParThreshold;
}

// ---------------------------------------------------------------------------
// N-D transforms — ndfft.go, rfft2.go.

static i64 lineBase(i64 c, const i64* shape, const i64* stride, int rank, int ax) {  // ndfft.go:144-155
  i64 base = 0;
  for (int a = rank - 1; a >= 0; a--) {
    if (a == ax) continue;
    i64 s = shape[a];
    base += (c % s) * stride[a];
    c /= s;
  }
  return base;
}

// transformAxis, ndfft.go:102-136.
static void transformAxis(cx* out, i64 total, const i64* shape, const i64* stride, int rank,
                          int ax, bool inverse) {
  i64 n = shape[ax];
  if (n == 1) return;
  i64 st = stride[ax];
  const Plan* plan = cachedPlan(n);
  i64 lineCount = total / n;
  auto work = [&](i64 lo, i64 hi) {
    std::vector<cx> buf(n), res(n);
    for (i64 c = lo; c < hi; c++) {
      i64 base = lineBase(c, shape, stride, rank, ax);
      for (i64 i = 0; i < n; i++) buf[i] = out[base + i * st];
      if (inverse) planIFFT(plan, res.data(), buf.data());
      else planFFT(plan, res.data(), buf.data());
      for (i64 i = 0; i < n; i++) out[base + i * st] = res[i];
    }
  };
  if (parallelizeLines(lineCount, n)) parChunks(lineCount, work);
  else work(0, lineCount);
}

// transformN, ndfft.go:66-90 (validation done by the caller).
static void transformN(cx* out, i64 total, const i64* shape, int rank, bool inverse) {
  if (total <= 1) return;
  std::vector<i64> stride(rank);
  i64 acc = 1;
  for (int ax = rank - 1; ax >= 0; ax--) {
    stride[ax] = acc;
    acc *= shape[ax];
  }
  for (int ax = 0; ax < rank; ax++) transformAxis(out, total, shape, stride.data(), rank, ax, inverse);
}

// RFFT2, rfft2.go:22-74.
static void rfft2(const double* data, i64 rows, i64 cols, cx* half) {
  i64 rcols = cols / 2 + 1;
  const RealPlan* rp = cachedRealPlan(cols);
  auto rowWork = [&](i64 lo, i64 hi) {
    std::vector<double> row(cols);
    std::vector<cx> dst(rcols);
    for (i64 r = lo; r < hi; r++) {
      std::memcpy(row.data(), data + r * cols, sizeof(double) * cols);
      realRFFT(rp, dst.data(), row.data());
      std::memcpy(half + r * rcols, dst.data(), sizeof(cx) * rcols);
    }
  };
  if (parallelizeLines(rows, cols)) parChunks(rows, rowWork);
  else rowWork(0, rows);
  const Plan* cp = cachedPlan(rows);
  auto colWork = [&](i64 lo, i64 hi) {
    std::vector<cx> col(rows), dst(rows);
    for (i64 c = lo; c < hi; c++) {
      for (i64 r = 0; r < rows; r++) col[r] = half[r * rcols + c];
      planFFT(cp, dst.data(), col.data());
      for (i64 r = 0; r < rows; r++) half[r * rcols + c] = dst[r];
    }
  };
  if (parallelizeLines(rcols, rows)) parChunks(rcols, colWork);
  else colWork(0, rcols);
}

// IRFFT2, rfft2.go:87-138.
static void irfft2(const cx* data, i64 datalen, i64 rows, i64 cols, double* out) {
  i64 rcols = cols / 2 + 1;
  std::vector<cx> half(rows * rcols, mk(0, 0));
  for (i64 i = 0; i < rows * rcols && i < datalen; i++) half[i] = data[i];
  const Plan* cp = cachedPlan(rows);
  auto colWork = [&](i64 lo, i64 hi) {
    std::vector<cx> col(rows), dst(rows);
    for (i64 c = lo; c < hi; c++) {
      for (i64 r = 0; r < rows; r++) col[r] = half[r * rcols + c];
      planIFFT(cp, dst.data(), col.data());
      for (i64 r = 0; r < rows; r++) half[r * rcols + c] = dst[r];
    }
  };
  if (parallelizeLines(rcols, rows)) parChunks(rcols, colWork);
  else colWork(0, rcols);
  const RealPlan* rp = cachedRealPlan(cols);
  auto rowWork = [&](i64 lo, i64 hi) {
    std::vector<double> dst(cols);
    for (i64 r = lo; r < hi; r++) {
      realIRFFT(rp, dst.data(), half.data() + r * rcols, rcols);
      std::memcpy(out + r * cols, dst.data(), sizeof(double) * cols);
    }
  };
  if (parallelizeLines(rows, cols)) parChunks(rows, rowWork);
  else rowWork(0, rows);
}

// ---------------------------------------------------------------------------
// PSD / Spectrogram / windows / frequency axes — freq.go, window.go.

static void psd(const double* x, i64 n, double d, double* out) {  // freq.go:69-90
  if (n == 0) return;
  i64 m = n / 2 + 1;
  std::vector<cx> spec(m);
  realRFFT(cachedRealPlan(n), spec.data(), x);
  double scale = d / (double)n;
  for (i64 k = 0; k < m; k++) {
    double re = spec[k].re, im = spec[k].im;
    double p = (re * re + im * im) * scale;
    if (k != 0 && !(n % 2 == 0 && k == m - 1)) p *= 2;
    out[k] = p;
  }
}

static void cosineWindow(i64 n, const double* a, int na, double* w) {  // window.go:59-78
  if (n <= 0) return;
  if (n == 1) { w[0] = 1; return; }
  double denom = (double)(n - 1);
  const double twoPi = 2 * kPi;
  for (i64 k = 0; k < n; k++) {
    double s = a[0];
    double sign = -1.0;
    for (int j = 1; j < na; j++) {
      s += sign * a[j] * std::cos(twoPi * (double)j * (double)k / denom);
      sign = -sign;
    }
    w[k] = s;
  }
}

}  // namespace oracle

// ===========================================================================
// C ABI for the ctypes test harness and the CPU baseline. Complex data is
// interleaved (re, im) float64.
using namespace oracle;

extern "C" {

void or_set_workers(int w) { g_parWorkers = w < 1 ? 1 : w; }
int or_get_workers(void) { return g_parWorkers; }

// Engine a length routes to (plan.go:42-73): 0 trivial, 1 iterative pow2,
// 2 mixed-radix, 3 Rader, 4 Bluestein.
int or_route(i64 n) {
  Plan* p = cachedPlan(n);
  if (p->it) return 1;
  if (p->ct) return 2;
  if (p->rader) return 3;
  if (p->bluestein) return 4;
  return 0;
}

int or_factorize(i64 n, int* out, int cap) {
  std::vector<int> f = factorize(n);
  int k = (int)std::min<size_t>(f.size(), (size_t)cap);
  for (int i = 0; i < k; i++) out[i] = f[i];
  return (int)f.size();
}
int or_factors_are_small(i64 n) { return factorsAreSmall(n); }
i64 or_next_smooth_conv(i64 lo) { return nextSmoothConv(lo); }
int or_conv_cost(i64 m, double* cost) { return convCost(m, cost); }
i64 or_best_conv_len(i64 lo) { return bestConvLen(lo); }
i64 or_primitive_root(i64 p) { return primitiveRoot(p); }
int or_is_prime(i64 n) { return isPrime(n); }
i64 or_mod_pow(i64 b, i64 e, i64 m) { return modPow(b, e, m); }

// Rader plan data: perm/iperm need n-1 entries each.
int or_rader_info(i64 n, i64* perm, i64* iperm, i64* cl, int* cyclic) {
  if (n < 3 || !isPrime(n)) return -1;
  RaderPlan* p = newRaderPlan(n);
  for (i64 i = 0; i < n - 1; i++) {
    if (perm) perm[i] = p->perm[i];
    if (iperm) iperm[i] = p->iperm[i];
  }
  *cl = p->cl;
  *cyclic = p->cyclic;
  delete p;
  return 0;
}

// Bluestein convolution length and chirp phase indices j^2 mod 2n.
i64 or_bluestein_info(i64 n, i64* phase) {
  i64 m = 1;
  while (m < 2 * n - 1) m <<= 1;
  if (phase)
    for (i64 j = 0; j < n; j++) phase[j] = (j * j) % (2 * n);
  return m;
}

void or_bitrev_for_radices(i64 n, const int* radices, int nr, i64* out) {
  std::vector<int> r(radices, radices + nr);
  std::vector<i64> rev = bitReversalForRadices(n, r);
  std::memcpy(out, rev.data(), sizeof(i64) * n);
}
// Iterative-engine radix schedule for a power of two (iterative.go:65-75).
int or_it_radices(i64 n, int* out) {
  int log2 = 0;
  while (((i64)1 << log2) < n) log2++;
  int k = 0;
  if (log2 % 2 == 1) out[k++] = 2;
  for (int i = 0; i < log2 / 2; i++) out[k++] = 4;
  return k;
}

void or_bit_reverse(i64 n, double* x) { bitReverse((cx*)x, n); }
void or_radix2_rec(i64 n, double* x, int inverse) { radix2Rec((cx*)x, n, inverse != 0); }
void or_go_sincos(double x, double* s, double* c) { go_sincos(x, s, c); }
// The per-stage recurrence twiddles kernels.Radix2 uses for a stage of the
// given length: w[k], k < length/2 (kernels.go:44-55).
void or_radix2_stage_twiddles(i64 length, int inverse, double* out) {
  double sign = inverse ? 1.0 : -1.0;
  double ang = sign * 2 * kPi / (double)length;
  double s, c;
  go_sincos(ang, &s, &c);
  cx wlen = mk(c, s), w = mk(1, 0);
  for (i64 k = 0; k < length / 2; k++) {
    ((cx*)out)[k] = w;
    w = mul(w, wlen);
  }
}

void or_twiddle_table(i64 n, double* out) {
  std::vector<cx> t = twiddleTable(n);
  std::memcpy(out, t.data(), sizeof(cx) * n);
}

// Plan.FFT / Plan.IFFT on one length-n line (dst may alias src).
int or_fft(i64 n, const double* src, double* dst, int inverse) {
  if (n < 0) return -1;
  if (n == 0) return 0;
  const Plan* p = cachedPlan(n);
  if (inverse) planIFFT(p, (cx*)dst, (const cx*)src);
  else planFFT(p, (cx*)dst, (const cx*)src);
  return 0;
}

// A batch of independent lines; `threads` > 1 splits the batch into contiguous
// chunks (an extension: the reference has no batch API and loops serially).
int or_fft_batch(i64 n, i64 batch, const double* src, double* dst, int inverse, int threads) {
  if (n <= 0 || batch < 0) return -1;
  const Plan* p = cachedPlan(n);
  auto body = [&](i64 lo, i64 hi) {
    for (i64 b = lo; b < hi; b++) {
      const cx* s = (const cx*)src + b * n;
      cx* d = (cx*)dst + b * n;
      if (inverse) planIFFT(p, d, s);
      else planFFT(p, d, s);
    }
  };
  if (threads <= 1 || batch <= 1) { body(0, batch); return 0; }
  i64 w = std::min<i64>(threads, batch), chunk = (batch + w - 1) / w;
  std::vector<std::thread> ts;
  for (i64 lo = 0; lo < batch; lo += chunk) ts.emplace_back(body, lo, std::min(lo + chunk, batch));
  for (auto& t : ts) t.join();
  return 0;
}

// RealPlan.RFFT: out holds n/2+1 complex bins.
int or_rfft(i64 n, const double* x, double* out) {
  if (n < 0) return -1;
  if (n == 0) return 0;
  realRFFT(cachedRealPlan(n), (cx*)out, x);
  return 0;
}

// RealPlan.IRFFT: reads min(nbins, n/2+1) bins.
int or_irfft(i64 n, const double* spec, i64 nbins, double* out) {
  if (n <= 0) return 0;
  realIRFFT(cachedRealPlan(n), out, (const cx*)spec, nbins);
  return 0;
}

// FFTN / IFFTN. Returns -1 for a non-positive axis, -2 for a product mismatch
// (ndfft.go:50-62).
int or_fftn(const i64* shape, int rank, i64 total, const double* src, double* dst, int inverse) {
  i64 prod = 1;
  for (int i = 0; i < rank; i++) {
    if (shape[i] <= 0) return -1;
    prod *= shape[i];
  }
  if (prod != total) return -2;
  std::memmove(dst, src, sizeof(cx) * total);
  transformN((cx*)dst, total, shape, rank, inverse != 0);
  return 0;
}

int or_rfft2(i64 rows, i64 cols, i64 total, const double* x, double* out) {
  if (rows <= 0 || cols <= 0) return -1;
  if (rows * cols != total) return -2;
  rfft2(x, rows, cols, (cx*)out);
  return 0;
}

int or_irfft2(i64 rows, i64 cols, const double* spec, i64 nbins, double* out) {
  if (rows <= 0 || cols <= 0) return -1;
  irfft2((const cx*)spec, nbins, rows, cols, out);
  return 0;
}

int or_psd(i64 n, const double* x, double d, double* out) {
  if (n < 0) return -1;
  psd(x, n, d, out);
  return 0;
}

// Spectrogram, freq.go:107-127. Returns the frame count (or -1/-2/-3 for the
// three argument panics); out (frames x (segment/2+1)) may be null to count.
// `threads` > 1 splits frames over threads (an extension: the reference is serial).
i64 or_spectrogram(i64 n, const double* x, i64 segment, i64 overlap, const double* window,
                   i64 window_len, double d, double* out, int threads) {
  if (segment <= 0) return -1;
  if (window_len != segment) return -2;
  if (overlap < 0 || overlap >= segment) return -3;
  i64 step = segment - overlap;
  i64 frames = 0;
  for (i64 start = 0; start + segment <= n; start += step) frames++;
  if (!out) return frames;
  i64 bins = segment / 2 + 1;
  auto body = [&](i64 lo, i64 hi) {
    std::vector<double> seg(segment);
    for (i64 t = lo; t < hi; t++) {
      i64 start = t * step;
      for (i64 i = 0; i < segment; i++) seg[i] = x[start + i] * window[i];
      psd(seg.data(), segment, d, out + t * bins);
    }
  };
  if (threads <= 1 || frames <= 1) { body(0, frames); return frames; }
  i64 w = std::min<i64>(threads, frames), chunk = (frames + w - 1) / w;
  std::vector<std::thread> ts;
  for (i64 lo = 0; lo < frames; lo += chunk) ts.emplace_back(body, lo, std::min(lo + chunk, frames));
  for (auto& t : ts) t.join();
  return frames;
}

// Windows, window.go: kind 0 Hann, 1 Hamming, 2 Blackman, 3 BlackmanHarris,
// 4 Bartlett.
int or_window(int kind, i64 n, double* out) {
  static const double hann[2] = {0.5, 0.5}, hamming[2] = {0.54, 0.46},
                      blackman[3] = {0.42, 0.5, 0.08},
                      bh[4] = {0.35875, 0.48829, 0.14128, 0.01168};
  if (n <= 0) return 0;
  switch (kind) {
    case 0: cosineWindow(n, hann, 2, out); break;
    case 1: cosineWindow(n, hamming, 2, out); break;
    case 2: cosineWindow(n, blackman, 3, out); break;
    case 3: cosineWindow(n, bh, 4, out); break;
    case 4: {  // Bartlett, window.go:86-100
      if (n == 1) { out[0] = 1; break; }
      double denom = (double)(n - 1), half = denom / 2;
      for (i64 k = 0; k < n; k++) out[k] = (2 / denom) * (half - std::fabs((double)k - half));
      break;
    }
    default: return -1;
  }
  return 0;
}

void or_fftfreq(i64 n, double d, double* out) {  // freq.go:17-35
  if (n <= 0) return;
  double val = 1.0 / (d * (double)n);
  i64 half = (n - 1) / 2;
  for (i64 i = 0; i <= half; i++) out[i] = (double)i * val;
  i64 j = -(n / 2);
  for (i64 i = half + 1; i < n; i++) { out[i] = (double)j * val; j++; }
}

void or_rfftfreq(i64 n, double d, double* out) {  // freq.go:45-56
  if (n <= 0) return;
  i64 m = n / 2 + 1;
  double val = 1.0 / (d * (double)n);
  for (i64 i = 0; i < m; i++) out[i] = (double)i * val;
}

// Independent check: O(n^2) DFT in long double with exact integer phase
// reduction (k*j mod n). Not a restatement of the reference; it pins it.
int or_naive_dft(i64 n, const double* src, double* dst, int inverse) {
  const long double pi = 3.141592653589793238462643383279502884L;
  long double sign = inverse ? 1.0L : -1.0L;
  std::vector<long double> c(n), s(n);
  for (i64 p = 0; p < n; p++) {
    long double a = sign * 2.0L * pi * (long double)p / (long double)n;
    c[p] = cosl(a);
    s[p] = sinl(a);
  }
  for (i64 k = 0; k < n; k++) {
    long double re = 0, im = 0;
    for (i64 j = 0; j < n; j++) {
      i64 p = (i64)((__int128)k * j % n);
      long double xr = src[2 * j], xi = src[2 * j + 1];
      re += xr * c[p] - xi * s[p];
      im += xr * s[p] + xi * c[p];
    }
    dst[2 * k] = (double)re;
    dst[2 * k + 1] = (double)im;
  }
  return 0;
}

}  // extern "C"
