// gfx950 transform kernels: LDS-staged Stockham autosort FFTs.
//
// Layout conventions: complex data is interleaved (re, im) in the working
// precision (double2 for the plain complex128 path, float2 for the complex64
// path). A workgroup owns C columns of R points; LDS holds point r of column c
// at lds[c * Rp + lp(r, psh)], with an odd column pitch Rp = R + (R >> psh) + 1
// and one pad slot after every 2^psh points (psh = 4 normally), so that
// neither walking across columns (load/store phases) nor the stride-radix
// writes of the first Stockham stages hit one LDS bank. A single-column
// convolution that only fits the 160 KB of LDS with less padding uses psh = 6.
// Every LDS stage reads all its butterfly inputs into registers, synchronises,
// then writes the outputs (in-place Stockham).
#include "kernels.h"

#include <hip/hip_runtime.h>

namespace gofft {

template <typename T> struct Vec2;
template <> struct Vec2<double> { typedef double2 type; };
template <> struct Vec2<float> { typedef float2 type; };

template <typename T2> __device__ __forceinline__ T2 mkc(decltype(T2{}.x) x, decltype(T2{}.x) y) {
  T2 r;
  r.x = x;
  r.y = y;
  return r;
}
template <typename T2> __device__ __forceinline__ T2 cadd(T2 a, T2 b) { return mkc<T2>(a.x + b.x, a.y + b.y); }
template <typename T2> __device__ __forceinline__ T2 csub(T2 a, T2 b) { return mkc<T2>(a.x - b.x, a.y - b.y); }
template <typename T2> __device__ __forceinline__ T2 cmul(T2 a, T2 b) {
  return mkc<T2>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T2> __device__ __forceinline__ T2 cscale(T2 a, decltype(T2{}.x) s) { return mkc<T2>(a.x * s, a.y * s); }
template <typename T2> __device__ __forceinline__ T2 mul_negi(T2 a) { return mkc<T2>(a.y, -a.x); }
template <typename T2> __device__ __forceinline__ T2 cconj(T2 a) { return mkc<T2>(a.x, -a.y); }

template <typename T2> __device__ __forceinline__ T2 from_d2(double2 w);
template <> __device__ __forceinline__ double2 from_d2<double2>(double2 w) { return w; }
template <> __device__ __forceinline__ float2 from_d2<float2>(double2 w) {
  return make_float2((float)w.x, (float)w.y);
}

// Line data that a launch reads once and writes once (the transform's input
// and output points, never the tables) can go through the non-temporal forms
// of the vector loads and stores (PassArgs::nt / ConvArgs::nt, chosen per
// launch by the planner). with_nt turns the launch's flag into a
// compile-time constant, so each access loop is straight-line code.
template <bool B> struct BoolC { static constexpr bool value = B; };
template <typename T2> struct NtVec;
template <> struct NtVec<float2> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct NtVec<double2> { typedef double type __attribute__((ext_vector_type(2))); };
template <bool NT, typename T2> __device__ __forceinline__ T2 ldl(const T2* p) {
  if constexpr (NT) {
    typedef typename NtVec<T2>::type V;
    const V v = __builtin_nontemporal_load(reinterpret_cast<const V*>(p));
    return mkc<T2>(v.x, v.y);
  } else {
    return *p;
  }
}
template <bool NT, typename T2> __device__ __forceinline__ void stl(T2* p, T2 v) {
  if constexpr (NT) {
    typedef typename NtVec<T2>::type V;
    V w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<V*>(p));
  } else {
    *p = v;
  }
}
template <typename F> __device__ __forceinline__ void with_nt(bool nt, F&& f) {
  if (nt)
    f(BoolC<true>{});
  else
    f(BoolC<false>{});
}

// ---------------------------------------------------------------------------
// Forward DFT butterflies in registers, natural order in and out.

template <typename T2> __device__ __forceinline__ void dft2(T2* v) {
  T2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <typename T2> __device__ __forceinline__ void dft4(T2& v0, T2& v1, T2& v2, T2& v3) {
  T2 t0 = cadd(v0, v2), t1 = csub(v0, v2), t2 = cadd(v1, v3), t3 = mul_negi(csub(v1, v3));
  v0 = cadd(t0, t2);
  v1 = cadd(t1, t3);
  v2 = csub(t0, t2);
  v3 = csub(t1, t3);
}

// cos/sin(2*pi*m/P) for m = 1 .. (P-1)/2.
template <int P> struct PrimeConst;
template <> struct PrimeConst<3> {
  __device__ static double c(int m) { return -0.499999999999999999973; }
  __device__ static double s(int m) { return 0.866025403784438646787; }
};
template <> struct PrimeConst<5> {
  __device__ static double c(int m) { return m == 1 ? 0.309016994374947424076 : -0.809016994374947424104; }
  __device__ static double s(int m) { return m == 1 ? 0.951056516295153572111 : 0.587785252292473129135; }
};
template <> struct PrimeConst<7> {
  __device__ static double c(int m) {
    return m == 1 ? 0.623489801858733530527 : m == 2 ? -0.222520933956314404342 : -0.900968867902419126320;
  }
  __device__ static double s(int m) {
    return m == 1 ? 0.781831482468029808730 : m == 2 ? 0.974927912181823606982 : 0.433883739117558120293;
  }
};
template <> struct PrimeConst<11> {
  __device__ static double c(int m) {
    const double t[5] = {0.841253532831181168863, 0.415415013001886425489, -0.142314838273285140460,
                         -0.654860733945285064127, -0.959492973614497389901};
    return t[m - 1];
  }
  __device__ static double s(int m) {
    const double t[5] = {0.540640817455597582101, 0.909631995354518371418, 0.989821441880932732359,
                         0.755749574354258283704, 0.281732556841429697652};
    return t[m - 1];
  }
};
template <> struct PrimeConst<13> {
  __device__ static double c(int m) {
    const double t[6] = {0.885456025653209895872, 0.568064746731155802487, 0.120536680255323053250,
                         -0.354604887042535626003, -0.748510748171101098684, -0.970941817426052027193};
    return t[m - 1];
  }
  __device__ static double s(int m) {
    const double t[6] = {0.464723172043768545663, 0.822983865893656394580, 0.992708874098053992835,
                         0.935016242685414823450, 0.663122658240795202322, 0.239315664287557766952};
    return t[m - 1];
  }
};

// Odd prime P: pair inputs (m, P-m) into sums S_m and differences D_m, then
// y_k = x0 + sum_m cos(2pi km/P) S_m - i * sum_m sin(2pi km/P) D_m and
// y_{P-k} the conjugate combination.
template <int P, typename T2> __device__ __forceinline__ void dft_odd(T2* v) {
  typedef decltype(T2{}.x) T;
  constexpr int H = (P - 1) / 2;
  T2 S[H], D[H];
#pragma unroll
  for (int m = 1; m <= H; ++m) {
    S[m - 1] = cadd(v[m], v[P - m]);
    D[m - 1] = csub(v[m], v[P - m]);
  }
  T2 x0 = v[0];
  T2 y0 = x0;
#pragma unroll
  for (int m = 0; m < H; ++m) y0 = cadd(y0, S[m]);
#pragma unroll
  for (int k = 1; k <= H; ++k) {
    T2 r = x0;
    T2 q = mkc<T2>(0, 0);
#pragma unroll
    for (int m = 1; m <= H; ++m) {
      int e = (k * m) % P;
      T c = (T)PrimeConst<P>::c(e <= H ? e : P - e);
      T s = (T)(e <= H ? PrimeConst<P>::s(e) : -PrimeConst<P>::s(P - e));
      r = mkc<T2>(r.x + c * S[m - 1].x, r.y + c * S[m - 1].y);
      q = mkc<T2>(q.x + s * D[m - 1].x, q.y + s * D[m - 1].y);
    }
    // -i * q
    T2 iq = mul_negi(q);
    v[k] = cadd(r, iq);
    v[P - k] = csub(r, iq);
  }
  v[0] = y0;
}

template <typename T2> __device__ __forceinline__ void dft8(T2* v) {
  typedef decltype(T2{}.x) T;
  const T r = (T)0.707106781186547524401;
  T2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  T2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4(e0, e1, e2, e3);
  dft4(o0, o1, o2, o3);
  o1 = mkc<T2>(r * (o1.x + o1.y), r * (o1.y - o1.x));
  o2 = mul_negi(o2);
  o3 = mkc<T2>(r * (o3.y - o3.x), -r * (o3.x + o3.y));
  v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
  v[1] = cadd(e1, o1); v[5] = csub(e1, o1);
  v[2] = cadd(e2, o2); v[6] = csub(e2, o2);
  v[3] = cadd(e3, o3); v[7] = csub(e3, o3);
}

template <typename T2> __device__ __forceinline__ void dft16(T2* v) {
  typedef decltype(T2{}.x) T;
  // n = 4*n1 + n2: DFT4 over n1, twiddle W16^(n2*k1), DFT4 over n2.
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) dft4(v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]);
  const T c1 = (T)0.923879532511286756101, s1 = (T)0.382683432365089771723;
  const T r2 = (T)0.707106781186547524401;
  // position of A[n2][k1] is n2 + 4*k1
  // n2=1: k1=1,2,3 -> W^1, W^2, W^3
  v[5] = mkc<T2>(c1 * v[5].x + s1 * v[5].y, c1 * v[5].y - s1 * v[5].x);
  v[9] = mkc<T2>(r2 * (v[9].x + v[9].y), r2 * (v[9].y - v[9].x));
  v[13] = mkc<T2>(s1 * v[13].x + c1 * v[13].y, s1 * v[13].y - c1 * v[13].x);
  // n2=2: W^2, W^4, W^6
  v[6] = mkc<T2>(r2 * (v[6].x + v[6].y), r2 * (v[6].y - v[6].x));
  v[10] = mul_negi(v[10]);
  v[14] = mkc<T2>(r2 * (v[14].y - v[14].x), -r2 * (v[14].x + v[14].y));
  // n2=3: W^3, W^6, W^9
  v[7] = mkc<T2>(s1 * v[7].x + c1 * v[7].y, s1 * v[7].y - c1 * v[7].x);
  v[11] = mkc<T2>(r2 * (v[11].y - v[11].x), -r2 * (v[11].x + v[11].y));
  // W^9 = exp(-2pi i 9/16) = -c1 + i s1
  v[15] = mkc<T2>(-c1 * v[15].x - s1 * v[15].y, -c1 * v[15].y + s1 * v[15].x);
  T2 y[16];
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    T2 a0 = v[4 * k1], a1 = v[4 * k1 + 1], a2 = v[4 * k1 + 2], a3 = v[4 * k1 + 3];
    dft4(a0, a1, a2, a3);
    y[k1] = a0;
    y[k1 + 4] = a1;
    y[k1 + 8] = a2;
    y[k1 + 12] = a3;
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = y[i];
}

template <int RAD, typename T2> __device__ __forceinline__ void dft(T2* v);

constexpr int cinv(int a, int m) {  // a^-1 mod m (gcd(a, m) = 1)
  int r = 1;
  while ((a * r) % m != 1) ++r;
  return r;
}

// Prime-factor (Good-Thomas) DFT of N = A*B, gcd(A, B) = 1, with no twiddles:
// input n = (B*n1 + A*n2) mod N, output k = (k1*B*(B^-1 mod A) +
// k2*A*(A^-1 mod B)) mod N; A-point DFTs over n1, then B-point DFTs over n2.
template <int A, int B, typename T2> __device__ __forceinline__ void dft_pfa(T2* v) {
  constexpr int N = A * B;
  constexpr int EA = B * cinv(B % A, A) % N, EB = A * cinv(A % B, B) % N;
  T2 t[N];
#pragma unroll
  for (int n2 = 0; n2 < B; ++n2) {
    T2 a[A];
#pragma unroll
    for (int n1 = 0; n1 < A; ++n1) a[n1] = v[(B * n1 + A * n2) % N];
    dft<A>(a);
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) t[k1 * B + n2] = a[k1];
  }
#pragma unroll
  for (int k1 = 0; k1 < A; ++k1) {
    T2 b[B];
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) b[n2] = t[k1 * B + n2];
    dft<B>(b);
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) v[(k1 * EA + k2 * EB) % N] = b[k2];
  }
}

template <int RAD, typename T2> __device__ __forceinline__ void dft(T2* v) {
  if constexpr (RAD == 2) dft2(v);
  else if constexpr (RAD == 4) dft4(v[0], v[1], v[2], v[3]);
  else if constexpr (RAD == 8) dft8(v);
  else if constexpr (RAD == 16) dft16(v);
  else if constexpr (RAD == 6) dft_pfa<2, 3>(v);
  else if constexpr (RAD == 10) dft_pfa<2, 5>(v);
  else if constexpr (RAD == 12) dft_pfa<4, 3>(v);
  else if constexpr (RAD == 14) dft_pfa<2, 7>(v);
  else if constexpr (RAD == 15) dft_pfa<3, 5>(v);
  else dft_odd<RAD>(v);
}

// ---------------------------------------------------------------------------
// One LDS stage over all columns of the workgroup. Thread (c, lt) owns the
// butterflies j = lt, lt + tpc, ... of column c.

__device__ __forceinline__ int lp(int p, int sh) { return p + (p >> sh); }

// The complex64 convolution's wide form (24 points per thread, a single
// column of up to 20480 points in the whole LDS) is unpadded: its stages run
// the odd radices first, so every stage's write stride is odd or at least
// a wavefront wide, and the reads walk consecutive butterflies. Without
// pads every LDS address is a base plus a constant step.
template <int MAXP> __device__ __forceinline__ int lpm(int p, int sh) { return MAXP == 24 ? p : lp(p, sh); }

template <typename T2, int RAD, int MAXP>
__device__ __forceinline__ void lds_stage(T2* lds, const SubFFT& f, const Stage& s, const T2* stw,
                                          int c, int lt, int cdig) {
  constexpr int NB = MAXP / RAD > 0 ? MAXP / RAD : 1;
  constexpr bool NOPAD = MAXP == 24;
  T2* col = lds + c * f.Rp;
  const int nb = s.nb, ns = s.ns, tpc = f.tpc, sh = f.psh;
  // Uniform per stage: reads j + i*nb and writes o + i*ns are affine when
  // nb (ns) is a multiple of the pad block; the first stage's writes (ns = 1)
  // when RAD divides the pad block; always without pads.
  const bool rd_aff = NOPAD || (nb & ((1 << sh) - 1)) == 0;
  const bool wr_aff = NOPAD || (ns & ((1 << sh) - 1)) == 0 || (ns == 1 && ((1 << sh) % RAD) == 0);
  const int rd_step = NOPAD ? nb : nb + (nb >> sh);
  const int wr_step = ns == 1 ? 1 : NOPAD ? ns : ns + (ns >> sh);
  T2 v[NB][RAD];
#pragma unroll
  for (int m = 0; m < NB; ++m) {
    int j = lt + m * tpc;
    if (j < nb) {
      if (rd_aff) {
        const T2* p = col + lpm<MAXP>(j, sh);
#pragma unroll
        for (int i = 0; i < RAD; ++i) v[m][i] = p[i * rd_step];
      } else {
#pragma unroll
        for (int i = 0; i < RAD; ++i) v[m][i] = col[lpm<MAXP>(j + i * nb, sh)];
      }
      if (ns > 1 || s.cmul) {
        int q = (int)fastdiv_q((uint32_t)j, s.nsd);
        int k = j - q * ns;
        const T2* tw = stw + s.tw_off + k * s.kmul + cdig * s.cmul;
#pragma unroll
        for (int i = 1; i < RAD; ++i) v[m][i] = cmul(v[m][i], tw[i - 1]);
      }
      dft<RAD>(v[m]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < NB; ++m) {
    int j = lt + m * tpc;
    if (j < nb) {
      int q = (int)fastdiv_q((uint32_t)j, s.nsd);
      int k = j - q * ns;
      const int o = q * ns * RAD + k;
      if (wr_aff) {
        T2* p = col + lpm<MAXP>(o, sh);
#pragma unroll
        for (int i = 0; i < RAD; ++i) p[i * wr_step] = v[m][i];
      } else {
#pragma unroll
        for (int i = 0; i < RAD; ++i) col[lpm<MAXP>(o + i * ns, sh)] = v[m][i];
      }
    }
  }
  __syncthreads();
}

constexpr int cbrev(int x, int bits) {
  int r = 0;
  for (int b = 0; b < bits; ++b) r = (r << 1) | ((x >> b) & 1);
  return r;
}
constexpr int clog2(int x) { return x <= 1 ? 0 : 1 + clog2(x / 2); }

// TT = 2^t consecutive radix-2 layers of the reference's iterative transform
// (kernels.Radix2), done in registers with the reference's per-layer twiddle
// tables. Layer u pairs registers (i, i + TT/2^(u+1)) in place, using entry
// (k + v*ns)*Nsg + cdig of the table of layer length L = Nsg*ns*2^(u+1),
// v = bitrev_u(i >> (t-u)); register i ends as output bitrev_t(i). This is
// the same set of butterflies, operands and twiddles as t separate layers.
// The entry is read from a copy laid out per column digit (make_sub_r2).
template <typename T2, int TT, int MAXP>
__device__ __forceinline__ void lds_stage_rec(T2* lds, const SubFFT& f, const Stage& s, const T2* stw,
                                              int c, int lt, int cdig) {
  constexpr int t = clog2(TT);
  constexpr int NB = MAXP / TT > 0 ? MAXP / TT : 1;
  T2* col = lds + c * f.Rp;
  const int nb = s.nb, ns = s.ns, tpc = f.tpc, Nsg = s.kmul;
  T2 v[NB][TT];
#pragma unroll
  for (int m = 0; m < NB; ++m) {
    const int j = lt + m * tpc;
    if (j < nb) {
#pragma unroll
      for (int i = 0; i < TT; ++i) v[m][i] = col[lp(j + i * nb, f.psh)];
      const int q = (int)fastdiv_q((uint32_t)j, s.nsd);
      const int k = j - q * ns;
#pragma unroll
      for (int u = 0; u < t; ++u) {
        const int h = TT >> (u + 1);
        // layer u's table (see make_sub_r2): column digit block, then i
        const int off = s.tw_off + Nsg * ns * ((1 << u) - 1) + cdig * (ns << u);
#pragma unroll
        for (int i = 0; i < TT; ++i) {
          if (i & h) continue;
          const int vv = cbrev(i >> (t - u), u);
          const T2 w = stw[off + k + vv * ns];
          const T2 b = cmul(v[m][i + h], w);
          v[m][i + h] = csub(v[m][i], b);
          v[m][i] = cadd(v[m][i], b);
        }
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < NB; ++m) {
    const int j = lt + m * tpc;
    if (j < nb) {
      const int q = (int)fastdiv_q((uint32_t)j, s.nsd);
      const int k = j - q * ns;
      const int o = q * ns * TT + k;
#pragma unroll
      for (int i = 0; i < TT; ++i) col[lp(o + cbrev(i, t) * ns, f.psh)] = v[m][i];
    }
  }
  __syncthreads();
}

// Runs stages [s0, s1) of the R-point sub-FFT (forward) on every column in
// LDS (all of them by default). cdig is the
// calling thread's column twiddle digit (only read by stages with cmul != 0).
template <typename T2, int MAXP>
__device__ __forceinline__ void lds_fft(T2* lds, const SubFFT& f, const T2* stw, int cdig = 0, int s0 = 0, int s1 = -1) {
  const int tid = threadIdx.x;
  const int c = tid >> f.log2tpc;
  const int lt = tid & (f.tpc - 1);
  if (s1 < 0) s1 = f.nstages;
  for (int si = s0; si < s1; ++si) {
    const Stage& s = f.st[si];
    switch (s.radix) {
      case 2: lds_stage<T2, 2, MAXP>(lds, f, s, stw, c, lt, cdig); break;
      case 3: lds_stage<T2, 3, MAXP>(lds, f, s, stw, c, lt, cdig); break;
      case 4: lds_stage<T2, 4, MAXP>(lds, f, s, stw, c, lt, cdig); break;
      case 5: lds_stage<T2, 5, MAXP>(lds, f, s, stw, c, lt, cdig); break;
      case 7: lds_stage<T2, 7, MAXP>(lds, f, s, stw, c, lt, cdig); break;
      case 8: lds_stage<T2, 8, MAXP>(lds, f, s, stw, c, lt, cdig); break;
      case 11: lds_stage<T2, 11, MAXP>(lds, f, s, stw, c, lt, cdig); break;
      case 13: lds_stage<T2, 13, MAXP>(lds, f, s, stw, c, lt, cdig); break;
      // composite prime-factor radices: planned for the 16-point form and
      // the complex64 convolution's wide form
      case 6: if constexpr (MAXP == 16 || MAXP == 24) lds_stage<T2, 6, MAXP>(lds, f, s, stw, c, lt, cdig); break;
      case 10: if constexpr (MAXP == 16 || MAXP == 24) lds_stage<T2, 10, MAXP>(lds, f, s, stw, c, lt, cdig); break;
      case 12: if constexpr (MAXP == 16 || MAXP == 24) lds_stage<T2, 12, MAXP>(lds, f, s, stw, c, lt, cdig); break;
      // grouped radix-2 layers: only the 16-point form is planned with them
      case 102: if constexpr (MAXP == 16) lds_stage_rec<T2, 2, MAXP>(lds, f, s, stw, c, lt, cdig); break;
      case 104: if constexpr (MAXP == 16) lds_stage_rec<T2, 4, MAXP>(lds, f, s, stw, c, lt, cdig); break;
      case 108: if constexpr (MAXP == 16) lds_stage_rec<T2, 8, MAXP>(lds, f, s, stw, c, lt, cdig); break;
      case 116: if constexpr (MAXP == 16) lds_stage_rec<T2, 16, MAXP>(lds, f, s, stw, c, lt, cdig); break;
      default: lds_stage<T2, 16, MAXP>(lds, f, s, stw, c, lt, cdig); break;
    }
  }
}

// lds_fft for sub-FFTs whose stages are all radix 2, 4, 8 or 16 (fewer
// live registers than the general dispatch).
template <typename T2, int MAXP>
__device__ __forceinline__ void lds_fft_pow2(T2* lds, const SubFFT& f, const T2* stw) {
  const int tid = threadIdx.x;
  const int c = tid >> f.log2tpc;
  const int lt = tid & (f.tpc - 1);
  for (int si = 0; si < f.nstages; ++si) {
    const Stage& s = f.st[si];
    switch (s.radix) {
      case 2: lds_stage<T2, 2, MAXP>(lds, f, s, stw, c, lt, 0); break;
      case 4: lds_stage<T2, 4, MAXP>(lds, f, s, stw, c, lt, 0); break;
      case 8: lds_stage<T2, 8, MAXP>(lds, f, s, stw, c, lt, 0); break;
      default: if constexpr (MAXP >= 16) lds_stage<T2, 16, MAXP>(lds, f, s, stw, c, lt, 0); break;
    }
  }
}

// Column g's digits: returns the in-line part sum(d_i * a_i, i < 3); *line
// gets the line part (offset table entry or d3 * a3), *d3 the line index.
__device__ __forceinline__ unsigned long long udiv(unsigned long long n, long long d,
                                                   unsigned long long M) {
  return M ? __umul64hi(n << 8, M) : n / (unsigned long long)d;
}

__device__ __forceinline__ long long col_digits(const ColMap& m, long long g, long long* dtw,
                                               int tw_digit, const long long* off, long long* line,
                                               long long* d3out) {
  unsigned long long t = (unsigned long long)g;
  unsigned long long q = udiv(t, m.D0, m.M0);
  const long long d0 = (long long)(t - q * (unsigned long long)m.D0);
  t = q;
  q = udiv(t, m.D1, m.M1);
  const long long d1 = (long long)(t - q * (unsigned long long)m.D1);
  t = q;
  q = udiv(t, m.D2, m.M2);
  const long long d2 = (long long)(t - q * (unsigned long long)m.D2);
  const long long d3 = (long long)q;
  if (dtw) *dtw = tw_digit == 0 ? d0 : d1;
  if (line) *line = off ? off[d3] : d3 * m.a3;
  if (d3out) *d3out = d3;
  return d0 * m.a0 + d1 * m.a1 + d2 * m.a2;
}

__device__ __forceinline__ double2 pass_twiddle(const double2* hi, const double2* lo, long long e,
                                                int shift) {
  double2 a = hi[e >> shift];
  double2 b = lo[e & ((1ll << shift) - 1)];
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ double2 dmul2(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Pass twiddles W_M^(r k) of a thread's points r = r0 + u * step (u = 0, 1,
// ...) in the column whose twiddle digit is k. TWM selects how they are
// formed (PassArgs::tw_mode):
//  0: FP64 two-level tables, FP64 product and running product, each twiddle
//     rounded once to the data precision (every complex128 pass, and the
//     complex64 default);
//  1: FP32 two-level tables, FP32 product and FP32 running product;
//  2: two-level tables of hi + lo float pairs, each twiddle formed directly
//     in double-float arithmetic (error-free products by FMA), rounded once.
template <typename T2, int TWM> struct PassTw;

template <typename T2> struct PassTw<T2, 0> {
  double2 w = make_double2(1.0, 0.0), st = make_double2(1.0, 0.0);
  __device__ __forceinline__ void init(const PassArgs& a, long long k, int r0, int step) {
    if (k) {
      w = pass_twiddle(a.ptw_hi, a.ptw_lo, (long long)r0 * k, a.tw_shift);
      st = pass_twiddle(a.ptw_hi, a.ptw_lo, (long long)step * k, a.tw_shift);
    }
  }
  __device__ __forceinline__ T2 next() {
    const T2 r = from_d2<T2>(w);
    w = dmul2(w, st);
    return r;
  }
};

template <> struct PassTw<float2, 1> {
  float2 w = make_float2(1.f, 0.f), st = make_float2(1.f, 0.f);
  __device__ __forceinline__ static float2 lookup(const PassArgs& a, long long e) {
    const float2* hi = reinterpret_cast<const float2*>(a.ptwx_hi);
    const float2* lo = reinterpret_cast<const float2*>(a.ptwx_lo);
    return cmul(hi[e >> a.tw_shift], lo[e & ((1ll << a.tw_shift) - 1)]);
  }
  __device__ __forceinline__ void init(const PassArgs& a, long long k, int r0, int step) {
    if (k) {
      w = lookup(a, (long long)r0 * k);
      st = lookup(a, (long long)step * k);
    }
  }
  __device__ __forceinline__ float2 next() {
    const float2 r = w;
    w = cmul(w, st);
    return r;
  }
};

#pragma clang fp contract(off)
// (A_hi + A_lo)(B_hi + B_lo) rounded to float2: the products of the high
// parts exactly (FMA residuals), their sums with their rounding errors, plus
// the cross terms; error about 2^-44 relative before the final rounding.
__device__ __forceinline__ float2 ddf_cmul(float4 A, float4 B) {
  const float p1 = A.x * B.x, e1 = __fmaf_rn(A.x, B.x, -p1);
  const float p2 = A.y * B.y, e2 = __fmaf_rn(A.y, B.y, -p2);
  const float sr = p1 - p2, br = sr - p1, er = (p1 - (sr - br)) + (-p2 - br);
  const float cr = ((e1 - e2) + er) + (A.x * B.z + A.z * B.x - A.y * B.w - A.w * B.y);
  const float p3 = A.x * B.y, e3 = __fmaf_rn(A.x, B.y, -p3);
  const float p4 = A.y * B.x, e4 = __fmaf_rn(A.y, B.x, -p4);
  const float si = p3 + p4, bi = si - p3, ei = (p3 - (si - bi)) + (p4 - bi);
  const float ci = ((e3 + e4) + ei) + (A.x * B.w + A.z * B.y + A.y * B.z + A.w * B.x);
  return make_float2(sr + cr, si + ci);
}
#pragma clang fp contract(on)

template <> struct PassTw<float2, 2> {
  long long e = 0, de = 0;
  const float4* hi = nullptr;
  const float4* lo = nullptr;
  int sh = 0;
  __device__ __forceinline__ void init(const PassArgs& a, long long k, int r0, int step) {
    e = (long long)r0 * k;
    de = (long long)step * k;
    hi = reinterpret_cast<const float4*>(a.ptwx_hi);
    lo = reinterpret_cast<const float4*>(a.ptwx_lo);
    sh = a.tw_shift;
  }
  __device__ __forceinline__ float2 next() {
    const float2 r = ddf_cmul(hi[e >> sh], lo[e & ((1ll << sh) - 1)]);
    e += de;
    return r;
  }
};

extern __shared__ unsigned char g_dyn_lds[];

// ---------------------------------------------------------------------------
// Generic pass: load C columns (optional pass twiddle W_M^(r*k), optional
// conjugation), R-point FFT in LDS, store (optional conjugation and scale).
// Middle launch of a convolution chain whose forward pass and inverse first
// pass are both 256 = 16 x 16 points, 16 threads per column, no hooks. Thread
// (c, lt) holds the inputs lt + 16 i of forward stage 2's butterfly lt, whose
// outputs lt + 16 i are the spectrum points to multiply and, after it, the
// inputs of the inverse's stage-1 butterfly lt. So forward stage 2, the
// multiply and inverse stage 1 stay in registers: two LDS round trips and
// three barriers instead of six and seven. Same arithmetic as the general
// middle launch.
template <typename T>
__device__ __forceinline__ void mid256(const PassArgs& a, typename Vec2<T>::type* lds, const long long* s_ib,
                                       const long long* s_il, const long long* s_ob, const long long* s_ol,
                                       const long long* s_tk, const long long* s_mk, const long long* s_ln) {
  typedef typename Vec2<T>::type T2;
  const SubFFT& f = a.f;
  const int tid = threadIdx.x, C = f.C, Rp = f.Rp, sh = f.psh;
  T2 v[16];
  // load: fixed column per thread (either layout), points r0 + 16 u
  {
    const int ic = a.im.rowmajor ? tid >> 4 : tid & (C - 1);
    const int ir0 = a.im.rowmajor ? tid & 15 : tid >> f.log2C;
    const long long base = s_ib[ic];
    const T2* in = reinterpret_cast<const T2*>(a.in);
    const T2* p = in + (base >= 0 ? base + s_il[ic] : 0) + (long long)ir0 * a.im.g;
    const long long st = 16 * a.im.g;
    with_nt(a.nt & kNtLoad, [&](auto nt) {
#pragma unroll
      for (int u = 0; u < 16; ++u) v[u] = base >= 0 ? ldl<decltype(nt)::value>(p + u * st) : mkc<T2>(0, 0);
    });
    double2 tw_w = make_double2(1.0, 0.0), tw_step = make_double2(1.0, 0.0);
    const long long k = a.tw_M ? s_tk[ic] : 0;
    if (k) {
      tw_w = pass_twiddle(a.ptw_hi, a.ptw_lo, (long long)ir0 * k, a.tw_shift);
      tw_step = pass_twiddle(a.ptw_hi, a.ptw_lo, 16 * k, a.tw_shift);
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      if (k) {
        v[u] = cmul(v[u], from_d2<T2>(tw_w));
        tw_w = dmul2(tw_w, tw_step);
      }
      if (a.conj_in) v[u].y = -v[u].y;
    }
    // forward stage 1 (ns = 1): butterfly ir0, outputs ir0 * 16 + i
    dft<16>(v);
    T2* q = lds + ic * Rp + lp(ir0 * 16, sh);
#pragma unroll
    for (int i = 0; i < 16; ++i) q[i] = v[i];
  }
  __syncthreads();
  const int c = tid >> 4, lt = tid & 15;
  T2* const col = lds + c * Rp;
  const int step = 16 + (16 >> sh);  // lp(lt + 16 i) = lp(lt) + i * step (sh <= 4)
  // forward stage 2 (ns = 16): butterfly lt, inputs and outputs lt + 16 i
  {
    const Stage& s2 = f.st[1];
    const T2* tw = reinterpret_cast<const T2*>(a.stw) + s2.tw_off + lt * s2.kmul;
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = col[lp(lt, sh) + i * step];
#pragma unroll
    for (int i = 1; i < 16; ++i) v[i] = cmul(v[i], tw[i - 1]);
    dft<16>(v);
  }
  // spectrum multiply at forward output positions k = s_mk[c] + r * mm.g,
  // r = lt + 16 i (mm.g >= 1, so k = 0 only at lt = i = 0 of a column with
  // s_mk = 0). Every lane loads its 16 entries before using any; a tail
  // column (s_mk = 0) reads valid entries and discards them.
  const bool live = s_ob[c] >= 0;
  {
    if (live && a.dc && lt == 0 && s_mk[c] == 0) reinterpret_cast<T2*>(a.dc)[s_ln[c]] = v[0];  // Rader: A[0]
    const T2* spec = reinterpret_cast<const T2*>(a.mid_spec) + s_mk[c] + (long long)lt * a.mm.g;
    const long long sst = 16 * a.mm.g;
    T2 w[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) w[i] = spec[i * sst];
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = live ? cconj(cmul(v[i], w[i])) : mkc<T2>(0, 0);
  }
  // inverse stage 1 (ns = 1): butterfly lt has inputs lt + 16 i, the values
  // just computed; outputs lt * 16 + i
  dft<16>(v);
  __syncthreads();  // every thread has read its forward stage-2 inputs
  {
    T2* q = col + lp(lt * 16, sh);
#pragma unroll
    for (int i = 0; i < 16; ++i) q[i] = v[i];
  }
  __syncthreads();
  // inverse stage 2 (ns = 16): butterfly lt, outputs lt + 16 i, stored
  {
    const Stage& s2 = a.f2.st[1];
    const T2* tw = reinterpret_cast<const T2*>(a.stw2) + s2.tw_off + lt * s2.kmul;
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = col[lp(lt, sh) + i * step];
#pragma unroll
    for (int i = 1; i < 16; ++i) v[i] = cmul(v[i], tw[i - 1]);
    dft<16>(v);
  }
  if (!live) return;
  T2* out = reinterpret_cast<T2*>(a.out) + s_ob[c] + s_ol[c] + (long long)lt * a.om.g;
  const long long st = 16 * a.om.g;
  const T sc = (T)a.scale;
  with_nt(a.nt & kNtStore, [&](auto nt) {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      T2 w = v[i];
      if (a.conj_out) w.y = -w.y;
      if (a.scale != 1.0) w = cscale(w, sc);
      stl<decltype(nt)::value>(out + i * st, w);
    }
  });
}

// Two-stage register form of a pass whose sub-FFT is R = A x 16 points, radix
// A first, then radix 16, with 16 threads per column. Thread (c, r0) loads
// the points r0 + 16 u (u < A): exactly the inputs of stage-1 butterfly r0
// (nb = 16), so stage 1 runs from the load registers. Stage 2 (ns = A) gives
// butterfly j < A of a column the outputs j + A i, which go straight to
// memory. One LDS round trip and one barrier instead of three and five.
// Handles the pass twiddles, conjugation, scale, the Bluestein chirp hooks
// (multiply on load / conjugate-multiply on store, zero past nvalid) and, in
// form HK = 2, the Rader hooks (gather by the generator permutation on load;
// on store, x0 plus the conjugated convolution scattered by the inverse
// permutation, and X[0] = x0 + A[0]). Form HK = 1 takes the chirp hooks only
// and issues all of a thread's store-side chirp loads before any is used
// (more registers, fewer round trips).
template <typename T, int A, int HK>
__device__ __forceinline__ void reg2(const PassArgs& a, typename Vec2<T>::type* lds, const long long* s_ib,
                                     const long long* s_il, const long long* s_ob, const long long* s_ol,
                                     const long long* s_tk, const long long* s_ln,
                                     const typename Vec2<T>::type* s_x0) {
  typedef typename Vec2<T>::type T2;
  const SubFFT& f = a.f;
  const int tid = threadIdx.x, C = f.C, Rp = f.Rp, sh = f.psh;
  T2 v[16];
  {
    const int ic = a.im.rowmajor ? tid >> 4 : tid & (C - 1);
    const int ir0 = a.im.rowmajor ? tid & 15 : tid >> f.log2C;
    const long long base = s_ib[ic];
    const long long il = s_il[ic];
    const long long g = a.im.g;
    const T2* in = reinterpret_cast<const T2*>(a.in);
    T2 w[A];
    if (a.in_hook == 0) {
      const T2* p = in + (base >= 0 ? base + il : 0) + (long long)ir0 * g;
      with_nt(a.nt & kNtLoad, [&](auto nt) {
#pragma unroll
        for (int u = 0; u < A; ++u) v[u] = base >= 0 ? ldl<decltype(nt)::value>(p + u * 16 * g) : mkc<T2>(0, 0);
      });
    } else if (a.in_hook == 1) {
      const T2* hw = reinterpret_cast<const T2*>(a.hook_w);
#pragma unroll
      for (int u = 0; u < A; ++u) {
        const long long i = il + (long long)(ir0 + 16 * u) * g;
        const bool live = base >= 0 && i < a.nvalid;
        v[u] = live ? in[base + i] : mkc<T2>(0, 0);
        w[u] = live ? hw[i] : mkc<T2>(0, 0);
      }
    } else if constexpr (HK == 2) {
#pragma unroll
      for (int u = 0; u < A; ++u) {
        const long long i = il + (long long)(ir0 + 16 * u) * g;
        const bool live = base >= 0 && i < a.nvalid;
        v[u] = live ? in[base + (long long)a.hook_perm[i]] : mkc<T2>(0, 0);
      }
    }
    double2 tw_w = make_double2(1.0, 0.0), tw_step = make_double2(1.0, 0.0);
    const long long k = a.tw_M ? s_tk[ic] : 0;
    if (k) {
      tw_w = pass_twiddle(a.ptw_hi, a.ptw_lo, (long long)ir0 * k, a.tw_shift);
      tw_step = pass_twiddle(a.ptw_hi, a.ptw_lo, 16 * k, a.tw_shift);
    }
    // same order as the general form: pass twiddle, conjugation, hook
#pragma unroll
    for (int u = 0; u < A; ++u) {
      if (k) {
        v[u] = cmul(v[u], from_d2<T2>(tw_w));
        tw_w = dmul2(tw_w, tw_step);
      }
      if (a.conj_in) v[u].y = -v[u].y;
      if (a.in_hook == 1) v[u] = cmul(v[u], w[u]);
    }
    // stage 1 (ns = 1): butterfly ir0, outputs ir0 * A + i
    dft<A>(v);
    T2* col = lds + ic * Rp;
#pragma unroll
    for (int i = 0; i < A; ++i) col[lp(ir0 * A + i, sh)] = v[i];
  }
  __syncthreads();
  // stage 2 (radix 16, ns = nb = A): butterfly j of column oc
  const int oc = a.om.rowmajor ? tid >> 4 : tid & (C - 1);
  const int j = a.om.rowmajor ? tid & 15 : tid >> f.log2C;
  const long long obase = s_ob[oc];
  if (j >= A || obase < 0) return;
  {
    const T2* col = lds + oc * Rp;
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = col[lp(j + A * i, sh)];
    const Stage& s2 = f.st[1];
    const T2* tw = reinterpret_cast<const T2*>(a.stw) + s2.tw_off + j * s2.kmul;
#pragma unroll
    for (int i = 1; i < 16; ++i) v[i] = cmul(v[i], tw[i - 1]);
    dft<16>(v);
  }
  T2* out = reinterpret_cast<T2*>(a.out) + obase;
  const long long ol = s_ol[oc];
  const long long g = a.om.g;
  const T sc = (T)a.scale;
  if (a.out_hook == 0) {
    const bool do_scale = a.scale != 1.0;
    with_nt(a.nt & kNtStore, [&](auto nt) {
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        T2 x = v[i];
        if (a.conj_out) x.y = -x.y;
        if (do_scale) x = cscale(x, sc);
        stl<decltype(nt)::value>(out + ol + (long long)(j + A * i) * g, x);
      }
    });
  } else if (HK == 1) {
    // all chirp loads first (index clamped to 0 past nvalid), then the
    // stores of the points below nvalid
    const T2* hw = reinterpret_cast<const T2*>(a.hook_w);
    T2 w[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const long long p = ol + (long long)(j + A * i) * g;
      w[i] = hw[p < a.nvalid ? p : 0];
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const long long p = ol + (long long)(j + A * i) * g;
      if (p < a.nvalid) {
        T2 x = cmul(cconj(v[i]), w[i]);
        if (a.conj_out) x.y = -x.y;
        out[p] = cscale(x, sc);
      }
    }
  } else if (a.out_hook == 1) {
    const T2* hw = reinterpret_cast<const T2*>(a.hook_w);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const long long p = ol + (long long)(j + A * i) * g;
      if (p < a.nvalid) {
        T2 x = cmul(cconj(v[i]), hw[p]);
        if (a.conj_out) x.y = -x.y;
        out[p] = cscale(x, sc);
      }
    }
  } else {
    const T2 x0 = s_x0[oc];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const long long p = ol + (long long)(j + A * i) * g;
      const long long idx = p - a.conv_off;
      if (idx >= 0 && idx < a.nvalid) {
        T2 y = cadd(x0, cconj(v[i]));
        if (a.conj_out) y.y = -y.y;
        out[a.hook_perm[idx]] = cscale(y, sc);
      }
      if (p == 0) {
        T2 y = cadd(x0, reinterpret_cast<const T2*>(a.dc)[s_ln[oc]]);
        if (a.conj_out) y.y = -y.y;
        out[0] = cscale(y, sc);
      }
    }
  }
}

// The real forward untangle (realplan.go:138-169) of the half-length
// transform's last pass, from LDS to the bins. Workgroup slot c < C/2 holds
// column j = s_mk[c] (0 <= j <= Ns/2) and slot c + C/2 its mirror Ns - j, so
// output k = j + r*Ns has its partner m - k at point R-1-r of the mirror
// (columns 0 and Ns/2 are their own mirrors: points R-r and R-1-r). Each pair
// is untangled once and both bins are stored; k = 0 gives DC and Nyquist.
// Threads walk the columns fastest, so stores form contiguous runs.
template <typename T>
__device__ __forceinline__ void untangle_store(const PassArgs& a, const typename Vec2<T>::type* lds,
                                               const long long* s_mk, const long long* s_ln) {
  typedef typename Vec2<T>::type T2;
  const SubFFT& f = a.f;
  const int H = f.C >> 1, log2H = f.log2C - 1, R = f.R, Rp = f.Rp, sh = f.psh;
  const long long m = a.ru_m, Ns = m / R;
  const T2* utw = reinterpret_cast<const T2*>(a.ru_utw);
  auto ut = [&](long long k) -> T2 {
    return a.ru_uhi ? from_d2<T2>(pass_twiddle(a.ru_uhi, a.ru_ulo, k, a.ru_ushift)) : utw[k];
  };
  const T half = (T)0.5;
  auto untangle = [&](T2 za, T2 zb, T2 w) {
    const T xer = (za.x + zb.x) * half, xei = (za.y - zb.y) * half;
    const T xor_ = (za.y + zb.y) * half, xoi = -(za.x - zb.x) * half;
    return mkc<T2>(xer + (w.x * xor_ - w.y * xoi), xei + (w.x * xoi + w.y * xor_));
  };
  auto emit = [&](long long line, long long k, T2 X) {
    const bool edge = k == 0 || k == m;
    if (a.ru_mode == 1) {
      T p = (X.x * X.x + X.y * X.y) * (T)a.ru_psd;
      if (!edge) p *= (T)2;
      reinterpret_cast<T*>(a.out)[line * a.ru_ostride + k] = p;
    } else {
      if (a.ru_mode == 2) X = edge ? mkc<T2>(X.x * (T)a.ru_w_edge, 0) : cscale(X, (T)a.ru_w_mid);
      reinterpret_cast<T2*>(a.out)[line * a.ru_ostride + k] = X;
    }
  };
  for (int idx = threadIdx.x; idx < H * R; idx += blockDim.x) {
    const int c = idx & (H - 1), r = idx >> log2H;
    const long long j = s_mk[c];
    if (j < 0) continue;
    int rp, cp;
    if (j == 0) {
      if (2 * r > R) continue;
      rp = r == 0 ? 0 : R - r;
      cp = c;
    } else if (2 * j == Ns) {
      if (r > R - 1 - r) continue;
      rp = R - 1 - r;
      cp = c;
    } else {
      rp = R - 1 - r;
      cp = c + H;
    }
    const T2 zk = lds[c * Rp + lp(r, sh)];
    const long long line = s_ln[c];
    const long long k = j + (long long)r * Ns;
    if (k == 0) {
      emit(line, 0, mkc<T2>(zk.x + zk.y, 0));
      emit(line, m, mkc<T2>(zk.x - zk.y, 0));
      continue;
    }
    const T2 zp = lds[cp * Rp + lp(rp, sh)];
    const long long kp = m - k;
    emit(line, k, untangle(zk, zp, ut(k)));
    if (kp != k) emit(line, kp, untangle(zp, zk, ut(kp)));
  }
}

// The real inverse pack (realplan.go:221-272) of the half-length transform's
// first pass, in LDS: slots hold raw bins X[j + r*M'] (M' = m/R columns per
// line) with the same column pairing as untangle_store, and each pair
// (k, m - k) becomes Z[k], Z[m - k] in place (Z[0] from X[0] and the Nyquist
// bin X[m], loaded into s_x0). Bins 1..m-1 are scaled by w_mid; Z is
// conjugated when ru_conj (the inverse as conj(FFT(conj))).
template <typename T>
__device__ __forceinline__ void pack_step(const PassArgs& a, typename Vec2<T>::type* lds, const long long* s_mk,
                                          const typename Vec2<T>::type* s_x0) {
  typedef typename Vec2<T>::type T2;
  const SubFFT& f = a.f;
  const int H = f.C >> 1, log2H = f.log2C - 1, R = f.R, Rp = f.Rp, sh = f.psh;
  const long long m = a.ru_m, Mc = m / R;
  const T2* utw = reinterpret_cast<const T2*>(a.ru_utw);
  auto ut = [&](long long k) -> T2 {
    return a.ru_uhi ? from_d2<T2>(pass_twiddle(a.ru_uhi, a.ru_ulo, k, a.ru_ushift)) : utw[k];
  };
  const T half = (T)0.5, wm = (T)a.ru_w_mid;
  auto pack = [&](T2 xk, T2 xmk, T2 w) {
    xk = cscale(xk, wm);
    xmk = cscale(xmk, wm);
    const T xer = (xk.x + xmk.x) * half, xei = (xk.y - xmk.y) * half;
    const T dr = (xk.x - xmk.x) * half, di = (xk.y + xmk.y) * half;
    const T xor_ = w.x * dr + w.y * di, xoi = w.x * di - w.y * dr;
    T2 z = mkc<T2>(xer - xoi, xei + xor_);
    if (a.ru_conj) z.y = -z.y;
    return z;
  };
  for (int idx = threadIdx.x; idx < H * R; idx += blockDim.x) {
    const int c = idx & (H - 1), r = idx >> log2H;
    const long long j = s_mk[c];
    if (j < 0) continue;
    int rp, cp;
    if (j == 0) {
      if (2 * r > R) continue;
      rp = r == 0 ? 0 : R - r;
      cp = c;
    } else if (2 * j == Mc) {
      if (r > R - 1 - r) continue;
      rp = R - 1 - r;
      cp = c;
    } else {
      rp = R - 1 - r;
      cp = c + H;
    }
    T2* const pk = lds + c * Rp + lp(r, sh);
    const long long k = j + (long long)r * Mc;
    if (k == 0) {
      const T x0 = (*pk).x, xm = s_x0[c].x;
      T2 z = mkc<T2>((x0 + xm) * half, (x0 - xm) * half);
      if (a.ru_conj) z.y = -z.y;
      *pk = z;
      continue;
    }
    T2* const pp = lds + cp * Rp + lp(rp, sh);
    const T2 xk = *pk, xmk = *pp;
    const long long kp = m - k;
    *pk = pack(xk, xmk, ut(k));
    if (kp != k) *pp = pack(xmk, xk, ut(kp));
  }
}

// First radix of the two-stage register form of each MODE (0: not that
// form), and its hook form (reg2's HK).
constexpr int reg2_radix(int mode) {
  return mode == 3 || mode == 6 ? 10 : mode == 4 || mode == 7 ? 12 : mode == 5 ? 16 : 0;
}
constexpr int reg2_hk(int mode) { return mode >= 5 ? 1 : 2; }

// MODE 0: plain passes. MODE 1 (EXT): the convolution-chain features
// (load/store hooks, the middle launch's spectrum multiply and second FFT);
// plain passes use the form without them, which compiles to fewer registers.
// MODE 2: a middle launch whose two FFTs are both 256 = 16 x 16 points with
// 16 threads per column and no hooks (mid256 above). MODES 3, 4: the
// two-stage register form (reg2 above) with first radix 10, 12, every hook
// (HK = 2); MODES 6, 7, 5: first radix 10, 12, 16 with the chirp hooks'
// loads issued up front (HK = 1), for complex64 Bluestein chains. MODE 11:
// the last pass of a real forward transform's half-length FFT with the
// untangle fused into its store (untangle_store above). MODE 12: the first
// pass of a real inverse's half-length FFT with the pack fused into its load
// (pack_step above).
// LB: the largest workgroup the instantiation serves. The complex128 forms
// that spill under the 128-register cap of a 1024-thread workgroup (0, 3
// and 4) also have an LB = 256 instantiation for passes of at most 256
// threads, which holds them in registers (126-131 VGPRs).
//
// pass_body is the workgroup's work for column group `blk` of a launch of
// `nblk` groups; k_pass runs it for blockIdx.x, the work-queue kernel for
// the groups it pops. sh holds the per-column tables in static LDS.
template <typename T2>
struct PassShared {
  long long ib[kMaxCols], il[kMaxCols], ob[kMaxCols], ol[kMaxCols], tk[kMaxCols];
  long long mk[kMaxCols], ln[kMaxCols];
  T2 x0[kMaxCols];
};

template <typename T, int MAXP, int MODE, int TWM>
__device__ __forceinline__ void pass_body(const PassArgs& a, long long blk, long long nblk,
                                          PassShared<typename Vec2<T>::type>& sh) {
  constexpr bool EXT = MODE >= 1;
  typedef typename Vec2<T>::type T2;
  T2* lds = reinterpret_cast<T2*>(g_dyn_lds);
  long long* const s_ib = sh.ib;
  long long* const s_il = sh.il;
  long long* const s_ob = sh.ob;
  long long* const s_ol = sh.ol;
  long long* const s_tk = sh.tk;
  long long* const s_mk = sh.mk;
  long long* const s_ln = sh.ln;
  T2* const s_x0 = sh.x0;
  const SubFFT& f = a.f;
  const int in_hook = EXT ? a.in_hook : 0;
  const int out_hook = EXT ? a.out_hook : 0;
  const void* const mid_spec = EXT ? a.mid_spec : nullptr;
  const int tid = threadIdx.x;
  const int C = f.C, R = f.R, Rp = f.Rp;
  // XCD-aware order: the hardware deals workgroup b to XCD b mod 8. When
  // consecutive column groups share cache lines (xcd_group > 1: they split
  // 128-byte runs, or read and write unaligned lines), workgroup b takes
  // group (b mod 8) * q + b / 8 (q = G / 8; the last G mod 8 keep their
  // index), so consecutive groups run on one XCD and share its L2.
  if (a.xcd_group > 1) {
    const long long G = nblk, q = G >> 3;
    if (blk < (q << 3)) blk = (blk & 7) * q + (blk >> 3);
  }
  const long long g0 = blk * C;
  if constexpr (MODE == 11 || MODE == 12) {
    // paired columns: slot c < C/2 takes j = w*C/2 + c (j <= Ns/2), slot
    // c + C/2 its mirror Ns - j (0 < j < Ns/2); Ns = m/R columns per line
    if (tid < C) {
      const int H = C >> 1;
      const long long line = blk / a.ru_wpl, w = blk - line * a.ru_wpl;
      const long long Ns = a.ru_m / R;
      const long long jl = w * H + (tid & (H - 1));
      long long j = -1;
      if (tid < H) {
        if (2 * jl <= Ns) j = jl;
      } else if (jl >= 1 && 2 * jl < Ns) {
        j = Ns - jl;
      }
      if (j >= 0) {
        long long k = 0, lb = 0, d3 = 0;
        s_il[tid] = col_digits(a.im, line * Ns + j, &k, a.tw_digit, a.in_off, &lb, &d3);
        s_ib[tid] = lb;
        s_tk[tid] = k;
        s_ln[tid] = d3;
      } else {
        s_ib[tid] = -1;
        s_il[tid] = 0;
        s_tk[tid] = 0;
        s_ln[tid] = 0;
      }
      s_mk[tid] = tid < H ? j : -1;
      s_x0[tid] = mkc<T2>(0, 0);
      if constexpr (MODE == 12) {
        // the first pass stores like any pass; column 0 also needs X[m]
        long long lb = 0;
        s_ol[tid] = j >= 0 ? col_digits(a.om, line * Ns + j, nullptr, 0, nullptr, &lb, nullptr) : 0;
        s_ob[tid] = j >= 0 ? lb : -1;
        if (j == 0 && a.ru_m < a.nvalid)
          s_x0[tid] = reinterpret_cast<const T2*>(a.in)[s_ib[tid] + s_il[tid] + a.ru_m];
      } else {
        s_ob[tid] = -1;
        s_ol[tid] = 0;
      }
    }
  } else if (tid < C) {
    long long g = g0 + tid;
    if (g < a.ncols) {
      long long k = 0, lb = 0, d3 = 0;
      s_il[tid] = col_digits(a.im, g, &k, a.tw_digit, a.in_off, &lb, &d3);
      s_ib[tid] = lb;
      s_ol[tid] = col_digits(a.om, g, nullptr, 0, a.out_off, &lb, nullptr);
      s_ob[tid] = lb;
      s_tk[tid] = k;
      s_ln[tid] = d3;
      if (mid_spec) s_mk[tid] = col_digits(a.mm, g, nullptr, 0, nullptr, &lb, nullptr);
      // Rader: the column holding point 0 of its line (in-line offset 0)
      // saves x[0] for the last launch, which may write over the input.
      if (in_hook == 2 && s_il[tid] == 0)
        reinterpret_cast<T2*>(a.x0buf)[d3] = reinterpret_cast<const T2*>(a.in)[s_ib[tid]];
      if (out_hook == 2) {
        T2 x0 = reinterpret_cast<const T2*>(a.x0buf)[d3];
        if (a.x0_conj) x0.y = -x0.y;
        s_x0[tid] = x0;
      }
    } else {
      // Tail columns past ncols: marked invalid (base -1); every other
      // per-column value is defined so that no path can index with garbage.
      s_ib[tid] = -1;
      s_ob[tid] = -1;
      s_il[tid] = 0;
      s_ol[tid] = 0;
      s_tk[tid] = 0;
      s_mk[tid] = 0;
      s_ln[tid] = 0;
      s_x0[tid] = mkc<T2>(0, 0);
    }
  }
  __syncthreads();
  if constexpr (MODE == 2) {
    mid256<T>(a, lds, s_ib, s_il, s_ob, s_ol, s_tk, s_mk, s_ln);
    return;
  }
  if constexpr (reg2_radix(MODE) > 0) {
    reg2<T, reg2_radix(MODE), reg2_hk(MODE)>(a, lds, s_ib, s_il, s_ob, s_ol, s_tk, s_ln, s_x0);
    return;
  }
  const T2* in = reinterpret_cast<const T2*>(a.in);
  // Every thread keeps one column and visits its points r0 + u * tpc
  // (R <= MAXP * tpc, guaranteed by the planner), so its addresses are one
  // base plus a fixed step. Walking across columns (adjacent columns
  // contiguous) the column is tid mod C and r0 = tid / C; walking along
  // columns (a column's points contiguous) the column is tid / tpc and
  // r0 = tid mod tpc. Loads and stores each pick the mapping of their layout.
  const int frs = f.tpc;
  const int ic = a.im.rowmajor ? tid >> f.log2tpc : tid & (C - 1);
  const int ir0 = a.im.rowmajor ? tid & (frs - 1) : tid >> f.log2C;
  const int oc = a.om.rowmajor ? tid >> f.log2tpc : tid & (C - 1);
  const int or0 = a.om.rowmajor ? tid & (frs - 1) : tid >> f.log2C;
  // Issue every global load of this thread before using any of them, so that
  // MAXP loads per lane are in flight at once.
  T2 buf[MAXP];
  {
    const long long base = s_ib[ic];
    const long long il = s_il[ic];
    if (base >= 0 && in_hook == 0) {
      const T2* p = in + base + il + (long long)ir0 * a.im.g;
      const long long st = (long long)frs * a.im.g;
      with_nt(a.nt & kNtLoad, [&](auto nt) {
#pragma unroll
        for (int u = 0; u < MAXP; ++u)
          buf[u] = ir0 + u * frs < R ? ldl<decltype(nt)::value>(p + u * st) : mkc<T2>(0, 0);
      });
    } else {
#pragma unroll
      for (int u = 0; u < MAXP; ++u) {
        const int r = ir0 + u * frs;
        buf[u] = mkc<T2>(0, 0);
        const long long i = il + (long long)r * a.im.g;
        if (base >= 0 && r < R && i < a.nvalid)
          buf[u] = in[base + (in_hook == 2 ? (long long)a.hook_perm[i] : i)];  // hooks 1, 2, 3
      }
    }
  }
  // Pass twiddles W_M^(r*k) for this thread's column k digit and points
  // r = ir0 + u * tpc (PassTw: by default one base and one step lookup, then
  // an FP64 running product).
  PassTw<T2, TWM> ptw;
  if (a.tw_M) ptw.init(a, s_tk[ic], ir0, frs);
  // The LDS slots of a column's points r0 + u * tpc are affine in u when tpc
  // is a whole number of pad blocks.
  const bool lds_aff = (frs & ((1 << f.psh) - 1)) == 0;
  const int lstep = frs + (frs >> f.psh);
  T2* const icol = lds + ic * Rp;
  T2* const ocol = lds + oc * Rp;
  // When a thread's loaded points are exactly the inputs j + i * nb (nb = tpc)
  // of first-stage butterfly j = r0, that stage runs from the load registers;
  // likewise a last stage whose outputs k + i * ns (ns = tpc) are the
  // thread's store points writes straight to memory (radix-16 stages only,
  // so that the register arrays stay static).
  const bool first_reg = !EXT && f.st[0].radix == 16 && R == 16 * frs;
  const bool last_reg = !EXT && MODE != 11 && f.nstages >= 2 && f.st[f.nstages - 1].radix == 16 && R == 16 * frs;
#pragma unroll
  for (int u = 0; u < MAXP; ++u) {
    const int r = ir0 + u * frs;
    if (r < R) {
      T2 v = buf[u];
      if (a.tw_M) v = cmul(v, ptw.next());
      if (a.conj_in) v.y = -v.y;
      if (in_hook == 1 && s_ib[ic] >= 0) {
        const long long i = s_il[ic] + (long long)r * a.im.g;
        if (i < a.nvalid) v = cmul(v, reinterpret_cast<const T2*>(a.hook_w)[i]);
      }
      if (first_reg)
        buf[u] = v;
      else if (lds_aff)
        icol[lp(ir0, f.psh) + u * lstep] = v;
      else
        icol[lp(r, f.psh)] = v;
    }
  }
  if (first_reg) {
    // stage 0 (ns = 1, no twiddles): outputs of butterfly ir0 at ir0 * 16 + i,
    // inside one pad block
    dft<16>(buf);
    T2* p = icol + lp(ir0 * 16, f.psh);
#pragma unroll
    for (int i = 0; i < 16; ++i) p[i] = buf[i];
  }
  __syncthreads();
  if constexpr (MODE == 12) {
    pack_step<T>(a, lds, s_mk, s_x0);
    __syncthreads();
  }
  lds_fft<T2, MAXP>(lds, f, reinterpret_cast<const T2*>(a.stw), (int)s_tk[tid >> f.log2tpc], first_reg ? 1 : 0,
                    last_reg ? f.nstages - 1 : f.nstages);
  if constexpr (MODE == 11) {
    untangle_store<T>(a, lds, s_mk, s_ln);  // every stage above ends with a barrier
    return;
  }
  if (mid_spec) {
    // forward output position k of point r of column c: s_mk[c] + r * mm.g
    const T2* spec = reinterpret_cast<const T2*>(mid_spec);
    const int c = tid >> f.log2tpc;
    const int lt = tid & (f.tpc - 1);
    if (s_ob[c] >= 0) {
      // batches of SB points per thread, table loads issued first
      constexpr int SB = sizeof(T) == 8 ? 4 : 8;
      const int tpc = f.tpc;
      for (int r0 = lt; r0 < R; r0 += SB * tpc) {
        T2 w[SB];
#pragma unroll
        for (int b = 0; b < SB; ++b) {
          const int r = r0 + b * tpc;
          w[b] = r < R ? spec[s_mk[c] + (long long)r * a.mm.g] : mkc<T2>(0, 0);
        }
#pragma unroll
        for (int b = 0; b < SB; ++b) {
          const int r = r0 + b * tpc;
          if (r < R) {
            const long long k = s_mk[c] + (long long)r * a.mm.g;
            T2 v = lds[c * Rp + lp(r, f.psh)];
            if (a.dc && k == 0) reinterpret_cast<T2*>(a.dc)[s_ln[c]] = v;  // Rader: A[0]
            lds[c * Rp + lp(r, f.psh)] = cconj(cmul(v, w[b]));
          }
        }
      }
    }
    __syncthreads();
    lds_fft<T2, MAXP>(lds, a.f2, reinterpret_cast<const T2*>(a.stw2), 0);
  }
  T2* out = reinterpret_cast<T2*>(a.out);
  const T sc = (T)a.scale;
  const bool do_scale = a.scale != 1.0;
  const long long obase = s_ob[oc];
  if (last_reg) {
    // last stage from LDS into registers, then straight to memory: butterfly
    // j = or0 (< ns), inputs and outputs or0 + i * tpc
    const Stage& sl = f.st[f.nstages - 1];
    T2 v[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = lds_aff ? ocol[lp(or0, f.psh) + i * lstep] : ocol[lp(or0 + i * frs, f.psh)];
    const T2* tw = reinterpret_cast<const T2*>(a.stw) + sl.tw_off + or0 * sl.kmul;
#pragma unroll
    for (int i = 1; i < 16; ++i) v[i] = cmul(v[i], tw[i - 1]);
    dft<16>(v);
    if (obase >= 0) {
      T2* po = out + obase + s_ol[oc] + (long long)or0 * a.om.g;
      const long long st = (long long)frs * a.om.g;
      with_nt(a.nt & kNtStore, [&](auto nt) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          T2 w = v[i];
          if (a.conj_out) w.y = -w.y;
          if (do_scale) w = cscale(w, sc);
          stl<decltype(nt)::value>(po + i * st, w);
        }
      });
    }
    return;
  }
  if (obase < 0) return;
  if (out_hook == 0) {
    T2* po = out + obase + s_ol[oc] + (long long)or0 * a.om.g;
    const long long st = (long long)frs * a.om.g;
    with_nt(a.nt & kNtStore, [&](auto nt) {
#pragma unroll
      for (int u = 0; u < MAXP; ++u) {
        const int r = or0 + u * frs;
        if (r < R) {
          T2 v = lds_aff ? ocol[lp(or0, f.psh) + u * lstep] : ocol[lp(r, f.psh)];
          if (a.conj_out) v.y = -v.y;
          if (do_scale) v = cscale(v, sc);
          stl<decltype(nt)::value>(po + u * st, v);
        }
      }
    });
    return;
  }
  const T2 x0 = out_hook == 2 ? s_x0[oc] : mkc<T2>(0, 0);
  // Hooked stores in batches of SB points: the chirp values or permutation
  // entries of a batch are loaded (index clamped to 0 where unused) before
  // any is used.
  constexpr int SB = sizeof(T) == 8 ? 4 : 8;
#pragma unroll
  for (int u0 = 0; u0 < MAXP; u0 += SB) {
    T2 w[SB];
    int ip[SB];
#pragma unroll
    for (int b = 0; b < SB; ++b) {
      const int r = or0 + (u0 + b) * frs;
      const long long p = s_ol[oc] + (long long)r * a.om.g;
      if (out_hook == 1) {
        w[b] = reinterpret_cast<const T2*>(a.hook_w)[r < R && p < a.nvalid ? p : 0];
      } else {
        const long long idx = p - a.conv_off;
        ip[b] = r < R && idx >= 0 && idx < a.nvalid ? a.hook_perm[idx] : 0;
      }
    }
#pragma unroll
    for (int b = 0; b < SB; ++b) {
      const int r = or0 + (u0 + b) * frs;
      if (r < R) {
        T2 v = ocol[lp(r, f.psh)];
        const long long p = s_ol[oc] + (long long)r * a.om.g;
        if (out_hook == 1) {
          if (p < a.nvalid) {
            v = cmul(cconj(v), w[b]);
            if (a.conj_out) v.y = -v.y;
            out[obase + p] = cscale(v, sc);
          }
        } else {
          const long long idx = p - a.conv_off;
          if (idx >= 0 && idx < a.nvalid) {
            T2 y = cadd(x0, cconj(v));
            if (a.conj_out) y.y = -y.y;
            out[obase + ip[b]] = cscale(y, sc);
          }
          if (p == 0) {
            T2 y = cadd(x0, reinterpret_cast<const T2*>(a.dc)[s_ln[oc]]);
            if (a.conj_out) y.y = -y.y;
            out[obase] = cscale(y, sc);
          }
        }
      }
    }
  }
}

template <typename T, int MAXP, int MODE, int LB = 1024, int TWM = 0>
__global__ void __launch_bounds__(LB) k_pass(PassArgs a) {
  __shared__ PassShared<typename Vec2<T>::type> sh;
  pass_body<T, MAXP, MODE, TWM>(a, blockIdx.x, gridDim.x, sh);
}

// ---------------------------------------------------------------------------
// Real forward, single pass: line l holds 2R real samples; Z = FFT_R(pairs),
// then the untangle to R+1 bins (mode 0) or the one-sided PSD (mode 1).

// Untangle of the transformed column `col` (line l) into R + 1 bins, and
// their stores: X_k = (Z_k + conj Z_(R-k)) / 2 + W_n^k (Z_k - conj Z_(R-k)) / 2i.
template <typename T, typename T2>
__device__ __forceinline__ void real_fwd_bins(const RealFwdArgs& a, const T2* col, const T2* utw, long long l,
                                              int lt, int tpc, int R, int sh) {
  auto bin = [&](int k) -> T2 {
    if (k == 0 || k == R) {
      const T2 z0 = col[0];
      return mkc<T2>(k == 0 ? z0.x + z0.y : z0.x - z0.y, 0);
    }
    const T2 zk = col[lp(k, sh)], zmk = col[lp(R - k, sh)];
    const T xer = (zk.x + zmk.x) * (T)0.5;
    const T xei = (zk.y - zmk.y) * (T)0.5;
    const T xor_ = (zk.y + zmk.y) * (T)0.5;
    const T xoi = -(zk.x - zmk.x) * (T)0.5;
    const T2 w = utw[k];
    const T tr = w.x * xor_ - w.y * xoi;
    const T ti = w.x * xoi + w.y * xor_;
    return mkc<T2>(xer + tr, xei + ti);
  };
  auto finish = [&](int k, T2 X) -> T2 {  // mode 2 weights; mode 1 power in .x
    const bool edge = k == 0 || k == R;  // imaginary part exactly 0 at DC and Nyquist
    if (a.mode == 2) return edge ? mkc<T2>(X.x * (T)a.w_edge, 0) : cscale(X, (T)a.w_mid);
    if (a.mode == 1) {
      T pw = (X.x * X.x + X.y * X.y) * (T)a.psd_scale;
      if (!edge) pw *= (T)2;
      return mkc<T2>(pw, 0);
    }
    return X;
  };
  // Bins k0 and k0 + 1 per thread and step, written by one store: adjacent
  // lanes cover adjacent pairs (the rows need not be aligned to the pair;
  // gfx950 vector stores need only element alignment).
  typedef T Tu2 __attribute__((ext_vector_type(2), aligned(sizeof(T))));
  typedef T Tu4 __attribute__((ext_vector_type(4), aligned(sizeof(T))));
  for (int k0 = 2 * lt; k0 <= R; k0 += 2 * tpc) {
    const bool two = k0 + 1 <= R;
    const T2 X0 = finish(k0, bin(k0));
    const T2 X1 = two ? finish(k0 + 1, bin(k0 + 1)) : mkc<T2>(0, 0);
    if (a.mode == 1) {
      T* o = reinterpret_cast<T*>(a.out) + l * a.out_stride + k0;
      if (two) {
        Tu2 v;
        v.x = X0.x;
        v.y = X1.x;
        *reinterpret_cast<Tu2*>(o) = v;
      } else {
        o[0] = X0.x;
      }
    } else {
      T2* o = reinterpret_cast<T2*>(a.out) + l * a.out_stride + k0;
      if (two) {
        Tu4 v;
        v.x = X0.x;
        v.y = X0.y;
        v.z = X1.x;
        v.w = X1.y;
        *reinterpret_cast<Tu4*>(o) = v;
      } else {
        o[0] = X0;
      }
    }
  }
}

// Loads line l's sample pairs r = lt + u * tpc into buf (zeros past R and
// for lines past the batch).
template <typename T, typename T2, int MAXP>
__device__ __forceinline__ void real_fwd_load(const RealFwdArgs& a, long long l, int lt, int tpc, int R,
                                              T2 (&buf)[MAXP]) {
  const bool live = l < a.nlines;
  const T* p = reinterpret_cast<const T*>(a.in) + (live ? l * a.in_stride : 0) + 2 * (long long)lt;
  with_nt(a.nt & kNtLoad, [&](auto nt) {
#pragma unroll
    for (int u = 0; u < MAXP; ++u) {
      const int r = lt + u * tpc;
      buf[u] = mkc<T2>(0, 0);
      if (live && r < R) {
        const T* q = p + 2 * (long long)u * tpc;
        buf[u] = a.vec_in ? ldl<decltype(nt)::value>(reinterpret_cast<const T2*>(q)) : mkc<T2>(q[0], q[1]);
      }
    }
  });
}

// Windowed sample pairs into the thread's column of LDS.
template <typename T2, int MAXP>
__device__ __forceinline__ void real_fwd_put(T2* col, const T2 (&buf)[MAXP], const T2* win, bool has_win, int lt,
                                             int tpc, int R, int sh) {
  // lp(lt + u * tpc) is affine in u when tpc is a whole number of pad blocks
  const bool aff = (tpc & ((1 << sh) - 1)) == 0;
  const int step = tpc + (tpc >> sh);
#pragma unroll
  for (int u = 0; u < MAXP; ++u) {
    const int r = lt + u * tpc;
    if (r < R) {
      T2 v = buf[u];
      if (has_win) {
        const T2 w = win[r];
        v = mkc<T2>(v.x * w.x, v.y * w.y);
      }
      if (aff)
        col[lp(lt, sh) + u * step] = v;
      else
        col[lp(r, sh)] = v;
    }
  }
}

template <typename T, int MAXP>
__global__ void __launch_bounds__(1024) k_real_fwd(RealFwdArgs a) {
  typedef typename Vec2<T>::type T2;
  T2* lds = reinterpret_cast<T2*>(g_dyn_lds);
  const SubFFT& f = a.f;
  const int tid = threadIdx.x;
  const int R = f.R, Rp = f.Rp, tpc = f.tpc, sh = f.psh;
  // Each thread keeps one line: c = tid / tpc, points r = lt + u * tpc
  // (adjacent lanes read adjacent sample pairs).
  const int c = tid >> f.log2tpc;
  const int lt = tid & (tpc - 1);
  const long long l = (long long)blockIdx.x * f.C + c;
  T2* const col = lds + c * Rp;
  T2 buf[MAXP];
  real_fwd_load<T, T2, MAXP>(a, l, lt, tpc, R, buf);
  real_fwd_put<T2, MAXP>(col, buf, reinterpret_cast<const T2*>(a.window), a.window != nullptr, lt, tpc, R, sh);
  __syncthreads();
  lds_fft<T2, MAXP>(lds, f, reinterpret_cast<const T2*>(a.stw));
  if (l >= a.nlines) return;
  real_fwd_bins<T, T2>(a, col, reinterpret_cast<const T2*>(a.utw), l, lt, tpc, R, sh);
}

// Table form of k_real_fwd (RealFwdArgs::lds_tab): the stage and untangle
// twiddles and the window are read from LDS, so the only global loads inside
// the loop are the next line group's samples, which stay in flight while the
// current group is transformed.
template <typename T, int MAXP>
__global__ void __launch_bounds__(512) k_real_fwd_tab(RealFwdArgs a) {
  typedef typename Vec2<T>::type T2;
  T2* lds = reinterpret_cast<T2*>(g_dyn_lds);
  const SubFFT& f = a.f;
  const int tid = threadIdx.x;
  const int R = f.R, Rp = f.Rp, tpc = f.tpc, sh = f.psh;
  const int c = tid >> f.log2tpc;
  const int lt = tid & (tpc - 1);
  const long long ngroups = (a.nlines + f.C - 1) / f.C;
  const long long gstep = gridDim.x;
  long long g = blockIdx.x;
  T2* const col = lds + c * Rp;
  T2 buf[MAXP];
  real_fwd_load<T, T2, MAXP>(a, g * f.C + c, lt, tpc, R, buf);
  // tables after the columns: stage twiddles, untangle twiddles, window
  T2* const tab = lds + f.C * Rp;
  const bool has_win = a.window != nullptr;
  {
    const T2* stw = reinterpret_cast<const T2*>(a.stw);
    const T2* utw = reinterpret_cast<const T2*>(a.utw);
    const T2* win = reinterpret_cast<const T2*>(a.window);
    const int nthr = blockDim.x;
    for (int i = tid; i < a.stw_n; i += nthr) tab[i] = stw[i];
    for (int i = tid; i <= R; i += nthr) tab[a.stw_n + i] = utw[i];
    if (has_win)
      for (int i = tid; i < R; i += nthr) tab[a.stw_n + R + 1 + i] = win[i];
  }
  const T2* const stw_l = tab;
  const T2* const utw_l = tab + a.stw_n;
  const T2* const win_l = tab + a.stw_n + R + 1;
  __syncthreads();
  for (; g < ngroups; g += gstep) {
    real_fwd_put<T2, MAXP>(col, buf, win_l, has_win, lt, tpc, R, sh);
    const long long l = g * f.C + c;
    if (g + gstep < ngroups) real_fwd_load<T, T2, MAXP>(a, (g + gstep) * f.C + c, lt, tpc, R, buf);
    __syncthreads();
    lds_fft_pow2<T2, MAXP>(lds, f, stw_l);
    if (l < a.nlines) real_fwd_bins<T, T2>(a, col, utw_l, l, lt, tpc, R, sh);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Real inverse, single pass: R+1 bins -> packed spectrum Z -> IFFT_R (via
// conj(FFT(conj))) -> 2R real samples scaled by 1/R.
// Loads line l's bins r = lt + u * tpc < R into buf, bins past in_bins and
// lines past the batch as zero, every bin but DC times w_mid; *xm gets the
// Nyquist bin's real part (lt = 0 only).
template <typename T, typename T2, int MAXP>
__device__ __forceinline__ void real_inv_load(const RealInvArgs& a, long long l, int lt, int tpc, int R,
                                              T2 (&buf)[MAXP], T* xm) {
  const bool live = l < a.nlines;
  const T2* in = reinterpret_cast<const T2*>(a.in) + (live ? l * a.in_stride : 0);
#pragma unroll
  for (int u = 0; u < MAXP; ++u) {
    const int r = lt + u * tpc;
    buf[u] = mkc<T2>(0, 0);
    if (live && r < R && r < a.in_bins) buf[u] = in[r];
    if (r > 0) buf[u] = cscale(buf[u], (T)a.w_mid);
  }
  *xm = 0;
  if (lt == 0 && live && R < a.in_bins) *xm = in[R].x;
}

template <typename T2, int MAXP>
__device__ __forceinline__ void real_inv_put(T2* X, const T2 (&buf)[MAXP], int lt, int tpc, int R, int sh) {
  const bool aff = (tpc & ((1 << sh) - 1)) == 0;
  const int step = tpc + (tpc >> sh);
#pragma unroll
  for (int u = 0; u < MAXP; ++u) {
    const int r = lt + u * tpc;
    if (r < R) {
      if (aff)
        X[lp(lt, sh) + u * step] = buf[u];
      else
        X[lp(r, sh)] = buf[u];
    }
  }
}

// Pairs (k, R-k) for k = 0 .. R/2, each owned by one thread; k = 0 pairs DC
// with the Nyquist bin's real part xm. Leaves conj(Z) in X, Z the packed
// half-length spectrum, for the inverse by a forward transform.
template <typename T, typename T2, typename XM>
__device__ __forceinline__ void real_inv_pack(T2* X, const T2* utw, XM nyq, int lt, int tpc, int R, int sh) {
  for (int k = lt; k <= R / 2; k += tpc) {
    if (k == 0) {
      const T x0 = X[0].x;
      const T xm = nyq();
      X[0] = mkc<T2>((x0 + xm) * (T)0.5, -((x0 - xm) * (T)0.5));
      continue;
    }
    const int mk_ = R - k;
    T2 xk = X[lp(k, sh)], xmk = X[lp(mk_, sh)];
    // Z[k] = xe + i*xo with xe = (Xk + conj Xmk)/2, d = (Xk - conj Xmk)/2,
    // xo = conj(W^k) d; Z[m-k] symmetric.
    T2 w = utw[k];
    T xer = (xk.x + xmk.x) * (T)0.5, xei = (xk.y - xmk.y) * (T)0.5;
    T dr = (xk.x - xmk.x) * (T)0.5, di = (xk.y + xmk.y) * (T)0.5;
    T xor_ = w.x * dr + w.y * di, xoi = w.x * di - w.y * dr;
    T2 zk = mkc<T2>(xer - xoi, xei + xor_);
    if (k != mk_) {
      T2 zmk = mkc<T2>(xer + xoi, -(xei - xor_));
      X[lp(mk_, sh)] = cconj(zmk);
    }
    X[lp(k, sh)] = cconj(zk);
  }
}

// Samples of line l from the transformed column: conj, times 1/m.
template <typename T, typename T2>
__device__ __forceinline__ void real_inv_store(const RealInvArgs& a, const T2* X, long long l, int lt, int tpc,
                                               int R, int sh) {
  // out_stride is even (n = 2R), so each sample pair is one aligned T2
  T2* out = reinterpret_cast<T2*>(reinterpret_cast<T*>(a.out) + l * a.out_stride);
  const T sc = (T)a.scale;
  with_nt(a.nt & kNtStore, [&](auto nt) {
    for (int r = lt; r < R; r += tpc) {
      const T2 z = X[lp(r, sh)];
      stl<decltype(nt)::value>(out + r, mkc<T2>(z.x * sc, -z.y * sc));
    }
  });
}

template <typename T, int MAXP>
__global__ void __launch_bounds__(1024) k_real_inv(RealInvArgs a) {
  typedef typename Vec2<T>::type T2;
  T2* lds = reinterpret_cast<T2*>(g_dyn_lds);
  const SubFFT& f = a.f;
  const int tid = threadIdx.x;
  const int R = f.R, Rp = f.Rp, tpc = f.tpc, sh = f.psh;
  // Each thread keeps one line: c = tid / tpc, points r = lt + u * tpc.
  const int c = tid >> f.log2tpc;
  const int lt = tid & (tpc - 1);
  const long long l = (long long)blockIdx.x * f.C + c;
  const bool live = l < a.nlines;
  const T2* in = reinterpret_cast<const T2*>(a.in) + (live ? l * a.in_stride : 0);
  const T2* utw = reinterpret_cast<const T2*>(a.utw);
  T2* const X = lds + c * Rp;
  T2 buf[MAXP];
#pragma unroll
  for (int u = 0; u < MAXP; ++u) {
    const int r = lt + u * tpc;
    buf[u] = mkc<T2>(0, 0);
    if (live && r < R && r < a.in_bins) buf[u] = in[r];
    if (r > 0) buf[u] = cscale(buf[u], (T)a.w_mid);
  }
  real_inv_put<T2, MAXP>(X, buf, lt, tpc, R, sh);
  __syncthreads();
  real_inv_pack<T, T2>(
      X, utw, [&] { return live && R < a.in_bins ? in[R].x : (T)0; }, lt, tpc, R, sh);
  __syncthreads();
  lds_fft<T2, MAXP>(lds, f, reinterpret_cast<const T2*>(a.stw));
  if (!live) return;
  real_inv_store<T, T2>(a, X, l, lt, tpc, R, sh);
}

// Table form of k_real_inv (RealInvArgs::lds_tab), as k_real_fwd_tab: the
// stage and pack twiddles in LDS, the next line group's bins loaded while
// the current group is transformed.
template <typename T, int MAXP>
__global__ void __launch_bounds__(512) k_real_inv_tab(RealInvArgs a) {
  typedef typename Vec2<T>::type T2;
  T2* lds = reinterpret_cast<T2*>(g_dyn_lds);
  const SubFFT& f = a.f;
  const int tid = threadIdx.x;
  const int R = f.R, Rp = f.Rp, tpc = f.tpc, sh = f.psh;
  const int c = tid >> f.log2tpc;
  const int lt = tid & (tpc - 1);
  const long long ngroups = (a.nlines + f.C - 1) / f.C;
  const long long gstep = gridDim.x;
  long long g = blockIdx.x;
  T2* const X = lds + c * Rp;
  T2 buf[MAXP];
  T xm;
  real_inv_load<T, T2, MAXP>(a, g * f.C + c, lt, tpc, R, buf, &xm);
  // tables after the columns: stage twiddles, then pack twiddles k <= R/2
  T2* const tab = lds + f.C * Rp;
  {
    const T2* stw = reinterpret_cast<const T2*>(a.stw);
    const T2* utw = reinterpret_cast<const T2*>(a.utw);
    const int nthr = blockDim.x;
    for (int i = tid; i < a.stw_n; i += nthr) tab[i] = stw[i];
    for (int i = tid; i <= R / 2; i += nthr) tab[a.stw_n + i] = utw[i];
  }
  const T2* const stw_l = tab;
  const T2* const utw_l = tab + a.stw_n;
  __syncthreads();
  for (; g < ngroups; g += gstep) {
    real_inv_put<T2, MAXP>(X, buf, lt, tpc, R, sh);
    const T xmc = xm;
    const long long l = g * f.C + c;
    if (g + gstep < ngroups) real_inv_load<T, T2, MAXP>(a, (g + gstep) * f.C + c, lt, tpc, R, buf, &xm);
    __syncthreads();
    real_inv_pack<T, T2>(X, utw_l, [&] { return xmc; }, lt, tpc, R, sh);
    __syncthreads();
    lds_fft_pow2<T2, MAXP>(lds, f, stw_l);
    if (l < a.nlines) real_inv_store<T, T2>(a, X, l, lt, tpc, R, sh);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Fused convolution (Rader / Bluestein), C transforms per workgroup.
// NT: the largest workgroup the instantiation serves (complex128 also has
// NT = 256, which lifts the 128-register cap it spills under at 1024).
// conv_body is the workgroup's work for line group `blk`; k_conv runs it for
// blockIdx.x, the work-queue kernel for the groups it pops.
template <typename T, int MAXP>
__device__ __forceinline__ void conv_body(const ConvArgs& a, long long blk) {
  typedef typename Vec2<T>::type T2;
  T2* lds = reinterpret_cast<T2*>(g_dyn_lds);
  const SubFFT& f = a.f;
  const int tid = threadIdx.x;
  const int C = f.C, R = f.R, Rp = f.Rp;
  const long long l0 = blk * C;
  const T2* in = reinterpret_cast<const T2*>(a.in);
  const T2* chirp = reinterpret_cast<const T2*>(a.chirp);
  const int n = a.n;
  const int q = n - 1;
  // Each thread keeps one line, c = tid / tpc, and loads its points
  // r = lt + u * tpc in rounds of LB buffered loads (one round unless R
  // exceeds 16 * tpc, which only the 24-point form allows).
  const int c = tid >> f.log2tpc;
  const int lt = tid & (f.tpc - 1);
  const int tpc = f.tpc;
  // Rader's x0 = x[0] of this thread's line, read with the other input loads:
  // every read of the input ends before the first barrier and every store of
  // the output starts after the last, so an output aliasing the input is safe.
  T2 x0 = mkc<T2>(0, 0);
  {
    constexpr int LB = 16;
    const long long l = l0 + c;
    const bool live = l < a.nlines;
    const T2* x = in + (live ? (a.in_off ? a.in_off[l] : l * a.in_stride) : 0);
    if (live && a.kind != 2) {
      x0 = x[0];
      if (a.conj_in) x0.y = -x0.y;
    }
    const int rounds = MAXP == LB ? 1 : (R + LB * tpc - 1) / (LB * tpc);
    if (a.kind == 2) {
      // Bluestein: x[r] times the chirp, zero past n
      for (int rd = 0; rd < rounds; ++rd) {
        const int r0 = lt + rd * LB * tpc;
        T2 buf[LB];
#pragma unroll
        for (int u = 0; u < LB; ++u) {
          const int r = r0 + u * tpc;
          buf[u] = mkc<T2>(0, 0);
          if (live && r < n) buf[u] = x[r];
        }
#pragma unroll
        for (int u = 0; u < LB; ++u) {
          const int r = r0 + u * tpc;
          if (r < R) {
            T2 v = buf[u];
            if (a.conj_in) v.y = -v.y;
            if (r < n) v = cmul(v, chirp[r]);
            lds[c * Rp + lpm<MAXP>(r, f.psh)] = v;
          }
        }
      }
    } else {
      // Rader: point k of the convolution input is x[perm[k]] (k < q, zero
      // beyond). The line is read in order, x[i] for i = 1..q, each value
      // placed at k = dlog[i - 1]: contiguous loads instead of a gather
      // that touches a different cache line per point.
      for (int r = q + lt; r < R; r += tpc) lds[c * Rp + lpm<MAXP>(r, f.psh)] = mkc<T2>(0, 0);
      for (int rd = 0; rd < rounds; ++rd) {
        const int i0 = 1 + lt + rd * LB * tpc;
        T2 buf[LB];
        int pos[LB];
#pragma unroll
        for (int u = 0; u < LB; ++u) {
          const int i = i0 + u * tpc;
          buf[u] = mkc<T2>(0, 0);
          pos[u] = 0;
          if (live && i < n) {
            buf[u] = x[i];
            pos[u] = a.dlog[i - 1];
          }
        }
#pragma unroll
        for (int u = 0; u < LB; ++u) {
          const int i = i0 + u * tpc;
          if (i < n) {
            T2 v = buf[u];
            if (a.conj_in) v.y = -v.y;
            lds[c * Rp + lpm<MAXP>(pos[u], f.psh)] = v;
          }
        }
      }
    }
  }
  __syncthreads();
  const T2* stw = reinterpret_cast<const T2*>(a.stw);
  lds_fft<T2, MAXP>(lds, f, stw);
  // Rader DC: X[0] = x[0] + sum of the gathered samples, and that sum is the
  // forward transform's A[0]. The thread with lt == 0 reads it before it
  // multiplies point 0 itself.
  const T2 colsum = a.kind != 2 && lt == 0 ? lds[c * Rp] : mkc<T2>(0, 0);
  // Spectrum multiply in batches of SB points per thread: every table load
  // of a batch is issued before any is used.
  constexpr int SB = 8;
  const T2* spec = reinterpret_cast<const T2*>(a.spec);
  for (int r0 = lt; r0 < R; r0 += SB * tpc) {
    T2 w[SB];
#pragma unroll
    for (int b = 0; b < SB; ++b) {
      const int r = r0 + b * tpc;
      w[b] = r < R ? spec[r] : mkc<T2>(0, 0);
    }
#pragma unroll
    for (int b = 0; b < SB; ++b) {
      const int r = r0 + b * tpc;
      if (r < R) {
        T2* q = lds + c * Rp + lpm<MAXP>(r, f.psh);
        *q = cconj(cmul(*q, w[b]));
      }
    }
  }
  __syncthreads();
  lds_fft<T2, MAXP>(lds, f, stw);
  const long long l = l0 + c;
  const bool live = l < a.nlines;
  T2* out = reinterpret_cast<T2*>(a.out) + (live ? (a.out_off ? a.out_off[l] : l * a.out_stride) : 0);
  const T sc = (T)a.scale;
  // Output in batches of SB points per thread, table loads (chirp or
  // inverse permutation) issued first.
  if (a.kind == 2) {
    for (int k0 = lt; k0 < n && live; k0 += SB * tpc) {
      T2 w[SB];
#pragma unroll
      for (int b = 0; b < SB; ++b) {
        const int k = k0 + b * tpc;
        w[b] = k < n ? chirp[k] : mkc<T2>(0, 0);
      }
#pragma unroll
      for (int b = 0; b < SB; ++b) {
        const int k = k0 + b * tpc;
        if (k < n) {
          T2 v = cmul(cconj(lds[c * Rp + lpm<MAXP>(k, f.psh)]), w[b]);
          if (a.conj_out) v.y = -v.y;
          out[k] = cscale(v, sc);
        }
      }
    }
  } else {
    // X[p] = x0 + conv[k + off] for the k with iperm[k] = p (p = 1..q),
    // written in output order: each thread gathers its outputs' entries
    // from LDS through oinv[p - 1] = k, so the stores to the line are
    // contiguous rather than scattered by the permutation (which left
    // partly written cache lines and about 3.4x the line's bytes of HBM
    // writes).
    const int off = a.kind == 1 ? q : 0;
    for (int p0 = lt; p0 < q && live; p0 += SB * tpc) {
      int kk[SB];
#pragma unroll
      for (int b = 0; b < SB; ++b) {
        const int p = p0 + b * tpc;
        kk[b] = p < q ? a.oinv[p] : 0;
      }
#pragma unroll
      for (int b = 0; b < SB; ++b) {
        const int p = p0 + b * tpc;
        if (p < q) {
          T2 v = cadd(x0, cconj(lds[c * Rp + lpm<MAXP>(kk[b] + off, f.psh)]));
          if (a.conj_out) v.y = -v.y;
          out[p + 1] = cscale(v, sc);
        }
      }
    }
    if (lt == 0 && live) {
      T2 v = cadd(x0, colsum);
      if (a.conj_out) v.y = -v.y;
      out[0] = cscale(v, sc);
    }
  }
}

template <typename T, int MAXP, int NT = 1024>
__global__ void __launch_bounds__(NT) k_conv(ConvArgs a) {
  conv_body<T, MAXP>(a, blockIdx.x);
}

// Everything below but the work-queue section is compiled in this file's
// own translation unit; queue.hip compiles the file again with
// GOFFT_QUEUE_TU defined for the work-queue kernels alone (the kernel
// templates above are instantiated only where they are launched).
#ifndef GOFFT_QUEUE_TU

// ---------------------------------------------------------------------------
// Element-wise helpers.

__device__ __forceinline__ long long line_base(const long long* off, long long l, long long stride) {
  return off ? off[l] : l * stride;
}

template <typename T>
__global__ void k_cmul_table(const void* in_, void* out_, const void* tab_, long long n,
                             long long nlines, long long is, long long os, int conj_in, int conj_out,
                             double scale, const long long* ioff, const long long* ooff) {
  typedef typename Vec2<T>::type T2;
  const T2* in = reinterpret_cast<const T2*>(in_);
  const T2* tab = reinterpret_cast<const T2*>(tab_);
  T2* out = reinterpret_cast<T2*>(out_);
  long long total = n * nlines;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    long long l = i / n, k = i - l * n;
    T2 v = in[line_base(ioff, l, is) + k];
    if (conj_in) v.y = -v.y;
    v = cmul(v, tab[k]);
    if (conj_out) v.y = -v.y;
    if (scale != 1.0) v = cscale(v, (T)scale);
    out[line_base(ooff, l, os) + k] = v;
  }
}

template <typename T>
__global__ void k_rader_gather(const void* in_, void* out_, const int* perm, long long q, long long cl,
                               long long nlines, long long is, long long os, int conj_in,
                               const long long* ioff) {
  typedef typename Vec2<T>::type T2;
  const T2* in = reinterpret_cast<const T2*>(in_);
  T2* out = reinterpret_cast<T2*>(out_);
  long long total = cl * nlines;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    long long l = i / cl, k = i - l * cl;
    T2 v = mkc<T2>(0, 0);
    if (k < q) {
      v = in[line_base(ioff, l, is) + perm[k]];
      if (conj_in) v.y = -v.y;
    }
    out[l * os + k] = v;
  }
}

// One workgroup per line: scatter x0 + conv, and the DC sum by a tree.
template <typename T>
__global__ void __launch_bounds__(256) k_rader_scatter(const void* x_, const void* conv_, void* out_,
                                                       const int* iperm, long long n, long long off,
                                                       long long xs, long long cs, long long os,
                                                       int conj_x, int conj_out, double scale,
                                                       const long long* xoff, const long long* ooff) {
  typedef typename Vec2<T>::type T2;
  __shared__ T2 red[256];
  const long long l = blockIdx.x;
  const T2* x = reinterpret_cast<const T2*>(x_) + line_base(xoff, l, xs);
  const T2* conv = reinterpret_cast<const T2*>(conv_) + l * cs;
  T2* out = reinterpret_cast<T2*>(out_) + line_base(ooff, l, os);
  T2 x0 = x[0];
  const T sc = (T)scale;
  T2 acc = mkc<T2>(0, 0);
  for (long long i = threadIdx.x; i < n; i += blockDim.x) acc = cadd(acc, x[i]);
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = cadd(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  // With conj_x the convolution ran on conj(x), so x0 and the DC sum enter
  // conjugated; conj_out conjugates every output.
  T2 sum = red[0];
  if (conj_x) {
    x0.y = -x0.y;
    sum.y = -sum.y;
  }
  for (long long k = threadIdx.x; k < n - 1; k += blockDim.x) {
    T2 v = cadd(x0, conv[k + off]);
    if (conj_out) v.y = -v.y;
    out[iperm[k]] = cscale(v, sc);
  }
  if (threadIdx.x == 0) {
    if (conj_out) sum.y = -sum.y;
    out[0] = cscale(sum, sc);
  }
}

template <typename T>
__global__ void k_chirp_in(const void* in_, void* out_, const void* w_, long long n, long long m,
                           long long nlines, long long is, long long os, int conj_in,
                           const long long* ioff) {
  typedef typename Vec2<T>::type T2;
  const T2* in = reinterpret_cast<const T2*>(in_);
  const T2* w = reinterpret_cast<const T2*>(w_);
  T2* out = reinterpret_cast<T2*>(out_);
  long long total = m * nlines;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    long long l = i / m, j = i - l * m;
    T2 v = mkc<T2>(0, 0);
    if (j < n) {
      v = in[line_base(ioff, l, is) + j];
      if (conj_in) v.y = -v.y;
      v = cmul(v, w[j]);
    }
    out[l * os + j] = v;
  }
}

template <typename T>
__global__ void k_untangle(const void* Z_, void* out_, const void* utw_, long long m, long long nlines,
                           long long zs, long long os, int mode, double psd_scale, double w_edge,
                           double w_mid) {
  typedef typename Vec2<T>::type T2;
  const T2* Zall = reinterpret_cast<const T2*>(Z_);
  const T2* utw = reinterpret_cast<const T2*>(utw_);
  long long nb = m + 1, total = nb * nlines;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    long long l = i / nb, k = i - l * nb;
    const T2* Z = Zall + l * zs;
    T2 X;
    if (k == 0 || k == m) {
      T2 z0 = Z[0];
      X = mkc<T2>(k == 0 ? z0.x + z0.y : z0.x - z0.y, 0);
    } else {
      T2 zk = Z[k], zmk = Z[m - k];
      T xer = (zk.x + zmk.x) * (T)0.5, xei = (zk.y - zmk.y) * (T)0.5;
      T xor_ = (zk.y + zmk.y) * (T)0.5, xoi = -(zk.x - zmk.x) * (T)0.5;
      T2 w = utw[k];
      X = mkc<T2>(xer + (w.x * xor_ - w.y * xoi), xei + (w.x * xoi + w.y * xor_));
    }
    if (mode == 0) {
      reinterpret_cast<T2*>(out_)[l * os + k] = X;
    } else if (mode == 2) {
      const bool edge = k == 0 || k == m;  // imaginary part exactly 0 at DC and Nyquist
      reinterpret_cast<T2*>(out_)[l * os + k] = edge ? mkc<T2>(X.x * (T)w_edge, 0) : cscale(X, (T)w_mid);
    } else {
      T p = (X.x * X.x + X.y * X.y) * (T)psd_scale;
      if (k != 0 && k != m) p *= (T)2;
      reinterpret_cast<T*>(out_)[l * os + k] = p;
    }
  }
}

template <typename T>
__global__ void k_herm_pack(const void* X_, void* Z_, const void* utw_, long long m, long long nlines,
                            long long xs, long long bins, long long zs, double w_mid) {
  typedef typename Vec2<T>::type T2;
  const T2* Xall = reinterpret_cast<const T2*>(X_);
  const T2* utw = reinterpret_cast<const T2*>(utw_);
  T2* Zall = reinterpret_cast<T2*>(Z_);
  long long total = m * nlines;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    long long l = i / m, k = i - l * m;
    const T2* X = Xall + l * xs;
    T2 z;
    if (k == 0) {
      T x0 = bins > 0 ? X[0].x : (T)0;
      T xm = m < bins ? X[m].x : (T)0;
      z = mkc<T2>((x0 + xm) * (T)0.5, (x0 - xm) * (T)0.5);
    } else {
      T2 xk = k < bins ? cscale(X[k], (T)w_mid) : mkc<T2>(0, 0);
      T2 xmk = (m - k) < bins ? cscale(X[m - k], (T)w_mid) : mkc<T2>(0, 0);
      T xer = (xk.x + xmk.x) * (T)0.5, xei = (xk.y - xmk.y) * (T)0.5;
      T dr = (xk.x - xmk.x) * (T)0.5, di = (xk.y + xmk.y) * (T)0.5;
      T2 w = utw[k];
      T xor_ = w.x * dr + w.y * di, xoi = w.x * di - w.y * dr;
      z = mkc<T2>(xer - xoi, xei + xor_);
    }
    Zall[l * zs + k] = z;
  }
}


template <typename T>
__global__ void k_promote(const T* in, void* out_, long long n, long long nlines, long long is,
                          long long os) {
  typedef typename Vec2<T>::type T2;
  T2* out = reinterpret_cast<T2*>(out_);
  long long total = n * nlines;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    long long l = i / n, k = i - l * n;
    out[l * os + k] = mkc<T2>(in[l * is + k], 0);
  }
}

template <typename T>
__global__ void k_copy_lines(const void* in_, void* out_, long long n, long long nlines, long long is,
                             long long os) {
  typedef typename Vec2<T>::type T2;
  const T2* in = reinterpret_cast<const T2*>(in_);
  T2* out = reinterpret_cast<T2*>(out_);
  long long total = n * nlines;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    long long l = i / n, k = i - l * n;
    out[l * os + k] = in[l * is + k];
  }
}

// full[k] = src[k] (k < bins, k <= n/2), full[n-k] = conj(src[k]) for 0 < k < n-k
// (realplan.go:197-213).
template <typename T>
__global__ void k_herm_full(const void* in_, void* out_, long long n, long long nlines, long long is,
                            long long bins, long long os, double w_mid) {
  typedef typename Vec2<T>::type T2;
  const T2* in = reinterpret_cast<const T2*>(in_);
  T2* out = reinterpret_cast<T2*>(out_);
  long long half = n / 2 + 1;
  long long total = n * nlines;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    long long l = i / n, k = i - l * n;
    T2 v = mkc<T2>(0, 0);
    if (k < half) {
      if (k < bins) v = in[l * is + k];
    } else {
      long long kk = n - k;  // 0 < kk < n - kk
      if (kk < bins) v = cconj(in[l * is + kk]);
    }
    if (k != 0) v = cscale(v, (T)w_mid);
    out[l * os + k] = v;
  }
}

template <typename T>
__global__ void k_weight_bins(const void* in_, void* out_, long long bins, long long n, long long nlines,
                              long long is, long long os, double w_edge, double w_mid) {
  typedef typename Vec2<T>::type T2;
  const T2* in = reinterpret_cast<const T2*>(in_);
  T2* out = reinterpret_cast<T2*>(out_);
  long long total = bins * nlines;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    long long l = i / bins, k = i - l * bins;
    const bool edge = k == 0 || (n % 2 == 0 && k == n / 2);
    const T2 v = in[l * is + k];
    out[l * os + k] = edge ? mkc<T2>(v.x * (T)w_edge, 0) : cscale(v, (T)w_mid);
  }
}

template <typename T>
__global__ void k_take_real(const void* in_, T* out, long long n, long long nlines, long long is,
                            long long os, double scale) {
  typedef typename Vec2<T>::type T2;
  const T2* in = reinterpret_cast<const T2*>(in_);
  long long total = n * nlines;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    long long l = i / n, k = i - l * n;
    out[l * os + k] = in[l * is + k].x * (T)scale;
  }
}

template <typename T>
__global__ void k_power(const void* in_, T* out, long long bins, long long n, long long nlines,
                        long long is, long long os, double scale) {
  typedef typename Vec2<T>::type T2;
  const T2* in = reinterpret_cast<const T2*>(in_);
  long long total = bins * nlines;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    long long l = i / bins, k = i - l * bins;
    T2 X = in[l * is + k];
    T p = (X.x * X.x + X.y * X.y) * (T)scale;
    if (k != 0 && !(n % 2 == 0 && k == bins - 1)) p *= (T)2;
    out[l * os + k] = p;
  }
}

template <typename T>
__global__ void k_frames(const T* x, const T* w, T* out, long long seg, long long step,
                         long long frames) {
  long long total = seg * frames;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    long long t = i / seg, k = i - t * seg;
    out[i] = x[t * step + k] * w[k];
  }
}

template <typename T>
__global__ void k_axis_copy(const void* in_, void* out_, long long n, long long st, long long lines,
                            int dir) {
  typedef typename Vec2<T>::type T2;
  const T2* in = reinterpret_cast<const T2*>(in_);
  T2* out = reinterpret_cast<T2*>(out_);
  long long total = n * lines;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    // Iterate in strided-array order so both sides stay mostly coalesced.
    long long inner = i % st;
    long long t = i / st;
    long long k = t % n;
    long long outer = t / n;
    long long strided = outer * n * st + k * st + inner;
    long long contig = (outer * st + inner) * n + k;
    if (dir == 0) out[contig] = in[strided];
    else out[strided] = in[contig];
  }
}

template <typename T>
__global__ void k_fill(T* out, long long n, double value) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = (T)value;
}

template <typename T>
__global__ void k_scale(const T* in, T* out, long long n, double scale) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = in[i] * (T)scale;
}

// Spectral filter loss helpers (gofft_filter2_loss_grad).
// S[b*half + i] *= H[i] (or conj(H[i])), b < nimg.
template <typename T>
__global__ void k_filter_mul(void* S_, const void* H_, long long half, long long nimg, int conj_h) {
  typedef typename Vec2<T>::type T2;
  T2* S = reinterpret_cast<T2*>(S_);
  const T2* H = reinterpret_cast<const T2*>(H_);
  const long long total = half * nimg;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    T2 h = H[i % half];
    if (conj_h) h.y = -h.y;
    S[i] = cmul(S[i], h);
  }
}

// r = r - y over nimg images of `per` reals; block (x, b) also writes the sum
// of its r^2 to partial[b * gridDim.x + x] and stores 2 r (the loss's
// gradient with respect to r). Block sums are tree reductions in a fixed
// order, so the loss is the same for any number of images per call.
template <typename T>
__global__ void __launch_bounds__(256) k_residual(T* r, const T* y, long long per, T* partial) {
  __shared__ T red[256];
  const long long b = blockIdx.y;
  T acc = 0;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < per; i += (long long)gridDim.x * 256) {
    const T v = r[b * per + i] - y[b * per + i];
    acc += v * v;
    r[b * per + i] = v + v;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[b * gridDim.x + blockIdx.x] = red[0];
}

// loss[b] = sum of partial[b * nparts .. + nparts) in order.
template <typename T>
__global__ void k_sum_parts(const T* partial, long long nparts, long long nimg, T* loss) {
  const long long b = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (b >= nimg) return;
  T acc = 0;
  for (long long k = 0; k < nparts; k++) acc += partial[b * nparts + k];
  loss[b] = acc;
}

// ---------------------------------------------------------------------------
// Launchers.

static inline unsigned grid_for(long long total, int block) {
  long long g = (total + block - 1) / block;
  if (g > 1048576) g = 1048576;
  if (g < 1) g = 1;
  return (unsigned)g;
}

template <typename T> static size_t elt2() { return 2 * sizeof(T); }

// The 16-points-per-thread form of a kernel (the only one except k_conv's
// complex64 24-point form).
#define GOFFT_LAUNCH16(KERNEL, T, ARGS, BYTES, THREADS, GRID, STREAM) \
  hipLaunchKernelGGL((KERNEL<T, 16>), dim3(GRID), dim3(THREADS), BYTES, STREAM, ARGS)

int pass_form(int prec, const PassArgs& a) {
  if (a.real_last || a.real_first) return a.real_last ? 11 : 12;
  const bool ext = a.in_hook || a.out_hook || a.mid_spec;
  int mode = a.mid256 && prec != 0 ? 2 : ext ? 1 : 0;  // form 2: complex64 only
  // The two-stage register form: R = A x 16 planned radix A first, 16
  // threads per column, plain stage twiddles, no spectrum multiply. Complex64
  // Bluestein chains take the HK = 1 forms (they spill in complex128); the
  // Rader hooks run in the register form in complex128 only (measured slower
  // than form 1 in complex64). A plain 16 x 16 pass keeps form 0, which
  // already runs both of its stages from registers.
  const SubFFT& f = a.f;
  const int A = f.st[0].radix;
  if (mode <= 1 && f.nstages == 2 && f.tpc == 16 && f.st[1].radix == 16 && f.R == 16 * A &&
      f.st[0].cmul == 0 && f.st[1].cmul == 0 && !a.mid_spec) {
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

hipError_t launch_pass(int prec, int maxp, const PassArgs& a, int threads, hipStream_t s) {
  (void)maxp;
  const int mode = pass_form(prec, a);
  const long long groups = a.real_last || a.real_first ? a.ru_lines * a.ru_wpl : (a.ncols + a.f.C - 1) / a.f.C;
#define GOFFT_PASS(T, M) hipLaunchKernelGGL((k_pass<T, 16, M>), dim3(groups), dim3(threads), bytes, s, a)
#define GOFFT_PASS_LB(T, M, LB) hipLaunchKernelGGL((k_pass<T, 16, M, LB>), dim3(groups), dim3(threads), bytes, s, a)
  if (prec == 0) {
    // complex128 has forms 0, 1, 3, 4, 11, 12 (the others are complex64 only)
    const size_t bytes = (size_t)a.f.C * a.f.Rp * elt2<double>();
    if (threads <= 256 && (mode == 0 || mode == 3 || mode == 4)) {
      switch (mode) {
        case 3: GOFFT_PASS_LB(double, 3, 256); break;
        case 4: GOFFT_PASS_LB(double, 4, 256); break;
        default: GOFFT_PASS_LB(double, 0, 256); break;
      }
    } else if (threads <= 512 && (mode == 0 || mode == 3 || mode == 4)) {
      switch (mode) {
        case 3: GOFFT_PASS_LB(double, 3, 512); break;
        case 4: GOFFT_PASS_LB(double, 4, 512); break;
        default: GOFFT_PASS_LB(double, 0, 512); break;
      }
    } else {
      switch (mode) {
        case 1: GOFFT_PASS(double, 1); break;
        case 3: GOFFT_PASS(double, 3); break;
        case 4: GOFFT_PASS(double, 4); break;
        case 11: GOFFT_PASS(double, 11); break;
        case 12: GOFFT_PASS(double, 12); break;
        default: GOFFT_PASS(double, 0); break;
      }
    }
  } else if (a.tw_M && a.tw_mode != 0 && (mode == 0 || mode == 11 || mode == 12)) {
    // the complex64 pass-twiddle variants (PassTw), for the plain passes and
    // the real transforms' fused first and last passes
    const size_t bytes = (size_t)a.f.C * a.f.Rp * elt2<float>();
#define GOFFT_PASS_TW(M, TW) hipLaunchKernelGGL((k_pass<float, 16, M, 1024, TW>), dim3(groups), dim3(threads), bytes, s, a)
    if (a.tw_mode == 1) {
      if (mode == 11) GOFFT_PASS_TW(11, 1);
      else if (mode == 12) GOFFT_PASS_TW(12, 1);
      else GOFFT_PASS_TW(0, 1);
    } else {
      if (mode == 11) GOFFT_PASS_TW(11, 2);
      else if (mode == 12) GOFFT_PASS_TW(12, 2);
      else GOFFT_PASS_TW(0, 2);
    }
#undef GOFFT_PASS_TW
  } else {
    const size_t bytes = (size_t)a.f.C * a.f.Rp * elt2<float>();
    switch (mode) {
      case 1: GOFFT_PASS(float, 1); break;
      case 2: GOFFT_PASS(float, 2); break;
      case 3: GOFFT_PASS(float, 3); break;
      case 4: GOFFT_PASS(float, 4); break;
      case 5: GOFFT_PASS(float, 5); break;
      case 6: GOFFT_PASS(float, 6); break;
      case 7: GOFFT_PASS(float, 7); break;
      case 11: GOFFT_PASS(float, 11); break;
      case 12: GOFFT_PASS(float, 12); break;
      default: GOFFT_PASS(float, 0); break;
    }
  }
#undef GOFFT_PASS
#undef GOFFT_PASS_LB
  return hipGetLastError();
}

hipError_t launch_real_fwd(int prec, int maxp, const RealFwdArgs& a, int threads, hipStream_t s) {
  (void)maxp;
  long long groups = (a.nlines + a.f.C - 1) / a.f.C;
  if (prec == 0) {
    size_t bytes = (size_t)a.f.C * a.f.Rp * elt2<double>();
    GOFFT_LAUNCH16(k_real_fwd, double, a, bytes, threads, groups, s);
  } else {
    size_t bytes = (size_t)a.f.C * a.f.Rp * elt2<float>();
    if (a.lds_tab) {
      bytes += (size_t)(a.stw_n + 2 * a.f.R + 1) * elt2<float>();
      hipLaunchKernelGGL((k_real_fwd_tab<float, 8>), dim3(a.wgs < groups ? a.wgs : groups), dim3(threads), bytes, s, a);
    } else {
      GOFFT_LAUNCH16(k_real_fwd, float, a, bytes, threads, groups, s);
    }
  }
  return hipGetLastError();
}

int real_kernel_vgprs(int prec, int dir) {
  hipFuncAttributes fa{}, ia{};
  hipError_t e1, e2;
  if (prec == 0) {
    e1 = hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&k_real_fwd<double, 16>));
    e2 = hipFuncGetAttributes(&ia, reinterpret_cast<const void*>(&k_real_inv<double, 16>));
  } else {
    e1 = hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&k_real_fwd<float, 16>));
    e2 = hipFuncGetAttributes(&ia, reinterpret_cast<const void*>(&k_real_inv<float, 16>));
  }
  if (e1 != hipSuccess || e2 != hipSuccess) return 0;
  if (dir == 0) return fa.numRegs;
  if (dir == 1) return ia.numRegs;
  return fa.numRegs > ia.numRegs ? fa.numRegs : ia.numRegs;
}

int real_tab_vgprs(int inverse) {
  hipFuncAttributes fa{};
  const void* k = inverse ? reinterpret_cast<const void*>(&k_real_inv_tab<float, 8>)
                          : reinterpret_cast<const void*>(&k_real_fwd_tab<float, 8>);
  if (hipFuncGetAttributes(&fa, k) != hipSuccess) return 0;
  return fa.numRegs;
}

hipError_t launch_real_inv(int prec, int maxp, const RealInvArgs& a, int threads, hipStream_t s) {
  (void)maxp;
  long long groups = (a.nlines + a.f.C - 1) / a.f.C;
  if (prec == 0) {
    size_t bytes = (size_t)a.f.C * a.f.Rp * elt2<double>();
    GOFFT_LAUNCH16(k_real_inv, double, a, bytes, threads, groups, s);
  } else {
    size_t bytes = (size_t)a.f.C * a.f.Rp * elt2<float>();
    if (a.lds_tab) {
      bytes += (size_t)(a.stw_n + a.f.R / 2 + 1) * elt2<float>();
      hipLaunchKernelGGL((k_real_inv_tab<float, 8>), dim3(a.wgs < groups ? a.wgs : groups), dim3(threads), bytes, s, a);
    } else {
      GOFFT_LAUNCH16(k_real_inv, float, a, bytes, threads, groups, s);
    }
  }
  return hipGetLastError();
}

hipError_t launch_conv(int prec, int maxp, const ConvArgs& a, int threads, hipStream_t s) {
  long long groups = (a.nlines + a.f.C - 1) / a.f.C;
  if (prec == 0) {
    size_t bytes = (size_t)a.f.C * a.f.Rp * elt2<double>();
    if (threads <= 256)
      hipLaunchKernelGGL((k_conv<double, 16, 256>), dim3(groups), dim3(threads), bytes, s, a);
    else if (threads <= 512)
      hipLaunchKernelGGL((k_conv<double, 16, 512>), dim3(groups), dim3(threads), bytes, s, a);
    else
      GOFFT_LAUNCH16(k_conv, double, a, bytes, threads, groups, s);
  } else {
    size_t bytes = (size_t)a.f.C * a.f.Rp * elt2<float>();
    if (maxp == 24)
      hipLaunchKernelGGL((k_conv<float, 24>), dim3(groups), dim3(threads), bytes, s, a);
    else
      GOFFT_LAUNCH16(k_conv, float, a, bytes, threads, groups, s);
  }
  return hipGetLastError();
}

hipError_t launch_cmul_table(int prec, const void* in, void* out, const void* tab, long long n,
                             long long nlines, long long is, long long os, int conj_in, int conj_out,
                             double scale, const long long* ioff, const long long* ooff,
                             hipStream_t s) {
  unsigned g = grid_for(n * nlines, 256);
  if (prec == 0)
    hipLaunchKernelGGL(k_cmul_table<double>, dim3(g), dim3(256), 0, s, in, out, tab, n, nlines, is, os,
                       conj_in, conj_out, scale, ioff, ooff);
  else
    hipLaunchKernelGGL(k_cmul_table<float>, dim3(g), dim3(256), 0, s, in, out, tab, n, nlines, is, os,
                       conj_in, conj_out, scale, ioff, ooff);
  return hipGetLastError();
}

hipError_t launch_rader_gather(int prec, const void* in, void* out, const int* perm, long long q,
                               long long cl, long long nlines, long long is, long long os,
                               int conj_in, const long long* ioff, hipStream_t s) {
  unsigned g = grid_for(cl * nlines, 256);
  if (prec == 0)
    hipLaunchKernelGGL(k_rader_gather<double>, dim3(g), dim3(256), 0, s, in, out, perm, q, cl, nlines,
                       is, os, conj_in, ioff);
  else
    hipLaunchKernelGGL(k_rader_gather<float>, dim3(g), dim3(256), 0, s, in, out, perm, q, cl, nlines,
                       is, os, conj_in, ioff);
  return hipGetLastError();
}

hipError_t launch_rader_scatter(int prec, const void* x, const void* conv, void* out, const int* iperm,
                                long long n, long long off, long long nlines, long long xs,
                                long long cs, long long os, int conj_x, int conj_out, double scale,
                                const long long* xoff, const long long* ooff, hipStream_t s) {
  if (nlines <= 0) return hipSuccess;
  if (prec == 0)
    hipLaunchKernelGGL(k_rader_scatter<double>, dim3((unsigned)nlines), dim3(256), 0, s, x, conv, out,
                       iperm, n, off, xs, cs, os, conj_x, conj_out, scale, xoff, ooff);
  else
    hipLaunchKernelGGL(k_rader_scatter<float>, dim3((unsigned)nlines), dim3(256), 0, s, x, conv, out,
                       iperm, n, off, xs, cs, os, conj_x, conj_out, scale, xoff, ooff);
  return hipGetLastError();
}

hipError_t launch_chirp_in(int prec, const void* in, void* out, const void* w, long long n,
                           long long m, long long nlines, long long is, long long os, int conj_in,
                           const long long* ioff, hipStream_t s) {
  unsigned g = grid_for(m * nlines, 256);
  if (prec == 0)
    hipLaunchKernelGGL(k_chirp_in<double>, dim3(g), dim3(256), 0, s, in, out, w, n, m, nlines, is, os,
                       conj_in, ioff);
  else
    hipLaunchKernelGGL(k_chirp_in<float>, dim3(g), dim3(256), 0, s, in, out, w, n, m, nlines, is, os,
                       conj_in, ioff);
  return hipGetLastError();
}

hipError_t launch_untangle(int prec, const void* Z, void* out, const void* utw, long long m,
                           long long nlines, long long zs, long long os, int mode,
                           double psd_scale, double w_edge, double w_mid, hipStream_t s) {
  unsigned g = grid_for((m + 1) * nlines, 256);
  if (prec == 0)
    hipLaunchKernelGGL(k_untangle<double>, dim3(g), dim3(256), 0, s, Z, out, utw, m, nlines, zs, os,
                       mode, psd_scale, w_edge, w_mid);
  else
    hipLaunchKernelGGL(k_untangle<float>, dim3(g), dim3(256), 0, s, Z, out, utw, m, nlines, zs, os,
                       mode, psd_scale, w_edge, w_mid);
  return hipGetLastError();
}

hipError_t launch_herm_pack(int prec, const void* X, void* Z, const void* utw, long long m,
                            long long nlines, long long xs, long long bins, long long zs,
                            double w_mid, hipStream_t s) {
  unsigned g = grid_for(m * nlines, 256);
  if (prec == 0)
    hipLaunchKernelGGL(k_herm_pack<double>, dim3(g), dim3(256), 0, s, X, Z, utw, m, nlines, xs, bins,
                       zs, w_mid);
  else
    hipLaunchKernelGGL(k_herm_pack<float>, dim3(g), dim3(256), 0, s, X, Z, utw, m, nlines, xs, bins,
                       zs, w_mid);
  return hipGetLastError();
}


#define GOFFT_LAUNCH_EW(KERNEL, TOTAL, ...)                                                    \
  do {                                                                                       \
    unsigned g_ = grid_for(TOTAL, 256);                                                      \
    if (prec == 0)                                                                           \
      hipLaunchKernelGGL(KERNEL<double>, dim3(g_), dim3(256), 0, s, __VA_ARGS__);            \
    else                                                                                     \
      hipLaunchKernelGGL(KERNEL<float>, dim3(g_), dim3(256), 0, s, __VA_ARGS__);             \
  } while (0)

hipError_t launch_promote(int prec, const void* in, void* out, long long n, long long nlines,
                          long long is, long long os, hipStream_t s) {
  unsigned g = grid_for(n * nlines, 256);
  if (prec == 0)
    hipLaunchKernelGGL(k_promote<double>, dim3(g), dim3(256), 0, s, (const double*)in, out, n, nlines,
                       is, os);
  else
    hipLaunchKernelGGL(k_promote<float>, dim3(g), dim3(256), 0, s, (const float*)in, out, n, nlines,
                       is, os);
  return hipGetLastError();
}

hipError_t launch_copy_lines(int prec, const void* in, void* out, long long n, long long nlines,
                             long long is, long long os, hipStream_t s) {
  GOFFT_LAUNCH_EW(k_copy_lines, n * nlines, in, out, n, nlines, is, os);
  return hipGetLastError();
}

hipError_t launch_herm_full(int prec, const void* in, void* out, long long n, long long nlines,
                            long long is, long long bins, long long os, double w_mid, hipStream_t s) {
  GOFFT_LAUNCH_EW(k_herm_full, n * nlines, in, out, n, nlines, is, bins, os, w_mid);
  return hipGetLastError();
}

hipError_t launch_weight_bins(int prec, const void* in, void* out, long long bins, long long n,
                              long long nlines, long long is, long long os, double w_edge,
                              double w_mid, hipStream_t s) {
  GOFFT_LAUNCH_EW(k_weight_bins, bins * nlines, in, out, bins, n, nlines, is, os, w_edge, w_mid);
  return hipGetLastError();
}

hipError_t launch_take_real(int prec, const void* in, void* out, long long n, long long nlines,
                            long long is, long long os, double scale, hipStream_t s) {
  unsigned g = grid_for(n * nlines, 256);
  if (prec == 0)
    hipLaunchKernelGGL(k_take_real<double>, dim3(g), dim3(256), 0, s, in, (double*)out, n, nlines, is,
                       os, scale);
  else
    hipLaunchKernelGGL(k_take_real<float>, dim3(g), dim3(256), 0, s, in, (float*)out, n, nlines, is,
                       os, scale);
  return hipGetLastError();
}

hipError_t launch_power(int prec, const void* in, void* out, long long bins, long long n,
                        long long nlines, long long is, long long os, double scale, hipStream_t s) {
  unsigned g = grid_for(bins * nlines, 256);
  if (prec == 0)
    hipLaunchKernelGGL(k_power<double>, dim3(g), dim3(256), 0, s, in, (double*)out, bins, n, nlines, is,
                       os, scale);
  else
    hipLaunchKernelGGL(k_power<float>, dim3(g), dim3(256), 0, s, in, (float*)out, bins, n, nlines, is,
                       os, scale);
  return hipGetLastError();
}

hipError_t launch_frames(int prec, const void* x, const void* w, void* out, long long seg,
                         long long step, long long frames, hipStream_t s) {
  unsigned g = grid_for(seg * frames, 256);
  if (prec == 0)
    hipLaunchKernelGGL(k_frames<double>, dim3(g), dim3(256), 0, s, (const double*)x, (const double*)w,
                       (double*)out, seg, step, frames);
  else
    hipLaunchKernelGGL(k_frames<float>, dim3(g), dim3(256), 0, s, (const float*)x, (const float*)w,
                       (float*)out, seg, step, frames);
  return hipGetLastError();
}

hipError_t launch_axis_copy(int prec, const void* in, void* out, long long n, long long st,
                            long long lines, int dir, hipStream_t s) {
  GOFFT_LAUNCH_EW(k_axis_copy, n * lines, in, out, n, st, lines, dir);
  return hipGetLastError();
}

hipError_t launch_filter_mul(int prec, void* S, const void* H, long long half, long long nimg, int conj_h,
                             hipStream_t s) {
  unsigned g = grid_for(half * nimg, 256);
  if (prec == 0)
    hipLaunchKernelGGL(k_filter_mul<double>, dim3(g), dim3(256), 0, s, S, H, half, nimg, conj_h);
  else
    hipLaunchKernelGGL(k_filter_mul<float>, dim3(g), dim3(256), 0, s, S, H, half, nimg, conj_h);
  return hipGetLastError();
}

long long residual_parts(long long per) {
  long long p = (per + 256 * 8 - 1) / (256 * 8);
  return p < 1 ? 1 : p > 1024 ? 1024 : p;
}

hipError_t launch_residual_loss(int prec, void* r, const void* y, long long per, long long nimg, void* partial,
                                void* loss, hipStream_t s) {
  if (nimg <= 0) return hipSuccess;
  const long long parts = residual_parts(per);
  const unsigned lg = (unsigned)((nimg + 255) / 256);
  if (prec == 0) {
    hipLaunchKernelGGL(k_residual<double>, dim3((unsigned)parts, (unsigned)nimg), dim3(256), 0, s, (double*)r,
                       (const double*)y, per, (double*)partial);
    hipLaunchKernelGGL(k_sum_parts<double>, dim3(lg), dim3(256), 0, s, (const double*)partial, parts, nimg,
                       (double*)loss);
  } else {
    hipLaunchKernelGGL(k_residual<float>, dim3((unsigned)parts, (unsigned)nimg), dim3(256), 0, s, (float*)r,
                       (const float*)y, per, (float*)partial);
    hipLaunchKernelGGL(k_sum_parts<float>, dim3(lg), dim3(256), 0, s, (const float*)partial, parts, nimg,
                       (float*)loss);
  }
  return hipGetLastError();
}

hipError_t launch_fill(int prec, void* out, long long n, double value, hipStream_t s) {
  unsigned g = grid_for(n, 256);
  if (prec == 0)
    hipLaunchKernelGGL(k_fill<double>, dim3(g), dim3(256), 0, s, (double*)out, n, value);
  else
    hipLaunchKernelGGL(k_fill<float>, dim3(g), dim3(256), 0, s, (float*)out, n, value);
  return hipGetLastError();
}

hipError_t launch_scale(int prec, const void* in, void* out, long long n, double scale,
                        hipStream_t s) {
  unsigned g = grid_for(n, 256);
  if (prec == 0)
    hipLaunchKernelGGL(k_scale<double>, dim3(g), dim3(256), 0, s, (const double*)in, (double*)out, n,
                       scale);
  else
    hipLaunchKernelGGL(k_scale<float>, dim3(g), dim3(256), 0, s, (const float*)in, (float*)out, n,
                       scale);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Probe: one radix-16 DFT stage on complex64 data held in registers, applied
// `reps` times, as VALU butterflies (method 0) or as a dense DFT-matrix
// product on the matrix cores (method 1). Each application is the unitary
// DFT (scaled by 1/4), so four of them give the input back.
//
// Method 1 writes the 16-point complex DFT as a 32 x 32 real matrix W (rows
// and columns: 16 real parts, then 16 imaginary parts) and applies it to 32
// transforms per wave with sixteen v_mfma_f32_32x32x2f32 (K = 2 each). The
// instruction's result D (32 x 32) sits in 16 registers per lane: lane l,
// register r holds row 8*(r/4) + 4*(l/32) + r%4 of column l%32. Slice q of
// the K dimension is taken to be the rows that register q holds in the two
// lane halves, so lane half h supplies its own register q as the B operand
// and W[l%32][row(q, h)] as the A operand: the result comes out in the
// layout the next product reads, and repeated stages need no exchange.

__global__ void __launch_bounds__(256) k_probe_dft16_valu(const float2* in, float2* out, long long transforms,
                                                          int reps) {
  const long long t = blockIdx.x * 256ll + threadIdx.x;
  if (t >= transforms) return;
  float2 v[16];
  const float2* p = in + t * 16;
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = p[i];
  for (int r = 0; r < reps; ++r) {
    dft<16>(v);
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = cscale(v[i], 0.25f);
  }
  float2* o = out + t * 16;
#pragma unroll
  for (int i = 0; i < 16; ++i) o[i] = v[i];
}

typedef float f32x16 __attribute__((ext_vector_type(16)));

__global__ void __launch_bounds__(256) k_probe_dft16_mfma(const float2* in, float2* out, long long transforms,
                                                          int reps, const float* wmat) {
  const int lane = threadIdx.x & 63;
  const long long t0 = ((blockIdx.x * 256ll + threadIdx.x) >> 6) * 32;
  if (t0 >= transforms) return;  // whole waves only
  const int j = lane & 31, h = lane >> 5;
  float a[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) a[q] = wmat[lane * 16 + q];
  const long long t = t0 + j;
  const bool live = t < transforms;
  f32x16 d;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int pt = 8 * (q >> 2) + 4 * h + (q & 3);
    const float2 x = live ? in[t * 16 + pt] : make_float2(0.f, 0.f);
    d[q] = x.x;
    d[q + 8] = x.y;
  }
  for (int r = 0; r < reps; ++r) {
    f32x16 acc = {};
#pragma unroll
    for (int q = 0; q < 16; ++q) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[q], d[q], acc, 0, 0, 0);
    d = acc;
  }
  if (!live) return;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int pt = 8 * (q >> 2) + 4 * h + (q & 3);
    out[t * 16 + pt] = make_float2(d[q], d[q + 8]);
  }
}

hipError_t launch_probe_dft16(int method, const void* in, void* out, long long transforms, int reps,
                              const float* wmat, hipStream_t s) {
  if (method == 0) {
    const unsigned g = (unsigned)((transforms + 255) / 256);
    hipLaunchKernelGGL(k_probe_dft16_valu, dim3(g), dim3(256), 0, s, (const float2*)in, (float2*)out,
                       transforms, reps);
  } else {
    const unsigned g = (unsigned)((transforms + 127) / 128);  // 4 waves of 32 transforms
    hipLaunchKernelGGL(k_probe_dft16_mfma, dim3(g), dim3(256), 0, s, (const float2*)in, (float2*)out,
                       transforms, reps, wmat);
  }
  return hipGetLastError();
}

#else  // GOFFT_QUEUE_TU

// ---------------------------------------------------------------------------
// Persistent work queue (mixed-length plans). Every workgroup loops: thread 0
// takes the next ticket from the device counter, the workgroup finds the
// stage holding it and runs that stage's pass or convolution body on the
// column (line) group ticket - start. Workgroups are NT threads with one
// dynamic LDS size; the host re-plans each stage's columns (lines) per
// workgroup so that they fill NT threads.
//
// Ordering between stages: a workgroup that moves on from a stage whose
// items others wait for (signal) first makes its stores visible at agent
// scope -- every wave waits for its stores to complete, the workgroup
// meets at a barrier, and thread 0 releases (agent scope) and adds the
// number of that stage's items it finished to done[stage]. A workgroup
// entering a stage with a dependency waits, once, until done[dep] reaches
// the dep's item count (thread 0 polls with acquire loads, within a
// wall-clock budget), then meets at a barrier. Tickets
// are taken in stage order and dependencies point to earlier stages, so the
// lowest unfinished ticket never waits on unfinished work: the queue always
// drains, and a wait that outlives the budget means a planning error, which
// sets *error and stops that workgroup (and, through the flag, the others).
// complex64: at most 128 registers, so that 16 waves share a CU at any
// workgroup size.
template <typename T, int NT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 1)))
k_queue(QueueArgs q) {
  typedef typename Vec2<T>::type T2;
  // The pass stages' per-column tables live in dynamic LDS past the largest
  // pass stage's columns, so that a fused convolution can have all of it.
  PassShared<T2>& sh = *reinterpret_cast<PassShared<T2>*>(g_dyn_lds + q.sh_off);
  __shared__ long long s_ticket;
  __shared__ int s_fail;
  const int tid = threadIdx.x;
  int cur = -1;       // stage of the previous item
  unsigned pend = 0;  // items of a signalling stage finished since its last count
  for (;;) {
    if (tid == 0)
      s_ticket = (long long)__hip_atomic_fetch_add(q.ticket, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    // (tickets stay below 2^31: the host plans at most 2^30 items)
    const long long t = __builtin_amdgcn_readfirstlane((int)s_ticket);
    // (every read of s_ticket ends before thread 0 can write it again)
    __syncthreads();
    int si = cur < 0 ? 0 : cur;
    while (si < q.nstages && t >= q.st[si].start + q.st[si].nitems) ++si;
    if (si != cur) {
      if (pend) {
        // every wave waits until its stores have reached the L2, then one
        // agent-scope release (an L2 write-back) covers the workgroup
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (tid == 0) {
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
          __hip_atomic_fetch_add(q.done + cur, pend, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        pend = 0;
      }
      if (si >= q.nstages) break;
      cur = si;
      const int d = q.st[si].dep;
      if (d >= 0) {
        if (tid == 0) {
          const unsigned need = (unsigned)(q.st[d].nitems + q.need_extra);
          const unsigned long long t0 = wall_clock64();
          int fail = 0;
          while (__hip_atomic_load(q.done + d, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < need) {
            if (__hip_atomic_load(q.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
              fail = 1;
              break;
            }
            if ((long long)(wall_clock64() - t0) > q.spin_limit) {
              __hip_atomic_store(q.error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              fail = 1;
              break;
            }
            __builtin_amdgcn_s_sleep(2);
          }
          s_fail = fail;
        }
        // (thread 0's acquire invalidated the CU's and the XCD's caches,
        // which the whole workgroup reads through)
        __syncthreads();
        if (s_fail) break;
      }
    }
    const QueueStage& st = q.st[si];
    const long long blk = t - st.start;
    // The descriptors are read through the constant address space, which
    // lets the bodies load their fields with scalar loads as they load a
    // kernel's own arguments (nothing writes them during the kernel).
    typedef __attribute__((address_space(4))) const char cchar;
    cchar* dp = (cchar*)q.descs + st.desc;
    if (st.kind == 0) {
      const PassArgs& a = *(const PassArgs*)dp;
      switch (st.form) {
        case 1: pass_body<T, 16, 1, 0>(a, blk, st.nitems, sh); break;
        case 3: pass_body<T, 16, 3, 0>(a, blk, st.nitems, sh); break;
        case 4: pass_body<T, 16, 4, 0>(a, blk, st.nitems, sh); break;
        default:
          if constexpr (sizeof(T) == 4) {
            if (st.form == 2) { pass_body<T, 16, 2, 0>(a, blk, st.nitems, sh); break; }
            if (st.form == 5) { pass_body<T, 16, 5, 0>(a, blk, st.nitems, sh); break; }
            if (st.form == 6) { pass_body<T, 16, 6, 0>(a, blk, st.nitems, sh); break; }
            if (st.form == 7) { pass_body<T, 16, 7, 0>(a, blk, st.nitems, sh); break; }
          }
          pass_body<T, 16, 0, 0>(a, blk, st.nitems, sh);
          break;
      }
    } else {
      const ConvArgs& a = *(const ConvArgs*)dp;
      if constexpr (sizeof(T) == 4) {
        if (st.form == 24) conv_body<T, 24>(a, blk);
        else conv_body<T, 16>(a, blk);
      } else {
        conv_body<T, 16>(a, blk);
      }
    }
    if (st.signal) ++pend;
  }
}

// Copies the plan's descriptor template into the exec's workspace, one
// workgroup per descriptor, and sets its pointer fields for this exec;
// workgroup 0 also zeroes the queue's counters.
__global__ void __launch_bounds__(256) k_queue_setup(QueueSetupArgs a) {
  const QueueDesc& d = a.d[blockIdx.x];
  const long long* src = reinterpret_cast<const long long*>(a.tmpl + d.off);
  long long* dst = reinterpret_cast<long long*>(a.descs + d.off);
  for (long long i = threadIdx.x; i < d.bytes / 8; i += blockDim.x) dst[i] = src[i];
  __syncthreads();
  if ((int)threadIdx.x < d.nreloc) {
    const int k = threadIdx.x;
    *reinterpret_cast<const char**>(a.descs + d.off + d.field[k]) = a.base[d.sel[k]] + d.add[k];
  }
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < a.nctr; i += blockDim.x) a.ctr[i] = 0u;
}

unsigned queue_pass_forms(int prec) { return prec == 0 ? 0x1bu : 0xffu; }  // complex128: 0, 1, 3, 4

bool queue_conv_form(int prec, int maxp) { return maxp == 16 || (prec != 0 && maxp == 24); }

size_t queue_pass_tables(int prec) { return prec == 0 ? sizeof(PassShared<double2>) : sizeof(PassShared<float2>); }

size_t queue_static_lds() { return 64; }

// Queue kernel instantiations: complex64 with 256, 512 or 1024 threads,
// complex128 with 256 or 512.
template <typename F>
static bool queue_kernel(int prec, int threads, F&& f) {
  if (prec == 0) {
    if (threads == 256) return f(reinterpret_cast<const void*>(&k_queue<double, 256>));
    if (threads == 512) return f(reinterpret_cast<const void*>(&k_queue<double, 512>));
  } else {
    if (threads == 256) return f(reinterpret_cast<const void*>(&k_queue<float, 256>));
    if (threads == 512) return f(reinterpret_cast<const void*>(&k_queue<float, 512>));
    if (threads == 1024) return f(reinterpret_cast<const void*>(&k_queue<float, 1024>));
  }
  return false;
}

int queue_blocks_per_cu(int prec, int threads, size_t lds) {
  int nb = 0;
  queue_kernel(prec, threads, [&](const void* k) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, threads, lds) != hipSuccess) nb = 0;
    return true;
  });
  (void)hipGetLastError();
  return nb;
}

hipError_t launch_queue_setup(const QueueSetupArgs& a, int ndesc, hipStream_t s) {
  hipLaunchKernelGGL(k_queue_setup, dim3(ndesc), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_queue(int prec, const QueueArgs& a, int threads, size_t lds, long long grid, hipStream_t s) {
  if (prec == 0) {
    if (threads == 256) hipLaunchKernelGGL((k_queue<double, 256>), dim3(grid), dim3(256), lds, s, a);
    else if (threads == 512) hipLaunchKernelGGL((k_queue<double, 512>), dim3(grid), dim3(512), lds, s, a);
    else return hipErrorInvalidValue;
  } else {
    if (threads == 256) hipLaunchKernelGGL((k_queue<float, 256>), dim3(grid), dim3(256), lds, s, a);
    else if (threads == 512) hipLaunchKernelGGL((k_queue<float, 512>), dim3(grid), dim3(512), lds, s, a);
    else if (threads == 1024) hipLaunchKernelGGL((k_queue<float, 1024>), dim3(grid), dim3(1024), lds, s, a);
    else return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

#endif  // GOFFT_QUEUE_TU

}  // namespace gofft
