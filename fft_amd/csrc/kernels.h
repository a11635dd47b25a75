// Shared host/device declarations for the gfx950 transform kernels.
//
// Every transform is built from one kernel shape: a workgroup stages C
// "columns" of R points each into LDS, runs an R-point Stockham autosort FFT
// over them with a runtime radix schedule (2/3/4/5/7/8/11/13/16), and writes the
// columns back. How columns map to HBM addresses is a 4-digit mixed-radix
// decomposition of the flat column index, which expresses batched single-pass
// transforms, the passes of a multi-pass (global Stockham) transform, and the
// strided axis passes of FFTN/RFFT2 with the same code.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace gofft {

constexpr int kMaxStages = 16;   // LDS stages per sub-FFT (R <= 2^16 needs <= 16)
constexpr int kMaxCols = 64;     // columns per workgroup
// Static LDS of k_pass: seven 8-byte per-column tables and one complex
// (at most 16 bytes) per column.
constexpr int kPassStaticLds = kMaxCols * (7 * 8 + 16);

// Division by a runtime constant d for n < 2^20, d < 2^20 (see fastdiv_make).
struct FastDiv {
  uint32_t d;
  uint32_t pad;
  uint64_t m;
};

inline FastDiv fastdiv_make(uint32_t d) {
  FastDiv f;
  f.d = d;
  f.pad = 0;
  f.m = ((1ull << 40) + d - 1) / d;
  return f;
}

__host__ __device__ inline uint32_t fastdiv_q(uint32_t n, const FastDiv& f) {
  return (uint32_t)(((uint64_t)n * f.m) >> 40);
}

// One LDS stage of an R-point sub-FFT: radix-`radix` Stockham DIT over
// sub-transforms of size `ns` (product of the earlier radices).
struct Stage {
  int radix;
  int ns;
  int nb;      // R / radix butterflies per column
  int tw_off;  // offset of this stage's twiddles in the stage table
  int kmul;    // twiddles of butterfly k start at tw_off + k*kmul + cdig*cmul (cdig: the
  int cmul;    // column's twiddle digit); generic stages kmul = radix-1, cmul = 0;
               // grouped radix-2 stages kmul = the pass's Ns, cmul = 1
  FastDiv nsd;
};

// Column addressing: flat column g -> digits d0 = g % D0, d1 = (g / D0) % D1,
// d2 = (g / (D0*D1)) % D2, d3 = rest. Element r of column g sits at
// sum(d_i * a_i) + r * g_stride.
// PassArgs::nt bits
constexpr int kNtLoad = 1;
constexpr int kNtStore = 2;

struct ColMap {
  long long D0, D1, D2;
  long long a0, a1, a2, a3;
  long long g;
  int rowmajor;  // 1: walk each column contiguously (g == 1); 0: walk across columns
  int pad;
  // Division magic for D0..D2: q = umul64hi(n << 8, M) = floor(n / D) exactly
  // for n, D < 2^28 (M = ceil(2^56 / D)); M = 0 selects plain division.
  unsigned long long M0, M1, M2;
};

inline void colmap_magic(ColMap& m, long long max_cols) {
  auto mk = [&](long long d) -> unsigned long long {
    if (d <= 0 || d >= (1ll << 28) || max_cols >= (1ll << 28)) return 0;
    return ((1ull << 56) + (unsigned long long)d - 1) / (unsigned long long)d;
  };
  m.M0 = mk(m.D0);
  m.M1 = mk(m.D1);
  m.M2 = mk(m.D2);
}

// Sub-FFT description shared by all kernels.
struct SubFFT {
  int R;         // points per column
  int Rp;        // LDS column pitch R + (R >> psh) + 1
  int psh;       // LDS pad shift: point r of a column sits at r + (r >> psh)
  int C;         // columns per workgroup (power of two)
  int log2C;
  int tpc;       // threads per column (power of two)
  int log2tpc;
  int nstages;
  FastDiv Rd;
  Stage st[kMaxStages];
};

struct PassArgs {
  const void* in;
  void* out;
  const void* stw;          // stage twiddles, precision of the data
  const double2* ptw_hi;    // pass twiddle W_M^e = hi[e >> s] * lo[e & (2^s - 1)]
  const double2* ptw_lo;
  // tw_mode 1 (FP32) and 2 (hi + lo float) of complex64 plain and real
  // first/last passes: the same two-level tables as float2 / float4 entries
  int tw_mode;
  const void* ptwx_hi;
  const void* ptwx_lo;
  long long ncols;
  long long tw_M;           // 0: no pass twiddle
  int tw_shift;
  int tw_digit;             // which column digit is the twiddle's k
  int conj_in, conj_out;
  double scale;             // applied on store when != 1
  const long long* in_off;  // optional: digit d3 -> element offset (mixed-length batches)
  const long long* out_off;
  ColMap im, om;
  SubFFT f;
  // --- convolution hooks (Rader / Bluestein as three fused passes) -------
  // in_hook 1: element i of a line loads x[i] * w[i] (i < nvalid, else 0)
  // in_hook 2: element i loads x[perm[i]]          (i < nvalid, else 0); x[0]
  //            goes to x0buf[line]
  // in_hook 3: element i loads x[i]                 (i < nvalid, else 0)
  // out_hook 1: output p = conj(v) * w[p] for p < nvalid (others dropped)
  // out_hook 2: output p -> x0 + conj(v) at iperm[p - conv_off] when
  //             0 <= p - conv_off < nvalid, plus X[0] = x0 + dc[line] at p = 0
  int in_hook, out_hook;
  long long nvalid, conv_off;
  const void* hook_w;
  const int* hook_perm;
  // Rader: per-line x0 = x[0] of the caller's line, saved by the first launch
  // (in_hook 2: the column holding line point 0 stores it) and read by the
  // last (out_hook 2), so that an output that aliases the input is safe.
  void* x0buf;
  int x0_conj;                     // out_hook 2: x0 enters conjugated (inverse)
  void* dc;                        // Rader: per-line A[0], written by the mid pass
  // mid pass: after the first sub-FFT, multiply by mid_spec at the forward
  // output position (mm map), conjugate, run the second sub-FFT f2.
  const void* mid_spec;
  ColMap mm;
  SubFFT f2;
  const void* stw2;         // stage twiddles of f2
  int mid256;               // f and f2 are both 256 = 16 x 16 with 16 threads per column,
                            // no hooks, output walked along columns (kernels.hip mid256)
  // Real forward untangle fused into the last pass of the half-length
  // transform (kernels.hip untangle_store): workgroup w of a line takes
  // columns j = w*C/2 .. and their mirrors Ns - j, whose outputs hold the
  // pairs (k, m - k).
  int real_last;            // 1: this launch is that form
  // Real inverse pack (realplan.go:221-272) fused into the first pass of the
  // half-length transform (kernels.hip pack_step): same column pairing, the
  // bins read through load hook 3 (zero past nvalid = bins), Z conjugated
  // for the inverse when ru_conj.
  int real_first;           // 1: this launch is that form
  int xcd_group;            // > 1: consecutive column groups share cache lines; keep them on one XCD
  // kNtLoad: the input points go through non-temporal loads; kNtStore: the
  // output points through non-temporal stores. Set only where the launch's
  // lines cover whole 128-byte runs that no other workgroup touches (plan
  // workspaces, aligned caller lines): there the streamed data no longer
  // displaces the tables in L2. Where neighbouring workgroups share cache
  // lines it costs a second fetch or a partial-line write.
  int nt;
  int ru_conj;
  long long ru_m;           // m = n/2 complex points per line
  long long ru_lines;       // lines of this launch
  long long ru_wpl;         // workgroups per line
  const void* ru_utw;       // W_n^k, k = 0..m
  // or, when ru_uhi is set, W_n^k = ru_uhi[k >> ru_ushift] * ru_ulo[k mod
  // 2^ru_ushift], formed in FP64 (two tables of about sqrt(m) entries)
  const double2* ru_uhi;
  const double2* ru_ulo;
  int ru_ushift;
  long long ru_ostride;     // output elements between lines
  int ru_mode;              // 0 bins, 1 one-sided PSD, 2 bins weighted (DC/Nyquist, others)
  double ru_psd, ru_w_edge, ru_w_mid;  // (real_first: w_mid scales bins 1..m-1)
};

// Real-input forward (RFFT / PSD / spectrogram), single pass: each column is
// one real line of 2R samples seen as R complex pairs.
struct RealFwdArgs {
  const void* in;           // real samples (T)
  void* out;                // complex bins (m+1 per line) or real PSD bins
  const void* stw;
  const void* utw;          // untangle twiddles W_n^k, k = 0..R (precision T)
  const void* window;       // 2R window samples or null
  long long nlines;
  long long in_stride;      // real samples between consecutive lines
  long long out_stride;     // output elements between consecutive lines
  int mode;                 // 0 complex bins, 1 PSD, 2 weighted complex bins
  int odd_n;                // unused for even n (reserved)
  double psd_scale;         // d / n
  double w_edge, w_mid;     // mode 2: DC/Nyquist bins times w_edge, others times w_mid
  int vec_in;               // sample pairs are 2T-aligned (even stride, aligned base)
  int nt;                   // kNtLoad: the lines are read through non-temporal loads
  // Table form (complex64): the stage twiddles (stw_n entries), the R + 1
  // untangle twiddles and the window are copied to LDS once per workgroup,
  // and `wgs` workgroups loop over the line groups, loading the next
  // group's samples while the current one is transformed.
  int lds_tab;
  int stw_n;
  long long wgs;
  FastDiv nbd;              // division by R + 1 (bins per line)
  SubFFT f;
};

// Real-output inverse (IRFFT), single pass.
struct RealInvArgs {
  const void* in;           // complex bins
  void* out;                // real samples
  const void* stw;
  const void* utw;
  long long nlines;
  long long in_stride;      // complex elements between consecutive lines
  long long in_bins;        // bins present per line (missing ones read as 0)
  long long out_stride;     // real samples between consecutive lines
  double scale;             // 1/m
  double w_mid;             // weight of the bins other than DC and Nyquist (1: IRFFT)
  int nt;                   // kNtStore: the lines are written through non-temporal stores
  // Table form (complex64), as RealFwdArgs: stage twiddles (stw_n entries)
  // and the R/2 + 1 pack twiddles in LDS, `wgs` looping workgroups.
  int lds_tab;
  int stw_n;
  long long wgs;
  FastDiv halfd;            // division by R/2 + 1 (bin pairs per line)
  SubFFT f;
};

// Fused single-workgroup cyclic convolution (Rader / Bluestein when the
// convolution length fits in LDS): gather/chirp -> FFT -> *B -> IFFT -> scatter.
struct ConvArgs {
  const void* in;
  void* out;
  const void* stw;
  const void* spec;         // kernel spectrum B (length R, 1/R folded in)
  const void* chirp;        // Bluestein chirp (length n) or null
  const int* perm;          // Rader input gather perm[k] (length q) or null
  const int* dlog;          // Rader input order: dlog[i - 1] = k with perm[k] = i, i = 1..q
  const int* oinv;          // Rader output order: oinv[p - 1] = k with iperm[k] = p, p = 1..q
  const long long* in_off;  // optional per-line element offsets (else l * stride)
  const long long* out_off;
  long long nlines;
  long long in_stride, out_stride;
  int n;                    // transform length
  int kind;                 // 0 Rader cyclic, 1 Rader linear (read conv[k+q]), 2 Bluestein
  int conj_in, conj_out;    // conjugate the input / the output (inverse = both)
  double scale;             // post scale (1/n for IFFT)
  SubFFT f;
};

// Persistent work queue of a mixed-length plan: one kernel whose workgroups
// pop items -- the column groups of a pass, or the line groups of a fused
// convolution -- of several launches ("stages") in order from a device
// counter. A stage whose `dep` is set waits until every item of that earlier
// stage has finished; items are popped in stage order and every dependency
// points back, so the lowest unfinished item never waits and the queue
// drains. Waits are bounded in wall-clock time and end in an error flag.
constexpr int kMaxQueueStages = 48;
struct QueueStage {
  long long start;   // first item of the stage
  long long nitems;  // its items
  long long desc;    // byte offset of its PassArgs / ConvArgs in the descriptor block
  int kind;          // 0 pass (k_pass form `form`), 1 fused convolution (points per thread `form`)
  int form;
  int dep;           // stage that must finish first, or -1
  int signal;        // 1: a later stage waits on this one (its finished items are counted)
};
struct QueueArgs {
  const char* descs;         // descriptor block with the exec's pointers resolved
  unsigned long long* ticket;
  unsigned* error;           // set to 1 when a wait exceeds spin_limit
  unsigned* done;            // finished items per stage
  long long nitems;
  long long spin_limit;      // ticks of the 100 MHz wall clock a wait may take
  long long sh_off;          // byte offset in dynamic LDS of the pass stages' per-column tables
  int nstages;
  int need_extra;            // added to every wait's item count (0; tests of the time-out path)
  QueueStage st[kMaxQueueStages];
};
// One descriptor of the block: copied from the plan's template and its
// pointer fields (up to four) set to base[sel] + add (sel 1 source, 2
// destination, 3 workspace) for the exec at hand.
struct QueueDesc {
  long long off, bytes;
  int nreloc;
  int field[4];              // byte offsets of the pointer fields in the descriptor
  int sel[4];
  long long add[4];
};
struct QueueSetupArgs {
  const char* tmpl;
  char* descs;
  const QueueDesc* d;
  const char* base[4];
  unsigned* ctr;             // counters to zero (ticket, error flag, stage counts)
  int nctr;
};
// Forms of k_pass that the queue kernel of precision prec runs (bit f set
// for form f), and the points per thread of its fused convolutions.
unsigned queue_pass_forms(int prec);
bool queue_conv_form(int prec, int maxp);
// The k_pass form launch_pass picks for a, twiddle variants aside.
int pass_form(int prec, const PassArgs& a);
// Workgroups of the queue kernel (prec, threads) that one CU holds with
// `lds` bytes of dynamic LDS each (0 when it does not fit).
int queue_blocks_per_cu(int prec, int threads, size_t lds);
// Bytes of the per-column tables that a pass stage of the queue kernel keeps
// in dynamic LDS after its columns (at QueueArgs::sh_off), and of the queue
// kernel's static LDS.
size_t queue_pass_tables(int prec);
size_t queue_static_lds();
hipError_t launch_queue_setup(const QueueSetupArgs& a, int ndesc, hipStream_t s);
hipError_t launch_queue(int prec, const QueueArgs& a, int threads, size_t lds, long long grid,
                        hipStream_t s);

// Launchers (defined in kernels.hip). Each returns a hipError_t.
hipError_t launch_pass(int prec, int maxp, const PassArgs& a, int threads, hipStream_t s);
hipError_t launch_real_fwd(int prec, int maxp, const RealFwdArgs& a, int threads, hipStream_t s);
hipError_t launch_real_inv(int prec, int maxp, const RealInvArgs& a, int threads, hipStream_t s);
// Vector registers per thread of the single-pass real kernels at precision
// prec, from the code object: dir 0 k_real_fwd, 1 k_real_inv, -1 the larger.
int real_kernel_vgprs(int prec, int dir = -1);
// Registers per thread of the complex64 table forms of the real forward
// (inverse = 0) and inverse (inverse = 1) kernels.
int real_tab_vgprs(int inverse);
hipError_t launch_conv(int prec, int maxp, const ConvArgs& a, int threads, hipStream_t s);

// Element-wise helpers for the multi-pass convolution and real paths.
// out[l*os + k] = in[l*is + k] * tab[k] (complex), k < n; conj on in/out optional.
// Optional per-line offset tables (ioff/ooff) replace l*is / l*os.
hipError_t launch_cmul_table(int prec, const void* in, void* out, const void* tab, long long n,
                             long long nlines, long long is, long long os, int conj_in, int conj_out,
                             double scale, const long long* ioff, const long long* ooff,
                             hipStream_t s);
// Rader gather: out[l*os + k] = k < q ? in[l*is + perm[k]] : 0, k < cl; conj optional.
hipError_t launch_rader_gather(int prec, const void* in, void* out, const int* perm, long long q,
                               long long cl, long long nlines, long long is, long long os,
                               int conj_in, const long long* ioff, hipStream_t s);
// Rader scatter: out[l*os + iperm[k]] = x0 + conv[l*cs + k + off], out[l*os] = sum(x).
hipError_t launch_rader_scatter(int prec, const void* x, const void* conv, void* out, const int* iperm,
                                long long n, long long off, long long nlines, long long xs,
                                long long cs, long long os, int conj_x, int conj_out, double scale,
                                const long long* xoff, const long long* ooff, hipStream_t s);
// Bluestein chirp in: out[l*os + j] = j < n ? in[l*is + j] * w[j] : 0, j < m.
hipError_t launch_chirp_in(int prec, const void* in, void* out, const void* w, long long n,
                           long long m, long long nlines, long long is, long long os, int conj_in,
                           const long long* ioff, hipStream_t s);
// RFFT untangle of a packed m-point spectrum Z into m+1 bins (mode 1: PSD,
// mode 2: bins weighted w_edge at DC and Nyquist, w_mid elsewhere).
hipError_t launch_untangle(int prec, const void* Z, void* out, const void* utw, long long m,
                           long long nlines, long long zs, long long os, int mode,
                           double psd_scale, double w_edge, double w_mid, hipStream_t s);
// IRFFT pre-pass: Hermitian half spectrum -> packed m-point spectrum Z, the
// bins other than DC and Nyquist weighted by w_mid (1 for the IRFFT).
hipError_t launch_herm_pack(int prec, const void* X, void* Z, const void* utw, long long m,
                            long long nlines, long long xs, long long bins, long long zs,
                            double w_mid, hipStream_t s);
// out[l*os + k] = in[l*is + k] * w_mid, k < bins, except DC and (even n) k = n/2:
// (re(in) * w_edge, 0).
hipError_t launch_weight_bins(int prec, const void* in, void* out, long long bins, long long n,
                              long long nlines, long long is, long long os, double w_edge,
                              double w_mid, hipStream_t s);
// Real lines -> complex lines (imaginary 0): out[l*os + i] = (in[l*is + i], 0), i < n.
hipError_t launch_promote(int prec, const void* in, void* out, long long n, long long nlines,
                          long long is, long long os, hipStream_t s);
// Complex line copy: out[l*os + k] = in[l*is + k], k < n.
hipError_t launch_copy_lines(int prec, const void* in, void* out, long long n, long long nlines,
                             long long is, long long os, hipStream_t s);
// Hermitian mirror of `bins` supplied bins into a full length-n spectrum,
// the bins other than DC weighted by w_mid (1 for the IRFFT; n is odd).
hipError_t launch_herm_full(int prec, const void* in, void* out, long long n, long long nlines,
                            long long is, long long bins, long long os, double w_mid, hipStream_t s);
// Real part of complex lines, optionally scaled: out[l*os + i] = re(in[l*is + i]) * scale.
hipError_t launch_take_real(int prec, const void* in, void* out, long long n, long long nlines,
                            long long is, long long os, double scale, hipStream_t s);
// One-sided PSD of bins: out = |X|^2 * scale, doubled except DC (and k = n/2 when n even).
hipError_t launch_power(int prec, const void* in, void* out, long long bins, long long n,
                        long long nlines, long long is, long long os, double scale, hipStream_t s);
// Spectrogram frames: out[t*seg + i] = x[t*step + i] * w[i].
hipError_t launch_frames(int prec, const void* x, const void* w, void* out, long long seg,
                         long long step, long long frames, hipStream_t s);
// Strided axis lines <-> contiguous lines: line l = (outer, inner), element i at
// outer*n*st + inner + i*st. dir 0 gathers to contiguous, 1 scatters back.
hipError_t launch_axis_copy(int prec, const void* in, void* out, long long n, long long st,
                            long long lines, int dir, hipStream_t s);
// Spectral filter loss: S[b*half + i] *= H[i] (conj(H[i]) when conj_h), b < nimg.
hipError_t launch_filter_mul(int prec, void* S, const void* H, long long half, long long nimg, int conj_h,
                             hipStream_t s);
// Partial sums per image that launch_residual_loss writes (its workspace
// holds nimg * residual_parts(per) reals).
long long residual_parts(long long per);
// r = r - y (nimg images of `per` reals), loss[b] = sum of (r - y)^2 over
// image b, then r = 2 (r - y); partial is workspace (see residual_parts).
hipError_t launch_residual_loss(int prec, void* r, const void* y, long long per, long long nimg, void* partial,
                                void* loss, hipStream_t s);
// Fill n real elements of T with a value (graph-capturable, unlike a memset).
hipError_t launch_fill(int prec, void* out, long long n, double value, hipStream_t s);
// Radix-16 stage probe (method 0 VALU butterflies, 1 matrix cores), kernels.hip.
hipError_t launch_probe_dft16(int method, const void* in, void* out, long long transforms, int reps,
                              const float* wmat, hipStream_t s);
// Scale a real or complex array in place or out of place (n elements of T).
hipError_t launch_scale(int prec, const void* in, void* out, long long n, double scale,
                        hipStream_t s);

}  // namespace gofft
