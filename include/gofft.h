/* gofft.h — C ABI of libgofft_hip.so, the gfx950 (MI355X) backend for the
 * go-fft transform path.
 *
 * The reference (go-fft, pure Go) has no FFI; its drop-in boundary is its public
 * Go API. Each entry point below names the reference function(s) it replaces
 * (file:line in the reference tree). A Go package binds these over cgo; the
 * Python test harness binds them over ctypes.
 *
 * Conventions
 *  - Complex data is interleaved (re, im): complex128 = 2 x double,
 *    complex64 = 2 x float. Real data is double or float.
 *  - `gofft_exec_*` take DEVICE pointers and a hipStream_t (as void*, may be
 *    NULL for the default stream); they enqueue work and return without
 *    synchronising. `gofft_*_host` take HOST pointers, stage through device
 *    memory and return after the result is in host memory.
 *  - Forward transforms are unnormalised; inverses divide by n (by the product
 *    of the axis lengths for N-D), as the reference does (doc.go:21-26).
 *  - For complex-to-complex transforms (c2c, mixed-length and N-D C2C plans)
 *    and their VJPs, dst may alias src (plan.go:88-90). The real transforms
 *    (r2c, c2r, PSD, Spectrogram, RFFT2/IRFFT2) and their VJPs return
 *    GOFFT_ERR_INVALID_ARG when the input and output ranges overlap: the
 *    reference's real calls cannot alias either, their slices having
 *    different element types (realplan.go:77, 181). src is never modified
 *    unless it aliases dst.
 *  - Every call returns GOFFT_OK (0) or a negative status; gofft_last_error()
 *    returns a thread-local message for the last failure on this thread.
 *  - Transform and VJP calls on a plan with data reject a NULL data pointer,
 *    and a direction other than GOFFT_FORWARD / GOFFT_INVERSE, with
 *    GOFFT_ERR_INVALID_ARG before any work is enqueued (an inverse given
 *    zero bins may pass NULL for its input).
 *  - Plans are immutable after creation and safe to execute concurrently from
 *    several threads provided each concurrent device call uses its own
 *    stream; the plan keeps one workspace per stream (plan.go:12-14,
 *    realplan.go:10-11). Host-staged calls need no stream: each concurrent
 *    call gets its own stream, device buffers and pinned staging from the
 *    plan, made on first use and reused, and waits for that stream only.
 */
#ifndef GOFFT_H
#define GOFFT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define GOFFT_OK 0
#define GOFFT_ERR_INVALID_ARG (-1)
#define GOFFT_ERR_SHAPE (-2)
#define GOFFT_ERR_OOM (-3)
#define GOFFT_ERR_HIP (-4)
#define GOFFT_ERR_UNSUPPORTED (-5)

/* element types */
#define GOFFT_C128 0 /* complex128 / float64: the plain drop-in */
#define GOFFT_C64 1  /* complex64 / float32: the reduced-precision batched form */

/* directions */
#define GOFFT_FORWARD 0
#define GOFFT_INVERSE 1

/* plan flags */
#define GOFFT_FLAG_NONE 0u
/* Bluestein lengths use the reference's recurrence twiddles for the
 * power-of-two convolution (internal/kernels/kernels.go:38-59) so the result
 * reproduces the reference's rounding; set by default for GOFFT_C128. */
#define GOFFT_FLAG_REFERENCE_TWIDDLES 1u
/* Force accurate (directly evaluated) twiddles everywhere. */
#define GOFFT_FLAG_ACCURATE_TWIDDLES 2u
/* Time the planner's alternatives on the current device at plan creation and
 * keep the fastest: convolution lengths of Rader and Bluestein (the role of
 * the reference's bestConvLen, plan.go:192-259) and columns per workgroup of
 * multi-pass transforms. Results are cached per (route, length, lines,
 * precision), and in the file named by GOFFT_TUNE_CACHE if set. The env var
 * GOFFT_AUTOTUNE=1 sets this flag for every plan. No reference counterpart. */
#define GOFFT_FLAG_AUTOTUNE 4u
/* complex64 only: how the pass twiddles of multi-pass transforms are formed
 * (no reference counterpart; the reference is FP64, plan.go:319-327). By
 * default they come from FP64 two-level tables built from exact integer
 * phases, multiplied and stepped in FP64 and rounded once to FP32.
 * TWIDDLES_FP32 uses FP32 tables and FP32 products; TWIDDLES_SPLIT uses
 * tables of hi + lo float pairs and forms each twiddle in double-float
 * arithmetic. They apply to the plain passes and the real transforms' fused
 * first and last passes; convolution chains keep the default. */
#define GOFFT_FLAG_TWIDDLES_FP32 8u
#define GOFFT_FLAG_TWIDDLES_SPLIT 16u
/* Mixed-length plans (gofft_plan_c2c_many): run the length groups as the
 * stages of one persistent kernel over a device work queue instead of
 * launches on several streams. Each group's launches become stages, re-planned
 * to one workgroup size; workgroups take column or line groups in stage order
 * from a device counter, and a stage that needs an earlier one's results waits
 * (bounded in time) for all of its items. Groups the queue kernel cannot run
 * (twiddle-form flags, workgroups beyond its LDS share) stay on the streams
 * beside it. Same results as without the flag, bit for bit. The env var
 * GOFFT_QUEUE=1 sets it for every plan. No reference counterpart (the
 * reference runs transforms one by one). */
#define GOFFT_FLAG_QUEUE 32u

/* N-D kinds */
#define GOFFT_ND_C2C 0
#define GOFFT_ND_R2C 1 /* RFFT2/IRFFT2 semantics on the last axis */

typedef struct gofft_plan gofft_plan;

/* Library set-up and diagnostics. gofft_init selects the current HIP device
 * and reads its limits; it is called implicitly by every plan constructor. */
int gofft_init(void);
const char* gofft_last_error(void);
const char* gofft_version(void);
/* Engine a length routes to (plan.go:42-73): 0 trivial, 1 power of two,
 * 2 mixed-radix, 3 Rader, 4 Bluestein. */
int gofft_route(int64_t n);

/* --- plan construction ------------------------------------------------- */

/* NewPlan (plan.go:42-73) for `batch` independent contiguous lines of length n.
 * Replaces Plan.FFT / Plan.IFFT (plan.go:91-109) and FFT/IFFT (fft.go:14-40). */
int gofft_plan_c2c(gofft_plan** plan, int64_t n, int64_t batch, int dtype, unsigned flags);

/* Many transforms of uneven lengths in one plan: line i has length lengths[i]
 * and starts at element offset sum(lengths[0..i-1]) of both src and dst. No
 * reference counterpart (the reference loops over Plan.FFT per length). */
int gofft_plan_c2c_many(gofft_plan** plan, const int64_t* lengths, int64_t count, int dtype,
                        unsigned flags);

/* NewRealPlan (realplan.go:37-58) for `batch` lines of n real samples; line i
 * of the spectrum holds n/2+1 bins. Replaces RealPlan.RFFT / RealPlan.IRFFT
 * (realplan.go:77-192) and RFFT/IRFFT (rfft.go:19-45). */
int gofft_plan_r2c(gofft_plan** plan, int64_t n, int64_t batch, int dtype, unsigned flags);

/* FFTN/IFFTN/FFT2/IFFT2 (ndfft.go:22-45) for kind GOFFT_ND_C2C, and
 * RFFT2/IRFFT2 (rfft2.go:22-138) for kind GOFFT_ND_R2C (rank 2 only), over
 * `batch` contiguous arrays of the given row-major shape. rank 0 is the
 * identity. A non-positive axis returns GOFFT_ERR_SHAPE with the reference's
 * panic text "fft: shape lengths must be positive" (ndfft.go:54). */
int gofft_plan_nd(gofft_plan** plan, const int64_t* shape, int rank, int kind, int64_t batch,
                  int dtype, unsigned flags);

/* PSD (freq.go:69-90) of `batch` real lines of length n with sample spacing d;
 * each output line holds n/2+1 densities. */
int gofft_plan_psd(gofft_plan** plan, int64_t n, int64_t batch, double d, int dtype,
                   unsigned flags);

/* Spectrogram (freq.go:107-127) of a real signal of `length` samples: frames of
 * `segment` samples every `segment - overlap`, multiplied by `window` (host
 * array of `segment` doubles, e.g. Hann, window.go:20-22), one-sided PSD each.
 * Errors carry the reference's panic texts (freq.go:109,112,115). */
int gofft_plan_spectrogram(gofft_plan** plan, int64_t length, int64_t segment, int64_t overlap,
                           const double* window, int64_t window_len, double d, int dtype,
                           unsigned flags);

int gofft_plan_destroy(gofft_plan* plan);

/* Plan queries: transform length, frames (spectrogram), element counts of one
 * input and one output array, and the workspace the plan holds per stream. */
int64_t gofft_plan_len(const gofft_plan* plan);
int64_t gofft_plan_frames(const gofft_plan* plan);
int64_t gofft_plan_in_elems(const gofft_plan* plan, int direction);
int64_t gofft_plan_out_elems(const gofft_plan* plan, int direction);
int64_t gofft_plan_workspace_bytes(const gofft_plan* plan);
/* Number of kernel launches one forward execution enqueues. */
int gofft_plan_num_launches(const gofft_plan* plan, int direction);

/* Text description of the plan's engines and launches (pass lengths, columns
   and threads per workgroup, LDS stage radices), one line per engine, written
   NUL-terminated into buf (truncated to len). Returns the length needed,
   terminator included. No reference counterpart (diagnostics). */
int64_t gofft_plan_describe(const gofft_plan* plan, char* buf, int64_t len);

/* Kernel families counted by gofft_plan_kernel_work. */
enum {
  GOFFT_KERNEL_PASS = 0, /* Stockham pass kernel (one pass of a transform or convolution) */
  GOFFT_KERNEL_CONV = 1, /* fused single-kernel Rader / Bluestein transform */
  GOFFT_KERNEL_REAL = 2, /* fused single-kernel real-input transform */
  GOFFT_KERNEL_AUX = 3,  /* elementwise kernels (copies, packing, scaling) */
  GOFFT_KERNEL_COUNT = 4
};

/* Launches of family `kernel` in one exec of the plan in `direction`, and the
   data bytes those launches read and write: each launch's input lines and
   output lines once, tables excluded. For a measurement's roofline (bytes over
   the launches' measured time). No reference counterpart. */
int gofft_plan_kernel_work(const gofft_plan* p, int direction, int kernel, int64_t* launches,
                           int64_t* bytes);

/* Integer index tables of the plan's transform of length n, for checking
   them bit for bit (no reference counterpart; diagnostics). Copies up to
   `cap` entries into out (may be NULL) and returns the table's length, or a
   negative status. The permutations are read back from the device copies
   the kernels index with; the chirp indices are the ones the plan's chirp
   tables were built from. */
enum {
  GOFFT_TABLE_RADER_GENERATOR = 0, /* {g}: smallest primitive root mod n (rader.go:178-197) */
  GOFFT_TABLE_RADER_PERM = 1,      /* perm[k] = g^k mod n, k < n-1 (rader.go:78-82) */
  GOFFT_TABLE_RADER_IPERM = 2,     /* iperm[k] = perm[(n-1-k) mod (n-1)] (rader.go:83-86) */
  GOFFT_TABLE_RADER_CONV = 3,      /* {convolution length, 1 if cyclic} */
  GOFFT_TABLE_CHIRP_INDEX = 4,     /* j^2 mod 2n, j < n (bluestein.go:35-41) */
  GOFFT_TABLE_BLUESTEIN_CONV = 5   /* {convolution length m} */
};
int64_t gofft_plan_index_table(const gofft_plan* plan, int64_t n, int table, int64_t* out, int64_t cap);

/* Measurement probe (no reference counterpart; no plan): applies the unitary
   16-point forward DFT (scaled by 1/4) `reps` times to each of `transforms`
   complex64 lines of 16 points (device memory, out may alias in) in
   registers, as VALU butterflies (method 0) or as a dense 32 x 32 real
   DFT-matrix product on the matrix cores (method 1). It exists to time one
   radix-16 stage both ways; no transform uses method 1. */
int gofft_probe_dft16(int method, const void* in, void* out, int64_t transforms, int reps, void* stream);

/* --- execution on device memory -------------------------------------------- */

/* c2c / c2c_many / nd-C2C plans. */
int gofft_exec_c2c(const gofft_plan* plan, const void* src, void* dst, int direction,
                   void* stream);
/* r2c plans: forward real -> n/2+1 bins per line. */
int gofft_exec_r2c(const gofft_plan* plan, const void* src, void* dst, void* stream);
/* r2c plans: inverse; each input line holds `bins` bins (missing bins read as
 * zero, the imaginary parts of DC and Nyquist are ignored; realplan.go:221-250). */
int gofft_exec_c2r(const gofft_plan* plan, const void* src, int64_t bins, void* dst, void* stream);
/* nd-R2C plans: RFFT2 (direction 0) / IRFFT2 (direction 1); for the inverse
 * `bins` is the number of complex bins supplied per array (rfft2.go:99-106). */
int gofft_exec_nd_r2c(const gofft_plan* plan, const void* src, int64_t bins, void* dst,
                      int direction, void* stream);
/* psd and spectrogram plans: real in, real densities out. */
int gofft_exec_psd(const gofft_plan* plan, const void* src, void* dst, void* stream);
int gofft_exec_spectrogram(const gofft_plan* plan, const void* src, void* dst, void* stream);

/* gofft_exec_c2c without any allocation: the plan's workspace comes from a
 * static device array of GOFFT_NOALLOC_SCRATCH_BYTES per device (held until
 * the process ends) instead of the per-stream pool, which allocates on a
 * stream's first use. No counterpart in the reference; for callers that
 * cannot allocate (code under graph capture, tight serving loops). Plans
 * whose workspace (times the side lanes of a mixed-length plan) exceeds the
 * scratch return GOFFT_ERR_UNSUPPORTED. Every call on one device shares that
 * device's scratch: calls that overlap on one device race. Call it with the
 * plan's device current. */
#define GOFFT_NOALLOC_SCRATCH_BYTES (64ll << 20)
int gofft_exec_c2c_noalloc(const gofft_plan* plan, const void* src, void* dst, int direction, void* stream);

/* --- vector-Jacobian products (no reference counterpart) ------------------ */

/* For y = FFT(x) (complex, unnormalised) and an upstream gradient gy = dL/dconj(y)
 * in the convention where dL = Re<gy, dy>, the gradient w.r.t. x is
 * gx = n * IFFT(gy) = conj(FFT(conj(gy))). direction selects which transform's
 * VJP: GOFFT_FORWARD for y = FFT(x), GOFFT_INVERSE for y = IFFT(x). */
int gofft_vjp_c2c(const gofft_plan* plan, const void* gy, void* gx, int direction, void* stream);
/* For y = RFFT(x) (n/2+1 bins per line, real x): gx = real part of the adjoint,
 * i.e. n * IRFFT(w * gy) with w = 1 at DC and Nyquist and 1/2 elsewhere. */
int gofft_vjp_r2c(const gofft_plan* plan, const void* gy, void* gx, void* stream);
/* For x = IRFFT(X) (real output): gX = RFFT(gx) / n with interior bins doubled. */
int gofft_vjp_c2r(const gofft_plan* plan, const void* gx, void* gX, void* stream);
/* nd-R2C plans: VJP of RFFT2 (direction 0) and IRFFT2 (direction 1). */
int gofft_vjp_nd_r2c(const gofft_plan* plan, const void* g_out, void* g_in, int direction,
                     void* stream);

/* Loss and gradient of a spectral filter, for each array b of an nd-R2C
 * plan (no reference counterpart):
 *   loss[b] = || IRFFT2(RFFT2(x_b) * H) - y_b ||^2,   gx_b = d loss[b] / d x_b,
 * with H (rows x (cols/2+1) complex, shared by every array) multiplying the
 * half spectrum. The gradient runs the adjoint transforms (as
 * gofft_vjp_nd_r2c) on 2 (IRFFT2(RFFT2(x) H) - y) with conj(H) between them.
 * The arrays are processed `chunk` at a time: the spectrum and residual of
 * only that many arrays are held, formed from x again for each chunk rather
 * than stored for the whole batch; chunk <= 0 takes the largest chunk that
 * fits 80% of the free device memory (the whole batch when it fits). The
 * results do not depend on the chunk. Device pointers: x, y, gx of
 * batch*rows*cols reals, loss of batch reals. */
int gofft_filter2_loss_grad(const gofft_plan* plan, const void* x, const void* H, const void* y, void* loss,
                            void* gx, int64_t chunk, void* stream);

/* --- host convenience (reference slice API) -------------------------------- */

/* The same operations on HOST buffers (ordinary pageable memory, e.g. Go
 * slices): the input is copied through the plan's pinned staging buffers to
 * device memory, the transform runs on a stream of the call's own, and the
 * output comes back the same way; the call returns when the output is in
 * dst and waits for nothing else. dst may alias src here for every
 * transform (the device copies are separate). Element counts as for the
 * device calls. */
/* Plan.FFT / Plan.IFFT (plan.go:91-109), FFT/IFFT (fft.go:14-40), and
 * FFTN/IFFTN/FFT2/IFFT2 (ndfft.go:22-45) with an nd-C2C plan. */
int gofft_exec_c2c_host(const gofft_plan* plan, const void* src, void* dst, int direction);
/* RealPlan.RFFT (realplan.go:77-116) / RFFT (rfft.go:19-26). */
int gofft_exec_r2c_host(const gofft_plan* plan, const void* src, void* dst);
/* RealPlan.IRFFT (realplan.go:181-192) / IRFFT (rfft.go:39-45). */
int gofft_exec_c2r_host(const gofft_plan* plan, const void* src, int64_t bins, void* dst);
/* RFFT2 / IRFFT2 (rfft2.go:22-138). */
int gofft_exec_nd_r2c_host(const gofft_plan* plan, const void* src, int64_t bins, void* dst, int direction);
/* PSD (freq.go:69-90). */
int gofft_exec_psd_host(const gofft_plan* plan, const void* src, void* dst);
/* Spectrogram (freq.go:107-127): frames x (segment/2+1) densities, row-major. */
int gofft_exec_spectrogram_host(const gofft_plan* plan, const void* src, void* dst);
/* Host-staged VJPs (no reference counterpart): as gofft_vjp_c2c, _r2c, _c2r
   and _nd_r2c, with host buffers. */
int gofft_vjp_c2c_host(const gofft_plan* plan, const void* gy, void* gx, int direction);
int gofft_vjp_r2c_host(const gofft_plan* plan, const void* gy, void* gx);
int gofft_vjp_c2r_host(const gofft_plan* plan, const void* gx, void* gX);
int gofft_vjp_nd_r2c_host(const gofft_plan* plan, const void* g_out, void* g_in, int direction);

/* --- graphs (no reference counterpart) ------------------------------------ */

/* Operations a graph can capture (the exec and VJP calls above). */
#define GOFFT_OP_C2C 0         /* gofft_exec_c2c */
#define GOFFT_OP_R2C 1         /* gofft_exec_r2c */
#define GOFFT_OP_C2R 2         /* gofft_exec_c2r */
#define GOFFT_OP_ND_R2C 3      /* gofft_exec_nd_r2c */
#define GOFFT_OP_PSD 4         /* gofft_exec_psd */
#define GOFFT_OP_SPECTROGRAM 5 /* gofft_exec_spectrogram */
#define GOFFT_OP_VJP_C2C 6     /* gofft_vjp_c2c */
#define GOFFT_OP_VJP_R2C 7     /* gofft_vjp_r2c */
#define GOFFT_OP_VJP_C2R 8     /* gofft_vjp_c2r */
#define GOFFT_OP_VJP_ND_R2C 9  /* gofft_vjp_nd_r2c */

/* One call of operation `op` on the plan with fixed device buffers (and
   `bins` / `direction` where the op takes them), captured on `stream` (not
   the null stream) as a hipGraph: a launch replays every kernel of that call
   (and a mixed-length plan's stream fork and join) at once. Each graph owns
   its workspace, allocated here, so replays of one graph may overlap execs of
   the plan and replays of other graphs, but not each other. The plan must
   outlive the graph. */
typedef struct gofft_graph gofft_graph;
int gofft_graph_capture(gofft_graph** graph, const gofft_plan* plan, int op, const void* src, int64_t bins,
                        void* dst, int direction, void* stream);
/* gofft_graph_capture of GOFFT_OP_C2C. */
int gofft_graph_c2c(gofft_graph** graph, const gofft_plan* plan, const void* src, void* dst, int direction,
                    void* stream);
int gofft_graph_launch(const gofft_graph* graph, void* stream);
int gofft_graph_destroy(gofft_graph* graph);

/* Synchronise the device (for callers without their own stream handling). */
int gofft_sync(void);

#ifdef __cplusplus
}
#endif

#endif /* GOFFT_H */
