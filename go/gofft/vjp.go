package gofft

/*
#include "gofft.h"
*/
import "C"

import "runtime"

// Vector-Jacobian products (no reference counterpart). For a real loss L the
// gradient with respect to a complex array z is dL/dRe z + i dL/dIm z; each
// function maps the gradient of a transform's output to that of its input
// through the library's adjoint kernels.

// FFTVJP is the gradient through y = FFT(x): n * IFFT(gy).
func FFTVJP(gy []complex128) []complex128 { return vjpC2C(gy, C.GOFFT_FORWARD) }

// IFFTVJP is the gradient through y = IFFT(x): FFT(gy) / n.
func IFFTVJP(gy []complex128) []complex128 { return vjpC2C(gy, C.GOFFT_INVERSE) }

func vjpC2C(gy []complex128, dir C.int) []complex128 {
	out := make([]complex128, len(gy))
	if len(gy) <= 1 {
		copy(out, gy)
		return out
	}
	p := cachedPlan(len(gy))
	call(func() C.int { return C.gofft_vjp_c2c_host(p.h.p, cptr(gy), cptr(out), dir) })
	runtime.KeepAlive(p)
	return out
}

// FFTNVJP and IFFTNVJP are the gradients through FFTN and IFFTN.
func FFTNVJP(gy []complex128, shape []int) []complex128  { return vjpN(gy, shape, C.GOFFT_FORWARD) }
func IFFTNVJP(gy []complex128, shape []int) []complex128 { return vjpN(gy, shape, C.GOFFT_INVERSE) }

func vjpN(gy []complex128, shape []int, dir C.int) []complex128 {
	checkShape(shape, len(gy))
	out := make([]complex128, len(gy))
	if len(gy) <= 1 {
		copy(out, gy)
		return out
	}
	h := ndPlan(shape, C.GOFFT_ND_C2C)
	call(func() C.int { return C.gofft_vjp_c2c_host(h.p, cptr(gy), cptr(out), dir) })
	runtime.KeepAlive(h)
	return out
}

// RFFTVJP is the gradient through y = RFFT(x) for a real x of length n
// (len(gy) = n/2+1): n * IRFFT(w * gy), w = 1 at DC and Nyquist and 1/2 at
// the bins the inverse counts twice.
func RFFTVJP(gy []complex128, n int) []float64 {
	if n <= 0 {
		return []float64{}
	}
	if len(gy) != n/2+1 {
		panic("fft: the gradient of RFFT has n/2+1 bins")
	}
	out := make([]float64, n)
	if n == 1 {
		out[0] = real(gy[0])
		return out
	}
	p := cachedRealPlan(n)
	call(func() C.int { return C.gofft_vjp_r2c_host(p.h.p, cptr(gy), rptr(out)) })
	runtime.KeepAlive(p)
	return out
}

// IRFFTVJP is the gradient through x = IRFFT(X, len(gx)): RFFT(gx) / n with
// the bins other than DC and Nyquist doubled (their imaginary parts exact
// zeros at DC and Nyquist).
func IRFFTVJP(gx []float64) []complex128 {
	n := len(gx)
	if n == 0 {
		return []complex128{}
	}
	out := make([]complex128, n/2+1)
	if n == 1 {
		out[0] = complex(gx[0], 0)
		return out
	}
	p := cachedRealPlan(n)
	call(func() C.int { return C.gofft_vjp_c2r_host(p.h.p, rptr(gx), cptr(out)) })
	runtime.KeepAlive(p)
	return out
}

// RFFT2VJP is the gradient through y = RFFT2(x, shape) (gy holds rows x
// (cols/2+1) bins); IRFFT2VJP the gradient through x = IRFFT2(X, shape).
func RFFT2VJP(gy []complex128, shape [2]int) []float64 {
	rows, cols := shape[0], shape[1]
	if rows <= 0 || cols <= 0 {
		panic("fft: shape lengths must be positive")
	}
	if len(gy) != rows*(cols/2+1) {
		panic("fft: shape product does not match len(data)")
	}
	out := make([]float64, rows*cols)
	h := ndPlan(shape[:], C.GOFFT_ND_R2C)
	call(func() C.int { return C.gofft_vjp_nd_r2c_host(h.p, cptr(gy), rptr(out), C.GOFFT_FORWARD) })
	runtime.KeepAlive(h)
	return out
}

func IRFFT2VJP(gx []float64, shape [2]int) []complex128 {
	rows, cols := shape[0], shape[1]
	if rows <= 0 || cols <= 0 {
		panic("fft: shape lengths must be positive")
	}
	if len(gx) != rows*cols {
		panic("fft: shape product does not match len(data)")
	}
	out := make([]complex128, rows*(cols/2+1))
	h := ndPlan(shape[:], C.GOFFT_ND_R2C)
	call(func() C.int { return C.gofft_vjp_nd_r2c_host(h.p, rptr(gx), cptr(out), C.GOFFT_INVERSE) })
	runtime.KeepAlive(h)
	return out
}
