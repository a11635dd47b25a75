package gofft

/*
#include "gofft.h"
*/
import "C"

import "runtime"

// checkShape applies the reference's shape validation and panic texts
// (ndfft.go:50-62) and returns the product of the axes.
func checkShape(shape []int, total int) int {
	prod := 1
	for _, s := range shape {
		if s <= 0 {
			panic("fft: shape lengths must be positive")
		}
		prod *= s
	}
	if prod != total {
		panic("fft: shape product does not match len(data)")
	}
	return prod
}

func cshape(shape []int) []C.int64_t {
	cs := make([]C.int64_t, len(shape))
	for i, s := range shape {
		cs[i] = C.int64_t(s)
	}
	return cs
}

func ndPlan(shape []int, kind C.int) *handle {
	cs := cshape(shape)
	return newHandle(func(pp **C.gofft_plan) C.int {
		var sp *C.int64_t
		if len(cs) > 0 {
			sp = &cs[0]
		}
		return C.gofft_plan_nd(pp, sp, C.int(len(cs)), kind, 1, C.GOFFT_C128, 0)
	})
}

func transformN(data []complex128, shape []int, dir C.int) []complex128 {
	checkShape(shape, len(data))
	out := make([]complex128, len(data))
	if len(data) <= 1 {
		copy(out, data)
		return out
	}
	h := ndPlan(shape, C.GOFFT_ND_C2C)
	call(func() C.int { return C.gofft_exec_c2c_host(h.p, cptr(data), cptr(out), dir) })
	runtime.KeepAlive(h)
	return out
}

// FFTN returns the N-dimensional DFT of a row-major array (ndfft.go:22-24).
// An empty shape is the identity (a copy); a non-positive axis or a product
// other than len(data) panics with the reference's texts.
func FFTN(data []complex128, shape []int) []complex128 {
	return transformN(data, shape, C.GOFFT_FORWARD)
}

// IFFTN is the inverse of FFTN, normalised by the product of the axes
// (ndfft.go:28-33).
func IFFTN(data []complex128, shape []int) []complex128 {
	return transformN(data, shape, C.GOFFT_INVERSE)
}

// FFT2 is FFTN over two axes (ndfft.go:35-39).
func FFT2(data []complex128, shape [2]int) []complex128 { return FFTN(data, shape[:]) }

// IFFT2 is IFFTN over two axes (ndfft.go:41-45).
func IFFT2(data []complex128, shape [2]int) []complex128 { return IFFTN(data, shape[:]) }

// RFFT2 returns the rows x (cols/2+1) half spectrum of a real rows x cols
// array (rfft2.go:14-74).
func RFFT2(data []float64, shape [2]int) []complex128 {
	rows, cols := shape[0], shape[1]
	if rows <= 0 || cols <= 0 {
		panic("fft: shape lengths must be positive")
	}
	if rows*cols != len(data) {
		panic("fft: shape product does not match len(data)")
	}
	out := make([]complex128, rows*(cols/2+1))
	h := ndPlan(shape[:], C.GOFFT_ND_R2C)
	call(func() C.int { return C.gofft_exec_nd_r2c_host(h.p, rptr(data), 0, cptr(out), C.GOFFT_FORWARD) })
	runtime.KeepAlive(h)
	return out
}

// IRFFT2 inverts RFFT2 to a real rows x cols array normalised by 1/(rows
// cols) (rfft2.go:76-138); a short spectrum is zero padded, extra bins are
// ignored.
func IRFFT2(data []complex128, shape [2]int) []float64 {
	rows, cols := shape[0], shape[1]
	if rows <= 0 || cols <= 0 {
		panic("fft: shape lengths must be positive")
	}
	out := make([]float64, rows*cols)
	bins := len(data)
	if half := rows * (cols/2 + 1); bins > half {
		bins = half
	}
	h := ndPlan(shape[:], C.GOFFT_ND_R2C)
	call(func() C.int {
		return C.gofft_exec_nd_r2c_host(h.p, cptr(data[:bins]), C.int64_t(bins), rptr(out), C.GOFFT_INVERSE)
	})
	runtime.KeepAlive(h)
	return out
}
