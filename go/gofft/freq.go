package gofft

/*
#include "gofft.h"
*/
import "C"

import (
	"runtime"
	"unsafe"
)

// FFTFreq returns the sample frequencies of an n-point DFT with sample
// spacing d (freq.go:10-29): k / (n d) for k = 0..ceil(n/2)-1, then the
// negative frequencies. n <= 0 gives an empty non-nil slice.
func FFTFreq(n int, d float64) []float64 {
	if n <= 0 {
		return []float64{}
	}
	out := make([]float64, n)
	scale := 1.0 / (d * float64(n))
	pos := (n-1)/2 + 1
	for i := 0; i < pos; i++ {
		out[i] = float64(i) * scale
	}
	for i := pos; i < n; i++ {
		out[i] = float64(i-n) * scale
	}
	return out
}

// RFFTFreq returns the n/2+1 non-negative frequencies of RFFT (freq.go:38-56).
func RFFTFreq(n int, d float64) []float64 {
	if n <= 0 {
		return []float64{}
	}
	out := make([]float64, n/2+1)
	scale := 1.0 / (d * float64(n))
	for i := range out {
		out[i] = float64(i) * scale
	}
	return out
}

// PSD returns the one-sided power spectral density of x with sample spacing
// d (freq.go:58-90): |X_k|^2 d / n, doubled except at DC and (even n) the
// Nyquist bin. Empty input gives an empty non-nil slice.
func PSD(x []float64, d float64) []float64 {
	n := len(x)
	if n == 0 {
		return []float64{}
	}
	out := make([]float64, n/2+1)
	if n == 1 {
		out[0] = x[0] * x[0] * d
		return out
	}
	h := newHandle(func(pp **C.gofft_plan) C.int {
		return C.gofft_plan_psd(pp, C.int64_t(n), 1, C.double(d), C.GOFFT_C128, 0)
	})
	call(func() C.int { return C.gofft_exec_psd_host(h.p, rptr(x), rptr(out)) })
	runtime.KeepAlive(h)
	return out
}

// Spectrogram returns the PSD of every window-weighted frame of `segment`
// samples taken every segment-overlap samples (freq.go:92-127), or nil when
// x holds no whole frame. It panics with the reference's texts when segment
// <= 0, len(window) != segment, or overlap is outside [0, segment).
func Spectrogram(x []float64, segment, overlap int, window []float64, d float64) [][]float64 {
	if segment <= 0 {
		panic("fft: segment length must be positive")
	}
	if len(window) != segment {
		panic("fft: window length must equal segment length")
	}
	if overlap < 0 || overlap >= segment {
		panic("fft: overlap must satisfy 0 <= overlap < segment")
	}
	step := segment - overlap
	if len(x) < segment {
		return nil
	}
	frames := (len(x)-segment)/step + 1
	bins := segment/2 + 1
	flat := make([]float64, frames*bins)
	if segment == 1 {
		for t := 0; t < frames; t++ {
			v := x[t*step] * window[0]
			flat[t] = v * v * d
		}
	} else {
		h := newHandle(func(pp **C.gofft_plan) C.int {
			return C.gofft_plan_spectrogram(pp, C.int64_t(len(x)), C.int64_t(segment), C.int64_t(overlap),
				(*C.double)(unsafe.Pointer(&window[0])), C.int64_t(len(window)), C.double(d), C.GOFFT_C128, 0)
		})
		call(func() C.int { return C.gofft_exec_spectrogram_host(h.p, rptr(x), rptr(flat)) })
		runtime.KeepAlive(h)
	}
	out := make([][]float64, frames)
	for t := range out {
		out[t] = flat[t*bins : (t+1)*bins : (t+1)*bins]
	}
	return out
}
