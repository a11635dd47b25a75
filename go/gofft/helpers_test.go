package gofft

import (
	"math"
	"math/cmplx"
	"testing"
)

// The reference tests' tolerance (fft_test.go: 1e-9 absolute).
const tol = 1e-9

// dft is the O(n^2) transform in float64 with exact integer phases j*k mod n.
func dft(x []complex128, inverse bool) []complex128 {
	n := len(x)
	out := make([]complex128, n)
	sign := -1.0
	if inverse {
		sign = 1
	}
	for k := 0; k < n; k++ {
		var acc complex128
		for j := 0; j < n; j++ {
			ph := float64((j * k) % n)
			acc += x[j] * cmplx.Rect(1, sign*2*math.Pi*ph/float64(n))
		}
		if inverse {
			acc /= complex(float64(n), 0)
		}
		out[k] = acc
	}
	return out
}

// dftN applies dft along every axis of a row-major array.
func dftN(x []complex128, shape []int) []complex128 {
	out := append([]complex128(nil), x...)
	stride := 1
	for ax := len(shape) - 1; ax >= 0; ax-- {
		n := shape[ax]
		outer := len(out) / (n * stride)
		line := make([]complex128, n)
		for o := 0; o < outer; o++ {
			for in := 0; in < stride; in++ {
				base := o*n*stride + in
				for i := 0; i < n; i++ {
					line[i] = out[base+i*stride]
				}
				y := dft(line, false)
				for i := 0; i < n; i++ {
					out[base+i*stride] = y[i]
				}
			}
		}
		stride *= n
	}
	return out
}

func near(t *testing.T, what string, got, want []complex128) {
	t.Helper()
	if len(got) != len(want) {
		t.Fatalf("%s: %d values, want %d", what, len(got), len(want))
	}
	for i := range got {
		if d := cmplx.Abs(got[i] - want[i]); d > tol {
			t.Fatalf("%s[%d] = %v, want %v (off by %g)", what, i, got[i], want[i], d)
		}
	}
}

func nearReal(t *testing.T, what string, got, want []float64) {
	t.Helper()
	if len(got) != len(want) {
		t.Fatalf("%s: %d values, want %d", what, len(got), len(want))
	}
	for i := range got {
		if d := math.Abs(got[i] - want[i]); d > tol {
			t.Fatalf("%s[%d] = %.12g, want %.12g (off by %g)", what, i, got[i], want[i], d)
		}
	}
}

// mustPanic runs f and fails unless it panics with the message want.
func mustPanic(t *testing.T, what, want string, f func()) {
	t.Helper()
	defer func() {
		r := recover()
		if r == nil {
			t.Fatalf("%s: no panic", what)
		}
		if s, ok := r.(string); !ok || s != want {
			t.Fatalf("%s: panic %v, want %q", what, r, want)
		}
	}()
	f()
}

// Deterministic test signals (the shapes of the reference tests' inputs).
func complexSignal(n int) []complex128 {
	x := make([]complex128, n)
	for i := range x {
		x[i] = complex(math.Sin(0.7*float64(i))+float64(i%3), math.Cos(0.3*float64(i))-float64(i%5))
	}
	return x
}

func realSignal(n int) []float64 {
	x := make([]float64, n)
	for i := range x {
		x[i] = math.Sin(0.7*float64(i)) + 0.3*math.Cos(1.9*float64(i)) + float64(i%4)
	}
	return x
}

func promote(x []float64) []complex128 {
	c := make([]complex128, len(x))
	for i, v := range x {
		c[i] = complex(v, 0)
	}
	return c
}
