package gofft

import (
	"math"
	"math/cmplx"
	"testing"
)

// The semantics of the reference's fft_test.go and plan_test.go, run through
// the GPU backend.

// every engine: powers of two, 13-smooth composites, Rader primes, Bluestein
var lengths = []int{2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 15, 16, 17, 19, 23, 25, 27, 31, 32, 34, 45, 49,
	64, 100, 101, 121, 128, 169, 289, 641, 769, 1000, 1009, 1296, 2017}

func TestFFTMatchesDFT(t *testing.T) {
	for _, n := range lengths {
		x := complexSignal(n)
		near(t, Route(n), FFT(x), dft(x, false))
		near(t, Route(n)+" inverse", IFFT(x), dft(x, true))
	}
}

func TestRoundTrips(t *testing.T) {
	for _, n := range append([]int{1, 63}, lengths...) {
		x := complexSignal(n)
		near(t, "IFFT(FFT(x))", IFFT(FFT(x)), x)
	}
}

func TestImpulseConstantAndTone(t *testing.T) {
	for _, n := range []int{4, 6, 8, 9, 13, 15, 23, 25, 49} {
		imp := make([]complex128, n)
		imp[0] = 1
		ones := make([]complex128, n)
		for i := range ones {
			ones[i] = 1
		}
		near(t, "impulse", FFT(imp), ones)
		spike := make([]complex128, n)
		spike[0] = complex(float64(n), 0)
		near(t, "constant", FFT(ones), spike)
	}
	n, f := 16, 3
	x := make([]complex128, n)
	for i := range x {
		x[i] = cmplx.Rect(1, 2*math.Pi*float64(f*i)/float64(n))
	}
	for k, v := range FFT(x) {
		want := 0.0
		if k == f {
			want = float64(n)
		}
		if math.Abs(cmplx.Abs(v)-want) > tol {
			t.Fatalf("tone bin %d: |X| = %g, want %g", k, cmplx.Abs(v), want)
		}
	}
}

func TestLinearity(t *testing.T) {
	for _, n := range []int{10, 1009} {
		a, b := complex(2, -1), complex(0.5, 3)
		x, y := complexSignal(n), make([]complex128, n)
		for i := range y {
			y[i] = complex(math.Cos(float64(i)), float64(i%4))
		}
		mix := make([]complex128, n)
		for i := range mix {
			mix[i] = a*x[i] + b*y[i]
		}
		fx, fy := FFT(x), FFT(y)
		want := make([]complex128, n)
		for i := range want {
			want[i] = a*fx[i] + b*fy[i]
		}
		near(t, "linearity", FFT(mix), want)
	}
}

func TestFFTRealAndEdges(t *testing.T) {
	for _, n := range []int{1, 5, 8, 12} {
		r := realSignal(n)
		near(t, "FFTReal", FFTReal(r), FFT(promote(r)))
	}
	if got := FFT([]complex128{}); got == nil || len(got) != 0 {
		t.Fatalf("FFT of empty: %v", got)
	}
	if got := IFFT([]complex128{}); got == nil || len(got) != 0 {
		t.Fatalf("IFFT of empty: %v", got)
	}
	if got := FFTReal([]float64{}); got == nil || len(got) != 0 {
		t.Fatalf("FFTReal of empty: %v", got)
	}
	one := []complex128{complex(7, -3)}
	if got := FFT(one); len(got) != 1 || got[0] != one[0] {
		t.Fatalf("FFT of one value: %v", got)
	}
	if got := IFFT(one); len(got) != 1 || got[0] != one[0] {
		t.Fatalf("IFFT of one value: %v", got)
	}
}

func TestInputsUnchanged(t *testing.T) {
	for _, n := range []int{8, 10, 1009} {
		x := complexSignal(n)
		keep := append([]complex128(nil), x...)
		FFT(x)
		IFFT(x)
		for i := range x {
			if x[i] != keep[i] {
				t.Fatalf("n=%d: input changed at %d", n, i)
			}
		}
	}
}

func TestPlanReuseAliasAndTrivial(t *testing.T) {
	for _, n := range []int{8, 9, 12, 23, 1000, 1296, 10007} {
		p := NewPlan(n)
		if p.Len() != n {
			t.Fatalf("Len() = %d, want %d", p.Len(), n)
		}
		x := complexSignal(n)
		y := p.FFT(make([]complex128, n), x)
		if n <= 2000 {
			near(t, "plan", y, dft(x, false))
		}
		near(t, "plan round trip", p.IFFT(make([]complex128, n), y), x)
		near(t, "second call", p.FFT(make([]complex128, n), x), y)
		z := append([]complex128(nil), x...)
		p.FFT(z, z) // dst aliasing src
		near(t, "in place", z, y)
	}
	if got := NewPlan(0).FFT([]complex128{}, []complex128{}); len(got) != 0 {
		t.Fatalf("n=0: %v", got)
	}
	one := []complex128{complex(7, -3)}
	if got := NewPlan(1).IFFT(make([]complex128, 1), one); got[0] != one[0] {
		t.Fatalf("n=1: %v", got)
	}
}

func TestRouting(t *testing.T) {
	// plan.go:42-73: trivial, powers of two, 13-smooth, primes >= 700, others
	want := map[int]string{0: "trivial", 1: "trivial", 4096: "iterative", 6720: "mixedradix", 169: "mixedradix",
		1009: "rader", 10007: "rader", 769: "rader", 641: "bluestein", 17: "bluestein", 20014: "bluestein"}
	for n, r := range want {
		if got := Route(n); got != r {
			t.Fatalf("Route(%d) = %s, want %s", n, got, r)
		}
	}
}
