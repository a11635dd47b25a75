package gofft

import (
	"math"
	"math/cmplx"
	"testing"
)

// The semantics of the reference's rfft_test.go and realplan_test.go.

var realLengths = []int{2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 15, 16, 17, 20, 24, 31, 32, 36, 64, 100, 128, 1000, 2018}

func TestRFFTIsTheLowerHalfOfTheFullSpectrum(t *testing.T) {
	for _, n := range realLengths {
		x := realSignal(n)
		got := RFFT(x)
		if len(got) != n/2+1 {
			t.Fatalf("n=%d: %d bins, want %d", n, len(got), n/2+1)
		}
		near(t, "RFFT", got, FFTReal(x)[:n/2+1])
	}
}

func TestRFFTSymmetryAndRealEdges(t *testing.T) {
	n := 12
	x := realSignal(n)
	half, full := RFFT(x), FFTReal(x)
	for k := 1; k < n-k; k++ {
		if d := cmplx.Abs(full[n-k] - cmplx.Conj(half[k])); d > tol {
			t.Fatalf("bin %d is not the mirror (off by %g)", k, d)
		}
	}
	if math.Abs(imag(half[0])) > tol || math.Abs(imag(half[n/2])) > tol {
		t.Fatalf("DC %v or Nyquist %v not real", half[0], half[n/2])
	}
}

func TestIRFFTRoundTripAndMirror(t *testing.T) {
	for _, n := range append([]int{1}, realLengths...) {
		x := realSignal(n)
		nearReal(t, "IRFFT(RFFT(x))", IRFFT(RFFT(x), n), x)
	}
	// a hand-made half spectrum against the inverse of its Hermitian mirror
	n := 6
	half := []complex128{3, complex(1, -2), complex(-0.5, 0.4), 2}
	full := []complex128{half[0], half[1], half[2], half[3], cmplx.Conj(half[2]), cmplx.Conj(half[1])}
	inv := IFFT(full)
	got := IRFFT(half, n)
	for i := range got {
		if math.Abs(got[i]-real(inv[i])) > tol {
			t.Fatalf("index %d: %g, want %g", i, got[i], real(inv[i]))
		}
	}
	// the imaginary parts of DC and Nyquist are ignored
	odd := []complex128{complex(3, 5), complex(1, -2), complex(-0.5, 0.4), complex(2, -7)}
	nearReal(t, "DC/Nyquist imaginary parts", IRFFT(odd, n), got)
}

func TestRFFTOfACosine(t *testing.T) {
	n, f := 16, 3
	x := make([]float64, n)
	for i := range x {
		x[i] = math.Cos(2 * math.Pi * float64(f*i) / float64(n))
	}
	for k, v := range RFFT(x) {
		want := 0.0
		if k == f {
			want = float64(n) / 2
		}
		if math.Abs(cmplx.Abs(v)-want) > tol {
			t.Fatalf("bin %d: |X| = %g, want %g", k, cmplx.Abs(v), want)
		}
	}
}

func TestRealEdgeCases(t *testing.T) {
	if got := RFFT([]float64{}); got == nil || len(got) != 0 {
		t.Fatalf("RFFT of empty: %v", got)
	}
	if got := RFFT([]float64{5}); len(got) != 1 || got[0] != 5 {
		t.Fatalf("RFFT of one value: %v", got)
	}
	for _, n := range []int{0, -3} {
		if got := IRFFT([]complex128{1}, n); got == nil || len(got) != 0 {
			t.Fatalf("IRFFT n=%d: %v", n, got)
		}
	}
	if got := IRFFT([]complex128{4}, 1); len(got) != 1 || math.Abs(got[0]-4) > tol {
		t.Fatalf("IRFFT n=1: %v", got)
	}
	// a short spectrum reads as zero padded
	nearReal(t, "short spectrum", IRFFT([]complex128{2}, 4), IRFFT([]complex128{2, 0, 0}, 4))
	x := realSignal(8)
	keep := append([]float64(nil), x...)
	spec := RFFT(x)
	skeep := append([]complex128(nil), spec...)
	IRFFT(spec, 8)
	for i := range x {
		if x[i] != keep[i] {
			t.Fatalf("RFFT changed its input at %d", i)
		}
	}
	for i := range spec {
		if spec[i] != skeep[i] {
			t.Fatalf("IRFFT changed its input at %d", i)
		}
	}
}

func TestRealPlanReuse(t *testing.T) {
	for _, n := range []int{8, 10, 15, 1024, 20014} {
		p := NewRealPlan(n)
		if p.Len() != n {
			t.Fatalf("Len() = %d", p.Len())
		}
		x := realSignal(n)
		X := p.RFFT(make([]complex128, n/2+1), x)
		near(t, "RealPlan.RFFT", p.RFFT(make([]complex128, n/2+1), x), X)
		nearReal(t, "RealPlan.IRFFT", p.IRFFT(make([]float64, n), X), x)
	}
	if got := NewRealPlan(0).RFFT([]complex128{}, []float64{}); len(got) != 0 {
		t.Fatalf("n=0: %v", got)
	}
	if got := NewRealPlan(1).RFFT(make([]complex128, 1), []float64{2.5}); got[0] != 2.5 {
		t.Fatalf("n=1: %v", got)
	}
}
