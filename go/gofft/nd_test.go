package gofft

import (
	"math"
	"math/cmplx"
	"testing"
)

// The semantics of the reference's ndfft_test.go and rfft2_test.go.

func grid(shape []int) []complex128 {
	total := 1
	for _, s := range shape {
		total *= s
	}
	x := make([]complex128, total)
	for i := range x {
		x[i] = complex(math.Sin(0.7*float64(i))+float64(i%3), math.Cos(0.3*float64(i))-float64(i%2))
	}
	return x
}

func TestFFTNMatchesSeparableDFT(t *testing.T) {
	for _, shape := range [][]int{{2, 2}, {3, 4}, {4, 3}, {1, 5}, {5, 1}, {2, 3, 4}, {3, 1, 2}, {2, 2, 2, 2}, {6, 5},
		{16, 1009}, {9, 23}} {
		x := grid(shape)
		near(t, "FFTN", FFTN(x, shape), dftN(x, shape))
		near(t, "IFFTN(FFTN(x))", IFFTN(FFTN(x, shape), shape), x)
	}
	x := grid([]int{4, 6})
	near(t, "FFT2", FFT2(x, [2]int{4, 6}), FFTN(x, []int{4, 6}))
	near(t, "IFFT2", IFFT2(FFT2(x, [2]int{4, 6}), [2]int{4, 6}), x)
}

func TestFFTNClosedForms(t *testing.T) {
	ones := make([]complex128, 12)
	for i := range ones {
		ones[i] = 1
	}
	spike := make([]complex128, 12)
	spike[0] = 12
	near(t, "constant", FFTN(ones, []int{3, 4}), spike)
	imp := make([]complex128, 16)
	imp[0] = 1
	flat := make([]complex128, 16)
	for i := range flat {
		flat[i] = 1
	}
	near(t, "impulse", FFTN(imp, []int{4, 4}), flat)
	x := []complex128{complex(3, -2)}
	if got := FFTN(x, []int{}); len(got) != 1 || got[0] != x[0] {
		t.Fatalf("empty shape: %v", got)
	}
	if got := IFFTN(x, []int{}); len(got) != 1 || got[0] != x[0] {
		t.Fatalf("empty shape inverse: %v", got)
	}
}

func TestFFTNPanics(t *testing.T) {
	mustPanic(t, "product", "fft: shape product does not match len(data)", func() { FFTN(make([]complex128, 10), []int{3, 4}) })
	mustPanic(t, "zero axis", "fft: shape lengths must be positive", func() { FFTN(nil, []int{0, 4}) })
	mustPanic(t, "negative axis", "fft: shape lengths must be positive", func() { FFTN(make([]complex128, 4), []int{-1, 4}) })
}

func realGrid(rows, cols int) []float64 {
	x := make([]float64, rows*cols)
	for i := range x {
		x[i] = math.Sin(0.7*float64(i)) + 0.3*math.Cos(1.9*float64(i)) + float64(i%4)
	}
	return x
}

func TestRFFT2MatchesFFT2(t *testing.T) {
	for _, sh := range [][2]int{{4, 6}, {3, 4}, {5, 8}, {2, 2}, {4, 5}, {1, 6}, {6, 1}, {7, 3}, {64, 100}} {
		rows, cols := sh[0], sh[1]
		rc := cols/2 + 1
		x := realGrid(rows, cols)
		full := FFT2(promote(x), sh)
		got := RFFT2(x, sh)
		if len(got) != rows*rc {
			t.Fatalf("%v: %d bins", sh, len(got))
		}
		for r := 0; r < rows; r++ {
			for c := 0; c < rc; c++ {
				if d := cmplx.Abs(got[r*rc+c] - full[r*cols+c]); d > tol {
					t.Fatalf("%v bin (%d,%d) off by %g", sh, r, c, d)
				}
			}
		}
		nearReal(t, "IRFFT2(RFFT2(x))", IRFFT2(got, sh), x)
	}
}

func TestRFFT2EdgesAndPanics(t *testing.T) {
	ones := make([]float64, 12)
	for i := range ones {
		ones[i] = 1
	}
	X := RFFT2(ones, [2]int{3, 4})
	for i, v := range X {
		want := 0.0
		if i == 0 {
			want = 12
		}
		if cmplx.Abs(v-complex(want, 0)) > tol {
			t.Fatalf("bin %d: %v", i, v)
		}
	}
	sh := [2]int{2, 4}
	full := RFFT2(realGrid(2, 4), sh)
	short := full[:len(full)-1]
	padded := append(append([]complex128(nil), short...), 0)
	nearReal(t, "short spectrum", IRFFT2(short, sh), IRFFT2(padded, sh))
	mustPanic(t, "rows 0", "fft: shape lengths must be positive", func() { RFFT2(make([]float64, 4), [2]int{0, 4}) })
	mustPanic(t, "cols < 0", "fft: shape lengths must be positive", func() { RFFT2(make([]float64, 4), [2]int{2, -4}) })
	mustPanic(t, "product", "fft: shape product does not match len(data)", func() { RFFT2(make([]float64, 5), [2]int{2, 4}) })
	mustPanic(t, "inverse rows 0", "fft: shape lengths must be positive", func() { IRFFT2(make([]complex128, 4), [2]int{0, 4}) })
}
