package gofft

import (
	"math"
	"testing"
)

// The semantics of the reference's freq_test.go and window_test.go.

func TestFrequencies(t *testing.T) {
	nearReal(t, "FFTFreq(8,1)", FFTFreq(8, 1), []float64{0, 0.125, 0.25, 0.375, -0.5, -0.375, -0.25, -0.125})
	nearReal(t, "FFTFreq(5,1)", FFTFreq(5, 1), []float64{0, 0.2, 0.4, -0.4, -0.2})
	nearReal(t, "FFTFreq(4,0.5)", FFTFreq(4, 0.5), []float64{0, 0.5, -1, -0.5})
	nearReal(t, "FFTFreq(1,1)", FFTFreq(1, 1), []float64{0})
	nearReal(t, "RFFTFreq(8,1)", RFFTFreq(8, 1), []float64{0, 0.125, 0.25, 0.375, 0.5})
	nearReal(t, "RFFTFreq(5,1)", RFFTFreq(5, 1), []float64{0, 0.2, 0.4})
	nearReal(t, "RFFTFreq(4,0.25)", RFFTFreq(4, 0.25), []float64{0, 1, 2})
	for _, n := range []int{0, -2} {
		if got := FFTFreq(n, 1); got == nil || len(got) != 0 {
			t.Fatalf("FFTFreq(%d): %v", n, got)
		}
		if got := RFFTFreq(n, 1); got == nil || len(got) != 0 {
			t.Fatalf("RFFTFreq(%d): %v", n, got)
		}
	}
}

// Parseval: the PSD integrated over frequency equals the mean square.
func TestPSDParseval(t *testing.T) {
	for _, n := range []int{32, 15, 4096, 1 << 16} {
		d := 0.1
		x := make([]float64, n)
		for i := range x {
			x[i] = math.Sin(2*math.Pi*3*float64(i)/float64(n)) + 0.5*math.Cos(2*math.Pi*7*float64(i)/float64(n)) + float64(i%3)
		}
		p := PSD(x, d)
		if len(p) != n/2+1 {
			t.Fatalf("n=%d: %d densities", n, len(p))
		}
		f := RFFTFreq(n, d)
		integral, ms := 0.0, 0.0
		for _, v := range p {
			integral += v * (f[1] - f[0])
		}
		for _, v := range x {
			ms += v * v
		}
		ms /= float64(n)
		if math.Abs(integral-ms) > 1e-9*math.Max(1, ms) {
			t.Fatalf("n=%d: integral %.12g, mean square %.12g", n, integral, ms)
		}
	}
	if got := PSD([]float64{}, 1); got == nil || len(got) != 0 {
		t.Fatalf("PSD of empty: %v", got)
	}
}

func TestSpectrogram(t *testing.T) {
	x := make([]float64, 20)
	for i := range x {
		x[i] = math.Sin(float64(i))
	}
	win := Hann(8)
	frames := Spectrogram(x, 8, 4, win, 1)
	if len(frames) != 4 {
		t.Fatalf("%d frames, want 4", len(frames))
	}
	for k, fr := range frames {
		seg := make([]float64, 8)
		for i := range seg {
			seg[i] = x[4*k+i] * win[i]
		}
		nearReal(t, "frame", fr, PSD(seg, 1))
	}
	if got := Spectrogram([]float64{1, 2, 3}, 8, 0, Hann(8), 1); got != nil {
		t.Fatalf("no whole frame: %v", got)
	}
	mustPanic(t, "segment", "fft: segment length must be positive", func() { Spectrogram(nil, 0, 0, nil, 1) })
	mustPanic(t, "window", "fft: window length must equal segment length", func() { Spectrogram(nil, 4, 0, Hann(3), 1) })
	mustPanic(t, "overlap < 0", "fft: overlap must satisfy 0 <= overlap < segment", func() { Spectrogram(nil, 4, -1, Hann(4), 1) })
	mustPanic(t, "overlap = segment", "fft: overlap must satisfy 0 <= overlap < segment", func() { Spectrogram(nil, 4, 4, Hann(4), 1) })
}

func TestWindows(t *testing.T) {
	nearReal(t, "Hann(5)", Hann(5), []float64{0, 0.5, 1, 0.5, 0})
	nearReal(t, "Bartlett(5)", Bartlett(5), []float64{0, 0.5, 1, 0.5, 0})
	nearReal(t, "Bartlett(4)", Bartlett(4), []float64{0, 2.0 / 3, 2.0 / 3, 0})
	ham := Hamming(5)
	if math.Abs(ham[0]-0.08) > tol || math.Abs(ham[2]-1) > tol {
		t.Fatalf("Hamming(5) = %v", ham)
	}
	b := Blackman(7)
	if math.Abs(b[0]) > tol || math.Abs(b[3]-1) > tol {
		t.Fatalf("Blackman(7) = %v", b)
	}
	bh := BlackmanHarris(8)
	if math.Abs(bh[0]-(0.35875-0.48829+0.14128-0.01168)) > tol {
		t.Fatalf("BlackmanHarris(8)[0] = %g", bh[0])
	}
	for name, w := range map[string]func(int) []float64{"Hann": Hann, "Hamming": Hamming, "Blackman": Blackman,
		"BlackmanHarris": BlackmanHarris, "Bartlett": Bartlett} {
		if got := w(0); got == nil || len(got) != 0 {
			t.Fatalf("%s(0): %v", name, got)
		}
		if got := w(1); len(got) != 1 || got[0] != 1 {
			t.Fatalf("%s(1): %v", name, got)
		}
		v := w(9)
		for i := range v {
			if math.Abs(v[i]-v[8-i]) > tol {
				t.Fatalf("%s(9) not symmetric at %d", name, i)
			}
		}
	}
}
