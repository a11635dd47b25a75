package gofft

import "math"

// sumOfCosines is the shared form of the generalised cosine windows
// (window.go:53-78): w[k] = a0 - a1 cos(t) + a2 cos(2t) - ..., t = 2 pi k /
// (n-1), symmetric; n == 1 gives [1], n <= 0 an empty non-nil slice.
func sumOfCosines(n int, a ...float64) []float64 {
	if n <= 0 {
		return []float64{}
	}
	if n == 1 {
		return []float64{1}
	}
	w := make([]float64, n)
	for k := range w {
		t := 2 * math.Pi * float64(k) / float64(n-1)
		v, sign := a[0], -1.0
		for j := 1; j < len(a); j++ {
			v += sign * a[j] * math.Cos(float64(j)*t)
			sign = -sign
		}
		w[k] = v
	}
	return w
}

// Hann is the symmetric Hann window (window.go:14-22).
func Hann(n int) []float64 { return sumOfCosines(n, 0.5, 0.5) }

// Hamming is the symmetric Hamming window (window.go:24-31).
func Hamming(n int) []float64 { return sumOfCosines(n, 0.54, 0.46) }

// Blackman is the symmetric Blackman window (window.go:33-40).
func Blackman(n int) []float64 { return sumOfCosines(n, 0.42, 0.5, 0.08) }

// BlackmanHarris is the symmetric 4-term Blackman-Harris window
// (window.go:42-51).
func BlackmanHarris(n int) []float64 { return sumOfCosines(n, 0.35875, 0.48829, 0.14128, 0.01168) }

// Bartlett is the triangular window, 0 at both ends and 1 at the centre
// (window.go:80-100).
func Bartlett(n int) []float64 {
	if n <= 0 {
		return []float64{}
	}
	if n == 1 {
		return []float64{1}
	}
	w := make([]float64, n)
	for k := range w {
		w[k] = 1 - math.Abs(2*float64(k)/float64(n-1)-1)
	}
	return w
}
