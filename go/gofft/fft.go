package gofft

// FFT returns the forward DFT of x (fft.go:14-23): unnormalised, the input
// unmodified, an empty non-nil slice for empty input and a copy for length 1.
func FFT(x []complex128) []complex128 {
	out := make([]complex128, len(x))
	if len(x) <= 1 {
		copy(out, x)
		return out
	}
	return cachedPlan(len(x)).FFT(out, x)
}

// IFFT returns the inverse DFT of x normalised by 1/n (fft.go:25-40).
func IFFT(x []complex128) []complex128 {
	out := make([]complex128, len(x))
	if len(x) <= 1 {
		copy(out, x)
		return out
	}
	return cachedPlan(len(x)).IFFT(out, x)
}

// FFTReal returns the full complex spectrum of a real signal (fft.go:42-51).
func FFTReal(x []float64) []complex128 {
	c := make([]complex128, len(x))
	for i, v := range x {
		c[i] = complex(v, 0)
	}
	return FFT(c)
}

// RFFT returns the n/2+1 non-redundant bins of a real signal (rfft.go:14-26).
func RFFT(x []float64) []complex128 {
	if len(x) == 0 {
		return []complex128{}
	}
	dst := make([]complex128, len(x)/2+1)
	return cachedRealPlan(len(x)).RFFT(dst, x)
}

// IRFFT returns the real signal of length n with half spectrum `spectrum`,
// normalised by 1/n (rfft.go:28-45); n <= 0 gives an empty non-nil slice.
func IRFFT(spectrum []complex128, n int) []float64 {
	if n <= 0 {
		return []float64{}
	}
	dst := make([]float64, n)
	return cachedRealPlan(n).IRFFT(dst, spectrum)
}
