package gofft

import (
	"math"
	"math/cmplx"
	"sync"
	"testing"
)

// dot is <a, b> = sum conj(a) b.
func dot(a, b []complex128) complex128 {
	var s complex128
	for i := range a {
		s += cmplx.Conj(a[i]) * b[i]
	}
	return s
}

// The adjoint identity <F x, g> = <x, F^H g> for the complex transforms.
func TestVJPAdjointIdentity(t *testing.T) {
	for _, n := range []int{16, 1009, 20014} {
		x, g := complexSignal(n), grid([]int{n})
		lhs := dot(g, FFT(x))
		rhs := dot(FFTVJP(g), x)
		if cmplx.Abs(lhs-rhs) > 1e-9*cmplx.Abs(lhs) {
			t.Fatalf("FFT n=%d: %v vs %v", n, lhs, rhs)
		}
		lhs = dot(g, IFFT(x))
		rhs = dot(IFFTVJP(g), x)
		if cmplx.Abs(lhs-rhs) > 1e-9*cmplx.Abs(lhs) {
			t.Fatalf("IFFT n=%d: %v vs %v", n, lhs, rhs)
		}
	}
}

// Central differences of L(x) = Re <g, RFFT(x)> against RFFTVJP.
func TestRFFTVJPFiniteDifferences(t *testing.T) {
	for _, n := range []int{8, 17, 1000} {
		x := realSignal(n)
		g := grid([]int{n/2 + 1})
		loss := func(v []float64) float64 { return real(dot(g, RFFT(v))) }
		grad := RFFTVJP(g, n)
		for _, i := range []int{0, n / 3, n - 1} {
			const eps = 1e-3
			xp := append([]float64(nil), x...)
			xm := append([]float64(nil), x...)
			xp[i] += eps
			xm[i] -= eps
			fd := (loss(xp) - loss(xm)) / (2 * eps)
			if math.Abs(fd-grad[i]) > 1e-6*math.Max(1, math.Abs(fd)) {
				t.Fatalf("n=%d i=%d: %g vs %g", n, i, fd, grad[i])
			}
		}
	}
}

// One plan from several goroutines at once (plan.go:12-14).
func TestConcurrentUseOfOnePlan(t *testing.T) {
	n := 4096
	p := NewPlan(n)
	inputs := make([][]complex128, 4)
	want := make([][]complex128, 4)
	for k := range inputs {
		inputs[k] = grid([]int{n})
		for i := range inputs[k] {
			inputs[k][i] *= complex(float64(k+1), 0)
		}
		want[k] = p.FFT(make([]complex128, n), inputs[k])
	}
	var wg sync.WaitGroup
	errs := make(chan string, 64)
	for k := range inputs {
		wg.Add(1)
		go func(k int) {
			defer wg.Done()
			for r := 0; r < 8; r++ {
				got := p.FFT(make([]complex128, n), inputs[k])
				for i := range got {
					if got[i] != want[k][i] {
						errs <- "mismatch"
						return
					}
				}
			}
		}(k)
	}
	wg.Wait()
	close(errs)
	for e := range errs {
		t.Fatal(e)
	}
}
