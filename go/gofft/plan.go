package gofft

/*
#include "gofft.h"
*/
import "C"

import (
	"runtime"
	"sync"
)

// Plan is a precomputed transform of one length (plan.go:19-40): immutable
// after NewPlan and safe for concurrent use.
type Plan struct {
	n int
	h *handle
}

// NewPlan builds a complex128 plan for length n (plan.go:42-73). Lengths
// route as the reference's do: powers of two and 13-smooth lengths to
// Stockham kernels, primes of at least 700 to Rader, others to Bluestein
// (reproducing the reference's Bluestein rounding).
func NewPlan(n int) *Plan {
	p := &Plan{n: n}
	if n > 1 {
		p.h = newHandle(func(pp **C.gofft_plan) C.int {
			return C.gofft_plan_c2c(pp, C.int64_t(n), 1, C.GOFFT_C128, 0)
		})
	}
	return p
}

// Len reports the transform length the plan was built for.
func (p *Plan) Len() int { return p.n }

// FFT writes the forward DFT of src into dst and returns dst. dst and src
// must each have length Len(); dst may alias src; src is not modified unless
// it aliases dst (plan.go:88-94).
func (p *Plan) FFT(dst, src []complex128) []complex128 {
	p.execute(dst, src, C.GOFFT_FORWARD)
	return dst
}

// IFFT writes the inverse DFT of src into dst, normalised by 1/n, and
// returns dst (plan.go:96-109).
func (p *Plan) IFFT(dst, src []complex128) []complex128 {
	p.execute(dst, src, C.GOFFT_INVERSE)
	return dst
}

func (p *Plan) execute(dst, src []complex128, dir C.int) {
	n := p.n
	if n <= 1 {
		copy(dst, src)
		return
	}
	if len(dst) < n || len(src) < n {
		panic("fft: dst and src must have length Len()")
	}
	call(func() C.int {
		return C.gofft_exec_c2c_host(p.h.p, cptr(src), cptr(dst), dir)
	})
	runtime.KeepAlive(p)
}

// The plan caches of the package functions (plan.go:281-317): look up under
// the lock, build outside it, and keep the first plan stored for a length.
var (
	planMu    sync.Mutex
	planCache = map[int]*Plan{}
)

func cachedPlan(n int) *Plan {
	planMu.Lock()
	p, ok := planCache[n]
	planMu.Unlock()
	if ok {
		return p
	}
	p = NewPlan(n)
	planMu.Lock()
	defer planMu.Unlock()
	if q, ok := planCache[n]; ok {
		return q
	}
	planCache[n] = p
	return p
}

// RealPlan is a precomputed real-input transform of one length
// (realplan.go:14-35), safe for concurrent use.
type RealPlan struct {
	n int
	h *handle
}

// NewRealPlan builds a float64 real-input plan for length n
// (realplan.go:37-58).
func NewRealPlan(n int) *RealPlan {
	p := &RealPlan{n: n}
	if n > 1 {
		p.h = newHandle(func(pp **C.gofft_plan) C.int {
			return C.gofft_plan_r2c(pp, C.int64_t(n), 1, C.GOFFT_C128, 0)
		})
	}
	return p
}

// Len reports the real-input length the plan was built for.
func (p *RealPlan) Len() int { return p.n }

// RFFT writes the n/2+1 non-redundant bins of the real signal src into dst
// and returns dst[:n/2+1] (realplan.go:74-116). src is not modified.
func (p *RealPlan) RFFT(dst []complex128, src []float64) []complex128 {
	n := p.n
	if n == 0 {
		return dst[:0]
	}
	if n == 1 {
		dst[0] = complex(src[0], 0)
		return dst[:1]
	}
	if len(src) < n || len(dst) < n/2+1 {
		panic("fft: src must have length Len() and dst Len()/2+1")
	}
	call(func() C.int { return C.gofft_exec_r2c_host(p.h.p, rptr(src), cptr(dst)) })
	runtime.KeepAlive(p)
	return dst[:n/2+1]
}

// IRFFT writes the real signal of length n whose half spectrum is src into
// dst and returns dst (realplan.go:171-192): min(len(src), n/2+1) bins are
// read, missing bins count as zero, and the imaginary parts of the DC and
// Nyquist bins are ignored; the result is normalised by 1/n.
func (p *RealPlan) IRFFT(dst []float64, src []complex128) []float64 {
	n := p.n
	if n <= 0 {
		return dst[:0]
	}
	if n == 1 {
		dst[0] = 0
		if len(src) > 0 {
			dst[0] = real(src[0])
		}
		return dst[:1]
	}
	if len(dst) < n {
		panic("fft: dst must have length Len()")
	}
	bins := len(src)
	if bins > n/2+1 {
		bins = n/2 + 1
	}
	call(func() C.int {
		return C.gofft_exec_c2r_host(p.h.p, cptr(src[:bins]), C.int64_t(bins), rptr(dst))
	})
	runtime.KeepAlive(p)
	return dst[:n]
}

var (
	realPlanMu    sync.Mutex
	realPlanCache = map[int]*RealPlan{}
)

// cachedRealPlan mirrors rfft.go:47-64 (built under its own lock).
func cachedRealPlan(n int) *RealPlan {
	realPlanMu.Lock()
	defer realPlanMu.Unlock()
	p, ok := realPlanCache[n]
	if !ok {
		p = NewRealPlan(n)
		realPlanCache[n] = p
	}
	return p
}
