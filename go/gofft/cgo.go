// Package gofft is the go-fft transform API (NewPlan, Plan.FFT/IFFT,
// NewRealPlan, RealPlan.RFFT/IRFFT, FFT/IFFT/FFTReal, RFFT/IRFFT,
// FFTN/IFFTN/FFT2/IFFT2, RFFT2/IRFFT2, PSD, Spectrogram, the windows and the
// frequency helpers) running on one MI355X through libgofft_hip.so, the
// gfx950 HIP backend, over cgo. Names, argument meanings, normalisation and
// panic texts follow the reference package (plan.go, realplan.go, fft.go,
// rfft.go, ndfft.go, rfft2.go, freq.go, window.go); the plain path is
// complex128 / float64 throughout.
//
// Every transform stages the caller's slices through the library's pinned
// buffers and device memory and returns when the result is in the Go slice;
// no Go pointer is kept past a call. Plans are immutable and safe for
// concurrent use: concurrent calls on one plan run on separate HIP streams.
package gofft

/*
#cgo CFLAGS: -I${SRCDIR}/../../include
#cgo LDFLAGS: -L${SRCDIR}/../../fft_amd -lgofft_hip -Wl,-rpath,${SRCDIR}/../../fft_amd
#include <stdlib.h>
#include "gofft.h"
*/
import "C"

import (
	"runtime"
	"unsafe"
)

// call runs one library call on a locked OS thread and panics with the
// library's message on failure: the message is thread-local, so it must be
// read on the thread that made the call.
func call(f func() C.int) {
	runtime.LockOSThread()
	defer runtime.UnlockOSThread()
	if rc := f(); rc != C.GOFFT_OK {
		panic(C.GoString(C.gofft_last_error()))
	}
}

func cptr(s []complex128) unsafe.Pointer {
	if len(s) == 0 {
		return nil
	}
	return unsafe.Pointer(&s[0])
}

func rptr(s []float64) unsafe.Pointer {
	if len(s) == 0 {
		return nil
	}
	return unsafe.Pointer(&s[0])
}

// handle owns one library plan and destroys it when collected.
type handle struct {
	p *C.gofft_plan
}

func newHandle(build func(pp **C.gofft_plan) C.int) *handle {
	h := &handle{}
	var p *C.gofft_plan
	call(func() C.int { return build(&p) })
	h.p = p
	runtime.SetFinalizer(h, func(h *handle) { C.gofft_plan_destroy(h.p) })
	return h
}

// Route names the engine a length routes to, as the reference's NewPlan
// chooses (plan.go:42-73): "trivial", "iterative", "mixedradix", "rader" or
// "bluestein".
func Route(n int) string {
	return [...]string{"trivial", "iterative", "mixedradix", "rader", "bluestein"}[C.gofft_route(C.int64_t(n))]
}
