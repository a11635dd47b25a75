"""CPU: the library's host code under AddressSanitizer and
UndefinedBehaviorSanitizer, against a host stand-in for the HIP runtime
(tests/hostsim): plans of every kind and many shapes with every kernel launch
checked against the kernels' assumptions (workgroup shape, LDS footprint,
points per thread, radix schedule), the out-of-memory and launch-failure
paths driven by injected failures (status codes, gofft_last_error texts, no
leaks), a corrupt autotuner cache file, argument errors, and the Rader and
Bluestein index tables against their closed forms."""
import os
import subprocess

from conftest import ROOT

HOSTSIM = os.path.join(ROOT, "tests", "hostsim")


def test_host_code_under_sanitizers():
    b = subprocess.run(["make", "-s", "-j4"], cwd=HOSTSIM, capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stdout[-3000:] + b.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(HOSTSIM, "_build", "hostsim")], cwd=HOSTSIM, capture_output=True, text=True,
                       timeout=900, env=env)
    fails = [l for l in r.stdout.splitlines() if l.startswith("FAIL")]
    assert r.returncode == 0 and not fails, (fails[:20], r.stdout[-2000:], r.stderr[-4000:])
    assert r.stdout.count("\nok ") > 300
