"""GPU energy telemetry (ROCm SMI energy counter behind hsps_energy_*): the
counter advances while the device works (it is updated about once a second,
so the test spans whole update intervals), and the energy per streamed byte
is in a physical range.  No reference counterpart to compare values against
(the reference reads NVML / RAPL)."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_energy_counter_advances_under_load(H, ctx):
    try:
        E = H.Energy(0)
    except H.HspsError as e:
        pytest.skip(f"no GPU energy counter on this device: {e}")
    info = E.info()
    assert info["name"] != ""
    n = 8 * 1024 * 1024
    x, y = ctx.array(np.ones(n)), ctx.array(n=n)
    L = H.lib()
    reps = 0

    def work():
        nonlocal reps
        for _ in range(20):
            H.check(L.hsps_axpy_f64(ctx.h, n, 1.0, x.ptr, y.ptr))
        ctx.sync()
        reps += 20

    _, _, tp = E.read()
    while True:  # start on a counter update seen while working
        work()
        j0, _, t0 = E.read()
        if t0 != tp:
            break
    r0 = reps
    while True:
        work()
        j1, _, t1 = E.read()
        if t1 - t0 >= 2.0:
            break
    dj, dt = j1 - j0, t1 - t0
    watts = dj / dt
    assert dj > 0 and 50.0 < watts < 2000.0, (dj, dt, watts)
    # J per streamed byte (whole board, static power included): 1 pJ .. 10 nJ
    jb = dj / ((reps - r0) * 3.0 * n * 8)
    assert 1e-12 < jb < 1e-8, jb
    np.testing.assert_allclose(y.download()[:4], reps * np.ones(4))
    E.close()


def test_energy_reports_unsupported_cleanly(H):
    """An absent device index is an error code, not a crash."""
    h = H.C.c_void_p()
    rc = H.lib().hsps_energy_open(9999, H.C.byref(h))
    assert rc < 0 and not h.value
