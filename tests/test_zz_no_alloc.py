"""xf_run_steps_noalloc (the stepper entry point that allocates nothing; its
device memory is one static array per GPU) against xf_run_steps on the same
inputs: the same kernels, so the step log, the state, r and dr must be equal
bit for bit.  Two calls in a row, the second on a smaller input, so the
second call finds the first one's rows and log in the shared array."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _capi():
    from xflat_amd import capi
    if capi.device_count() < 1:
        pytest.fail("no HIP device visible")
    return capi


def _pair(capi, p, psi0, prec, steps, shape=None):
    """The same steps through both entry points; returns both results."""
    out = {}
    for name in ("alloc", "noalloc"):
        with capi.Context(p, precision=prec) as ctx:
            ctx.upload(psi0, p.R0, 1e-3)
            if shape:
                ctx.set_launch(*shape)
            launch = ctx.launch()                  # the shape is chosen at set-up
            run = ctx.run if name == "alloc" else ctx.run_noalloc
            info, rec = run(steps, log=True)
            out[name] = (info, rec, *ctx.download(), launch)
    return out


def _same(out):
    (ia, ra, pa, r_a, d_a, la), (ib, rb, pb, r_b, d_b, lb) = out["alloc"], out["noalloc"]
    assert la == lb
    assert (ia.steps, ia.accepted, ia.total_steps, ia.stop_reason) == (ib.steps, ib.accepted, ib.total_steps, ib.stop_reason)
    assert np.array_equal(ra, rb), "step logs differ"
    assert r_a == r_b and d_a == d_b
    assert np.array_equal(pa, pb), "states differ"
    return ia


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noalloc_equals_run_steps_twice_second_smaller(prec):
    capi = _capi()
    from xflat_amd.config import bulb_params
    from xflat_amd.seed import seeded_psi
    pr = capi.XF_FP64 if prec == "fp64" else capi.XF_FP32
    cases = [(bulb_params(Abins=200, Ebins=100, R0=10), 80),     # first call
             (bulb_params(Abins=24, Ebins=16, R0=20), 40)]       # second call, smaller input
    for k, (p, steps) in enumerate(cases):
        print(f"[noalloc] {prec} call {k}: {p.abins} x {p.ebins}, {steps} steps", flush=True)
        psi0 = seeded_psi(p.abins, p.ebins, seed=300 + k, amp=1e-6)
        info = _same(_pair(capi, p, psi0, pr, steps))
        assert info.steps == steps


def test_noalloc_streaming_shape_and_stop():
    """A streaming launch (mode 4, several tiles per workgroup) and a dump-trigger
    stop, through both entry points."""
    capi = _capi()
    from xflat_amd.config import bulb_params
    from xflat_amd.seed import seeded_psi
    p = bulb_params(Abins=120, Ebins=40, R0=10)
    psi0 = seeded_psi(120, 40, seed=77, amp=1e-6)
    print("[noalloc] mode 4, grid 8", flush=True)
    out = {}
    stop = capi.XfStop(0.0, 0.0, 7, 0)
    for name in ("alloc", "noalloc"):
        with capi.Context(p, precision=capi.XF_FP64, grid=8) as ctx:
            ctx.upload(psi0, 10.0, 1e-3)
            ctx.set_launch(4, 2, capi.block_for(2, 40, 2), 2)
            run = ctx.run if name == "alloc" else ctx.run_noalloc
            info, rec = run(60, log=True, stop=stop)
            out[name] = (info, rec, *ctx.download(), ctx.launch())
    info = _same(out)
    assert info.stop_reason == 2 and info.steps < 60


def test_noalloc_rejects_what_the_array_cannot_hold():
    capi = _capi()
    from xflat_amd.config import bulb_params
    p = bulb_params(Abins=24, Ebins=16, R0=20)
    with capi.Context(p, precision=capi.XF_FP64) as ctx:
        with pytest.raises(capi.XfError) as e:           # no launch shape chosen at set-up
            ctx.run_noalloc(5)
        assert e.value.code == -1
        ctx.launch()
        with pytest.raises(capi.XfError) as e:           # a log longer than the array's
            ctx.run_noalloc(4097, log=True)
        assert e.value.code == -1
        info, _ = ctx.run_noalloc(5)
        assert info.steps == 5
    # more workgroups than the array has exchange rows for
    q = bulb_params(Abins=4000, Ebins=8, R0=10)
    with capi.Context(q, precision=capi.XF_FP64) as ctx:
        ctx.set_launch(2, 1, 64, 1)
        grid = ctx.launch()["grid"]
        print(f"[noalloc] streaming grid {grid}", flush=True)
        assert grid > 1024
        with pytest.raises(capi.XfError) as e:
            ctx.run_noalloc(5)
        assert e.value.code == -1
        info, _ = ctx.run(5)                              # the allocating entry point still runs it
        assert info.steps == 5
