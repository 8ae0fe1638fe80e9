"""mts_error_norms_noalloc, called through the C ABI, against mts_error_norms.

The allocation-free entry point keeps its partials and their reduction in
the library's fixed device array; its L-inf, L2 and argmax must equal the
instance-buffer entry point's bit for bit (the same partials, reduced in the
same order).  Two calls in a row, the second on a smaller grid, so the
second reads none of the first call's partials; then a third call on the
first instance again, and an FP32 instance.
"""
import ctypes as C

import pytest

pytestmark = pytest.mark.gpu


def _norms(mts, inst, t, fn):
    linf, l2 = C.c_double(), C.c_double()
    am = (C.c_int * 3)()
    rc = getattr(mts.lib(), fn)(inst._h, t, C.byref(linf), C.byref(l2), am)
    assert rc == 0, mts.lib().mts_last_error().decode()
    return linf.value, l2.value, (am[0], am[1], am[2])


def _instance(mts, ctx, n, method, steps, precision=None):
    phys = mts.physics("sincos", 10.0, 1.0)
    st = mts.Setup(mts.grid(n), mts.surface("ellipsoid", 0.7, 0.5, 0.4, 1 - 2 ** -20), phys,
                   mib=2)
    dt = 2.0 / (max(n) if isinstance(n, tuple) else n - 1)
    inst = (mts.Instance(ctx, st, method, dt) if precision is None
            else mts.Instance(ctx, st, method, dt, precision=precision))
    inst.initialize(0.0)
    _, t_end = inst.run(0.0, steps * dt)
    return inst, t_end


def test_noalloc_norms_match_the_instance_buffer_form(mts, ctx):
    print("case: no-alloc norms, 73^3 LOD-CN then (21, 23, 19) TS then 73^3 again", flush=True)
    big, tb = _instance(mts, ctx, 73, "lod-cn", 3)
    small, ts_ = _instance(mts, ctx, (21, 23, 19), "ts", 2)
    # 73^3 nodes use 1,517 partials, the small grid 36
    for inst, t in ((big, tb), (small, ts_), (big, tb)):
        ref = _norms(mts, inst, t, "mts_error_norms")
        got = _norms(mts, inst, t, "mts_error_norms_noalloc")
        print(f"  {ref} vs {got}", flush=True)
        assert got == ref
    big.close()
    small.close()


def test_noalloc_norms_fp32_instance(mts, ctx):
    print("case: no-alloc norms, FP32 LOD-CN 41^3", flush=True)
    inst, t = _instance(mts, ctx, 41, "lod-cn", 2, precision="fp32")
    assert _norms(mts, inst, t, "mts_error_norms_noalloc") == _norms(mts, inst, t,
                                                                       "mts_error_norms")
    inst.close()
