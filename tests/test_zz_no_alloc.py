"""dk_apply_noalloc (the entry point that allocates no device memory) against
dk_apply, through the C ABI.

The two entry points share their kernels; dk_apply_noalloc only keeps the
norm's block partials in the library's fixed scratch array instead of the
lane's buffer, so values and norms agree bit for bit (the arithmetic and its
order are the same).  Two calls in a row, the second on a smaller state,
check that the shared array carries nothing from one call into the next.
The file is collected last: it runs alone in a fresh process and after the
rest of the suite.
"""
import numpy as np
import pytest

from doki_amd import workloads as W

pytestmark = pytest.mark.gpu


def _gates(rng, n):
    """One-qubit gates on every target class (lane, half-wave, pair kernels),
    with controls and anticontrols, a 3-qubit tile gate and a 5-qubit dense
    (MFMA) gate."""
    out = []
    for t in (0, 3, 7, n - 1):
        out.append(([t], W.gate_u3(0.4 + t, 1.3, 2.2), (), ()))
    out.append(([2], W.gate_rx(0.9), ((5,)), ((n - 2,))))
    out.append(([1, 6, n - 3], W.haar_unitary(rng, 3), (4,), ()))
    out.append(([0, 2, 5, 8, 9], W.haar_unitary(rng, 5), (), ()))
    return out


def _both(capi, psi, targets, u, ctrl, actrl, dtype):
    n = int(np.log2(psi.size))
    src = capi.from_host(psi)
    if dtype == capi.C64:
        src = capi.convert(src, capi.C64)
    g = capi.Gate(u)
    ref = capi.apply(src, g, targets, ctrl, actrl)           # allocating form
    out = capi.new_typed(n, dtype)                           # caller-owned
    got = capi.apply_noalloc(src, g, targets, ctrl, actrl, out=out)
    assert got is out
    return ref, got, src, g


@pytest.mark.parametrize("dtype", [0, 1])
def test_noalloc_equals_apply_two_calls(capi, dtype):
    rng = W.rng_for(8080 + dtype)
    for n in (20, 12):   # the second call on a smaller state
        psi = W.random_normalised_state(900 + n, n)
        for targets, u, ctrl, actrl in _gates(rng, n):
            ref, got, src, g = _both(capi, psi, targets, u, ctrl, actrl, dtype)
            rv, rn = ref.raw()
            gv, gn = got.raw()
            assert np.array_equal(rv, gv), (n, targets)
            assert rn == gn, (n, targets, rn, gn)
            # in place (out == in) gives the same values again
            capi.apply_noalloc(src, g, targets, ctrl, actrl)
            sv, sn = src.raw()
            assert np.array_equal(sv, rv) and sn == rn, (n, targets)


def test_noalloc_limits(capi):
    psi = W.random_normalised_state(5, 10)
    src = capi.from_host(psi)
    g = capi.Gate(W.gate_h())
    with pytest.raises(capi.DokiHipError) as e:
        capi.apply_noalloc(src, g, [0], out=capi.new(11))
    assert e.value.code == 5 and "qubit count" in str(e.value)
    with pytest.raises(capi.DokiHipError) as e:
        capi.apply_noalloc(src, g, [10])
    assert e.value.code == 5 and "Target qubit out of range" in str(e.value)
    big = capi.new_typed(27, capi.C64)   # 1 GiB, above the scratch's limit
    with pytest.raises(capi.DokiHipError) as e:
        capi.apply_noalloc(big, g, [0])
    assert e.value.code == 5 and "fixed scratch" in str(e.value)
    del big
    # the limit itself is accepted
    top = capi.new_typed(26, capi.C64)
    capi.apply_noalloc(top, g, [25])
    p = capi.prob(top, 25)
    assert abs(p - 0.5) <= 1e-6
