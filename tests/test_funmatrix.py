"""The lazy matrix algebra of the `doki` extension (funmatrix_*), host side.

Nodes built from funmatrix_create / identity / statezero / densityzero /
hadamard and the operators over them are evaluated on the host, element by
element, in the order the reference's getters use (src/doki/funmatrix.c
getitem and the node functions), so results are compared bit-for-bit with
the reference's own outputs stored in golden/funmatrix.* (written by
golden/make_golden.py from the reference's C sources).

projection is pinned by mathematics instead (a row mask): the reference's
compiled parsers for projection and apply are not usable (see
make_golden.funmatrix_fixture).  apply nodes are evaluated on the GPU (one
gate application over the input column), so only their validation runs
here; their values, and chains whose leaves are density matrices of GPU
registries, are covered in test_gpu_funmatrix.py and test_gpu_parity.py.
"""
import json

import numpy as np
import pytest

import doki_amd
import fm_exprs
from helpers import GOLDEN

META = json.load(open(f"{GOLDEN}/funmatrix.json"))
ARR = np.load(f"{GOLDEN}/funmatrix.npz")


@pytest.fixture(scope="module")
def d():
    if not doki_amd.built():
        from doki_amd import build
        build.build()
    return doki_amd.doki


def F(d, name, *a):
    return getattr(d, "funmatrix_" + name)(*a, False)


def dense(d, fm):
    r, c = d.funmatrix_shape(fm, False)
    return np.array([[d.funmatrix_get(fm, i, j, False) for j in range(c)]
                     for i in range(r)], dtype=np.complex128)


def same_bits(a, b):
    """Equal as IEEE doubles (signed zeros too), element by element."""
    a = np.asarray(a, dtype=np.complex128).view(np.float64)
    b = np.asarray(b, dtype=np.complex128).view(np.float64)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


@pytest.mark.parametrize("seed", META["seeds"])
def test_chains_match_reference(d, seed):
    chains = fm_exprs.build(d, fm_exprs.random_inputs(seed))
    assert len(chains) >= 30
    for name, fm in chains.items():
        key = f"s{seed}_{name}"
        want = META["chains"][key]
        shape, mem, vals, tr = fm_exprs.snapshot(d, fm)
        assert list(shape) == want["shape"], key
        assert same_bits(vals, ARR[key]), key
        assert tr == want["trace"], key
        if want["mem"] is not None:
            assert mem == want["mem"], key
        else:
            # the reference raises SystemError for nodes over custom
            # matrices; this module returns its byte count
            assert isinstance(mem, int) and mem > 0, key


def test_error_and_edge_cases_match_reference(d):
    got = fm_exprs.run_error_cases(d)
    assert got == META["errors"]


def test_custom_matrix_mem(d):
    """Node record + 2-D matrix record + the element array."""
    assert F(d, "mem", F(d, "create", [[1, 2], [3, 4]])) == 112 + 40 + 4 * 16
    a = F(d, "create", [[1, 2], [3, 4]])
    assert F(d, "mem", F(d, "add", a, a)) == 112 + 2 * (112 + 40 + 64)


@pytest.mark.parametrize("q", [0, 1, 2])
@pytest.mark.parametrize("value", [False, True])
def test_projection_is_row_mask(d, q, value):
    inp = fm_exprs.random_inputs(11)
    ch = fm_exprs.build(d, inp)
    for name in ("B", "kron", "mm", "dag", "I2"):
        base = dense(d, ch[name])
        p = F(d, "projection", ch[name], q, value)
        keep = ((np.arange(base.shape[0]) >> q) & 1) == int(value)
        want = np.where(keep[:, None], base, 0)
        assert same_bits(dense(d, p), want), (name, q, value)
        assert F(d, "shape", p) == base.shape
    # projecting onto both values splits the matrix
    b = dense(d, ch["B"])
    p0 = dense(d, F(d, "projection", ch["B"], q, False))
    p1 = dense(d, F(d, "projection", ch["B"], q, True))
    assert np.array_equal(p0 + p1, b)


def test_projection_past_row_bits_and_trace(d):
    b = F(d, "create", [[1, 2], [3, 4]])
    assert np.array_equal(dense(d, F(d, "projection", b, 5, False)),
                          dense(d, b))
    assert not dense(d, F(d, "projection", b, 5, True)).any()
    assert F(d, "trace", F(d, "projection", b, 0, True)) == 4


@pytest.mark.parametrize("call,msg", [
    (lambda d, st, g1, g2: F(d, "apply", F(d, "create", [[1, 2]]), g1, [0],
                             None, None), "registry is not a column vector"),
    (lambda d, st, g1, g2: F(d, "apply", st, F(d, "create", [[1, 2, 3, 4]]),
                             [0], None, None),
     "gates have to be square matrices"),
    (lambda d, st, g1, g2: F(d, "apply", F(d, "statezero", 1), g2, [0, 1],
                             None, None),
     "gate is too big for this state vector"),
    (lambda d, st, g1, g2: F(d, "apply", st, g2, [0], None, None),
     "Wrong number of targets specified for that gate"),
    (lambda d, st, g1, g2: F(d, "apply", st, g1, 0, None, None),
     "target_list must be a list"),
    (lambda d, st, g1, g2: F(d, "apply", st, g1, [0], [1], None),
     "control_set must be a set or None"),
    (lambda d, st, g1, g2: F(d, "apply", st, g1, [0], None, [1]),
     "anticontrol_set must be a set or None"),
    (lambda d, st, g1, g2: F(d, "apply", st, g1, [0], {5}, None),
     "Control qubit out of range"),
    (lambda d, st, g1, g2: F(d, "apply", st, g1, [0], {1}, {1}),
     "A control cannot also be an anticontrol"),
    (lambda d, st, g1, g2: F(d, "apply", st, g1, [1], {1}, None),
     "A target cannot also be a control or an anticontrol"),
    (lambda d, st, g1, g2: F(d, "apply", st, g1, [3], None, None),
     "Target qubit out of range"),
    (lambda d, st, g1, g2: F(d, "apply", st, g1, ["a"], None, None),
     "target_list must be a list of qubit ids (unsigned integers)"),
    (lambda d, st, g1, g2: F(d, "apply", 5, g1, [0], None, None),
     "NULL pointer to registry"),
    (lambda d, st, g1, g2: F(d, "apply", st, 5, [0], None, None),
     "NULL pointer to gate"),
])
def test_apply_validation(d, call, msg):
    """Texts and order of the reference's checks,
    src/doki/doki.c:2228-2400."""
    st = F(d, "statezero", 2)
    g1 = F(d, "hadamard", 1)
    g2 = F(d, "hadamard", 2)
    with pytest.raises(d.error) as ei:
        call(d, st, g1, g2)
    assert str(ei.value) == msg


def test_nodes_outlive_their_operands(d):
    """Every node holds its operands: dropping the Python handles of the
    inputs leaves the chain readable."""
    inp = fm_exprs.random_inputs(0)
    ch = fm_exprs.build(d, inp)
    deep = ch["deep"]
    want = ARR["s0_deep"]
    del ch
    assert same_bits(dense(d, deep), want)


def test_live_reference_more_seeds(d):
    """With the reference's extension built from its sources (oracle/_ref),
    the same chains on fresh seeds agree bit-for-bit as well."""
    from oracle import oracle as O
    ref = O.load_reference()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    for seed in (21, 22):
        inp = fm_exprs.random_inputs(seed)
        ours, theirs = fm_exprs.build(d, inp), fm_exprs.build(ref, inp)
        for name in ours:
            a = fm_exprs.snapshot(d, ours[name])
            b = fm_exprs.snapshot(ref, theirs[name])
            assert a[0] == b[0] and a[3] == b[3], name
            assert same_bits(a[2], b[2]), name
            if b[1] is not None:
                assert a[1] == b[1], name
