"""Expression sets over the lazy matrix API (funmatrix_*), shared by the
golden-fixture writer and the tests, so that both build the same chains on
any module exposing that API.

Matrices enter through funmatrix_create with square inputs or a single row,
the shapes the reference's create stores in range (src/doki/doki.c:1306-1377
writes element (i, j) to i * rows + j).
"""
import numpy as np


def random_inputs(seed):
    rng = np.random.default_rng(seed)

    def rnd(r, c):
        z = rng.standard_normal((r, c)) + 1j * rng.standard_normal((r, c))
        return [[complex(x) for x in row] for row in z]

    return {"M2": rnd(2, 2), "M4": rnd(4, 4), "V": rnd(1, 4),
            "s1": complex(*rng.standard_normal(2)),
            "s2": complex(*rng.standard_normal(2))}


def build(mod, inp):
    """Named chains of every node kind except projection and apply."""
    def F(name, *a):
        return getattr(mod, "funmatrix_" + name)(*a, False)

    A = F("create", inp["M2"])
    B = F("create", inp["M4"])
    H = F("hadamard", 2)
    Id = F("identity", 1)
    out = {"A": A, "B": B, "H1": F("hadamard", 1), "H2": H,
           "H3": F("hadamard", 3), "I2": F("identity", 2),
           "sz": F("statezero", 2), "dz": F("densityzero", 2)}
    out["cu"] = F("addcontrol", A)
    out["cu2"] = F("addcontrol", out["cu"])
    out["ek"] = F("eyekron", A, 1, 1)
    out["ek2"] = F("eyekron", H, 0, 1)
    out["kron"] = F("kron", A, H)
    out["add"] = F("add", B, H)
    out["sub"] = F("sub", B, F("kron", A, Id))
    out["mm"] = F("matmul", B, H)
    out["ew"] = F("ewmul", B, H)
    out["smul"] = F("scalar_mul", out["mm"], inp["s1"])
    out["sdiv"] = F("scalar_div", out["ew"], inp["s2"])
    out["smul_int"] = F("scalar_mul", B, 3)
    out["sdiv_f"] = F("scalar_div", B, 0.25)
    out["tr"] = F("transpose", out["mm"])
    out["dag"] = F("dagger", F("scalar_mul", B, 2j))
    out["dagdag"] = F("dagger", out["dag"])
    out["trsum"] = F("add", F("transpose", B), F("dagger", B))
    out["pt"] = F("partialtrace", out["mm"], 1)
    out["pt0"] = F("partialtrace", out["kron"], 0)
    out["pt_ek"] = F("partialtrace", F("eyekron", B, 1, 0), 2)
    out["rowT"] = F("transpose", F("create", inp["V"]))
    out["row_mm"] = F("matmul", F("create", inp["V"]), B)
    out["outer"] = F("ewmul", F("create", inp["V"]), F("statezero", 2))
    out["outer2"] = F("ewmul", F("statezero", 2), F("create", inp["V"]))
    out["kron_mm"] = F("matmul", F("kron", H, A), F("kron", B, Id))
    out["cu_dag"] = F("dagger", F("addcontrol", F("matmul", A, out["H1"])))
    out["deep"] = F("sub", F("matmul", out["dag"], out["smul"]),
                    F("kron", F("transpose", A), F("scalar_mul", A, 0.5)))
    return out


def snapshot(mod, fm):
    """shape, mem (None where the call raises), every element, and the
    trace (or the text of its error)."""
    r, c = mod.funmatrix_shape(fm, False)
    try:
        mem = mod.funmatrix_mem(fm, False)
    except SystemError:
        mem = None
    vals = np.array([[mod.funmatrix_get(fm, i, j, False) for j in range(c)]
                     for i in range(r)], dtype=np.complex128)
    try:
        t = mod.funmatrix_trace(fm, False)
        tr = [t.real, t.imag]
    except Exception as e:  # noqa: BLE001
        tr = str(e)
    return (r, c), mem, vals, tr


def error_cases(mod):
    """Calls with invalid or edge-case arguments; each returns a value that
    JSON can hold or raises."""
    def F(name, *a):
        return getattr(mod, "funmatrix_" + name)(*a, False)

    def shape(fm):
        return list(F("shape", fm))

    def cx(z):
        return [z.real, z.imag]

    M2 = F("create", [[1, 2], [3, 4]])
    M4 = F("create", [[1, 2, 3, 4]] * 4)
    R = F("create", [[1, 2, 3, 4]])
    # a product whose element (0, 0) fails at k = 1 (NaN factor) after a
    # partial sum: leaves above it carry that partial value on
    N = F("create", [[1, 2], [float("nan"), 4]])
    MN = F("matmul", F("scalar_mul", M2, 1.5 - 0.5j), N)

    def dense(fm):
        r, c = F("shape", fm)
        return [[cx(F("get", fm, i, j)) for j in range(c)] for i in range(r)]

    return {
        "mm_fail": lambda: cx(F("get", MN, 0, 0)),
        "mm_fail_trace": lambda: cx(F("trace", MN)),
        "pt_partial": lambda: cx(F("get", F("partialtrace", MN, 0), 0, 0)),
        "cu_partial": lambda: cx(F("get", F("addcontrol", MN), 2, 2)),
        "cu_partial_ok": lambda: cx(F("get", F("addcontrol", MN), 2, 3)),
        "ek_partial": lambda: cx(F("get", F("eyekron", MN, 1, 1), 0, 0)),
        "pt_oob_values": lambda: dense(F("partialtrace", M4, 2)),
        "pt_oob_T_values": lambda: dense(F("partialtrace",
                                           F("transpose", M4), 3)),
        "pt_oob_trace": lambda: cx(F("trace", F("partialtrace", M4, 2))),
        "pt_nan_first": lambda: cx(F("get", F("partialtrace", N, 0), 0, 0)),
        "kron_nan": lambda: cx(F("get", F("kron", M2, N), 2, 0)),
        "get_row_oob": lambda: cx(F("get", M2, 2, 0)),
        "get_col_oob": lambda: cx(F("get", M2, 0, 2)),
        "get_T_oob": lambda: cx(F("get", F("transpose", R), 3, 0)),
        "get_T_ok": lambda: cx(F("get", F("transpose", R), 0, 3)),
        "get_not_capsule": lambda: cx(F("get", 5, 0, 0)),
        "add_mismatch": lambda: shape(F("add", M2, M4)),
        "sub_mismatch": lambda: shape(F("sub", M2, M4)),
        "matmul_mismatch": lambda: shape(F("matmul", M2, M4)),
        "matmul_row_col": lambda: shape(F("matmul", R, F("statezero", 2))),
        "ewmul_mismatch": lambda: shape(F("ewmul", M2, M4)),
        "trace_row": lambda: cx(F("trace", R)),
        "pt_nonsquare": lambda: shape(F("partialtrace", R, 0)),
        "pt_oob": lambda: shape(F("partialtrace", M4, 2)),
        "pt_2x2": lambda: shape(F("partialtrace", M2, 0)),
        "smul_str": lambda: shape(F("scalar_mul", M2, "a")),
        "sdiv_str": lambda: shape(F("scalar_div", M2, "a")),
        "sdiv_zero": lambda: shape(F("scalar_div", M2, 0)),
        "create_int": lambda: shape(F("create", 5)),
        "create_empty": lambda: shape(F("create", [])),
        "create_bad_elem": lambda: shape(F("create", [[1, "a"], [1, 2]])),
        "create_ragged": lambda: shape(F("create", [[1, 2], [3]])),
        "identity0": lambda: shape(F("identity", 0)),
        "hadamard0": lambda: shape(F("hadamard", 0)),
        "statezero0": lambda: shape(F("statezero", 0)),
        "densityzero0": lambda: shape(F("densityzero", 0)),
        "addcontrol_row": lambda: shape(F("addcontrol", R)),
        "eyekron00": lambda: shape(F("eyekron", M2, 0, 0)),
        "shape_not_capsule": lambda: shape("x"),
        "trace_pt_chain": lambda: cx(F("trace", F("partialtrace",
                                                   F("kron", M2, M2), 1))),
    }


def run_error_cases(mod):
    out = {}
    for name, fn in error_cases(mod).items():
        try:
            out[name] = ["ok", fn()]
        except Exception as e:  # noqa: BLE001
            out[name] = [type(e).__module__ + "." + type(e).__name__, str(e)]
    return out
