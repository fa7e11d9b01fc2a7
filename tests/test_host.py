"""Host-side checks that need no GPU: the C ABI library loads and exports every
declared entry point, the engine's MatrixMarket reader matches the reference's
readMatrixMarket (src/common/matrix_market.cpp:28-144), and the workload
generators produce what they promise."""
import os
import re
import subprocess

import numpy as np
import pytest

import cases
import oracle as O
import spgemm_amd as S
import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "spgemm_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(spg_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = S.lib()
    names = declared_symbols()
    assert len(names) >= 18
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(S.EXPORTED) <= set(names)
    nm = subprocess.run(["nm", "-D", "--defined-only", S.LIB_PATH], capture_output=True, text=True).stdout
    for n in names:
        assert re.search(rf"\bT {n}$", nm, re.M), n


def test_status_strings():
    lib = S.lib()
    for code in range(11):
        assert lib.spg_status_string(code)


def test_create_without_device_reports_enodevice():
    import ctypes as C
    h = C.c_void_p()
    s = S.lib().spg_create(C.byref(h), 0)
    if s == 0:
        S.lib().spg_destroy(h)
        pytest.skip("a GPU is present")
    assert s == 7  # SPG_ENODEVICE


def test_server_and_batch_need_a_device():
    import ctypes as C
    h = C.c_void_p()
    if S.lib().spg_create(C.byref(h), 0) == 0:
        S.lib().spg_destroy(h)
        pytest.skip("a GPU is present")
    with pytest.raises(S.SpgError) as e:
        S.Server(0, budget_bytes=1 << 30)
    assert e.value.name == "SPG_ENODEVICE"


def test_cpp_shim_builds_and_fails_loudly_without_device(tmp_path):
    """The C++17 shim program links against the library; with no GPU the
    Engine constructor throws (no silent CPU path)."""
    import ctypes as C
    shim = os.path.join(os.path.dirname(S.LIB_PATH), "shim_test")
    assert os.path.exists(shim)
    h = C.c_void_p()
    if S.lib().spg_create(C.byref(h), 0) == 0:
        S.lib().spg_destroy(h)
        pytest.skip("a GPU is present")
    a = cases.data_path("cage8.mtx")
    r = subprocess.run([shim, a, a, str(tmp_path / "c.csr")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "device" in r.stderr.lower(), r.stderr
    assert not (tmp_path / "c.csr").exists()


def test_mm_reader_matches_reference(golden):
    for rel, g in golden["mm_read"].items():
        path = cases.data_path(rel)
        if g["ok"]:
            m = S.read_mm(path)
            assert O.csr_digest(O.Csr(m.rows, m.cols, m.row_ptr, m.col, m.val)) == g["digest"], rel
        else:
            with pytest.raises(S.SpgError) as e:
                S.read_mm(path)
            assert e.value.name == "SPG_EPARSE"
            assert g["error"] in str(e.value), rel


MM_EDGE = {
    "d_exponent": "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1.5D2\n2 2 -2.5d-3\n",
    "pattern_sym": "%%MatrixMarket matrix coordinate pattern symmetric\n3 3 3\n1 1\n3 1\n2 3\n",
    "skew": "%%MatrixMarket matrix coordinate real skew-symmetric\n3 3 2\n2 1 4.0\n3 2 -1.0\n",
    "hermitian": "%%MatrixMarket matrix coordinate real hermitian\n2 2 2\n1 1 3\n2 1 5\n",
    "comments_blank": "%%MatrixMarket matrix coordinate integer general\n% c\n\n  \n2 3 3\n% mid\n1 3 7\n\n2 1 -4\n1 1 0\n% end\n",
    "dups_zeros_order": "%%MatrixMarket matrix coordinate real general\n2 4 5\n1 4 1\n1 2 2\n1 4 0\n2 1 3\n1 2 -2\n",
    "upper_banner": "%%MatrixMarket MATRIX Coordinate REAL General\n1 1 1\n1 1 2\n",
    "bad_extra_token": "%%MatrixMarket matrix coordinate real general\n1 1 1\n1 1 2 3\n",
    "bad_too_many": "%%MatrixMarket matrix coordinate real general\n1 1 1\n1 1 2\n1 1 3\n",
    "bad_too_few": "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 2\n",
    "bad_index": "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 2\n",
    "bad_skew_diag": "%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n1 1 2\n",
    "bad_array": "%%MatrixMarket matrix array real general\n1 1\n2\n",
    "bad_complex": "%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 2 0\n",
    "bad_dims": "%%MatrixMarket matrix coordinate real general\n0 2 0\n",
    "bad_banner_extra": "%%MatrixMarket matrix coordinate real general extra\n1 1 0\n",
    "bad_empty": "",
}


@pytest.mark.parametrize("name", sorted(MM_EDGE))
def test_mm_reader_edge_cases_agree(name, tmp_path):
    path = str(tmp_path / f"{name}.mtx")
    with open(path, "w") as f:
        f.write(MM_EDGE[name])
    try:
        want = O.mm_read(path)
        want_err = None
    except O.OracleError as e:
        want, want_err = None, str(e)
    if want_err is None:
        got = S.read_mm(path)
        assert O.csr_digest(O.Csr(got.rows, got.cols, got.row_ptr, got.col, got.val)) == O.csr_digest(want)
    else:
        with pytest.raises(S.SpgError):
            S.read_mm(path)
    if O.ref_available():  # the reference binary, where it was built
        out = str(tmp_path / "ref.csr")
        try:
            O.ref_run("read", path, out)
            assert want_err is None and O.csr_digest(O.read_csr(out)) == O.csr_digest(want), name
        except O.OracleError:
            assert want_err is not None, name


def test_mm_reader_keeps_file_order():
    # tests/matrix_market_contract.cpp: order, duplicates and explicit zeros kept
    m = S.read_mm(cases.data_path("tiny_symmetric.mtx"))
    assert m.row_ptr.tolist() == [0, 2, 4, 6]
    assert m.col.tolist() == [0, 1, 0, 2, 1, 2]
    assert m.val.tolist() == [2.0, 3.0, 3.0, 4.0, 4.0, 5.0]


def test_uniform_generator():
    m = W.uniform_random(1000, 5000, 16, seed=3)
    assert m.nnz == 16000
    c = m.col.reshape(1000, 16)
    assert np.all(np.diff(c, axis=1) > 0) and c.min() >= 0 and c.max() < 5000
    assert np.all(m.val >= -1) and np.all(m.val < 1)
    assert O.csr_digest(O.Csr(m.rows, m.cols, m.row_ptr, m.col, m.val)) == \
        O.csr_digest(O.Csr(*(lambda x: (x.rows, x.cols, x.row_ptr, x.col, x.val))(W.uniform_random(1000, 5000, 16, seed=3))))


def test_banded_generator_matches_case_formula():
    a = W.banded(300, 250, 8)
    b = cases.banded(300, 250, 8)
    assert np.array_equal(a.col, b.col) and np.array_equal(a.val, b.val)


def test_stencil_and_aggregation():
    n = 6
    A = W.stencil27(n)
    assert A.rows == n ** 3 and A.nnz == (3 * n - 2) ** 3
    rs = np.add.reduceat(A.val, A.row_ptr[:-1])
    interior = np.zeros((n, n, n), bool)
    interior[1:-1, 1:-1, 1:-1] = True
    assert np.all(rs[interior.reshape(-1)] == 0)
    P = W.aggregation(n)
    assert P.cols == 27 and P.nnz == n ** 3
    R = W.transpose(P)
    assert R.rows == 27 and np.all(np.diff(R.row_ptr) == 8)


def test_products_and_bytes_model():
    a = W.uniform_random(100, 200, 4, seed=1)
    b = W.uniform_random(200, 300, 7, seed=2)
    assert W.products(a, b) == 100 * 4 * 7
    assert W.row_products(a, b).sum() == 2800
    assert W.streaming_bytes(a, 200, 2800, 10) == 8 * 101 + 12 * 400 + 16 * 400 + 12 * 2800 + 8 * 101 + 120


def _big_mm(path, symmetric, bad_at=None, extra=0, missing=0, seed=5):
    """A MatrixMarket file large enough for many reader threads: comments and
    blank lines between entries, D exponents, '+' signs and duplicates."""
    rng = np.random.default_rng(seed)
    n, ents = 4000, 400_000
    r = rng.integers(1, n + 1, ents)
    c = rng.integers(1, n + 1, ents)
    if symmetric:
        r, c = np.maximum(r, c), np.minimum(r, c)
    v = rng.standard_normal(ents)
    kind = "symmetric" if symmetric else "general"
    lines = [f"%%MatrixMarket matrix coordinate real {kind}", "% generated", f"{n} {n} {ents - missing}"]
    for i in range(ents - missing):
        if i % 9973 == 0:
            lines.append("% a comment between entries")
        if i % 7919 == 0:
            lines.append("   ")
        val = f"{v[i]:.17g}"
        if i % 5 == 0:
            val = val.replace("e", "D")
        if i % 11 == 0 and not val.startswith("-"):
            val = "+" + val
        lines.append(f"{r[i]} {c[i]} {val}" if i != bad_at else f"{r[i]} {c[i]} {val} 7")
    for i in range(extra):  # undeclared
        lines.append(f"1 1 {i}.5")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


@pytest.mark.parametrize("symmetric", [False, True])
def test_parallel_mm_reader_matches_oracle(tmp_path, symmetric):
    path = str(tmp_path / "big.mtx")
    _big_mm(path, symmetric)
    assert os.path.getsize(path) > 9 << 20  # several reader chunks
    got = S.read_mm(path)
    want = O.mm_read(path)
    assert O.csr_digest(O.Csr(got.rows, got.cols, got.row_ptr, got.col, got.val)) == O.csr_digest(want)
    # binary CSR round trip
    bpath = str(tmp_path / "big.csr")
    S.write_csr(got, bpath)
    back = S.read_csr(bpath)
    assert back.row_ptr.tolist() == got.row_ptr.tolist() and back.col.tolist() == got.col.tolist()
    assert back.val.view(np.uint64).tolist() == got.val.view(np.uint64).tolist()


def test_parallel_mm_reader_errors_in_late_chunks(tmp_path):
    cases_ = [("late_bad", dict(bad_at=390_000), "invalid Matrix Market entry"),
              ("extra", dict(extra=3), "more Matrix Market entries than declared"),
              ("missing", dict(missing=2), "unexpected end of Matrix Market entries")]
    for name, kw, msg in cases_:
        path = str(tmp_path / f"{name}.mtx")
        _big_mm(path, False, **kw)
        if name == "missing":  # declare two more than present
            txt = open(path).read().replace("\n4000 4000 399998\n", "\n4000 4000 400000\n", 1)
            open(path, "w").write(txt)
        with pytest.raises(S.SpgError) as e:
            S.read_mm(path)
        assert msg in str(e.value), name
        with pytest.raises(O.OracleError) as e2:
            O.mm_read(path)
    bad = str(tmp_path / "trunc.csr")
    good = S.Csr.make(2, 2, [0, 1, 2], [0, 1], [1.0, 2.0])
    S.write_csr(good, bad)
    raw = open(bad, "rb").read()
    open(bad, "wb").write(raw[:-3])
    with pytest.raises(S.SpgError):
        S.read_csr(bad)
