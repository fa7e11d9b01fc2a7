"""Mint tests/golden/golden.json from the reference's own functions.

Runs oracle/_ref/refdrv (built from the reference sources by
oracle/build_ref.sh) on every case of cases.py and records, per case:
P (scalarMultiplicationCount), nnz(C), the SHA-256 digest of C from the
reference's spgemm() at -np 1 (checked equal to serialSpgemm's), the digest of
buildPartialReference's row counts and its finite flag; for tiny products the
full arrays (values as IEEE bit patterns).  Also the digest of
readMatrixMarket's CSR for every data file, and of trident::LocalAccumulator's
result for each staged case.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import cases  # noqa: E402
from oracle import csr_digest, read_csr, ref_available, ref_run, write_csr  # noqa: E402


def counts_digest(counts: np.ndarray) -> str:
    return hashlib.sha256(np.asarray(counts, dtype="<i8").tobytes()).hexdigest()


def main():
    if not ref_available():
        sys.exit("oracle/_ref/refdrv missing: run oracle/build_ref.sh first")
    tmp = tempfile.mkdtemp(dir=HERE, prefix="_tmp")
    try:
        out = {"products": {}, "mm_read": {}, "accumulate": {}}
        all_cases = [(n, a, b) for n, a, b in cases.mm_cases()] + cases.synthetic_cases()
        all_cases += [(f"smoke_{s}",) + cases.smoke_case(s) for s in range(100)]
        all_cases += [(f"dyadic_{i}", a, b) for i, (a, b) in enumerate(cases.dyadic_cases())]
        for name, a, b in all_cases:
            pa, pb = os.path.join(tmp, "a.csr"), os.path.join(tmp, "b.csr")
            write_csr(pa, a)
            write_csr(pb, b)
            pc, ps, pp = (os.path.join(tmp, x) for x in ("c.csr", "s.csr", "p.bin"))
            ref_run("spgemm", pa, pb, pc)
            ref_run("serial", pa, pb, ps)
            ref_run("partial", pa, pb, pp)
            c, s = read_csr(pc), read_csr(ps)
            assert csr_digest(c) == csr_digest(s), name
            raw = open(pp, "rb").read()
            finite = int(np.frombuffer(raw[:4], dtype="<i4")[0])
            counts = np.frombuffer(raw[4:], dtype="<i8")
            rec = {"rows": a.rows, "cols": b.cols, "products": int(ref_run("count", pa, pb).strip()),
                   "c_nnz": c.nnz, "digest": csr_digest(c), "partial_digest": counts_digest(counts),
                   "partial_finite": finite, "partial_total": int(counts.sum())}
            if c.nnz <= 64 and a.rows <= 64:
                rec["row_ptr"] = c.row_ptr.tolist()
                rec["col"] = c.col.tolist()
                rec["val_bits"] = [f"{x:016x}" for x in c.val.view("<u8").tolist()]
            out["products"][name] = rec
        for root, _, files in os.walk(cases.DATA):
            for f in sorted(files):
                if not f.endswith(".mtx"):
                    continue
                path = os.path.join(root, f)
                rel = os.path.relpath(path, cases.DATA)
                pm = os.path.join(tmp, "m.csr")
                try:
                    ref_run("read", path, pm)
                    m = read_csr(pm)
                    out["mm_read"][rel] = {"ok": True, "digest": csr_digest(m), "rows": m.rows, "cols": m.cols,
                                           "nnz": m.nnz}
                except Exception as e:  # the reference rejects the file
                    msg = str(e).split("refdrv: ", 1)[-1].split(": /", 1)[0]
                    out["mm_read"][rel] = {"ok": False, "error": msg[:200]}
        for name, rows, cols, stages in cases.accumulate_cases():
            args = []
            for i, (a, b) in enumerate(stages):
                pa, pb = os.path.join(tmp, f"sa{i}.csr"), os.path.join(tmp, f"sb{i}.csr")
                write_csr(pa, a)
                write_csr(pb, b)
                args += [pa, pb]
            pc = os.path.join(tmp, "acc.csr")
            ref_run("accumulate", pc, rows, cols, *args)
            c = read_csr(pc)
            out["accumulate"][name] = {"c_nnz": c.nnz, "digest": csr_digest(c)}
        with open(os.path.join(HERE, "golden.json"), "w") as f:
            json.dump(out, f, indent=1, sort_keys=True)
        print(f"golden.json: {len(out['products'])} products, {len(out['mm_read'])} files, "
              f"{len(out['accumulate'])} staged products")
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    main()
