"""Write the golden fixtures of tests/golden/ from the reference itself.

    python tests/golden/make_golden.py

Runs the reference library (oracle/_ref/libxss_ref.so, built by
`make -C oracle ref` from the reference's own sources) on seeded inputs and
stores inputs and outputs as small compressed .npz files (no pickles):

  golden_sort.npz    qsort asc/desc (hasnan where the input has NaNs) and
                     argsort: input, sorted output, gathered keys of argsort
  golden_select.npz  qselect / partial_qsort / argselect at several k
  golden_kv.npz      keyvalue_qsort / keyvalue_select / keyvalue_partial_sort

Keys are compared by the rules of the reference's own tests (sorted keys
bit-exact, +-0 and NaN runs by value, payloads per equal-key run), so the
fixtures pin the parts of the output the reference determines.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import xss_testlib as T  # noqa: E402

SIZES = [1, 2, 7, 64, 300]
DISTS = ["random", "dups", "reverse", "smallrange", "nan", "rand_with_max_and_nan"]
KV = [("u32", "u32"), ("f32", "u64"), ("u64", "u64"), ("i32", "f64"), ("f64", "i64"), ("i64", "u32")]


def main():
    R = T.REF()
    if R is None:
        raise SystemExit("oracle/_ref/libxss_ref.so is missing: run `make -C oracle ref`")
    sort, sel, kv = {}, {}, {}
    seed = 0
    for tag in T.TAGS:
        for dist in DISTS:
            for n in SIZES:
                seed += 1
                a = T.gen(tag, dist, n, 90000 + seed)
                hasnan = dist in T.NAN_DISTS
                key = f"{tag}|{dist}|{n}|{int(hasnan)}"
                sort[key + "|in"] = a
                for desc in (0, 1):
                    sort[key + f"|qsort{desc}"] = R.qsort(a, tag, hasnan, desc)
                    idx = R.argsort(a, tag, hasnan, desc)
                    sort[key + f"|argkeys{desc}"] = a[idx.astype(np.int64)]
                if n >= 2:
                    for k in sorted({n // 2, n - 1}):
                        for desc in (0, 1):
                            sel[key + f"|qselect{k}|{desc}"] = R.qselect(a, tag, k, hasnan, desc)
                            sel[key + f"|partial{k + 1}|{desc}"] = R.partial_qsort(a, tag, k + 1, hasnan, desc)[: k + 1]
                        idx = R.argselect(a, tag, k, hasnan)
                        sel[key + f"|argselect{k}"] = a[idx.astype(np.int64)]
                    sel[key + "|in"] = a
    for ktag, vtag in KV:
        for dist in ["random", "dups", "nan_payloads", "signed_zeros"]:
            for n in [2, 64, 300]:
                seed += 1
                k = T.gen(ktag, dist, n, 95000 + seed)
                v = np.arange(n).astype(T.TAGS[vtag])
                hasnan = dist in T.NAN_DISTS
                key = f"{ktag}|{vtag}|{dist}|{n}|{int(hasnan)}"
                kv[key + "|k"], kv[key + "|v"] = k, v
                for desc in (0, 1):
                    ok, ov = R.keyvalue_qsort(k, v, ktag, vtag, hasnan, desc)
                    kv[key + f"|qsort{desc}|k"], kv[key + f"|qsort{desc}|v"] = ok, ov
                    kk = n // 2
                    ok, ov = R.keyvalue_select(k, v, ktag, vtag, kk, hasnan, desc)
                    kv[key + f"|select{kk}|{desc}|k"], kv[key + f"|select{kk}|{desc}|v"] = ok, ov
                    ok, ov = R.keyvalue_partial_sort(k, v, ktag, vtag, kk + 1, hasnan, desc)
                    kv[key + f"|partial{kk + 1}|{desc}|k"] = ok[: kk + 1]
    np.savez_compressed(os.path.join(HERE, "golden_sort.npz"), **sort)
    np.savez_compressed(os.path.join(HERE, "golden_select.npz"), **sel)
    np.savez_compressed(os.path.join(HERE, "golden_kv.npz"), **kv)
    print("isa:", T.ref().xssref_isa().decode(), "| sort", len(sort), "select", len(sel), "kv", len(kv))


if __name__ == "__main__":
    main()
