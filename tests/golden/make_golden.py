"""Writes the golden fixtures of tests/golden/ (run in the build container).

Inputs come from the reference's own workload generator
(test/include/data_generator.hpp of the reference, compiled in place by
`make -C oracle ref` into oracle/_ref/ref_datagen), reproducing the set-up of
the reference driver test/test.cu:92-132 (u32 keys and values, cardinality =
n * pct / 100, distribution 0 uniform / 1 zipf(0.9) / 2 self-similar(0.9)).
Expected outputs: the reference's only printed observable, the group count,
which must equal the cardinality (test/test.cu:111,163-168), plus the sorted
(key, COUNT, SUM mod 2^32) rows from the exact oracle.
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle as O  # noqa: E402

CASES = [(4096, 10, 0), (4096, 10, 1), (4096, 10, 2), (20000, 5, 0), (20000, 50, 1)]


def main():
    O.build()
    for n, pct, dist in CASES:
        with tempfile.TemporaryDirectory() as td:
            keys, vals = O.ref_datagen(n, pct, dist, os.path.join(td, "kv.bin"))
        r = O.groupby(keys, [vals], ["count", ("sum", 0)])
        out = os.path.join(HERE, f"refgen_n{n}_c{pct}_d{dist}.npz")
        np.savez_compressed(out, keys=keys, vals=vals, cardinality=np.int64(n * pct // 100),
                            out_keys=r["keys"], out_count=r["aggs"][0],
                            out_sum_u32=(r["aggs"][1].astype(np.uint64) & np.uint64(0xFFFFFFFF)).astype(np.uint32))
        print(out, r["n_groups"])


if __name__ == "__main__":
    main()
