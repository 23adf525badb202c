"""Writes the golden fixtures of tests/golden/ (run in the build container).

Inputs come from the reference's own workload generator
(test/include/data_generator.hpp of the reference, compiled in place by
`make -C oracle ref` into oracle/_ref/ref_datagen), reproducing the set-up of
the reference driver test/test.cu:92-132 (u32 keys and values, cardinality =
n * pct / 100, distribution 0 uniform / 1 zipf(0.9) / 2 self-similar(0.9)).
Expected outputs: the reference's only printed observable, the group count,
which must equal the cardinality (test/test.cu:111,163-168), plus the sorted
(key, COUNT, SUM mod 2^32) rows, computed here by a numpy sort and segmented
reduce -- independent of the oracle, which the fixtures then check.
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle as O  # noqa: E402

CASES = [(4096, 10, 0), (4096, 10, 1), (4096, 10, 2), (20000, 5, 0), (20000, 50, 1)]


def brute_force(keys, vals):
    """Sorted distinct keys, their row counts and the u32 wrap-around sums of
    their values (the reference's atomicAdd on u32, kernel.cu:48,152,334,340)."""
    o = np.argsort(keys, kind="stable")
    k = keys[o]
    v = vals[o].astype(np.uint64)
    starts = np.flatnonzero(np.r_[True, k[1:] != k[:-1]])
    counts = np.diff(np.r_[starts, k.size]).astype(np.int64)
    sums = (np.add.reduceat(v, starts) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    return k[starts], counts, sums


def main():
    O.build()
    for n, pct, dist in CASES:
        with tempfile.TemporaryDirectory() as td:
            keys, vals = O.ref_datagen(n, pct, dist, os.path.join(td, "kv.bin"))
        ok, oc, osum = brute_force(keys, vals)
        out = os.path.join(HERE, f"refgen_n{n}_c{pct}_d{dist}.npz")
        np.savez_compressed(out, keys=keys, vals=vals, cardinality=np.int64(n * pct // 100),
                            out_keys=ok, out_count=oc, out_sum_u32=osum)
        print(out, ok.size)


if __name__ == "__main__":
    main()
