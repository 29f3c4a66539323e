"""Writes tests/golden/kat.json: known-answer vectors for the m-discriminator.

Inputs come from the reference's own spec tests (haskell_code/tests.hs:5 and
haskell_code/testsSegm.hs:4,7 of asbjornvg/mdiscriminator) plus seeded
workloads.  Expected outputs are produced by the reference's own SeqMdiscr
(cuda_code/SeqMdiscr.h:10-57, compiled from source into
oracle/_ref/libseqref.so by oracle/Makefile); the segmented expectation applies
it per segment, as MdiscrSegmTest.cu:21-38 validates.  The small vectors also
agree with hand derivation from the spec (haskell_code/Mdiscr.hs:9-54).

    make -C oracle && python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from tests.helpers import fnv_fast  # noqa: E402
from mdiscriminator_amd import workloads as W  # noqa: E402


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libseqref.so not built (needs the reference sources)")
    kat = {"source": "reference SeqMdiscr (oracle/_ref) on reference spec inputs", "plain": [],
           "segmented": [], "seeded": []}
    arr = [5, 4, 2, 3, 7, 8, 10, 4, 1, 9, 11, 12, 6]  # haskell_code/tests.hs:5
    for m in (5, 4, 3, 2, 1):                          # haskell_code/tests.hs:26-34
        out, sizes = oracle.ref_seq_mdiscr(np.array(arr, dtype=np.uint32), m)
        kat["plain"].append({"input": arr, "m": m, "out": out.tolist(), "sizes": sizes.tolist()})
    # Trailing empty classes: the reference would also write sizes[n].
    tr = [0, 4, 8, 1, 5, 9]
    out, sizes = oracle.ref_seq_mdiscr(np.array(tr, dtype=np.uint32), 4)
    kat["plain"].append({"input": tr, "m": 4, "out": out.tolist(), "sizes": sizes.tolist()})

    sarr = [5, 4, 2, 3, 7, 8, 6, 4, 1, 9, 11, 12, 10]  # haskell_code/testsSegm.hs:4
    ssz = [4, 0, 0, 0, 2, 0, 1, 3, 0, 0, 3, 0, 0]      # haskell_code/testsSegm.hs:7
    sout, snew = [], []
    i = 0
    while i < len(sarr):
        L = ssz[i]
        o, s = oracle.ref_seq_mdiscr(np.array(sarr[i:i + L], dtype=np.uint32), 4)
        sout += o.tolist()
        snew += s.tolist()
        i += L
    kat["segmented"].append({"input": sarr, "segment_sizes": ssz, "m": 4, "out": sout,
                             "new_sizes": snew})

    for name, n, m, seed, spec in [
        ("uniform 2^20 m=4 (mod)", 1 << 20, 4, W.SEED_BASE + 1, True),
        ("uniform 2^20 m=8 (mod)", 1 << 20, 8, W.SEED_BASE + 2, True),
        ("uniform 2^20 m=7", 1 << 20, 7, W.SEED_BASE + 7, False),
        ("uniform 2^20 m=64", 1 << 20, 64, W.SEED_BASE + 3, False),
        ("uniform 100003 m=13", 100003, 13, W.SEED_BASE + 13, False),
    ]:
        x = W.uniform_u31(n, seed)
        out, sizes = oracle.ref_seq_mdiscr(x, m, specialized=spec)
        counts = np.bincount(x % m, minlength=m).astype(np.uint32)
        kat["seeded"].append({
            "name": name, "n": n, "m": m, "seed": seed, "generator": "uniform_u31",
            "input_head": x[:8].tolist(), "out_head": out[:16].tolist(),
            "out_tail": out[-16:].tolist(), "counts": counts.tolist(),
            "out_digest": fnv_fast(out), "sizes_digest": fnv_fast(sizes)})
    with open(os.path.join(HERE, "kat.json"), "w") as f:
        json.dump(kat, f, separators=(",", ":"))
        f.write("\n")
    print("wrote", os.path.join(HERE, "kat.json"))


if __name__ == "__main__":
    main()
