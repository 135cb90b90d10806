"""Write golden vectors for the expression analysis and reshape.

Runs oracle/_ref/ref_probe, which oracle/Makefile compiles from the
reference's own headers (src/Basics/ContractInfo.h, Utils.h), over the
expressions of test/test_kernel.cpp:306-356, the contraction configurations,
permutation cases and every error branch of ContractInfo.h:79-328.  Output:
tests/golden/contract_info.json and tests/golden/reshape.json.

Run from the repository root after `make -C oracle`.
"""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
PROBE = os.path.join(ROOT, "oracle", "_ref", "ref_probe")
OUT = os.path.join(ROOT, "tests", "golden")

# (expr, sizes A, sizes B, sizes C)
CASES = [
    # test/test_kernel.cpp:306-336
    ("ik, kj -> ij", [8, 5], [5, 6], [8, 6]),
    ("jik, kmj -> im", [6, 5, 8], [8, 4, 6], [5, 4]),
    ("ikl, lkmj -> mij", [6, 5, 7], [7, 5, 6, 6], [6, 6, 6]),
    ("ijml, lmk -> jki", [5, 6, 8, 9], [9, 8, 7], [6, 7, 5]),
    ("ij, jk -> ik", [2, 2], [2, 2], [2, 2]),
    ("ij, jk -> ik", [5, 5], [5, 5], [5, 5]),
    ("ij, jk -> ik", [10, 10], [10, 10], [10, 10]),
    ("ij, jk -> ki", [4, 4], [4, 4], [4, 4]),
    # test/test_kernel.cpp:338-356
    ("ikl, ljk -> ij", [8, 8, 8], [8, 8, 8], [8, 8]),
    ("kil, ljk -> ij", [8, 8, 8], [8, 8, 8], [8, 8]),
    ("ikl, ljk -> ji", [5, 5, 5], [5, 5, 5], [5, 5]),
    ("ik, kj -> ij", [16, 16], [16, 16], [16, 16]),
    # contraction configurations (block level and element level)
    ("ijk,kl->ijl", [16, 16, 16], [16, 16], [16, 16, 16]),
    ("ijkl,klm->ijm", [5, 7, 9, 11], [9, 11, 13], [5, 7, 13]),
    ("ijkl,klm->ijm", [44, 43, 45, 44], [45, 44, 46], [44, 43, 46]),
    ("ik,kj->ij", [32, 32], [32, 32], [32, 32]),
    ("ijk,km->ijm", [7, 9, 11], [11, 13], [7, 9, 13]),
    # permutations (transposed / interleaved operands)
    ("ki,kj->ij", [3, 4], [3, 5], [4, 5]),
    ("ik,jk->ij", [4, 3], [5, 3], [4, 5]),
    ("ik,kj->ji", [4, 3], [3, 5], [5, 4]),
    ("ik,kjl->ilj", [2, 3], [3, 4, 5], [2, 5, 4]),
    ("akbl,lcka->bc", [2, 3, 4, 5], [5, 6, 3, 2], [4, 6]),
    ("a_b,b_c->a_c", [2, 3, 4], [4, 3, 5], [2, 3, 5]),
    ("pqrs,srqt->tp", [2, 3, 4, 5], [5, 4, 3, 6], [6, 2]),
    # error branches, ContractInfo.h:79-328
    ("", [2, 2], [2, 2], [2, 2]),
    ("ij,jk->i#", [2, 2], [2, 2], [2, 2]),
    (",jk->ik", [2, 2], [2, 2], [2, 2]),
    ("ij", [2, 2], [2, 2], [2, 2]),
    ("ij jk->ik", [2, 2], [2, 2], [2, 2]),
    ("ij,", [2, 2], [2, 2], [2, 2]),
    ("ij,jk", [2, 2], [2, 2], [2, 2]),
    ("ij,jk,ik", [2, 2], [2, 2], [2, 2]),
    ("ij,jk->", [2, 2], [2, 2], [2, 2]),
    ("ij,jk->ik,", [2, 2], [2, 2], [2, 2]),
    ("ii,jk->ik", [2, 2], [2, 2], [2, 2]),
    ("ij,jj->ij", [2, 2], [2, 2], [2, 2]),
    ("ij,jk->iki", [2, 2], [2, 2], [2, 2, 2]),
    ("ij,ij->ij", [2, 2], [2, 2], [2, 2]),
    ("ij,kl->ijkl", [2, 2], [2, 2], [2, 2, 2, 2]),
    ("ij,jk->i", [2, 2], [2, 2], [2]),
    ("ij,jk->ikl", [2, 2], [2, 2], [2, 2, 2]),
    ("ij,jk->ik", [2, 3], [2, 4], [2, 4]),
    ("ij,jk->ik", [2, 3], [3, 4], [5, 4]),
    ("ij,jk->ik", [2, 3], [3, 4], [2, 5]),
    ("ij -> ik", [2, 2], [2, 2], [2, 2]),
    ("-ij,jk->ik", [2, 2], [2, 2], [2, 2]),
    # the reference accepts an output index absent from A and B (count check
    # only, ContractInfo.h:257); the build rejects it (documented divergence)
    ("ij,jk->il", [4, 3], [3, 4], [4, 4]),
]

RESHAPES = [
    ([2, 3, 4], [2, 0, 1]),
    ([2, 3, 4], [0, 1, 2]),
    ([2, 3, 4], [1, 2, 0]),
    ([5, 6, 8, 9], [1, 3, 0, 2]),
    ([3, 4], [1, 0]),
    ([2, 3, 4, 5], [3, 2, 1, 0]),
]


def run(lines):
    p = subprocess.run([PROBE], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True)
    return [json.loads(l) for l in p.stdout.splitlines() if l.strip()]


def main():
    if not os.path.exists(PROBE):
        sys.exit("oracle/_ref/ref_probe missing: run `make -C oracle` with the reference present")
    lines = ["I|%s|%s|%s|%s" % (e, ",".join(map(str, a)), ",".join(map(str, b)), ",".join(map(str, c)))
             for e, a, b, c in CASES]
    recs = run(lines)
    assert len(recs) == len(CASES), (len(recs), len(CASES))
    with open(os.path.join(OUT, "contract_info.json"), "w") as f:
        json.dump({"source": "oracle/_ref/ref_probe (reference src/Basics/ContractInfo.h)", "cases": recs}, f,
                  indent=0)
    rrecs = run(["R|%s|%s" % (",".join(map(str, s)), ",".join(map(str, o))) for s, o in RESHAPES])
    with open(os.path.join(OUT, "reshape.json"), "w") as f:
        json.dump({"source": "oracle/_ref/ref_probe (reference src/Basics/Utils.h reshape)", "cases": rrecs}, f,
                  indent=0)
    print("wrote %d contract-info and %d reshape vectors" % (len(recs), len(rrecs)))


if __name__ == "__main__":
    main()
