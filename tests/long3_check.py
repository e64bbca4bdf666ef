"""Three-factor long-line steps against the FP64 oracle, in full.

Run by tests/test_gpu_long_lines.py in a child process with
MDFT_LARGE_SPLIT3_LOGM=15, so lines of padded length 2^15 and above take the
three-factor path (A1, A2, C; a Bluestein line twice, through the product)
that otherwise serves only padded lengths 2^29 and 2^30, at sizes the oracle
checks element by element.  Prints one line per case; exits non-zero on the
first mismatch.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from tests.helpers import check_against_fp64, rand_c64  # noqa: E402

CASES = [(32768, 2), (2, 32768), (16385, 2), (3, 20000), (40000, 1), (65536, 1), (9000, 5), (1, 1 << 20),
         (300001, 1)]


def main():
    assert os.environ.get("MDFT_LARGE_SPLIT3_LOGM") == "15", "run with MDFT_LARGE_SPLIT3_LOGM=15"
    from math_amd import mdft as M

    M.init(0)
    for (w, h) in CASES:
        rng = np.random.default_rng(w * 3 + h)
        x = rand_c64(rng, (h, w))
        p = M.Plan(w, h)
        for inv in (False, True):
            d = np.ascontiguousarray(x.copy())
            p.exec(inv, d)
            m = check_against_fp64(d, x, inv, "three-factor")
            print(f"three-factor {w}x{h} inv={inv}: nl2 {m['nl2']:.2e}", flush=True)
        p.close()
    print("ok", flush=True)


if __name__ == "__main__":
    main()
