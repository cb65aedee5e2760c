"""SHA-256 digests of the reference C backend's halton_qrng outputs on the
seeded cases of tests/test_halton.py (reference compiled from its own sources
by `make -C oracle ref` into oracle/_ref, never committed).  The digests let
the tests compare against the reference's results where the reference is
absent.

Usage:  python tests/golden/halton_ref.py"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = os.path.dirname(HERE)
sys.path.insert(0, os.path.dirname(TESTS))
sys.path.insert(0, TESTS)


def main():
    import _clib
    import test_halton as th
    ref = _clib.reference()
    if ref is None:
        sys.exit("reference library not built (make -C oracle ref)")
    table = {}
    for case in th.CASES:
        randu, dvec = th.inputs(*case)
        table["/".join(str(v) for v in case)] = _clib.digest(th.call(ref, "halton_qrng", *case[:4], randu, dvec))
    with open(os.path.join(HERE, "halton_ref.json"), "w") as f:
        json.dump({"cases": "n/d/n0/generalized/seed (inputs as tests/test_halton.py:inputs)",
                   "digests": table}, f, indent=1)
    print("wrote %d digests" % len(table))


if __name__ == "__main__":
    main()
