"""SHA-256 digests of the reference C backend's outputs on the seeded "medium"
cases of tests/cases.py.  The reference library is compiled from the
reference's own sources by `make -C oracle ref` (into oracle/_ref, never
committed); the digests let the GPU parity tests compare against the
reference's own results on a box where the reference is absent."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = os.path.dirname(HERE)
sys.path.insert(0, os.path.dirname(TESTS))
sys.path.insert(0, TESTS)

DIGEST_SEEDS = (11, 12)


def write_digests(ref_so, out_json):
    import _clib
    import cases
    from qmctoolscl_amd._api import KERNEL_NAMES

    if not os.path.exists(ref_so):
        sys.exit("reference library not built at %s (make -C oracle ref)" % ref_so)
    lib = _clib.load(ref_so)
    table = {}
    for name in KERNEL_NAMES:
        for seed in DIGEST_SEEDS:
            case = cases.make_case(name, seed=seed, size="medium")
            _clib.call_ref_abi(lib, name, case)
            table["%s/%d" % (name, seed)] = {o: _clib.digest(case[o]) for o in cases.outputs_of(name)}
    with open(out_json, "w") as f:
        json.dump({"size": "medium", "seeds": list(DIGEST_SEEDS), "digests": table}, f, indent=1)
    print("wrote %d digests" % len(table))


if __name__ == "__main__":
    write_digests(os.path.join(os.path.dirname(TESTS), "oracle", "_ref", "libqmcref.so"),
                  os.path.join(HERE, "ref_digests.json"))
