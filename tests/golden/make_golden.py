"""Write the golden fixtures the tests check against.

Run in a container where the reference tree is mounted (default
/root/reference).  Output is data only, never reference source text:

* doctest_vectors.json -- every array the reference's doctest suite
  (docs/index.md) displays as the value of a bare variable, in document
  order, as {"line", "name", "dtype", "shape", "values"}; plus the bit rows
  that dnb2_get_linear_scramble_matrix prints (docs/index.md:278-336).
  The doctests' inputs are either literal arrays (restated in the tests) or
  draws from numpy.random.Generator(SFC64(7)) in document order.

* ref_digests.json -- SHA-256 digests of the reference C backend's outputs
  (compiled from its sources into oracle/_ref by oracle/Makefile) on seeded,
  moderately sized inputs, so the GPU box, where the reference is absent, can
  still compare against the reference's own results.

Usage:  python tests/golden/make_golden.py [--ref /root/reference]
"""
import argparse
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))


def _parse_array_text(text):
    text = text.replace("<BLANKLINE>", "")
    ns = {"array": np.array, "uint64": np.uint64, "float64": np.float64}
    arr = eval(text, {"__builtins__": {}}, ns)  # text is a numpy repr: array([...], dtype=...)
    return np.asarray(arr)


def extract_doctest_arrays(index_md):
    lines = open(index_md).read().splitlines()
    out = []
    i = 0
    bare = re.compile(r"^>>> ([A-Za-z_][A-Za-z_0-9]*)\s*$")
    while i < len(lines):
        m = bare.match(lines[i])
        if not m:
            i += 1
            continue
        name, lineno = m.group(1), i + 1
        j = i + 1
        body = []
        while j < len(lines) and not lines[j].startswith(">>>") and not lines[j].startswith("```"):
            body.append(lines[j])
            j += 1
        text = "\n".join(body).strip()
        if text.startswith("array("):
            arr = _parse_array_text(text)
            dtype = "uint64" if "dtype=uint64" in text else "float64"
            arr = arr.astype(np.uint64 if dtype == "uint64" else np.float64)
            out.append({
                "line": lineno,
                "name": name,
                "dtype": dtype,
                "shape": list(arr.shape),
                "values": [int(v) for v in arr.ravel()] if dtype == "uint64"
                          else [float(v) for v in arr.ravel()],
            })
        i = j
    return out


def extract_printed_bit_rows(index_md):
    """Rows printed by dnb2_get_linear_scramble_matrix(print_mats=True)."""
    lines = open(index_md).read().splitlines()
    rows = []
    start = None
    for k, ln in enumerate(lines):
        if ln.startswith("S with shape"):
            start = k
            break
    if start is None:
        return None
    for ln in lines[start + 1:]:
        if ln.startswith(">>>"):
            break
        s = ln.strip()
        if re.fullmatch(r"[01]+", s):
            rows.append(s)
    return {"line": start + 1, "header": lines[start].strip(), "rows": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--skip-digests", action="store_true")
    args = ap.parse_args()
    index_md = os.path.join(args.ref, "docs", "index.md")
    if not os.path.exists(index_md):
        sys.exit("reference docs not found at %s" % index_md)
    payload = {
        "source": "docs/index.md of QMCToolsCL 1.2.1 (doctest displays)",
        "rng": "numpy.random.Generator(numpy.random.SFC64(7)), consumed in document order",
        "arrays": extract_doctest_arrays(index_md),
        "dnb2_S_printed": extract_printed_bit_rows(index_md),
    }
    with open(os.path.join(HERE, "doctest_vectors.json"), "w") as f:
        json.dump(payload, f, indent=0)
    print("wrote %d arrays" % len(payload["arrays"]))
    if not args.skip_digests:
        sys.path.insert(0, HERE)
        import ref_digests  # noqa: E402  (sibling script)
        ref_digests.write_digests(os.path.join(REPO, "oracle", "_ref", "libqmcref.so"),
                                  os.path.join(HERE, "ref_digests.json"))


if __name__ == "__main__":
    main()
