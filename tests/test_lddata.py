"""LDData readers (qmctoolscl_amd/lddata.py) against the reference's own way of
loading the same files (np.loadtxt(..., dtype=np.uint64, skiprows=7 / 6),
_problems.py:8,17), on local files written in that layout; then the loaded
inputs through the generators, on the GPU, against the oracle."""
import numpy as np
import pytest

import _clib
from qmctoolscl_amd import lddata


def write(path, header, rows):
    with open(path, "w") as f:
        for h in header:
            f.write(h + "\n")
        for r in rows:
            f.write(" ".join(str(int(v)) for v in r) + "\n")


def net_rows(d, m, t, seed):
    rng = np.random.default_rng(seed)
    rows = []
    for j in range(d):
        rows.append([(1 << (t - 1 - k)) | (int(rng.integers(0, 1 << 62)) & ((1 << k) - 1)) << (t - k) for k in range(m)])
    return rows


DNET_HEADER = ["# dnet", "# base 2 digital net, made-up matrices", "# 2", "# 12 dimensions", "# 32 columns",
               "# 64 bits", "# end of header"]
LAT_HEADER = ["# lattice", "# made-up rank-1 vector", "# 12 dimensions", "# 1048576 points", "# base 2",
              "# end of header"]


def test_dnet_reader_matches_loadtxt(tmp_path):
    rows = net_rows(12, 32, 64, 1)
    p = tmp_path / "net.txt"
    write(p, DNET_HEADER, rows)
    C, header = lddata.read_dnet(str(p))
    want = np.loadtxt(str(p), dtype=np.uint64, skiprows=7)
    assert C.shape == (1, 12, 32) and C.dtype == np.uint64
    assert header == DNET_HEADER
    assert np.array_equal(C[0], np.array(rows, dtype=np.uint64))   # all 64 bits
    assert np.array_equal(C[0], want)
    C5, _ = lddata.read_dnet(str(p), d=5)
    assert np.array_equal(C5[0], C[0][:5])
    with pytest.raises(AssertionError):
        lddata.read_dnet(str(p), d=13)
    Ch, _ = lddata.read_dnet(str(p), skiprows=None)  # header found by its '#' lines
    assert np.array_equal(Ch, C)


def test_lattice_reader_matches_loadtxt(tmp_path):
    rng = np.random.default_rng(2)
    g = [1] + [int(2 * rng.integers(0, 2 ** 19) + 1) for _ in range(11)]
    p = tmp_path / "lat.txt"
    write(p, LAT_HEADER, [[v] for v in g])
    gg, header = lddata.read_lattice(str(p))
    want = np.loadtxt(str(p), dtype=np.uint64, skiprows=6)
    assert gg.shape == (1, 12) and np.array_equal(gg[0], want) and header == LAT_HEADER
    g3, _ = lddata.read_lattice(str(p), d=3)
    assert np.array_equal(g3[0], want[:3])


def test_readers_reject_malformed(tmp_path):
    p = tmp_path / "bad.txt"
    write(p, DNET_HEADER, [[1, 2, 3], [4, 5]])
    with pytest.raises(ValueError):
        lddata.read_dnet(str(p))
    write(p, LAT_HEADER, [[1, 2]])
    with pytest.raises(ValueError):
        lddata.read_lattice(str(p))
    write(p, LAT_HEADER, [[1 << 64]])
    with pytest.raises(ValueError):
        lddata.read_lattice(str(p))


@pytest.mark.gpu
def test_loaded_inputs_generate_on_gpu(tmp_path):
    """The reference's own uses (_problems.py:7-23): dnb2_gen_gray on the first d
    rows of a dnet file, lat_gen_gray on the first d entries of a lattice file."""
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    import qmctoolscl_amd as q
    U = np.uint64
    p = tmp_path / "net.txt"
    write(p, DNET_HEADER, net_rows(12, 32, 64, 3))
    n, d = 1 << 12, 10
    C, _ = lddata.read_dnet(str(p), d=d)
    mmax = C.shape[2]
    xb = np.empty((1, n, d), dtype=np.uint64)
    q.dnb2_gen_gray(U(1), U(n), U(d), U(0), U(mmax), C, xb)
    want = np.empty_like(xb)
    _clib.call_oracle("dnb2_gen_gray", dict(r=1, n=n, d=d, n_start=0, mmax=mmax, C=C, xb=want))
    assert np.array_equal(xb, want)
    p = tmp_path / "lat.txt"
    rng = np.random.default_rng(4)
    write(p, LAT_HEADER, [[1]] + [[int(2 * rng.integers(0, 2 ** 19) + 1)] for _ in range(11)])
    g, _ = lddata.read_lattice(str(p), d=d)
    x = np.empty((1, n, d))
    q.lat_gen_gray(U(1), U(n), U(d), U(0), g, x)
    wx = np.empty_like(x)
    _clib.call_oracle("lat_gen_gray", dict(r=1, n=n, d=d, n_start=0, g=g, x=wx))
    assert np.array_equal(x.view(np.uint64), wx.view(np.uint64))
