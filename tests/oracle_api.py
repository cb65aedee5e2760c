"""The CPU oracle behind the same Python call surface as qmctoolscl
(positional arguments without batch sizes, outputs written in place, returns a
timing pair), so one replay script can drive either the oracle or the HIP
product.  Test infrastructure only."""
import os
import sys
import time

import numpy as np

import _clib
from qmctoolscl_amd._api import SIGNATURES
from qmctoolscl_amd import rand_funcs

sys.path.insert(0, os.path.join(_clib.REPO, "oracle"))
import nus_oracle  # noqa: E402


class OracleAPI:
    name = "oracle"

    def __init__(self):
        self._lib = _clib.oracle()
        for fname in SIGNATURES:
            setattr(self, fname, self._make(fname))
        for helper in ("random_tbit_uint64s", "random_uint64_permutations", "lat_get_shifts",
                       "dnb2_get_digital_shifts", "dnb2_get_linear_scramble_matrix",
                       "gdn_get_linear_scramble_matrix", "gdn_get_halton_generating_matrix",
                       "gdn_get_digital_shifts", "gdn_get_digital_permutations"):
            setattr(self, helper, getattr(rand_funcs, helper))
        self.NUSNode_dnb2 = nus_oracle.Node2
        self.NUSNode_gdn = nus_oracle.NodeB

    def _make(self, fname):
        names = [s[0] for s in SIGNATURES[fname]]

        def call(*args, **kwargs):
            assert len(args) == len(names), (fname, len(args))
            case = dict(zip(names, args))
            t0, p0 = time.perf_counter(), time.process_time()
            _clib.call_oracle(fname, case, self._lib)
            return time.perf_counter() - t0, time.process_time() - p0
        call.__name__ = fname
        return call

    def dnb2_nested_uniform_scramble(self, r, n, d, r_x, tmax, tmax_new, rngs, root_nodes, xb, xrb):
        nus_oracle.dnb2_nus(r, n, d, r_x, tmax, tmax_new, rngs, root_nodes, xb, xrb)
        return 0.0, 0.0

    def gdn_nested_uniform_scramble(self, r, n, d, r_x, r_b, tmax, tmax_new, rngs, root_nodes,
                                    bases, xdig, xrdig):
        nus_oracle.gdn_nus(r, n, d, r_x, r_b, tmax, tmax_new, rngs, root_nodes, bases, xdig, xrdig)
        return 0.0, 0.0


def as_numpy(a):
    return a if isinstance(a, np.ndarray) else a.cpu().numpy()
