"""qmctoolscl_amd: QMCToolsCL's point-generation and transform path on MI355X.

Drop-in for the `qmctoolscl` call surface on that path:

* the 23 kernels (lat_*, dnb2_*, gdn_*, fwht_1d_radix2, fft_bro_1d_radix2,
  ifft_bro_1d_radix2) with the reference's positional arguments, in-place
  outputs and (t_perf, t_process) return -- see kernels.py;
* the host-side random draws (lat_get_shifts, dnb2_get_*, gdn_get_*,
  random_tbit_uint64s, random_uint64_permutations) -- rand_funcs.py;
* nested uniform scrambling with persistent trees (NUSNode_dnb2,
  NUSNode_gdn, dnb2_nested_uniform_scramble, gdn_nested_uniform_scramble);
* the generalized Halton generator halton_qrng_c (halton.py).

All compute runs in hand-written gfx950 HIP kernels (libqmctoolscl_amd.so,
C ABI in include/qmctoolscl_amd.h); there is no CPU fallback.
"""
__version__ = "0.1.0"

from .rand_funcs import (  # noqa: F401
    random_tbit_uint64s, random_uint64_permutations, lat_get_shifts, dnb2_get_digital_shifts,
    dnb2_get_linear_scramble_matrix, gdn_get_linear_scramble_matrix, gdn_get_halton_generating_matrix,
    gdn_get_digital_shifts, gdn_get_digital_permutations,
)
from .kernels import *  # noqa: F401,F403
from .nus import (  # noqa: F401
    NUSNode_dnb2, NUSNode_gdn, dnb2_nested_uniform_scramble, gdn_nested_uniform_scramble,
)
from .halton import halton_qrng, halton_qrng_c  # noqa: F401
from ._api import SIGNATURES, KERNEL_NAMES  # noqa: F401
from ._lib import QMCError  # noqa: F401
