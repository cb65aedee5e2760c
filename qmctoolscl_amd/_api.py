"""Signature table of the 23 kernels on the QMC point-generation and transform path.

Each entry lists the arguments in the reference's positional order, i.e. the
order of the C prototypes in qmctoolscl/c_funcs/<name>.c with the three batch
sizes removed (the reference Python wrapper inserts them itself,
qmctoolscl/util.py:107; the generated ctypes argtypes are built from the same
prototypes, pdm_build.py:54-80).

kind:  "u64"  scalar unsigned 64-bit
       "pu64" pointer to uint64 array
       "pf64" pointer to float64 array
role:  "in", "out" (fully written), "out_partial" (natural-order outputs, which
       leave rows untouched when the index range is not closed under the Gray
       map, exactly as the reference does), "inout" (in place)
"""

SIGNATURES = {
    # lattices -- lat_gen_linear.c:3-13, lat_gen_natural.c:3-14,
    # lat_gen_gray.c:3-14, lat_shift_mod_1.c:3-15
    "lat_gen_linear": [("r", "u64"), ("n", "u64"), ("d", "u64"),
                       ("g", "pu64", "in"), ("x", "pf64", "out")],
    "lat_gen_natural": [("r", "u64"), ("n", "u64"), ("d", "u64"), ("n_start", "u64"),
                        ("g", "pu64", "in"), ("x", "pf64", "out_partial")],
    "lat_gen_gray": [("r", "u64"), ("n", "u64"), ("d", "u64"), ("n_start", "u64"),
                     ("g", "pu64", "in"), ("x", "pf64", "out")],
    "lat_shift_mod_1": [("r", "u64"), ("n", "u64"), ("d", "u64"), ("r_x", "u64"),
                        ("x", "pf64", "in"), ("shifts", "pf64", "in"), ("xr", "pf64", "out")],
    # base-2 digital nets -- dnb2_*.c:3-16
    "dnb2_gmat_lsb_to_msb": [("r", "u64"), ("d", "u64"), ("mmax", "u64"),
                             ("tmaxes", "pu64", "in"), ("C_lsb", "pu64", "in"),
                             ("C_msb", "pu64", "out")],
    "dnb2_linear_matrix_scramble": [("r", "u64"), ("d", "u64"), ("mmax", "u64"),
                                    ("r_C", "u64"), ("tmax_new", "u64"),
                                    ("S", "pu64", "in"), ("C", "pu64", "in"),
                                    ("C_lms", "pu64", "out")],
    "dnb2_gen_natural": [("r", "u64"), ("n", "u64"), ("d", "u64"), ("n_start", "u64"),
                         ("mmax", "u64"), ("C", "pu64", "in"), ("xb", "pu64", "out_partial")],
    "dnb2_gen_gray": [("r", "u64"), ("n", "u64"), ("d", "u64"), ("n_start", "u64"),
                      ("mmax", "u64"), ("C", "pu64", "in"), ("xb", "pu64", "out")],
    "dnb2_digital_shift": [("r", "u64"), ("n", "u64"), ("d", "u64"), ("r_x", "u64"),
                           ("lshifts", "pu64", "in"), ("xb", "pu64", "in"),
                           ("shiftsb", "pu64", "in"), ("xrb", "pu64", "out")],
    "dnb2_integer_to_float": [("r", "u64"), ("n", "u64"), ("d", "u64"),
                              ("tmaxes", "pu64", "in"), ("xb", "pu64", "in"),
                              ("x", "pf64", "out")],
    "dnb2_interlace": [("r", "u64"), ("d_alpha", "u64"), ("mmax", "u64"), ("d", "u64"),
                       ("tmax", "u64"), ("tmax_alpha", "u64"), ("alpha", "u64"),
                       ("C", "pu64", "in"), ("C_alpha", "pu64", "out")],
    "dnb2_undo_interlace": [("r", "u64"), ("d", "u64"), ("mmax", "u64"), ("d_alpha", "u64"),
                            ("tmax", "u64"), ("tmax_alpha", "u64"), ("alpha", "u64"),
                            ("C_alpha", "pu64", "in"), ("C", "pu64", "out")],
    # generalized digital nets -- gdn_*.c:3-19
    "gdn_gen_natural": [("r", "u64"), ("n", "u64"), ("d", "u64"), ("r_b", "u64"),
                        ("mmax", "u64"), ("tmax", "u64"), ("n_start", "u64"),
                        ("bases", "pu64", "in"), ("C", "pu64", "in"),
                        ("xdig", "pu64", "out")],
    "gdn_gen_natural_same_base": [("r", "u64"), ("n", "u64"), ("d", "u64"), ("mmax", "u64"),
                                  ("tmax", "u64"), ("n_start", "u64"), ("b", "u64"),
                                  ("C", "pu64", "in"), ("xdig", "pu64", "out")],
    "gdn_linear_matrix_scramble": [("r", "u64"), ("d", "u64"), ("mmax", "u64"),
                                   ("r_C", "u64"), ("r_b", "u64"), ("tmax", "u64"),
                                   ("tmax_new", "u64"), ("bases", "pu64", "in"),
                                   ("S", "pu64", "in"), ("C", "pu64", "in"),
                                   ("C_lms", "pu64", "out")],
    "gdn_digital_permutation": [("r", "u64"), ("n", "u64"), ("d", "u64"), ("r_x", "u64"),
                                ("r_b", "u64"), ("tmax", "u64"), ("tmax_new", "u64"),
                                ("bmax", "u64"), ("perms", "pu64", "in"),
                                ("xdig", "pu64", "in"), ("xdig_new", "pu64", "out")],
    "gdn_digital_shift": [("r", "u64"), ("n", "u64"), ("d", "u64"), ("r_x", "u64"),
                          ("r_b", "u64"), ("tmax", "u64"), ("tmax_new", "u64"),
                          ("bases", "pu64", "in"), ("shifts", "pu64", "in"),
                          ("xdig", "pu64", "in"), ("xdig_new", "pu64", "out")],
    "gdn_integer_to_float": [("r", "u64"), ("n", "u64"), ("d", "u64"), ("r_b", "u64"),
                             ("tmax", "u64"), ("bases", "pu64", "in"),
                             ("xdig", "pu64", "in"), ("x", "pf64", "out")],
    "gdn_interlace": [("r", "u64"), ("d_alpha", "u64"), ("mmax", "u64"), ("d", "u64"),
                      ("tmax", "u64"), ("tmax_alpha", "u64"), ("alpha", "u64"),
                      ("C", "pu64", "in"), ("C_alpha", "pu64", "out")],
    "gdn_undo_interlace": [("r", "u64"), ("d", "u64"), ("mmax", "u64"), ("d_alpha", "u64"),
                           ("tmax", "u64"), ("tmax_alpha", "u64"), ("alpha", "u64"),
                           ("C_alpha", "pu64", "in"), ("C", "pu64", "out")],
    # transforms -- fwht_1d_radix2.c:3-14, fft_bro_1d_radix2.c:3-17, ifft_bro_1d_radix2.c:3-17
    "fwht_1d_radix2": [("d1", "u64"), ("d2", "u64"), ("n_half", "u64"),
                       ("x", "pf64", "inout")],
    "fft_bro_1d_radix2": [("d1", "u64"), ("d2", "u64"), ("n_half", "u64"),
                          ("twiddler", "pf64", "out"), ("twiddlei", "pf64", "out"),
                          ("xr", "pf64", "inout"), ("xi", "pf64", "inout")],
    "ifft_bro_1d_radix2": [("d1", "u64"), ("d2", "u64"), ("n_half", "u64"),
                           ("twiddler", "pf64", "out"), ("twiddlei", "pf64", "out"),
                           ("xr", "pf64", "inout"), ("xi", "pf64", "inout")],
}

KERNEL_NAMES = tuple(SIGNATURES)


def arg_kind(spec):
    return spec[1]


def arg_role(spec):
    return spec[2] if len(spec) > 2 else "scalar"
