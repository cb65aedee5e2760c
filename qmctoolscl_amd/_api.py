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

# Elements of every array argument a call reads or writes, from the sizes in
# the reference's prototype comments and index expressions (c_funcs/<name>.c:
# e.g. lat_shift_mod_1.c:11-13, dnb2_digital_shift.c:32 for lshifts[l % r_x],
# gdn_linear_matrix_scramble.c:33 for bases[(l % r_b) * d + j]).  The wrappers
# reject a smaller array before any launch.
SIZES = {
    "lat_gen_linear": {"g": "r*d", "x": "r*n*d"},
    "lat_gen_natural": {"g": "r*d", "x": "r*n*d"},
    "lat_gen_gray": {"g": "r*d", "x": "r*n*d"},
    "lat_shift_mod_1": {"x": "r_x*n*d", "shifts": "r*d", "xr": "r*n*d"},
    "dnb2_gmat_lsb_to_msb": {"tmaxes": "r", "C_lsb": "r*d*mmax", "C_msb": "r*d*mmax"},
    "dnb2_linear_matrix_scramble": {"S": "r*d*tmax_new", "C": "r_C*d*mmax", "C_lms": "r*d*mmax"},
    "dnb2_gen_natural": {"C": "r*d*mmax", "xb": "r*n*d"},
    "dnb2_gen_gray": {"C": "r*d*mmax", "xb": "r*n*d"},
    "dnb2_digital_shift": {"lshifts": "r_x", "xb": "r_x*n*d", "shiftsb": "r*d", "xrb": "r*n*d"},
    "dnb2_integer_to_float": {"tmaxes": "r", "xb": "r*n*d", "x": "r*n*d"},
    "dnb2_interlace": {"C": "r*d*mmax", "C_alpha": "r*d_alpha*mmax"},
    "dnb2_undo_interlace": {"C_alpha": "r*d_alpha*mmax", "C": "r*d*mmax"},
    "gdn_gen_natural": {"bases": "r_b*d", "C": "r*d*mmax*tmax", "xdig": "r*n*d*tmax"},
    "gdn_gen_natural_same_base": {"C": "r*d*mmax*tmax", "xdig": "r*n*d*tmax"},
    "gdn_linear_matrix_scramble": {"bases": "r_b*d", "S": "r*d*tmax_new*tmax", "C": "r_C*d*mmax*tmax",
                                   "C_lms": "r*d*mmax*tmax_new"},
    "gdn_digital_permutation": {"perms": "r*d*tmax_new*bmax", "xdig": "r_x*n*d*tmax", "xdig_new": "r*n*d*tmax_new"},
    "gdn_digital_shift": {"bases": "r_b*d", "shifts": "r*d*tmax_new", "xdig": "r_x*n*d*tmax",
                          "xdig_new": "r*n*d*tmax_new"},
    "gdn_integer_to_float": {"bases": "r_b*d", "xdig": "r*n*d*tmax", "x": "r*n*d"},
    "gdn_interlace": {"C": "r*d*mmax*tmax", "C_alpha": "r*d_alpha*mmax*tmax_alpha"},
    "gdn_undo_interlace": {"C_alpha": "r*d_alpha*mmax*tmax_alpha", "C": "r*d*mmax*tmax"},
    "fwht_1d_radix2": {"x": "d1*d2*2*n_half"},
    "fft_bro_1d_radix2": {"twiddler": "2*n_half", "twiddlei": "2*n_half", "xr": "d1*d2*2*n_half",
                          "xi": "d1*d2*2*n_half"},
    "ifft_bro_1d_radix2": {"twiddler": "2*n_half", "twiddlei": "2*n_half", "xr": "d1*d2*2*n_half",
                           "xi": "d1*d2*2*n_half"},
}


def required_elements(name, scalars):
    """{array argument: elements the call touches} for a call of `name` with
    the scalar arguments `scalars` (a dict by argument name)."""
    out = {}
    for arg, expr in SIZES[name].items():
        v = 1
        for f in expr.split("*"):
            v *= int(f) if f.isdigit() else int(scalars[f])
        out[arg] = v
    return out


def arg_kind(spec):
    return spec[1]


def arg_role(spec):
    return spec[2] if len(spec) > 2 else "scalar"
