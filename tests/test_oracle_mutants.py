"""Mutation tests of the oracle pins (CPU only).

Each mutant is a copy of oracle/skm_oracle.c with one plausible mistake
planted by a text substitution, compiled into a temporary directory. The pins
of test_oracle_pins_weighted.py must fail on every mutant; the unmodified
source must pass them. This shows the pins constrain the formulas they are
meant to fix (F applied twice with the right exponents, the LCC neighbourhood
weight's exponent and support, the stride lattice's anchor, the variance
rule), rather than restating them.
"""
import os
import shutil
import subprocess

import pytest

import oracle as orc
import test_oracle_pins_weighted as pins

SRC = os.path.join(os.path.dirname(orc.__file__), "skm_oracle.c")
HDR = os.path.join(os.path.dirname(orc.__file__), "skm_oracle.h")

MUTANTS = {
    # F applied once instead of in both factors (non-stationary)
    "ns_F_once": [("double c0 = F * skmo_wendland_c0(sq / pr->sigma_c0);\n    return c4 * c0;  /* F appears",
                   "double c0 = skmo_wendland_c0(sq / pr->sigma_c0);\n    return c4 * c0;  /* F appears")],
    # wrong exponent of sigma_P in F
    "ns_F_exponent": [("pow(fabs(sigmaP), 0.25)", "pow(fabs(sigmaP), 0.5)")],
    # anisotropic F applied once
    "aniso_F_once": [("    double c0 = F * skmo_wendland_c0(sq / pr->sigma_c0);\n    return c4 * c0;\n}",
                      "    double c0 = skmo_wendland_c0(sq / pr->sigma_c0);\n    return c4 * c0;\n}")],
    # anisotropic F with the wrong power of det((S_P + S_Q)/2)
    "aniso_F_exponent": [("* pow(skmo_determinant(n, m), -0.5)", "* pow(skmo_determinant(n, m), -0.25)")],
    # LCC: kw exponent 8 instead of 10 in the value
    "lcc_kw_exponent": [("double kw = neighborhood_weight(n, x, y, sty.support, 10.0);\n                if (kw > 0.0 && lcc_sample",
                         "double kw = neighborhood_weight(n, x, y, sty.support, 8.0);\n                if (kw > 0.0 && lcc_sample")],
    # LCC: the centre's support in kw instead of the neighbour's
    "lcc_kw_centre_support": [("double kw = neighborhood_weight(n, x, y, sty.support, 10.0);\n                if (kw > 0.0 && lcc_sample",
                               "double kw = neighborhood_weight(n, x, y, stx->support, 10.0);\n                if (kw > 0.0 && lcc_sample")],
    # LCC: stride lattice anchored at the centre instead of the box's lower bound
    "lcc_centred_lattice": [("    tight_window(&pr->fixed, pr->dim, x, support, lo, hi);\n}",
                             "    tight_window(&pr->fixed, pr->dim, x, support, lo, hi);\n"
                             "    for (int d = 0; d < pr->dim; ++d) {\n"
                             "        int c = (int)(x[d] + 0.5), sn = pr->subsample_neighborhood;\n"
                             "        lo[d] = c - ((c - lo[d]) / sn) * sn;\n"
                             "    }\n}")],
    # bilateral: intensity Gaussian without the factor 1/2
    "bilateral_intensity_half": [("exp(-di * di / (2.0 * sigma_i * sigma_i))", "exp(-di * di / (sigma_i * sigma_i))")],
    # bilateral: derivative with w_j only (assumes symmetric weights, as the reference's RD does)
    "bilateral_weights": [("2.0 * b * (wj + pr->cpw[ii])", "4.0 * b * wj")],
    # LCC: cc whenever one variance is positive
    "lcc_one_variance": [("        if (d1 > 0 && d2 > 0) {\n            cc = s.sfm",
                          "        if (d1 > 0 || d2 > 0) {\n            cc = s.sfm")],
}

CHECKS = [pins.check_nonstationary_closed_forms, pins.check_anisotropic_closed_forms,
          pins.check_symmetry_and_gram, pins.check_lcc_against_brute_force, pins.check_lcc_zero_variance,
          pins.check_bilateral_closed_forms, pins.check_bilateral_fd]


def _build(tmp_path, name, subs):
    src = open(SRC).read()
    for old, new in subs:
        assert src.count(old) == 1, (name, old)
        src = src.replace(old, new)
    d = tmp_path / name
    d.mkdir()
    shutil.copy(HDR, d / "skm_oracle.h")
    (d / "skm_oracle.c").write_text(src)
    so = d / "libmut.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-std=c99", "-shared", "-o", str(so), str(d / "skm_oracle.c"), "-lm"],
                   check=True)
    return orc.load_path(str(so))


def _failures(lib):
    failed = []
    for chk in CHECKS:
        try:
            chk(lib)
        except Exception:  # noqa: BLE001 -- any failure of a pin counts (NaN in eigvalsh included)
            failed.append(chk.__name__)
    return failed


def test_unmodified_copy_passes_every_pin(tmp_path):
    assert _failures(_build(tmp_path, "clean", [])) == []


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_pins_catch_mutant(tmp_path, name):
    failed = _failures(_build(tmp_path, name, MUTANTS[name]))
    assert failed, f"mutant {name} passes every pin"
