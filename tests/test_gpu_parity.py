"""HIP path (libskm.so, through the C ABI) against the FP64 oracle on small inputs.

Tolerances (FP32 per pair, FP64 final sums):
  metric value       relative 1e-5
  derivative         max-norm relative 1e-4  (max|d_gpu - d_oracle| / max|d_oracle|)
  regularizer values relative 1e-5
  displacement       max-norm relative 1e-5
"""
import numpy as np
import pytest

import oracle as orc
from helpers import Case, rel_err, scalar_rel
from skmimageregistration_amd import MEASURE

pytestmark = pytest.mark.gpu

VALUE_TOL = 1e-5
DERIV_TOL = 1e-4


def _compare(case, measure, regs=(0.0, 0.0, 0.0), pg_scaling=1.0, calls=1, value_only=False, tuning=None):
    ev = case.gpu()
    if tuning:
        ev.set_tuning(**tuning)
    if regs[0] > 0:
        ev.SetRegularizerRKHS(regs[0])
    if regs[1] > 0:
        ev.SetRegularizerRD(regs[1])
    if regs[2] > 0:
        ev.SetRegularizerPG(regs[2], pg_scaling)
    df = case.df_image()
    for _ in range(calls - 1):
        ev.getValue(measure, case.cp, case.cpw, df, do_resampling=True)
    if value_only:
        res = ev.getValue(measure, case.cp, case.cpw, df, do_resampling=True)
        der = None
    else:
        res, der = ev.getValueAndDerivative(measure, case.cp, case.cpw, df, do_resampling=True)
    offsets = None
    if case.sub > 1:
        offsets = case.offsets(calls)
        np.testing.assert_array_equal(ev.cube_offsets(), offsets)  # bit-exact PCG32 jitter
    orc = case.oracle(offsets)
    ores, oder = orc.evaluate(int(measure), not value_only, *regs, pg_scaling)
    assert res.numPixelsCountedGpu == ores["pixels_counted"]
    assert scalar_rel(res.measure, ores["measure"]) < VALUE_TOL, (res.measure, ores["measure"])
    for name, gv in (("reg_rkhs", res.reg_value_rkhs), ("reg_rd", res.reg_value_rd), ("reg_pg", res.reg_value_pg)):
        assert scalar_rel(gv, ores[name]) < VALUE_TOL, (name, gv, ores[name])
    if der is not None:
        assert rel_err(der, oder) < DERIV_TOL, rel_err(der, oder)
    ev.close()


@pytest.mark.parametrize("measure", [MEASURE.MSE, MEASURE.NCC])
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_2d_modes(measure, mode):
    _compare(Case(dim=2, mode=mode, seed=3), measure)


@pytest.mark.parametrize("measure", [MEASURE.MSE, MEASURE.NCC])
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_3d_modes(measure, mode):
    _compare(Case(dim=3, shape=(18, 20, 22), cp_spacing=5.0, sigma=12.0, mode=mode, seed=5), measure)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("sub", [2, 3])
@pytest.mark.parametrize("measure", [MEASURE.MSE, MEASURE.NCC])
def test_subsampled(dim, sub, measure):
    shape = (48, 52) if dim == 2 else (18, 20, 22)
    _compare(Case(dim=dim, shape=shape, cp_spacing=5.0, sigma=12.0, subsample=sub, seed=11), measure, calls=2)


@pytest.mark.parametrize("dim", [2, 3])
def test_exact_windows(dim):
    shape = (48, 52) if dim == 2 else (18, 20, 22)
    _compare(Case(dim=dim, shape=shape, cp_spacing=5.0, sigma=12.0, subsample=2, exact=True, seed=13), MEASURE.MSE)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("mode", [0, 1])
def test_regularizers(dim, mode):
    shape = (40, 44) if dim == 2 else (16, 18, 20)
    case = Case(dim=dim, shape=shape, cp_spacing=4.0, sigma=10.0, mode=mode, seed=17, coef_scale=0.8)
    _compare(case, MEASURE.MSE, regs=(0.3, 0.2, 0.1), pg_scaling=0.7)
    _compare(case, MEASURE.NCC, regs=(0.3, 0.2, 0.1), pg_scaling=1.3, value_only=True)


@pytest.mark.parametrize("dim", [2, 3])
def test_previous_displacement(dim):
    shape = (40, 44) if dim == 2 else (16, 18, 20)
    _compare(Case(dim=dim, shape=shape, cp_spacing=5.0, sigma=12.0, with_df=True, seed=19), MEASURE.NCC)


@pytest.mark.parametrize("dim", [2, 3])
def test_geometry(dim):
    origin = (-3.5, 2.25, 1.0)[:dim]
    spacing = (1.25, 0.8, 1.5)[:dim]
    shape = (36, 40) if dim == 2 else (16, 18, 20)
    _compare(Case(dim=dim, shape=shape, cp_spacing=6.0, sigma=14.0, origin=origin, spacing=spacing, seed=23),
             MEASURE.MSE)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_displacement_field(dim, mode):
    shape = (40, 44) if dim == 2 else (16, 18, 20)
    case = Case(dim=dim, shape=shape, cp_spacing=5.0, sigma=12.0, mode=mode, with_df=True, seed=29)
    ev = case.gpu()
    u = ev.evaluateDisplacementField(case.cp, case.cpw, case.df_image())
    ref = case.oracle().displacement()
    assert rel_err(u, ref) < 1e-5
    ev.close()


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("exact", [False, True])
def test_lcc(dim, exact):
    shape = (28, 30) if dim == 2 else (12, 13, 14)
    case = Case(dim=dim, shape=shape, cp_spacing=4.0, sigma=7.0, subsample=2, subsample_neighborhood=2,
                exact=exact, seed=31, coef_scale=0.8)
    _compare(case, MEASURE.LCC)


@pytest.mark.parametrize("sub,subn,spacing", [(3, 3, (0.8, 1.25, 1.0)), (2, 1, (1.0, 1.0, 1.5)), (1, 2, (1.2, 0.9, 1.1))])
def test_lcc_exact_geometry(sub, subn, spacing):
    """The exact-gradient LCC derivative's per-voxel gather (stationary
    kernels) with anisotropic voxel spacing, subsample 1, 2 and 3 and
    neighbourhood strides 1-3: the reach masks and the support ranges per axis
    in voxel units, the cell index of a position, the jitter match."""
    case = Case(dim=3, shape=(14, 15, 16), cp_spacing=4.0, sigma=6.5, subsample=sub, subsample_neighborhood=subn,
                exact=True, seed=33, coef_scale=0.8, spacing=spacing, origin=(1.5, -2.0, 0.25))
    _compare(case, MEASURE.LCC)


@pytest.mark.parametrize("mode", [1, 2])
def test_lcc_weighted(mode):
    case = Case(dim=3, shape=(12, 13, 14), cp_spacing=4.0, sigma=7.0, mode=mode, subsample=2,
                subsample_neighborhood=2, seed=37, coef_scale=0.8)
    _compare(case, MEASURE.LCC)


def test_value_matches_value_and_derivative():
    """GetValue == GetValueAndDerivative value (itkFiniteDifferenceTest.cpp:188-195)."""
    case = Case(dim=3, shape=(18, 20, 22), cp_spacing=5.0, sigma=12.0, subsample=2, seed=41)
    ev = case.gpu()
    for measure in (MEASURE.MSE, MEASURE.NCC, MEASURE.LCC):
        r1, _ = ev.getValueAndDerivative(measure, case.cp, case.cpw, do_resampling=False)
        r2 = ev.getValue(measure, case.cp, case.cpw, do_resampling=False)
        assert r1.measure == r2.measure
    ev.close()


@pytest.mark.parametrize("tuning", [dict(gather_split=4), dict(value_split=1), dict(value_split=4),
                                    dict(value_split=16), dict(gather_split=4, value_split=8),
                                    dict(gather_regions=1), dict(gather_regions=8, gather_split=4)],
                         ids=lambda t: ",".join(f"{k}={v}" for k, v in t.items()))
@pytest.mark.parametrize("mode", [0, 2])
def test_launch_shapes(tuning, mode):
    """Every launch shape skm_set_tuning can select gives the oracle's result."""
    case = Case(dim=3, shape=(18, 20, 22), cp_spacing=5.0, sigma=12.0, mode=mode, subsample=2, seed=43)
    _compare(case, MEASURE.NCC, tuning=tuning)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("mode", [0, 2])
def test_bilateral_regularizer(dim, mode):
    """The bilateral term (no reference counterpart) against the oracle's
    restatement: its value and the derivative it adds to MSE's, with RD on
    too, non-uniform control-point weights and spatial sigma 2 control-point
    spacings (cut at 6 spacings, wider than the grid in places)."""
    shape = (40, 44) if dim == 2 else (16, 18, 20)
    case = Case(dim=dim, shape=shape, cp_spacing=4.0, sigma=10.0, mode=mode, seed=47, coef_scale=0.8)
    ev = case.gpu()
    ev.SetRegularizerRD(0.2)
    ev.SetRegularizerBilateral(0.3, 8.0, 0.1)
    res, der = ev.getValueAndDerivative(MEASURE.MSE, case.cp, case.cpw)
    ev.close()
    pr = case.oracle()
    ores, oder = pr.evaluate(int(MEASURE.MSE), True, 0.0, 0.2, 0.0)
    bval, bder = pr.bilateral(0.3, 8.0, 0.1)
    assert bval > 0
    assert scalar_rel(res.reg_value_bilateral, bval) < VALUE_TOL, (res.reg_value_bilateral, bval)
    assert scalar_rel(res.reg_value_rd, ores["reg_rd"]) < VALUE_TOL
    assert scalar_rel(res.measure, ores["measure"]) < VALUE_TOL
    assert rel_err(der, oder + bder) < DERIV_TOL, rel_err(der, oder + bder)
    # the device optimizer's value assembly adds the weighted bilateral value
    ev = case.gpu()
    ev.SetRegularizerRD(0.2)
    ev.SetRegularizerBilateral(0.3, 8.0, 0.1)
    soa = np.ascontiguousarray(case.cp.T).reshape(-1)
    _, recs = ev.optimize(soa, metric=MEASURE.MSE, iterations=1, cpwData=case.cpw, use_negative_gradient=True)
    ev.close()
    expected = ores["measure"] + 0.2 * ores["reg_rd"] + 0.3 * bval
    assert scalar_rel(recs[0][0], expected) < VALUE_TOL, (recs[0][0], expected)


@pytest.mark.parametrize("dim,measure", [(2, MEASURE.MSE), (3, MEASURE.NCC), (3, MEASURE.LCC)])
def test_moving_coefficients_as_the_reference_passes_them(dim, measure):
    """SKM_FLAG_MOVING_COEFFICIENTS: the (n+3)^D prefiltered coefficients
    that the reference's TransformAdapter hands to GpuKernelEvaluator
    (itkTransformAdapter.hxx:435-461), here from the oracle's FP64 prefilter
    rounded to FP32, give the oracle's result and the raw-image path's."""
    case = Case(dim=dim, shape=(40, 44) if dim == 2 else (16, 18, 20), cp_spacing=5.0, sigma=12.0, seed=49,
                subsample_neighborhood=2 if measure == MEASURE.LCC else 1)
    coeffs = orc.bspline_coefficients(case.moving).astype(np.float32)
    ev_raw = case.gpu()
    ev_cf = case.gpu(moving_coefficients=coeffs)
    r_raw, d_raw = ev_raw.getValueAndDerivative(measure, case.cp, case.cpw)
    r_cf, d_cf = ev_cf.getValueAndDerivative(measure, case.cp, case.cpw)
    ev_raw.close()
    ev_cf.close()
    ores, oder = case.oracle().evaluate(int(measure), True)
    assert r_cf.numPixelsCountedGpu == ores["pixels_counted"]
    assert scalar_rel(r_cf.measure, ores["measure"]) < VALUE_TOL
    assert rel_err(d_cf, oder) < DERIV_TOL
    assert scalar_rel(r_cf.measure, r_raw.measure) < 1e-6
    assert rel_err(d_cf, d_raw) < 1e-5
    with pytest.raises(ValueError):
        case.gpu(moving_coefficients=coeffs[1:])
