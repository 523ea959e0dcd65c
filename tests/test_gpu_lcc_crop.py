"""LCC at the sliding-motion workload's own parameters (support 20, control
points every 5 voxels, subsample 2, neighbourhood stride 6, so 7^3
neighbours per centre) on volumes small enough for the OpenMP FP64 oracle:
both derivative modes against the oracle; the reference-window derivative
in both of its forms, regrouped per neighbour voxel (the same bits on a
repeated call; also at the coarser levels' parameters) and as the centre walk at forced workgroup
splits (1, 2 and 64 workgroups per control-point tile, the split that the
full-size workload's 51x51x26 grid and the coarse levels' grids select); and
the exact-mode derivative's bits on a repeated call.

Tolerances: value 1e-5 relative, derivative 1e-4 of its largest entry.
"""
import os

import numpy as np
import pytest

import oracle as orc
from helpers import Case, rel_err, scalar_rel
from skmimageregistration_amd import MEASURE, synthetic

pytestmark = pytest.mark.gpu

VALUE_TOL, DERIV_TOL = 1e-5, 1e-4
SUPPORT, SPACING, SUB, SUBN = 20.0, 5.0, 2, 6


def _threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def _sliding_case(shape, mode, exact):
    case = Case(dim=3, shape=shape, cp_spacing=SPACING, sigma=SUPPORT, mode=mode, subsample=SUB,
                subsample_neighborhood=SUBN, seed=21, exact=exact, coef_scale=1.0)
    mov, fix, _ = synthetic.sliding_3d(shape, shift=2.0, breathing=3.0)
    case.moving, case.fixed = mov, fix
    return case


def _oracle(case):
    orc.load(parallel=True).skmo_set_threads(_threads())
    res, der = case.oracle(case.offsets(1), parallel=True).evaluate(int(MEASURE.LCC), True)
    return res, der


def _check(ev, case, ores, oder, label):
    res, der = ev.getValueAndDerivative(MEASURE.LCC, case.cp, case.cpw, do_resampling=False)
    np.testing.assert_array_equal(ev.cube_offsets(), case.offsets(1))
    assert res.numPixelsCountedGpu == ores["pixels_counted"], label
    assert scalar_rel(res.measure, ores["measure"]) < VALUE_TOL, (label, res.measure, ores["measure"])
    err = rel_err(der, oder)
    print(f"{label}: value {res.measure:.9f} (oracle {ores['measure']:.9f}) derivative rel {err:.2e}", flush=True)
    assert err < DERIV_TOL, (label, err)
    return res, der


def test_lcc_reference_windows_cfg_parameters_forced_splits():
    case = _sliding_case((48, 48, 48), 0, False)
    print("LCC reference windows, 48^3, support 20, stride 6, subsample 2", flush=True)
    ores, oder = _oracle(case)
    ev = case.gpu()
    ev.getValue(MEASURE.LCC, case.cp, case.cpw)  # first draw of the offsets
    # the sum regrouped per neighbour voxel, twice (same bits)
    ev.set_tuning(lcc_ref_walk=2)
    r1, d1 = _check(ev, case, ores, oder, "voxel tables")
    r2, d2 = _check(ev, case, ores, oder, "voxel tables, repeated")
    assert r1.measure == r2.measure and np.array_equal(d1, d2), "voxel-table derivative differs between runs"
    # the centre walk at forced workgroup splits
    for chunks in (0, 1, 2, 64):
        ev.set_tuning(lcc_ref_chunks=chunks, lcc_ref_walk=1)
        _, dw = _check(ev, case, ores, oder, f"centre walk, chunks={chunks or 'auto'}")
        assert rel_err(d1, dw) < 1e-5
    ev.close()


@pytest.mark.parametrize("support,spacing,sub,subn", [(40.0, 10.0, 3, 12), (80.0, 20.0, 4, 23)])
def test_lcc_reference_windows_coarse_level_parameters(support, spacing, sub, subn):
    """The sliding-motion registration's coarser levels (support 40 and 80,
    control points every support / 4, subsample 3 and 4, strides 12 and 23):
    reach rows wider than 64 positions and control-point cells split into
    voxel pieces; the voxel tables and the centre walk against the oracle."""
    case = Case(dim=3, shape=(48, 48, 48), cp_spacing=spacing, sigma=support, mode=0, subsample=sub,
                subsample_neighborhood=subn, seed=23, exact=False, coef_scale=1.0)
    mov, fix, _ = synthetic.sliding_3d((48, 48, 48), shift=2.0, breathing=3.0)
    case.moving, case.fixed = mov, fix
    print(f"LCC reference windows, 48^3, support {support:g}, stride {subn}, subsample {sub}", flush=True)
    ores, oder = _oracle(case)
    ev = case.gpu()
    ev.getValue(MEASURE.LCC, case.cp, case.cpw)
    ev.set_tuning(lcc_ref_walk=2)
    _, d1 = _check(ev, case, ores, oder, "voxel tables")
    ev.set_tuning(lcc_ref_walk=1)
    _, dw = _check(ev, case, ores, oder, "centre walk")
    assert rel_err(d1, dw) < 1e-5
    ev.close()


def test_lcc_exact_cfg_parameters_repeatable():
    case = _sliding_case((48, 48, 48), 0, True)
    print("LCC exact gradient, 48^3, support 20, stride 6, subsample 2", flush=True)
    ores, oder = _oracle(case)
    ev = case.gpu()
    ev.getValue(MEASURE.LCC, case.cp, case.cpw)
    r1, d1 = _check(ev, case, ores, oder, "exact")
    r2, d2 = _check(ev, case, ores, oder, "exact, repeated")
    assert r1.measure == r2.measure and np.array_equal(d1, d2), "exact LCC derivative differs between runs"
    ev.close()


@pytest.mark.parametrize("exact", [False, True])
def test_lcc_tensor_kernel_cfg_parameters(exact):
    """Anisotropic kernel: the reference's weighted-mode quirks (sigma(y)
    twice in the 3D derivative's neighbour warp, the neighbour's support in
    kw) at the workload's window sizes."""
    case = _sliding_case((32, 32, 32), 2, exact)
    print(f"LCC tensor kernel, 32^3, exact={exact}", flush=True)
    ores, oder = _oracle(case)
    ev = case.gpu()
    ev.getValue(MEASURE.LCC, case.cp, case.cpw)
    _check(ev, case, ores, oder, f"tensor exact={exact}")
    ev.close()
