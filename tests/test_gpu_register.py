"""The multi-scale registration driver (skm_register, the scale loop of
src/SKMReg.cxx:84-525) against the same loop restated in Python over the
package's building blocks, plus restart equivalence and recovery of a known
warp.

The Python restatement calls the same library entry points in the same order
with the same inputs, so the displacement must agree to float rounding of the
host-side preprocessing (1e-6 of its range); the preprocessing itself
(ThresholdImage / RescaleImage, include/itkUtils.h:45-112) is restated here in
numpy from ITK's filter definitions.
"""
import numpy as np
import pytest

from skmimageregistration_amd import (MEASURE, ComputeAnisotropicTensor, GaussianSmoothing, GpuKernelEvaluator, Grid,
                                      Image, Kernel, RegistrationConfig, _capi, control_point_grid,
                                      register_multiscale, synthetic)

pytestmark = pytest.mark.gpu


def threshold_image(img, thresh_max, thresh_min):
    """ThresholdImage: two ITK ThresholdOutside passes (itkUtils.h:45-100)."""
    x = img.astype(np.float64)
    below_max = x[x < thresh_max].max() if (x < thresh_max).any() else -np.finfo(np.float64).max
    above_min = x[x > thresh_min].min() if (x > thresh_min).any() else np.finfo(np.float64).max
    lo, hi = x.min(), x.max()
    y = img.astype(np.float32).copy()
    y[(x < lo) | (x > below_max)] = np.float32(below_max)
    x2 = y.astype(np.float64)
    y[(x2 < above_min) | (x2 > hi)] = np.float32(above_min)
    return y


def rescale_image(img):
    """RescaleIntensityImageFilter to [0, 1] (itkUtils.h:103-112)."""
    x = img.astype(np.float64)
    lo, hi = x.min(), x.max()
    scale = 1.0 / (hi - lo) if hi != lo else (1.0 / hi if hi != 0 else 0.0)
    return np.clip(x * scale - lo * scale, 0.0, 1.0).astype(np.float32)


def python_loop(reference, target, cfg):
    dim = target.dim
    ref, tar = reference.data.astype(np.float32), target.data.astype(np.float32)
    th = cfg["intensity_thresholds"]
    if len(th) == 2:
        ref = threshold_image(ref, th[1], th[0])
        tar = threshold_image(tar, th[1], th[0])
    ref, tar = rescale_image(ref), rescale_image(tar)
    measure = {"mse": MEASURE.MSE, "ncc": MEASURE.NCC, "lcc": MEASURE.LCC}[cfg["metric"]]
    prev = None
    for s in range(int(cfg.CURRENT("num_scales", 0))):
        var = cfg.CURRENT("smoothing_variance", s)
        rs = GaussianSmoothing(Image(ref, reference.origin, reference.spacing), var).data if var > 1e-5 else ref
        ts = GaussianSmoothing(Image(tar, target.origin, target.spacing), var).data if var > 1e-5 else tar
        s4, s0 = cfg.CURRENT("sigma_C4", s), cfg.CURRENT("sigma_C0", s)
        tw = cfg.CURRENT("tensor_weights", s)
        if tw > 0:
            ten, tilt, mw, mcov = ComputeAnisotropicTensor(Image(ts, target.origin, target.spacing), s4, tw, 2)
            kernel = Kernel(s0, s4, useWeightTensor=True, maximumWeight=mw, maximumEigenvalueCovariance=mcov,
                            wTensor=(Grid(target.size(), target.origin, target.spacing), ten))
        else:
            kernel = Kernel(s0, s4)
        sp = min(s4, s0) / cfg.CURRENT("cp_spacing_scaler", 0)
        cpg = control_point_grid(Image(rs, reference.origin, reference.spacing), (sp,) * dim)
        ev = GpuKernelEvaluator(cpg.count() * dim, cfg.CURRENT("sampling_rate", s),
                                cfg.CURRENT("neighborhood_sampling", s), kernel,
                                Image(rs, reference.origin, reference.spacing),
                                Image(ts, target.origin, target.spacing), cpg)
        if cfg.CURRENT("reg_rd", s) > 0:
            ev.SetRegularizerRD(cfg.CURRENT("reg_rd", s), cfg.CURRENT("reg_rd_scaling", s))
        if cfg.CURRENT("reg_bilateral", s) > 0:
            ev.SetRegularizerBilateral(cfg.CURRENT("reg_bilateral", s), cfg.CURRENT("reg_bilateral_spatial_sigma", s),
                                       cfg.CURRENT("reg_bilateral_intensity_sigma", s))
        prev_img = None if prev is None else Image(prev, target.origin, target.spacing, dim=dim)
        ev.SetDisplacementField(prev_img)
        regs = {k: float(cfg.CURRENT(k, s)) for k in ("reg_l1", "reg_l21", "reg_l2")}
        only = [k for k in ("reg_l1", "reg_l21", "reg_l2", "reg_rkhs", "reg_rd", "reg_pg", "reg_bilateral")
                if cfg.CURRENT(k, s) > 0]
        orthant = only in (["reg_l1"], ["reg_l21"])
        params, recs = ev.optimize(np.zeros(cpg.count() * dim, np.float32), metric=measure, **regs,
                                   kind=_capi.SKM_OPT_ADAPTIVE, a=cfg.CURRENT("initial_step_size", s), A=20.0,
                                   alpha=1.0, fmax=1e7, fmin=-1e6, omega=1.0, orthant_projection=orthant,
                                   iterations=int(cfg.CURRENT("num_function_evaluations", s)), return_best=True)
        cp_aos = params.reshape(dim, -1).T.copy()
        prev = ev.evaluateDisplacementField(cp_aos, np.ones(cpg.count(), np.float32),
                                            prev_img).reshape(tar.shape + (dim,))
        ev.close()
    return prev


def _case3d():
    mov, fix, disp = synthetic.ellipsoids_3d((36, 40, 44), n=10, smooth=1.5, max_disp=2.5, noise=0.01,
                                             seeds=(5, 6, 7))
    return Image(mov), Image(fix), disp


@pytest.mark.parametrize("metric,tensor,bilateral", [("ncc", 400.0, 0.0), ("mse", 0.0, 0.0), ("ncc", 400.0, 1e-3)])
def test_driver_matches_python_loop(metric, tensor, bilateral):
    reference, target, _ = _case3d()
    # thresholds inside the intensity range so that ThresholdImage clips
    cfg = RegistrationConfig(sigma_C4=[16, 8], sigma_C0=[16, 8], tensor_weights=[tensor], num_scales=2,
                             num_function_evaluations=[15, 10], initial_step_size=[0.5], sampling_rate=[2, 1],
                             neighborhood_sampling=[1], smoothing_variance=[2, 0], reg_rd=[1e-3],
                             reg_bilateral=[bilateral], reg_bilateral_spatial_sigma=[8.0],
                             reg_bilateral_intensity_sigma=[0.2], intensity_thresholds=[0.05, 0.8], metric=metric)
    print(f"driver vs python loop metric={metric} tensor={tensor} bilateral={bilateral}", flush=True)
    df, reports = register_multiscale(reference, target, cfg)
    ref_df = python_loop(reference, target, cfg)
    assert np.isfinite(df).all() and np.abs(df).max() > 0
    np.testing.assert_allclose(df, ref_df, rtol=0, atol=1e-6 * max(1.0, np.abs(ref_df).max()))
    for r in reports:
        assert r["iterations"] > 0 and r["value"] <= r["first_value"]


def test_restart_continues_the_same_registration():
    reference, target, _ = _case3d()
    base = dict(sigma_C4=[16, 8], sigma_C0=[16, 8], tensor_weights=[0], num_function_evaluations=[12],
                initial_step_size=[0.5], sampling_rate=[1], smoothing_variance=[1, 0], metric="ncc")
    full, _ = register_multiscale(reference, target, RegistrationConfig(num_scales=2, **base))
    first, _ = register_multiscale(reference, target, RegistrationConfig(num_scales=1, **base))
    resumed, reps = register_multiscale(reference, target, RegistrationConfig(num_scales=2, restart_at=2, **base),
                                        restart_displacement=first)
    assert reps[0]["iterations"] == 0 and reps[1]["iterations"] == 12
    np.testing.assert_array_equal(resumed, full)


def test_registration_recovers_the_warp():
    reference, target, disp = _case3d()
    cfg = RegistrationConfig(sigma_C4=[16, 8], sigma_C0=[16, 8], tensor_weights=[0], num_scales=2,
                             num_function_evaluations=[60, 60], initial_step_size=[1000.0], sampling_rate=[1],
                             smoothing_variance=[1, 0], metric="ncc")
    df, reports = register_multiscale(reference, target, cfg)
    print("reports", reports, flush=True)
    for r in reports:
        assert r["value"] < r["first_value"]
    inner = (slice(6, -6),) * 3
    err_before = np.abs(disp[inner]).mean()
    err_after = np.abs(df[inner] - disp[inner]).mean()
    print(f"mean |d| {err_before:.3f} -> mean |u - d| {err_after:.3f}", flush=True)
    assert err_after < 0.7 * err_before


@pytest.mark.parametrize("thresholds", [[0.05, 0.8], []], ids=["clipping", "none"])
def test_preprocessed_images_equal_the_host_restatement(tmp_path, thresholds):
    """The driver thresholds and rescales on the device; its
    reference-/target-preprocessed files must equal ThresholdImage and
    RescaleImage restated on the host, bit for bit (comparisons and the
    rescale map in double, one rounding per operation)."""
    from skmimageregistration_amd import read_image
    reference, target, _ = _case3d()
    rng = np.random.default_rng(8)
    # intensities off [0, 1] and with repeats, so that both bounds matter
    ref = (reference.data * 700.0 - 120.0 + rng.normal(size=reference.data.shape) * 3.0).astype(np.float32)
    tar = (target.data * 650.0 - 90.0).astype(np.float32)
    cfg = RegistrationConfig(sigma_C4=[16], sigma_C0=[16], tensor_weights=[0], num_scales=1,
                             num_function_evaluations=[2], initial_step_size=[0.5], sampling_rate=[2],
                             smoothing_variance=[0], metric="ncc",
                             intensity_thresholds=[t * 700.0 - 100.0 for t in thresholds])
    print(f"preprocessing thresholds={thresholds}", flush=True)
    register_multiscale(Image(ref), Image(tar), cfg, temp_directory=str(tmp_path))
    th = cfg["intensity_thresholds"]
    for name, img in (("reference", ref), ("target", tar)):
        want = threshold_image(img, th[1], th[0]) if len(th) == 2 else img
        want = rescale_image(want)
        got = read_image(str(tmp_path / f"{name}-preprocessed.vtk")).data
        np.testing.assert_array_equal(got, want)


def test_exact_gradient_option_reaches_every_level():
    """exact_gradient (no SKMReg counterpart) gives the level's evaluator the
    exact-gradient windows: the LCC registration starts from the same value
    (the windows change only the derivative), then moves differently from the
    reference-window one, and keeps its best iterate."""
    mov, fix, _ = synthetic.sliding_3d((24, 32, 32), shift=2.0, breathing=2.0)
    base = dict(sigma_C4=[12.0], sigma_C0=[12.0], tensor_weights=[0.0], num_scales=1, num_function_evaluations=[5],
                initial_step_size=[1000.0], sampling_rate=[2], neighborhood_sampling=[4], smoothing_variance=[0],
                reg_rd=[1e-3], metric="lcc", intensity_thresholds=[])
    df_ref, rep_ref = register_multiscale(Image(mov), Image(fix), RegistrationConfig(**base))
    df_x, rep_x = register_multiscale(Image(mov), Image(fix), RegistrationConfig(exact_gradient=[1], **base))
    assert np.isfinite(df_x).all() and rep_x[0]["first_value"] == rep_ref[0]["first_value"]
    assert not np.array_equal(df_x, df_ref)
    assert rep_x[0]["value"] <= rep_x[0]["first_value"]
