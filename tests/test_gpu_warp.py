"""GPU parity of the warped outputs (skm_warp_image) against
oracle/warp_oracle.py, and SKMReg's file flow (register_files: inputs read
from VTK files, every result file written into temp_directory, restart from
df-full) checked file by file against the oracle and the returned field.

Tolerances: the warp computes positions and weights in double from the same
float inputs as the oracle and stores float, so values agree to float
rounding (1e-6 of the image range).  The Jacobian determinant is computed in
double from the float control-point image (1e-9 relative).
"""
import json
import os

import numpy as np
import pytest

import warp_oracle as wo
from skmimageregistration_amd import (Grid, Image, RegistrationConfig, read_image, register_files,
                                      register_multiscale, synthetic, warp_image, write_image)
from skmimageregistration_amd._capi import SkmError

pytestmark = pytest.mark.gpu


def _smooth_field(shape, dim, seed, amp):
    rng = np.random.default_rng(seed)
    from scipy import ndimage
    f = np.stack([ndimage.gaussian_filter(rng.normal(size=shape), 2.0) for _ in range(dim)], axis=-1)
    return (amp * f / np.abs(f).max()).astype(np.float32)


@pytest.mark.parametrize("dim", [2, 3])
def test_linear_warp_matches_oracle(dim):
    shape_i = (20, 22, 24)[:dim] if dim == 3 else (30, 34)
    rng = np.random.default_rng(dim)
    img = rng.uniform(0, 1, size=shape_i).astype(np.float32)
    gi = ((0.5, -1.0, 2.0)[:dim], (1.0, 1.2, 0.9)[:dim])
    shape_d = tuple(n // 2 + 3 for n in shape_i)
    gd = ((-1.0, 0.0, 1.5)[:dim], (2.1, 2.3, 1.9)[:dim])
    df = _smooth_field(shape_d, dim, 10 + dim, 3.0)
    shape_o = tuple(n + 2 for n in shape_i)
    go = ((0.0, -2.0, 1.0)[:dim], (1.1, 0.95, 1.0)[:dim])
    print(f"linear warp dim={dim}", flush=True)
    out = warp_image(Image(img, *gi), Image(df, *gd, dim=dim), "linear",
                     out_grid=Grid(tuple(reversed(shape_o)), *go), default_value=-1.0)
    ref = wo.warp_linear(img, gi, df, gd, shape_o, go, default=-1.0)
    assert out.data.shape == shape_o
    np.testing.assert_allclose(out.data, ref, rtol=0, atol=1e-6)
    assert (ref == -1.0).any() and (ref != -1.0).sum() > 0.5 * ref.size


@pytest.mark.parametrize("dim", [2, 3])
def test_bspline3_warp_matches_oracle(dim):
    shape_i = (18, 20, 22)[:dim] if dim == 3 else (26, 30)
    rng = np.random.default_rng(20 + dim)
    img = rng.uniform(0, 1, size=shape_i).astype(np.float32)
    gi = ((0.0,) * dim, (1.0, 1.1, 0.8)[:dim])
    shape_d = tuple(n + 1 for n in shape_i)
    gd = ((-0.5, 0.25, 0.0)[:dim], (1.0, 1.0, 0.9)[:dim])
    df = _smooth_field(shape_d, dim, 30 + dim, 2.5)
    print(f"bspline3 warp dim={dim}", flush=True)
    out = warp_image(Image(img, *gi), Image(df, *gd, dim=dim), "bspline3")
    ref = wo.warp_bspline3(img, gi, df, gd)
    np.testing.assert_allclose(out.data, ref, rtol=0, atol=1e-6)
    with pytest.raises(SkmError):
        warp_image(Image(img, *gi), Image(df, *gd, dim=dim), "bspline3",
                   out_grid=Grid(tuple(reversed(shape_i)), *gi))


def _write_inputs(tmp, reference, target, disp):
    write_image(os.path.join(tmp, "ref.vtk"), reference)
    write_image(os.path.join(tmp, "tar.vtk"), target)
    # landmarks: target points and where the known warp sends them
    rng = np.random.default_rng(3)
    n = 6
    pts = np.stack([rng.uniform(8, s - 8, size=n) for s in reversed(target.data.shape)], axis=-1)
    idx = np.round(pts).astype(int)
    moved = pts + disp[idx[:, 2], idx[:, 1], idx[:, 0]]
    for name, p in (("tar_lm.vtk", pts), ("ref_lm.vtk", moved)):
        with open(os.path.join(tmp, name), "w") as f:
            f.write(f"# vtk DataFile Version 3.0\nvtk output\nASCII\nDATASET POLYDATA\nPOINTS {n} float\n")
            f.write("\n".join(" ".join(f"{v:.6f}" for v in q) for q in p) + "\n")
    return moved, pts


def test_register_files_writes_skmreg_results(tmp_path):
    mov, fix, disp = synthetic.ellipsoids_3d((24, 28, 32), n=8, smooth=1.5, max_disp=2.0, noise=0.01,
                                             seeds=(5, 6, 7))
    tmp = str(tmp_path)
    out_dir = os.path.join(tmp, "out")
    ref_lm, tar_lm = _write_inputs(tmp, Image(mov), Image(fix), disp)
    cfg_file = os.path.join(tmp, "config.json")
    json.dump({"reference_filename": os.path.join(tmp, "ref.vtk"), "target_filename": os.path.join(tmp, "tar.vtk"),
               "reference_landmarks": os.path.join(tmp, "ref_lm.vtk"),
               "target_landmarks": os.path.join(tmp, "tar_lm.vtk"), "temp_directory": out_dir,
               "sigma_C4": [12, 8], "sigma_C0": [12, 8], "tensor_weights": [200, 100], "num_scales": 2,
               "num_function_evaluations": [12, 8], "initial_step_size": [0.5], "sampling_rate": [2, 1],
               "neighborhood_sampling": [1], "smoothing_variance": [1, 0], "metric": "ncc", "verbosity": 2},
              open(cfg_file, "w"))
    cfg = RegistrationConfig.from_json(cfg_file)
    print("register_files 3D 32x28x24, 2 levels", flush=True)
    df, reports = register_files(cfg)
    names = ["reference-preprocessed", "target-preprocessed", "df", "warped"] + \
        [f"{k}-{s}" for s in (0, 1) for k in ("covariance_tensor", "df-full", "cp_image", "cp_jacdet",
                                              "smoothed_reference", "smoothed_target", "df", "warped")]
    for n in names:
        assert os.path.exists(os.path.join(out_dir, n + ".vtk")), n
    geom = ((0.0,) * 3, (1.0,) * 3)
    dfile = read_image(os.path.join(out_dir, "df.vtk"))
    np.testing.assert_array_equal(dfile.data, df)
    np.testing.assert_array_equal(read_image(os.path.join(out_dir, "df-full-1.vtk")).data, df)
    ref_pre = read_image(os.path.join(out_dir, "reference-preprocessed.vtk")).data
    assert ref_pre.min() == 0.0 and ref_pre.max() == 1.0
    warped = read_image(os.path.join(out_dir, "warped.vtk")).data
    np.testing.assert_allclose(warped, wo.warp_linear(ref_pre, geom, df, geom, df.shape[:3], geom),
                               rtol=0, atol=1e-6)
    warped1 = read_image(os.path.join(out_dir, "warped-1.vtk")).data
    np.testing.assert_allclose(warped1, wo.warp_bspline3(ref_pre, geom, df, geom), rtol=0, atol=1e-6)
    for s in (0, 1):
        cpi = read_image(os.path.join(out_dir, f"cp_image-{s}.vtk"))
        jac = read_image(os.path.join(out_dir, f"cp_jacdet-{s}.vtk")).data
        np.testing.assert_allclose(jac, wo.jacobian_determinant(cpi.data, cpi.spacing), rtol=1e-6, atol=1e-6)
        dfs = read_image(os.path.join(out_dir, f"df-full-{s}.vtk")).data
        assert reports[s]["tre"] == pytest.approx(wo.tre(dfs, geom, ref_lm, tar_lm), rel=1e-5)
        zeros = 100.0 * np.mean(cpi.data == 0.0)
        assert reports[s]["zero_parameters"] == pytest.approx(zeros)
    # the level's own field plus the coarser levels' field is df-full
    df0 = read_image(os.path.join(out_dir, "df-full-0.vtk")).data
    own1 = read_image(os.path.join(out_dir, "df-1.vtk")).data
    np.testing.assert_allclose(own1 + df0, df, rtol=0, atol=1e-5 * max(1.0, np.abs(df).max()))
    assert all(np.isfinite(r["tre"]) and r["tre"] >= 0.0 for r in reports)

    # restart at level 2 from the df-full-0.vtk just written
    print("restart_at 2 from df-full-0.vtk", flush=True)
    cfg2 = RegistrationConfig.from_json(cfg_file)
    cfg2.values["restart_at"] = [2]
    df_r, reps_r = register_files(cfg2)
    assert reps_r[0]["iterations"] == 0 and reps_r[1]["iterations"] == 8
    df_m, _ = register_multiscale(read_image(os.path.join(tmp, "ref.vtk")), read_image(os.path.join(tmp, "tar.vtk")),
                                  cfg2, restart_displacement=df0)
    np.testing.assert_array_equal(df_r, df_m)
    np.testing.assert_allclose(df_r, df, rtol=0, atol=1e-6 * max(1.0, np.abs(df).max()))
