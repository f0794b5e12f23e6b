"""The reconstruction programs (laminocam_amd/bin/: forward_scalar,
recon_lamino, recon_tomo; src/radon.cpp, src/recon.cpp, src/mbir2d.cpp of the
reference) end to end on the GPU, each against the same computation made
through the Python API on the same inputs, bitwise:

- forward_scalar: a TIFF volume padded by pad_factor, projected at two tilts
  over generated angles, cropped back: against laminocam_amd.forward on the
  padded volume.
- recon_lamino: forward_scalar's TIFF stack and angles file through a TOML
  configuration (Split-Bregman): against laminocam_amd.config +
  drivers.MBIR.
- recon_tomo: an HDF5 file (/images, /coords/alpha in degrees), CG and
  Split-Bregman: against config + prepare_sinogram + drivers.MBIR2D.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu

PROGS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "laminocam_amd", "bin")


def _prog(name, cfg_path, cwd):
    exe = os.path.join(PROGS, name)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (run __graft_entry__.build())")
    r = subprocess.run([exe, str(cfg_path)], capture_output=True, text=True, timeout=240, cwd=cwd)
    print(r.stdout[-2000:])
    print(r.stderr[-2000:])
    assert r.returncode == 0, f"{name} exited with {r.returncode}"
    return r


def _volume(shape, seed):
    from laminocam_amd import synth
    return synth.ellipsoid_phantom(shape, 8, synth.Stream(seed)).astype(np.float32)


def _forward_scalar_expected(lc, vol, pad_factor, amin, amax, n, gamma_deg):
    """radon.cpp's arithmetic: float32 angles, double-precision radians rounded to float"""
    step = (np.float32(amax) - np.float32(amin)) / np.float32(n)
    ang = np.float32(amin) + np.arange(n, dtype=np.float32) * step
    theta = (ang.astype(np.float64) * np.pi / 180.0).astype(np.float32)
    gamma = np.float32(np.float64(np.float32(gamma_deg)) * np.pi / 180.0)

    def pad(d):
        m = int(np.float32(d) * np.float32(pad_factor))
        return m - 1 if m % 2 == 0 else m
    pd = tuple(pad(d) for d in vol.shape)
    big = np.zeros(pd, np.float32)
    o = [(p - d) // 2 for p, d in zip(pd, vol.shape)]
    big[o[0]:o[0] + vol.shape[0], o[1]:o[1] + vol.shape[1], o[2]:o[2] + vol.shape[2]] = vol
    g = lc.PolarGrid(theta.astype(np.float64), np.full(n, gamma, np.float64), np.zeros(n), pd[1], pd[2])
    proj = lc.forward(big, g)
    c1, c2 = (pd[1] - vol.shape[1]) // 2, (pd[2] - vol.shape[2]) // 2
    return ang, proj[:, c1:c1 + vol.shape[1], c2:c2 + vol.shape[2]]


def test_forward_scalar_then_recon_lamino(tmp_path):
    import laminocam_amd as lc
    from laminocam_amd import config as cfg
    from laminocam_amd import drivers
    from laminocam_amd import io as lio
    announce("forward_scalar 9x45x45 pad 1.5, 36 angles, gamma 0 and 30; recon_lamino SB")
    vol = _volume((9, 45, 45), 31)
    lio.tiff_write(str(tmp_path / "volume.tiff"), vol)
    (tmp_path / "fwd.toml").write_text(
        '[input]\nfilename = "volume.tiff"\n'
        '[output]\nbasedir = "out"\npad_factor = 1.5\nproj_angles = { min = -90, max = 90, n = 36 }\n'
        '[[projections]]\ngamma = 0\nfilename = "g0.tiff"\n'
        '[[projections]]\ngamma = 30\nfilename = "g30.tiff"\n')
    _prog("forward_scalar", tmp_path / "fwd.toml", tmp_path)
    for gdeg, name in ((0.0, "g0.tiff"), (30.0, "g30.tiff")):
        ang, exp = _forward_scalar_expected(lc, vol, 1.5, -90, 90, 36, gdeg)
        got = lio.tiff_read(str(tmp_path / "out" / name))
        assert got.shape == exp.shape == (36, 45, 45)
        assert np.array_equal(got, exp), f"gamma {gdeg}: max |diff| {np.abs(got - exp).max():.3e}"
    txt = np.loadtxt(str(tmp_path / "out" / "angles.txt"))
    assert np.allclose(txt, ang, atol=1e-4)

    (tmp_path / "lam.toml").write_text(
        '[[input]]\nfilename = "out/g30.tiff"\nangles = "out/angles.txt"\ngamma = -30\n'
        '[output]\nfilename = "rec/recon.tiff"\n'
        '[recon_params]\nmax_iters = 6\nrecon_dims = [9, 45, 45]\n'
        '[recon_params.regularizer]\nmethod = "split_bregman"\n'
        '[recon_params.regularizer.split_bregman]\nlambda = 0.05\nmu = 10.0\ninner_iters = 2\n')
    r = _prog("recon_lamino", tmp_path / "lam.toml", tmp_path)
    assert "Saved reconstruction to rec/recon.tiff" in r.stdout
    got = lio.tiff_read(str(tmp_path / "rec" / "recon.tiff"))
    conf = cfg.read_toml_file(str(tmp_path / "lam.toml"))
    ds = cfg.parse_input_datasets(conf, str(tmp_path))
    par = cfg.parse_recon_params(conf)
    exp = drivers.MBIR(ds, par)
    assert got.shape == exp.shape == (9, 45, 45)
    assert np.array_equal(got, exp), f"max |diff| {np.abs(got - exp).max():.3e}"
    # the reconstruction resembles the volume the projections were made from
    c = float(np.sum(got * vol) / (np.linalg.norm(got) * np.linalg.norm(vol)))
    print(f"cosine similarity to the volume {c:.3f}")
    assert c > 0.5


@pytest.mark.parametrize("solver", ["cg", "sb"])
def test_recon_tomo_from_hdf5(tmp_path, solver):
    import h5write
    import laminocam_amd as lc
    from laminocam_amd import config as cfg
    from laminocam_amd import drivers
    from laminocam_amd import io as lio
    announce(f"recon_tomo HDF5 5 rows x 61 cols, 90 angles, {solver}")
    from laminocam_amd import synth
    rng = synth.Stream(32)
    ang_deg = np.arange(90) * 2.0
    g = lc.PolarGrid2D(np.deg2rad(ang_deg), 61)
    vol = np.stack([synth.ellipse_phantom(61, 6, rng) for _ in range(5)]).astype(np.float32)
    sino = lc.forward2d(vol, g)                       # {rows, angles, cols}
    projs = np.ascontiguousarray(sino.transpose(1, 0, 2))   # {angles, rows, cols}
    h5write.write(str(tmp_path / "scan.h5"), {"images": projs, "coords/alpha": ang_deg})
    reg = ('[recon_params.regularizer]\nmethod = "split_bregman"\n'
           '[recon_params.regularizer.split_bregman]\nlambda = 0.05\nmu = 10.0\n') if solver == "sb" else ""
    (tmp_path / "tomo.toml").write_text(
        '[[input]]\nfilename = "scan.h5"\n'
        '[output]\nfilename = "recon2d.tiff"\n'
        f'[recon_params]\nmax_iters = 12\nrecon_dims = [5, 61, 61]\n{reg}')
    _prog("recon_tomo", tmp_path / "tomo.toml", tmp_path)
    got = lio.tiff_read(str(tmp_path / "recon2d.tiff"))
    conf = cfg.read_toml_file(str(tmp_path / "tomo.toml"))
    ds = cfg.parse_input_datasets(conf, str(tmp_path))
    par = cfg.parse_recon_params(conf)
    s = drivers.prepare_sinogram(ds)
    exp = drivers.MBIR2D(s.data, s.theta, par)
    assert got.shape == exp.shape == (5, 61, 61)
    assert np.array_equal(got, exp), f"max |diff| {np.abs(got - exp).max():.3e}"
    # the radial samples cover the disk of k-space and the ellipses' sharp
    # edges reach past it: CG from zeros gives about 0.73 here, as the FP64
    # oracle's CG does on the same data
    c = float(np.sum(got * vol) / (np.linalg.norm(got) * np.linalg.norm(vol)))
    print(f"cosine similarity to the phantom {c:.3f}")
    assert c > 0.6
