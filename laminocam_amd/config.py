"""The configuration layer of the reconstruction programs (include/config.h:45-489),
for Python callers; the C++ programs use include/laminocam_io.hpp, which has the
same keys, defaults, checks and messages.

  read_toml_file, resolve_path, warn_unknown_keys        config.h:45-89
  angles_to_radians, read_angles_file, read_shifts_file  config.h:91-141
  parse_input_datasets                                   config.h:160-331
  parse_recon_params, warn_laminography_dims             config.h:333-417
  parse_output_params, OutputParams, dump_config         config.h:419-489,
                                                         recon_params.h:84-93

Configurations are TOML ([[input]] datasets, [output], [recon_params]);
numbers are handled in float32, as the reference's parse_input_datasets<float>
and its float parameter fields do.  Errors raise RuntimeError with the
reference's messages.
"""
from __future__ import annotations

import math
import os
import sys
from dataclasses import dataclass, field

import numpy as np

from . import io as lio
from .drivers import Dataset, ReconParams, Regularizer

__all__ = ["read_toml_file", "resolve_path", "warn_unknown_keys", "angles_to_radians",
           "read_angles_file", "read_shifts_file", "parse_input_datasets", "parse_recon_params",
           "warn_laminography_dims", "OutputParams", "parse_output_params", "dump_config",
           "LAMINO_CONFIG_TEMPLATE", "TOMO_CONFIG_TEMPLATE"]

F32 = np.float32
PI32 = F32(math.pi)


def read_toml_file(filepath) -> dict:
    import tomli
    if not os.path.exists(filepath):
        raise RuntimeError(f"TOML file does not exist: {filepath}")
    try:
        with open(filepath, "rb") as f:
            return tomli.load(f)
    except tomli.TOMLDecodeError as e:
        raise RuntimeError(f"Failed to parse TOML file '{filepath}': {e}") from None


def resolve_path(p: str, base_dir) -> str:
    """A relative path is taken relative to the config file's directory when
    the file exists there, else as given."""
    if not os.path.isabs(p) and base_dir:
        cand = os.path.join(base_dir, p)
        if os.path.exists(cand):
            return cand
    return p


def warn_unknown_keys(tbl: dict, allowed, section: str) -> None:
    for k in tbl:
        if k not in allowed:
            print(f"\033[33mWarning\033[0m: unknown key '{k}' in {section} (ignored)", file=sys.stderr)


def angles_to_radians(angles, units: str):
    a = np.asarray(angles, dtype=F32)
    if units == "deg":
        return (a * PI32) / F32(180)
    if units != "rad":
        raise RuntimeError(f"[[input]] 'angle_units' must be \"deg\" or \"rad\", got '{units}'")
    return a.copy()


def _numbers(path):
    vals = []
    with open(path) as f:
        for tok in f.read().split():
            try:
                vals.append(F32(float(tok)))
            except ValueError:
                break   # the stream stops at the first token that is not a number
    return vals


def read_angles_file(filepath, units="deg"):
    """Angles, one per whitespace-separated number, in `units`; radians out (float32)."""
    if not os.path.exists(filepath):
        raise RuntimeError(f"Could not open angles file: {filepath}")
    a = _numbers(filepath)
    if not a:
        raise RuntimeError(f"No angles found in file: {filepath}")
    return angles_to_radians(a, units)


def read_shifts_file(filepath):
    """Pairs "dx dy" (pixels), {n, 2} float32."""
    if not os.path.exists(filepath):
        raise RuntimeError(f"Could not open shifts file: {filepath}")
    v = _numbers(filepath)
    v = v[: len(v) // 2 * 2]
    if not v:
        raise RuntimeError(f"No shifts found in file: {filepath}")
    return np.asarray(v, F32).reshape(-1, 2)


def _as_f32(v):
    """toml++ value<float>: integers and floats convert, anything else is absent"""
    if isinstance(v, bool) or not isinstance(v, (int, float)):
        return None
    return F32(v)


def _as_size(v):
    """value<size_t>: a non-negative integer (or a whole float) in range"""
    if isinstance(v, bool):
        return None
    if isinstance(v, float) and v.is_integer():
        v = int(v)
    if isinstance(v, int) and 0 <= v < 2 ** 64:
        return v
    return None


def _deg2rad(v):
    return (F32(v) * PI32) / F32(180)


def parse_input_datasets(config: dict, base_dir="") -> list:
    """The [[input]] datasets (include/config.h:160-331): TIFF stacks with an
    angles text file, or HDF5 ('projs_dataset', default /images; 'angles',
    default /coords/alpha); non-finite values become 0; gamma (degrees, sign
    flipped) and beta default to 0; shifts from a 'shifts' file or a
    broadcast 'cor-offset' pair; an 'alignment' TOML replaces gamma, beta,
    angles and shifts.  Returns Datasets with float32 projections, angles and
    {n, 2} shifts."""
    inputs = config.get("input")
    if not isinstance(inputs, list):
        raise RuntimeError("Missing [[input]] array in TOML file")
    out = []
    for t in inputs:
        if not isinstance(t, dict):
            raise RuntimeError("Invalid [[input]] entry")
        if "filename" not in t:
            raise RuntimeError("[[input]] entry must have a 'filename' field")
        warn_unknown_keys(t, ("filename", "angles", "angle_units", "gamma", "beta", "shifts",
                              "cor-offset", "alignment", "projs_dataset"), "[[input]]")
        filename = t["filename"]
        if not isinstance(filename, str):
            raise RuntimeError("[[input]] 'filename' must be a string")
        filename = resolve_path(filename, base_dir)
        units = t["angle_units"] if isinstance(t.get("angle_units"), str) else "deg"
        if not os.path.exists(filename):
            raise RuntimeError(f"Projection file does not exist: {filename}")
        g = _as_f32(t.get("gamma"))
        b = _as_f32(t.get("beta"))
        gamma = _deg2rad(-(g if g is not None else F32(0)))
        beta = _deg2rad(b if b is not None else F32(0))

        shifts = None
        if "shifts" in t:
            if not isinstance(t["shifts"], str):
                raise RuntimeError("[[input]] 'shifts' must be a string path")
            sp = resolve_path(t["shifts"], base_dir)
            if not os.path.exists(sp):
                raise RuntimeError(f"Shifts file does not exist: {sp}")
            shifts = read_shifts_file(sp)
        elif "cor-offset" in t:
            off = t["cor-offset"]
            if not isinstance(off, list) or len(off) != 2:
                raise RuntimeError("[[input]] 'cor-offset' must be in form [dx, dy]")
            vals = [_as_f32(v) for v in off]
            if any(v is None for v in vals):
                raise RuntimeError("[[input]] 'cor-offset' must be an array of numbers")
            shifts = np.asarray([vals], F32)

        ext = os.path.splitext(filename)[1].lower()
        if ext in (".tiff", ".tif"):
            ap = t.get("angles")
            if not isinstance(ap, str):
                raise RuntimeError("[[input]] 'angles' is required for TIFF files")
            ap = resolve_path(ap, base_dir)
            if not os.path.exists(ap):
                raise RuntimeError(f"Angles file does not exist: {ap}")
            projs = lio.tiff_read(filename)
            angles = read_angles_file(ap, units)
            if projs.shape[0] != len(angles):
                raise RuntimeError(f"'{filename}' has {projs.shape[0]} pages but '{ap}' has "
                                   f"{len(angles)} entries")
        elif ext in (".h5", ".hdf5"):
            ads = t["angles"] if isinstance(t.get("angles"), str) else "/coords/alpha"
            pds = t["projs_dataset"] if isinstance(t.get("projs_dataset"), str) else "/images"
            projs = lio.h5_read_images(filename, pds)
            angles = angles_to_radians(lio.h5_read_angles(filename, ads), units)
            if projs.shape[0] != len(angles):
                raise RuntimeError(f"'{filename}' images has {projs.shape[0]} slices but angles "
                                   f"dataset '{ads}' has {len(angles)} entries")
        else:
            raise RuntimeError(f"Unsupported file extension '{ext}': {filename}")
        projs = lio.mask_infs_nans(projs)
        n = len(angles)
        if shifts is None:
            shifts = np.zeros((n, 2), F32)
        elif len(shifts) == 1:
            shifts = np.repeat(shifts, n, axis=0)
        elif len(shifts) != n:
            raise RuntimeError(f"shifts file has {len(shifts)} entries but {n} projections were loaded")

        if "alignment" in t:
            if not isinstance(t["alignment"], str):
                raise RuntimeError("[[input]] 'alignment' must be a string path")
            al = resolve_path(t["alignment"], base_dir)
            if not os.path.exists(al):
                raise RuntimeError(f"Alignment file does not exist: {al}")
            at = read_toml_file(al)
            gv = _as_f32(at.get("gamma_deg"))
            if gv is not None:
                gamma = _deg2rad(-gv)
            bv = _as_f32(at.get("beta_deg"))
            if bv is not None:
                beta = _deg2rad(bv)
            ca = at.get("corrected_angles_deg")
            if not isinstance(ca, list):
                raise RuntimeError("Alignment TOML missing 'corrected_angles_deg'")
            vals = [_as_f32(v) for v in ca]
            if any(v is None for v in vals):
                raise RuntimeError("Alignment TOML 'corrected_angles_deg' must hold numbers")
            angles = (np.asarray(vals, F32) * PI32) / F32(180)
            if isinstance(at.get("shifts_px"), list):
                pairs = []
                for e in at["shifts_px"]:
                    if not isinstance(e, list) or len(e) < 2:
                        raise RuntimeError("Alignment TOML 'shifts_px' entries must be [dx, dy] pairs")
                    dx, dy = _as_f32(e[0]), _as_f32(e[1])
                    if dx is None or dy is None:
                        raise RuntimeError("Alignment TOML 'shifts_px' entries must be [dx, dy] pairs")
                    pairs.append((dx, dy))
                shifts = np.asarray(pairs, F32).reshape(-1, 2)
            if len(shifts) != len(angles):
                raise RuntimeError(f"Alignment TOML: shifts_px has {len(shifts)} entries but "
                                   f"corrected_angles_deg has {len(angles)}")
        out.append(Dataset(projs=projs, angles=angles, gamma=float(gamma), beta=float(beta),
                           shifts=shifts))
    return out


def parse_recon_params(config: dict) -> ReconParams:
    """[recon_params] (include/config.h:333-400): max_iters (50), recon_dims
    (three positive integers, even ones reduced by 1), tol and xtol (1e-5),
    num_gpus, chunk_slices, streamed, and [recon_params.regularizer] with
    method "split_bregman" and its lambda (0.1), mu (10) and inner_iters (1)."""
    recon = config.get("recon_params")
    if not isinstance(recon, dict):
        raise RuntimeError("Missing [recon_params] section in config file")
    warn_unknown_keys(recon, ("max_iters", "tol", "xtol", "recon_dims", "regularizer", "num_gpus",
                              "chunk_slices", "streamed"), "[recon_params]")
    p = ReconParams()

    def size(v, d):
        x = _as_size(v)
        return d if x is None else x

    def f32(v, d):
        x = _as_f32(v)
        return float(F32(d) if x is None else x)

    p.maxIters = size(recon.get("max_iters"), 50)
    dims = recon.get("recon_dims")
    if not isinstance(dims, list) or len(dims) != 3:
        raise RuntimeError("[recon_params] 'recon_dims' must be an array of three integers")
    rd = []
    for i, v in enumerate(dims):
        x = size(v, 0)
        if x == 0:
            raise RuntimeError(f"[recon_params] 'recon_dims[{i}]' must be a positive integer")
        rd.append(x - 1 if x % 2 == 0 else x)
    p.recon_dims = tuple(rd)
    p.tol = f32(recon.get("tol"), 1e-5)
    p.xtol = f32(recon.get("xtol"), 1e-5)
    p.numGpus = size(recon.get("num_gpus"), 0)
    p.chunkSlices = size(recon.get("chunk_slices"), 0)
    p.streamed = recon.get("streamed") if isinstance(recon.get("streamed"), bool) else False
    reg = recon.get("regularizer")
    if isinstance(reg, dict):
        warn_unknown_keys(reg, ("method", "split_bregman"), "[recon_params.regularizer]")
        method = reg["method"] if isinstance(reg.get("method"), str) else "split_bregman"
        if method != "split_bregman":
            raise RuntimeError("[recon_params] 'regularizer' must be 'split_bregman'")
        p.regularizer = Regularizer.SPLIT_BREGMAN
        sb = reg.get("split_bregman")
        if not isinstance(sb, dict):
            raise RuntimeError("Missing [recon_params.regularizer.split_bregman] section in config file")
        warn_unknown_keys(sb, ("lambda", "mu", "inner_iters"), "[recon_params.regularizer.split_bregman]")
        p.lam = f32(sb.get("lambda"), 0.1)
        p.mu = f32(sb.get("mu"), 10.0)
        p.innerIters = size(sb.get("inner_iters"), 1)
    return p


def warn_laminography_dims(p: ReconParams) -> None:
    ratio = F32(p.recon_dims[0]) / F32(p.recon_dims[2])
    if ratio > F32(0.15):
        print(f"\033[31mWarning\033[0m: recon_dims[0] / recon_dims[2] = {ratio:.2f} > 0.15.\n"
              "laminography thickness (recon_dims[0]) is expected to be much smaller than the "
              "in-plane dimensions (recon_dims[1], recon_dims[2]).", file=sys.stderr)


@dataclass
class OutputParams:
    filepath: str = "./recon.tiff"
    formats: list = field(default_factory=lambda: ["tiff"])

    def has_format(self, fmt: str) -> bool:
        return fmt in self.formats


def parse_output_params(config: dict) -> OutputParams:
    """[output] (include/config.h:419-475): filename (./recon.tiff) and formats
    ("tiff" and / or "vti"; default tiff; duplicates dropped in order)."""
    out = config.get("output")
    if out is None:
        raise RuntimeError("Missing [output] section in config file")
    if isinstance(out, dict):
        warn_unknown_keys(out, ("filename", "formats"), "[output]")
    else:
        out = {}
    fp = out["filename"] if isinstance(out.get("filename"), str) else "./recon.tiff"
    fa = out.get("formats")
    formats = []
    if isinstance(fa, list):
        for f in fa:
            if not isinstance(f, str):
                raise RuntimeError("[output] 'formats' array must contain strings")
            if f not in ("tiff", "vti"):
                raise RuntimeError(f"[output] invalid format '{f}'. Must be 'tiff' or 'vti'")
            formats.append(f)
    else:
        formats = ["tiff"]
    uniq = []
    for f in formats:
        if f not in uniq:
            uniq.append(f)
    return OutputParams(fp, uniq)


def dump_config(template: str, filepath) -> None:
    try:
        with open(filepath, "w") as f:
            f.write(template)
    except OSError:
        raise RuntimeError(f"Could not open file for writing: {filepath}") from None


LAMINO_CONFIG_TEMPLATE = """# recon_lamino: laminography reconstruction of one or more tilt series.
# Run: recon_lamino config.toml

[[input]]
filename = "projections.tiff"     # a .tif/.tiff stack (one page per projection) or .h5/.hdf5
angles = "angles.txt"             # TIFF: text file of angles; HDF5: angles dataset (default /coords/alpha)
angle_units = "deg"               # "deg" (default) or "rad"
gamma = 30.0                      # tilt of the rotation axis, degrees
# beta = 0.0                      # second tilt, degrees
# shifts = "shifts.txt"           # per-projection "dx dy" in pixels
# cor-offset = [0.0, 0.0]         # one shift for every projection
# alignment = "alignment.toml"    # gamma_deg, beta_deg, corrected_angles_deg, shifts_px
# projs_dataset = "/images"       # HDF5 projections dataset

[output]
filename = "recon.tiff"
formats = ["tiff"]

[recon_params]
max_iters = 50
tol = 1e-5
xtol = 1e-5
recon_dims = [41, 511, 511]       # (slices, y, x); even sizes are reduced by one

# omit this table for an unregularised least-squares solve
[recon_params.regularizer]
method = "split_bregman"

[recon_params.regularizer.split_bregman]
lambda = 0.1
mu = 10.0
inner_iters = 1
"""

TOMO_CONFIG_TEMPLATE = """# recon_tomo: slice-by-slice tomography about a vertical axis.
# Run: recon_tomo config.toml

[[input]]
filename = "projections.h5"       # .h5/.hdf5 or a .tif/.tiff stack
# angles = "angles.txt"           # TIFF: text file of angles (required)
# projs_dataset = "/images"       # HDF5 projections dataset (default)
# angles = "/coords/alpha"        # HDF5 angles dataset (default)
angle_units = "deg"               # "deg" (default) or "rad"
# gamma and beta must be 0 here; use recon_lamino for tilted axes

[output]
filename = "recon2d.tiff"
formats = ["tiff"]

[recon_params]
max_iters = 50
tol = 1e-5
xtol = 1e-5
recon_dims = [511, 511, 511]      # (slices, y, x); slices = detector rows; y = x

[recon_params.regularizer]
method = "split_bregman"

[recon_params.regularizer.split_bregman]
lambda = 0.1
mu = 10.0
inner_iters = 1
"""
