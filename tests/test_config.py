"""The configuration layer of the reconstruction programs, on the CPU.

- include/laminocam_toml.hpp (the C++ TOML reader standing in for toml++)
  against tomli, an independent TOML 1.0 parser: the same tables, keys,
  types and values for a corpus of documents (the reference's own
  configuration files when its tree is present), and both reject the same
  malformed documents.
- include/laminocam_io.hpp's parse_input_datasets / parse_recon_params /
  parse_output_params (through tests/cxx/test_config.cpp) against
  laminocam_amd.config on the same configurations: TIFF and HDF5 inputs,
  angle units, cor-offset, shifts files, alignment files.
- the reference's rules (include/config.h): even recon_dims reduced by one,
  defaults, the error messages of malformed sections, example configurations.
"""
import json
import math
import os
import subprocess

import numpy as np
import pytest

from laminocam_amd import config as cfg
from laminocam_amd import io as lio

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cxx", "_bin")
PROGS = os.path.join(HERE, "..", "laminocam_amd", "bin")
REF = "/root/reference"


def _run(prog, *args, cwd=None):
    exe = os.path.join(BIN, prog)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (run __graft_entry__.build())")
    return subprocess.run([exe, *args], capture_output=True, text=True, timeout=60, cwd=cwd)


def _tag(v):
    """tomli's value in the tagged form test_toml prints"""
    import datetime
    if isinstance(v, dict):
        return {k: _tag(x) for k, x in v.items()}
    if isinstance(v, list):
        return [_tag(x) for x in v]
    if isinstance(v, bool):
        return {"type": "bool", "value": "true" if v else "false"}
    if isinstance(v, int):
        return {"type": "integer", "value": str(v)}
    if isinstance(v, float):
        return {"type": "float", "value": v}
    if isinstance(v, str):
        return {"type": "string", "value": v}
    if isinstance(v, (datetime.date, datetime.time)):
        return {"type": "datetime"}
    raise TypeError(type(v))


def _norm(t):
    """C++ output: floats as numbers, datetimes without their text"""
    if isinstance(t, dict) and set(t) == {"type", "value"} and isinstance(t["type"], str):
        if t["type"] == "float":
            return {"type": "float", "value": float(t["value"])}
        if t["type"] == "datetime":
            return {"type": "datetime"}
        return t
    if isinstance(t, dict):
        return {k: _norm(x) for k, x in t.items()}
    if isinstance(t, list):
        return [_norm(x) for x in t]
    return t


def _same(a, b):
    if isinstance(a, dict) and isinstance(b, dict):
        return a.keys() == b.keys() and all(_same(a[k], b[k]) for k in a)
    if isinstance(a, list) and isinstance(b, list):
        return len(a) == len(b) and all(_same(x, y) for x, y in zip(a, b))
    if isinstance(a, float) and isinstance(b, float):
        return (math.isnan(a) and math.isnan(b)) or a == b
    return a == b


VALID = [
    'a = 1\nb = "two"\nc = 3.5\nd = true\n',
    '# comment\n[t]\nx = -7 # trailing\ny = +3\n[t.sub]\nz = "q"\n',
    '[[arr]]\nn = 1\n[[arr]]\nn = 2\n[arr.sub]\nk = 3\n[[arr]]\n',
    'dotted.key.here = 1\ndotted.key.there = 2\n"quoted key" = 3\n\'lit key\' = 4\n',
    'inl = { a = 1, b = { c = [1, 2] }, "d.e" = "x" }\nempty = {}\n',
    'arr = [ 1, 2,\n 3, # c\n ]\nmixed = [1, "a", 2.5, [true], {x = 1}]\nnested = [[1, 2], [3]]\n',
    'hex = 0xDEAD_beef\noct = 0o755\nbin = 0b1101\nbig = 9_223_372_036_854_775_807\nneg = -9223372036854775808\n',
    'f1 = 1e-5\nf2 = -0.0\nf3 = 6.02E+23\nf4 = 1_000.5\nf5 = inf\nf6 = -inf\nf7 = nan\nf8 = 5e2_2\n',
    's1 = "tab\\tnew\\nquote\\"back\\\\u\\u00e9U\\U0001F600"\ns2 = \'C:\\\\raw\\path\'\n',
    's3 = """\nline1\n  line2 \\\n   joined"""\ns4 = \'\'\'\nraw \\n text\n\'\'\'\ns5 = """ends with quotes"\"\"\"\n',
    'd1 = 1979-05-27\nd2 = 1979-05-27T07:32:00Z\nd3 = 1979-05-27 07:32:00.999+01:00\nt1 = 07:32:00\n',
    '[a.b.c]\nx = 1\n[a]\ny = 2\n',
    '[x]\n[x.y]\n[x.z]\nk = 1\n',
    'key = "value"   \r\nother = 1\r\n',
    '[[input]]\nfilename = "p.tiff"\nangles = "a.txt"\ngamma = 30.0\ncor-offset = [1.5, -2]\n'
    '[output]\nfilename = "o.tiff"\nformats = ["tiff", "vti", "tiff"]\n'
    '[recon_params]\nrecon_dims = [41, 380, 41]\n[recon_params.regularizer]\nmethod = "split_bregman"\n'
    '[recon_params.regularizer.split_bregman]\nlambda = 0.05\n',
]

INVALID = [
    'a = 1\na = 2\n',                       # duplicate key
    '[t]\n[t]\n',                            # table defined twice
    'a = \n',                                # missing value
    'a = "unterminated\n',
    'a = 01\n',                              # leading zero
    'a = 1__0\n',
    'a = [1, 2\n',
    'x = { a = 1 }\n[x]\n',                  # inline table reopened
    'a.b = 1\n[a.b]\n',
    'a = 1\n[a]\n',
    'arr = [1]\n[[arr]]\n',                  # static array extended as an array of tables
    'a = "bad \\q escape"\n',
    'k = 1 2\n',
    '[t\n',
    'a = 3.\n',
    'a = .5\n',
    'a = +0x10\n',
]


@pytest.mark.parametrize("i", range(len(VALID)))
def test_toml_reader_matches_tomli(tmp_path, i):
    import tomli
    p = tmp_path / "doc.toml"
    p.write_bytes(VALID[i].encode())
    r = _run("test_toml", str(p))
    assert r.returncode == 0, r.stdout + r.stderr
    mine = _norm(json.loads(r.stdout))
    ref = _tag(tomli.loads(VALID[i]))
    assert _same(mine, ref), f"\n{json.dumps(mine)}\n{json.dumps(ref, default=str)}"


@pytest.mark.parametrize("i", range(len(INVALID)))
def test_toml_reader_rejects_what_tomli_rejects(tmp_path, i):
    import tomli
    with pytest.raises(tomli.TOMLDecodeError):
        tomli.loads(INVALID[i])
    p = tmp_path / "bad.toml"
    p.write_bytes(INVALID[i].encode())
    r = _run("test_toml", str(p))
    assert r.returncode == 2 and r.stdout.startswith("ERROR:"), (INVALID[i], r.stdout)


REF_TOML = ["config_template_tomo.toml", "config_template_lamino.toml", "synthetic/config.toml",
            "cliff.toml"]


@pytest.mark.parametrize("name", REF_TOML)
def test_toml_reader_on_the_reference_configs(name):
    import tomli
    path = os.path.join(REF, name)
    if not os.path.exists(path):
        pytest.skip("the reference tree is not present")
    r = _run("test_toml", path)
    assert r.returncode == 0, r.stdout
    with open(path, "rb") as f:
        ref = _tag(tomli.load(f))
    assert _same(_norm(json.loads(r.stdout)), ref)


def test_example_configurations(tmp_path):
    # the programs write the same example configuration as the Python layer
    for prog, text in (("recon_lamino", cfg.LAMINO_CONFIG_TEMPLATE),
                       ("recon_tomo", cfg.TOMO_CONFIG_TEMPLATE)):
        exe = os.path.join(PROGS, prog)
        if not os.path.exists(exe):
            pytest.fail(f"{exe} is not built (run __graft_entry__.build())")
        r = subprocess.run([exe], capture_output=True, text=True, timeout=60, cwd=tmp_path)
        assert r.returncode == 1 and "Usage:" in r.stderr
        assert (tmp_path / "config_template.toml").read_text() == text
        c = cfg.read_toml_file(str(tmp_path / "config_template.toml"))
        p = cfg.parse_recon_params(c)
        assert p.regularizer == cfg.Regularizer.SPLIT_BREGMAN and p.maxIters == 50


# ---- configurations over real files ----

@pytest.fixture
def files(tmp_path):
    import h5write
    rng = np.random.default_rng(5)
    proj = rng.standard_normal((6, 8, 10)).astype(np.float32)
    proj[0, 1, 1] = np.nan
    proj[2, 3, 4] = np.inf
    lio.tiff_write(str(tmp_path / "projs.tiff"), proj)
    (tmp_path / "angles.txt").write_text("\n".join(str(a) for a in (-60, -30.5, 0, 15, 30.25, 60)) + "\n")
    (tmp_path / "shifts.txt").write_text("\n".join(f"{0.5 * i} {-0.25 * i}" for i in range(6)))
    h5write.write(str(tmp_path / "data.h5"), {"exchange/data": proj[:5],
                                              "exchange/theta": np.linspace(-1.0, 1.0, 5)})
    (tmp_path / "align.toml").write_text(
        "gamma_deg = 29.5\nbeta_deg = -0.25\ncorrected_angles_deg = [0, 61, 119, 181, 240, 299]\n"
        "shifts_px = [[1, 2], [3, 4], [5, 6], [7, 8], [9, 10], [11, 12]]\n")
    return tmp_path, proj


CONFIGS = {
    "tiff_cor": """
[[input]]
filename = "projs.tiff"
angles = "angles.txt"
gamma = 30.0
cor-offset = [1.5, -2]
[output]
filename = "out/recon.tiff"
formats = ["tiff", "vti", "tiff"]
[recon_params]
max_iters = 7
recon_dims = [9, 40, 41]
""",
    "tiff_shifts_rad_sb": """
[[input]]
filename = "projs.tiff"
angles = "angles.txt"
angle_units = "rad"
beta = 1
shifts = "shifts.txt"
[output]
[recon_params]
recon_dims = [3, 11, 11]
tol = 1e-3
num_gpus = 2
chunk_slices = 4
streamed = true
[recon_params.regularizer]
[recon_params.regularizer.split_bregman]
lambda = 0.25
mu = 5
inner_iters = 3
""",
    "h5_alignment_two_inputs": """
[[input]]
filename = "data.h5"
projs_dataset = "/exchange/data"
angles = "/exchange/theta"
angle_units = "rad"
[[input]]
filename = "projs.tiff"
angles = "angles.txt"
alignment = "align.toml"
[output]
filename = "r.tiff"
[recon_params]
recon_dims = [5, 21, 21]
""",
}


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_cxx_and_python_read_a_configuration_alike(files, name):
    d, proj = files
    p = d / f"{name}.toml"
    p.write_text(CONFIGS[name])
    r = _run("test_config", str(p))
    assert r.returncode == 0, r.stdout + r.stderr
    c = json.loads(r.stdout)
    conf = cfg.read_toml_file(str(p))
    ds = cfg.parse_input_datasets(conf, str(d))
    par = cfg.parse_recon_params(conf)
    out = cfg.parse_output_params(conf)
    assert len(ds) == len(c["datasets"])
    for x, y in zip(ds, c["datasets"]):
        assert list(x.projs.shape) == y["shape"]
        assert np.all(np.isfinite(x.projs))
        assert math.isclose(float(np.sum(x.projs, dtype=np.float64)), y["sum"], rel_tol=1e-9, abs_tol=1e-9)
        assert np.array_equal(x.angles, np.asarray(y["angles"], np.float32))
        assert np.float32(x.gamma) == np.float32(y["gamma"]) and np.float32(x.beta) == np.float32(y["beta"])
        assert np.array_equal(x.shifts, np.asarray(y["shifts"], np.float32).reshape(-1, 2))
    q = c["params"]
    assert list(par.recon_dims) == q["recon_dims"] and par.maxIters == q["maxIters"]
    assert int(par.regularizer) == q["regularizer"] and par.innerIters == q["innerIters"]
    for k, v in (("lam", "lambda"), ("mu", "mu"), ("tol", "tol"), ("xtol", "xtol")):
        assert np.float32(getattr(par, k)) == np.float32(q[v])
    assert (par.numGpus, par.chunkSlices, par.streamed) == (q["numGpus"], q["chunkSlices"], q["streamed"])
    assert out.filepath == c["output"]["filepath"] and out.formats == c["output"]["formats"]


def test_reference_rules(files):
    d, proj = files
    conf = cfg.read_toml_file(str(_write(d, CONFIGS["tiff_cor"])))
    ds = cfg.parse_input_datasets(conf, str(d))[0]
    # TIFF pages read with even sizes made odd; non-finite values zeroed
    assert ds.projs.shape == (6, 7, 9) and ds.projs[0, 1, 1] == 0 and ds.projs[2, 3, 4] == 0
    assert np.array_equal(ds.projs[1], proj[1, :7, :9])
    # degrees to radians in float32; gamma's sign flipped; cor-offset broadcast
    assert np.array_equal(ds.angles, (np.float32([-60, -30.5, 0, 15, 30.25, 60]) * np.float32(math.pi))
                          / np.float32(180))
    assert np.float32(ds.gamma) == (np.float32(-30.0) * np.float32(math.pi)) / np.float32(180)
    assert ds.shifts.shape == (6, 2) and np.all(ds.shifts == [1.5, -2])
    par = cfg.parse_recon_params(conf)
    assert par.recon_dims == (9, 39, 41) and par.regularizer == cfg.Regularizer.UNCONSTRAINED
    assert math.isclose(par.tol, 1e-5, rel_tol=1e-6)
    out = cfg.parse_output_params(conf)
    assert out.formats == ["tiff", "vti"] and out.has_format("vti")
    # alignment overrides angles, tilts and shifts
    conf = cfg.read_toml_file(str(_write(d, CONFIGS["h5_alignment_two_inputs"])))
    h5, al = cfg.parse_input_datasets(conf, str(d))
    assert h5.projs.shape == (5, 8, 10) and np.array_equal(h5.angles, np.linspace(-1, 1, 5).astype(np.float32))
    assert np.float32(al.gamma) == (np.float32(-29.5) * np.float32(math.pi)) / np.float32(180)
    assert np.array_equal(al.shifts[2], [5, 6])


def _write(d, text, name="c.toml"):
    p = d / name
    p.write_text(text)
    return p


@pytest.mark.parametrize("text,msg", [
    ("[output]\n[recon_params]\nrecon_dims=[1,1,1]\n", "Missing \\[\\[input\\]\\] array"),
    ("[[input]]\nangles='a'\n", "must have a 'filename'"),
    ("[[input]]\nfilename='nothere.tiff'\n", "Projection file does not exist"),
    ("[[input]]\nfilename='projs.tiff'\n", "'angles' is required for TIFF"),
    ("[[input]]\nfilename='projs.tiff'\nangles='angles.txt'\nangle_units='grad'\n", "must be \"deg\" or \"rad\""),
    ("[[input]]\nfilename='projs.tiff'\nangles='angles.txt'\ncor-offset=[1]\n", "must be in form \\[dx, dy\\]"),
    ("[[input]]\nfilename='projs.tiff'\nangles='shifts.txt'\n", "has 6 pages but"),
    ("[[input]]\nfilename='angles.txt'\n", "Unsupported file extension '.txt'"),
])
def test_input_errors(files, text, msg):
    d, _ = files
    conf = cfg.read_toml_file(str(_write(d, text)))
    with pytest.raises(RuntimeError, match=msg):
        cfg.parse_input_datasets(conf, str(d))
    r = _run("test_config", str(_write(d, text + "[output]\n[recon_params]\nrecon_dims=[1,1,1]\n"
                                       if "[output]" not in text else text)))
    assert r.returncode == 2


@pytest.mark.parametrize("text,msg", [
    ("x = 1\n", "Missing \\[recon_params\\]"),
    ("[recon_params]\n", "'recon_dims' must be an array of three"),
    ("[recon_params]\nrecon_dims = [3, 0, 5]\n", "'recon_dims\\[1\\]' must be a positive"),
    ("[recon_params]\nrecon_dims = [3, 3, 3]\n[recon_params.regularizer]\nmethod = 'tv'\n", "must be 'split_bregman'"),
    ("[recon_params]\nrecon_dims = [3, 3, 3]\n[recon_params.regularizer]\n", "Missing \\[recon_params.regularizer.split_bregman\\]"),
])
def test_recon_params_errors(tmp_path, text, msg):
    conf = cfg.read_toml_file(str(_write(tmp_path, text)))
    with pytest.raises(RuntimeError, match=msg):
        cfg.parse_recon_params(conf)


def test_output_errors(tmp_path):
    with pytest.raises(RuntimeError, match="Missing \\[output\\]"):
        cfg.parse_output_params({})
    with pytest.raises(RuntimeError, match="invalid format 'png'"):
        cfg.parse_output_params({"output": {"formats": ["png"]}})
    assert cfg.parse_output_params({"output": {}}).filepath == "./recon.tiff"
    with pytest.raises(RuntimeError, match="Failed to parse TOML file"):
        cfg.read_toml_file(str(_write(tmp_path, "a = = 1\n")))


@pytest.mark.parametrize("extra,msg", [
    ("gamma = 30.0\n", "only supports vertical-axis tomography"),
    ("", "must be square in-plane"),
])
def test_recon_tomo_rejects_before_the_device(files, extra, msg):
    # recon_tomo's checks (src/mbir2d.cpp:119-154) run before any device work
    d, _ = files
    dims = "[5, 21, 21]" if extra else "[5, 21, 23]"
    p = _write(d, f"[[input]]\nfilename = 'projs.tiff'\nangles = 'angles.txt'\n{extra}"
                  f"[output]\n[recon_params]\nrecon_dims = {dims}\n")
    exe = os.path.join(PROGS, "recon_tomo")
    r = subprocess.run([exe, str(p)], capture_output=True, text=True, timeout=60, cwd=d)
    assert r.returncode == 1 and msg in r.stderr, r.stderr
