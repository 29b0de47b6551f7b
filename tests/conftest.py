import dataclasses
import importlib.util
import os
import shutil
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "timbir-python_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def load_pkg():
    if "timbir_python_amd" in sys.modules:
        return sys.modules["timbir_python_amd"]
    spec = importlib.util.spec_from_file_location(
        "timbir_python_amd", os.path.join(PKG_DIR, "__init__.py"),
        submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["timbir_python_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


load_pkg()
from timbir_python_amd import synth  # noqa: E402
import py_oracle  # noqa: E402

# small scans the oracle finishes in seconds
TINY = dataclasses.replace(synth.C1, name="tiny", N_xy=32, N_z=4, N_p=48, iterations=8,
                           num_threads=2, seed=5)
SMALL = dataclasses.replace(synth.C1, name="small", N_xy=64, N_z=8, N_p=96, iterations=20,
                            num_threads=4, seed=11)
SMALL4D = dataclasses.replace(synth.C3, name="small4d", N_xy=32, N_z=4, N_p=256, N_time=4,
                              interlaced_K=8, interlaced_N_theta=64, iterations=10,
                              num_threads=4, seed=13)


@pytest.fixture(scope="session")
def oracle32():
    return py_oracle.Oracle(False)


@pytest.fixture(scope="session")
def oracle64():
    return py_oracle.Oracle(True)


@pytest.fixture(scope="session")
def oracle64a():
    """FP64 arithmetic on the float build's system matrix."""
    return py_oracle.Oracle(True, float_amatrix=True)


_scan_cache = {}


def make_scan(cfg, tmp_root, tag=""):
    """Generate cfg into a fresh folder (cached per session) and return a copy path."""
    key = (cfg.name, cfg.seed, cfg.N_xy, cfg.N_z, cfg.N_p, cfg.N_time)
    if key not in _scan_cache:
        src = os.path.join(str(tmp_root), "scan_" + "_".join(str(k) for k in key))
        o64 = py_oracle.Oracle(True, float_amatrix=True)
        info = synth.generate(cfg, src, o64.projector(src, cfg.cli_args()))
        _scan_cache[key] = (src, info)
    src, info = _scan_cache[key]
    dst = os.path.join(str(tmp_root), f"run_{cfg.name}_{cfg.seed}_{tag}_{np.random.randint(1 << 30)}")
    shutil.copytree(src, dst)
    return dst, info


def as_double(path):
    """Copy of a run folder with every Real_t file widened to float64 -- what the
    reference's -DREAL_IS_DOUBLE build (and the f64/f64a oracles) read."""
    dst = path.rstrip("/") + "_f64"
    if os.path.isdir(dst):
        shutil.rmtree(dst)
    shutil.copytree(path, dst)
    for f in os.listdir(dst):
        if f.endswith(".bin"):
            a = np.fromfile(os.path.join(dst, f), np.float32)
            a.astype(np.float64).tofile(os.path.join(dst, f))
    return dst


def oracle_path(path, oracle):
    return as_double(path) if oracle.double else path


@pytest.fixture(scope="session")
def scan_root(tmp_path_factory):
    return tmp_path_factory.mktemp("scans")
