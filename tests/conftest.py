import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


def load_pkg():
    """Loads hpm-mvs_amd/ as the package hpm_mvs_amd and returns its bindings."""
    name = "hpm_mvs_amd"
    if name not in sys.modules:
        d = os.path.join(ROOT, "hpm-mvs_amd")
        spec = importlib.util.spec_from_file_location(name, os.path.join(d, "__init__.py"),
                                                      submodule_search_locations=[d])
        mod = importlib.util.module_from_spec(spec)
        sys.modules[name] = mod
        spec.loader.exec_module(mod)
    from hpm_mvs_amd import hpm

    return hpm


@pytest.fixture(scope="session")
def hpm():
    return load_pkg()


@pytest.fixture(scope="session")
def cpu():
    from oracle import cpu as oracle_cpu

    return oracle_cpu


def problem_inputs(hpm, cfg, ref):
    """Renders a scene and orders its views as a problem: the reference
    first, then the others nearest first (the pair.txt order)."""
    cams, imgs, gd, gn, low = hpm.render_scene(cfg)
    n = cfg.num_views
    order = [ref] + sorted([i for i in range(n) if i != ref], key=lambda i: (abs(i - ref), i))
    return ([cams[i] for i in order], [imgs[i].astype(np.float32) for i in order], gd, gn, low, order)


@pytest.fixture(scope="session")
def c1_small(hpm):
    """A small version of the 640x480 4-plane workload: 160x120, 1 ref + 4 src."""
    cfg = hpm.scene_config(seed=11, width=160, height=120, fx=125.0, fy=125.0, cx=80.0, cy=60.0)
    return cfg, problem_inputs(hpm, cfg, 2)


@pytest.fixture(scope="session")
def c1_full(hpm):
    cfg = hpm.workload_config("c1")
    return cfg, problem_inputs(hpm, cfg, 2)
