import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a HIP GPU (MI355X) and libscopyon_hip.so')
    config.addinivalue_line('markers', 'slow: CPU test that takes tens of seconds')


@pytest.fixture(scope='session')
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    from scopyon_amd import _lib
    _lib.lib()
    return torch.device('cuda', 0)


def oracle_params(overrides=None, rng=None):
    """Oracle parameter set from the default config plus nested overrides."""
    import copy
    from oracle import params as P
    cfg = P.default_config()

    def merge(d, o):
        for k, v in o.items():
            if isinstance(v, dict) and isinstance(d.get(k), dict):
                merge(d[k], v)
            else:
                d[k] = copy.deepcopy(v)
    if overrides:
        merge(cfg, overrides)
    return P.epifm_params(cfg, rng)


def product_config(overrides=None):
    import scopyon_amd
    cfg = scopyon_amd.DefaultConfiguration()
    if overrides:
        cfg.update({'default': overrides})
    return cfg


PRE = {'scale': 1.0, 'origin': [0.0, 0.0, 0.0], 'unit_x': [1.0, 0.0, 0.0],
       'unit_y': [0.0, 1.0, 0.0]}
