import sys
from pathlib import Path

import numpy as np
import pytest
import yaml

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long CPU oracle replay")


def golden_dir(material):
    return GOLDEN / material


def load_configs(material):
    d = golden_dir(material)
    with open(d / "sys_config.yml") as fh:
        cfg = yaml.safe_load(fh)
    with open(d / "simulation_parameters.yml") as fh:
        sim = yaml.safe_load(fh)
    cfg["input_coord_file_location"] = d / "POSCAR"
    return cfg, sim


@pytest.fixture(scope="session")
def hematite():
    from tests.systems import build_example
    return build_example("hematite")


@pytest.fixture(scope="session")
def bvo():
    from tests.systems import build_example
    return build_example("bvo")
