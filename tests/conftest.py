"""Shared test setup: the `gpu` marker, loaders for the rt-graph_amd package (its directory
name is not an identifier) and for the CPU oracle, and a GPU-presence check."""
from __future__ import annotations

import importlib.util
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def load_rtg():
    if "rtgraph_amd" in sys.modules:
        return sys.modules["rtgraph_amd"]
    pkg = ROOT / "rt-graph_amd"
    spec = importlib.util.spec_from_file_location("rtgraph_amd", pkg / "__init__.py",
                                                  submodule_search_locations=[str(pkg)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["rtgraph_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def load_oracle():
    from oracle import oracle  # noqa: WPS433
    return oracle


@pytest.fixture(scope="session")
def rtg():
    return load_rtg()


@pytest.fixture(scope="session")
def orc():
    return load_oracle()


@pytest.fixture(scope="session")
def ctx(rtg):
    c = rtg.Context(0)
    yield c
    c.close()
