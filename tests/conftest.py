"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on CPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import __graft_entry__ as entry  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")


@pytest.fixture(scope="session")
def pkg():
    return entry.load_package()


@pytest.fixture(scope="session")
def orc():
    return entry.load_oracle()


@pytest.fixture(scope="session")
def scenes(pkg):
    from importlib import import_module
    return import_module(entry.PKG_NAME + ".scenes")


@pytest.fixture(scope="session")
def c1(scenes):
    return scenes.c1()


@pytest.fixture(scope="session")
def ctx(pkg):
    with pkg.Context(timing=True) as c:
        yield c


def oracle_view(orc, job, **kw):
    args = dict(radius=job.radius, stride=job.stride, top_k=job.top_k)
    args.update(kw)
    return orc.View(job.ref, job.srcs, job.Kinv, job.A, job.b, job.dmin, job.dmax, **args)


def rel_within(a, b, tol=0.01):
    b = np.asarray(b, dtype=np.float64)
    return np.abs(np.asarray(a, dtype=np.float64) - b) <= tol * np.abs(b)
