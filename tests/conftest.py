import glob
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from __graft_entry__ import load_package  # noqa: E402

load_package()  # makes `import whale_jl_amd` work in every test module

GOLDEN = os.path.join(ROOT, "tests", "golden")
EXTREE = ("((MPOL:4.752,PPAT:4.752):0.292,(SMOE:4.457,(((OSAT:1.555,(ATHA:0.5548,CPAP:0.5548):1.0002):0"
          ".738,ATRI:2.293):1.225,(GBIL:3.178,PABI:3.178):0.34):0.939):0.587);")
KAT_INSERTS = [("wgd_1", "ATHA", "ATHA"), ("wgd_2", "ATHA", "ATRI")]
MULTREE = ("((MPOL:4.752,PPAT:4.752):0.292,((SMOE:4.0,PPAT:4.0):0.457,(((OSAT:1.555,(ATHA:0.5548,CPAP:0.5548)"
           ":1.0002):0.738,ATRI:2.293):1.225,(GBIL:3.178,PABI:3.178):0.34):0.939):0.587);")
KAT1 = -60.96367806571888
KAT2 = -592.0185620440255


def unpack_rates(rates, v):
    """Rates of the same kind as `rates` from a parameter vector laid out as
    rates.pack() lays it out."""
    v = [float(x) for x in v]
    nq = len(rates.q)
    if isinstance(rates.λ, (int, float)):
        return type(rates)(λ=v[0], μ=v[1], q=v[2:2 + nq], p=rates.p, η=v[-1])
    nl = len(rates.λ)
    return type(rates)(λ=v[:nl], μ=v[nl:2 * nl], q=v[2 * nl:2 * nl + nq], p=rates.p, η=v[-1])


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def example1_files():
    return sorted(glob.glob(os.path.join(GOLDEN, "example-1", "*.ale")))


def example1_texts():
    return [open(f).read() for f in example1_files()]


@pytest.fixture(scope="session")
def w():
    return load_package()
