import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "fast-er_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and libfaster_hip.so")
    config.addinivalue_line("markers", "slow: longer-running check")


@pytest.fixture(scope="session")
def oracle():
    import faster_oracle

    faster_oracle.lib()
    return faster_oracle


@pytest.fixture(scope="session")
def config5_fit():
    """1,000,000 x 1,000,000 synthetic linkage with three string variables, an
    exact house number and a numeric birth day (levels within 0 / 3 days),
    every value missing with probability 5%: the fitted Comparison (768
    patterns, levels [4, 4, 4, 3, 4]) and its tables. Needs the GPU."""
    import faster
    from faster.synthetic import add_numeric_and_missing, linkage_tables

    n = 1_000_000
    vf = ["first_name", "last_name", "street"]
    dfa, dfb, truth = linkage_tables(n, n, seed=0)
    fa, fb = add_numeric_and_missing(dfa, dfb, truth, seed=0, missing=0.05)
    print("1M x 1M with numeric + missing levels", flush=True)
    comp = faster.Comparison(fa, fb, vf, vf, ["house_number"], ["house_number"], Vars_Numeric_A=["birth_day"],
                             Vars_Numeric_B=["birth_day"])
    comp.fit(Numeric_Dist=(0.0, 3.0), Missing=True)
    return fa, fb, comp
