"""The committed C1 fixture (tests/golden/c1_64cube.json, written by
tests/golden/make_c1_golden.py): the workload generator, the host SZ-like codec
and the oracle's correction reproduce every recorded digest and count."""
import json
import os

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def golden():
    with open(os.path.join(HERE, "golden", "c1_64cube.json")) as fh:
        return json.load(fh)


def test_c1_oracle_matches_golden(golden):
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_c1_golden
    rec = make_c1_golden.record()
    assert rec == golden
