"""The oracle reproduces the committed fixtures (regression guard for the
checker itself)."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "singlepoints.json")))


@pytest.mark.parametrize("name", sorted(GOLD))
def test_oracle_reproduces_fixture(orc, name):
    g = GOLD[name]
    r = orc.singlepoint(np.array(g["num"]), np.array(g["xyz_bohr"]))
    assert abs(r["energy"] - g["energy"]) < 1e-9
    assert r["iterations"] == g["iterations"]
    assert np.max(np.abs(r["qsh"] - np.array(g["qsh"]))) < 1e-8
