"""The golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py from the FP64 oracle): the files match their
manifest's sha256, and the oracle still reproduces every stored array bit
for bit, so a change to the oracle that moves any result shows up here."""
import hashlib
import os
import sys

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)

from make_golden import run_oracle, scenes  # noqa: E402


def _manifest():
    out = {}
    for line in open(os.path.join(GOLDEN, "MANIFEST.txt")):
        h, name = line.split()
        out[name] = h
    return out


def test_manifest_hashes():
    m = _manifest()
    assert sorted(m) == sorted(f"{n}.npz" for n in scenes())
    for name, h in m.items():
        assert hashlib.sha256(open(os.path.join(GOLDEN, name), "rb").read()).hexdigest() == h, name


@pytest.mark.parametrize("name", list(scenes()))
def test_oracle_reproduces_fixture(name, oracle_threads):
    ref = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    got = run_oracle(name)
    assert sorted(ref.files) == sorted(got)
    for k in ref.files:
        assert np.array_equal(ref[k], got[k]), (name, k)
    assert np.abs(ref["ez"]).max() > 0
