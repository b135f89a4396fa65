import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def sysdir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("systems"))


@pytest.fixture(scope="session")
def make_system(sysdir):
    """make_system(n_orb, n_elec, seed, **gen) -> dict with in_path, fcidump_path, n_so, ..."""
    from sfci_amd import integrals as I
    cache = {}

    def _make(n_orb, n_elec, seed, **gen):
        key = (n_orb, n_elec, seed, tuple(sorted(gen.items())))
        if key not in cache:
            tag = "o%d-e%d-s%d-%d" % (n_orb, n_elec, seed, len(cache))
            fc = os.path.join(sysdir, tag + "_fcidump")
            inp = os.path.join(sysdir, tag + "_IN")
            I.synth_fcidump(fc, n_orb, n_elec, seed, **gen)
            I.convert_rhf(fc, inp)
            cache[key] = dict(in_path=inp, fcidump_path=fc, n_orb=n_orb, n_so=2 * n_orb, n_elec=n_elec,
                              seed=seed, init_hf=(1 << n_elec) - 1)
        return cache[key]

    return _make
