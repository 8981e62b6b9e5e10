import pytest


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (gfx950) and libdlu.so")
    config.addinivalue_line("markers", "slow: takes more than ~10 s on the CPU")
