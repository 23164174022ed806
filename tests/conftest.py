"""Shared pytest configuration: the `gpu` marker, import paths and fixtures.

CPU tests (-m "not gpu") cover the oracle against the CFRG golden vectors, the
derived constants, and the C-ABI library's loading/exports and host-side
validation.  GPU tests (-m gpu) are the parity tests proper: every call goes
through libcubbs_bbs_hip.so and is compared with the oracle or a fixture.
"""
import glob
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "cubbs-cred_amd"))
GOLDEN = os.path.join(ROOT, "tests", "golden", "cfrg")

R_ORDER = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcubbs_bbs_hip.so)")


def load_json(rel):
    with open(os.path.join(GOLDEN, rel)) as fh:
        return json.load(fh)


def signature_fixtures():
    return [load_json(os.path.relpath(p, GOLDEN)) for p in sorted(glob.glob(os.path.join(GOLDEN, "signature", "*.json")))]


def proof_fixtures():
    return [load_json(os.path.relpath(p, GOLDEN)) for p in sorted(glob.glob(os.path.join(GOLDEN, "proof", "*.json")))]


def trace_scalars(trace):
    rs = trace["random_scalars"]
    return [bytes.fromhex(rs[k]) for k in ("r1", "r2", "e_tilde", "r1_tilde", "r3_tilde")] + [
        bytes.fromhex(x) for x in rs["m_tilde_scalars"]]


@pytest.fixture(scope="session")
def oracle():
    import oracle_py
    oracle_py.lib()
    return oracle_py


@pytest.fixture(scope="session")
def gpu():
    import cubbs_hip
    rc, kind, hw = cubbs_hip.backend_info()
    if rc != 0 or not hw:
        pytest.fail("HIP device not usable (backend_info rc=%d); the GPU tests have no CPU fallback" % rc)
    return cubbs_hip
