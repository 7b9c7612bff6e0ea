"""The reference-side adapter (menos_amd/integration/: MENOS's SimulationEngine
over libmenos_hip.so, its registry entry and dense backend choice), run by the
native test oracle/Makefile builds against the reference's own headers and
objects (tests/native/adapter_differential.cpp). Host mode checks the backend
choice and registry rules with synthetic inputs; device mode checks the
adapter against the reference CPU engine through SimulationEngine::execute:
statevectors and counts bit for bit, raw terms within 1e-12, the monitor,
budgets and the ledger."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINARY = os.path.join(ROOT, "oracle", "_ref", "adapter_differential")


def _run(mode: str, timeout: int):
    if not os.path.exists(BINARY):
        pytest.skip("oracle/_ref/adapter_differential is not built (the reference sources were absent at build time)")
    r = subprocess.run([BINARY, mode], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 failed" in r.stdout, r.stdout
    return r.stdout


def test_backend_choice_and_registry_rules():
    out = _run("--host", 60)
    assert out.startswith("host:")


@pytest.mark.gpu
def test_adapter_matches_the_reference_cpu_engine():
    out = _run("--device", 240)
    print(out)
    assert out.startswith("device:")
