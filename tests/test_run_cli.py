"""The c3d_run program's input checking (no device needed: every check runs before the
first GPU call).  A deck follows init.cpp:361-378's keys; what the reference would abort
on, or loop on forever, is an input error with exit status 1."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "c3d-2015_amd", "c3d_run")

GOOD = {
    "SOLVER": "mn", "TFINAL": "0.1", "OUTDT": "1", "OUTPUT_DATA": "true",
    "OUTPUT_OPT_STATS": "false", "OUTPUT_TIMINGS": "false", "NX": "4", "NY": "4", "NZ": "4",
    "MPIX": "1", "MPIY": "1", "MPIZ": "1", "AX": "-1", "AY": "-1", "AZ": "-1", "BX": "1",
    "BY": "1", "BZ": "1", "MOMENT_ORDER": "2", "QUAD_ORDER": "8", "CFL_FACTOR": "0.5",
    "CROSS_SECTION": "1", "INIT_FLOOR": "1e-4", "GAUSS_SIGMA": "0.3", "INIT_COND": "gaussian",
    "THETA": "2", "OMP_THREADS": "1", "PN_USE_KINETIC_FLUX": "0", "MN_TOL": "1e-8",
    "MN_TOL_GAMMA": "1.1", "MN_COND_H_MAX": "1e10", "MN_MAX_ITER": "100",
    "MN_BATCHED": "true", "MN_USE_GAUNT_COEF": "false", "MN_USE_GAUNT_SPARSE": "false",
    "CUDA_KINETIC": "true", "CUDA_BATCHED": "true", "CUDA_BATCH_SIZE": "50",
}


def _run(tmp_path, deck, *args):
    if deck is not None:
        (tmp_path / "input.deck").write_text(
            "".join(f"{k} {v}   # comment\n" for k, v in deck.items()))
    return subprocess.run([EXE, *args], cwd=tmp_path, capture_output=True, text=True,
                          timeout=60)


@pytest.fixture(autouse=True)
def _built():
    if not os.path.exists(EXE):
        pytest.skip("c3d_run not built")


def test_no_deck(tmp_path):
    r = _run(tmp_path, None)
    assert r.returncode == 1 and "no input deck" in r.stdout


@pytest.mark.parametrize("key", ["TFINAL", "MPIZ", "THETA", "MN_USE_GAUNT_SPARSE", "BZ"])
def test_missing_key(tmp_path, key):
    deck = dict(GOOD)
    del deck[key]
    r = _run(tmp_path, deck)
    assert r.returncode == 1 and f"missing or malformed key {key}" in r.stdout


@pytest.mark.parametrize("key,value,msg", [
    ("SOLVER", "sn", "solver type does not exist"),
    ("INIT_COND", "square", "initial condition not supported"),
    ("TFINAL", "0", "TFINAL and OUTDT must be positive"),
    ("OUTDT", "-1", "TFINAL and OUTDT must be positive"),
    ("MOMENT_ORDER", "5", "MOMENT_ORDER must be 1..4"),
    ("NX", "1", "at least 2 cells"),
    ("BY", "-2", "AX < BX"),
    ("NY", "x", "at least 2 cells"),  # read as 0, as the reference's reader does
])
def test_bad_value(tmp_path, key, value, msg):
    deck = dict(GOOD)
    deck[key] = value
    r = _run(tmp_path, deck)
    assert r.returncode == 1 and msg in r.stdout, r.stdout


def test_pn_deck_needs_no_mn_keys_but_its_own(tmp_path):
    deck = {k: v for k, v in GOOD.items() if not k.startswith("MN_")}
    deck["SOLVER"] = "pn"
    del deck["PN_USE_KINETIC_FLUX"]
    r = _run(tmp_path, deck)
    assert r.returncode == 1 and "PN_USE_KINETIC_FLUX" in r.stdout


def test_usage(tmp_path):
    assert _run(tmp_path, GOOD, "--bogus").returncode == 2
    r = _run(tmp_path, GOOD, "--precision", "fp16")
    assert r.returncode == 1 and "unknown precision" in r.stdout
