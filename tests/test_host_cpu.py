"""The host-level drop-in library (libc3d_host.so, include/c3d_host.h) loads, exports the
reference's entry points under their C++ names, and reads input decks with the rules of
src/input_deck_reader.cpp:17-157 (no GPU needed)."""
import ctypes as C

import host_lib as H


def test_exports_reference_entry_points():
    L = H.load()
    missing = [k for k, v in H.CXX_SYMBOLS.items() if not hasattr(L, v)]
    missing += [s for s in H.C_SYMBOLS if not hasattr(L, s)]
    for g in ("g_u", "g_mnAlpha", "g_gX", "g_numMoments", "g_theta", "g_beta"):
        if not hasattr(L, g):
            missing.append(g)
    assert not missing, missing


DECK = """SOLVER mn       # Solver Type: pn or mn
TFINAL 0.1
OUTPUT_DATA TRUE
OUTPUT_TIMINGS false
NX 50
Mixed_Case_Key Value_With_Case
\tTABBED_KEY 7
TAB_IN_VALUE 3\t
MN_TOL 1.0e-4
# a comment line
   # indented comment
KEY_ONLY
TWO_VALUES 1 2
MN_BATCHED true             # trailing comment
"""


def _value(path, key):
    L = H.load()
    buf = C.create_string_buffer(128)
    ok = L.c3d_host_deck_value(str(path).encode(), key.encode(), buf)
    return buf.value.decode() if ok else None


def test_deck_reader_rules(tmp_path):
    p = tmp_path / "input.deck"
    p.write_text(DECK)
    assert _value(p, "SOLVER") == "mn"
    assert _value(p, "solver") == "mn"              # keys compare case-insensitively
    assert _value(p, "TFINAL") == "0.1"
    assert _value(p, "OUTPUT_DATA") == "true"       # values are lower-cased
    assert _value(p, "MIXED_CASE_KEY") == "value_with_case"
    assert _value(p, "TABBED_KEY") == "7"            # tab separates the key
    assert _value(p, "TAB_IN_VALUE") == "3\t"        # but not the value
    assert _value(p, "MN_TOL") == "1.0e-4"
    assert _value(p, "MN_BATCHED") == "true"
    assert _value(p, "TWO_VALUES") is None           # a third token drops the line
    # a key without a value keeps the previous record's value
    assert _value(p, "KEY_ONLY") == "1.0e-4"
    assert _value(p, "NOT_THERE") is None
    assert _value(tmp_path / "missing.deck", "SOLVER") is None


def test_tuning_table_lookup_and_candidates(tmp_path):
    """The tuner's candidate grid keeps the four-wave threshold at or below the tail switch
    and holds the tail knobs only (the column sweep has one launch shape at every order); the
    front end reads entries by (N, Q, precision) and reports absent ones as None."""
    import importlib.util
    import json
    import os
    from conftest import ROOT, product
    spec = importlib.util.spec_from_file_location(
        "c3d_tune", os.path.join(ROOT, "c3d-2015_amd", "tune.py"))
    tune = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(tune)
    mixed = list(tune.candidates("mixed", 25))
    cols = list(tune.candidates("mixed", 16))
    fp64 = list(tune.candidates("fp64", 16))
    assert all(c["tail_wide"] <= c["tail_cells"] for c in mixed + cols + fp64)
    assert all(set(c) == {"tail_cells", "tail_passes", "tail_wide"} for c in mixed + cols + fp64)
    P = product()
    path = tmp_path / "t.json"
    path.write_text(json.dumps({"3,24,mixed": {"tail_cells": 8192, "tail_passes": 32,
                                               "tail_wide": 512}}))
    assert P.tuning_for(3, 24, "mixed", str(path))["tail_wide"] == 512
    assert P.tuning_for(3, 24, "fp64", str(path)) is None
    assert P.tuning_for(3, 24, "mixed", str(tmp_path / "absent.json")) is None
