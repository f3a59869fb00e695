"""Plan files (asz_amd.plans) on two of the reference's saved plans (copied as
fixtures from plans/isabel_sz3_style.json, version 1, and
plans/transfer/isabel_wind_self/Uf24.json, version 2): parsing, the
rel_eb lookup, the version-2 round trip and the launcher's validation errors."""
import json
import os

import pytest

from asz_amd import plans

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "plans")
V1 = os.path.join(HERE, "isabel_sz3_style.json")
V2 = os.path.join(HERE, "isabel_wind_Uf24.json")


def _raw(path):
    with open(path) as fh:
        return json.load(fh)


@pytest.mark.parametrize("path", [V1, V2])
def test_parse_matches_the_raw_layers(path):
    doc = _raw(path)
    got = plans.load_plan(path)
    assert len(got) == len(doc["plans"])
    inv = {v: k for k, v in plans.TRAVERSAL_ORDERS.items()}
    for (rel, preds, fam), raw in zip(got, doc["plans"]):
        assert rel == raw["rel_eb"]
        assert fam == raw.get("predictor_family", "cubic")
        assert len(preds) == len(raw["layers"])
        for p, layer in zip(preds, raw["layers"]):
            assert p == f"cubic{layer['prediction_direction']}_{plans.TRAVERSAL_ORDERS[tuple(layer['traversal_order'])]}"
            assert list(inv[p[7:]]) == layer["traversal_order"]
        plan = plans.to_plan(preds)
        assert plan.layers == len(preds)


def test_version1_defaults_to_cubic_and_lookup():
    got = plans.load_plan(V1)
    assert all(f == "cubic" for _, _, f in got)
    rels = [r for r, _, _ in got]
    preds, fam = plans.predictors_for_rel_eb(got, rels[1], V1)
    assert preds == got[1][1] and fam == "cubic"
    with pytest.raises(ValueError):
        plans.predictors_for_rel_eb(got, 0.123456, V1)


def test_version2_round_trip(tmp_path):
    got = plans.load_plan(V2)
    entries = {r: (p, f) for r, p, f in got}
    out = tmp_path / "sub" / "plan.json"
    plans.write_plan(out, entries)
    again = plans.load_plan(out)
    assert sorted(again, reverse=True) == sorted(got, reverse=True)
    doc = _raw(out)
    assert doc["version"] == 2
    assert [p["rel_eb"] for p in doc["plans"]] == sorted([r for r, _, _ in got], reverse=True)
    assert not (tmp_path / "sub" / "plan.json.tmp").exists()


@pytest.mark.parametrize("doc", [
    [],
    {"version": 3, "plans": []},
    {"version": 1, "plans": [{"rel_eb": "x", "layers": [{"prediction_direction": 0, "traversal_order": [0, 1, 2]}]}]},
    {"version": 1, "plans": [{"rel_eb": 0.1, "layers": []}]},
    {"version": 1, "plans": [{"rel_eb": 0.1, "layers": [{"prediction_direction": 3, "traversal_order": [0, 1, 2]}]}]},
    {"version": 1, "plans": [{"rel_eb": 0.1, "layers": [{"prediction_direction": 0, "traversal_order": [2, 1, 0]}]}]},
    {"version": 1, "plans": [{"rel_eb": 0.1, "layers": [{"prediction_direction": 0, "traversal_order": [0, 1, 2],
                                                          "extra": 1}]}]},
    {"version": 2, "plans": [{"rel_eb": 0.1, "predictor_family": "quartic",
                              "layers": [{"prediction_direction": 0, "traversal_order": [0, 1, 2]}]}]},
])
def test_malformed_documents_raise(tmp_path, doc):
    path = tmp_path / "bad.json"
    path.write_text(json.dumps(doc))
    with pytest.raises(ValueError):
        plans.load_plan(path)


def test_unknown_predictor_cannot_be_written(tmp_path):
    with pytest.raises(ValueError):
        plans.write_plan(tmp_path / "p.json", {0.1: (["linear0_012"], "cubic")})
