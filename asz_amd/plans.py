"""Saved plan files (version 1 and 2 JSON), as the reference launcher reads and
writes them (t1_start_compressors.py:224-311).

A document is {"version": 1 | 2, "plans": [{"rel_eb": float,
"predictor_family": "cubic" | "quintic" (version 2; cubic when absent),
"layers": [{"prediction_direction": 0 | 1 | 2, "traversal_order": [i, j, k]},
...]}, ...]}; each layer is the predictor "cubic<direction>_<order>" with
order (1,2,0) -> "120", (0,2,1) -> "021", (0,1,2) -> "012".  The same
validation as the launcher's load_plan, with the errors raised as
ValueError; write_plan writes version 2, plans sorted by decreasing rel_eb,
through a temporary file and an atomic rename.
"""
from __future__ import annotations

import json
import math
import os
import re
from pathlib import Path

from . import codec

TRAVERSAL_ORDERS = {(1, 2, 0): "120", (0, 2, 1): "021", (0, 1, 2): "012"}


def load_plan(path):
    """[(rel_eb, [predictor, ...], family), ...] from a plan file (t1_start_compressors.py:224-248)."""
    with open(path, encoding="utf-8") as fh:
        doc = json.load(fh)
    if not isinstance(doc, dict) or doc.get("version") not in (1, 2) or not isinstance(doc.get("plans"), list):
        raise ValueError(f"malformed plan file: {path}")
    out = []
    for pid, plan in enumerate(doc["plans"]):
        if (not isinstance(plan, dict) or not isinstance(plan.get("rel_eb"), (int, float))
                or isinstance(plan.get("rel_eb"), bool) or not isinstance(plan.get("layers"), list)
                or not plan["layers"]):
            raise ValueError(f"plan {pid} is malformed: {path}")
        predictors = []
        for lid, layer in enumerate(plan["layers"]):
            if not isinstance(layer, dict) or set(layer) != {"prediction_direction", "traversal_order"}:
                raise ValueError(f"layer {lid} of plan {pid} is malformed: {path}")
            d = layer["prediction_direction"]
            trav = layer["traversal_order"]
            if not isinstance(d, int) or isinstance(d, bool) or d not in (0, 1, 2):
                raise ValueError(f"layer {lid} of plan {pid} has an invalid prediction direction: {path}")
            if not isinstance(trav, list) or tuple(trav) not in TRAVERSAL_ORDERS:
                raise ValueError(f"layer {lid} of plan {pid} has an invalid traversal order: {path}")
            predictors.append(f"cubic{d}_{TRAVERSAL_ORDERS[tuple(trav)]}")
        family = plan.get("predictor_family", "cubic")
        if family not in ("cubic", "quintic"):
            raise ValueError(f"plan {pid} has an invalid predictor_family: {path}")
        out.append((float(plan["rel_eb"]), predictors, family))
    return out


def predictors_for_rel_eb(plans, rel_eb, path=""):
    """The (predictors, family) of the one plan whose rel_eb equals rel_eb
    to 1e-12 relative (t1_start_compressors.py:250-254)."""
    matches = [(p, f) for r, p, f in plans if math.isclose(r, rel_eb, rel_tol=1e-12, abs_tol=0.0)]
    if len(matches) != 1:
        raise ValueError(f"{path} should hold exactly one plan with rel_eb={rel_eb:g}, found {len(matches)}")
    return matches[0]


def plan_document(entries):
    """Version-2 document from {rel_eb: (predictors, family)} (t1_start_compressors.py:289-300)."""
    plans = []
    for rel_eb, (predictors, family) in sorted(entries.items(), reverse=True):
        layers = []
        for pr in predictors:
            m = re.fullmatch(r"cubic([012])_(120|021|012)", pr)
            if m is None:
                raise ValueError(f"cannot write unknown predictor: {pr}")
            layers.append({"prediction_direction": int(m.group(1)), "traversal_order": [int(v) for v in m.group(2)]})
        plans.append({"rel_eb": rel_eb, "predictor_family": family, "layers": layers})
    return {"version": 2, "plans": plans}


def write_plan(path, entries):
    """Write {rel_eb: (predictors, family)} as a version-2 plan file (t1_start_compressors.py:302-309)."""
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    tmp = path.with_name(path.name + ".tmp")
    with open(tmp, "w", encoding="utf-8") as fh:
        json.dump(plan_document(entries), fh, indent=2)
        fh.write("\n")
    os.replace(tmp, path)


def to_plan(predictors):
    """The C ABI plan (asz_plan) of a predictor list."""
    return codec.make_plan(predictors)
