"""Multi-field compression as a captured HIP graph (ASZ_OPT_GRAPHS).

A batch chunk's device sequence (field descriptors, pivots, every round with
its t selection and folds) is captured once per (plan, field count, buffers)
and replayed for later chunks: the replays must give exactly the streams and
reconstructions of the launches issued one by one, for new field values each
time (the descriptors are copied in by the graph itself), with eager work and
synchronisations of the program between the replays."""
import numpy as np
import pytest
import torch

from asz_amd import codec, fields

pytestmark = pytest.mark.gpu

PLAN = ["cubic0_012", "cubic1_012", "cubic2_012", "cubic0_021", "cubic1_120", "cubic2_012", "cubic0_012",
        "cubic1_021", "cubic2_120", "cubic0_012", "cubic1_012", "cubic2_021", "cubic0_120", "cubic1_012"]
SHAPES = [(48, 56, 80)] * 6 + [(40, 44, 52)] * 4


@pytest.fixture(scope="module")
def ctx():
    c = codec.Context(0)
    yield c
    c.set_graphs(True)
    c.close()


def _batch(ctx, flds, graphs, tsel, concurrency):
    ctx.set_graphs(graphs)
    work = [g.clone() for g in flds]
    streams, _ = ctx.compress_batch(work, [codec.make_plan(PLAN)] * len(flds),
                                    codec.make_params(rel_eb=1e-3, block_pivot_ratio=3, family="cubic"), tsel=tsel,
                                    concurrency=concurrency)
    torch.cuda.synchronize()
    st = ctx.stats()
    return streams, [w.cpu().numpy() for w in work], int(st.graph_launched), int(st.graph_captured)


@pytest.mark.parametrize("tsel", ["grad", "gpu"])
def test_graph_replays_equal_eager(ctx, tsel):
    launched_total = captured_total = 0
    for call, seed0 in enumerate((50, 70, 90)):
        print(f"graph replay: call {call}, tsel {tsel}", flush=True)
        flds = [fields.config2_field(s, seed=seed0 + i, device="cuda").contiguous() for i, s in enumerate(SHAPES)]
        eager_s, eager_r, l0, _ = _batch(ctx, flds, False, tsel, 1)
        assert l0 == 0
        # eager work of the program between the graph launches
        single = fields.to_numpy(fields.config2_field((24, 30, 36), seed=seed0)).astype(np.float32)
        ctx.compress(single, codec.make_plan(PLAN[:6]), codec.make_params(rel_eb=1e-3, block_pivot_ratio=3))
        torch.cuda.synchronize()
        graph_s, graph_r, launched, captured = _batch(ctx, flds, True, tsel, 1)
        # one worker: a chunk per shape (6 and 4 fields), each a graph
        assert launched == 2
        if call > 0:
            assert captured == 0, "the second and later calls replay the captured graphs"
        launched_total += launched
        captured_total += captured
        for i in range(len(SHAPES)):
            assert graph_s[i] == eager_s[i], f"call {call}, field {i}: streams differ"
            assert np.array_equal(graph_r[i].view(np.uint32), eager_r[i].view(np.uint32))
    assert captured_total <= 2 and launched_total == 6


def test_graph_with_concurrent_workers(ctx):
    """Several worker contexts, each with its own graphs, on their own streams."""
    flds = [fields.config2_field(s, seed=110 + i, device="cuda").contiguous() for i, s in enumerate(SHAPES)]
    eager_s, eager_r, _, _ = _batch(ctx, flds, False, "grad", 3)
    for rep in range(2):
        graph_s, graph_r, launched, _ = _batch(ctx, flds, True, "grad", 3)
        assert launched >= 2
        for i in range(len(SHAPES)):
            assert graph_s[i] == eager_s[i], f"rep {rep}, field {i}: streams differ"
            assert np.array_equal(graph_r[i].view(np.uint32), eager_r[i].view(np.uint32))
