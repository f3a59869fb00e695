"""Decompression of a stream that is already in device memory
(asz_decompress_device): only the header bytes go to the host, the payload
is decoded where it lies.  The reconstruction must equal the host-stream
decompression bit for bit, for a copy in another device buffer and for the
stream the compressor left in the context's own buffer (decoded in place);
truncated and malformed device streams must raise, as host streams do."""
import numpy as np
import pytest
import torch

from asz_amd import codec, fields

pytestmark = pytest.mark.gpu

PLAN = ["cubic0_012", "cubic1_012", "cubic2_012", "cubic0_021", "cubic1_120", "cubic2_012", "cubic0_012",
        "cubic1_021", "cubic2_120", "cubic0_012", "cubic1_012", "cubic2_021", "cubic0_120", "cubic1_012"]
SHAPE = (96, 112, 128)


def _setup(ctx, seed=5):
    field = fields.config2_field(SHAPE, seed=seed, device="cuda").contiguous()
    plan = codec.make_plan(PLAN)
    params = codec.make_params(rel_eb=1e-3, block_pivot_ratio=3, family="cubic")
    ctx.set_stream_on_device(True)
    w = field.clone()
    r = ctx.compress_field(w.data_ptr(), SHAPE, plan, params, tsel="gpu", copy_stream=True)
    torch.cuda.synchronize()
    return field, plan, params, r.stream


def _bits(t):
    return t.cpu().numpy().view(np.uint32)


def test_device_stream_equals_host_stream():
    with codec.Context(0) as ctx:
        field, plan, params, stream = _setup(ctx)
        ref = ctx.decompress(stream, SHAPE, params).reshape(-1)
        N = ref.size
        # a copy in a device buffer of the caller, decoded by another context
        dev = torch.frombuffer(bytearray(stream), dtype=torch.uint8).cuda()
        out = torch.empty(N, dtype=torch.float32, device="cuda")
        with codec.Context(0) as ctx2:
            ctx2.decompress_device(dev.data_ptr(), len(stream), SHAPE, params, out.data_ptr())
        torch.cuda.synchronize()
        assert np.array_equal(_bits(out), ref.view(np.uint32))
        # the stream the compressor left in its own buffer, decoded in place, twice
        w = field.clone()
        ctx.compress_field(w.data_ptr(), SHAPE, plan, params, tsel="gpu", copy_stream=False)
        ptr, nbytes = ctx.stream_device()
        assert nbytes == len(stream)
        for rep in range(2):
            out2 = torch.full((N,), float("nan"), dtype=torch.float32, device="cuda")
            ctx.decompress_device(ptr, nbytes, SHAPE, params, out2.data_ptr())
            torch.cuda.synchronize()
            assert np.array_equal(_bits(out2), ref.view(np.uint32)), f"in place, call {rep}"
        assert ctx.stream_view().tobytes() == stream, "decoding in place leaves the stream as it was"


def test_device_stream_errors():
    with codec.Context(0) as ctx:
        _, _, params, stream = _setup(ctx, seed=6)
        N = int(np.prod(SHAPE))
        out = torch.empty(N, dtype=torch.float32, device="cuda")
        for bad in (stream[: len(stream) - 50], stream[:20], stream[:8] + bytes([0]) + stream[9:]):
            dev = torch.frombuffer(bytearray(bad), dtype=torch.uint8).cuda()
            with pytest.raises(codec.AszError):
                ctx.decompress_device(dev.data_ptr(), len(bad), SHAPE, params, out.data_ptr())
            with pytest.raises(codec.AszError):
                ctx.decompress(bad, SHAPE, params)
