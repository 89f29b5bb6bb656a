"""Command-stream replay on the device: a frame recorded to the deferred
command stream, written to a file, read back and replayed through
RiveRenderer renders the same pixels, bit for bit, as the draw stream
rendered directly; streams cut short, with an unknown opcode or with a
handle naming no resource are rejected before anything is flushed.
"""
import numpy as np
import pytest

import rivehip

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("config,draws", [(1, 200), (2, 1500)])
def test_replay_equals_direct(hip_ctx, tmp_path, config, draws):
    print(f"[cmd] config {config} draws {draws}", flush=True)
    s = rivehip.generate_stream(config, draw_count=draws)
    st = s.contents
    w, h = st.width, st.height
    t = hip_ctx.target(w, h)
    try:
        hip_ctx.render_stream(s, t)
        direct = hip_ctx.read_pixels(t, w, h)
        cmd = rivehip.CommandBuffer.record(s)
        cmd.save(tmp_path / "f.rivecmd")
        cmd.close()
        back = rivehip.CommandBuffer.load(tmp_path / "f.rivecmd")
        hip_ctx.render_commands(back, t)
        replayed = hip_ctx.read_pixels(t, w, h)
        back.close()
    finally:
        hip_ctx.free_target(t)
        rivehip.free_stream(s)
    assert direct.any()
    assert np.array_equal(direct, replayed)


def test_corrupt_streams_are_rejected(hip_ctx):
    print("[cmd] corrupt", flush=True)
    s = rivehip.generate_stream(1, draw_count=20)
    w, h = s.contents.width, s.contents.height
    cmd = rivehip.CommandBuffer.record(s)
    c, b = cmd.bytes()
    cmd.close()
    rivehip.free_stream(s)
    t = hip_ctx.target(w, h)
    try:
        bad = [c[:-3],                                    # truncated payload
               c + bytes([200]),                          # unknown opcode
               c + bytes([25]) + (999).to_bytes(4, "little") + bytes(12)]  # drawPath of no path
        for cb in bad:
            x = rivehip.CommandBuffer.from_bytes(cb, b, w, h)
            with pytest.raises(RuntimeError):
                hip_ctx.render_commands(x, t)
            x.close()
    finally:
        hip_ctx.free_target(t)
