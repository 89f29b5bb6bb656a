"""Writes the full-size oracle frames the GPU parity tests compare against.

The oracle (oracle/liboracle.so, the CPU restatement of the reference's
renderer) renders the 1024x1024 200-path frame and the 1920x1080 5000-draw
frame of the seeded workload generator -- every logical flush -- and records,
per pixel, whether some draw covers it fractionally (an anti-aliased edge).
The 1080p frame takes minutes on the CPU, so it is rendered here once and
stored; tests/test_golden_frames.py checks that the stored frames still match
the generator and the oracle.

    python tests/golden/make_frame_fixtures.py

Outputs, per frame: frame_<name>.npy.xz (RGBA8, height x width x 4) and
edges_<name>.npy.xz (packed bits, row-major), plus frames.json with the
frame sizes, logical flush counts and a SHA-256 of the oracle's layout
records (spans, contours, path and paint records of every logical flush).
"""
import hashlib
import io
import json
import lzma
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "rive-runtime_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle_binding as ob  # noqa: E402
import rivehip  # noqa: E402

FRAMES = {"cfg1": (1, 1024, 1024, 200), "cfg2": (2, 1920, 1080, 5000)}


def layout_digest(f):
    """SHA-256 over every logical flush's spans, contours, PathData,
    PaintData and PaintAuxData records."""
    h = hashlib.sha256()
    keep = f.index
    for k in range(f.flush_count):
        f.select(k)
        for what in (0, 1, 2, 3, 4, 5):
            h.update(f.get(what).tobytes())
    f.select(keep)
    return h.hexdigest()


def save_xz(path, array):
    buf = io.BytesIO()
    np.save(buf, array, allow_pickle=False)
    with open(path, "wb") as fh:
        fh.write(lzma.compress(buf.getvalue(), preset=6))


def load_xz(path):
    with open(path, "rb") as fh:
        return np.load(io.BytesIO(lzma.decompress(fh.read())), allow_pickle=False)


def main(names):
    meta_path = os.path.join(HERE, "frames.json")
    meta = json.load(open(meta_path)) if os.path.exists(meta_path) else {}
    for name in names:
        cfg, w, h, n = FRAMES[name]
        s = rivehip.generate_stream(cfg, width=w, height=h, draw_count=n)
        f = ob.OracleFlush(s)
        img, edges, secs = f.render_edges()
        save_xz(os.path.join(HERE, f"frame_{name}.npy.xz"), img)
        save_xz(os.path.join(HERE, f"edges_{name}.npy.xz"), np.packbits(edges.ravel()))
        meta[name] = {"config": cfg, "width": w, "height": h, "draws": n,
                      "seed": hex(rivehip.CONFIG_SEEDS[cfg]), "logical_flushes": f.flush_count,
                      "layout_sha256": layout_digest(f), "edge_fraction": float(edges.mean())}
        print(f"{name}: {w}x{h}, {f.flush_count} logical flushes, rendered in {secs:.1f} s", flush=True)
        f.close()
        rivehip.free_stream(s)
    with open(meta_path, "w") as fh:
        json.dump(meta, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:] or list(FRAMES))
