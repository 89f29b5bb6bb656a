"""Copies the reference's image-decoder test assets into tests/golden/images.

The files are the inputs of tests/unit_tests/runtime/image_decoders_test.cpp
(their sizes and decoded dimensions are asserted there): three PNGs, the
baseline JPEG and the oversized JPEG.  bad.png (534 KB, a 58278x24566 header
over far too little image data) is kept as its first 4096 bytes, which hold
the header and the start of the image data: enough for the same outcome.
Run in a container where the reference is at /root/reference.
"""
import os
import shutil

SRC = "/root/reference/tests/unit_tests/assets"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "images")

if __name__ == "__main__":
    os.makedirs(DST, exist_ok=True)
    for name in ["placeholder.png", "batdude.png", "nomoon.png", "open_source.jpg", "bad.jpg"]:
        shutil.copyfile(os.path.join(SRC, name), os.path.join(DST, name))
    with open(os.path.join(SRC, "bad.png"), "rb") as f:
        head = f.read(4096)
    with open(os.path.join(DST, "bad_head.png"), "wb") as f:
        f.write(head)
    print("wrote", sorted(os.listdir(DST)))
