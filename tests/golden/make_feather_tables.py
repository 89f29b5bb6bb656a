"""Writes tests/golden/feather_tables.npz: the reference's two fp16 function
tables for feathering, g_gaussianIntegralTableF16 and
g_inverseGaussianIntegralTableF16 (renderer/src/gpu.cpp), as uint16 arrays.

They are data (512 texel values each), parsed from the reference's source
file where it lies; the oracle regenerates both tables from their
definitions and the tests compare the two.  Run in a container that has the
reference:  python tests/golden/make_feather_tables.py
"""
import os
import re
import sys

import numpy as np

REF = os.environ.get("RIVE_REFERENCE", "/root/reference")
SRC = os.path.join(REF, "renderer", "src", "gpu.cpp")


def parse_table(text, name):
    m = re.search(r"const uint16_t " + name + r"\[GAUSSIAN_TABLE_SIZE\] = \{(.*?)\};", text, re.S)
    if m is None:
        sys.exit("table %s not found in %s" % (name, SRC))
    vals = [int(v, 16) for v in re.findall(r"0x[0-9a-fA-F]+", m.group(1))]
    if len(vals) != 512:
        sys.exit("table %s has %d entries" % (name, len(vals)))
    return np.array(vals, np.uint16)


def main():
    text = open(SRC).read()
    gauss = parse_table(text, "g_gaussianIntegralTableF16")
    inverse = parse_table(text, "g_inverseGaussianIntegralTableF16")
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "feather_tables.npz")
    np.savez_compressed(out, gaussian_integral=gauss, inverse_gaussian_integral=inverse)
    print("wrote", out)


if __name__ == "__main__":
    main()
