"""Writes tests/golden/generators.json: the first 16 values of each synthetic
input generator of hifde3d_amd/problems.py (the definitions the device
generators follow), so that a change to either side is caught.

Usage: python tests/golden/make_generators.py
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from hifde3d_amd import problems as pb  # noqa: E402


def main():
    out = {
        "uniform_seed1": pb.uniform(1, 16).tolist(),
        "normal_seed2": pb.normal(2, 16).tolist(),
        "pow10_NC8_seed1_first16": pb.pow10_coeff(8, 1).ravel(order="F")[:16].tolist(),
        "gaussian_field_NC16_seed2_corr0.1_first16":
            pb.gaussian_field(16, 2, 0.1).ravel(order="F")[:16].tolist(),
        "rhs_normal_NC8_seed0_interior_first16":
            pb.rhs_normal(8, 0)[pb.interior_mask(8)][:16].tolist(),
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "generators.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
