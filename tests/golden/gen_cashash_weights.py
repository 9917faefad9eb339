"""Regenerates tests/golden/cashash_weights_seed11.npz from gen_cashash_weights.cpp.

Built with g++ and -ffp-contract=off (no fused multiply-add in the bias sum).
"""
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "gen")
        out = os.path.join(d, "w.bin")
        subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", exe,
                        os.path.join(HERE, "gen_cashash_weights.cpp")], check=True)
        subprocess.run([exe, out], check=True)
        raw = open(out, "rb").read()
    w = np.frombuffer(raw[:8192], np.int8).reshape(64, 128)
    b = np.frombuffer(raw[8192:], "<i4")
    np.savez(os.path.join(HERE, "cashash_weights_seed11.npz"), weights=w, bias_base=b,
             note=np.array("std::default_random_engine{11}, normal_distribution<float>, g++ 11.4 libstdc++"))


if __name__ == "__main__":
    main()
