"""Writes the seeded bubble fixtures (discrete.in format) of the four
generated workloads with the library's SplitMix64 generator
(zsc_generate_bubbles).  Run from the repository root:

    python tests/golden/make_bubbles.py
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from lbm_bench_gpu_amd import configs, zsc  # noqa: E402


def main():
    for i in (1, 2, 3, 4):
        w = configs.workload(i)
        path = os.path.join(HERE, f"discrete_seed{i}.in")
        zsc.write_bubbles(path, w.bubbles)
        print(path, len(w.bubbles))


if __name__ == "__main__":
    main()
