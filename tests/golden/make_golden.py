"""Writes tests/golden/inputs.json: SHA-256 digests of the seeded synthetic
inputs of the benchmark workloads, so that a change of generator (or of the
numpy / scipy arithmetic under it) shows up as a failing test rather than as
a silently different workload.

usage: python tests/golden/make_golden.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import __graft_entry__ as entry  # noqa: E402


def digests():
    pkg = entry.load_package()
    cf = pkg.configs
    out = {}
    _, g, inlet = cf.density_field(nx=1024, ny=512, seed=0)
    out["density_1024x512_seed0_gamma"] = cf.sha256(g)
    out["density_1024x512_inlet"] = cf.sha256(inlet)
    _, g, _ = cf.density_field(nx=512, ny=256, seed=0)
    out["density_512x256_seed0_gamma"] = cf.sha256(g)
    st = cf.thermal_stack()
    out["thermal_512x256_power_map_seed0"] = cf.sha256(st["q0"])
    out["batch48_sizes"] = [[c.nx, c.ny] for c, _, _ in cf.batch_sweep()]
    return out


if __name__ == "__main__":
    with open(os.path.join(HERE, "inputs.json"), "w") as f:
        json.dump(digests(), f, indent=1)
    print("wrote", os.path.join(HERE, "inputs.json"))
