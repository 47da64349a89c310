"""Writes the 4^4 ILDG fixture and its oracle flow trajectory.

conf_4x4x4x4.ildg: Haar hot start (seed 99) + 10 heatbath sweeps at beta=6.0
from the oracle's generator, written with the reference's LIME framing
(src/io/ildg.cpp:54-136) by tests/limefile.py.
flow_4x4x4x4.json: the oracle's topo_charge_flowed trajectory on it,
N_steps_gf = 21, N_rk_steps = 1, eps = 0.02 (t, Q, E and the plaquette, %.17g).
Run from the repository root:  python tests/golden/make_ildg_fixture.py
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import limefile  # noqa: E402
import oracle  # noqa: E402


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    L = 4
    f = oracle.hot_start(L, 99)
    oracle.heatbath(f, L, 6.0, 10, 99 + 0x5EED)
    limefile.write_ildg(os.path.join(here, "conf_4x4x4x4.ildg"), f, (4, 4, 4, 4))
    tqe, plaq, _ = oracle.flow_measure(f, L, 0.02, 21, 1)
    with open(os.path.join(here, "flow_4x4x4x4.json"), "w") as fp:
        json.dump({
            "lattice": [4, 4, 4, 4], "eps": 0.02, "N_steps_gf": 21, "N_rk_steps": 1,
            "t": ["%.17g" % v for v in tqe[:, 0]],
            "Q": ["%.17g" % v for v in tqe[:, 1]],
            "E": ["%.17g" % v for v in tqe[:, 2]],
            "P": ["%.17g" % v for v in plaq],
        }, fp, indent=1)
    print("wrote conf_4x4x4x4.ildg and flow_4x4x4x4.json")


if __name__ == "__main__":
    main()
