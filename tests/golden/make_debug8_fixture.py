"""Writes tests/golden/debug8_payoffs.json from the reference's own per-path payoff dump.

Source data: rBergomi/testing/debug8.txt~ of the reference repository -- 200 lines
"m payoff" printed by the single-thread driver RBergomiST::ComputePayoffRT_single
(rBergomi/RBergomi.cpp:174-180, loop rBergomi/text/tests.cpp:283-291) for the
deterministic inputs W1_i = W1perp_i = cos(i*m) (debugFillVector, rBergomi/aux.h:200-203)
and the test parameters of rBergomi/text/tests.cpp:27-37 (option I = 0).
The payoffs are printed with 6 significant digits (std::cout default precision).
Only the numbers are copied; this script is rerun only when the fixture must be rebuilt.
"""
import json
import os
import sys

REF = "/root/reference/rBergomi/testing/debug8.txt~"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rows = []
    with open(REF) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 2:
                rows.append((int(parts[0]), float(parts[1])))
    assert [m for m, _ in rows] == list(range(200)), "unexpected layout"
    fixture = {
        "source": "rBergomi/testing/debug8.txt~ (reference repository)",
        "inputs": "W1[i] = W1perp[i] = cos(i*m), i = 0..N-1, path m = 0..199",
        "params": {"xi": 0.07, "H": 0.07, "eta": 2.2, "rho": -0.9, "T": 0.5, "K": 0.8,
                   "N": 256},
        "significant_digits": 6,
        "m": [m for m, _ in rows],
        "payoff": [p for _, p in rows],
    }
    with open(os.path.join(HERE, "debug8_payoffs.json"), "w") as f:
        json.dump(fixture, f, indent=1)
    print("wrote", len(rows), "payoffs")


if __name__ == "__main__":
    sys.exit(main())
