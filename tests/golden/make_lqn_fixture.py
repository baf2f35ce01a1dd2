#!/usr/bin/env python3
"""Write lqn_reference.json: the level-symmetric constants the reference ships.

Reads /root/reference/src/sn_lqn_table.hpp as TEXT (the numbers of its
generated table, data only) and writes, per order, the level cosines and the
octant points' (triple, weight). The product generates its own table
(ubolt_amd/csrc/lqn_gen.py); tests/test_quadrature.py compares the two.

    python3 tests/golden/make_lqn_fixture.py
"""
import json
import os
import re

SRC = "/root/reference/src/sn_lqn_table.hpp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lqn_reference.json")


def main():
    text = open(SRC).read()
    table = {}
    for m in re.finditer(r"lqn_mu_s(\d+)\[\d+\]\s*=\s*\{([^}]*)\}", text):
        order = int(m.group(1))
        table.setdefault(order, {})["mu"] = [float(v) for v in m.group(2).replace("\n", " ").split(",") if v.strip()]
    for m in re.finditer(r"lqn_points_s(\d+)\[\d+\]\s*=\s*\{(.*?)\n\};", text, re.S):
        order = int(m.group(1))
        pts = re.findall(r"\{\s*(\d+),\s*(\d+),\s*(\d+),\s*([0-9eE.+-]+)\s*\}", m.group(2))
        table[order]["triples"] = [[int(a), int(b), int(c)] for a, b, c, _ in pts]
        table[order]["weights"] = [float(w) for *_, w in pts]
    with open(OUT, "w") as f:
        json.dump({str(k): table[k] for k in sorted(table)}, f, indent=1)
    print("wrote", OUT, "orders", sorted(table))


if __name__ == "__main__":
    main()
