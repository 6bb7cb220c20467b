"""Writes lcg_123456.json: the BSD LCG of libOMM/src/omm_rand.F90:37-55 from the NORAND seed
123456 (omm_rand.F90:28), restated independently of the oracle in pure Python integers."""
import json
import os

x = 123456
states, values = [], []
for _ in range(64):
    x = (1103515245 * x + 12345) % (2 ** 31)
    states.append(x)
    values.append(x / 2 ** 31)
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "lcg_123456.json"), "w") as f:
    json.dump({"seed": 123456, "states": states, "values": values}, f, indent=1)
