"""Write oracle vectors for the highest azimuthal orders the HIP fill supports.

    PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_oracle_high_orders.py

The FP64 oracle (oracle/bor_oracle.py) needs about a minute per truncation
order here, too long for a GPU test, so its unmasked CFIE blocks for a small
sphere at m_max = 100 and 126 (seven and eight 16-order tiles of the far
fill, 15 and 16 eight-order chunks of the near band) are computed once and
kept as a fixture.  Only a spread of orders is stored: the first and last
order of every tile, both signs.
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))

import bor_oracle as bo  # noqa: E402

N_ELEMS = 8
FREQ = 3.5e8
ALPHA = 0.5


def orders(m_max: int):
    ms = set()
    for a in range(0, m_max + 1, 16):
        ms.update({a, min(a + 15, m_max)})
    ms.add(m_max)
    return sorted({s * a for a in ms for s in (1, -1)})


def main():
    nodes = bo.sphere_nodes(0.5, N_ELEMS)
    out = {"nodes": nodes, "freq_hz": FREQ, "alpha": ALPHA}
    for mm in (100, 126):
        sysm = bo.ModeSystem(nodes, FREQ, mm, "cfie", ALPHA, streaming=False)
        ms = orders(mm)
        out[f"m_{mm}"] = np.array(ms)
        for m in ms:
            out[f"A_{mm}_{m}"] = ALPHA * sysm.efie_full(m) + (1 - ALPHA) * bo.ETA0 * sysm.mfie_full(m)
        print(mm, len(ms), flush=True)
    np.savez_compressed(os.path.join(HERE, "oracle_sphere8_high_orders.npz"), **out)


if __name__ == "__main__":
    main()
