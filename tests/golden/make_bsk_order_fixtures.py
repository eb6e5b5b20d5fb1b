"""Writes bsk_order_fixtures.npz: the reference's BSK NTT-domain order for several radix
cuts, computed by the reference's own function (sw/ntt/ntt_gf64_bsk_order.py:9-35).

Only the cut [5,6] is pinned independently (it equals this engine's NTT-domain order,
checked against the NTT's closed form); the other cuts have no fixture in the reference,
so their vectors come from running that function here.  Run from the repository root:
    python tests/golden/make_bsk_order_fixtures.py /root/reference
"""
import importlib.util
import os
import sys

import numpy as np

CUTS = [[5, 6], [6, 5], [4, 7], [3, 4, 4], [2, 3, 3, 3], [3, 3, 3, 2], [1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1], [11],
        [3, 3], [2, 2, 2]]


def main(ref_root):
    path = os.path.join(ref_root, "sw", "ntt", "ntt_gf64_bsk_order.py")
    spec = importlib.util.spec_from_file_location("ntt_gf64_bsk_order", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    out = {}
    for cut in CUTS:
        out["cut_" + "_".join(map(str, cut))] = np.array(mod.bsk_order(cut), dtype=np.uint32)
    here = os.path.dirname(os.path.abspath(__file__))
    np.savez_compressed(os.path.join(here, "bsk_order_fixtures.npz"), **out)
    print({k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
