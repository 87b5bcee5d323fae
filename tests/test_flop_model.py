"""The generated HGP kernels' flop model (pychem_amd/csrc/eri_flops.json),
priced like the reference's VRR/contract/HRR loops, reproduces the per-class
figures derived from the reference's recurrence schedule."""
import json
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# per primitive quartet (VRR + contraction) and per contracted quartet (HRR),
# from Methods/integrals.py:73-191 priced as two_electron_vrr.c:92-109,
# two_electron_contract.c:45, two_electron_hrr.c:82
REFERENCE = {"0000": (3, 0), "1000": (27, 0), "1100": (114, 18), "1010": (129, 0),
             "1111": (1461, 324), "2121": (None, 1224)}


def test_flop_model_matches_reference_schedule():
    d = json.load(open(os.path.join(ROOT, "pychem_amd", "csrc", "eri_flops.json")))
    assert len(d) == 21
    for k, (prim, hrr) in REFERENCE.items():
        if prim is not None:
            assert d[k]["vrr"] + d[k]["contract"] == prim, k
        assert d[k]["hrr"] == hrr, k
    # d classes: the generator emits only the components the targets need, the
    # reference computes whole intermediate classes; within 10 %
    for k, prim in (("2000", 114), ("2200", 647), ("2110", 1192), ("2121", 7014), ("2222", 27687)):
        mine = d[k]["vrr"] + d[k]["contract"]
        assert 0.9 * prim <= mine <= prim, (k, mine, prim)
