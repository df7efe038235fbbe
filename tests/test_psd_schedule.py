"""The round-robin schedule of the PSD projection's register-resident
eigenvector rows (psd_project_kernel in pdhcg_amd/csrc/dev_cones.hpp).

Lane L holds the rows (up[L], dn[L]) of pair L.  After each round the up rows
move one lane right, the down rows one lane left, lane 0 keeps its up row,
lane 1 takes lane 0's down row and the last lane takes its own up row (a
single pair stays in place).  This restates those moves and checks the
properties the Jacobi sweep relies on: every round pairs all kp indices
disjointly, every pair of indices meets exactly once in kp - 1 rounds, and the
positions are back at their start after a sweep.  (The rotations themselves
are checked against the oracle by the GPU PSD tests at orders 97..128.)
"""
import itertools


def _move(up, dn):
    np_ = len(up)
    nu = [up[0] if L == 0 else (dn[0] if L == 1 else up[L - 1]) for L in range(np_)]
    nd = [((up[L] if np_ > 1 else dn[L]) if L == np_ - 1 else dn[L + 1]) for L in range(np_)]
    return nu, nd


def test_every_pair_once_per_sweep_and_positions_restored():
    for kp in range(2, 130, 2):
        np_ = kp // 2
        up = [2 * L for L in range(np_)]
        dn = [2 * L + 1 for L in range(np_)]
        seen = set()
        for _ in range(kp - 1):
            assert sorted(up + dn) == list(range(kp)), kp  # a partition of the indices
            for a, b in zip(up, dn):
                pair = (min(a, b), max(a, b))
                assert pair not in seen, (kp, pair)
                seen.add(pair)
            up, dn = _move(up, dn)
        assert seen == set(itertools.combinations(range(kp), 2)), kp
        assert up == [2 * L for L in range(np_)] and dn == [2 * L + 1 for L in range(np_)], kp
