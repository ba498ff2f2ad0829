"""The LOCAL temperature protocols (tests/local_protocol.py) against the
reference's own runs with NEQ = NMESS = 0 (tests/golden/ref_local_indep_init,
ref_local_carry_init: L = 6, H = 4, HOT start, T = 4, 6, 8, two configs,
IRINIT = 5): with no sweeps, Local_nematic_OP writes every site's start spin
and local nematic value at each temperature of config 1, and cfen the mean
start energy of the two configs per temperature, so the reseed rule, the
carried start after the vacancy draws and the downward walk are fixed site by
site."""
import os

import numpy as np
import pytest

import oracle as O
from local_protocol import local_nematic, read_locnem, start_spins, temperatures
from refdata import find, read_table

L, H, IRINIT, NCONF = 6, 4, 5, 2


@pytest.mark.parametrize("protocol", ["INDEP", "CARRY"])
def test_local_starts_match_reference(golden_dir, protocol):
    gd = os.path.join(golden_dir, f"ref_local_{protocol.lower()}_init")
    temps = temperatures(4.0, 8.0, 2.0, "HOT")
    assert np.array_equal(temps, [8.0, 6.0, 4.0])
    ntemp = temps.size
    en = read_table(find(gd, "cfen"))
    assert np.array_equal(en[:, 0], temps)
    starts = []
    for itemp in range(1, ntemp + 1):
        s = [start_spins(L, H, "HOT", protocol, IRINIT, iconf, ntemp, itemp) for iconf in (1, 2)]
        starts.append(s[0])
        head, spin, nem = read_locnem(os.path.join(gd, f"locnem{itemp:04d}.dat"))
        assert head["Temperature"] == [temps[itemp - 1]]
        assert np.array_equal(spin, s[0]) and np.array_equal(nem, local_nematic(L, H, s[0]))
        # cfen: the mean over the configs of each start's energy per site
        e = [O.OracleLattice(L, H, O.corr_field(L, L, H, 3.0, IRINIT + k, True, 2.0)).measurement_ref(s[k])[1]
             for k in range(NCONF)]
        assert abs(en[itemp - 1, 1] - np.mean(e)) < 1e-14
    distinct = len({st.tobytes() for st in starts})
    assert distinct == (ntemp if protocol == "INDEP" else 1)
