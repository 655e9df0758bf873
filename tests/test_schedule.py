"""Launch schedule (a4): same step counts as the reference's generated host
launchers, parsed from compiled/float/*_host.cu into golden/schedule.json."""
import json
import os

import pytest

import pyan5d

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "schedule.json")
TABLE = json.load(open(GOLD))


def reference_schedule(T, bt):
    """Evaluate the parsed reference tables (host.cu:72-77 main loop + remainder)."""
    tab = TABLE[str(bt)]
    nat = (T + bt - 1) // bt
    mismatch = (T % 2) != (nat % 2)
    padr = 1 if (mismatch and T % bt < 2) else 0
    out = [bt] * (T // bt - padr)
    r = T % bt
    if mismatch:
        out += tab["mismatch"][str(r)]
    elif r:
        out += tab["match"][str(r)]
    return out


@pytest.mark.parametrize("bt", sorted(int(b) for b in TABLE))
def test_schedule_matches_reference_tables(bt):
    for T in range(1, 4 * bt + 3):
        assert pyan5d.schedule(T, bt) == reference_schedule(T, bt), (T, bt)


@pytest.mark.parametrize("bt", range(1, 17))
def test_schedule_invariants(bt):
    for T in range(0, 60):
        s = pyan5d.schedule(T, bt)
        assert sum(s) == T
        assert len(s) % 2 == T % 2  # u_T lands in buffer T%2
        assert all(1 <= k <= bt for k in s)


def test_schedule_rejects_bad_args():
    with pytest.raises(pyan5d.An5dError):
        pyan5d.schedule(5, 0)
