"""The oracle against the reference's own known-answer unit tests (hand-computed expected
collisions, moves and ranks from test/units/simulation_cpu/*.cpp, kept as data in
tests/golden/reference_kat.json)."""
import pytest

import kat_replay
from oracle_backend import OracleBackend

CASES = kat_replay.kat_cases()


def test_fixture_covers_the_reference_suites():
    names = [c["name"] for c in CASES]
    assert sum(n.startswith("Simulation 0") for n in names) == 12
    assert sum(n.startswith("Detect LEF-BAR") for n in names) == 5
    assert sum(n.split(" ")[0] in ("Detect", "Process") and "LEF-LEF" in n for n in names) == 4
    assert sum(n.startswith("LEFs ranking") for n in names) == 2
    assert sum(n.startswith("Adjust LEF moves") for n in names) == 2


@pytest.mark.parametrize("case", CASES, ids=[c["name"][:40] for c in CASES])
def test_oracle_matches_reference_kat(case):
    bad = kat_replay.replay(case, OracleBackend())
    assert not bad, bad
