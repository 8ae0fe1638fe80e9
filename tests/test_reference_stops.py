"""Where the reference stops on its own, against the reference build.

tests/golden/ref_stops.json (written by tests/golden/make_stops.py from the
reference compiled from its own sources) records, per geometry, whether one
reference step completed and, when it did not, the message of its exit(0)
(the set-up of its tangential-jump approximations, or its first step's
Eta_* / Tau_*).  mts_setup_reference_stops must predict both: an exit exactly
when the reference stopped, with the same message.  Host setup only.
"""
import json
import os

import pytest

from conftest import GOLDEN_DIR

CASES = json.load(open(os.path.join(GOLDEN_DIR, "ref_stops.json")))


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_reference_stops_match_the_reference(mts, case):
    phys = mts.physics("const", 10.0, 1.0)
    st = mts.Setup(mts.grid(case["n"] + 1), mts.surface(*case["surface"]), phys, mib=1)
    r = st.reference_stops()
    exits = r["count"][0] + r["count"][1] + r["count"][3]
    assert case["returncode"] == 0
    if case["completed"]:
        assert exits == 0 and r["first_kind"] == -1, r
    else:
        assert exits > 0 and r["first_kind"] in (0, 1, 3), r
        assert r["message"] == case["last_line"]


def test_reference_stops_report_is_consistent(mts):
    """The first exit names a crossing of its direction; counts are per crossing."""
    phys = mts.physics("const", 10.0, 1.0)
    st = mts.Setup(mts.grid(65), mts.surface("ellipsoid", 0.7, 0.5, 0.4, 1 - 2 ** -20), phys,
                   mib=1)
    r = st.reference_stops()
    info = st.info()
    assert r["first_kind"] >= 0
    assert 0 <= r["first_dir"] <= 2
    assert 0 <= r["first_index"] < info.n_cross[r["first_dir"]]
    total = sum(info.n_cross)
    assert all(0 <= c <= 2 * total for c in r["count"])
