"""Pins the oracle's stateful heuristics with hand-derived known-answer streams.

The region fill-level increment and reset (Scatter.hpp:612-617, 727-735), the
access-block cursor (:620-629, 733-734), the two-pass checklevel switch (:553,
634) and _firstFreePageBased (:796-801, 885) decide where a serialized stream
lands once pages are in use.  The reference holds no vectors for them, so the
streams in oracle_kats.py derive every offset and every metadata value from
the text of Scatter.hpp.  The faithful oracle must reproduce them; each
test-only mutant build of the oracle (scatter_oracle.c, ORC_MUTANT), which
breaks one heuristic the way a misreading would, must fail at least one.
"""
import pytest

from oracle_kats import KATS, run_on_oracle

MUTANTS = {
    1: "no region fill-level increment",
    2: "fill-level increment capped below regionsize",
    3: "no region reset on free",
    4: "region reset when the page empties instead of at P/2/c",
    5: "no access-block cursor advance",
    6: "cursor advance without the > guard",
    7: "no atomicMin of the cursor on free",
    8: "no second pass",
    9: "first pass gated at regionsize instead of 3/4 of it",
    10: "no _firstFreePageBased CAS",
    11: "no _firstFreePageBased atomicMax on free",
}


@pytest.mark.parametrize("reset", [False, True], ids=["keep", "resetfreedpages"])
@pytest.mark.parametrize("kat", KATS, ids=[k.name for k in KATS])
def test_oracle_matches_hand_derived_stream(oracle_mod, kat, reset):
    assert run_on_oracle(oracle_mod, kat, reset) == []


def test_every_mutant_is_targeted():
    targeted = {m for k in KATS for m in k.mutants}
    assert targeted == set(MUTANTS)


@pytest.mark.parametrize("mutant", sorted(MUTANTS), ids=[MUTANTS[m] for m in sorted(MUTANTS)])
def test_mutant_oracle_fails_its_streams(oracle_mod, mutant):
    lib = oracle_mod.mutant_lib(mutant)
    for kat in KATS:
        if mutant in kat.mutants:
            for reset in (False, True):
                diffs = run_on_oracle(oracle_mod, kat, reset, library=lib)
                assert diffs, f"mutant {mutant} ({MUTANTS[mutant]}) passes '{kat.name}'"
