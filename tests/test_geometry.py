"""The implicit lattice geometry (host only): esp_plan_model reports the implicit layout only
after the partner of every site and slot computed from the (group, slot) records -- sub-cell
shifts stored mod M, wrapped by an unsigned minimum -- has matched the explicit neighbour table
built the reference's way (jij_get, jij.F90:123-446).  Small supercells put whole periods
inside one shift (M = 1: every shift is 0 mod M), larger ones need the wrap at the faces."""
import pytest

from espins_amd import lattice as Lt
from espins_amd.engine import plan_model

CASES = [(k, L, J) for k in ("sc", "bcc", "fcc") for L in (2, 3, 4, 6, 8, 12)
         for J in ((1.0,), (1.0, 0.3), (1.0, 0.3, 0.1))]


@pytest.mark.parametrize("kind,L,J", CASES)
def test_implicit_partners_match_explicit_tables(kind, L, J):
    m = Lt.cubic_model_fast(kind, L, J=J)
    ncol, col, info = plan_model(m)
    assert info["layout"] == "implicit"
    assert info["N"] == m.N and ncol >= 2


def test_aperiodic_colouring_keeps_tables():
    """A greedy (non-periodic) colouring has no group structure: the tables stay."""
    m = Lt.cubic_model_fast("fcc", 5, J=(1.0, 0.3, 0.1))
    _, _, info = plan_model(m)
    assert info["layout"] != "implicit"
