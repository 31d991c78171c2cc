"""The build's smoke check (one H·v, a ground state and KPM moments at N=10
against the oracle) as a GPU test, so a change that breaks it shows here."""
import pytest

pytestmark = pytest.mark.gpu


def test_smoke_runs(capsys):
    import __graft_entry__ as ge
    ge.smoke()
    assert 'smoke ok' in capsys.readouterr().out
