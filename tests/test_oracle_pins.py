"""The oracle passes every pinning check (tests/pins.py)."""
import pytest

import pins


@pytest.mark.parametrize("check", pins.ALL_CHECKS, ids=lambda c: c.__name__)
def test_pin(orc, check):
    failures = check(orc)
    assert not failures, "\n".join(failures)
