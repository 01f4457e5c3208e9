"""The reference's end-to-end suites (restated in tests/reference_suite.py) on two
programs: the reference's own main program on the CPU (which shows that the restated
checks hold for the reference) and this package's command line on the GPU."""
from __future__ import annotations

import pytest

from tests import reference_suite as rs


@pytest.mark.parametrize("check", rs.CHECKS, ids=rs.ids(rs.CHECKS))
def test_reference_main(check, ref_main_run):
    rs.apply(check, ref_main_run)


@pytest.mark.gpu
@pytest.mark.parametrize("check", rs.CHECKS + rs.SLOW_CHECKS, ids=rs.ids(rs.CHECKS + rs.SLOW_CHECKS))
def test_cli(check, cli_run):
    rs.apply(check, cli_run)
