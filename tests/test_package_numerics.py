"""The package's numerics tests (parament_amd/test/test_numerics.py, after the reference's
test_numerics.py), collected here so that a run of this directory covers them; all need a GPU."""
import pytest

from parament_amd.test.test_numerics import *  # noqa: F401,F403

pytestmark = pytest.mark.gpu
