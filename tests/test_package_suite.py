"""The tests shipped inside the package (parament_amd/test: the reference package's test_error,
test_wrapper and test_numerics modules restated), collected here as well so that a run of this
directory covers them.  Their own ``gpu`` marks decide where each runs."""
from parament_amd.test.test_error import *  # noqa: F401,F403
from parament_amd.test.test_wrapper import *  # noqa: F401,F403
