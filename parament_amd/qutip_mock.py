"""Stand-in for qutip.Qobj when qutip is not installed
(reference: src/python/pyparament/parament/qutip_mock.py:15-16)."""


class Qobj:
    pass
