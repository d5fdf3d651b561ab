"""Exception types of the step, named as in ``betse/exceptions.py``."""


class BetseException(Exception):
    """Base of every exception this package raises."""


class BetseSimException(BetseException):
    """Simulation-level error (``betse/exceptions.py`` ``BetseSimException``)."""


class BetseSimUnstableException(BetseSimException):
    """A NaN appeared in Vmem or a cell concentration.

    Raised where the reference raises it from ``check_v`` / ``no_negs``
    (``sim_toolbox.py:332-344, 475-503``; class at ``exceptions.py:625``).
    ``step`` is the 0-based index, within the phase, of the step that produced
    the NaN, or ``None`` if unknown.
    """

    def __init__(self, message: str, step=None):
        super().__init__(message)
        self.step = step


class BetseNativeError(BetseException):
    """The HIP library rejected a call (invalid argument or HIP error)."""
