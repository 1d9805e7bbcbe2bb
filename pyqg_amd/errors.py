"""Exceptions of the model layer (pyqg/errors.py:1-10)."""


class PYQGError(Exception):
    pass


class DiagnosticNotFilledError(PYQGError):
    """Raised when a diagnostic is read before anything was accumulated."""

    def __init__(self, dname):
        super().__init__(f"Diagnostic {dname} has not been filled. "
                         "Please adjust tavestart and taveint.")
