"""SIGINT handling for long solves (reference: opymize/tools/util.py:4-39)."""
import signal


class GracefulInterruptHandler(object):
    """Context manager that turns SIGINT into a flag the solver polls, so an
    interrupted solve returns its last checked state instead of raising."""

    def __init__(self, sig=signal.SIGINT):
        self.sig = sig

    def __enter__(self):
        self.interrupted = False
        self.released = False
        try:
            self.original_handler = signal.getsignal(self.sig)
            signal.signal(self.sig, self.handle)
            self._installed = True
        except ValueError:  # not in the main thread: nothing to intercept
            self._installed = False
        return self

    def __exit__(self, type, value, tb):
        self.release()

    def handle(self, signum, frame):
        self.release()
        self.interrupted = True

    def release(self):
        if self.released:
            return False
        if self._installed:
            # a handler installed outside Python reads back as None; restore the default then
            orig = self.original_handler if self.original_handler is not None else signal.SIG_DFL
            signal.signal(self.sig, orig)
        self.released = True
        return True
