"""Minimal logger with pyscf's verbosity levels (the (T) path logs one line).

The reference routes this through pyscf.lib.logger (pyscfad/lib/logger.py:50-88);
the levels and the ``new_logger(obj, verbose)`` idiom are kept so callers that
pass ``verbose=logger.NOTE`` work unchanged.
"""
import sys
import time

QUIET, ERROR, WARN, NOTE, INFO, DEBUG = 0, 1, 2, 3, 4, 5


class Logger:
    def __init__(self, stdout=None, verbose=NOTE):
        self.stdout = stdout if stdout is not None else sys.stdout
        self.verbose = verbose
        self._t0 = (time.process_time(), time.perf_counter())

    def _write(self, level, msg, *args):
        if self.verbose >= level:
            self.stdout.write((msg % args if args else msg) + '\n')
            self.stdout.flush()

    def note(self, msg, *args):
        self._write(NOTE, msg, *args)

    def info(self, msg, *args):
        self._write(INFO, msg, *args)

    def warn(self, msg, *args):
        self._write(WARN, 'WARN: ' + msg, *args)

    def debug(self, msg, *args):
        self._write(DEBUG, msg, *args)

    def timer(self, msg, t0=None):
        t1 = (time.process_time(), time.perf_counter())
        c0, w0 = t0 if t0 is not None else self._t0
        self.info('    CPU time for %s %9.2f sec, wall time %9.2f sec', msg, t1[0] - c0, t1[1] - w0)
        self._t0 = t1
        return t1


def new_logger(obj=None, verbose=None):
    if isinstance(verbose, Logger):
        return verbose
    if verbose is None:
        verbose = getattr(obj, 'verbose', NOTE)
    return Logger(getattr(obj, 'stdout', None), verbose)
