"""Logger singleton; level from the ``LOG`` environment variable.

Mirrors ``bitgen/log.py:18-93`` (``MyLogger.debug/info/warning/error``) using the
standard library only (colorlog is not installed here).
"""
import logging
import os
import sys


class _Logger:
    def __init__(self):
        self._log = logging.getLogger("bitgen_amd")
        if not self._log.handlers:
            h = logging.StreamHandler(sys.stderr)
            h.setFormatter(logging.Formatter("[%(levelname)s] %(message)s"))
            self._log.addHandler(h)
        self._log.setLevel(os.environ.get("LOG", "WARNING").upper())
        self._log.propagate = False

    def debug(self, msg):
        self._log.debug(msg)

    def info(self, msg):
        self._log.info(msg)

    def warning(self, msg):
        self._log.warning(msg)

    def error(self, msg):
        self._log.error(msg)


MyLogger = _Logger()
