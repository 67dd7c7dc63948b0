"""IR producer entry points with the reference's signatures
(``bitgen/tool/pcre2pablo.py:67-101``).

The reference shells out to icgrep-1.0; here the in-process emitter
``regex2pablo`` writes the pablo file.  ``pcrefile2pablo`` returns an object
with ``returncode``/``stdout``/``stderr`` like the subprocess result the
reference's callers inspect (``bitgen/from_icgrep.py:437-442``).
"""
from dataclasses import dataclass

from .regex2pablo import regexes_to_pablo, RegexSyntaxError
from .. import config as cfg


@dataclass
class ProducerResult:
    returncode: int
    stdout: str = ""
    stderr: str = ""


def _star_mode():
    return cfg.get_config("star_mode", "matchstar")


def pcre2pablo(pcre, pablo):
    with open(pablo, "w") as f:
        f.write(regexes_to_pablo([pcre], star_mode=_star_mode()))
    return pablo


def pcrefile2pablo(pcre_file_path, pablo):
    """``pcre_file_path``: one regex per line without slashes (the group file
    the reference writes at from_icgrep.py:429-434)."""
    try:
        with open(pcre_file_path, "r", encoding="latin-1") as f:
            regexes = [ln.rstrip("\n") for ln in f if ln.strip()]
        text = regexes_to_pablo(regexes, star_mode=_star_mode())
    except (OSError, RegexSyntaxError) as e:
        return ProducerResult(1, "", str(e))
    with open(pablo, "w") as f:
        f.write(text)
    return ProducerResult(0, pablo, "")
