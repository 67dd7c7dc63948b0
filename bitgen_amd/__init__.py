"""bitgen_amd — MI355X-native (gfx950) backend for BitGen's bitstream regex compiler.

The package mirrors the reference's import surface for the hot path
(``bitgen.config``, ``bitgen.codegen.create_generator``, ``bitgen.tool.timer``,
``bitgen.from_icgrep``; see bitgen/__init__.py and bitgen/from_icgrep.py in the
reference) and adds the ``hip`` backend: pablo bitstream IR -> passes -> HIP
kernels for gfx950 loaded through the C ABI in ``libbitgen_hip.so``.
"""

__version__ = "0.1.0"
