from .gen import Generator  # noqa: F401
from .gen_hip import HipGenerator, GroupProgram, HipCodegenError  # noqa: F401
from .create_gen import create_generator  # noqa: F401
