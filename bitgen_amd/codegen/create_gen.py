"""Backend factory (``bitgen/codegen/create_gen.py:6-23``) with the ``hip``
target added.  The reference's cuda/torch/python generators are not part of
this package; asking for them raises."""
from .gen import Generator
from .gen_hip import HipGenerator


def create_generator(type: str):
    if type == "hip":
        return HipGenerator
    if type == "bitstream":
        return Generator
    if type in ("cuda", "torch", "python"):
        raise ValueError(f"backend {type!r} is not provided by bitgen_amd; use --backend=hip")
    raise ValueError("Unknown lower type.")
