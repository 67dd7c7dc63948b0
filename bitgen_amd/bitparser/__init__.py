from .bitstream_parser import BitstreamParser, PabloSyntaxError  # noqa: F401
