from .timer import global_timer  # noqa: F401
