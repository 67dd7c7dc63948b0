from .pass_manager import PassManager  # noqa: F401
from .star_rewrite import star_rewrite_pass  # noqa: F401
from .cc_advance import cc_advance_pass  # noqa: F401
from .remove_alias import remove_alias_pass  # noqa: F401
from .short_circuit import short_circuit_pass  # noqa: F401
from .graph_break import graph_break_pass  # noqa: F401
from .inst_stats import inst_stats_pass, inst_counts  # noqa: F401
