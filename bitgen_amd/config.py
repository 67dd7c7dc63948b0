"""Process-global configuration dictionary.

Same contract as the reference's ``bitgen/config.py:15-39``: a module-level
dict read with ``get_config(key, default)`` and written with ``set_config``.
The CLI (``bitgen_amd.run_bitgen``) fills it from argparse flags and every
stage of the pipeline reads it.
"""


class MyConfig:
    def __init__(self):
        self.data = {}

    def __str__(self):
        return str(self.data)

    __repr__ = __str__


GLOBAL_CONFIG = MyConfig()


def get_config(key, default=None):
    return GLOBAL_CONFIG.data.get(key, default)


def set_config(key, value):
    GLOBAL_CONFIG.data[key] = value


def set_from_dict(config_dict):
    GLOBAL_CONFIG.data = dict(config_dict)


def reset_config():
    GLOBAL_CONFIG.data = {}


def get_config_dict():
    return GLOBAL_CONFIG.data


def print_config():
    print(GLOBAL_CONFIG.data)


# Defaults of the hip backend's own knobs.  Reference flags keep the defaults
# that scripts/run_bitgen.py:298-531 gives them.
HIP_DEFAULTS = {
    "backend": "hip",
    "pass_cc_advanced": 1,
    "pass_cc_advanced_max": 8,
    "pass_cc_advanced_merge_cc": 1,
    "pass_short_circuit": 1,
    "pass_short_circuit_start": 2,
    "pass_short_circuit_interval": 1,
    "pass_short_circuit_interval_disable": 0,
    "pass_short_circuit_syncpoint": 1,
    "pass_graph_break": -1,
    "pass_inst_stats": 0,
    "pass_inst_stats_dyn": 0,
    "pass_star_rewrite": 1,
    "block_size": (256, 1, 1),
    "multi_input": 1,
    "split_input": 1,
    "repeat_input": 1,
    "warmup_iters": 4,
    "exec_iters": 8,
}


def apply_hip_defaults(overwrite=False):
    for k, v in HIP_DEFAULTS.items():
        if overwrite or k not in GLOBAL_CONFIG.data:
            GLOBAL_CONFIG.data[k] = v
