"""Generator base: the on-disk layout of a generated kernel module.

Contract kept from the reference (``bitgen/codegen/gen.py:8-37``), because
``from_icgrep.init_kernel`` imports what the generator writes: every file of
a compiled pattern set lives in ``<cwd>/<cfg["__cached_module_path"]>/
kernels/``, and ``kernels/__init__.py`` imports each wrapper module by name
(``import kernels.kernel_0 as kernel_0``)."""
import os

from .. import config as cfg


class Generator:
    def __init__(self, insts_list, var_name_map_list, regex_list=None):
        if len(insts_list) != len(var_name_map_list):
            raise ValueError(f"{len(insts_list)} programs but {len(var_name_map_list)} variable maps")
        if regex_list is not None and not cfg.get_config("use_cached_code") and len(regex_list) != len(insts_list):
            raise ValueError(f"{len(insts_list)} programs but {len(regex_list)} regex groups")
        self.insts_list = insts_list
        self.var_name_map_list = var_name_map_list
        self.regex_list = regex_list

    def close(self):
        pass

    def kernel_dir(self):
        """The module's ``kernels/`` directory (created on first use)."""
        base = cfg.get_config("__cached_module_path")
        if base is None:
            raise RuntimeError("__cached_module_path is not set (from_icgrep.init_cached_module_path sets it)")
        d = os.path.join(os.getcwd(), base, "kernels")
        os.makedirs(d, exist_ok=True)
        return d

    def get_kernel_path(self, kernel_file_name):
        return os.path.join(self.kernel_dir(), kernel_file_name)

    def gen_kernel_module_file(self, kernel_file_name_list):
        """Write ``kernels/__init__.py`` importing the given wrapper files."""
        stems = [os.path.splitext(name)[0] for name in kernel_file_name_list]
        text = "".join(f"import kernels.{s} as {s}\n" for s in stems)
        with open(self.get_kernel_path("__init__.py"), "w") as f:
            f.write(text)
