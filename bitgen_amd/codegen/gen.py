"""Generator base (``bitgen/codegen/gen.py:8-37``): kernel cache paths under
``cfg["__cached_module_path"]/kernels/`` and the ``kernels/__init__.py``
module writer."""
import os

from .. import config as cfg


class Generator:
    def __init__(self, insts_list, var_name_map_list, regex_list=None):
        assert len(insts_list) == len(var_name_map_list)
        if regex_list is not None and not cfg.get_config("use_cached_code"):
            assert len(insts_list) == len(regex_list)
        self.insts_list = insts_list
        self.var_name_map_list = var_name_map_list
        self.regex_list = regex_list

    def close(self):
        pass

    def get_kernel_path(self, kernel_file_name):
        base = cfg.get_config("__cached_module_path")
        assert base is not None, "__cached_module_path is not set"
        path = os.path.join(os.getcwd(), base, "kernels", kernel_file_name)
        os.makedirs(os.path.dirname(path), exist_ok=True)
        return path

    def gen_kernel_module_file(self, kernel_file_name_list):
        path = self.get_kernel_path("__init__.py")
        with open(path, "w") as f:
            for name in kernel_file_name_list:
                stem = name.split(".")[0]
                f.write(f"import kernels.{stem} as {stem}\n")

    def add_lines(self, file_handle, lines, indent=""):
        if isinstance(lines, list):
            file_handle.write("".join(indent + line + "\n" for line in lines))
        elif isinstance(lines, str):
            file_handle.write(indent + lines + "\n")
        else:
            raise NotImplementedError
