"""Pass sequencing (``bitgen/passes/pass_manager.py:3-21``)."""


class PassManager:
    def __init__(self):
        self.passes = []

    def add_pass(self, p, enabled=True):
        if enabled:
            self.passes.append(p)

    def print_passes(self):
        for p in self.passes:
            print(p.pass_name)

    def run(self, insts, var_name_map):
        for p in self.passes:
            insts, var_name_map = p.run(insts, var_name_map)
        return insts, var_name_map
