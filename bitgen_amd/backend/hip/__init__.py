"""hip backend runtime: ctypes C ABI (runtime.py) and the generated-module
contract (template.py)."""
