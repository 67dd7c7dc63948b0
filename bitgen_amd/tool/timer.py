"""Timing table with the reference's row schema.

Rows carry (group, exe, app, name, exec_num, duration, avg_duration, ...) as in
``bitgen/tool/timer.py:58-85`` so the reference's CSV consumers keep working.
Timing brackets synchronise the device on both sides like
``TimerContextManager`` (timer.py:17-31); the synchronisation callback is
injectable so the table itself needs no GPU.
"""
import os
import time
from datetime import datetime


def _default_sync():
    try:
        import torch

        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:  # torch missing: host-only timing
        pass


class Timer:
    def __init__(self):
        self.timing_data = []
        self.group = None
        self.exe = None
        self.app = None
        self.params = None
        self.sync = _default_sync
        self.inst_result = {}
        self.verbose = False

    class TimerContextManager:
        def __init__(self, store, name="unnamed"):
            self.store = store
            self.name = name

        def __enter__(self):
            self.store.sync()
            self.start_time = time.perf_counter() * 1000
            return self

        def __exit__(self, exc_type, exc_val, exc_tb):
            self.store.sync()
            self.end_time = time.perf_counter() * 1000
            self.duration = self.end_time - self.start_time
            self.store.add_timing(self.name, self.duration)

    def set_group(self, group_name):
        self.group = group_name

    def set_exe(self, exe):
        self.exe = exe

    def set_app(self, app):
        self.app = app

    def set_params(self, params):
        self.params = params

    def time(self, name="unnamed"):
        return Timer.TimerContextManager(self, name)

    def _match(self, entry, name):
        if entry["name"] != name or entry["group"] != self.group:
            return False
        if entry["exe"] != self.exe or entry["app"] != self.app:
            return False
        if self.params:
            for k, v in zip(*self.params):
                if entry.get(k) != v:
                    return False
        return True

    def add_timing(self, name, duration):
        if self.verbose:
            print(f"Timer: {name} took {duration:.4f} ms")
        for entry in self.timing_data:
            if self._match(entry, name):
                entry["duration"] += duration
                entry["exec_num"] += 1
                entry["avg_duration"] = entry["duration"] / entry["exec_num"]
                return
        entry = {
            "group": self.group,
            "exe": self.exe,
            "app": self.app,
            "name": name,
            "exec_num": 1,
            "duration": duration,
            "avg_duration": duration,
        }
        if self.params:
            for k, v in zip(*self.params):
                entry[k] = v
        self.timing_data.append(entry)

    def append_data(self, name, data_name, data_value):
        for entry in self.timing_data:
            if self._match(entry, name):
                entry[data_name] = data_value
                return
        raise ValueError(f"Entry {name} not found")

    def get_value(self, name, data_name="duration"):
        for entry in self.timing_data:
            if self._match(entry, name):
                return entry[data_name]
        raise ValueError(f"Entry {name} not found")

    def add_row(self, row):
        self.timing_data.append(row)

    def get_data(self):
        return self.timing_data

    def display_timings(self):
        if not self.timing_data:
            print("No timing data collected.")
            return
        import pandas as pd

        pd.set_option("display.max_columns", 40)
        pd.set_option("display.width", 1000)
        print("\n----- Timing Data -----")
        print(pd.DataFrame(self.timing_data))
        print("-----------------------\n")

    def save_to_file(self, dir_path):
        import pandas as pd

        os.makedirs(dir_path, exist_ok=True)
        stamp = datetime.now().strftime("%Y%m%d-%H%M%S")
        path = os.path.join(dir_path, f"{stamp}.csv")
        pd.DataFrame(self.timing_data).to_csv(path, index=False)
        return path

    def reset_timings(self):
        self.timing_data = []


global_timer = Timer()
