"""PCRE file reader (``bitgen/tool/pcre.py:5-41``): one ``/regex/`` per line."""
import random


def read_regex_from_pcre_file(regex_file):
    regexes = []
    with open(regex_file, "r", encoding="latin-1") as f:
        for line in f:
            regex = line.strip()
            if not regex:
                continue
            first, last = regex.find("/"), regex.rfind("/")
            if first == 0 and last > first:
                regexes.append(regex[first + 1:last])
            else:
                raise ValueError(f"Regex requires slashes: {regex}")
    return regexes


def select_regex(regexes, regex_num=-1, random_select=False):
    if regex_num is not None and regex_num > 0:
        if random_select:
            random.seed(42)
            regexes = random.sample(regexes, regex_num)
        else:
            regexes = regexes[:regex_num]
    return regexes
