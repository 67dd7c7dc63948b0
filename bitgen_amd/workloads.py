"""Seeded synthetic pattern sets and inputs for the benchmark workloads.

Workloads (sizes as named; every generator is deterministic in its seed):

* ``plain_small``: 8 regexes (printable literals of 3-12 bytes, classes such
  as [a-z0-9], 2-3-way alternation, {1,4} repeats, one [^\\n]* each) in 4
  groups over 10,000 bytes of printable ASCII with 1% LF and ~20 planted
  matches per regex.
* ``ids``: 1,000 regexes (600 literals of 4-32 bytes, 300 literal + class +
  bounded repeat, 100 with .* or [^\\r\\n]*) in 16 groups over 100 MiB (70%
  printable, 30% uniform bytes, planted hits at 1e-4 per byte).
* ``genomic``: 512 motifs over {A,C,G,T} with IUPAC classes ([AG], [CT],
  [ACGT]), length 8-24, half gapped m1.{0,20}m2, in 32 groups over DNA text
  (runs of N, LF every 80 bytes).
* ``batched``: pattern sets of 64/256/1,024 ``ids``-mix regexes and buffers
  with log-normal lengths (median 16 KiB, clipped to 1 KiB-8 MiB).
* ``serving``: 8 resident ``ids``-mix sets and Poisson arrivals of jobs with
  log-normal buffers (median 4 KiB, max 64 MiB).
"""
import functools
import random

import numpy as np

from .tool.regex2pablo import parse_regex, CC, Seq, Alt, Rep, Start, End, LF

PRINTABLE = bytes(range(0x20, 0x7F))
_META = set(b"\\^$.|?*+()[]{}/")


def lit(bs):
    """Regex text matching the bytes literally."""
    out = []
    for b in bs:
        if b in _META or b < 0x20 or b > 0x7E:
            out.append(f"\\x{b:02x}")
        else:
            out.append(chr(b))
    return "".join(out)


def _rand_lit(rng, lo, hi, alphabet=PRINTABLE):
    return bytes(rng.choice(alphabet) for _ in range(rng.randint(lo, hi)))


_CLASSES = ["[a-z0-9]", "[A-Z]", "[0-9]", "[a-f0-9]", "[a-z]", "[A-Za-z]", "\\d", "[xyz]", "[^ ]", "\\w"]


# ---- regex generators ----------------------------------------------------------
def regex_plain_small(rng):
    parts = [lit(_rand_lit(rng, 3, 12))]
    k = rng.randint(2, 3)
    parts.append("(" + "|".join(lit(_rand_lit(rng, 2, 5)) for _ in range(k)) + ")")
    parts.append(rng.choice(_CLASSES) + "{1,4}")
    parts.append("[^\\n]*")
    parts.append(lit(_rand_lit(rng, 3, 8)))
    return "".join(parts)


def regex_ids(rng, kind):
    if kind == "literal":
        return lit(_rand_lit(rng, 4, 32))
    if kind == "class":
        a = lit(_rand_lit(rng, 3, 12))
        c = rng.choice(_CLASSES)
        lo = rng.randint(1, 3)
        hi = lo + rng.randint(0, 4)
        return f"{a}{c}{{{lo},{hi}}}" + lit(_rand_lit(rng, 1, 6))
    star = rng.choice([".*", "[^\\r\\n]*"])
    return lit(_rand_lit(rng, 3, 10)) + star + lit(_rand_lit(rng, 3, 10))


def ids_regexes(n, seed, mix=(0.6, 0.3, 0.1)):
    rng = random.Random(seed)
    n_lit = round(n * mix[0])
    n_cls = round(n * mix[1])
    kinds = ["literal"] * n_lit + ["class"] * n_cls + ["star"] * (n - n_lit - n_cls)
    rng.shuffle(kinds)
    return [regex_ids(rng, k) for k in kinds]


_IUPAC = ["[AG]", "[CT]", "[ACGT]"]


def _motif(rng, length):
    out = []
    for _ in range(length):
        if rng.random() < 0.15:
            out.append(rng.choice(_IUPAC))
        else:
            out.append(rng.choice("ACGT"))
    return "".join(out)


def genomic_regexes(n, seed):
    rng = random.Random(seed)
    regs = []
    for i in range(n):
        if i % 2 == 0:
            regs.append(_motif(rng, rng.randint(8, 24)))
        else:
            l1 = rng.randint(4, 12)
            l2 = rng.randint(4, 12)
            regs.append(_motif(rng, l1) + ".{0,20}" + _motif(rng, l2))
    return regs


# ---- sampling from a regex's language ----------------------------------------
@functools.lru_cache(maxsize=None)
def _cc_choices(mask):
    return [b for b in range(256) if (mask >> b) & 1 and b != LF]


@functools.lru_cache(maxsize=4096)
def _parsed(regex):
    return parse_regex(regex)


def _sample(node, rng, out):
    if isinstance(node, CC):
        out.append(rng.choice(_cc_choices(node.mask)))
    elif isinstance(node, Seq):
        for it in node.items:
            _sample(it, rng, out)
    elif isinstance(node, Alt):
        _sample(rng.choice(node.items), rng, out)
    elif isinstance(node, Rep):
        hi = node.hi if node.hi is not None else node.lo + 5
        for _ in range(rng.randint(node.lo, min(hi, node.lo + 3))):
            _sample(node.node, rng, out)
    elif isinstance(node, (Start, End)):
        pass
    else:
        raise TypeError(node)


def sample_match(regex, rng):
    out = []
    _sample(_parsed(regex), rng, out)
    return bytes(out)


def plant(buf, regexes, count, rng, avoid_overlap=True):
    """Write ``count`` samples of random regexes into buf (numpy uint8)."""
    n = buf.size
    used = []
    for _ in range(count):
        r = regexes[rng.randrange(len(regexes))]
        s = sample_match(r, rng)
        if not s or len(s) >= n:
            continue
        for _attempt in range(20):
            off = rng.randrange(0, n - len(s))
            if not avoid_overlap or all(off + len(s) <= a or off >= b for a, b in used):
                buf[off:off + len(s)] = np.frombuffer(s, dtype=np.uint8)
                if avoid_overlap:
                    used.append((off, off + len(s)))
                break
    return buf


# ---- inputs -------------------------------------------------------------------
def text_printable(n, seed, lf_rate=0.01):
    g = np.random.default_rng(seed)
    buf = g.integers(0x20, 0x7F, size=n, dtype=np.uint8)
    buf[g.random(n) < lf_rate] = LF
    return buf


def text_ids(n, seed, printable=0.7):
    g = np.random.default_rng(seed)
    buf = g.integers(0x20, 0x7F, size=n, dtype=np.uint8)
    mask = g.random(n) >= printable
    buf[mask] = g.integers(0, 256, size=int(mask.sum()), dtype=np.uint8)
    return buf


def _ids_lut():
    """Byte for each 16-bit draw u: u < 77*256 (30.1%) gives the uniform byte
    u & 0xFF; the other 45,824 values spread over the 95 printable bytes
    (each 482 or 483 of them)."""
    u = np.arange(65536)
    lut = np.empty(65536, dtype=np.uint8)
    wild = u < 77 * 256
    lut[wild] = u[wild] & 0xFF
    lut[~wild] = 0x20 + ((u[~wild] - 77 * 256) * 95) // (65536 - 77 * 256)
    return lut


def text_ids_fast(n, seed, chunk=1 << 24, workers=8):
    """The ids byte mix (about 70% printable, 30% uniform bytes) from one
    16-bit draw per position through a lookup table, in 16 MiB chunks drawn
    by independent generators (SeedSequence children of ``seed``) on a few
    threads: gigabytes in seconds."""
    import concurrent.futures

    lut = _ids_lut()
    out = np.empty(n, dtype=np.uint8)
    starts = list(range(0, n, chunk))
    seeds = np.random.SeedSequence(seed).spawn(len(starts))

    def fill(k):
        o = starts[k]
        m = min(chunk, n - o)
        idx = np.frombuffer(np.random.default_rng(seeds[k]).bytes(2 * m), dtype=np.uint16)
        out[o:o + m] = lut[idx]

    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as ex:
        list(ex.map(fill, range(len(starts))))
    return out


def text_dna(n, seed, n_run_rate=0.001, line=80):
    """i.i.d. uniform bases, runs of N (length 10-100, starting at about
    n_run_rate/55 of the positions), LF every ``line`` bytes.  Generated in
    64 MiB chunks of uint8 draws so 1 GiB needs no larger temporaries."""
    g = np.random.default_rng(seed)
    lut = np.frombuffer(b"ACGT", dtype=np.uint8)
    buf = np.empty(n, dtype=np.uint8)
    step = 1 << 26
    for o in range(0, n, step):
        m = min(step, n - o)
        buf[o:o + m] = lut[g.integers(0, 4, size=m, dtype=np.uint8)]
    k = int(g.binomial(n, n_run_rate / 55.0)) if n else 0
    starts = np.sort(g.integers(0, max(n, 1), size=k))
    lens = g.integers(10, 101, size=k)
    for s, ln in zip(starts, lens):
        buf[s:s + ln] = ord("N")
    buf[line - 1::line] = LF
    return buf


# ---- workloads -----------------------------------------------------------------
def plain_small(n_bytes=10_000):
    rng = random.Random(1)
    regexes = [regex_plain_small(rng) for _ in range(8)]
    buf = text_printable(n_bytes, 2)
    plant(buf, regexes, 20 * len(regexes), random.Random(2))
    return regexes, 4, buf


def ids(n_bytes=100 * 1024 * 1024, n_regex=1000, n_groups=16, hit_rate=1e-4):
    regexes = ids_regexes(n_regex, 3)
    buf = text_ids(n_bytes, 4)
    plant(buf, regexes, max(1, int(n_bytes * hit_rate)), random.Random(4), avoid_overlap=False)
    return regexes, n_groups, buf


def genomic(n_bytes=1024 * 1024 * 1024, n_regex=512, n_groups=32, hit_rate=1e-6):
    regexes = genomic_regexes(n_regex, 5)
    buf = text_dna(n_bytes, 6)
    plant(buf, regexes, max(1, int(n_bytes * hit_rate)), random.Random(6), avoid_overlap=False)
    return regexes, n_groups, buf


def lognormal_sizes(n, seed, median, lo, hi, total=None, sigma=1.5):
    """Log-normal lengths with the given median, clipped to [lo, hi].  With
    ``total``, sigma is solved (bisection) so the clipped sum is ~total while
    the median stays fixed."""
    z = np.random.default_rng(seed).standard_normal(n)

    def draw(sg):
        return np.clip(np.exp(np.log(median) + sg * z), lo, hi).astype(np.int64)

    if total is not None:
        a, b = 0.01, 8.0
        for _ in range(60):
            m = (a + b) / 2
            if draw(m).sum() < total:
                a = m
            else:
                b = m
        sigma = (a + b) / 2
    return draw(sigma)


def batched_sets(set_sizes=(64, 256, 1024)):
    """Regexes of the batched workload's pattern sets (ids mix, seeds 7, 8, 9)."""
    return [ids_regexes(n, 7 + i) for i, n in enumerate(set_sizes)]


SERVED_SET_SIZES = (32, 64, 128, 256, 512, 1024, 1536, 2048)


def served_sets(set_sizes=SERVED_SET_SIZES):
    """Regexes of the 8 resident pattern sets of the serving workload (ids
    mix, seeds 100 + i), each balanced into groups of about 64 regexes."""
    return [ids_regexes(n, 100 + i) for i, n in enumerate(set_sizes)]


def batched(n_buffers=4096, total_bytes=2 * 2**30, set_sizes=(64, 256, 1024), seed=10):
    """Pattern sets of the ids mix (seeds 7, 8, 9) and buffers with log-normal
    lengths (median 16 KiB, clipped to 1 KiB-8 MiB, scaled to ~total_bytes).
    Jobs: every (set, buffer) pair."""
    sets = batched_sets(set_sizes)
    sizes = lognormal_sizes(n_buffers, seed, 16 * 1024, 1024, 8 * 2**20, total=total_bytes)
    g = np.random.default_rng(seed)
    # one text for all buffers; buffer i is a view of its stretch
    total = int(sizes.sum())
    text = text_ids_fast(total, seed)
    # planted hits (1e-4 per byte) of the regexes of set i % 3, drawn from a
    # pool of 4,096 samples of each set's languages
    rng = random.Random(seed)
    pools = [[np.frombuffer(sample_match(rs[rng.randrange(len(rs))], rng), dtype=np.uint8) for _ in range(4096)]
             for rs in sets]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    buffers = []
    for i, n in enumerate(sizes.tolist()):
        buf = text[offs[i]:offs[i + 1]]
        k = max(1, int(n * 1e-4))
        pos = g.integers(0, max(1, n - 40), size=k)
        pick = g.integers(0, 4096, size=k)
        pool = pools[i % len(sets)]
        for p, s_ in zip(pos.tolist(), pick.tolist()):
            s_ = pool[s_][:n - p]
            buf[p:p + s_.size] = s_
        buffers.append(buf)
    jobs = [(si, b) for si in range(len(sets)) for b in range(n_buffers)]
    return sets, buffers, jobs


def groups_for(n_regex, per_group=64):
    """Groups for a set size: about 64 regexes per group."""
    return max(1, (n_regex + per_group - 1) // per_group)


def serving(n_jobs, rate, seed=11, set_sizes=(32, 64, 128, 256, 512, 1024, 1536, 2048), median=4096,
            max_bytes=64 * 2**20):
    """Poisson arrivals at ``rate`` jobs/s: (arrival time, set index, buffer)."""
    g = np.random.default_rng(seed)
    gaps = g.exponential(1.0 / rate, size=n_jobs)
    arrivals = np.cumsum(gaps)
    sizes = lognormal_sizes(n_jobs, seed + 1, median, 64, max_bytes)
    set_idx = g.integers(0, len(set_sizes), size=n_jobs)
    bufs = [text_ids(int(n), seed * 7919 + i) for i, n in enumerate(sizes)]
    return arrivals, set_idx, bufs
