"""Generated-module contract of the hip backend.

Same contract as the reference's ``bitgen/backend/cuda/template/template.py``
(``init_kernel`` at :124-149): ``init_kernel(regex_func_num,
input_stream_tensor, basic_stream_tensor, cfg)`` returns ``(result,
tmp_streams, kernel_handle, run, dyn_stats)`` where

* ``result``: torch.uint64 ``[1, M, G, W]`` match-end streams, LSB-first words
  (``to_reference_layout`` gives the reference's u32 ``[1, M, G, 2W]``);
* ``tmp_streams``: the carry look-back status words ``[M, K, T*C]`` (uint32
  stored as int32, K kernels) -- the fused hip kernels spill no streams;
* ``kernel_handle``: a ``HipGroupSet`` (``.free()`` unloads the modules);
* ``run(basic_stream_tensor)``: launches every (input, group) pair on the
  current torch stream and synchronises it, like CUDA_Kernel.run
  (cuda_kernel.py:411-427).  Stores are plain (not ``|=``), so the caller
  need not zero ``result`` first;
* ``dyn_stats``: None, or (kernels compiled with ``inst_stats_dyn``, the
  reference's ``--pass-inst-stats-dyn``) int64 ``[1, M, K, 1 + T]`` filled by
  every ``run``: ``[..., 0]`` passes summed over the kernel's tiles (1 per
  tile, more where carries forced replays; the reference's iteration
  count), ``[..., 1 + t]`` zero / byte-pair tests that fell through in tile
  t's exact pass (the dynamic work per tile; T = the most tiles of any
  kernel, shorter rows are zero-padded).
"""
import os

from . import runtime as rt

WAVES_PER_BLOCK = 4


def _torch():
    import torch

    return torch


class HipGroupSet:
    """Loaded kernels of one pattern set plus their launch state.  A kernel
    computes one group, or several consecutive groups merged into one
    program (meta "groups"); results are always per group."""

    def __init__(self, metas, base_dir, device=0):
        torch = _torch()
        rt.init(device)
        self.device = device
        self.metas = metas
        self.modules = []
        self.kernels = []
        for m in metas:
            path = m["hsaco"] if os.path.isabs(m["hsaco"]) else os.path.join(base_dir, m["hsaco"])
            mod = rt.Module(path)
            self.modules.append(mod)
            if m.get("kind") == "stream":
                # a group run at stream level: one kernel per step of its plan
                self.kernels.append([mod.kernel(sm["kernel"]) for sm in m["step_kernels"]])
            else:
                self.kernels.append(mod.kernel(m["kernel"]))
        self._workspaces = {}
        # the set's byte-pair filter pre-pass (kernels compiled with shared
        # filters): its kernel, a tile's answer words, one answer buffer per
        # stream length
        f = next((m["filter"] for m in metas if m.get("filter")), None)
        self.filter_row = 0
        self.filter_kernel = None
        self.filter_grid_blocks = 0  # 0 = the pre-pass blocks that fit at once
        self._filters = {}
        if f is not None:
            fpath = f["hsaco"] if os.path.isabs(f["hsaco"]) else os.path.join(base_dir, f["hsaco"])
            fmod = rt.Module(fpath)
            self.modules.append(fmod)
            self.filter_kernel = fmod.kernel(f["kernel"])
            self.filter_row = int(f["row"])
            self.filter_meta = next(m for m in metas if m.get("filter_row"))
        # first group of each kernel (metas without "groups" hold one group)
        self.first_group = []
        g = 0
        for m in metas:
            self.first_group.append(m.get("first_group", g))
            g = self.first_group[-1] + m.get("n_out", 1)
        dev = torch.device("cuda", device)
        # one control block per kernel (tile queues, error flag, launch epoch):
        # launches of a kernel are ordered on one stream, kernels never share one
        self.ctl = torch.zeros((len(metas), rt.CTL_WORDS), dtype=torch.int32, device=dev)
        self.n_queues = 0  # 0 = the runtime's default
        self.take = 0  # tiles per queue take, 0 = the runtime's default
        self.grid_blocks = 0  # 0 = as many blocks as fit on the device

    @property
    def n_kernels(self):
        return len(self.metas)

    @property
    def n_groups(self):
        return sum(m.get("n_out", 1) for m in self.metas)

    @staticmethod
    def n_tiles(meta, n_words):
        s0 = 64 - meta["HR"]
        if n_words <= s0:
            return 1
        return 1 + (n_words - s0 + meta["S"] - 1) // meta["S"]

    def status_words(self, n_words):
        per = []
        for m in self.metas:
            for sm in m.get("step_kernels", [m]):
                per.append(self.n_tiles(sm, n_words) * sm["status_slots"])
        return max(per)

    def alloc_status(self, n_inputs, n_words):
        """Zeroed look-back status words [n_inputs, kernels, words] for buffers
        of n_words words.  A status buffer belongs to this group set: its words
        are tagged with the epochs of this set's control blocks."""
        torch = _torch()
        return torch.zeros((n_inputs, self.n_kernels, self.status_words(n_words)), dtype=torch.int32,
                           device=torch.device("cuda", self.device))

    def _filter_buffer(self, n_words):
        torch = _torch()
        buf = self._filters.get(n_words)
        if buf is None:
            n = self.n_tiles(self.filter_meta, n_words) * self.filter_row
            buf = torch.empty(max(1, n), dtype=torch.int32, device=torch.device("cuda", self.device))
            self._filters[n_words] = buf
        return buf

    def launch_filter(self, basis_ptr, n_words, stream):
        """Enqueue the set's byte-pair filter pre-pass over one buffer, before
        its group kernels on the same stream (no-op for a set without shared
        filters).  The answers land in a buffer per stream length, so the
        kernels of one buffer must follow before the next buffer's pre-pass
        (stream order keeps that)."""
        if self.filter_kernel is None:
            return
        buf = self._filter_buffer(n_words)
        rt.filter_launch(self.filter_kernel, basis_ptr, n_words, self.n_tiles(self.filter_meta, n_words),
                         buf.data_ptr(), stream, self.filter_grid_blocks)

    def launch(self, k, basis_ptr, result_ptr, status_ptr, n_words, stream):
        """Enqueue kernel k over one buffer (asynchronous).  ``result_ptr`` is
        the stream of the kernel's first group; its j-th group's stream
        follows n_words words later for each j.  A kernel with shared
        byte-pair answers reads those of the last launch_filter of this
        stream length."""
        meta = self.metas[k]
        args = rt.GroupArgs()
        if meta.get("filter_row"):
            if n_words not in self._filters:
                raise rt.HipRuntimeError(f"kernel {k} reads shared byte-pair answers: call launch_filter first")
            args.filter = self._filters[n_words].data_ptr()
        args.basis = basis_ptr
        args.result = result_ptr
        args.status = status_ptr
        args.ctl = self.ctl[k].data_ptr()
        args.n_words = n_words
        args.n_tiles = self.n_tiles(meta, n_words)
        args.n_queues = self.n_queues
        args.n_carry = meta["status_slots"]
        args.take = self.take
        rt.group_launch(self.kernels[k], args, stream, self.grid_blocks)

    def run(self, basis, result, status, stream=None):
        """basis u64 [S=1, M, 8, W]; result u64 [1, M, G, W]; status int32 [M, kernels, *]."""
        torch = _torch()
        if basis.dim() != 4 or basis.dtype != torch.uint64 or basis.size(2) != 8:
            raise ValueError(f"basis must be uint64 [1, M, 8, W], got {tuple(basis.shape)} {basis.dtype}")
        if basis.size(0) != 1:
            raise NotImplementedError("split inputs (dimension 0 > 1) are not supported")
        M, W = basis.size(1), basis.size(3)
        G, K = self.n_groups, self.n_kernels
        if tuple(result.shape) != (1, M, G, W) or result.dtype != torch.uint64:
            raise ValueError(f"result must be uint64 [1, {M}, {G}, {W}]")
        if status.dim() != 3 or status.size(0) != M or status.size(1) != K:
            raise ValueError(f"status must be int32 [{M}, {K}, words] (alloc_status)")
        if not (basis.is_contiguous() and result.is_contiguous() and status.is_contiguous()):
            raise ValueError("tensors must be contiguous")
        s = rt.current_stream_ptr(stream)
        for m in range(M):
            bptr = basis.data_ptr() + m * 8 * W * 8
            self.launch_filter(bptr, W, s)
            for k in range(K):
                rptr = result.data_ptr() + (m * G + self.first_group[k]) * W * 8
                sptr = status.data_ptr() + (m * K + k) * status.size(2) * 4
                if self.metas[k].get("kind") == "stream":
                    self._run_stream(k, basis[0, m], result[0, m, self.first_group[k]], sptr, W, stream)
                else:
                    self.launch(k, bptr, rptr, sptr, W, s)

    # -- groups run at stream level (graph breaks, while loops) -----------------
    # Loop guard: a while loop whose condition never empties is an error, as
    # in the oracle (a stream of n positions cannot need more iterations for
    # the monotone loops regexes produce; the bound is generous)
    LOOP_GUARD_FACTOR = 4

    def _workspace(self, k, W):
        torch = _torch()
        key = (k, W)
        ws = self._workspaces.get(key)
        if ws is None:
            ws = torch.zeros((self.metas[k]["n_slots"], W), dtype=torch.uint64,
                             device=torch.device("cuda", self.device))
            self._workspaces[key] = ws
        return ws

    def _run_stream(self, k, basis_m, result_row, status_ptr, W, stream=None):
        """Run group k's plan over one input: the basis is copied into
        workspace slots 0-7, each kernel step reads workspace slots and writes
        fresh ones, loops run while their condition slot has a set bit (one
        host check per iteration), and the result slot is copied out.  All on
        the given (or current) stream; the loop checks synchronise it."""
        torch = _torch()
        meta = self.metas[k]
        ws = self._workspace(k, W)
        st = stream if stream is not None else torch.cuda.current_stream(self.device)
        sptr = rt.current_stream_ptr(st)
        kernels = self.kernels[k]
        ctl = self.ctl[k].data_ptr()
        guard = self.LOOP_GUARD_FACTOR * 64 * W + 16

        def launch(i, out_base):
            sm = meta["step_kernels"][i]
            args = rt.GroupArgs()
            args.basis = ws.data_ptr()
            args.result = ws.data_ptr() + out_base * W * 8
            args.status = status_ptr
            # the steps share the group's control block: each launch advances
            # its epoch, so no step ever reads another step's status words
            args.ctl = ctl
            args.n_words = W
            args.n_tiles = self.n_tiles(sm, W)
            args.n_queues = self.n_queues
            args.n_carry = sm["status_slots"]
            args.take = self.take
            rt.group_launch(kernels[i], args, sptr, self.grid_blocks)

        def execute(plan):
            for step in plan:
                op = step[0]
                if op == "kernel":
                    launch(step[1], step[2])
                elif op == "copy":
                    ws[step[2]].copy_(ws[step[1]])
                elif op == "zero":
                    ws[step[1]].zero_()
                elif op == "loop":
                    n = 0
                    # (torch has no any() for uint64 on the device: view as int64)
                    while bool(ws[step[1]].view(torch.int64).any()):
                        n += 1
                        if n > guard:
                            raise rt.HipRuntimeError(f"while loop of group {meta['gid']} did not end "
                                                     f"after {guard} iterations")
                        execute(step[2])
                else:
                    raise ValueError(f"unknown plan step {op!r}")

        with torch.cuda.stream(st):
            ws[:8].copy_(basis_m)
            execute(meta["plan"])
            result_row.copy_(ws[meta["result_slot"]])

    def dyn_stats(self, status, n_words):
        """Per-tile statistics of kernels compiled with ``inst_stats_dyn``
        (see the module docstring): int64 [M, K, 1 + T] on the device, from
        the statistics slot (index n_carry) of each tile's status words."""
        torch = _torch()
        M = status.size(0)
        tiles = [self.n_tiles(m, n_words) if m.get("dyn_stats") else 0 for m in self.metas]
        out = torch.zeros((M, self.n_kernels, 1 + max(tiles + [0])), dtype=torch.int64, device=status.device)
        for k, m in enumerate(self.metas):
            T = tiles[k]
            if not T:
                continue
            C = m["status_slots"]
            words = status[:, k, :T * C].reshape(M, T, C)[:, :, C - 1].to(torch.int64) & 0xFFFFFFFF
            out[:, k, 1:1 + T] = words & 0xFFFFF
            out[:, k, 0] = (words >> 20).sum(dim=1)
        return out

    def copy_errors(self, stream=None):
        """Enqueue a copy of the kernels' error flags into pinned host memory
        (after the launches on the same stream); ``raise_errors`` reads it once
        the stream has synchronised, without a blocking copy of its own."""
        torch = _torch()
        if getattr(self, "_err_host", None) is None:
            self._err_host = torch.zeros(self.n_kernels, dtype=torch.int32, pin_memory=True)
        with torch.cuda.stream(stream or torch.cuda.current_stream(self.device)):
            self._err_host.copy_(self.ctl[:, rt.CTL_ERROR], non_blocking=True)

    def raise_errors(self, flags=None):
        """Raise if a carry look-back gave up; the flags are cleared first, so
        that later runs report only their own errors."""
        flags = self._err_host if flags is None else flags
        if int(flags.abs().sum()) != 0:
            bad = [k for k in range(self.n_kernels) if int(flags[k]) != 0]
            self.ctl[:, rt.CTL_ERROR].zero_()
            flags.zero_()
            raise rt.HipRuntimeError(f"carry look-back timed out in kernels {bad}; results are invalid")

    def check_errors(self):
        """Raise if any launch's carry look-back gave up (call after a sync)."""
        self.raise_errors(self.ctl[:, rt.CTL_ERROR].cpu())

    def info(self):
        out = []
        for mod, m in zip(self.modules, self.metas):
            names = [sm["kernel"] for sm in m["step_kernels"]] if m.get("kind") == "stream" else [m["kernel"]]
            infos = [mod.info(n) for n in names]
            out.append({f: max(i[f] for i in infos) for f in infos[0]})
        return out

    def free(self):
        for mod in self.modules:
            mod.free()
        self.modules = []
        self.kernels = []


def init_kernel(groups_meta, base_dir, regex_func_num, input_stream_tensor, basic_stream_tensor=None, cfg={}):
    torch = _torch()
    if basic_stream_tensor is None:
        raise ValueError("the hip backend needs the transposed basis tensor")
    basis = basic_stream_tensor
    if basis.dim() != 4 or basis.dtype != torch.uint64:
        raise ValueError("basic_stream_tensor must be uint64 [1, M, 8, W]")
    device = basis.device.index if basis.device.index is not None else 0
    ks = HipGroupSet(groups_meta, base_dir, device)
    if regex_func_num != ks.n_groups:
        raise ValueError(f"regex_func_num {regex_func_num} != {ks.n_groups} generated groups")
    M, W = basis.size(1), basis.size(3)
    result = torch.zeros((1, M, ks.n_groups, W), dtype=torch.uint64, device=basis.device)
    status = ks.alloc_status(M, W)
    dyn = None
    if any(m.get("dyn_stats") for m in groups_meta):
        dyn = ks.dyn_stats(status, W).unsqueeze(0)

    def run(basic_stream_tensor):
        ks.run(basic_stream_tensor, result, status)
        if dyn is not None:
            dyn[0].copy_(ks.dyn_stats(status, W))
        ks.copy_errors()
        rt.stream_sync(rt.current_stream_ptr())
        ks.raise_errors()

    return result, status, ks, run, dyn
