"""A transcript executed many times on resident device tensors, with its
segments captured as hipGraphs.

The reference runtime executes a transcript once, op by op
(python_execution/lattica_heal_runtime/runtime.py:10-93).  A serving process
runs the same program over and over, so `TranscriptProgram` keeps what a run
leaves on the device and replays:

* run 1 executes every entry eagerly, which also prepares the per-table device
  data of the backend (NTT plans, validated permutation tables);
* host_to_device of data equal to the previous run's upload reuses the
  resident tensor; changed data of the same shape is copied into the resident
  buffer;
* a segment (SEGMENT_START .. SEGMENT_END) whose entries are all compute or
  view ops is captured as a hipGraph in run 2 and replayed from run 3 on:
  tensors allocated by `empty`/`zeros` inside it keep the buffers of run 1,
  view ops (reshape, expand, ...) are re-applied on the host to fresh tensor
  objects, and every compute op runs inside the graph;
* a graph is dropped (the segment runs eagerly and is captured again) when a
  tensor it reads was re-uploaded under a new identity, or when changed data
  was copied into a tensor that the segment uses as a table (moduli,
  twiddles, permutations), since those are prepared once per buffer version.

Device_to_host and host_to_device entries always run eagerly and are never
inside a graph; a segment containing them, FREE_DEVICE_TENSOR, or an op that
allocates internally (contiguous, moveaxis, pad_single_axis) is not captured.
"""
from __future__ import annotations

import time

import torch

from .datatypes import DeviceOpArgType, ExecutionTranscriptOpType as K
from .runtime import VerificationError, _arg_value

_VIEW_OPS = {"reshape", "expand", "squeeze", "unsqueeze", "flatten", "new_reference", "get_slice"}
_ALLOC_OPS = {"empty", "zeros"}
_COMPUTE_OPS = {
    "_modmul_ttt", "_modmul_ttc", "_modmul_tct", "_modmul_tcc", "_modsum_ttt", "_modsum_ttc", "_modsum_tct",
    "_modsum_tcc", "_mod_tt", "_mod_tc", "_mod_ct", "_modneg_tt", "_modneg_tc", "axis_modsum", "modmul_axis_sum",
    "ntt", "intt", "take_along_axis", "permute", "apply_g_decomp", "abs", "set_const_val",
}
# argument positions that are tables prepared once per buffer version
_TABLE_ARGS = {
    "ntt": (1, 2, 3, 4, 5, 6), "intt": (1, 2, 3, 4, 5, 6), "take_along_axis": (1,), "permute": (1,),
    "_modmul_ttt": (2,), "_modmul_tct": (2,), "_modsum_ttt": (2,), "_modsum_tct": (2,), "axis_modsum": (2,),
    "modmul_axis_sum": (2,), "_modneg_tt": (1,),
}


def _names(arg):
    if arg.arg_type == DeviceOpArgType.DEVICE_TENSOR:
        return [arg.value.inf_name]
    if arg.arg_type == DeviceOpArgType.SHAPE:
        out = []
        for a in arg.value:
            out += _names(a)
        return out
    return []


class _Segment:
    def __init__(self, start, end, name, ops):
        self.start, self.end, self.name = start, end, name
        self.ops = ops  # entry indices of DEVICE_OPs
        self.capturable = True
        self.graph = None
        self.pins = None  # the backend's record of the plans the graph uses
        self.inputs = {}  # name -> (buffer id, offset, dims, strides) at capture
        self.out_arg = {}  # entry index -> position of the returned object among the op's args
        self.tables = set()  # names read in a table role
        self.defined = set()  # names produced inside the segment


class TranscriptProgram:
    def __init__(self, engine, transcript, graphs: bool = True):
        self.engine = engine
        self.entries = list(transcript)
        self.graphs = graphs
        self.runs = 0
        self.uploads = {}  # entry index -> (host tensor copy, device tensor, dims, strides)
        self.allocs = {}  # entry index -> (buffer-owning tensor, dims) from run 1
        self.segments = {}
        self._parse()

    # ---------------------------------------------------------------- setup
    def _parse(self):
        start, name, ops, plain = None, None, [], True
        for i, e in enumerate(self.entries):
            kind, payload = e[0], e[1] if len(e) > 1 else None
            if kind == K.SEGMENT_START:
                start, name, ops, plain = i, payload, [], True
            elif kind == K.SEGMENT_END and start is not None:
                seg = _Segment(start, i, name, ops)
                seg.capturable = plain
                for j in ops:
                    op = self.entries[j][1]
                    if op.name not in _VIEW_OPS | _ALLOC_OPS | _COMPUTE_OPS:
                        seg.capturable = False
                    for pos in _TABLE_ARGS.get(op.name, ()):
                        if pos < len(op.args):
                            seg.tables.update(_names(op.args[pos]))
                    if op.out is not None and op.out.value is not None:
                        seg.defined.add(op.out.value.inf_name)
                self.segments[start] = seg
                start = None
            elif start is not None:
                if kind == K.DEVICE_OP:
                    ops.append(i)
                elif kind is not None:
                    plain = False  # FREE_DEVICE_TENSOR inside the segment

    def _sfx(self, t):
        return self.engine._sfx(t)

    def _alias(self, t, dims, strides, offset):
        return getattr(self.engine.m, f"alias_view_{self._sfx(t)}")(t, list(dims), list(strides), offset)

    # ---------------------------------------------------------------- entries
    def _host_to_device(self, i, op, refs, changed):
        host = op.args[0].value.tensor
        host = host if isinstance(host, torch.Tensor) else torch.as_tensor(host)
        dtype = op.out.value.dtype
        name = op.out.value.inf_name
        prev = self.uploads.get(i)
        if prev is not None:
            h0, dev, dims, strides = prev
            same_layout = host.dtype == h0.dtype and list(host.shape) == list(h0.shape) and \
                host.stride() == h0.stride()
            if same_layout and torch.equal(host, h0):
                refs[name] = self._alias(dev, dims, strides, 0)
                return
            if same_layout and host.to(dtype).stride() == tuple(strides):
                getattr(self.engine.m, f"copy_from_host_{self._sfx(dev)}")(dev, host.to(dtype))
                self.uploads[i] = (host.clone(), dev, dims, strides)
                changed.add(name)
                refs[name] = self._alias(dev, dims, strides, 0)
                return
        dev = self.engine.host_to_device(host, dtype)
        self.uploads[i] = (host.clone(), dev, list(dev.shape), list(dev.strides))
        refs[name] = self._alias(dev, list(dev.shape), list(dev.strides), 0)

    def _eager_op(self, i, op, refs, verify, changed):
        if op.name == "host_to_device":
            self._host_to_device(i, op, refs, changed)
            return
        if op.name == "device_to_host":
            host = self.engine.device_to_host(refs[op.args[0].value.inf_name])
            if verify:
                if not bool((host == torch.as_tensor(op.out.value.tensor)).all()):
                    raise VerificationError(f"device_to_host of {op.args[0].value.inf_name!r} differs")
            return
        args = [_arg_value(refs, a) for a in op.args]
        out = getattr(self.engine, op.name)(*args)
        if op.name in _ALLOC_OPS and i not in self.allocs:
            self.allocs[i] = (out, list(out.shape))
        if op.out is not None and op.out.value is not None:
            refs[op.out.value.inf_name] = out

    def _alloc_view(self, i, op, refs):
        """empty/zeros inside a captured segment: the run-1 buffer under a fresh object."""
        owner, dims = self.allocs[i]
        strides, st = [], 1
        for d in reversed(dims):
            strides.insert(0, st)
            st *= d
        t = self._alias(owner, dims, strides, 0)
        refs[op.out.value.inf_name] = t
        return t

    # ---------------------------------------------------------------- segments
    def _segment_inputs(self, seg, refs):
        """Names the segment reads before (or without) defining them itself."""
        inputs, defined = set(), set()
        for j in seg.ops:
            op = self.entries[j][1]
            for a in op.args:
                inputs.update(n for n in _names(a) if n not in defined)
            if op.out is not None and op.out.value is not None:
                defined.add(op.out.value.inf_name)
        return {n for n in inputs if n in refs}

    def _capture(self, seg, refs):
        m = self.engine.m
        seg.inputs = {n: (refs[n].buffer_id, refs[n].offset, list(refs[n].shape), list(refs[n].strides))
                      for n in self._segment_inputs(seg, refs)}
        seg.out_arg = {}
        torch.cuda.synchronize()
        stream = torch.cuda.Stream()
        g = torch.cuda.CUDAGraph()
        old = m.get_stream()
        m.set_stream(stream.cuda_stream)
        m.capture_begin()
        try:
            with torch.cuda.graph(g, stream=stream):
                for j in seg.ops:
                    op = self.entries[j][1]
                    if op.name in _ALLOC_OPS:
                        t = self._alloc_view(j, op, refs)
                        if op.name == "zeros":
                            getattr(m, f"set_const_val_{self._sfx(t)}")(t, 0)
                        continue
                    args = [_arg_value(refs, a) for a in op.args]
                    out = getattr(self.engine, op.name)(*args)
                    if op.out is not None and op.out.value is not None:
                        refs[op.out.value.inf_name] = out
                        if op.name not in _VIEW_OPS:
                            pos = [k for k, a in enumerate(args) if a is out]
                            seg.out_arg[j] = pos[0] if pos else None
        finally:
            pins = m.capture_end()
            m.set_stream(old)
        seg.graph, seg.pins = g, pins
        g.replay()

    def _replay(self, seg, refs):
        # inputs get fresh objects with their capture-time metadata, so view
        # ops re-applied below start from the same state as in the capture
        for n, (_, offset, dims, strides) in seg.inputs.items():
            refs[n] = self._alias(refs[n], dims, strides, offset)
        for j in seg.ops:
            op = self.entries[j][1]
            if op.name in _ALLOC_OPS:
                self._alloc_view(j, op, refs)
            elif op.name in _VIEW_OPS:
                args = [_arg_value(refs, a) for a in op.args]
                out = getattr(self.engine, op.name)(*args)
                if op.out is not None and op.out.value is not None:
                    refs[op.out.value.inf_name] = out
            elif op.out is not None and op.out.value is not None:
                pos = seg.out_arg.get(j)
                if pos is None:
                    raise RuntimeError(f"segment {seg.name!r}: cannot rebind the result of {op.name}")
                refs[op.out.value.inf_name] = _arg_value(refs, op.args[pos])
        seg.graph.replay()

    @staticmethod
    def _drop_graph(seg):
        """Releases a segment's graph, then the plans it kept (after it ran)."""
        torch.cuda.synchronize()
        seg.graph = None
        seg.pins = None

    def _graph_usable(self, seg, refs, changed):
        """The graph still addresses every input: same buffer, offset and layout,
        and no table it reads was rewritten in this run."""
        if seg.graph is None:
            return False
        for n, (bid, offset, dims, strides) in seg.inputs.items():
            t = refs.get(n)
            if t is None or (t.buffer_id, t.offset, list(t.shape), list(t.strides)) != (bid, offset, dims, strides):
                return False
        return not (seg.tables & changed)

    # ---------------------------------------------------------------- run
    def _input_state(self, seg, refs):
        return {n: (refs[n].buffer_id, refs[n].offset, list(refs[n].shape), list(refs[n].strides))
                for n in self._segment_inputs(seg, refs)}

    def run(self, verify: bool = False) -> dict:
        """One pass over the transcript.  A capturable segment runs eagerly until
        two consecutive runs see the same inputs (so every table the backend
        prepares is already prepared), is then captured, and replayed while its
        inputs keep their buffers and no table it reads is rewritten."""
        refs = {}
        changed = set()
        stats = {"ops": 0, "graph_replays": 0, "captures": 0, "eager_segments": 0, "segments": []}
        t0 = time.perf_counter()
        i = 0
        n = len(self.entries)
        while i < n:
            entry = self.entries[i]
            kind, payload = entry[0], entry[1] if len(entry) > 1 else None
            seg = self.segments.get(i) if kind == K.SEGMENT_START else None
            if seg is not None:
                ts = time.perf_counter()
                mode = "eager"
                if self.graphs and seg.capturable:
                    state = self._input_state(seg, refs)
                    if seg.graph is not None and (seg.tables & changed):
                        self._drop_graph(seg)  # the graph holds the old tables' prepared data
                    if seg.graph is not None and state == seg.inputs:
                        mode = "replay"
                    elif getattr(seg, "eager_inputs", None) == state and not (seg.tables & changed):
                        mode = "capture"
                    else:
                        seg.eager_inputs = state
                if mode == "replay":
                    self._replay(seg, refs)
                    stats["graph_replays"] += 1
                elif mode == "capture":
                    self._capture(seg, refs)
                    stats["captures"] += 1
                else:
                    for j in seg.ops:
                        self._eager_op(j, self.entries[j][1], refs, verify, changed)
                    stats["eager_segments"] += 1
                stats["ops"] += len(seg.ops)
                stats["segments"].append((seg.name, mode, time.perf_counter() - ts))
                i = seg.end + 1
                continue
            if kind == K.DEVICE_OP:
                self._eager_op(i, payload, refs, verify, changed)
                stats["ops"] += 1
            elif kind == K.FREE_DEVICE_TENSOR:
                refs.pop(payload.tensor_name, None)
            elif kind not in (K.SEGMENT_START, K.SEGMENT_END, None):
                raise ValueError(f"Unknown op type: {kind}")
            i += 1
        torch.cuda.synchronize()
        self.runs += 1
        stats["elapsed_s"] = time.perf_counter() - t0
        return stats
