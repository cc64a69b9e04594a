"""Replays a HEAL instruction transcript on the MI355X backend.

Same execution model as the reference runtime
(python_execution/lattica_heal_runtime/runtime.py:10-93): a linear list of
(op type, payload) entries, a name -> device tensor table, host_to_device /
device_to_host with optional exact verification, FREE_DEVICE_TENSOR dropping
a name.  Segments are timed (device-synchronised) when `timing` is set.
"""
from __future__ import annotations

import time

import torch

from .datatypes import DeviceOp, DeviceOpArgType, ExecutionTranscriptOpType


class VerificationError(AssertionError):
    pass


def _arg_value(refs, arg):
    t = arg.arg_type
    if t == DeviceOpArgType.DEVICE_TENSOR:
        return refs[arg.value.inf_name]
    if t == DeviceOpArgType.HOST_TENSOR:
        raise ValueError("HostTensor arguments are only valid for host_to_device")
    if t == DeviceOpArgType.SHAPE:
        return tuple(_arg_value(refs, a) for a in arg.value)
    if t in (DeviceOpArgType.INT, DeviceOpArgType.TENSOR_TYPE):
        return arg.value
    if t == DeviceOpArgType.NONE:
        return None
    if t == DeviceOpArgType.SLICE:
        v = arg.value
        return v if isinstance(v, slice) else slice(v.start, v.stop, v.step)
    if t == DeviceOpArgType.ELLIPSIS:
        return Ellipsis
    raise ValueError(f"Unknown arg type: {t}")


def _run_device_op(engine, refs, op: DeviceOp, verify: bool):
    if op.name == "host_to_device":
        refs[op.out.value.inf_name] = engine.host_to_device(op.args[0].value.tensor, op.out.value.dtype)
        return
    if op.name == "device_to_host":
        host = engine.device_to_host(refs[op.args[0].value.inf_name])
        if verify:
            expected = op.out.value.tensor
            if not bool((host == torch.as_tensor(expected)).all()):
                raise VerificationError(f"device_to_host of {op.args[0].value.inf_name!r} differs from the transcript")
        return
    args = [_arg_value(refs, a) for a in op.args]
    out = getattr(engine, op.name)(*args)
    if op.out is not None and op.out.value is not None:
        refs[op.out.value.inf_name] = out


def run_transcript(engine, transcript, verify: bool = False, timing: bool = False, sync=None) -> dict:
    """Runs every entry; returns {"ops": n, "elapsed_s": t, "segments": [(name, seconds), ...]}."""
    refs: dict = {}
    stats = {"ops": 0, "segments": []}
    seg_name, seg_t0 = None, None
    sync = sync or (lambda: None)
    t0 = time.perf_counter()
    for entry in transcript:
        kind, payload = entry[0], entry[1] if len(entry) > 1 else None
        if kind == ExecutionTranscriptOpType.SEGMENT_START:
            if timing:
                sync()
            seg_name, seg_t0 = payload, time.perf_counter()
        elif kind == ExecutionTranscriptOpType.SEGMENT_END:
            if timing and seg_t0 is not None:
                sync()
                stats["segments"].append((seg_name, time.perf_counter() - seg_t0))
            seg_name, seg_t0 = None, None
        elif kind == ExecutionTranscriptOpType.DEVICE_OP:
            _run_device_op(engine, refs, payload, verify)
            stats["ops"] += 1
        elif kind == ExecutionTranscriptOpType.FREE_DEVICE_TENSOR:
            del refs[payload.tensor_name]
        elif kind is None:
            continue
        else:
            raise ValueError(f"Unknown op type: {kind}")
    sync()
    stats["elapsed_s"] = time.perf_counter() - t0
    return stats
