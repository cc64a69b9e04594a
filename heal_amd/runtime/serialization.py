"""JSON codec of HEAL transcripts.

Objects are dicts tagged with "__type__"; tensors are base64 payloads of
torch.save / numpy.save (format of the reference's
python_execution/lattica_heal_runtime/serialization.py:15-148).  Payloads are
decoded only with loaders that execute nothing from the file:
torch.load(weights_only=True) and numpy.load(allow_pickle=False).
"""
from __future__ import annotations

import base64
import io
import json

import numpy as np
import torch

from .datatypes import (DeviceOp, DeviceOpArg, DeviceOpArgType, DeviceTensorPointer, ExecutionTranscript,
                        ExecutionTranscriptOpType, FreeDeviceTensor, HostTensor)


def encode_tensor(tensor) -> dict:
    buf = io.BytesIO()
    if isinstance(tensor, torch.Tensor):
        torch.save(tensor, buf)
        tag = "torch"
    elif isinstance(tensor, np.ndarray):
        np.save(buf, tensor, allow_pickle=False)
        tag = "numpy"
    else:
        raise TypeError(f"cannot encode {type(tensor)}")
    return {"type": tag, "data": base64.b64encode(buf.getvalue()).decode("ascii")}


def decode_tensor(obj: dict):
    raw = io.BytesIO(base64.b64decode(obj["data"]))
    if obj["type"] == "torch":
        return torch.load(raw, weights_only=True)
    if obj["type"] == "numpy":
        return np.load(raw, allow_pickle=False)
    raise ValueError(f"unknown tensor payload type {obj['type']!r}")


def _dtype_name(dt) -> str:
    return str(dt).split(".")[-1]


class TranscriptEncoder(json.JSONEncoder):
    def default(self, o):
        if isinstance(o, DeviceOpArgType):
            return {"__type__": "DeviceOpArgType", "value": o.name}
        if isinstance(o, ExecutionTranscriptOpType):
            return {"__type__": "ExecutionTranscriptOpType", "value": o.name}
        if isinstance(o, DeviceTensorPointer):
            return {"__type__": "DeviceTensorPointer", "dtype": str(o.dtype), "inf_name": o.inf_name}
        if isinstance(o, HostTensor):
            return {"__type__": "HostTensor", "tensor_base64": encode_tensor(o.tensor)}
        if isinstance(o, DeviceOpArg):
            return {"__type__": "DeviceOpArg", "arg_type": o.arg_type, "value": o.value}
        if isinstance(o, DeviceOp):
            return {"__type__": "DeviceOp", "name": o.name, "args": o.args, "out": o.out}
        if isinstance(o, FreeDeviceTensor):
            return {"__type__": "FreeDeviceTensor", "tensor_name": o.tensor_name}
        if isinstance(o, ExecutionTranscript):
            return {"__type__": "ExecutionTranscript", "transcript": o.transcript}
        if isinstance(o, slice):
            return {"__type__": "slice", "start": o.start, "stop": o.stop, "step": o.step}
        if o is Ellipsis:
            return {"__type__": "ellipsis"}
        if isinstance(o, torch.dtype):
            return {"__type__": "dtype", "value": _dtype_name(o)}
        if isinstance(o, (complex, np.complexfloating)):
            return {"__type__": "complex", "real": float(o.real), "imag": float(o.imag)}
        return super().default(o)


def _decode(d: dict):
    t = d.get("__type__")
    if t is None:
        return d
    if t == "DeviceOpArgType":
        return DeviceOpArgType[d["value"]]
    if t == "ExecutionTranscriptOpType":
        return ExecutionTranscriptOpType[d["value"]] if d["value"] in ExecutionTranscriptOpType.__members__ else None
    if t == "DeviceTensorPointer":
        return DeviceTensorPointer(getattr(torch, d["dtype"].split(".")[-1]), d["inf_name"])
    if t == "HostTensor":
        return HostTensor(decode_tensor(d["tensor_base64"]))
    if t == "DeviceOpArg":
        return DeviceOpArg(d["arg_type"], d["value"])
    if t == "DeviceOp":
        return DeviceOp(d["name"], d["args"], d["out"])
    if t == "FreeDeviceTensor":
        return FreeDeviceTensor(d["tensor_name"])
    if t == "slice":
        return slice(d["start"], d["stop"], d["step"])
    if t == "ellipsis":
        return Ellipsis
    if t == "ExecutionTranscript":
        return ExecutionTranscript([tuple(item) for item in d["transcript"]])
    if t == "dtype":
        return getattr(torch, d["value"])
    if t == "complex":
        return complex(d["real"], d["imag"])
    raise ValueError(f"unknown transcript object type {t!r}")


def dumps(obj) -> str:
    return json.dumps(obj, cls=TranscriptEncoder)


def loads(text: str):
    return json.loads(text, object_hook=_decode)


def save_transcript_to_json(transcript, filename: str) -> None:
    with open(filename, "w") as f:
        json.dump(transcript, f, cls=TranscriptEncoder)


def load_transcript_from_json(filename: str):
    with open(filename) as f:
        return json.load(f, object_hook=_decode)
