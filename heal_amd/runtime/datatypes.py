"""Transcript data types.

The class names, enum member names and enum values are the on-disk vocabulary
of HEAL instruction transcripts (reference:
python_execution/lattica_heal_runtime/datatypes.py:4-41); they are kept
identical so that transcripts written for the reference load here.
"""
from __future__ import annotations

from enum import Enum
from typing import Any


class DeviceTensorPointer:
    """Names a device tensor of the transcript (dtype + unique name)."""

    def __init__(self, dtype=None, inf_name: str | None = None):
        self.dtype = dtype
        self.inf_name = inf_name


class HostTensor:
    def __init__(self, tensor=None):
        self.tensor = tensor


class DeviceOpArgType(Enum):
    HOST_TENSOR = "HostTensor"
    DEVICE_TENSOR = "DeviceTensor"
    SHAPE = "Shape"
    INT = "int"
    NONE = "None"
    TENSOR_TYPE = "tensor_type"
    SLICE = "slice"
    ELLIPSIS = "Ellipsis"


class DeviceOpArg:
    def __init__(self, arg_type: DeviceOpArgType | None = None, value: Any = None):
        self.arg_type = arg_type
        self.value = value


class DeviceOp:
    def __init__(self, name: str | None = None, args: list | None = None, out: DeviceOpArg | None = None):
        self.name = name
        self.args = args if args is not None else []
        self.out = out


class FreeDeviceTensor:
    def __init__(self, tensor_name: str):
        self.tensor_name = tensor_name


class ExecutionTranscriptOpType(Enum):
    SEGMENT_START = "SegmentStart"
    SEGMENT_END = "SegmentEnd"
    DEVICE_OP = "DeviceOp"
    FREE_DEVICE_TENSOR = "FreeDeviceTensor"


class ExecutionTranscript:
    def __init__(self, transcript: list | None = None):
        self.transcript = transcript if transcript is not None else []
