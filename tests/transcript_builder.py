"""Builds small HEAL instruction transcripts for tests (test infrastructure)."""
from __future__ import annotations

import torch

from heal_amd.runtime.datatypes import (DeviceOp, DeviceOpArg, DeviceOpArgType as A, DeviceTensorPointer,
                                        ExecutionTranscript, ExecutionTranscriptOpType as K, FreeDeviceTensor,
                                        HostTensor)


class Builder:
    def __init__(self):
        self.ops = []
        self.n = 0

    def _ptr(self, dtype):
        self.n += 1
        return DeviceTensorPointer(dtype, f"t{self.n}")

    def segment(self, name):
        self.ops.append((K.SEGMENT_START, name))

    def end(self):
        self.ops.append((K.SEGMENT_END, None))

    def h2d(self, tensor):
        p = self._ptr(tensor.dtype)
        self.ops.append((K.DEVICE_OP, DeviceOp("host_to_device", [DeviceOpArg(A.HOST_TENSOR, HostTensor(tensor))],
                                               DeviceOpArg(A.DEVICE_TENSOR, p))))
        return p

    def empty(self, shape, dtype=torch.int64):
        p = self._ptr(dtype)
        self.ops.append((K.DEVICE_OP, DeviceOp("empty", [DeviceOpArg(A.SHAPE, [DeviceOpArg(A.INT, s) for s in shape]),
                                                         DeviceOpArg(A.TENSOR_TYPE, dtype)],
                                               DeviceOpArg(A.DEVICE_TENSOR, p))))
        return p

    def zeros(self, shape, dtype=torch.int64):
        p = self._ptr(dtype)
        self.ops.append((K.DEVICE_OP, DeviceOp("zeros", [DeviceOpArg(A.SHAPE, [DeviceOpArg(A.INT, s) for s in shape]),
                                                         DeviceOpArg(A.TENSOR_TYPE, dtype)],
                                               DeviceOpArg(A.DEVICE_TENSOR, p))))
        return p

    def op(self, name, *args, out=None):
        conv = []
        for a in args:
            if isinstance(a, DeviceTensorPointer):
                conv.append(DeviceOpArg(A.DEVICE_TENSOR, a))
            elif a is None:
                conv.append(DeviceOpArg(A.NONE, None))
            elif isinstance(a, bool):
                conv.append(DeviceOpArg(A.INT, a))
            elif isinstance(a, int):
                conv.append(DeviceOpArg(A.INT, a))
            elif isinstance(a, (list, tuple)):
                conv.append(DeviceOpArg(A.SHAPE, [DeviceOpArg(A.INT, s) for s in a]))
            else:
                raise TypeError(a)
        self.ops.append((K.DEVICE_OP, DeviceOp(name, conv, DeviceOpArg(A.DEVICE_TENSOR, out) if out else None)))
        return out

    def d2h(self, ptr, expected):
        self.ops.append((K.DEVICE_OP, DeviceOp("device_to_host", [DeviceOpArg(A.DEVICE_TENSOR, ptr)],
                                               DeviceOpArg(A.HOST_TENSOR, HostTensor(expected)))))

    def free(self, ptr):
        self.ops.append((K.FREE_DEVICE_TENSOR, FreeDeviceTensor(ptr.inf_name)))

    def transcript(self):
        return ExecutionTranscript(list(self.ops))
