"""CPU tests of the transcript format: every tagged object type survives a
JSON round trip, tensor payloads keep dtype, values and strides, and unsafe
payloads are refused by the loaders."""
from __future__ import annotations

import base64
import io
import json
import pickle

import numpy as np
import pytest
import torch

from heal_amd.runtime import serialization as S
from heal_amd.runtime.datatypes import DeviceOp, ExecutionTranscriptOpType as K
from transcript_builder import Builder


def test_round_trip_all_object_types(tmp_path):
    b = Builder()
    b.segment("seg0")
    strided = torch.arange(24, dtype=torch.int64).reshape(2, 3, 4).permute(2, 0, 1)
    x = b.h2d(strided)
    y = b.empty([2, 3], torch.int32)
    b.op("axis_modsum", x, 1, y, y, out=y)
    b.d2h(y, np.arange(6, dtype=np.int32).reshape(2, 3))
    b.free(x)
    b.end()
    tr = b.transcript()
    tr.transcript.append((K.DEVICE_OP, DeviceOp("get_slice", [], None)))
    path = tmp_path / "t.json"
    S.save_transcript_to_json(tr, str(path))
    back = S.load_transcript_from_json(str(path))
    assert len(back.transcript) == len(tr.transcript)
    h2d = back.transcript[1][1]
    t = h2d.args[0].value.tensor
    assert torch.equal(t, strided) and t.stride() == strided.stride()
    d2h = back.transcript[4][1]
    assert isinstance(d2h.args[0].value.inf_name, str)
    assert np.array_equal(d2h.out.value.tensor, np.arange(6, dtype=np.int32).reshape(2, 3))
    assert back.transcript[0] == (K.SEGMENT_START, "seg0")
    assert back.transcript[2][1].args[1].value == [torch.int32] or back.transcript[2][1].args[1].value is torch.int32
    # slices, ellipsis, complex and dtypes through the generic codec
    obj = [slice(1, 5, 2), Ellipsis, 1 + 2j, torch.float64]
    assert S.loads(S.dumps(obj)) == obj


def test_pickled_payloads_are_refused():
    class Evil:
        def __reduce__(self):
            return (print, ("executed",))

    raw = base64.b64encode(pickle.dumps(Evil())).decode()
    with pytest.raises(Exception):
        S.decode_tensor({"type": "torch", "data": raw})
    buf = io.BytesIO()
    np.save(buf, np.array([Evil()], dtype=object), allow_pickle=True)
    with pytest.raises(ValueError):
        S.decode_tensor({"type": "numpy", "data": base64.b64encode(buf.getvalue()).decode()})
    with pytest.raises(ValueError):
        S.loads(json.dumps({"__type__": "NoSuchType"}))
