"""HEAL transcript runtime on the MI355X backend: the on-disk transcript format
(datatypes, serialization), op-name dispatch onto lattica_hw (dispatcher), the
replay loop (runtime) and programs replayed many times with their segments as
hipGraphs (program)."""
from .datatypes import (DeviceOp, DeviceOpArg, DeviceOpArgType, DeviceTensorPointer, ExecutionTranscript,  # noqa: F401
                        ExecutionTranscriptOpType, FreeDeviceTensor, HostTensor)
from .program import TranscriptProgram  # noqa: F401
from .runtime import VerificationError, run_transcript  # noqa: F401
from .serialization import dumps, load_transcript_from_json, loads, save_transcript_to_json  # noqa: F401
