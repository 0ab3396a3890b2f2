"""The C ABI library loads on the CPU, exports every function include/*.h declares, and its record
layouts agree with the Python mirror (the role of NativeMetalShaderABITests.swift:61-381).
No compute call is made here: that needs a GPU."""
import ctypes
import re
from pathlib import Path

import numpy as np

from bubblephysics_amd import abi

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for header in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(bp_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_function(hip_library):
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in sorted(names) if not hasattr(hip_library, n)]
    assert not missing, missing
    assert set(abi.SIGNATURES) == names, set(abi.SIGNATURES) ^ names


def test_record_layouts_match_the_header(hip_library):
    sizes = (ctypes.c_uint32 * 9)()
    hip_library.bp_abi_record_sizes(sizes)
    expected = [ctypes.sizeof(t) for t in (abi.Bubble, abi.Segment, abi.Body, abi.Motor, abi.Band, abi.WorldConfig,
                                          abi.FrameInfo, abi.BatchWorldDesc, abi.Field)]
    assert list(sizes) == expected
    assert abi.BUBBLE_DTYPE.itemsize == 72 and abi.SEGMENT_DTYPE.itemsize == 64 and abi.BODY_DTYPE.itemsize == 96
    assert abi.FIELD_DTYPE.itemsize == 64 and abi.CONTACT_DTYPE.itemsize == 64
    server = (ctypes.c_uint32 * 4)()
    hip_library.bp_server_record_sizes(server)
    assert list(server) == [ctypes.sizeof(t) for t in (abi.Request, abi.ServerResult, abi.ServerConfig, abi.ServerStats)]
    assert abi.REQUEST_DTYPE.itemsize == 48 and abi.RESULT_DTYPE.itemsize == 56


def test_staging_records_and_codes_match_the_header():
    # bp_stage_scene and bp_stage_rejection: the sizes bp_stager.cpp asserts at compile time.
    assert ctypes.sizeof(abi.StageScene) == 56 and ctypes.sizeof(abi.StageRejection) == 16
    text = (ROOT / "include" / "bubblephysics.h").read_text()
    for name in ("BP_PLACEMENT_OUTSIDE_REGION", "BP_PLACEMENT_INSIDE_RIGID_BODY", "BP_PLACEMENT_COINCIDENT_CENTER",
                 "BP_REJECT_INVALID_STRUCTURAL_OPERATION", "BP_REJECT_INVALID_PLACEMENT", "BP_REJECT_DUPLICATE_ID",
                 "BP_REJECT_UNKNOWN_OBJECT"):
        (value,) = re.findall(rf"\b{name} = (\d+)", text)
        assert int(value) == getattr(abi, name), name


def test_server_defaults(hip_library):
    cfg = abi.ServerConfig()
    hip_library.bp_server_config_default(ctypes.byref(cfg))
    assert cfg.memory_budget == 0 and cfg.streams >= 1 and cfg.max_inflight_per_class >= 1 and cfg.big_world == 4096


def test_defaults_are_the_reference_defaults(hip_library):
    cfg = abi.default_config()
    # NativeSolverConfiguration.init (NativeSolverInput.swift:128-143), NativeMetalConfiguration defaults.
    assert cfg.time_step == np.float32(1 / 60)
    assert cfg.contact_tolerance == np.float32(0.001) and cfg.separation_tolerance == np.float32(0.002)
    assert cfg.candidate_limit == 128 and cfg.substep_split == 1


def test_status_strings_cover_every_code(hip_library):
    for code, name in abi.STATUS_NAMES.items():
        assert hip_library.bp_status_string(code)
    assert hip_library.bp_status_string(999) == b"unknown status"


def test_rigid_solver_records_match_the_reference_layouts():
    # NativeRigidBody (160 B), NativeJointRecord (96 B), NativeRigidCorrection-sized frame records.
    assert ctypes.sizeof(abi.RigidBody) == 160 and abi.RIGID_BODY_DTYPE.itemsize == 160
    assert ctypes.sizeof(abi.Joint) == 96 and abi.JOINT_DTYPE.itemsize == 96
    assert ctypes.sizeof(abi.RigidFrameInfo) == 32 and ctypes.sizeof(abi.RigidConfig) == 40
    text = (ROOT / "include" / "bubblephysics.h").read_text()
    for name in ("BP_JOINT_DISTANCE", "BP_JOINT_HINGE", "BP_JOINT_WELD", "BP_JOINT_HAS_LIMITS", "BP_JOINT_HAS_MOTOR",
                 "BP_GUARD_NONE", "BP_GUARD_RIGID_ROTATION_BUDGET", "BP_CONTACT_RIGID"):
        (value,) = re.findall(rf"\b{name} = (\d+)", text)
        assert int(value) == getattr(abi, name), name
