"""The raymarcher's exit point floors and converts a float in one instruction
(meso::floor_i32 in mesoengine_amd/csrc/traverse.h, v_cvt_flr_i32_f32).  The
parity of every hit cell rests on it giving floor(x) as an int32; this checks
that on the device over every finite binary32 value, both signs
(2 x 2,139,095,040 bit patterns): against v_floor_f32 then v_cvt_i32_f32
(the two-instruction form it replaced), and, for the 2,650,800,129 patterns
in [-2^31, 2^31), against floor(x) computed exactly in integers from the bits.
The probe is tests/native/floor_convert_probe.hip, built by
__graft_entry__.build() into tests/native/_build/.
"""
import ctypes as C
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROBE = os.path.join(ROOT, "tests", "native", "_build", "libfloor_convert_probe.so")
IN_RANGE = 2650800129  # finite floats whose floor lies in [-2^31, 2^31)


@pytest.mark.gpu
def test_floor_convert_instruction_is_floor_over_all_finite_floats():
    assert os.path.exists(PROBE), "build the probe first: python -c 'import __graft_entry__ as g; g.build()'"
    lib = C.CDLL(PROBE)
    lib.floor_convert_probe.restype = C.c_int
    lib.floor_convert_probe.argtypes = [C.POINTER(C.c_ulonglong)]
    out = (C.c_ulonglong * 4)()
    print("floor_convert_probe: all finite floats", flush=True)
    assert lib.floor_convert_probe(out) == 0, "probe launch failed"
    vs_pair, vs_exact, checked, first = list(out)
    assert checked == IN_RANGE
    assert vs_pair == 0 and vs_exact == 0, (vs_pair, vs_exact, hex(first))
