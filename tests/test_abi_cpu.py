"""CPU-side checks of the C-ABI library: it loads, exports every entry point
declared under include/, and its host-side argument handling answers without
a device (no compute call is made here)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "**", "*.h"), recursive=True):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(?:AFAPI|AFHIP_KAPI)\s+[\w\s\*]+?\b(\w+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    import arrayfire_amd

    return arrayfire_amd.lib()


def test_headers_declare_the_reference_entry_points():
    names = declared_symbols()
    for n in ("af_convolve1", "af_convolve2", "af_convolve3", "af_convolve2_sep",
              "af_convolve2_nn", "af_convolve2_gradient_nn", "af_create_array",
              "af_get_data_ptr", "af_release_array", "af_retain_array", "af_get_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_list_matches_headers():
    import arrayfire_amd

    assert sorted(arrayfire_amd.EXPORTED) == declared_symbols()


def test_invalid_handle_is_arg_error_without_device(lib):
    # handles are validated before any device work
    out = ctypes.c_void_p()
    bogus = ctypes.c_void_p(0x1234)
    assert lib.af_convolve2(ctypes.byref(out), bogus, bogus, 0, 1) == 202
    assert lib.af_release_array(ctypes.c_void_p(0)) == 0


def test_enum_values_match_reference():
    # include/af/defines.h values the ABI depends on
    import arrayfire_amd as af

    assert (af.f32, af.c32, af.f64, af.c64, af.b8, af.s32, af.u32, af.u8, af.s64, af.u64,
            af.s16, af.u16, af.f16, af.s8) == tuple(range(14))
    assert (af.AF_CONV_AUTO, af.AF_CONV_SPATIAL, af.AF_CONV_FREQ) == (0, 1, 2)
    assert (af.AF_CONV_GRADIENT_DEFAULT, af.AF_CONV_GRADIENT_FILTER,
            af.AF_CONV_GRADIENT_DATA, af.AF_CONV_GRADIENT_BIAS) == (0, 1, 2, 3)
    hdr = open(os.path.join(ROOT, "include", "af", "defines.h")).read()
    for name, val in (("AF_ERR_ARG", 202), ("AF_ERR_SIZE", 203), ("AF_ERR_TYPE", 204),
                      ("AF_ERR_NOT_SUPPORTED", 301), ("AF_ERR_INTERNAL", 998)):
        assert re.search(rf"\b{name}\s*=\s*{val}\b", hdr)
