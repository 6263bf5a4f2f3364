"""CPU-side checks of the C ABI library: it loads, exports every entry point
include/vp8g.h declares, and fails cleanly where no gfx950 device exists."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "vp8g.h"
LIB = ROOT / "oxideav-vp8_amd" / "libvp8g.so"


def declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(vp8g_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared()
    for must in ("vp8g_build_arnr_altref", "vp8g_predict_inter_mb_batch",
                 "vp8g_mb_luma_sad_at_mv_batch", "vp8g_half_pixel_refine_luma_batch",
                 "vp8g_quarter_pixel_refine_luma_batch", "vp8g_ctx_create"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not LIB.exists():
        pytest.fail(f"{LIB} missing: run python oxideav-vp8_amd/build.py")
    lib = C.CDLL(str(LIB))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_wrapper_binds_every_export():
    import vp8g
    assert set(vp8g.EXPORTS) <= set(declared())
    lib = vp8g.load_library()
    for n in vp8g.EXPORTS:
        assert hasattr(lib, n)


def test_status_strings_and_null_context():
    import vp8g
    lib = vp8g.load_library()
    assert lib.vp8g_status_str(0) == b"ok"
    assert lib.vp8g_status_str(2) == b"reference dimensions mismatch"
    # Every entry point rejects a null context before touching anything.
    assert lib.vp8g_build_arnr_altref(None, None, 0, 0, 3, None, None, None, None) == vp8g.VP8G_E_ARG
    assert lib.vp8g_ctx_synchronize(None) == vp8g.VP8G_E_ARG
    out = C.c_void_p()
    assert lib.vp8g_ctx_create(0, None) == vp8g.VP8G_E_ARG


def test_no_device_is_reported_not_faked():
    """Without a gfx950 device the context cannot be created: there is no
    CPU fallback behind the ABI."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import vp8g
    with pytest.raises(vp8g.Vp8gError) as e:
        vp8g.Context(0)
    assert e.value.status == vp8g.VP8G_E_NO_DEVICE
