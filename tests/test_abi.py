"""The C-ABI library loads and exports every entry point include/*.h declares."""
import ctypes

import thousandthchicken_amd as tca


def test_library_exports_every_declared_symbol():
    names = tca.exported_symbols()
    assert "tc_launch_t1_decode_lockstep" in names and "tc_decoder_decode" in names
    L = ctypes.CDLL(tca.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_status_strings():
    L = tca.lib()
    for code in range(11):
        assert L.tc_status_string(code)


def test_host_entry_points_reject_bad_arguments():
    L = tca.lib()
    assert L.tc_image_probe(None, 0, None) == 1
    n = ctypes.c_int(0)
    assert L.tc_image_list_cblks(b"\x00\x01", 2, None, 0, ctypes.byref(n)) != 0
    assert L.tc_decoder_create(None, None) == 1
