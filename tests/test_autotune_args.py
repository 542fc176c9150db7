"""Host side of the launch-shape autotuner: mode switching, forced shapes and
the report, with their argument checks.  No kernel is launched."""
from snapi_amd import abi, autotune
from snapi_amd._native import lib


def test_mode_force_and_report_arguments():
    L = autotune._decl()
    old = autotune.mode()
    try:
        assert autotune.mode(0) == old
        assert autotune.mode() == 0
        assert autotune.mode(1) == 0
        assert L.snapi_autotune_mode(7) == abi.SNAPI_ERR_ARG
        assert L.snapi_autotune_force(3, 0) == abi.SNAPI_ERR_ARG
        assert L.snapi_autotune_force(0, -2) == abi.SNAPI_ERR_ARG
        assert L.snapi_autotune_force(0, 8) == abi.SNAPI_ERR_ARG
        autotune.force(autotune.SYNTH, 1)
        autotune.force(autotune.SYNTH, -1)
        autotune.clear()
        assert autotune.report() == []
        assert L.snapi_autotune_report(None, 16) == abi.SNAPI_ERR_ARG
    finally:
        autotune.mode(old)
    assert lib() is L
