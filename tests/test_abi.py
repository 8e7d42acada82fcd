"""The C-ABI library loads and exports every symbol include/qjx.h declares
(no device calls: this runs without a GPU)."""
import os
import re

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(REPO, "include", "qjx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(qjx_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = _declared()
    for n in ("qjx_ctx_create", "qjx_ensemble_init", "qjx_trans", "qjx_run_periods", "qjx_chars",
              "qjx_get_jumps", "qjx_run_config", "qjx_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol(qjx_lib):
    for n in _declared():
        assert hasattr(qjx_lib, n), n


def test_python_mirror_binds_every_symbol():
    from os_d_amd import qjx
    bound = {s[0] for s in qjx.SIGNATURES}
    assert set(_declared()) <= bound


def test_missing_device_is_an_error_not_a_fallback(qjx_lib):
    import torch  # noqa: F401  (only to know whether a device is visible)
    if torch.cuda.is_available():
        return
    from os_d_amd import qjx
    try:
        qjx.Engine(4, ex_deep=2)
    except qjx.QJXError as e:
        assert e.code in (-2,)
    else:
        raise AssertionError("engine must refuse to run without a HIP device")
