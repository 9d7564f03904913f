"""cbclib_v2_amd — the streak-rasterization hot path of cbclib_v2, MI355X-native.

Public surface (drop-ins for the reference's):

* ``cbclib_v2_amd.ndimage.draw_lines`` (cbclib_v2.ndimage.draw_lines), the
  pixel-list mode ``draw_pixels`` and ``accumulate_lines``
  (cbclib_v2.ndimage.accumulate_lines);
* ``cbclib_v2_amd.label.{label, Structure, LabelResult, CPLabelResult,
  labels, index}`` (cbclib_v2.label);
* ``cbclib_v2_amd.device`` — device selection (``'rocm'``);
* ``cbclib_v2_amd.autograd`` — differentiable rendering.

Every operation runs in the HIP library ``libcbx.so`` (C ABI in
``include/cbx.h``); there is no CPU implementation in this package.
"""
from . import device

__version__ = '0.1.0'

__all__ = ['device', 'ndimage', 'label', 'autograd']


def __getattr__(name):
    # lazy submodules: importing the package must not require torch's GPU
    if name in ('ndimage', 'label', 'autograd', 'batch'):
        import importlib
        return importlib.import_module(f'{__name__}.{name}')
    raise AttributeError(name)
